// wos_bluestein.hpp — Bluestein (chirp-z) transforms of any length L through
// power-of-two FFTs of length M >= 2L - 1 (fft1DBluestein,
// optics/fft.go:159-204), on gfx950.
//
// For direction sign s (-1 forward, +1 inverse) with chirp c[k] =
// exp(i s pi k^2 / L) (argument formed (s*pi)*k^2/L in FP64, as the
// reference does):
//   X[k] = c[k] * IFFT_M( FFT_M(x c) * FFT_M(b) )[k],  b[k] = conj(c[|k|]) wrapped,
// scaled by 1/L for the inverse. The FFT_M(b)/M spectrum is precomputed per
// (L, direction) on the device in FP64. Inside a kernel the whole chain
// stays in registers: the forward plan's input map and the inverse plan's
// output map are both the radix-16 slot map, so chirp -> FFT -> x B ->
// IFFT -> chirp needs no extra LDS round trip, and the fused column kernel
// chains a forward and an inverse Bluestein around the transfer function.
#pragma once
#include "wos_kernels.hpp"

namespace wos {

// Tables of one (L, direction): chirp[L] and bspec[M] (= FFT_M(b)/M).
struct BsTab {
  const void* chirp;
  const void* bspec;
};

// Bluestein transform of the values in v (slot map of first_rad<M, -1>:
// slot (b, r) holds position t + b T + r M/R0). On entry positions p >= L
// must be zero; on exit positions p < L hold X[p] * post (others are zero).
template <typename R, int M, int ES>
__device__ __forceinline__ void bluestein_block(cpx<R>* v, int t, int L, const BsTab& tab, R post,
                                                const cpx<R>* __restrict__ tw, cpx<R>* lds) {
  constexpr int V = Plan<M>::V;
  constexpr int R0 = first_rad<M, -1>();
  constexpr int RL = last_rad<M, -1>();
  static_assert(last_rad<M, +1>() == R0, "forward input and inverse output maps must coincide");
  const cpx<R>* __restrict__ chirp = (const cpx<R>*)tab.chirp;
  const cpx<R>* __restrict__ bspec = (const cpx<R>*)tab.bspec;
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int p = slot_pos<M, R0>(t, b, r);
      v[b * R0 + r] = p < L ? cmul(v[b * R0 + r], chirp[p]) : cpx<R>{R(0), R(0)};
    }
  fft_block<R, M, -1, ES>(v, t, tw, lds);
#pragma unroll
  for (int b = 0; b < V / RL; ++b)
#pragma unroll
    for (int r = 0; r < RL; ++r) v[b * RL + r] = cmul(v[b * RL + r], bspec[slot_pos<M, RL>(t, b, r)]);
  fft_block<R, M, +1, ES>(v, t, tw, lds);
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int p = slot_pos<M, R0>(t, b, r);
      v[b * R0 + r] = p < L ? cscale(cmul(v[b * R0 + r], chirp[p]), post) : cpx<R>{R(0), R(0)};
    }
}

// Row pass of length-L Bluestein transforms (L = a.len). Direction and
// scale come from the table and bs_post (1 or 1/L); a.scale_* multiplies
// the outputs as in k_rows.
struct BsRowArgs {
  RowArgs io;
  int len;
  BsTab tab;
  double post;
};

template <typename R, int M, int ROWS>
__global__ __launch_bounds__(ROWS* Plan<M>::T) void k_rows_bs(const BsRowArgs A) {
  const RowArgs& a = A.io;
  constexpr int T = Plan<M>::T, V = Plan<M>::V;
  constexpr int R0 = first_rad<M, -1>();
  constexpr int NT = ROWS * T;
  constexpr int LS = row_lds_stride<R>(M, ROWS);
  __shared__ cpx<R> lds[ROWS * LS];
  __shared__ double red[(NT + 63) / 64];
  const int tid = threadIdx.x;
  const int lr = tid % ROWS, t = tid / ROWS;
  const int bpu = (a.rows + ROWS - 1) / ROWS;
  const int unit = blockIdx.x / bpu;
  const int row = (blockIdx.x % bpu) * ROWS + lr;
  const bool active = row < a.rows && unit < a.units;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit;
  const int L = A.len;

  cpx<R> v[V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int p = slot_pos<M, R0>(t, b, r);
      const int q = p - a.in_col0;
      cpx<R> x = {R(0), R(0)};
      if (active && p < L && q >= 0 && q < a.in_cols) x = in[tiled_offset(q, row, a.in_ld, a.in_tw, a.in_tstride)];
      v[b * R0 + r] = x;
    }
  if (a.pw_in) {
    double pin = 0;
#pragma unroll
    for (int i = 0; i < V; ++i) pin += (double)v[i].x * (double)v[i].x + (double)v[i].y * (double)v[i].y;
    block_sum_store<NT>(pin, red, a.pw_in, blockIdx.x);
  }
  bluestein_block<R, M, 1>(v, t, L, A.tab, (R)A.post, (const cpx<R>*)a.tw, lds + lr * LS);
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  if (a.scale_im != 0 || a.scale_re != 1) {
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = cmul(v[i], sc);
  }
  double pout = 0;
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int p = slot_pos<M, R0>(t, b, r);
      const int q = p - a.out_col0;
      if (active && p < L && q >= 0 && q < a.out_cols) {
        const cpx<R> y = v[b * R0 + r];
        out[tiled_offset(q, row, a.out_ld, a.out_tw, a.out_tstride)] = y;
        pout += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
      }
    }
  if (a.pw_out) block_sum_store<NT>(pout, red, a.pw_out, blockIdx.x);
}

// Column pass of length-L Bluestein transforms over W-column tiles. MODE 0:
// one transform with table tab0 (scale post0, then a.scale); MODE 2:
// forward transform (tab0) -> transfer function at (column, row) -> inverse
// transform (tab1, scale post1).
struct BsColArgs {
  ColArgs io;
  int len;
  BsTab tab0, tab1;
  double post0, post1;
};

template <typename R, int M, int W, int MODE>
__global__ __launch_bounds__(W* Plan<M>::T) void k_cols_bs(const BsColArgs A) {
  const ColArgs& a = A.io;
  constexpr int V = Plan<M>::V;
  constexpr int R0 = first_rad<M, -1>();
  __shared__ cpx<R> lds[lds_len(M) * W];
  const int tid = threadIdx.x;
  const int c = tid % W, t = tid / W;
  const int tpu = a.cols / W;
  const int unit = blockIdx.x / tpu;
  const int tile = blockIdx.x % tpu;
  const int col = tile * W + c;
  cpx<R>* base = (cpx<R>*)a.data + (size_t)unit * a.unit + (a.tiled ? tile * a.tstride + c : col);
  const int ld = a.tiled ? W : a.ld;
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  const bool active = unit < a.units;
  const int L = A.len;

  cpx<R> v[V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int p = slot_pos<M, R0>(t, b, r);
      const int q = p - a.in_row0;
      cpx<R> x = {R(0), R(0)};
      if (active && p < L && q >= 0 && q < a.rows) x = base[q * ld];
      v[b * R0 + r] = x;
    }
  bluestein_block<R, M, W>(v, t, L, A.tab0, (R)A.post0, tw, lds + c);
  if constexpr (MODE == 2) {
    if (active) {
      const TFUnit& u = a.tf[unit];  // global: a local copy would be spilled for the out-of-line calls
#pragma unroll
      for (int b = 0; b < V / R0; ++b)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int p = slot_pos<M, R0>(t, b, r);
          if (p < L) v[b * R0 + r] = cmul(v[b * R0 + r], tf_eval<R>(u, col, p));
        }
    }
    bluestein_block<R, M, W>(v, t, L, A.tab1, (R)A.post1, tw, lds + c);
  }
  const R sc = (R)a.scale;
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int p = slot_pos<M, R0>(t, b, r);
      const int q = p - a.out_row0;
      if (active && p < L && q >= 0 && q < a.out_rows) base[q * ld] = cscale(v[b * R0 + r], sc);
    }
}

}  // namespace wos
