// wos_mixreg.hpp — register-resident mixed-radix row and column passes for sides
// whose transform splits into stages of equal work per thread, so that each
// thread keeps V values of a column in registers from the first global load
// to the last store (the power-of-two engine's structure), with one LDS
// exchange between stages. The LDS-resident mixed-radix kernels
// (wos_mixed.hpp) read and write every stage through LDS and leave threads
// idle on stages with fewer butterflies than threads.
//
// Built for N = 2000 = 10 * 10 * 10 * 2 (the zero-padded side of 1000^2
// asm_pad): V = 10 values on T = 200 threads per column, forward stages
// [10, 10, 10, 2] and inverse stages [2, 10, 10, 10], so the forward plan's
// last butterfly map is the inverse plan's first and forward -> transfer
// function -> inverse keeps the column in registers.
//
// Stockham stage (radix RAD, sub-transform length NS, butterfly j = t + b T):
//   k = j mod NS; v[r] = x[j + r N/RAD] * w^(r k), w = exp(dir 2 pi i/(NS RAD));
//   v = DFT_RAD(v);  y[(j - k) RAD + k + r NS] = v[r].
// Twiddles: forward stages 1..S-1 then inverse stages 1..S-1, entry [k][r-1]
// = exp(-2 pi i r k/(NS RAD)) (conjugated in the inverse direction), placed
// after the LDS-resident plan's table of the same side.
#pragma once
#include "wos_mixed.hpp"

namespace wos {

// Stage radices of the register-resident plans (forward order).
template <int N>
struct MRDef {
  static constexpr int S = 0;
};
template <>
struct MRDef<2000> {
  static constexpr int S = 4, V = 10;
  static constexpr bool F32 = false, F64 = true, F32C = false;  // FP32: LDS-resident
  static constexpr int rad(int s) { return s == 3 ? 2 : 10; }
};
template <>
struct MRDef<1536> {
  static constexpr int S = 4, V = 24;
  static constexpr bool F32 = true, F64 = false, F32C = true;
  static constexpr int rad(int s) { return s == 3 ? 3 : 8; }
};
template <>
struct MRDef<1000> {
  static constexpr int S = 3, V = 10;
  static constexpr bool F32 = true, F64 = true, F32C = true;
  static constexpr int rad(int s) { return 10; }
};

template <int N>
constexpr bool mr_built() { return MRDef<N>::S > 0; }
// The register-resident kernels run the 2000-point FP64 transforms: a
// 1000^2 asm_pad Jones propagation (two components) takes 0.133 ms against
// 0.178 ms on the LDS-resident kernels, at one column per column workgroup.
// In FP32 both passes are slower register-resident (rows 23 against 17 us
// in four-row workgroups, the column pass 52 against 48 us), so FP32 keeps
// the LDS-resident kernels there. Rows (F32, F64) and columns (F32C, F64)
// are chosen apart: each takes natural order in and out, so either pass
// may use either engine.
template <typename R, int N>
constexpr bool mr_use_rows() {
  if constexpr (mr_built<N>()) return sizeof(R) == 8 ? MRDef<N>::F64 : MRDef<N>::F32;
  else return false;
}
template <typename R, int N>
constexpr bool mr_use_cols() {
  if constexpr (mr_built<N>()) return sizeof(R) == 8 ? MRDef<N>::F64 : MRDef<N>::F32C;
  else return false;
}

template <int N>
struct MRPlan {
  static constexpr int S = MRDef<N>::S, V = MRDef<N>::V, T = N / V;
  static constexpr int fwd_rad(int s) { return MRDef<N>::rad(s); }
  static constexpr int inv_rad(int s) { return MRDef<N>::rad(S - 1 - s); }
  static constexpr int rad(int dir, int s) { return dir < 0 ? fwd_rad(s) : inv_rad(s); }
  static constexpr int ns(int dir, int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= rad(dir, i);
    return p;
  }
  // table offsets, after the LDS-resident plan's table (MPlan<N>::tw_total())
  static constexpr int tw_off(int dir, int s) {
    int o = MPlan<N>::tw_total();
    if (dir > 0)
      for (int i = 1; i < S; ++i) o += ns(-1, i) * (fwd_rad(i) - 1);
    for (int i = 1; i < s; ++i) o += ns(dir, i) * (rad(dir, i) - 1);
    return o;
  }
};

// twiddles and the DFT of stage S on the thread's butterflies
template <typename R, int N, int DIR, int S>
__device__ __forceinline__ void mr_compute(cpx<R>* v, int t, const cpx<R>* __restrict__ tw) {
  using P = MRPlan<N>;
  constexpr int RAD = P::rad(DIR, S), NS = P::ns(DIR, S), OFF = P::tw_off(DIR, S);
#pragma unroll
  for (int b = 0; b < P::V / RAD; ++b) {
    if constexpr (NS > 1) {
      const int k = (t + b * P::T) % NS;
      const cpx<R>* tk = tw + OFF + k * (RAD - 1);
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[b * RAD + r] = DIR < 0 ? cmul(v[b * RAD + r], tk[r - 1]) : cmulc(v[b * RAD + r], tk[r - 1]);
    }
    dft_mixed<RAD, DIR>(v + b * RAD);
  }
}

// the exchange after stage S: its outputs to LDS (element p at lds[lpad(p) ES]),
// then the inputs of stage S + 1
template <typename R, int N, int DIR, int S, int ES>
__device__ __forceinline__ void mr_exchange(cpx<R>* v, int t, cpx<R>* lds) {
  using P = MRPlan<N>;
  constexpr int RAD = P::rad(DIR, S), NS = P::ns(DIR, S), RN = P::rad(DIR, S + 1);
  __syncthreads();  // earlier readers of lds are done
#pragma unroll
  for (int b = 0; b < P::V / RAD; ++b) {
    const int j = t + b * P::T, k = j % NS;
#pragma unroll
    for (int r = 0; r < RAD; ++r) lds[lpad((j - k) * RAD + k + r * NS) * ES] = v[b * RAD + r];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < P::V / RN; ++b) {
    const int j = t + b * P::T;
#pragma unroll
    for (int r = 0; r < RN; ++r) v[b * RN + r] = lds[lpad(j + r * (N / RN)) * ES];
  }
}

template <typename R, int N, int DIR, int S, int ES>
__device__ __forceinline__ void mr_stages(cpx<R>* v, int t, const cpx<R>* __restrict__ tw, cpx<R>* lds) {
  if constexpr (S < MRPlan<N>::S) {
    mr_compute<R, N, DIR, S>(v, t, tw);
    if constexpr (S + 1 < MRPlan<N>::S) {
      mr_exchange<R, N, DIR, S, ES>(v, t, lds);
      mr_stages<R, N, DIR, S + 1, ES>(v, t, tw, lds);
    }
  }
}

// transform position held by slot (b, r) on entry to a stage of radix RAD,
// and on exit from the last stage of radix RAD
template <int N, int RAD>
__device__ __forceinline__ int mr_pos(int t, int b, int r) {
  return t + b * MRPlan<N>::T + r * (N / RAD);
}

// ---------------------------------------------------------------------------
// Row pass: ROWS rows per workgroup (thread t of row lr at tid = t ROWS + lr,
// so a wave load covers 64 / ROWS consecutive positions of each of its
// rows), RowArgs as the other row passes: embed window in_col0 / in_cols,
// output window out_col0 / out_cols, tiled or row-major layouts, complex
// output scale, per-workgroup FP64 power sums.
// ---------------------------------------------------------------------------
template <typename R>
constexpr int mr_rows() { return sizeof(R) == 4 ? 4 : 2; }  // 68 KB of LDS either way

template <typename R, int N, int DIR, int ROWS>
__global__ __launch_bounds__(ROWS * MRPlan<N>::T) void k_rows_mr(const RowArgs a) {
  using P = MRPlan<N>;
  constexpr int T = P::T, V = P::V, NT = ROWS * T;
  constexpr int LS = lds_len(N) | 1;  // odd stride between the rows' LDS areas
  __shared__ cpx<R> lds[ROWS * LS];
  __shared__ double red[(NT + 63) / 64];
  const int tid = threadIdx.x;
  const int lr = tid % ROWS, t = tid / ROWS;
  const int bpu = (a.rows + ROWS - 1) / ROWS;
  const int unit = blockIdx.x / bpu;
  const int row = (blockIdx.x % bpu) * ROWS + lr;
  const bool active = row < a.rows;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit;
  constexpr int R0 = P::rad(DIR, 0), RL = P::rad(DIR, P::S - 1);
  cpx<R> v[V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int q = mr_pos<N, R0>(t, b, r) - a.in_col0;
      cpx<R> x = {R(0), R(0)};
      if (active && q >= 0 && q < a.in_cols) x = in[tiled_offset(q, row, a.in_ld, a.in_tw, a.in_tstride)];
      v[b * R0 + r] = x;
    }
  if (a.pw_in) {  // uniform: FP64 power of the input (zeros outside the window add nothing)
    double pin = 0;
#pragma unroll
    for (int i = 0; i < V; ++i) pin += (double)v[i].x * (double)v[i].x + (double)v[i].y * (double)v[i].y;
    block_sum_store<NT>(pin, red, a.pw_in, blockIdx.x);
  }
  mr_stages<R, N, DIR, 0, 1>(v, t, (const cpx<R>*)a.tw, lds + lr * LS);
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  const bool scale = a.scale_im != 0 || a.scale_re != 1;
  double pout = 0;
#pragma unroll
  for (int b = 0; b < V / RL; ++b)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int q = mr_pos<N, RL>(t, b, r) - a.out_col0;
      if (active && q >= 0 && q < a.out_cols) {
        cpx<R> y = v[b * RL + r];
        if (scale) y = cmul(y, sc);
        pout += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
        out[tiled_offset(q, row, a.out_ld, a.out_tw, a.out_tstride)] = y;
      }
    }
  if (a.pw_out) block_sum_store<NT>(pout, red, a.pw_out, blockIdx.x);  // uniform
}

// ---------------------------------------------------------------------------
// Column pass over W adjacent columns, ColArgs as the other column passes
// (tiled or row-major, input window in_row0/rows, output crop
// out_row0/out_rows, scale, XCD-aware order). MODE 0: one transform in
// direction DIR; MODE 2: forward -> transfer function (any kind; mirrored
// through LDS when even in fy) -> inverse. Threads are interleaved over the
// columns (c = tid % W) so a row of the tile is read as one W-element segment.
// ---------------------------------------------------------------------------
template <typename R, int N, int W, int MODE, int DIR>
__global__ __launch_bounds__(W * MRPlan<N>::T) void k_cols_mr(const ColArgs a) {
  using P = MRPlan<N>;
  constexpr int T = P::T, V = P::V;
  __shared__ cpx<R> lds[lds_len(N) * W];
  const int tid = threadIdx.x;
  const int c = tid % W, t = tid / W;
  const int tpu = a.cols / W;
  const int bid = a.xcd_remap ? remap_xcd(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int unit = bid / tpu;
  const int col = (bid % tpu) * W + c;
  const int tl = a.tiled ? 1 << a.tlg : 0;
  const int ld = a.tiled ? tl : a.ld;
  cpx<R>* base =
      (cpx<R>*)a.data + (size_t)unit * a.unit + (a.tiled ? (size_t)(col >> a.tlg) * a.tstride + (col & (tl - 1)) : col);
  cpx<R>* lc = lds + c;
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  constexpr int DIR0 = MODE == 2 ? -1 : DIR;
  constexpr int R0 = P::rad(DIR0, 0);
  cpx<R> v[V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int q = mr_pos<N, R0>(t, b, r) - a.in_row0;
      cpx<R> x = {R(0), R(0)};
      if (q >= 0 && q < a.rows) x = base[(size_t)q * ld];
      v[b * R0 + r] = x;
    }
  mr_stages<R, N, DIR0, 0, W>(v, t, tw, lc);
  if constexpr (MODE == 2) {
    constexpr int RL = P::rad(-1, P::S - 1), NB = V / RL, V2 = V / 2;
    static_assert(RL == P::rad(1, 0), "fused forward -> inverse map");
    const TFUnit& u = a.tf[unit];
    const TF32Col cc = tf32_col(u, col);  // FP32 asm only
    // thread t holds rows p = t + T m in slot (m % NB) RL + m / NB
    auto tf_at = [&](int p) -> cpx<R> {
      if (u.kind == TF_ASM || u.kind == TF_CONJ_ASM) {
        if constexpr (sizeof(R) == 4) {
          return asm_tf_fp32_full<true>(u, cc, p);  // cc carries the conjugate's sign
        } else {
          cpx<double> h = asm_tf_fp64_inl(u, col, p);
          if (u.kind == TF_CONJ_ASM) h.y = -h.y;
          return h;
        }
      }
      return tf_eval_t<R, true>(u, col, p);
    };
    // even in fy: row N - p (thread T - t, m' = V-1-m; thread 0: m' = V-m)
    // takes row p's value, passed on through LDS
    const bool mirror = u.fy_off == 0.0 && (u.kind == TF_ASM || u.kind == TF_CONJ_ASM || u.kind == TF_FRESNEL ||
                                            u.kind == TF_BANDLIMIT);
    if (mirror) {
      const int pt = t == 0 ? 0 : T - t;
      __syncthreads();  // the forward transform's last readers of lds are done
#pragma unroll
      for (int m = 0; m < V2; ++m) {
        const cpx<R> h = tf_at(t + T * m);
        v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], h);
        lc[(m * T + t) * W] = h;
      }
      if (t == 0) {  // row N/2 is its own mirror
        constexpr int m = V2;
        v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], tf_at(T * m));
      }
      __syncthreads();
#pragma unroll
      for (int m = V2; m < V; ++m) {
        if (t == 0 && m == V2) continue;
        const int ms = t == 0 ? V - m : V - 1 - m;
        v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], lc[(ms * T + pt) * W]);
      }
    } else {
#pragma unroll
      for (int m = 0; m < V; ++m) v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], tf_at(t + T * m));
    }
    mr_stages<R, N, +1, 0, W>(v, t, tw, lc);
  }
  constexpr int DIRL = MODE == 2 ? +1 : DIR;
  constexpr int RL = P::rad(DIRL, P::S - 1);
  const R sc = (R)a.scale;
#pragma unroll
  for (int b = 0; b < V / RL; ++b)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int q = mr_pos<N, RL>(t, b, r) - a.out_row0;
      if (q >= 0 && q < a.out_rows) {
        cpx<R> y = v[b * RL + r];
        if (sc != R(1)) y = cscale(y, sc);
        base[(size_t)q * ld] = y;
      }
    }
}

}  // namespace wos
