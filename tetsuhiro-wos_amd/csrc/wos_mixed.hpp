// wos_mixed.hpp — mixed-radix (2, 3, 5, 7) Stockham FFT for grid sides
// N = 2^a 3^b 5^c 7^d that are not powers of two (config grids 384, 1000,
// 1536 and their zero-padded 768, 2000, 3072), in place of Bluestein's three
// power-of-two transforms of length >= 2N - 1.
//
// A transform lives in LDS for its whole duration. Stage s (radix RAD,
// sub-transform length NS) has N/RAD butterflies; thread t takes butterflies
// t, t + T, ... (the map changes from stage to stage, so there is no
// register-resident chain across stages as in the power-of-two kernels):
//   load:  v[r] = x[j + r N/RAD]                 (j the butterfly, k = j mod NS)
//   twiddle v[r] *= w^(r k), w = exp(dir 2 pi i/(NS RAD)), from a per-stage table
//   v = DFT_RAD(v)
//   store: y[(j - k) RAD + k + r NS] = v[r]
// with a barrier between the loads and the stores of a stage. The table
// layout is the power-of-two engine's: entry [k][r-1] of a stage at its
// offset, conjugated for the inverse direction.
#pragma once
#include "wos_kernels.hpp"

namespace wos {

// exp(DIR 2 pi i m/3) and exp(DIR 2 pi i m/5) butterflies
template <int DIR, typename R>
__device__ __forceinline__ void dft3(cpx<R>* v) {
  constexpr double S3 = 0.86602540378443864676;  // sin(2 pi/3)
  const cpx<R> a = v[1] + v[2], b = v[1] - v[2];
  const cpx<R> t = {v[0].x - R(0.5) * a.x, v[0].y - R(0.5) * a.y};
  const cpx<R> u = mul_i<DIR>(cscale(b, (R)S3));
  v[0] = v[0] + a;
  v[1] = t + u;
  v[2] = t - u;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft5(cpx<R>* v) {
  constexpr double C1 = 0.30901699437494742410;   // cos(2 pi/5)
  constexpr double C2 = -0.80901699437494742410;  // cos(4 pi/5)
  constexpr double S1 = 0.95105651629515357212;   // sin(2 pi/5)
  constexpr double S2 = 0.58778525229247312917;   // sin(4 pi/5)
  const cpx<R> a1 = v[1] + v[4], b1 = v[1] - v[4], a2 = v[2] + v[3], b2 = v[2] - v[3];
  const cpx<R> x0 = v[0];
  const cpx<R> t1 = {x0.x + (R)C1 * a1.x + (R)C2 * a2.x, x0.y + (R)C1 * a1.y + (R)C2 * a2.y};
  const cpx<R> t2 = {x0.x + (R)C2 * a1.x + (R)C1 * a2.x, x0.y + (R)C2 * a1.y + (R)C1 * a2.y};
  const cpx<R> u1 = mul_i<DIR>(cpx<R>{(R)S1 * b1.x + (R)S2 * b2.x, (R)S1 * b1.y + (R)S2 * b2.y});
  const cpx<R> u2 = mul_i<DIR>(cpx<R>{(R)S2 * b1.x - (R)S1 * b2.x, (R)S2 * b1.y - (R)S1 * b2.y});
  v[0] = x0 + a1 + a2;
  v[1] = t1 + u1;
  v[4] = t1 - u1;
  v[2] = t2 + u2;
  v[3] = t2 - u2;
}

// 7 points: with a_k = v_k + v_(7-k), b_k = v_k - v_(7-k) (k = 1..3),
// X_m = v_0 + sum_k cos(2 pi mk/7) a_k + DIR i sum_k sin(2 pi mk/7) b_k and
// X_(7-m) the same with the sine part negated (m = 1..3)
template <int DIR, typename R>
__device__ __forceinline__ void dft7(cpx<R>* v) {
  constexpr double C1 = 0.62348980185873353053, C2 = -0.22252093395631440429, C3 = -0.90096886790241912624;
  constexpr double S1 = 0.78183148246802980871, S2 = 0.97492791218182360702, S3 = 0.43388373911755812048;
  const cpx<R> a1 = v[1] + v[6], b1 = v[1] - v[6], a2 = v[2] + v[5], b2 = v[2] - v[5], a3 = v[3] + v[4],
               b3 = v[3] - v[4];
  const cpx<R> x0 = v[0];
  const cpx<R> t1 = {x0.x + (R)C1 * a1.x + (R)C2 * a2.x + (R)C3 * a3.x, x0.y + (R)C1 * a1.y + (R)C2 * a2.y + (R)C3 * a3.y};
  const cpx<R> t2 = {x0.x + (R)C2 * a1.x + (R)C3 * a2.x + (R)C1 * a3.x, x0.y + (R)C2 * a1.y + (R)C3 * a2.y + (R)C1 * a3.y};
  const cpx<R> t3 = {x0.x + (R)C3 * a1.x + (R)C1 * a2.x + (R)C2 * a3.x, x0.y + (R)C3 * a1.y + (R)C1 * a2.y + (R)C2 * a3.y};
  const cpx<R> u1 = mul_i<DIR>(
      cpx<R>{(R)S1 * b1.x + (R)S2 * b2.x + (R)S3 * b3.x, (R)S1 * b1.y + (R)S2 * b2.y + (R)S3 * b3.y});
  const cpx<R> u2 = mul_i<DIR>(
      cpx<R>{(R)S2 * b1.x - (R)S3 * b2.x - (R)S1 * b3.x, (R)S2 * b1.y - (R)S3 * b2.y - (R)S1 * b3.y});
  const cpx<R> u3 = mul_i<DIR>(
      cpx<R>{(R)S3 * b1.x - (R)S1 * b2.x + (R)S2 * b3.x, (R)S3 * b1.y - (R)S1 * b2.y + (R)S2 * b3.y});
  v[0] = x0 + a1 + a2 + a3;
  v[1] = t1 + u1;
  v[6] = t1 - u1;
  v[2] = t2 + u2;
  v[5] = t2 - u2;
  v[3] = t3 + u3;
  v[4] = t3 - u3;
}

// 10 points as 2 x 5: n = 2 n1 + n2; DFT-5 over n1 for each n2, twiddle
// W10^(n2 k1), DFT-2 over n2; X[k1 + 5 k2].
template <int DIR, typename R>
__device__ __forceinline__ void dft10(cpx<R>* v) {
  constexpr double C36 = 0.80901699437494742410, S36 = 0.58778525229247312917;  // cos, sin 36 degrees
  constexpr double C72 = 0.30901699437494742410, S72 = 0.95105651629515357212;  // cos, sin 72 degrees
  constexpr double cs[5] = {1, C36, C72, -C72, -C36}, sn[5] = {0, S36, S72, S72, S36};
  cpx<R> a[5], b[5];
#pragma unroll
  for (int n1 = 0; n1 < 5; ++n1) {
    a[n1] = v[2 * n1];
    b[n1] = v[2 * n1 + 1];
  }
  dft5<DIR>(a);
  dft5<DIR>(b);
#pragma unroll
  for (int k = 1; k < 5; ++k) {
    const R c = (R)cs[k], s = (R)(DIR < 0 ? -sn[k] : sn[k]);
    b[k] = {c * b[k].x - s * b[k].y, s * b[k].x + c * b[k].y};
  }
#pragma unroll
  for (int k1 = 0; k1 < 5; ++k1) {
    v[k1] = a[k1] + b[k1];
    v[k1 + 5] = a[k1] - b[k1];
  }
}

template <int RAD, int DIR, typename R>
__device__ __forceinline__ void dft_mixed(cpx<R>* v) {
  if constexpr (RAD == 3) dft3<DIR>(v);
  else if constexpr (RAD == 5) dft5<DIR>(v);
  else if constexpr (RAD == 7) dft7<DIR>(v);
  else if constexpr (RAD == 10) dft10<DIR>(v);
  else dft<RAD, DIR, R>(v);
}

// ---------------------------------------------------------------------------
// Mixed-radix plan: radix-16 stages while four factors of two remain, one
// stage for the leftover power of two (2, 4 or 8), then radix-3 and radix-5
// stages. T threads per transform.
// ---------------------------------------------------------------------------
constexpr int count_factor(int n, int f) { return n % f == 0 ? 1 + count_factor(n / f, f) : 0; }
constexpr int pow_int(int b, int e) { return e == 0 ? 1 : b * pow_int(b, e - 1); }
constexpr bool mixed_len(int n) {
  return n > 1 && pow_int(2, count_factor(n, 2)) * pow_int(3, count_factor(n, 3)) * pow_int(5, count_factor(n, 5)) == n;
}

template <int N>
struct MPlan {
  static_assert(mixed_len(N), "N = 2^a 3^b 5^c");
  static constexpr int A = count_factor(N, 2), B = count_factor(N, 3), C = count_factor(N, 5);
  static constexpr int N16 = A / 4, REM2 = 1 << (A % 4);
  static constexpr int S = N16 + (REM2 > 1 ? 1 : 0) + B + C;
  static constexpr int T = N >= 1500 ? 256 : 128;  // threads per transform
  static constexpr int rad(int s) {
    return s < N16 ? 16 : (REM2 > 1 && s == N16) ? REM2 : (s < N16 + (REM2 > 1 ? 1 : 0) + B ? 3 : 5);
  }
  static constexpr int ns(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= rad(i);
    return p;
  }
  static constexpr int tw_off(int s) {  // stages 1 .. S-1 (stage 0 has NS = 1)
    int o = 0;
    for (int i = 1; i < s; ++i) o += ns(i) * (rad(i) - 1);
    return o;
  }
  static constexpr int tw_total() { return tw_off(S); }
  static constexpr int bpt(int s) { return (N / rad(s) + T - 1) / T; }  // butterflies per thread
};

// One stage on a transform in LDS (element p at lds[lpad(p) * ES]).
// Every thread of the workgroup calls it (it holds two barriers).
template <typename R, int N, int DIR, int S, int ES>
__device__ __forceinline__ void mixed_stage(cpx<R>* lds, int t, const cpx<R>* __restrict__ tw) {
  using P = MPlan<N>;
  constexpr int RAD = P::rad(S), NS = P::ns(S), NB = N / RAD, BPT = P::bpt(S), T = P::T;
  constexpr int OFF = P::tw_off(S);
  cpx<R> v[BPT][RAD];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * T;
    if (NB % T == 0 || j < NB) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[b][r] = lds[lpad(j + r * NB) * ES];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * T;
    if (NB % T == 0 || j < NB) {
      const int k = j % NS;
      if constexpr (NS > 1) {
        const cpx<R>* tk = tw + OFF + k * (RAD - 1);
#pragma unroll
        for (int r = 1; r < RAD; ++r) v[b][r] = DIR < 0 ? cmul(v[b][r], tk[r - 1]) : cmulc(v[b][r], tk[r - 1]);
      }
      dft_mixed<RAD, DIR>(v[b]);
      const int o = (j - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) lds[lpad(o + r * NS) * ES] = v[b][r];
    }
  }
  __syncthreads();
}

template <typename R, int N, int DIR, int S, int ES>
__device__ __forceinline__ void mixed_stages(cpx<R>* lds, int t, const cpx<R>* __restrict__ tw) {
  if constexpr (S < MPlan<N>::S) {
    mixed_stage<R, N, DIR, S, ES>(lds, t, tw);
    mixed_stages<R, N, DIR, S + 1, ES>(lds, t, tw);
  }
}

// Full N-point transform of the LDS-resident data (natural order in and
// out); the caller synchronises before (data written) and gets a barrier
// after the last stage.
template <typename R, int N, int DIR, int ES>
__device__ __forceinline__ void mixed_fft(cpx<R>* lds, int t, const cpx<R>* __restrict__ tw) {
  mixed_stages<R, N, DIR, 0, ES>(lds, t, tw);
}

// ---------------------------------------------------------------------------
// Row pass: one row per workgroup (T threads), RowArgs as the power-of-two
// row kernels (embed window, tiled or row-major layouts, output scale,
// per-workgroup FP64 power sums).
// ---------------------------------------------------------------------------
template <typename R, int N, int DIR>
__global__ __launch_bounds__(MPlan<N>::T) void k_rows_mixed(const RowArgs a) {
  constexpr int T = MPlan<N>::T;
  __shared__ cpx<R> lds[lds_len(N)];
  __shared__ double red[(T + 63) / 64];
  const int t = threadIdx.x;
  const int unit = blockIdx.x / a.rows;
  const int row = blockIdx.x % a.rows;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit;
  double pin = 0;
  for (int p = t; p < N; p += T) {
    const int q = p - a.in_col0;
    cpx<R> x = {R(0), R(0)};
    if (q >= 0 && q < a.in_cols) x = in[tiled_offset(q, row, a.in_ld, a.in_tw, a.in_tstride)];
    pin += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    lds[lpad(p)] = x;
  }
  if (a.pw_in) block_sum_store<T>(pin, red, a.pw_in, blockIdx.x);  // uniform
  __syncthreads();
  mixed_fft<R, N, DIR, 1>(lds, t, (const cpx<R>*)a.tw);
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  const bool scale = a.scale_im != 0 || a.scale_re != 1;
  double pout = 0;
  for (int p = t; p < N; p += T) {
    const int q = p - a.out_col0;
    if (q >= 0 && q < a.out_cols) {
      cpx<R> y = lds[lpad(p)];
      if (scale) y = cmul(y, sc);
      pout += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
      out[tiled_offset(q, row, a.out_ld, a.out_tw, a.out_tstride)] = y;
    }
  }
  if (a.pw_out) block_sum_store<T>(pout, red, a.pw_out, blockIdx.x);  // uniform
}

// The transfer function at transform row p of column col for the mixed-radix
// column passes: the lean asm forms (FP32: per-column constants and the
// hardware sine/cosine of a phase reduced in FP64 turns; FP64: the exact
// form inline) for asm and conjugate asm, tf_eval for the other kinds.
template <typename R>
__device__ __forceinline__ cpx<R> mixed_tf(const TFUnit& u, const TF32Col& cc, int col, int p) {
  if (u.kind == TF_ASM || u.kind == TF_CONJ_ASM) {
    if constexpr (sizeof(R) == 4) {
      return asm_tf_fp32_full(u, cc, p);   // cc carries the conjugate's sign
    } else {
      cpx<double> h = asm_tf_fp64_inl(u, col, p);
      if (u.kind == TF_CONJ_ASM) h.y = -h.y;
      return h;
    }
  }
  return tf_eval<R>(u, col, p);
}

// ---------------------------------------------------------------------------
// Column pass over W adjacent columns (tiled intermediate or row-major).
// MODE 0: one transform in direction DIR, then the scale; MODE 2: forward
// transform -> transfer function (any TF kind, tf_eval) -> inverse
// transform, then the scale. In place. Threads are interleaved over the
// columns (c = tid % W) so a row of the tile is read as one W-element
// segment.
// ---------------------------------------------------------------------------
template <typename R, int N, int W, int MODE, int DIR>
__global__ __launch_bounds__(W * MPlan<N>::T) void k_cols_mixed(const ColArgs a) {
  constexpr int T = MPlan<N>::T;
  __shared__ cpx<R> lds[lds_len(N) * W];
  const int tid = threadIdx.x;
  const int c = tid % W, t = tid / W;
  const int tpu = a.cols / W;
  // workgroups narrower than the layout tile share its lines: the XCD-aware
  // order keeps the workgroups of one tile on one XCD (one L2)
  const int bid = a.xcd_remap ? remap_xcd(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int unit = bid / tpu;
  const int col = (bid % tpu) * W + c;
  const int tl = a.tiled ? 1 << a.tlg : 0;
  const int ld = a.tiled ? tl : a.ld;
  cpx<R>* base =
      (cpx<R>*)a.data + (size_t)unit * a.unit + (a.tiled ? (col >> a.tlg) * a.tstride + (col & (tl - 1)) : col);
  cpx<R>* lc = lds + c;
  for (int p = t; p < N; p += T) {
    const int q = p - a.in_row0;
    cpx<R> x = {R(0), R(0)};
    if (q >= 0 && q < a.rows) x = base[(size_t)q * ld];
    lc[lpad(p) * W] = x;
  }
  __syncthreads();
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  if constexpr (MODE == 2) {
    mixed_fft<R, N, -1, W>(lc, t, tw);
    const TFUnit& u = a.tf[unit];
    const TF32Col cc = tf32_col(u, col);   // FP32 asm only
    // transfer functions even in fy (no spectral shift): row N - p has the
    // exact negative of row p's frequency, so one evaluation serves both
    const bool mirror = u.fy_off == 0.0 && (u.kind == TF_ASM || u.kind == TF_CONJ_ASM || u.kind == TF_FRESNEL ||
                                            u.kind == TF_BANDLIMIT);
    if (mirror) {
      for (int p = t; p <= N / 2; p += T) {
        const cpx<R> h = mixed_tf<R>(u, cc, col, p);
        lc[lpad(p) * W] = cmul(lc[lpad(p) * W], h);
        if (p != 0 && 2 * p != N) lc[lpad(N - p) * W] = cmul(lc[lpad(N - p) * W], h);
      }
    } else {
      for (int p = t; p < N; p += T) lc[lpad(p) * W] = cmul(lc[lpad(p) * W], mixed_tf<R>(u, cc, col, p));
    }
    __syncthreads();
    mixed_fft<R, N, +1, W>(lc, t, tw);
  } else {
    mixed_fft<R, N, DIR, W>(lc, t, tw);
  }
  const R sc = (R)a.scale;
  for (int p = t; p < N; p += T) {
    const int q = p - a.out_row0;
    if (q >= 0 && q < a.out_rows) {
      cpx<R> y = lc[lpad(p) * W];
      if (sc != R(1)) y = cscale(y, sc);
      base[(size_t)q * ld] = y;
    }
  }
}

// ---------------------------------------------------------------------------
// Runtime-length variants for the 2^a 3^b 5^c sides without a compiled
// MPlan<N>, and for sides with factors of 7 (radix-7 stages after the
// radix-5 ones): the same plan, butterfly maps and twiddle tables, with the
// stage list passed as an argument and the transform in dynamic LDS.
// A thread holds at most mixed_bmax(RAD) butterflies of a radix-RAD stage;
// the host picks T so that ceil((N / RAD) / T) <= mixed_bmax(RAD) for every
// stage of the plan, so each butterfly has a thread.
// ---------------------------------------------------------------------------
// The stage radices travel packed, three bits a stage (mixed_rad_code), so
// the kernel indexes no argument array; NS and the table offset of a stage
// follow from the radices before it.
constexpr int kMixedMaxStages = 21;
constexpr int kMixedMaxThreads = 512;
struct MixedRt {
  int n, s, t;              // length, stage count, threads per transform
  unsigned long long code;  // radix code of stage i at bits 3i .. 3i+2
};
constexpr int mixed_rad_code(int rad) {
  return rad == 16 ? 0 : rad == 8 ? 1 : rad == 4 ? 2 : rad == 2 ? 3 : rad == 3 ? 4 : rad == 5 ? 5 : 6;
}
constexpr int mixed_bmax(int rad) { return rad >= 16 ? 1 : rad == 2 ? 4 : rad == 3 ? 3 : 2; }

template <typename R, int RAD, int DIR, int ES>
__device__ __forceinline__ void mixed_stage_rt(cpx<R>* lds, int t, int T, int N, int NS, const cpx<R>* __restrict__ tw) {
  constexpr int BM = mixed_bmax(RAD);
  const int NB = N / RAD;
  cpx<R> v[BM][RAD];
#pragma unroll
  for (int b = 0; b < BM; ++b) {
    const int j = t + b * T;
    if (j < NB) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[b][r] = lds[lpad(j + r * NB) * ES];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BM; ++b) {
    const int j = t + b * T;
    if (j < NB) {
      const int k = j % NS;
      if (NS > 1) {
        const cpx<R>* tk = tw + k * (RAD - 1);
#pragma unroll
        for (int r = 1; r < RAD; ++r) v[b][r] = DIR < 0 ? cmul(v[b][r], tk[r - 1]) : cmulc(v[b][r], tk[r - 1]);
      }
      dft_mixed<RAD, DIR>(v[b]);
      const int o = (j - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) lds[lpad(o + r * NS) * ES] = v[b][r];
    }
  }
  __syncthreads();
}

// the plan's stages in order; every thread of the workgroup takes the same
// branch (the radix is uniform), so the stage barriers are reached by all.
// R7: the plan has radix-7 stages (a separate kernel, so that the sides
// without them do not carry its registers)
template <typename R, int DIR, int ES, bool R7>
__device__ __forceinline__ void mixed_fft_rt(cpx<R>* lds, int t, const MixedRt& p, const cpx<R>* __restrict__ tw) {
  int NS = 1, off = 0;  // stage 0 (NS == 1) reads no twiddles
  for (int s = 0; s < p.s; ++s) {
    const cpx<R>* ts = tw + off;
    int rad;
    switch ((int)(p.code >> (3 * s)) & 7) {
      case 0: rad = 16; mixed_stage_rt<R, 16, DIR, ES>(lds, t, p.t, p.n, NS, ts); break;
      case 1: rad = 8; mixed_stage_rt<R, 8, DIR, ES>(lds, t, p.t, p.n, NS, ts); break;
      case 2: rad = 4; mixed_stage_rt<R, 4, DIR, ES>(lds, t, p.t, p.n, NS, ts); break;
      case 3: rad = 2; mixed_stage_rt<R, 2, DIR, ES>(lds, t, p.t, p.n, NS, ts); break;
      case 4: rad = 3; mixed_stage_rt<R, 3, DIR, ES>(lds, t, p.t, p.n, NS, ts); break;
      case 5: rad = 5; mixed_stage_rt<R, 5, DIR, ES>(lds, t, p.t, p.n, NS, ts); break;
      default:  // code 6, only in plans launched with R7
        rad = 7;
        if constexpr (R7) mixed_stage_rt<R, 7, DIR, ES>(lds, t, p.t, p.n, NS, ts);
        break;
    }
    if (s > 0) off += NS * (rad - 1);
    NS *= rad;
  }
}

// block reduction as block_sum_store with the workgroup size known at run time
__device__ __forceinline__ void block_sum_store_rt(double v, double* red, double* out, int slot) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < nw; ++i) s += red[i];
    out[slot] = s;
  }
}

// k_rows_mixed with the plan at run time: p.t threads, lds_len(p.n) elements
// of dynamic LDS
// FP64: capped at 168 VGPRs (3 waves per SIMD), a few bytes of spills,
// faster than 2 waves per SIMD; the column kernels measured slower under the
// same cap
template <typename R, int DIR, bool R7>
__global__ __launch_bounds__(kMixedMaxThreads, (sizeof(R) == 8 ? 3 : 1)) void k_rows_mixed_rt(const RowArgs a, const MixedRt p) {
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  cpx<R>* lds = (cpx<R>*)dyn_lds;
  __shared__ double red[kMixedMaxThreads / 64];
  const int N = p.n, T = p.t;
  const int t = threadIdx.x;
  const int unit = blockIdx.x / a.rows;
  const int row = blockIdx.x % a.rows;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit;
  double pin = 0;
  for (int q0 = t; q0 < N; q0 += T) {
    const int q = q0 - a.in_col0;
    cpx<R> x = {R(0), R(0)};
    if (q >= 0 && q < a.in_cols) x = in[tiled_offset(q, row, a.in_ld, a.in_tw, a.in_tstride)];
    pin += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    lds[lpad(q0)] = x;
  }
  if (a.pw_in) block_sum_store_rt(pin, red, a.pw_in, blockIdx.x);  // uniform
  __syncthreads();
  mixed_fft_rt<R, DIR, 1, R7>(lds, t, p, (const cpx<R>*)a.tw);
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  const bool scale = a.scale_im != 0 || a.scale_re != 1;
  double pout = 0;
  for (int q0 = t; q0 < N; q0 += T) {
    const int q = q0 - a.out_col0;
    if (q >= 0 && q < a.out_cols) {
      cpx<R> y = lds[lpad(q0)];
      if (scale) y = cmul(y, sc);
      pout += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
      out[tiled_offset(q, row, a.out_ld, a.out_tw, a.out_tstride)] = y;
    }
  }
  if (a.pw_out) block_sum_store_rt(pout, red, a.pw_out, blockIdx.x);  // uniform
}

// k_cols_mixed with the plan at run time: W * p.t threads,
// lds_len(p.n) * W elements of dynamic LDS
template <typename R, int W, int MODE, int DIR, bool R7>
__global__ __launch_bounds__(kMixedMaxThreads) void k_cols_mixed_rt(const ColArgs a, const MixedRt p) {
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  cpx<R>* lds = (cpx<R>*)dyn_lds;
  const int N = p.n, T = p.t;
  const int tid = threadIdx.x;
  const int c = tid % W, t = tid / W;
  const int tpu = a.cols / W;
  // workgroups narrower than the layout tile share its lines: the XCD-aware
  // order keeps the workgroups of one tile on one XCD (one L2)
  const int bid = a.xcd_remap ? remap_xcd(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int unit = bid / tpu;
  const int col = (bid % tpu) * W + c;
  const int tl = a.tiled ? 1 << a.tlg : 0;
  const int ld = a.tiled ? tl : a.ld;
  cpx<R>* base =
      (cpx<R>*)a.data + (size_t)unit * a.unit + (a.tiled ? (col >> a.tlg) * a.tstride + (col & (tl - 1)) : col);
  cpx<R>* lc = lds + c;
  for (int q0 = t; q0 < N; q0 += T) {
    const int q = q0 - a.in_row0;
    cpx<R> x = {R(0), R(0)};
    if (q >= 0 && q < a.rows) x = base[(size_t)q * ld];
    lc[lpad(q0) * W] = x;
  }
  __syncthreads();
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  if constexpr (MODE == 2) {
    mixed_fft_rt<R, -1, W, R7>(lc, t, p, tw);
    const TFUnit& u = a.tf[unit];
    const TF32Col cc = tf32_col(u, col);   // FP32 asm only
    const bool mirror = u.fy_off == 0.0 && (u.kind == TF_ASM || u.kind == TF_CONJ_ASM || u.kind == TF_FRESNEL ||
                                            u.kind == TF_BANDLIMIT);
    if (mirror) {
      for (int q0 = t; q0 <= N / 2; q0 += T) {
        const cpx<R> h = mixed_tf<R>(u, cc, col, q0);
        lc[lpad(q0) * W] = cmul(lc[lpad(q0) * W], h);
        if (q0 != 0 && 2 * q0 != N) lc[lpad(N - q0) * W] = cmul(lc[lpad(N - q0) * W], h);
      }
    } else {
      for (int q0 = t; q0 < N; q0 += T) lc[lpad(q0) * W] = cmul(lc[lpad(q0) * W], mixed_tf<R>(u, cc, col, q0));
    }
    __syncthreads();
    mixed_fft_rt<R, 1, W, R7>(lc, t, p, tw);
  } else {
    mixed_fft_rt<R, DIR, W, R7>(lc, t, p, tw);
  }
  const R sc = (R)a.scale;
  for (int q0 = t; q0 < N; q0 += T) {
    const int q = q0 - a.out_row0;
    if (q >= 0 && q < a.out_rows) {
      cpx<R> y = lc[lpad(q0) * W];
      if (sc != R(1)) y = cscale(y, sc);
      base[(size_t)q * ld] = y;
    }
  }
}

}  // namespace wos
