// wos_fft.hpp — device building blocks of the libwos_hip FFT engine (gfx950).
//
// Stockham autosort FFT over a power-of-two length N held by a block of
// threads. Each thread owns V = 16 complex values (4 butterflies of radix 4,
// 2 of radix 8, or 1 of radix 16), or V = 8 under the radix-8 schedule;
// stages exchange data through LDS, the first stage reads global memory
// directly and the last stage writes global memory directly, both fully
// coalesced. Twiddles come from per-stage tables
// generated on the device in FP64: for stage (RAD, Ns) entry [k][r-1] holds
// exp(-2 pi i r k / (Ns RAD)), so a butterfly reads its RAD-1 twiddles from
// one base address.
//
// Stockham stage (radix RAD, sub-transform length Ns, butterfly j):
//   k = j mod Ns; v[r] = x[j + r N/RAD] * w^(r k), w = exp(dir 2 pi i/(Ns RAD));
//   v = DFT_RAD(v);  y[(j - k) RAD + k + r Ns] = v[r].
// After the last stage (Ns RAD = N) the output is in natural order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wos {

template <typename R>
struct __align__(2 * sizeof(R)) cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ cpx<R> cmulc(cpx<R> a, cpx<R> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }

// multiply by DIR*i (DIR = -1 forward, +1 inverse)
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> mul_i(cpx<R> a) {
  if constexpr (DIR < 0) return {a.y, -a.x};
  else return {-a.y, a.x};
}

// exp(DIR * 2 pi i m / 16) constants
template <typename R>
struct W16 {
  static constexpr double C1 = 0.92387953251128675613;  // cos(pi/8)
  static constexpr double S1 = 0.38268343236508977173;  // sin(pi/8)
  static constexpr double H = 0.70710678118654752440;   // sqrt(1/2)
};

template <int DIR, typename R>
__device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft4(cpx<R>& v0, cpx<R>& v1, cpx<R>& v2, cpx<R>& v3) {
  cpx<R> t0 = v0 + v2, t1 = v0 - v2, t2 = v1 + v3, t3 = mul_i<DIR>(v1 - v3);
  v0 = t0 + t2;
  v2 = t0 - t2;
  v1 = t1 + t3;
  v3 = t1 - t3;
}

// c * x with c a compile-time FP64 constant: rounded to R, or (ACC) as an
// R hi + lo pair, so the product is not biased by the constant's rounding.
template <bool ACC, typename R>
__device__ __forceinline__ R cmulk(double c, R x) {
  const R hi = (R)c;
  if constexpr (ACC) {
    const R lo = (R)(c - (double)hi);
    return fma(hi, x, lo * x);
  } else {
    return hi * x;
  }
}

// multiply by exp(DIR * 2 pi i m / 8), m = 1, 3
template <int DIR, int M, typename R, bool ACC = false>
__device__ __forceinline__ cpx<R> w8(cpx<R> a) {
  const double c = W16<R>::H;
  if constexpr (M == 1) {
    if constexpr (DIR < 0) return {cmulk<ACC>(c, a.x + a.y), cmulk<ACC>(c, a.y - a.x)};
    else return {cmulk<ACC>(c, a.x - a.y), cmulk<ACC>(c, a.x + a.y)};
  } else {  // M == 3
    if constexpr (DIR < 0) return {cmulk<ACC>(c, a.y - a.x), -cmulk<ACC>(c, a.x + a.y)};
    else return {-cmulk<ACC>(c, a.x + a.y), cmulk<ACC>(c, a.x - a.y)};
  }
}

template <int DIR, typename R, bool ACC = false>
__device__ __forceinline__ void dft8(cpx<R>* v) {
  cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  o1 = w8<DIR, 1, R, ACC>(o1);
  o2 = mul_i<DIR>(o2);
  o3 = w8<DIR, 3, R, ACC>(o3);
  v[0] = e0 + o0;
  v[4] = e0 - o0;
  v[1] = e1 + o1;
  v[5] = e1 - o1;
  v[2] = e2 + o2;
  v[6] = e2 - o2;
  v[3] = e3 + o3;
  v[7] = e3 - o3;
}

// multiply by exp(DIR * 2 pi i m / 16) for a compile-time m in 0..15
template <int DIR, int M, typename R, bool ACC = false>
__device__ __forceinline__ cpx<R> w16(cpx<R> a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) return a;
  else if constexpr (m == 4) return mul_i<DIR>(a);
  else if constexpr (m == 8) return {-a.x, -a.y};
  else if constexpr (m == 12) return mul_i<-DIR>(a);
  else if constexpr (m == 2) return w8<DIR, 1, R, ACC>(a);
  else if constexpr (m == 6) return w8<DIR, 3, R, ACC>(a);
  else {
    // generic: cos/sin of 2 pi m / 16
    constexpr double C1 = W16<R>::C1, S1 = W16<R>::S1;
    constexpr double cs[16] = {1, C1, 0, S1, 0, -S1, 0, -C1, -1, -C1, 0, -S1, 0, S1, 0, C1};
    constexpr double sn[16] = {0, S1, 0, C1, 1, C1, 0, S1, 0, -S1, 0, -C1, -1, -C1, 0, -S1};
    const double c = cs[m];
    const double s = DIR < 0 ? -sn[m] : sn[m];
    return {cmulk<ACC>(c, a.x) - cmulk<ACC>(s, a.y), cmulk<ACC>(s, a.x) + cmulk<ACC>(c, a.y)};
  }
}

template <int DIR, typename R, bool ACC = false>
__device__ __forceinline__ void dft16(cpx<R>* v) {
  // n = 4 n1 + n2: radix-4 over n1 for each n2
  dft4<DIR>(v[0], v[4], v[8], v[12]);
  dft4<DIR>(v[1], v[5], v[9], v[13]);
  dft4<DIR>(v[2], v[6], v[10], v[14]);
  dft4<DIR>(v[3], v[7], v[11], v[15]);
  // slot 4 k1 + n2 holds X[k1][n2]; twiddle by W16^(n2 k1)
  v[5] = w16<DIR, 1, R, ACC>(v[5]);
  v[6] = w16<DIR, 2, R, ACC>(v[6]);
  v[7] = w16<DIR, 3, R, ACC>(v[7]);
  v[9] = w16<DIR, 2, R, ACC>(v[9]);
  v[10] = w16<DIR, 4, R, ACC>(v[10]);
  v[11] = w16<DIR, 6, R, ACC>(v[11]);
  v[13] = w16<DIR, 3, R, ACC>(v[13]);
  v[14] = w16<DIR, 6, R, ACC>(v[14]);
  v[15] = w16<DIR, 9, R, ACC>(v[15]);
  // radix-4 over n2 for each k1: result y[k1 + 4 k2] lands in slot 4 k1 + k2
  dft4<DIR>(v[0], v[1], v[2], v[3]);
  dft4<DIR>(v[4], v[5], v[6], v[7]);
  dft4<DIR>(v[8], v[9], v[10], v[11]);
  dft4<DIR>(v[12], v[13], v[14], v[15]);
  // transpose 4x4 so slot k holds y[k]
  cpx<R> t;
  t = v[1]; v[1] = v[4]; v[4] = t;
  t = v[2]; v[2] = v[8]; v[8] = t;
  t = v[3]; v[3] = v[12]; v[12] = t;
  t = v[6]; v[6] = v[9]; v[9] = t;
  t = v[7]; v[7] = v[13]; v[13] = t;
  t = v[11]; v[11] = v[14]; v[14] = t;
}

template <int RAD, int DIR, typename R, bool ACC = false>
__device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (RAD == 2) dft2<DIR>(v);
  else if constexpr (RAD == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
  else if constexpr (RAD == 8) dft8<DIR, R, ACC>(v);
  else if constexpr (RAD == 16) dft16<DIR, R, ACC>(v);
}

// ---------------------------------------------------------------------------
// Radix plan for a power-of-two length N: values per thread V, stage radices.
// VV selects the schedule: radix-16 stages with 16 values per thread, or
// radix-8 stages with 8 values per thread (twice the threads per transform,
// half the registers, one more LDS exchange per 2^12 points).
// Forward: [VV, VV, ..., rem]; inverse: [rem, VV, ..., VV], so the last
// forward stage and the first inverse stage share the same butterfly map and
// a fused forward->multiply->inverse keeps the data in registers.
// ---------------------------------------------------------------------------
constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// VV = 32: the radix-16 stages with 32 values per thread (two butterflies
// of each radix-16 stage per thread, half the threads per transform), for
// column passes that cover more columns per workgroup.
template <int N, int VV = 16>
struct Plan {
  static_assert(VV == 16 || VV == 8 || VV == 32, "radix-16 or radix-8 schedule");
  static constexpr int RV = VV == 32 ? 16 : VV;         // stage radix
  static constexpr int LOG = ilog2(N);
  static constexpr int LV = ilog2(RV);
  static constexpr int V = N >= VV ? VV : N;            // values per thread
  static constexpr int T = N / V;                        // threads per transform
  static constexpr int NF = LOG / LV;                    // full-radix (RV) stages
  static constexpr int REM = 1 << (LOG % LV);            // leftover radix (1: none)
  static constexpr int S = NF + (REM > 1 ? 1 : 0);       // number of stages
  // radix of stage s, forward order
  static constexpr int fwd_rad(int s) {
    return N < RV ? N : (s < NF ? RV : REM);
  }
  // radix of stage s, inverse order
  static constexpr int inv_rad(int s) {
    return N < RV ? N : (REM > 1 ? (s == 0 ? REM : RV) : RV);
  }
  // sub-transform length before stage s
  static constexpr int ns_fwd(int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= fwd_rad(i);
    return ns;
  }
  static constexpr int ns_inv(int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= inv_rad(i);
    return ns;
  }
  // twiddle-table offsets: forward stages 1..S-1, then inverse stages 1..S-1
  static constexpr int tw_off_fwd(int s) {
    int o = 0;
    for (int i = 1; i < s; ++i) o += ns_fwd(i) * (fwd_rad(i) - 1);
    return o;
  }
  static constexpr int tw_off_inv(int s) {
    int o = tw_off_fwd(S);
    for (int i = 1; i < s; ++i) o += ns_inv(i) * (inv_rad(i) - 1);
    return o;
  }
  static constexpr int tw_total() { return tw_off_inv(S); }
};

// LDS padding: one slot every 16 to break power-of-two bank strides.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
constexpr int lds_len(int n) { return n + n / 16 + 1; }

// LDS stride between the transforms of a multi-row workgroup (row = tid %
// rows). LDS serves 256 bytes (S elements) per cycle; the lanes of one such
// group are S/rows transform positions of each of the rows. Offsetting row
// r by r*S/rows elements (mod S) puts the rows of a group in disjoint bank
// ranges for both the unit-stride and the stride-17 (padded radix-16) access
// of a stage; with rows >= S the lanes of a group are rows at one position
// and an odd stride keeps them apart.
template <typename E>
constexpr int lds_stride_elems(int n, int rows) {
  constexpr int S = 256 / (int)sizeof(E);
  int l = lds_len(n);
  if (rows <= 1) return l;
  const int want = rows >= S ? 1 : S / rows;
  const int mod = rows >= S ? 2 : S;
  while (l % mod != want % mod) ++l;
  return l;
}
template <typename R>
constexpr int row_lds_stride(int n, int rows) { return lds_stride_elems<cpx<R>>(n, rows); }

// Apply twiddles and the radix-RAD DFT to the butterflies a thread owns.
// v[b*RAD + r] holds x[j_b + r N/RAD], j_b = t + b*T. The stage's twiddle
// table starts at tws + OFF; entry [k][r-1] = exp(-2 pi i r k/(NS RAD)),
// conjugated for the inverse direction.
// ACC: the table holds [hi, lo] pairs (2 entries per twiddle) and the
// product is v*hi + v*lo.
template <typename R, int N, int DIR, int RAD, int NS, int OFF, bool ACC = false, int VV = 16>
__device__ __forceinline__ void stage_compute(cpx<R>* v, int t, const cpx<R>* __restrict__ tws) {
  constexpr int V = Plan<N, VV>::V, T = Plan<N, VV>::T;
  constexpr int NB = V / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if constexpr (NS > 1) {
      const int j = t + b * T;
      const int k = j & (NS - 1);
      if constexpr (ACC) {
        const cpx<R>* tk = tws + 2 * (OFF + k * (RAD - 1));
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          const cpx<R> hi = tk[2 * (r - 1)], lo = tk[2 * (r - 1) + 1];
          const cpx<R> x = v[b * RAD + r];
          v[b * RAD + r] = DIR < 0 ? cmul(x, hi) + cmul(x, lo) : cmulc(x, hi) + cmulc(x, lo);
        }
      } else {
        const cpx<R>* tk = tws + OFF + k * (RAD - 1);
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          const cpx<R> w = tk[r - 1];
          v[b * RAD + r] = DIR < 0 ? cmul(v[b * RAD + r], w) : cmulc(v[b * RAD + r], w);
        }
      }
    }
    dft<RAD, DIR, R, ACC>(v + b * RAD);
  }
}

// LDS slot of transform position base + c where c = r*X (X = NS for stores,
// N/RAD for loads) is lpad(base) + c + (c >> 4): base's low four bits plus
// c mod 16 never carry past 16 in a Stockham stage, so the padding splits
// into a per-butterfly part and a compile-time part.

// Write a stage's outputs to LDS (Stockham positions), element stride ES
// between consecutive positions of one transform (ES = W for column tiles).
template <typename R, int N, int RAD, int NS, int ES, int VV = 16>
__device__ __forceinline__ void stage_store_lds(const cpx<R>* v, int t, cpx<R>* lds) {
  constexpr int V = Plan<N, VV>::V, T = Plan<N, VV>::T;
  constexpr int NB = V / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const int k = j & (NS - 1);
    cpx<R>* p = lds + lpad((j - k) * RAD + k) * ES;
#pragma unroll
    for (int r = 0; r < RAD; ++r) p[(r * NS + ((r * NS) >> 4)) * ES] = v[b * RAD + r];
  }
}

// One component (COMP 0: real, 1: imaginary) of a stage's outputs / the
// next stage's inputs, through an LDS array of R: the exchange of a split
// transform takes two rounds but half the LDS.
template <typename R, int N, int RAD, int NS, int ES, int COMP, int VV = 16>
__device__ __forceinline__ void stage_store_lds_c(const cpx<R>* v, int t, R* lds) {
  constexpr int V = Plan<N, VV>::V, T = Plan<N, VV>::T;
  constexpr int NB = V / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const int k = j & (NS - 1);
    R* p = lds + lpad((j - k) * RAD + k) * ES;
#pragma unroll
    for (int r = 0; r < RAD; ++r) p[(r * NS + ((r * NS) >> 4)) * ES] = COMP == 0 ? v[b * RAD + r].x : v[b * RAD + r].y;
  }
}

template <typename R, int N, int RAD, int ES, int COMP, int VV = 16>
__device__ __forceinline__ void stage_load_lds_c(cpx<R>* v, int t, const R* lds) {
  constexpr int V = Plan<N, VV>::V, T = Plan<N, VV>::T;
  constexpr int NB = V / RAD;
  constexpr int X = N / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const R* p = lds + lpad(j) * ES;
#pragma unroll
    for (int r = 0; r < RAD; ++r) {
      const R x = p[(r * X + ((r * X) >> 4)) * ES];
      if constexpr (COMP == 0) v[b * RAD + r].x = x;
      else v[b * RAD + r].y = x;
    }
  }
}

// Read the next stage's inputs from LDS: v[b*RAD + r] = x[j_b + r N/RAD].
template <typename R, int N, int RAD, int ES, int VV = 16>
__device__ __forceinline__ void stage_load_lds(cpx<R>* v, int t, const cpx<R>* lds) {
  constexpr int V = Plan<N, VV>::V, T = Plan<N, VV>::T;
  constexpr int NB = V / RAD;
  constexpr int X = N / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const cpx<R>* p = lds + lpad(j) * ES;
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[b * RAD + r] = p[(r * X + ((r * X) >> 4)) * ES];
  }
}

// Position (within the transform) held by slot (b, r) at the start of the
// first stage of radix RAD0, and at the end of the last stage of radix RADL.
template <int N, int RAD, int VV = 16>
__device__ __forceinline__ int slot_pos(int t, int b, int r) {
  return t + b * Plan<N, VV>::T + r * (N / RAD);
}

// Run stages [S0, S1) of the forward (DIR<0) or inverse (DIR>0) plan with
// LDS exchanges between them. On entry v holds the inputs of stage S0 in its
// butterfly order; on exit v holds the outputs of stage S1-1 in the order
// "slot (b, r) = position t + b T + r N/RAD_last" (valid for the last stage).
// TWB: table entry that tw points at (a table holding one direction's part
// of the twiddles starts at that direction's first offset).
template <typename R, int N, int DIR, int S, int SEND, int ES, bool ACC = false, bool SPLIT = false, int VV = 16,
          int TWB = 0>
__device__ __forceinline__ void run_stages(cpx<R>* v, int t, const cpx<R>* __restrict__ tw, cpx<R>* lds) {
  if constexpr (S < SEND) {
    constexpr int RAD = DIR < 0 ? Plan<N, VV>::fwd_rad(S) : Plan<N, VV>::inv_rad(S);
    constexpr int NS = DIR < 0 ? Plan<N, VV>::ns_fwd(S) : Plan<N, VV>::ns_inv(S);
    constexpr int OFF = (DIR < 0 ? Plan<N, VV>::tw_off_fwd(S) : Plan<N, VV>::tw_off_inv(S)) - TWB;
    stage_compute<R, N, DIR, RAD, NS, OFF, ACC, VV>(v, t, tw);
    if constexpr (S + 1 < SEND) {
      constexpr int RADN = DIR < 0 ? Plan<N, VV>::fwd_rad(S + 1) : Plan<N, VV>::inv_rad(S + 1);
      if constexpr (SPLIT) {  // lds holds lds_len(N) values of R per transform
        R* lr = (R*)lds;
        __syncthreads();  // previous readers of lds are done
        stage_store_lds_c<R, N, RAD, NS, ES, 0, VV>(v, t, lr);
        __syncthreads();
        stage_load_lds_c<R, N, RADN, ES, 0, VV>(v, t, lr);
        __syncthreads();
        stage_store_lds_c<R, N, RAD, NS, ES, 1, VV>(v, t, lr);
        __syncthreads();
        stage_load_lds_c<R, N, RADN, ES, 1, VV>(v, t, lr);
      } else {
        __syncthreads();  // previous readers of lds are done
        stage_store_lds<R, N, RAD, NS, ES, VV>(v, t, lds);
        __syncthreads();
        stage_load_lds<R, N, RADN, ES, VV>(v, t, lds);
      }
      run_stages<R, N, DIR, S + 1, SEND, ES, ACC, SPLIT, VV, TWB>(v, t, tw, lds);
    }
  }
}

// First-stage radix (input map) and last-stage radix (output map).
template <int N, int DIR, int VV = 16>
constexpr int first_rad() { return DIR < 0 ? Plan<N, VV>::fwd_rad(0) : Plan<N, VV>::inv_rad(0); }
template <int N, int DIR, int VV = 16>
constexpr int last_rad() {
  return DIR < 0 ? Plan<N, VV>::fwd_rad(Plan<N, VV>::S - 1) : Plan<N, VV>::inv_rad(Plan<N, VV>::S - 1);
}

// Full N-point transform of the values in v (input map of first_rad, output
// map of last_rad).
// ACC: compensated twiddles and constants (tw then points to a [hi, lo] table).
// SPLIT: exchanges real and imaginary parts in two rounds through half the
// LDS (lds then holds lds_len(N) values of R per transform, not complex).
template <typename R, int N, int DIR, int ES, bool ACC = false, bool SPLIT = false, int VV = 16, int TWB = 0>
__device__ __forceinline__ void fft_block(cpx<R>* v, int t, const cpx<R>* __restrict__ tw, cpx<R>* lds) {
  run_stages<R, N, DIR, 0, Plan<N, VV>::S, ES, ACC, SPLIT, VV, TWB>(v, t, tw, lds);
}

// First table entry and entry count of one direction's stage twiddles.
template <int N, int DIR, int VV = 16>
constexpr int tw_dir_base() { return DIR < 0 ? 0 : Plan<N, VV>::tw_off_fwd(Plan<N, VV>::S); }
template <int N, int DIR, int VV = 16>
constexpr int tw_dir_len() {
  return DIR < 0 ? Plan<N, VV>::tw_off_fwd(Plan<N, VV>::S) : Plan<N, VV>::tw_total() - Plan<N, VV>::tw_off_fwd(Plan<N, VV>::S);
}

// ---------------------------------------------------------------------------
// Twiddles held in registers as seeds. A kernel that transforms several rows
// per thread keeps no twiddle loads inside the transform, so loads it issues
// for the next rows stay in flight across the whole transform (vmcnt is an
// in-order counter: waiting on a twiddle load issued after them would wait
// for them too). Per butterfly block a thread keeps w^1 and, for radix 8 and
// 16, w^4 in FP64 (w = exp(-2 pi i k/(NS RAD))); the other powers are FP64
// products of these (at most four deep, error ~1e-15) rounded once to R.
// ---------------------------------------------------------------------------
template <typename R, int N, int DIR, int VV = 16>
struct TwReg {
  static constexpr int V = Plan<N, VV>::V;
  static constexpr int rad(int s) { return DIR < 0 ? Plan<N, VV>::fwd_rad(s) : Plan<N, VV>::inv_rad(s); }
  static constexpr int seeds(int s) { return rad(s) >= 8 ? 2 : 1; }
  // register index of stage s's first seed (stages 1 .. S-1)
  static constexpr int base(int s) {
    int c = 0;
    for (int i = 1; i < s; ++i) c += (V / rad(i)) * seeds(i);
    return c;
  }
  static constexpr int total() { return base(Plan<N, VV>::S); }
  cpx<double> w[total() > 0 ? total() : 1];
};

template <typename R, int N, int DIR, int S, int VV = 16>
__device__ __forceinline__ void tw_load_stages(TwReg<R, N, DIR, VV>& tr, int t, const cpx<double>* __restrict__ tws) {
  if constexpr (S < Plan<N, VV>::S) {
    constexpr int RAD = TwReg<R, N, DIR, VV>::rad(S);
    constexpr int NS = DIR < 0 ? Plan<N, VV>::ns_fwd(S) : Plan<N, VV>::ns_inv(S);
    constexpr int OFF = DIR < 0 ? Plan<N, VV>::tw_off_fwd(S) : Plan<N, VV>::tw_off_inv(S);
    constexpr int B = TwReg<R, N, DIR, VV>::base(S);
    constexpr int SE = TwReg<R, N, DIR, VV>::seeds(S);
    constexpr int NB = Plan<N, VV>::V / RAD;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int k = (t + b * Plan<N, VV>::T) & (NS - 1);
      tr.w[B + b * SE] = tws[OFF + k * (RAD - 1)];               // w^1
      if constexpr (SE == 2) tr.w[B + b * SE + 1] = tws[OFF + k * (RAD - 1) + 3];  // w^4
    }
    tw_load_stages<R, N, DIR, S + 1, VV>(tr, t, tws);
  }
}

// tws: the FP64 stage twiddle table of N points (any kernel precision)
template <typename R, int N, int DIR, int VV = 16>
__device__ __forceinline__ void tw_load(TwReg<R, N, DIR, VV>& tr, int t, const cpx<double>* __restrict__ tws) {
  tw_load_stages<R, N, DIR, 1, VV>(tr, t, tws);
}

// powers w^1 .. w^(RAD-1) of a block from its seeds, rounded to R
template <typename R, int RAD>
__device__ __forceinline__ void tw_expand(const cpx<double>* sd, cpx<R>* w) {
  const cpx<double> w1 = sd[0];
  w[1] = {(R)w1.x, (R)w1.y};
  if constexpr (RAD >= 4) {
    const cpx<double> w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    w[2] = {(R)w2.x, (R)w2.y};
    w[3] = {(R)w3.x, (R)w3.y};
    if constexpr (RAD >= 8) {
      const cpx<double> w4 = sd[1];
      const cpx<double> w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3);
      w[4] = {(R)w4.x, (R)w4.y};
      w[5] = {(R)w5.x, (R)w5.y};
      w[6] = {(R)w6.x, (R)w6.y};
      w[7] = {(R)w7.x, (R)w7.y};
      if constexpr (RAD == 16) {
        const cpx<double> w8 = cmul(w4, w4), w12 = cmul(w8, w4);
        const cpx<double> p[8] = {w8, cmul(w8, w1), cmul(w8, w2), cmul(w8, w3),
                                  w12, cmul(w12, w1), cmul(w12, w2), cmul(w12, w3)};
#pragma unroll
        for (int i = 0; i < 8; ++i) w[8 + i] = {(R)p[i].x, (R)p[i].y};
      }
    }
  }
}

template <typename R, int N, int DIR, int S, int ES, int VV = 16>
__device__ __forceinline__ void run_stages_reg(cpx<R>* v, int t, const TwReg<R, N, DIR, VV>& tr, cpx<R>* lds) {
  if constexpr (S < Plan<N, VV>::S) {
    constexpr int RAD = TwReg<R, N, DIR, VV>::rad(S);
    constexpr int NS = DIR < 0 ? Plan<N, VV>::ns_fwd(S) : Plan<N, VV>::ns_inv(S);
    constexpr int NB = Plan<N, VV>::V / RAD;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if constexpr (NS > 1) {
        constexpr int B = TwReg<R, N, DIR, VV>::base(S);
        constexpr int SE = TwReg<R, N, DIR, VV>::seeds(S);
        cpx<R> w[RAD];
        tw_expand<R, RAD>(tr.w + B + b * SE, w);
#pragma unroll
        for (int r = 1; r < RAD; ++r)
          v[b * RAD + r] = DIR < 0 ? cmul(v[b * RAD + r], w[r]) : cmulc(v[b * RAD + r], w[r]);
      }
      dft<RAD, DIR>(v + b * RAD);
    }
    if constexpr (S + 1 < Plan<N, VV>::S) {
      constexpr int RADN = TwReg<R, N, DIR, VV>::rad(S + 1);
      __syncthreads();
      stage_store_lds<R, N, RAD, NS, ES, VV>(v, t, lds);
      __syncthreads();
      stage_load_lds<R, N, RADN, ES, VV>(v, t, lds);
      run_stages_reg<R, N, DIR, S + 1, ES, VV>(v, t, tr, lds);
    }
  }
}

// fft_block with register twiddle seeds (same maps as fft_block).
template <typename R, int N, int DIR, int ES, int VV = 16>
__device__ __forceinline__ void fft_block_reg(cpx<R>* v, int t, const TwReg<R, N, DIR, VV>& tr, cpx<R>* lds) {
  run_stages_reg<R, N, DIR, 0, ES, VV>(v, t, tr, lds);
}

}  // namespace wos
