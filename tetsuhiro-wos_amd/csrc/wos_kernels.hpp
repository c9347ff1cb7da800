// wos_kernels.hpp — gfx950 kernels of libwos_hip.
//
// Data layout in HBM: each propagation unit (one Jones component at one
// wavelength / distance) is a row-major complex array (complex128 for FP64,
// complex64 for FP32), units contiguous. Power-of-two 2-D transforms run as
//   k_rows   : one LDS-staged Stockham row FFT per row (contiguous rows),
//   k_cols_tf: forward column FFT -> transfer function -> inverse column FFT
//              on a tile of W columns, fully in registers/LDS (one HBM pass),
//   k_rows   : inverse row FFT with the 1/M^2 scale.
// so one ASM propagation reads and writes the field three times.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wos_fft.hpp"

namespace wos {

// ---------------------------------------------------------------------------
// Transfer functions. The FP64 phase argument is formed exactly as the
// reference does (no FMA contraction) because kz*z reaches ~5e5 rad.
// ---------------------------------------------------------------------------
enum TFKind : int { TF_ASM = 0, TF_FRESNEL = 1, TF_TABLE = 2, TF_NONE = 3, TF_CONJ_ASM = 4 };

struct TFUnit {
  double wl, z, k, az;     // wavelength, distance, k = (2 pi)/wl, |z|
  double alpha, limit;     // Tikhonov alpha, evanescent truncation (nepers)
  double fx_off, fy_off;   // carrier offset (asm_shift*), else 0
  double den;              // (double)M * dx: freq(i) = i/den (i <= M/2) else (i-M)/den
  double fr_c;             // fresnel_tf: ((-pi*wl)*z)
  double pre_re, pre_im;   // fresnel_tf: cexpI(k*z)
  int M;                   // transform side
  int kind;                // TFKind
  int zero_ev, reg, backward, pad_;
  const void* table;       // TF_TABLE: M*M complex spectrum (same precision as data)
};

// Go math.Hypot: max * sqrt(1 + (min/max)^2)
__device__ __forceinline__ double go_hypot(double p, double q) {
#pragma clang fp contract(off)
  p = fabs(p);
  q = fabs(q);
  if (p < q) {
    double t = p;
    p = q;
    q = t;
  }
  if (p == 0) return 0;
  q = q / p;
  return p * sqrt(1 + q * q);
}

__device__ __forceinline__ double freq_of(int i, int M, double den) {
#pragma clang fp contract(off)
  return i <= M / 2 ? (double)i / den : (double)(i - M) / den;
}

// asmTF (optics/propagate.go:159-181) in the reference's operation order.
__device__ __forceinline__ cpx<double> asm_tf_fp64(const TFUnit& u, double fx, double fy) {
#pragma clang fp contract(off)
  double rho = go_hypot(u.wl * fx, u.wl * fy);
  if (rho <= 1) {
    double ph = u.k * u.z * sqrt((1 - rho) * (1 + rho));
    double s, c;
    sincos(ph, &s, &c);
    return {c, s};
  }
  double decay = u.k * u.az * sqrt((rho - 1) * (rho + 1));
  if (u.zero_ev || (u.limit > 0 && decay > u.limit)) return {0.0, 0.0};
  if (u.backward) {
    if (u.reg) {
      double a = exp(decay);
      return {a / (1 + u.alpha * u.alpha * a * a), 0.0};
    }
    return {exp(decay), 0.0};
  }
  return {exp(-decay), 0.0};
}

// FP32 bulk path: the phase is formed in FP64 and wrapped into [-pi, pi]
// in FP64 before rounding to FP32 for sincosf.
__device__ __forceinline__ float wrap_phase_f32(double ph) {
  const double inv2pi = 0.15915494309189533577;
  const double two_pi_hi = 6.28318530717958623200;   // fl(2 pi)
  const double two_pi_lo = 2.44929359829470635445e-16; // 2 pi - fl(2 pi)
  double nq = rint(ph * inv2pi);
  double r = fma(-nq, two_pi_hi, ph);
  r = fma(-nq, two_pi_lo, r);
  return (float)r;
}

__device__ __forceinline__ cpx<float> asm_tf_fp32(const TFUnit& u, double fx, double fy) {
  double ax = u.wl * fx, ay = u.wl * fy;
  double rho2 = ax * ax + ay * ay;
  if (rho2 <= 1) {
    double ph = u.k * u.z * sqrt(1 - rho2);
    float s, c;
    sincosf(wrap_phase_f32(ph), &s, &c);
    return {c, s};
  }
  double decay = u.k * u.az * sqrt(rho2 - 1);
  if (u.zero_ev || (u.limit > 0 && decay > u.limit)) return {0.f, 0.f};
  if (u.backward) {
    if (u.reg) {
      double a = exp(decay);
      return {(float)(a / (1 + u.alpha * u.alpha * a * a)), 0.f};
    }
    return {(float)exp(decay), 0.f};
  }
  return {(float)exp(-decay), 0.f};
}

// Fresnel TF (optics/propagate.go:511-513): cexpI(k z) * cexpI(-pi wl z f^2)
__device__ __forceinline__ cpx<double> fresnel_tf_fp64(const TFUnit& u, double fx, double fy) {
#pragma clang fp contract(off)
  double ph = u.fr_c * (fx * fx + fy * fy);
  double s, c;
  sincos(ph, &s, &c);
  return {u.pre_re * c - u.pre_im * s, u.pre_re * s + u.pre_im * c};
}
__device__ __forceinline__ cpx<float> fresnel_tf_fp32(const TFUnit& u, double fx, double fy) {
  double ph = u.fr_c * (fx * fx + fy * fy);
  float s, c;
  sincosf(wrap_phase_f32(ph), &s, &c);
  float pr = (float)u.pre_re, pi = (float)u.pre_im;
  return {pr * c - pi * s, pr * s + pi * c};
}

template <typename R>
__device__ __forceinline__ cpx<R> tf_eval(const TFUnit& u, int ix, int iy);

template <>
__device__ __forceinline__ cpx<double> tf_eval<double>(const TFUnit& u, int ix, int iy) {
#pragma clang fp contract(off)
  double fx = freq_of(ix, u.M, u.den) + u.fx_off;
  double fy = freq_of(iy, u.M, u.den) + u.fy_off;
  switch (u.kind) {
    case TF_ASM: return asm_tf_fp64(u, fx, fy);
    case TF_CONJ_ASM: {
      cpx<double> h = asm_tf_fp64(u, fx, fy);
      return {h.x, -h.y};
    }
    case TF_FRESNEL: return fresnel_tf_fp64(u, fx, fy);
    case TF_TABLE: return ((const cpx<double>*)u.table)[(size_t)iy * u.M + ix];
    default: return {1.0, 0.0};
  }
}

template <>
__device__ __forceinline__ cpx<float> tf_eval<float>(const TFUnit& u, int ix, int iy) {
  double fx = freq_of(ix, u.M, u.den) + u.fx_off;
  double fy = freq_of(iy, u.M, u.den) + u.fy_off;
  switch (u.kind) {
    case TF_ASM: return asm_tf_fp32(u, fx, fy);
    case TF_CONJ_ASM: {
      cpx<float> h = asm_tf_fp32(u, fx, fy);
      return {h.x, -h.y};
    }
    case TF_FRESNEL: return fresnel_tf_fp32(u, fx, fy);
    case TF_TABLE: return ((const cpx<float>*)u.table)[(size_t)iy * u.M + ix];
    default: return {1.f, 0.f};
  }
}

// ---------------------------------------------------------------------------
// Block reduction of one double per thread -> out[slot] (deterministic).
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ void block_sum_store(double v, double* red, double* out, int slot) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = (NT + 63) / 64;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) s += red[i];
    out[slot] = s;
  }
}

// ---------------------------------------------------------------------------
// Twiddle table tw[m] = exp(-2 pi i m / n), m in [0, n), built in FP64.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_twiddle(cpx<R>* tw, int n) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(-2.0 * (double)m / (double)n, &s, &c);
  tw[m] = {(R)c, (R)s};
}

// ---------------------------------------------------------------------------
// Row pass.
// ---------------------------------------------------------------------------
struct RowArgs {
  const void* in;
  void* out;
  long long in_unit, out_unit;  // elements between units
  int in_ld, out_ld;            // elements between rows
  int rows;                     // rows per unit processed
  int units;
  int in_col0, in_cols;         // position p reads in[p - in_col0] when in range, else 0
  int out_col0, out_cols;       // position p is stored to out[p - out_col0] when in range
  double scale_re, scale_im;    // complex factor applied to the outputs
  double* pw_in;                // per-block sum |in|^2 (nullable)
  double* pw_out;               // per-block sum |out|^2 (nullable)
  const void* tw;               // N-point twiddles
};

template <int N>
constexpr int rows_per_block() {
  return Plan<N>::T >= 256 ? 1 : 256 / Plan<N>::T;
}

template <typename R, int N, int DIR, int ROWS>
__global__ __launch_bounds__(ROWS* Plan<N>::T) void k_rows(const RowArgs a) {
  constexpr int T = Plan<N>::T, V = Plan<N>::V;
  constexpr int R0 = first_rad<N, DIR>(), RL = last_rad<N, DIR>();
  constexpr int NT = ROWS * T;
  __shared__ cpx<R> lds[ROWS * lds_len(N)];
  __shared__ double red[(NT + 63) / 64];
  const int tid = threadIdx.x;
  const int lr = tid / T, t = tid % T;
  const int bpu = (a.rows + ROWS - 1) / ROWS;
  const int unit = blockIdx.x / bpu;
  const int row = (blockIdx.x % bpu) * ROWS + lr;
  const bool active = row < a.rows && unit < a.units;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit + (size_t)row * a.in_ld;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit + (size_t)row * a.out_ld;
  const cpx<R>* tw = (const cpx<R>*)a.tw;

  cpx<R> v[V];
  double pin = 0;
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int q = slot_pos<N, R0>(t, b, r) - a.in_col0;
      cpx<R> x = {R(0), R(0)};
      if (active && q >= 0 && q < a.in_cols) x = in[q];
      v[b * R0 + r] = x;
      pin += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
  fft_block<R, N, DIR, 1>(v, t, tw, lds + lr * lds_len(N));
  double pout = 0;
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  const bool real_scale = a.scale_im == 0;
#pragma unroll
  for (int b = 0; b < V / RL; ++b)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int q = slot_pos<N, RL>(t, b, r) - a.out_col0;
      if (active && q >= 0 && q < a.out_cols) {
        cpx<R> y = real_scale ? cscale(v[b * RL + r], sc.x) : cmul(v[b * RL + r], sc);
        out[q] = y;
        pout += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
      }
    }
  if (a.pw_in) block_sum_store<NT>(pin, red, a.pw_in, blockIdx.x);
  if (a.pw_out) block_sum_store<NT>(pout, red, a.pw_out, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Column pass over a tile of W adjacent columns. MODE 0: plain transform
// (direction DIR, output scale); MODE 1: forward -> TF -> inverse (ASM step).
// ---------------------------------------------------------------------------
struct ColArgs {
  void* data;          // in place
  long long unit;      // elements between units
  int ld;              // elements between rows
  int cols;            // columns per unit (multiple of W)
  int units;
  int in_row0, rows;   // transform position p maps to data row p - in_row0 when in range (else 0)
  int out_row0, out_rows;  // position p is stored to row p - out_row0 when in range
  double scale;
  int xcd_remap;       // spread consecutive tiles over one XCD
  const TFUnit* tf;    // per unit (MODE 1)
  const void* tw;
};

__device__ __forceinline__ int remap_xcd(int bid, int nb) {
  // The dispatcher assigns workgroup i to XCD i % 8; give each XCD a
  // contiguous range of tiles so neighbouring column tiles share its L2.
  if ((nb & 7) != 0) return bid;
  return (bid & 7) * (nb >> 3) + (bid >> 3);
}

template <typename R, int N, int W, int MODE, int DIR>
__global__ __launch_bounds__(W* Plan<N>::T) void k_cols(const ColArgs a) {
  constexpr int T = Plan<N>::T, V = Plan<N>::V;
  __shared__ cpx<R> lds[lds_len(N) * W];
  const int tid = threadIdx.x;
  const int c = tid % W, t = tid / W;
  const int bid = a.xcd_remap ? remap_xcd(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int tpu = a.cols / W;
  const int unit = bid / tpu;
  const int col = (bid % tpu) * W + c;
  cpx<R>* base = (cpx<R>*)a.data + (size_t)unit * a.unit + col;
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  const bool active = unit < a.units;

  constexpr int DIR0 = MODE == 1 ? -1 : DIR;
  constexpr int R0 = first_rad<N, DIR0>();
  cpx<R> v[V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int q = slot_pos<N, R0>(t, b, r) - a.in_row0;
      cpx<R> x = {R(0), R(0)};
      if (active && q >= 0 && q < a.rows) x = base[(size_t)q * a.ld];
      v[b * R0 + r] = x;
    }
  fft_block<R, N, DIR0, W>(v, t, tw, lds + c);
  if constexpr (MODE == 1) {
    constexpr int RL = last_rad<N, -1>();
    if (active) {
      const TFUnit u = a.tf[unit];
#pragma unroll
      for (int b = 0; b < V / RL; ++b)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int p = slot_pos<N, RL>(t, b, r);
          v[b * RL + r] = cmul(v[b * RL + r], tf_eval<R>(u, col, p));
        }
    }
    fft_block<R, N, +1, W>(v, t, tw, lds + c);
  }
  constexpr int DIRL = MODE == 1 ? +1 : DIR;
  constexpr int RL = last_rad<N, DIRL>();
  const R sc = (R)a.scale;
#pragma unroll
  for (int b = 0; b < V / RL; ++b)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int q = slot_pos<N, RL>(t, b, r) - a.out_row0;
      if (active && q >= 0 && q < a.out_rows) base[(size_t)q * a.ld] = cscale(v[b * RL + r], sc);
    }
}

}  // namespace wos
