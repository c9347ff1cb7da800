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
#include <string.h>

#include "wos_fft.hpp"

namespace wos {

// ---------------------------------------------------------------------------
// Transfer functions. The FP64 phase argument is formed exactly as the
// reference does (no FMA contraction) because kz*z reaches ~5e5 rad.
// ---------------------------------------------------------------------------
// TF_ASM_BL: asm followed by ApplyBandlimit on the same grid (one transfer;
// the column pass also sums |H U|^2 for the propagation's power check).
enum TFKind : int {
  TF_ASM = 0, TF_FRESNEL = 1, TF_TABLE = 2, TF_NONE = 3, TF_CONJ_ASM = 4, TF_BANDLIMIT = 5, TF_ASM_BL = 6
};

struct TFUnit {
  double wl, z, k, az;     // wavelength, distance, k = (2 pi)/wl, |z|
  double alpha, limit;     // Tikhonov alpha, evanescent truncation (nepers)
  double fx_off, fy_off;   // carrier offset (asm_shift*), else 0
  double den;              // (double)M * dx: freq(i) = i/den (i <= M/2) else (i-M)/den
  double inv_den;          // 1/den (FP32 path only; the FP64 path divides exactly)
  double kz;               // k*z (FP32 path)
  double fr_c;             // fresnel_tf: ((-pi*wl)*z)
  double pre_re, pre_im;   // fresnel_tf: cexpI(k*z)
  int M;                   // transform side
  int kind;                // TFKind
  int zero_ev, reg, backward, pad_;
  const void* table;       // TF_TABLE: M*M complex spectrum (same precision as data)
  double bl_f0, bl_sig;    // TF_BANDLIMIT, TF_ASM_BL: roll-off start and width (1/m)
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

// sin and cos of an FP64 transfer-function phase. Cody-Waite reduction by
// pi/2: n = rint(x 2/pi), r = x - n pi/2 with pi/2 split into two doubles
// (P1 = fl(pi/2), P2 = fl(pi/2 - P1)), each step one FMA, so r carries an
// absolute error of about one ulp of r for |x| < 2^31 pi/2 (the quadrant
// fits an int); then the minimax polynomials of fdlibm's __kernel_sin and
// __kernel_cos (Sun Microsystems, k_sin.c / k_cos.c) on [-pi/4, pi/4], the
// cosine with fdlibm's compensated 1 - z/2. Measured against libm and
// 200-bit mpmath up to |x| = 1e9: absolute error <= 1.1e-16, <= 1.5 ulp. Larger
// or non-finite arguments take the library sincos.
__device__ __forceinline__ void sincos_tf64(double x, double* s, double* c) {
  if (!(fabs(x) < 2.0e9)) {
    sincos(x, s, c);
    return;
  }
  const double n = rint(x * 0.63661977236758134308);
  double r = fma(-n, 1.5707963267948966192, x);
  r = fma(-n, 6.123233995736766036e-17, r);
  const double z = r * r;
  const double ps = fma(z, fma(z, fma(z, fma(z, fma(z, 1.58969099521155010221e-10, -2.50507602534068634195e-08),
                                              2.75573137070700676789e-06), -1.98412698298579493134e-04),
                                8.33333333332248946124e-03), -1.66666666666666324348e-01);
  const double pc = fma(z, fma(z, fma(z, fma(z, fma(z, -1.13596475577881948265e-11, 2.08757232129817482790e-09),
                                              -2.75573143513906633035e-07), 2.48015872894767294178e-05),
                                -1.38888888888741095749e-03), 4.16666666666666019037e-02);
  const double sr = fma(r * z, ps, r);
  const double hz = 0.5 * z, w = 1.0 - hz;
  const double cr = w + (((1.0 - w) - hz) + z * z * pc);
  const int q = (int)n & 3;
  const double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// asmTF (optics/propagate.go:159-181) in the reference's operation order.
__device__ __forceinline__ cpx<double> asm_tf_fp64(const TFUnit& u, double fx, double fy) {
#pragma clang fp contract(off)
  double rho = go_hypot(u.wl * fx, u.wl * fy);
  if (rho <= 1) {
    double ph = u.k * u.z * sqrt((1 - rho) * (1 + rho));
    double s, c;
    sincos_tf64(ph, &s, &c);
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
// in FP64 before rounding to FP32.
__device__ __forceinline__ float wrap_phase_f32(double ph) {
  const double inv2pi = 0.15915494309189533577;
  const double two_pi_hi = 6.28318530717958623200;     // fl(2 pi)
  const double two_pi_lo = 2.44929359829470635445e-16;  // 2 pi - fl(2 pi)
  double nq = rint(ph * inv2pi);
  double r = fma(-nq, two_pi_hi, ph);
  r = fma(-nq, two_pi_lo, r);
  return (float)r;
}

// sin/cos for the FP32 path of a phase given in quarter turns (phase =
// q4 * pi/2): the integer quadrant and the remainder in [-1/2, 1/2] quarter
// turns are split off in FP64, the remainder is rounded to FP32 and the
// Cephes single-precision minimax polynomials (S. Moshier, Cephes Math
// Library) evaluate sin/cos on [-pi/4, pi/4]. Branch-free; exact argument
// reduction for |q4| < 2^31.
__device__ __forceinline__ void sincos_quarter_f32(double q4, float* s, float* c) {
  const double qn = rint(q4);
  const float yf = (float)(q4 - qn) * 1.57079632679489661923f;
  const int qi = (int)qn & 3;
  const float z = yf * yf;
  const float sp = fmaf(yf * z, fmaf(z, fmaf(z, -1.9515295891e-4f, 8.3321608736e-3f), -1.6666654611e-1f), yf);
  const float cp = fmaf(z * z, fmaf(z, fmaf(z, 2.443315711809948e-5f, -1.388731625493765e-3f), 4.166664568298827e-2f),
                        fmaf(-0.5f, z, 1.0f));
  float ss = (qi & 1) ? cp : sp;
  float cc = (qi & 1) ? sp : cp;
  ss = (qi & 2) ? -ss : ss;
  cc = ((qi + 1) & 2) ? -cc : cc;
  *s = ss;
  *c = cc;
}

// sin/cos of an FP64 phase in radians for the FP32 path.
__device__ __forceinline__ void sincos_phase_f32(double ph, float* s, float* c) {
  sincos_quarter_f32(ph * 0.63661977236758134308, s, c);  // 2/pi
}

// Evanescent branch of asmTF (optics/propagate.go:168-180), out of line.
__device__ __noinline__ double asm_evanescent(const TFUnit& u, double decay) {
  if (u.zero_ev || (u.limit > 0 && decay > u.limit)) return 0.0;
  if (u.backward) {
    if (u.reg) {
      double a = exp(decay);
      return a / (1 + u.alpha * u.alpha * a * a);
    }
    return exp(decay);
  }
  return exp(-decay);
}


// asm_evanescent inline, for kernels that must not call out.
__device__ __forceinline__ double asm_evanescent_inl(const TFUnit& u, double decay) {
  if (u.zero_ev || (u.limit > 0 && decay > u.limit)) return 0.0;
  if (u.backward) {
    const double a = exp(decay);
    return u.reg ? a / (1 + u.alpha * u.alpha * a * a) : a;
  }
  return exp(-decay);
}

// sqrt for the FP32 path's phase: hardware reciprocal square root refined by
// one Newton step (relative error ~1e-16, enough for a phase kz*s with
// kz < 1e7 to stay far below the FP32 rounding of the field). 0 for x <= 0.
__device__ __forceinline__ double sqrt_fast(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  const double g = x * r, h = 0.5 * r;
  const double d = fma(-g, g, x);
  const double s = fma(d, h, g);
  return x > 0 ? s : 0.0;
}

// sin/cos of a phase given in turns (FP64) for the FP32 path: the reduction
// to [-1/2, 1/2] turn is exact in FP64, the hardware sine and cosine take
// turns (max error 1.3e-7 absolute, measured against FP64 over the range).
__device__ __forceinline__ void sincos_turns_f32(double q, float* s, float* c) {
  const float r = (float)(q - rint(q));
  *s = __builtin_amdgcn_sinf(r);
  *c = __builtin_amdgcn_cosf(r);
}

// Per-column constants of the FP32-path asmTF.
struct TF32Col {
  double one_m_ax2;  // 1 - (wl*fx)^2
  double cwl, coff;  // wl*fy = row*cwl + coff
  double kq;         // k*z in turns: (k*z)/(2 pi)
  int col;           // transform column (for rows at the cut-off)
};

__device__ __forceinline__ TF32Col tf32_col(const TFUnit& u, int col) {
  const double fx = (double)(col <= u.M / 2 ? col : col - u.M) * u.inv_den + u.fx_off;
  const double ax = u.wl * fx;
  TF32Col c;
  c.one_m_ax2 = fma(-ax, ax, 1.0);
  c.cwl = u.wl * u.inv_den;
  c.coff = u.wl * u.fy_off;
  // conjugate transfer function (adjoint): the propagating phase changes
  // sign, the real evanescent factor does not
  c.kq = (u.kind == TF_CONJ_ASM ? -u.kz : u.kz) * 0.15915494309189533577;  // 1/(2 pi)
  c.col = col;
  return c;
}

// The FP32 path's x = 1 - rho^2 (formed from reciprocal-multiplied
// frequencies with FMAs) differs from the reference's 1 - rho^2 (Go's
// hypot of divided frequencies, optics/propagate.go:162-166) by a few
// 1e-16. Rows with |x| below this band take the reference's own test, so a
// bin lying exactly on rho = 1 (e.g. i^2 + j^2 = 300^2 on a 1000-point grid
// with wl = 0.3 n dx) gets the reference's branch: under evanescent zeroing
// the two branches differ by the whole value.
constexpr double kTfEdge = 1e-14;

// Factor of the FP32 path for a row with x = 1 - rho^2 < kTfEdge, which
// multiplies the propagating value formed from x (exactly (1, 0) where
// x <= 0): 1 where the reference propagates, else its evanescent factor
// (optics/propagate.go:168-180).
template <bool INL>
__device__ __forceinline__ float asm_tf_fp32_edge(const TFUnit& u, int col, int iy, double x) {
  if (x >= kTfEdge) return 1.f;
  if (x > -kTfEdge) {
    const double rho = go_hypot(u.wl * (freq_of(col, u.M, u.den) + u.fx_off),
                                u.wl * (freq_of(iy, u.M, u.den) + u.fy_off));
    if (rho <= 1) return 1.f;
    x = -((rho - 1) * (rho + 1));
  }
  const double decay = u.k * u.az * sqrt(-x);
  return (float)(INL ? asm_evanescent_inl(u, decay) : asm_evanescent(u, decay));
}

// Propagating asmTF of the FP32 path at transform row iy: exp(i kz
// sqrt(1 - rho^2)) with the phase in FP64; rows with rho > 1 return 1, and
// rows with x < kTfEdge raise *evan for the evanescent fix-up (not
// bit-exact with the reference, which only the FP64 path requires).
__device__ __forceinline__ cpx<float> asm_tf_fp32_prop(const TF32Col& c, int iy, int M, bool* evan) {
  const double ay = fma((double)(iy <= M / 2 ? iy : iy - M), c.cwl, c.coff);
  const double x = fma(-ay, ay, c.one_m_ax2);
  *evan |= x < kTfEdge;
  float sn, cs;
  sincos_turns_f32(c.kq * sqrt_fast(x), &sn, &cs);
  return {cs, sn};
}

// Fix-up factor of the FP32 path at row iy (asm_tf_fp32_edge), out of line.
__device__ __noinline__ float asm_tf_fp32_evan(const TFUnit& u, TF32Col c, int iy) {
  const double ay = fma((double)(iy <= u.M / 2 ? iy : iy - u.M), c.cwl, c.coff);
  return asm_tf_fp32_edge<false>(u, c.col, iy, fma(-ay, ay, c.one_m_ax2));
}

// Fix-up factor of the FP32 path at row iy, inline (see asm_tf_fp64_inl
// for why the column passes make no calls).
__device__ __forceinline__ float asm_tf_fp32_evan_inl(const TFUnit& u, const TF32Col& c, int iy) {
  const double ay = fma((double)(iy <= u.M / 2 ? iy : iy - u.M), c.cwl, c.coff);
  return asm_tf_fp32_edge<true>(u, c.col, iy, fma(-ay, ay, c.one_m_ax2));
}

// The FP32-path asmTF at row iy, evanescent rows included. INL: the
// fix-up inline, for the register-resident column passes; the
// LDS-resident mixed-radix passes keep it out of line (a smaller loop body
// there).
template <bool INL = false>
__device__ __forceinline__ cpx<float> asm_tf_fp32_full(const TFUnit& u, const TF32Col& c, int iy) {
  bool evan = false;
  cpx<float> h = asm_tf_fp32_prop(c, iy, u.M, &evan);
  if (evan) h = cscale(h, INL ? asm_tf_fp32_evan_inl(u, c, iy) : asm_tf_fp32_evan(u, c, iy));
  return h;
}

// Propagating asmTF of the FP32 path from ay = wl*fy directly (the caller
// has folded the row and formed ay): x <= 0 is clamped to a tiny positive
// value instead of selecting s = 0 afterwards, so such rows get a phase of
// 0 turns, whose FP32 remainder rounds to 0: h = (1, 0) exactly, for the
// evanescent fix-up to scale. *evan is raised for x < kTfEdge.
__device__ __forceinline__ cpx<float> asm_tf_fp32_prop_ay(const TF32Col& c, double ay, bool* evan) {
  const double x = fma(-ay, ay, c.one_m_ax2);
  *evan |= x < kTfEdge;
  const double xc = fmax(x, 1e-300);
  const double r = __builtin_amdgcn_rsq(xc);
  const double g = xc * r, h = 0.5 * r;
  float sn, cs;
  sincos_turns_f32(c.kq * fma(fma(-g, g, xc), h, g), &sn, &cs);
  return {cs, sn};
}

// Fix-up factor of the FP32 path at row iy with ay = wl*fy given.
__device__ __forceinline__ float asm_tf_fp32_evan_ay(const TFUnit& u, const TF32Col& c, double ay, int iy) {
  return asm_tf_fp32_edge<true>(u, c.col, iy, fma(-ay, ay, c.one_m_ax2));
}

// FP32-path asmTF at any (fx, fy); INL keeps the evanescent branch inline
// (fx, fy: the reference's frequencies; within kTfEdge of the cut-off the
// reference's test on them picks the branch)
template <bool INL>
__device__ __forceinline__ cpx<float> asm_tf_fp32_generic_t(const TFUnit& u, double fx, double fy) {
  double ax = u.wl * fx, ay = u.wl * fy;
  double rho2 = ax * ax + ay * ay;
  bool prop = rho2 <= 1;
  if (fabs(rho2 - 1) < kTfEdge) prop = go_hypot(ax, ay) <= 1;
  if (prop) {
    float s, c;
    sincos_phase_f32(u.k * u.z * sqrt(fmax(1 - rho2, 0.0)), &s, &c);
    return {c, s};
  }
  const double decay = u.k * u.az * sqrt(fmax(rho2 - 1, 0.0));
  return {(float)(INL ? asm_evanescent_inl(u, decay) : asm_evanescent(u, decay)), 0.f};
}
__device__ __noinline__ cpx<float> asm_tf_fp32_generic(const TFUnit& u, double fx, double fy) {
  return asm_tf_fp32_generic_t<false>(u, fx, fy);
}

// Fresnel TF (optics/propagate.go:511-513): cexpI(k z) * cexpI(-pi wl z f^2)
__device__ __forceinline__ cpx<double> fresnel_tf_fp64(const TFUnit& u, double fx, double fy) {
#pragma clang fp contract(off)
  double ph = u.fr_c * (fx * fx + fy * fy);
  double s, c;
  sincos_tf64(ph, &s, &c);
  return {u.pre_re * c - u.pre_im * s, u.pre_re * s + u.pre_im * c};
}
__device__ __forceinline__ cpx<float> fresnel_tf_fp32(const TFUnit& u, double fx, double fy) {
  double ph = u.fr_c * (fx * fx + fy * fy);
  float s, c;
  sincos_phase_f32(ph, &s, &c);
  float pr = (float)u.pre_re, pi = (float)u.pre_im;
  return {pr * c - pi * s, pr * s + pi * c};
}

// Exact FP64 asmTF at transform position (ix, iy), entirely inline: a
// kernel whose registers are full of transform values keeps them across it.
// A call, even on the rare evanescent rows, makes the register allocator
// keep the live transform values in callee-saved registers or spill them
// around the whole transfer-function section: the FP64 2048-point column
// pass spilled 144-160 B per lane with calls here, 20 B without, and runs
// 2.62 -> 2.24 ms at 2048^2 x 64. Both branches take one square root of
// the reference's factored forms (optics/propagate.go:166,169).
__device__ __forceinline__ cpx<double> asm_tf_fp64_inl(const TFUnit& u, int ix, int iy) {
#pragma clang fp contract(off)
  const double fx = freq_of(ix, u.M, u.den) + u.fx_off;
  const double fy = freq_of(iy, u.M, u.den) + u.fy_off;
  const double rho = go_hypot(u.wl * fx, u.wl * fy);
  const bool prop = rho <= 1;
  const double sq = sqrt(prop ? (1 - rho) * (1 + rho) : (rho - 1) * (rho + 1));
  if (prop) {
    const double ph = u.k * u.z * sq;
    double sn, cs;
    sincos_tf64(ph, &sn, &cs);
    return {cs, sn};
  }
  return {asm_evanescent_inl(u, u.k * u.az * sq), 0.0};
}

// Exact FP64 asmTF at transform position (ix, iy), out of line so the
// sixteen call sites of a column thread share one copy of the code.
__device__ __noinline__ cpx<double> asm_tf_fp64_at(const TFUnit& u, int ix, int iy) {
#pragma clang fp contract(off)
  double fx = freq_of(ix, u.M, u.den) + u.fx_off;
  double fy = freq_of(iy, u.M, u.den) + u.fy_off;
  return asm_tf_fp64(u, fx, fy);
}

// ApplyBandlimit's roll-off (optics/field.go:276-289): exp(-d^2) with
// d = (hypot(fx, fy) - f0)/sig beyond f0, else 1 (Go's order, no FMA).
__device__ __forceinline__ double bandlimit_factor(const TFUnit& u, double fx, double fy) {
#pragma clang fp contract(off)
  const double fr = go_hypot(fx, fy);
  if (!(fr > u.bl_f0)) return 1.0;
  const double d = (fr - u.bl_f0) / u.bl_sig;
  return exp(-d * d);
}

// Any transfer-function kind (asm, conjugate asm, fresnel, table, band
// limit). tf_eval is out of line (one copy of the switch for the LDS-resident
// kernels); tf_eval_t<R, true> is the same code inline, for the
// register-resident column passes, where a call would cost the registers of
// every live transform value.
template <typename R, bool INL>
__device__ __forceinline__ cpx<R> tf_eval_t(const TFUnit& u, int ix, int iy) {
#pragma clang fp contract(off)
  const double fx = freq_of(ix, u.M, u.den) + u.fx_off;
  const double fy = freq_of(iy, u.M, u.den) + u.fy_off;
  if constexpr (sizeof(R) == 8) {
    switch (u.kind) {
      case TF_ASM: return asm_tf_fp64(u, fx, fy);
      case TF_CONJ_ASM: {
        cpx<double> h = asm_tf_fp64(u, fx, fy);
        return {h.x, -h.y};
      }
      case TF_FRESNEL: return fresnel_tf_fp64(u, fx, fy);
      case TF_TABLE: return ((const cpx<double>*)u.table)[(size_t)iy * u.M + ix];
      case TF_BANDLIMIT: return {bandlimit_factor(u, fx, fy), 0.0};
      default: return {1.0, 0.0};
    }
  } else {
    switch (u.kind) {
      case TF_ASM: return asm_tf_fp32_generic_t<INL>(u, fx, fy);
      case TF_CONJ_ASM: {
        cpx<float> h = asm_tf_fp32_generic_t<INL>(u, fx, fy);
        return {h.x, -h.y};
      }
      case TF_FRESNEL: return fresnel_tf_fp32(u, fx, fy);
      case TF_TABLE: return ((const cpx<float>*)u.table)[(size_t)iy * u.M + ix];
      case TF_BANDLIMIT: return {(float)bandlimit_factor(u, fx, fy), 0.f};
      default: return {1.f, 0.f};
    }
  }
}

template <typename R>
__device__ cpx<R> tf_eval(const TFUnit& u, int ix, int iy);
template <>
__device__ __noinline__ cpx<double> tf_eval<double>(const TFUnit& u, int ix, int iy) {
  return tf_eval_t<double, false>(u, ix, iy);
}
template <>
__device__ __noinline__ cpx<float> tf_eval<float>(const TFUnit& u, int ix, int iy) {
  return tf_eval_t<float, false>(u, ix, iy);
}

// TF_ASM_BL at transform position (ix, iy): the asm transfer function h and
// ApplyBandlimit's roll-off factor (inline; used by the column passes).
template <typename R>
struct TfBand {
  cpx<R> h;
  R band;
};
template <typename R>
__device__ __forceinline__ TfBand<R> asm_bl_eval(const TFUnit& u, int ix, int iy) {
#pragma clang fp contract(off)
  const double fx = freq_of(ix, u.M, u.den), fy = freq_of(iy, u.M, u.den);
  if constexpr (sizeof(R) == 8) return {asm_tf_fp64(u, fx, fy), bandlimit_factor(u, fx, fy)};
  else return {asm_tf_fp32_generic_t<true>(u, fx, fy), (float)bandlimit_factor(u, fx, fy)};
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
// Stage twiddle table of one Stockham stage (NS, RAD), built in FP64:
// tw[off + k*(RAD-1) + r-1] = exp(-2 pi i r k / (NS RAD)), k < NS, 1 <= r < RAD.
// The angle -2 r k/(NS RAD) (in units of pi) is exact: NS*RAD is a power of two.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_stage_twiddles(cpx<R>* tw, int off, int ns, int rad, int acc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ns * (rad - 1)) return;
  const int k = e / (rad - 1), r = e % (rad - 1) + 1;
  double s, c;
  sincospi(-2.0 * (double)(r * k) / (double)(ns * rad), &s, &c);
  const cpx<R> hi = {(R)c, (R)s};
  if (acc) {  // [hi, lo] pair: lo = rounded remainder of the FP64 twiddle
    tw[2 * (off + e)] = hi;
    tw[2 * (off + e) + 1] = {(R)(c - (double)hi.x), (R)(s - (double)hi.y)};
  } else {
    tw[off + e] = hi;
  }
}

// ---------------------------------------------------------------------------
// Row pass.
//
// Intermediate layout. Between the forward and the inverse row pass the
// spectrum is stored column-tiled: element (row r, column p) of a unit at
//   (p / TW) * tstride + r * TW + p % TW      (TW a power of two),
// so the column pass streams each tile of TW columns as one contiguous block
// and the row passes move whole 128-byte lines (ROWS rows x TW columns of a
// tile are contiguous). tw_log2 < 0 selects plain row-major (ld) addressing.
// In-unit offsets are 32-bit (a unit holds at most 2^26 elements).
// ---------------------------------------------------------------------------
struct RowArgs {
  const void* in;
  void* out;
  long long in_unit, out_unit;  // elements between units
  int in_ld, out_ld;            // elements between rows (row-major)
  int in_tw, out_tw;            // log2 of the column-tile width of a tiled layout (-1: row-major)
  int in_tstride, out_tstride;  // elements between column tiles (tiled layout)
  int rows;                     // rows per unit processed
  int units;
  int in_col0, in_cols;         // position p reads in[p - in_col0] when in range, else 0
  int out_col0, out_cols;       // position p is stored to out[p - out_col0] when in range
  double scale_re, scale_im;    // complex factor applied to the outputs
  double* pw_in;                // per-block sum |in|^2 (nullable)
  double* pw_out;               // per-block sum |out|^2 (nullable)
  const void* tw;               // N-point twiddles
  const void* screen;           // SCR kernels: N x N row-major complex128 phase screen
  int groups;                   // row groups per workgroup of the pipelined kernel (<= 1: plain)
  int split;                    // one-group kernel with split (real, imaginary) LDS exchanges
  const void* tw64;             // FP64 twiddles of the N points (pipelined kernel)
  int radix;                    // stage schedule of tw/tw64: 0 or 16 radix-16, 8 radix-8
};

// RowArgs with row-major addressing on both sides and no embed window.
inline RowArgs row_args_init() {
  RowArgs a;
  memset(&a, 0, sizeof(a));
  a.in_tw = a.out_tw = -1;
  a.scale_re = 1.0;
  return a;
}

template <int N, int VV = 16>
constexpr int rows_per_block() {
  return Plan<N, VV>::T >= 256 ? 1 : 256 / Plan<N, VV>::T;
}

// Streaming accesses of the FFT passes: every field element is read once and
// written once per pass, so loads and stores are marked nontemporal (the
// stores gain in every kernel; nontemporal loads slow the pipelined row
// kernel's prefetch, which uses gld<false>).
template <bool NT = true, typename R>
__device__ __forceinline__ cpx<R> gld(const cpx<R>* p) {
  if constexpr (NT) {
    typedef R v2 __attribute__((ext_vector_type(2)));
    const v2 x = __builtin_nontemporal_load((const v2*)p);
    return {x.x, x.y};
  } else {
    return *p;
  }
}
template <typename R>
__device__ __forceinline__ void gst(cpx<R>* p, cpx<R> v) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  const v2 x = {v.x, v.y};
  __builtin_nontemporal_store(x, (v2*)p);
}

// Nontemporal raw-buffer accesses: the descriptor (base, size) is uniform
// and lives in SGPRs, each thread keeps one 32-bit VGPR offset, and the
// offset of each of its slots is a compile-time byte count passed as the
// scalar offset, where per-slot 64-bit addresses would stay live from the
// loads to the stores of an in-place pass. `bytes` bounds the per-thread
// offset only (the scalar offsets are added after the range check), so
// callers keep every access inside the extent themselves.
template <typename R>
struct BufIO {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ BufIO(const void* base, uint32_t bytes) {
    // word 3 0x00020000: raw 32-bit data format, as gfx9 buffer loads take it
    rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
  }
  __device__ __forceinline__ cpx<R> ld(uint32_t voff, uint32_t soff) const {
    if constexpr (sizeof(R) == 4) {
      typedef float f2 __attribute__((ext_vector_type(2)));
      // the b64 builtin's result is an 8-byte value: reinterpret, never convert
      const auto x = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff, 2);
      static_assert(sizeof(x) == 8, "64-bit buffer load");
      const f2 f = __builtin_bit_cast(f2, x);
      return {f.x, f.y};
    } else {
      typedef double d2 __attribute__((ext_vector_type(2)));
      const auto x = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, (int)soff, 2);
      static_assert(sizeof(x) == 16, "128-bit buffer load");
      const d2 d = __builtin_bit_cast(d2, x);
      return {(R)d.x, (R)d.y};
    }
  }
  __device__ __forceinline__ void st(uint32_t voff, uint32_t soff, cpx<R> v) const {
    if constexpr (sizeof(R) == 4) {
      typedef float f2 __attribute__((ext_vector_type(2)));
      typedef unsigned v2u __attribute__((__vector_size__(2 * sizeof(unsigned))));
      const f2 f = {(float)v.x, (float)v.y};
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, f), rs, (int)voff, (int)soff, 2);
    } else {
      typedef unsigned v4u __attribute__((__vector_size__(4 * sizeof(unsigned))));
      typedef double d2 __attribute__((ext_vector_type(2)));
      const d2 d = {(double)v.x, (double)v.y};
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, d), rs, (int)voff, (int)soff, 2);
    }
  }
};

__device__ __forceinline__ int tiled_offset(int q, int row, int ld, int lg, int tstride) {
  if (lg < 0) return row * ld + q;
  return (q >> lg) * tstride + (row << lg) + (q & ((1 << lg) - 1));
}

// Threads are interleaved over the ROWS rows of a workgroup (row = tid %
// ROWS), so one wavefront covers 64/ROWS consecutive positions of ROWS rows.
// SCR = 1 (with DIR = +1): inverse row FFT -> scale -> multiply by the
// phase screen -> forward row FFT, the split-step BPM screen between two
// propagation steps (the inverse output map is the forward input map, so
// the data stays in registers).
// SPLIT: the stage exchanges move real and imaginary parts in two rounds
// through half the LDS, so more workgroups fit a CU.
// Per-row LDS stride of the row pass in complex values (SPLIT: half of a
// stride in values of R, rounded up so rows start complex-aligned; the
// multi-row strides are even already).
template <typename R, int N, int ROWS, bool SPLIT, int VV = 16>
constexpr int rows_lds_stride() {
  return SPLIT ? (lds_stride_elems<R>(N, ROWS) + 1) / 2 : row_lds_stride<R>(N, ROWS);
}
template <typename R, int N, int ROWS, bool SPLIT, int VV = 16>
constexpr int rows_lds_elems() { return ROWS * rows_lds_stride<R, N, ROWS, SPLIT, VV>(); }

// threadIdx.x; OPQ: hidden from the optimiser, so that a body run in a loop
// recomputes what it derives from the thread index (addresses, twiddle
// indices) instead of hoisting it and holding it live across iterations.
template <bool OPQ, int NT>
__device__ __forceinline__ int thread_index() {
  int tid = threadIdx.x;
  if constexpr (OPQ) {
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);
  }
  return tid;
}

// The row pass of workgroup `blk` (k_rows is one launch of it; the
// persistent transfer kernel runs it as a task). lds: rows_lds_elems
// complex values; red: one double per wave.
template <typename R, int N, int DIR, int ROWS, int SCR = 0, bool SPLIT = false, int VV = 16, bool OPQ = false>
__device__ __forceinline__ void rows_body(const RowArgs& a, int blk, cpx<R>* lds, double* red) {
  constexpr int T = Plan<N, VV>::T, V = Plan<N, VV>::V;
  static_assert(SCR == 0 || DIR > 0, "the screen follows an inverse transform");
  constexpr int R0 = first_rad<N, DIR, VV>(), RL = SCR ? last_rad<N, -1, VV>() : last_rad<N, DIR, VV>();
  constexpr int NT = ROWS * T;
  constexpr int LS = rows_lds_stride<R, N, ROWS, SPLIT, VV>();
  const int tid = thread_index<OPQ, NT>();
  const int lr = tid % ROWS, t = tid / ROWS;
  const int bpu = (a.rows + ROWS - 1) / ROWS;
  const int unit = blk / bpu;
  const int row = (blk % bpu) * ROWS + lr;
  const bool active = row < a.rows && unit < a.units;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit;
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  // uniform: every row of the grid is whole and no embed window applies
  const bool full_in = (a.rows % ROWS) == 0 && a.in_col0 == 0 && a.in_cols == N;
  const bool full_out = (a.rows % ROWS) == 0 && a.out_col0 == 0 && a.out_cols == N;

  cpx<R> v[V];
  if (full_in) {
#pragma unroll
    for (int b = 0; b < V / R0; ++b)
#pragma unroll
      for (int r = 0; r < R0; ++r)
        v[b * R0 + r] = gld(in + tiled_offset(slot_pos<N, R0, VV>(t, b, r), row, a.in_ld, a.in_tw, a.in_tstride));
  } else {
#pragma unroll
    for (int b = 0; b < V / R0; ++b)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int q = slot_pos<N, R0, VV>(t, b, r) - a.in_col0;
        cpx<R> x = {R(0), R(0)};
        if (active && q >= 0 && q < a.in_cols) x = in[tiled_offset(q, row, a.in_ld, a.in_tw, a.in_tstride)];
        v[b * R0 + r] = x;
      }
  }
  if (a.pw_in) {  // uniform: FP64 power of the input (zeros outside the window add nothing)
    double pin = 0;
#pragma unroll
    for (int i = 0; i < V; ++i) pin += (double)v[i].x * (double)v[i].x + (double)v[i].y * (double)v[i].y;
    block_sum_store<NT>(pin, red, a.pw_in, blk);
  }
  constexpr bool ACC = SCR == 1 && sizeof(R) == 4;  // split-step slices: compensated twiddles
  fft_block<R, N, DIR, 1, ACC, SPLIT, VV>(v, t, tw, lds + lr * LS);
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  if (a.scale_im == 0) {
    if (a.scale_re != 1) {
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = cscale(v[i], sc.x);
    }
  } else {
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = cmul(v[i], sc);
  }
  if constexpr (SCR == 1) {
    constexpr int RS = last_rad<N, +1, VV>();
    static_assert(RS == first_rad<N, -1, VV>(), "screen map");
    // FP64 screen applied in FP64: the same table multiplies every slice
    const cpx<double>* scr = (const cpx<double>*)a.screen + (size_t)row * N;
#pragma unroll
    for (int b = 0; b < V / RS; ++b)
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        const cpx<double> x = {(double)v[b * RS + r].x, (double)v[b * RS + r].y};
        const cpx<double> y = cmul(x, scr[slot_pos<N, RS, VV>(t, b, r)]);
        v[b * RS + r] = {(R)y.x, (R)y.y};
      }
    fft_block<R, N, -1, 1, ACC, SPLIT, VV>(v, t, tw, lds + lr * LS);
  }
  if (full_out) {
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r)
        gst(out + tiled_offset(slot_pos<N, RL, VV>(t, b, r), row, a.out_ld, a.out_tw, a.out_tstride), v[b * RL + r]);
  } else {
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int q = slot_pos<N, RL, VV>(t, b, r) - a.out_col0;
        if (active && q >= 0 && q < a.out_cols) out[tiled_offset(q, row, a.out_ld, a.out_tw, a.out_tstride)] = v[b * RL + r];
      }
  }
  if (a.pw_out) {  // uniform: FP64 power of the stored outputs
    double pout = 0;
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int q = slot_pos<N, RL, VV>(t, b, r) - a.out_col0;
        const cpx<R> y = v[b * RL + r];
        if (active && q >= 0 && q < a.out_cols) pout += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
      }
    block_sum_store<NT>(pout, red, a.pw_out, blk);
  }
}

template <typename R, int N, int DIR, int ROWS, int SCR = 0, bool SPLIT = false, int VV = 16>
__global__ __launch_bounds__((ROWS * Plan<N, VV>::T)) void k_rows(const RowArgs a) {
  __shared__ cpx<R> lds[rows_lds_elems<R, N, ROWS, SPLIT, VV>()];
  __shared__ double red[(ROWS * Plan<N, VV>::T + 63) / 64];
  rows_body<R, N, DIR, ROWS, SCR, SPLIT, VV>(a, blockIdx.x, lds, red);
}

// Pipelined row pass for whole rows (no embed window, no power sums, no
// screen): each workgroup transforms G groups of ROWS rows. The loads of
// group g+1 are issued before the transform of group g, and the twiddles sit
// in registers, so those loads stay in flight across the transform and the
// stores: a CU keeps HBM requests outstanding while it computes, where the
// one-group kernel alternates between loading and computing.
template <typename R, int N, int DIR, int ROWS, int G, int VV = 16>
__global__ __launch_bounds__((ROWS * Plan<N, VV>::T)) void k_rows_pipe(const RowArgs a) {
  constexpr int V = Plan<N, VV>::V;
  constexpr int R0 = first_rad<N, DIR, VV>(), RL = last_rad<N, DIR, VV>();
  constexpr int LS = row_lds_stride<R>(N, ROWS);
  __shared__ cpx<R> lds[ROWS * LS];
  const int tid = threadIdx.x;
  const int lr = tid % ROWS, t = tid / ROWS;
  const int bpu = a.rows / (ROWS * G);
  const int unit = blockIdx.x / bpu;
  const int row0 = (blockIdx.x % bpu) * ROWS * G + lr;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit;
  TwReg<R, N, DIR, VV> tr;
  tw_load<R, N, DIR, VV>(tr, t, (const cpx<double>*)a.tw64);
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  const bool scale = a.scale_im != 0 || a.scale_re != 1;
  cpx<R> v[V], nx[V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r)
      v[b * R0 + r] = gld<false>(in + tiled_offset(slot_pos<N, R0, VV>(t, b, r), row0, a.in_ld, a.in_tw, a.in_tstride));
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int row = row0 + g * ROWS;
    if (g + 1 < G) {
#pragma unroll
      for (int b = 0; b < V / R0; ++b)
#pragma unroll
        for (int r = 0; r < R0; ++r)
          nx[b * R0 + r] =
              gld<false>(in + tiled_offset(slot_pos<N, R0, VV>(t, b, r), row + ROWS, a.in_ld, a.in_tw, a.in_tstride));
    }
    fft_block_reg<R, N, DIR, 1, VV>(v, t, tr, lds + lr * LS);
    if (scale) {
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = cmul(v[i], sc);
    }
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r)
        gst(out + tiled_offset(slot_pos<N, RL, VV>(t, b, r), row, a.out_ld, a.out_tw, a.out_tstride), v[b * RL + r]);
    if (g + 1 < G) {
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = nx[i];
    }
  }
}

// Pipelined row pass with split exchanges and its stage twiddles in LDS:
// the workgroup copies its direction's twiddle table into LDS once, so the
// stages read twiddles with LDS loads (a separate counter from the global
// loads in flight) and no twiddle stays in registers; the next group's
// loads are issued before the current transform, as in k_rows_pipe. The
// registers hold two groups' values (64 VGPRs in FP32) instead of FP64
// twiddle seeds, so four-row workgroups fit two per CU.
template <typename R, int N, int DIR, int ROWS, int G, int VV = 16>
__global__ __launch_bounds__((ROWS * Plan<N, VV>::T)) void k_rows_pipe_lds(const RowArgs a) {
  constexpr int V = Plan<N, VV>::V, NT = ROWS * Plan<N, VV>::T;
  constexpr int R0 = first_rad<N, DIR, VV>(), RL = last_rad<N, DIR, VV>();
  constexpr int LS = rows_lds_stride<R, N, ROWS, true, VV>();
  constexpr int TWB = tw_dir_base<N, DIR, VV>(), TWN = tw_dir_len<N, DIR, VV>();
  __shared__ cpx<R> lds[rows_lds_elems<R, N, ROWS, true, VV>()];
  __shared__ cpx<R> ltw[TWN > 0 ? TWN : 1];
  const int tid = threadIdx.x;
  const int lr = tid % ROWS, t = tid / ROWS;
  const int bpu = a.rows / (ROWS * G);
  const int unit = blockIdx.x / bpu;
  const int row0 = (blockIdx.x % bpu) * ROWS * G + lr;
  const cpx<R>* in = (const cpx<R>*)a.in + (size_t)unit * a.in_unit;
  cpx<R>* out = (cpx<R>*)a.out + (size_t)unit * a.out_unit;
  const cpx<R>* gtw = (const cpx<R>*)a.tw + TWB;
  for (int i = tid; i < TWN; i += NT) ltw[i] = gtw[i];
  const cpx<R> sc = {(R)a.scale_re, (R)a.scale_im};
  const bool scale = a.scale_im != 0 || a.scale_re != 1;
  cpx<R> v[V], nx[V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r)
      v[b * R0 + r] = gld<false>(in + tiled_offset(slot_pos<N, R0, VV>(t, b, r), row0, a.in_ld, a.in_tw, a.in_tstride));
  __syncthreads();  // the twiddle table is in LDS
#pragma unroll 1
  for (int g = 0; g < G; ++g) {
    // the thread index, opaque per group: addresses derived from it are
    // recomputed each group instead of held live across the loop
    int t = tid / ROWS;
    asm volatile("" : "+v"(t));
    __builtin_assume(t >= 0 && t < Plan<N, VV>::T);
    const int row = row0 + g * ROWS;
    if (g + 1 < G) {
#pragma unroll
      for (int b = 0; b < V / R0; ++b)
#pragma unroll
        for (int r = 0; r < R0; ++r)
          nx[b * R0 + r] =
              gld<false>(in + tiled_offset(slot_pos<N, R0, VV>(t, b, r), row + ROWS, a.in_ld, a.in_tw, a.in_tstride));
    }
    fft_block<R, N, DIR, 1, false, true, VV, TWB>(v, t, ltw, lds + lr * LS);
    if (scale) {
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = cmul(v[i], sc);
    }
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r)
        gst(out + tiled_offset(slot_pos<N, RL, VV>(t, b, r), row, a.out_ld, a.out_tw, a.out_tstride), v[b * RL + r]);
    if (g + 1 < G) {
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = nx[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Column pass over a tile of W adjacent columns. MODE 0: plain transform
// (direction DIR, output scale); MODE 1: forward -> asm TF -> inverse (the
// ASM step); MODE 2: forward -> any TF kind -> inverse; MODE 3: the
// split-step variant of MODE 1 (see below); MODE 4: asm TF -> inverse on a
// spectrum already transformed forward (src, src_unit elements between
// units: 0 when every unit reads one shared spectrum), written to data. With `tiled` the
// tile is one contiguous block (row stride W, tiles tstride apart); otherwise
// rows are ld apart in a row-major array.
// ---------------------------------------------------------------------------
struct ColArgs {
  void* data;          // in place
  long long unit;      // elements between units
  int ld;              // elements between rows (row-major)
  int tiled;           // column-tiled layout (tile width 2^tlg, a multiple of W); 0: row-major
  int tstride;         // elements between tiles (tiled)
  int cols;            // columns per unit (multiple of W)
  int units;
  int in_row0, rows;   // transform position p maps to data row p - in_row0 when in range (else 0)
  int out_row0, out_rows;  // position p is stored to row p - out_row0 when in range
  double scale;
  int xcd_remap;       // spread consecutive tiles over one XCD
  const TFUnit* tf;    // per unit (MODE 1, 2)
  const void* tw;
  int tf_square;       // MODE 3: apply H^2 (two merged propagation steps)
  double* pw3;         // MODE 3: per-workgroup sums |U|^2, |H U|^2, |H^2 U|^2; MODE 2 with
                       // TF_ASM_BL: per-workgroup sum |H U|^2 (nullable)
  int tlg;             // log2 of the layout's tile width (tiled); a workgroup covers W of its columns
  const void* htab;    // MODE 3: FP64 H table in W-column tile order (required)
  int radix;           // stage schedule of tw: 0 or 16 radix-16, 8 radix-8 (MODE 1 only)
  int split;           // MODE 1 and 4, radix-16: exchanges through half the LDS (real, then imaginary parts)
  const void* src;     // MODE 4: the forward spectrum, same layout as data
  long long src_unit;  // MODE 4: elements between the units of src (0: one spectrum for every unit)
};

__device__ __forceinline__ int remap_xcd(int bid, int nb) {
  // The dispatcher assigns workgroup i to XCD (x0 + i) % 8 (x0 carries over
  // from the previous launch), so workgroups of equal i % 8 share an XCD;
  // give each such class a contiguous range of tiles so neighbouring column
  // tiles share its L2.
  if ((nb & 7) != 0) return bid;
  return (bid & 7) * (nb >> 3) + (bid >> 3);
}

// CS: the stage exchanges move real and imaginary parts in two rounds
// through half the LDS, so more workgroups fit a CU; in FP64 the kernel is
// also held to 128 VGPRs (4 waves per SIMD), as registers would otherwise
// keep a second workgroup out.
// LDS of the column pass in complex values (CS: split exchanges, half).
template <typename R, int N, int W, bool CS>
constexpr int cols_lds_elems() { return CS ? (lds_len(N) * W + 1) / 2 : lds_len(N) * W; }

// The column pass of workgroup `blk` of a launch of `nblk` workgroups
// (k_cols is one launch of it; the persistent transfer kernel runs it as a
// task). lds: cols_lds_elems complex values.
template <typename R, int N, int W, int MODE, int DIR, int VV = 16, bool CS = false, bool OPQ = false>
__device__ __forceinline__ void cols_body(const ColArgs& a, int blk, int nblk, cpx<R>* lds) {
  constexpr int T = Plan<N, VV>::T, V = Plan<N, VV>::V;
  const int tid = thread_index<OPQ, W * T>();
  const int c = tid % W, t = tid / W;
  // this thread's column in LDS (CS: values of R, W apart)
  cpx<R>* const lc = CS ? (cpx<R>*)((R*)lds + c) : lds + c;
  const int bid = a.xcd_remap ? remap_xcd(blk, nblk) : blk;
  const int tpu = a.cols / W;
  const int unit = bid / tpu;
  const int tile = bid % tpu;
  const int col = tile * W + c;
  // layout tile of 2^tlg columns; the workgroup's W columns are one part of it
  const int tl = a.tiled ? 1 << a.tlg : 0;
  const size_t toff = a.tiled ? (size_t)(col >> a.tlg) * a.tstride + (col & (tl - 1)) : (size_t)col;
  cpx<R>* base = (cpx<R>*)a.data + (size_t)unit * a.unit + toff;
  // where the loads read: in place, or (MODE 4) the spectrum
  const cpx<R>* ibase = MODE == 4 ? (const cpx<R>*)a.src + (size_t)unit * a.src_unit + toff : base;
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  const bool active = unit < a.units;
  // uniform fast path: contiguous tile, every transform position present;
  // tiles as wide as the workgroup (the usual shape) get compile-time strides
  const bool fast = a.tiled && a.in_row0 == 0 && a.rows == N && a.out_row0 == 0 && a.out_rows == N && active;
  const bool fastw = fast && tl == W;

  constexpr int DIR0 = MODE >= 1 ? -1 : DIR;
  // MODE 4 starts where the forward transform ends: its loads take the
  // forward plan's output map (natural order), the inverse plan's input map
  constexpr int R0 = MODE == 4 ? last_rad<N, -1, VV>() : first_rad<N, DIR0, VV>();
  cpx<R> v[V];
  if (fast && !fastw) {
#pragma unroll
    for (int b = 0; b < V / R0; ++b)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[b * R0 + r] = gld(ibase + slot_pos<N, R0, VV>(t, b, r) * tl);
  } else if (fastw && sizeof(R) == 8) {
    // FP64: the workgroup's tile through one buffer descriptor (uniform
    // base), one VGPR offset per thread (2.71 -> 2.61 ms at 2048^2 x 64;
    // the FP32 kernel is faster with plain addresses)
    const cpx<R>* ib = (MODE == 4 ? (const cpx<R>*)a.src + (size_t)unit * a.src_unit
                                  : (const cpx<R>*)a.data + (size_t)unit * a.unit) +
                       (size_t)tile * a.tstride;
    const BufIO<R> io(ib, (uint32_t)(N * W * sizeof(cpx<R>)));
    const uint32_t vo = (uint32_t)(c + t * W) * sizeof(cpx<R>);
#pragma unroll
    for (int b = 0; b < V / R0; ++b)
#pragma unroll
      for (int r = 0; r < R0; ++r)
        v[b * R0 + r] = io.ld(vo, (uint32_t)((b * T + r * (N / R0)) * W * sizeof(cpx<R>)));
  } else if (fastw) {
#pragma unroll
    for (int b = 0; b < V / R0; ++b)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[b * R0 + r] = gld(ibase + slot_pos<N, R0, VV>(t, b, r) * W);
  } else {
    const int ld = a.tiled ? tl : a.ld;
#pragma unroll
    for (int b = 0; b < V / R0; ++b)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int q = slot_pos<N, R0, VV>(t, b, r) - a.in_row0;
        cpx<R> x = {R(0), R(0)};
        if (active && q >= 0 && q < a.rows) x = ibase[q * ld];
        v[b * R0 + r] = x;
      }
  }
  constexpr bool ACC = MODE == 3 && sizeof(R) == 4;  // split-step slices: compensated twiddles
  if constexpr (MODE != 4) fft_block<R, N, DIR0, W, ACC, CS, VV>(v, t, tw, lc);
  constexpr bool ASM = MODE == 1 || MODE == 4;   // the lean asm transfer function
  if constexpr (MODE == 1 || MODE == 2 || MODE == 4) {
    constexpr int RL = last_rad<N, -1, VV>();
    if (active) {
      const TFUnit& u = a.tf[unit];  // by reference to global memory: a local copy would take registers
      // Without a spectral shift H(fx, -fy) = H(fx, fy), bit for bit (the
      // frequency of row N-p is the exact negative of row p's). Thread t
      // holds rows p = t + T m (slot (b, r): m = b + r V/RL); row N - p sits
      // in thread T - t at m' = V-1-m (thread 0: its own m' = V-m). Every
      // thread evaluates H for its rows m < V/2, passes them on through LDS
      // (free between the two transforms) and takes the rest from its
      // partner: each wave evaluates half as many transfer-function values.
      // The asm modes take no spectral shift (fy_off == 0): the host sends
      // shifted transfers (asm_shift) to MODE 2, so these kernels carry no
      // second transfer-function path, whose registers would spill the
      // transform values of the first.
      constexpr bool MIRROR = ASM && T >= 2 && V >= 2;
      if constexpr (MIRROR) {
        constexpr int NB = V / RL, V2 = V / 2;
        cpx<R>* hx = lds + c;
        const int pt = (T - t) & (T - 1);
        const TF32Col cc = tf32_col(u, col);  // FP32 path only
        const bool conj = u.kind == TF_CONJ_ASM;
        auto tf_at = [&](int p) -> cpx<R> {
          if constexpr (sizeof(R) == 4) {
            return asm_tf_fp32_full<true>(u, cc, p);
          } else {
            cpx<double> h = asm_tf_fp64_inl(u, col, p);
            if (conj) h.y = -h.y;
            return h;
          }
        };
        // One round through LDS as complex values in both layouts: the V/2
        // values a thread passes on fill (N/2) W complex slots, which the
        // split layout's half-size array holds too.
        __syncthreads();  // all waves are done reading the forward transform's last exchange
        if constexpr (sizeof(R) == 4) {
          // The rows t + T m (m <= V/2) are at most N/2, the non-negative
          // half: no fold, and wl*fy advances by T wl/N from row to row (one
          // add; the sum of up to V/2 steps stays within a few ulp of ay,
          // a phase error below 1e-11 turns). Evanescent rows (rho > 1,
          // rare) come out as h = (1, 0) and are scaled and rewritten after
          // the loop, by every thread whose extreme rows reach past the
          // cut-off: |wl*fy| peaks at row 0 or at the end of the walk (row
          // N/2 in thread 0, with the same rounding as in the loop; beyond
          // N/2 in the others).
          const double ay0 = fma((double)t, cc.cwl, cc.coff), dT = (double)T * cc.cwl;
          bool unused = false;
          double ay = ay0;
#pragma unroll
          for (int m = 0; m < V2; ++m) {
            const cpx<R> h = asm_tf_fp32_prop_ay(cc, ay, &unused);
            v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], h);
            hx[(m * T + t) * W] = h;
            ay += dT;
          }
          if (t == 0) {  // row N/2 (Nyquist) is its own mirror
            constexpr int m = V2;
            v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], asm_tf_fp32_prop_ay(cc, ay, &unused));
          }
          if (fmin(fma(-ay, ay, cc.one_m_ax2), fma(-cc.coff, cc.coff, cc.one_m_ax2)) < kTfEdge) {
            // ay again from an opaque copy of ay0: reusing the main loop's
            // values would keep them live across it (spilled)
            double aye = ay0;
            asm volatile("" : "+v"(aye));
#pragma unroll
            for (int m = 0; m <= V2; ++m) {
              if (m == V2 && t != 0) break;
              const float e = asm_tf_fp32_evan_ay(u, cc, aye, t + T * m);
              v[(m % NB) * RL + m / NB] = cscale(v[(m % NB) * RL + m / NB], e);
              // the partner's copy: only rows whose factor is not 1 change
              if (m < V2 && e != 1.f) {
                bool unused2 = false;
                hx[(m * T + t) * W] = cscale(asm_tf_fp32_prop_ay(cc, aye, &unused2), e);
              }
              aye += dT;
            }
          }
        } else {
#pragma unroll
          for (int m = 0; m < V2; ++m) {
            const cpx<R> h = tf_at(t + T * m);
            v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], h);
            hx[(m * T + t) * W] = h;
          }
          if (t == 0) {  // row N/2 (Nyquist) is its own mirror
            constexpr int m = V2;
            v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], tf_at(T * m));
          }
        }
        __syncthreads();
#pragma unroll
        for (int m = V2; m < V; ++m) {
          if (t == 0 && m == V2) continue;
          const int ms = t == 0 ? V - m : V - 1 - m;
          v[(m % NB) * RL + m / NB] = cmul(v[(m % NB) * RL + m / NB], hx[(ms * T + pt) * W]);
        }
      } else {
      if constexpr (ASM && sizeof(R) == 4) {
        const TF32Col cc = tf32_col(u, col);
        bool evan = false;
#pragma unroll
        for (int b = 0; b < V / RL; ++b)
#pragma unroll
          for (int r = 0; r < RL; ++r)
            v[b * RL + r] = cmul(v[b * RL + r], asm_tf_fp32_prop(cc, slot_pos<N, RL, VV>(t, b, r), u.M, &evan));
        if (evan) {  // rare: rows past the evanescent cut-off
#pragma unroll
          for (int b = 0; b < V / RL; ++b)
#pragma unroll
            for (int r = 0; r < RL; ++r)
              v[b * RL + r] = cscale(v[b * RL + r], asm_tf_fp32_evan_inl(u, cc, slot_pos<N, RL, VV>(t, b, r)));
        }
      } else if (MODE == 2 && u.kind == TF_ASM_BL) {
        // propagation then band limit: U <- B (H U), with the per-workgroup
        // sum of |H U|^2 (the propagated power, by Parseval) when pw3 is set
        double s1 = 0;
#pragma unroll
        for (int b = 0; b < V / RL; ++b)
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const TfBand<R> hb = asm_bl_eval<R>(u, col, slot_pos<N, RL, VV>(t, b, r));
            const cpx<R> y = cmul(v[b * RL + r], hb.h);
            s1 += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
            v[b * RL + r] = cscale(y, hb.band);
          }
        if (a.pw3) {  // uniform
          __shared__ double redc[(W * Plan<N, VV>::T + 63) / 64];
          block_sum_store<W * Plan<N, VV>::T>(s1, redc, a.pw3, blk);
        }
      } else {
        const bool conj = ASM && u.kind == TF_CONJ_ASM;
#pragma unroll
        for (int b = 0; b < V / RL; ++b)
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int p = slot_pos<N, RL, VV>(t, b, r);
            cpx<R> h;
            if constexpr (ASM) {
              h = asm_tf_fp64_inl(u, col, p);
              if (conj) h.y = -h.y;
            } else {
              h = tf_eval_t<R, true>(u, col, p);
            }
            v[b * RL + r] = cmul(v[b * RL + r], h);
          }
      }
      }
    }
  }
  if constexpr (MODE == 3) {
    // Split-step column pass: H (tf_square = 0) or H^2 (two merged half
    // steps), H evaluated exactly in FP64 and applied in FP64 arithmetic for
    // both precisions -- the same H multiplies the field every slice, so an
    // FP32-rounded H would accumulate its rounding linearly over the slices.
    // Optional spectral power sums |U|^2, |H U|^2, |H^2 U|^2 per workgroup.
    constexpr int RL = last_rad<N, -1, VV>();
    double s0 = 0, s1 = 0, s2 = 0;
    const bool want_pw = a.pw3 != nullptr;
    if (active) {
      // H of this column from the per-run table (htab is required: evaluating
      // H here would hold its registers across the whole pass)
      const cpx<double>* ht = (const cpx<double>*)a.htab + (size_t)(col / W) * N * W + (col % W);
#pragma unroll
      for (int b = 0; b < V / RL; ++b)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int p = slot_pos<N, RL, VV>(t, b, r);
          const cpx<double> h = ht[(size_t)p * W];
          const cpx<double> x = {(double)v[b * RL + r].x, (double)v[b * RL + r].y};
          const cpx<double> y = cmul(x, h);
          cpx<double> out = y;
          if (a.tf_square) out = cmul(y, h);
          if (want_pw) {
            s0 += x.x * x.x + x.y * x.y;
            s1 += y.x * y.x + y.y * y.y;
            s2 += out.x * out.x + out.y * out.y;
          }
          v[b * RL + r] = {(R)out.x, (R)out.y};
        }
    }
    if (want_pw) {  // uniform
      __shared__ double red3[3][(W * Plan<N, VV>::T + 63) / 64];
      double vals[3] = {s0, s1, s2};
      const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double x = vals[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red3[q][wave] = x;
      }
      __syncthreads();
      if (tid < 3) {
        double x = 0;
        for (int w = 0; w < (W * Plan<N, VV>::T + 63) / 64; ++w) x += red3[tid][w];
        a.pw3[(size_t)blk * 3 + tid] = x;
      }
    }
  }
  if constexpr (MODE >= 1) {
    fft_block<R, N, +1, W, ACC, CS, VV>(v, t, tw, lc);
  }
  constexpr int DIRL = MODE >= 1 ? +1 : DIR;
  constexpr int RL = last_rad<N, DIRL, VV>();
  const R sc = (R)a.scale;
  if (sc != R(1)) {
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = cscale(v[i], sc);
  }
  if (fastw && sizeof(R) == 8) {
    const BufIO<R> io((cpx<R>*)a.data + (size_t)unit * a.unit + (size_t)tile * a.tstride,
                      (uint32_t)(N * W * sizeof(cpx<R>)));
    const uint32_t vo = (uint32_t)(c + t * W) * sizeof(cpx<R>);
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r) io.st(vo, (uint32_t)((b * T + r * (N / RL)) * W * sizeof(cpx<R>)), v[b * RL + r]);
  } else if (fastw) {
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r) gst(base + slot_pos<N, RL, VV>(t, b, r) * W, v[b * RL + r]);
  } else if (fast) {
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r) gst(base + slot_pos<N, RL, VV>(t, b, r) * tl, v[b * RL + r]);
  } else {
    const int ld = a.tiled ? tl : a.ld;
#pragma unroll
    for (int b = 0; b < V / RL; ++b)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int q = slot_pos<N, RL, VV>(t, b, r) - a.out_row0;
        if (active && q >= 0 && q < a.out_rows) base[q * ld] = v[b * RL + r];
      }
  }
}

template <typename R, int N, int W, int MODE, int DIR, int VV = 16, bool CS = false>
__global__ __launch_bounds__((W * Plan<N, VV>::T), (((CS && sizeof(R) == 8) || VV == 32) ? 4 : 1))
void k_cols(const ColArgs a) {
  __shared__ cpx<R> lds[cols_lds_elems<R, N, W, CS>()];
  cols_body<R, N, W, MODE, DIR, VV, CS>(a, blockIdx.x, gridDim.x, lds);
}

// Arguments of the persistent transfer kernel (wos_persist.hip): the three
// passes of one transfer step on every unit of a batch, run from per-XCD
// work queues in one launch.
struct PersistArgs {
  RowArgs fwd;
  ColArgs col;
  RowArgs inv;
  int units;       // units of the batch
  int rb, cb;      // row-pass and column-pass workgroups per unit
  int interleave;  // 1: a round's passes task by task (needs rb == cb); 0: pass after pass
  unsigned* ctr;   // counters, persist_ctr_words(units) words (layout in wos_persist.hip)
  int spin_limit;  // sleeps before a dependency wait gives up
};
// Counter words: one 128-byte line per XCD (ticket, abort flag), a line for
// the sticky status, one line per unit (forward rows done, columns done).
constexpr int kPersistLine = 32;
constexpr int kPersistStatus = 8 * kPersistLine;
constexpr int kPersistUnits = kPersistStatus + kPersistLine;
inline size_t persist_ctr_words(int units) { return (size_t)kPersistUnits + (size_t)kPersistLine * units; }

}  // namespace wos
