/*
 * wos_oracle.c — FP64 CPU restatement of the Tetsuhiro-WOS propagation path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libwos_hip. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function follows the reference Go source cited beside it (paths are
 * relative to the reference repository). Arithmetic keeps the Go expression
 * order; the file is compiled with -O2 -ffp-contract=off and without
 * -ffast-math, which matches the reference's amd64 build (GOAMD64=v1 emits no
 * FMA). Go evaluates constant sub-expressions such as 2*math.Pi exactly before
 * rounding; 2.0*M_PI gives the same double because scaling by 2 is exact.
 */
#define _GNU_SOURCE
#include "wos_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <omp.h>

/* ------------------------------------------------------------------ */
/* complex helpers: Go complex128 semantics (no FMA, same op order)     */
/* ------------------------------------------------------------------ */
static inline or_cplx cmul(or_cplx a, or_cplx b) {
    /* Go ssagen: real = ar*br - ai*bi ; imag = ar*bi + ai*br */
    or_cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static inline or_cplx cadd(or_cplx a, or_cplx b) { or_cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline or_cplx csub(or_cplx a, or_cplx b) { or_cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline or_cplx cre(double x) { or_cplx r = {x, 0.0}; return r; }

/* cexpI (optics/field.go:231-234): math.Sincos */
static inline or_cplx cexpI(double phi) {
    double s, c;
    sincos(phi, &s, &c);
    or_cplx r = {c, s};
    return r;
}

/* Go math.Hypot (both the portable and the amd64 assembly version compute
 * max * sqrt(1 + (min/max)^2)); the special Inf/NaN cases are kept. */
double or_hypot(double p, double q) {
    p = fabs(p);
    q = fabs(q);
    if (isinf(p) || isinf(q)) return INFINITY;
    if (isnan(p) || isnan(q)) return NAN;
    if (p < q) { double t = p; p = q; q = t; }
    if (p == 0) return 0;
    q = q / p;
    return p * sqrt(1 + q * q);
}

/* ------------------------------------------------------------------ */
/* parFor (optics/fft.go:91-122)                                        */
/* ------------------------------------------------------------------ */
static int g_threads = 0;  /* 0: all hardware threads */

void or_set_threads(int p) { g_threads = p; }
int or_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_num_procs(); }

/* ------------------------------------------------------------------ */
/* FFT plan cache (optics/fft.go:14-51)                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int n, bits;
    int* rev;
    or_cplx* tw;  /* tw[k] = exp(-2 pi i k/n), k = 0..n/2 */
} fft_plan;

static fft_plan* g_plans[40];
static pthread_mutex_t g_plan_mu = PTHREAD_MUTEX_INITIALIZER;

static fft_plan* plan_fft(int n) {
    int bits = 0;
    for (int m = n; m > 1; m >>= 1) bits++;
    pthread_mutex_lock(&g_plan_mu);
    fft_plan* p = g_plans[bits];
    if (!p) {
        p = (fft_plan*)calloc(1, sizeof(fft_plan));
        p->n = n;
        p->bits = bits;
        p->rev = (int*)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; i++) {
            int r = 0, v = i;
            for (int b = 0; b < bits; b++) { r = (r << 1) | (v & 1); v >>= 1; }
            p->rev[i] = r;
        }
        int half = n / 2;
        p->tw = (or_cplx*)malloc(sizeof(or_cplx) * (size_t)(half + 1));
        for (int k = 0; k <= half; k++) {
            double ang = (-2.0 * M_PI) * (double)k / (double)n;
            p->tw[k].re = cos(ang);
            p->tw[k].im = sin(ang);
        }
        g_plans[bits] = p;
    }
    pthread_mutex_unlock(&g_plan_mu);
    return p;
}

/* fft1D (optics/fft.go:55-87): in-place radix-2 DIT; inverse conjugates the
 * twiddles and multiplies by complex(1/n, 0). */
static void fft1d_plan(or_cplx* a, const fft_plan* plan, int inverse) {
    int n = plan->n;
    for (int i = 0; i < n; i++) {
        int j = plan->rev[i];
        if (i < j) { or_cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int size = 2; size <= n; size <<= 1) {
        int half = size / 2;
        int step = n / size;
        for (int i = 0; i < n; i += size) {
            for (int j = 0; j < half; j++) {
                or_cplx w = plan->tw[j * step];
                if (inverse) w.im = -w.im;
                or_cplx u = a[i + j];
                or_cplx v = cmul(a[i + j + half], w);
                a[i + j] = cadd(u, v);
                a[i + j + half] = csub(u, v);
            }
        }
    }
    if (inverse) {
        or_cplx inv = cre(1.0 / (double)n);
        for (int i = 0; i < n; i++) a[i] = cmul(a[i], inv);
    }
}

void or_fft1d_radix2(or_cplx* a, int n, int inverse) {
    if (n <= 1) return;
    fft1d_plan(a, plan_fft(n), inverse);
}

/* fft1DBluestein (optics/fft.go:159-204) */
void or_fft1d_bluestein(or_cplx* a, int n, int inverse) {
    if (n <= 1) return;
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    double sign = inverse ? 1.0 : -1.0;
    or_cplx* chirp = (or_cplx*)malloc(sizeof(or_cplx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        /* sign * math.Pi * float64(k*k) / float64(n) */
        chirp[k] = cexpI(sign * M_PI * (double)((int64_t)k * k) / (double)n);
    }
    or_cplx* a0 = (or_cplx*)calloc((size_t)m, sizeof(or_cplx));
    or_cplx* b0 = (or_cplx*)calloc((size_t)m, sizeof(or_cplx));
    for (int k = 0; k < n; k++) a0[k] = cmul(a[k], chirp[k]);
    b0[0] = cre(1.0);
    for (int k = 1; k < n; k++) {
        or_cplx ck = {chirp[k].re, -chirp[k].im};
        b0[k] = ck;
        b0[m - k] = ck;
    }
    fft_plan* plan = plan_fft(m);
    fft1d_plan(a0, plan, 0);
    fft1d_plan(b0, plan, 0);
    for (int k = 0; k < m; k++) a0[k] = cmul(a0[k], b0[k]);
    fft1d_plan(a0, plan, 1);
    for (int k = 0; k < n; k++) a[k] = cmul(a0[k], chirp[k]);
    if (inverse) {
        or_cplx inv = cre(1.0 / (double)n);
        for (int k = 0; k < n; k++) a[k] = cmul(a[k], inv);
    }
    free(chirp);
    free(a0);
    free(b0);
}

/* fft1DAnyCPU (optics/fft.go:146-153) */
void or_fft1d(or_cplx* a, int n, int inverse) {
    if (n <= 0) return;
    if ((n & (n - 1)) == 0) {
        or_fft1d_radix2(a, n, inverse);
        return;
    }
    or_fft1d_bluestein(a, n, inverse);
}

/* fft2DCPU (optics/fft.go:219-242): rows, then strided columns through a
 * per-worker scratch buffer; parFor runs serially for n < 64. */
void or_fft2d(or_cplx* a, int n, int inverse) {
    int P = or_get_threads();
    if (P > n) P = n;
    int par = (P > 1 && n >= 64);
    int chunk = par ? (n + P - 1) / P : n;
    if (par) omp_set_num_threads(P);
#pragma omp parallel for schedule(static, chunk) if (par)
    for (int r = 0; r < n; r++) or_fft1d(a + (size_t)r * n, n, inverse);
#pragma omp parallel if (par)
    {
        or_cplx* col = (or_cplx*)malloc(sizeof(or_cplx) * (size_t)n);
#pragma omp for schedule(static, chunk)
        for (int c = 0; c < n; c++) {
            for (int r = 0; r < n; r++) col[r] = a[(size_t)r * n + c];
            or_fft1d(col, n, inverse);
            for (int r = 0; r < n; r++) a[(size_t)r * n + c] = col[r];
        }
        free(col);
    }
}

/* ------------------------------------------------------------------ */
/* Field helpers (optics/field.go)                                      */
/* ------------------------------------------------------------------ */
/* freq (optics/field.go:297-303) */
double or_freq(int i, int n, double dx) {
    if (i <= n / 2) return (double)i / ((double)n * dx);
    return (double)(i - n) / ((double)n * dx);
}

/* X/Y (optics/field.go:59-62) */
static inline double coord(int i, int n, double dx) { return ((double)i - (double)n / 2) * dx; }

static inline double intensity_at(const or_cplx* ex, const or_cplx* ey, size_t idx) {
    double p = ex[idx].re * ex[idx].re + ex[idx].im * ex[idx].im;
    if (ey) p += ey[idx].re * ey[idx].re + ey[idx].im * ey[idx].im;
    return p;
}

/* Power (optics/field.go:77-83): serial sum, then * DX * DX */
double or_power(const or_cplx* ex, const or_cplx* ey, int n, double dx) {
    double p = 0;
    size_t nn = (size_t)n * n;
    for (size_t i = 0; i < nn; i++) p += intensity_at(ex, ey, i);
    return p * dx * dx;
}

static void warn_add(or_warnings* w, int code, double value) {
    if (!w) return;
    w->count[code]++;
    if (value != 0) w->value[code] = value;
}

/* ------------------------------------------------------------------ */
/* Transfer functions (optics/propagate.go)                             */
/* ------------------------------------------------------------------ */
typedef struct {
    double wl, z, k, az;
    int zero_ev, reg, backward;
    double alpha, limit;
} asm_tf;

/* evanescentPolicy (optics/propagate.go:186-199) */
static void evanescent_policy(const or_ctx* ctx, double z, int* zero, int* reg, double* alpha,
                              double* limit) {
    *limit = ctx->evanescent_limit;
    if (z < 0) {
        if (ctx->backward_regularize) {
            *alpha = ctx->tikhonov_alpha;
            if (*alpha <= 0) *alpha = 1e-3;
            *zero = 0;
            *reg = 1;
            return;
        }
        *zero = 1;
        *reg = 0;
        *alpha = 0;
        return;
    }
    *zero = ctx->evanescent_zero != 0;
    *reg = 0;
    *alpha = 0;
}

/* asmTF (optics/propagate.go:159-181) */
static asm_tf make_asm_tf(double wl, double z, int zero_ev, int reg, double alpha, double limit) {
    asm_tf t;
    t.wl = wl;
    t.z = z;
    t.k = 2.0 * M_PI / wl;
    t.az = fabs(z);
    t.backward = z < 0;
    t.zero_ev = zero_ev;
    t.reg = reg;
    t.alpha = alpha;
    t.limit = limit;
    return t;
}

static inline or_cplx asm_tf_eval(const asm_tf* t, double fx, double fy) {
    double rho = or_hypot(t->wl * fx, t->wl * fy);
    if (rho <= 1) return cexpI(t->k * t->z * sqrt((1 - rho) * (1 + rho)));
    double decay = t->k * t->az * sqrt((rho - 1) * (rho + 1));
    if (t->zero_ev || (t->limit > 0 && decay > t->limit)) return cre(0);
    if (t->backward) {
        if (t->reg) {
            double a = exp(decay);
            return cre(a / (1 + t->alpha * t->alpha * a * a));
        }
        return cre(exp(decay));
    }
    return cre(exp(-decay));
}

/* asmTF1D (optics/propagate.go:679-701) */
static inline or_cplx asm_tf1d_eval(const asm_tf* t, double fx) {
    double rho = fabs(t->wl * fx);
    if (rho <= 1) return cexpI(t->k * t->z * sqrt((1 - rho) * (1 + rho)));
    double decay = t->k * t->az * sqrt((rho - 1) * (rho + 1));
    if (t->zero_ev || (t->limit > 0 && decay > t->limit)) return cre(0);
    if (t->backward) {
        if (t->reg) {
            double a = exp(decay);
            return cre(a / (1 + t->alpha * t->alpha * a * a));
        }
        return cre(exp(decay));
    }
    return cre(exp(-decay));
}

/* asmWarn (optics/propagate.go:203-218) */
static void asm_warn(or_warnings* w, double z, double p_in, double p_out, int zero_ev, int reg) {
    if (p_in > 0 && p_out < p_in * (1 - 1e-8) && z > 0) warn_add(w, OR_WARN_EVANESCENT_FILTERED, (p_in - p_out) / p_in);
    if (z < 0) {
        if (reg) warn_add(w, OR_WARN_BACKWARD_REGULARIZED, 0);
        else if (zero_ev) warn_add(w, OR_WARN_BACKWARD_EVANESCENT, 0);
    }
}

/* spectralRadius (optics/propagate.go:270-320) */
double or_spectral_radius(const or_cplx* ex, const or_cplx* ey, int n, double dx, double frac) {
    if (frac <= 0 || frac >= 1) frac = 0.99;
    const int bins = 96;
    double fnyq = 1 / (2 * dx);
    if (n == 0 || dx <= 0) return 0;
    double hist[97];
    memset(hist, 0, sizeof(hist));
    double tot = 0;
    size_t nn = (size_t)n * n;
    or_cplx* a = (or_cplx*)malloc(sizeof(or_cplx) * nn);
    for (int comp = 0; comp < 2; comp++) {
        const or_cplx* c = comp == 0 ? ex : ey;
        if (!c) continue;
        memcpy(a, c, sizeof(or_cplx) * nn);
        or_fft2d(a, n, 0);
        for (int j = 0; j < n; j++) {
            double fy = or_freq(j, n, dx);
            for (int i = 0; i < n; i++) {
                or_cplx v = a[(size_t)j * n + i];
                double wgt = v.re * v.re + v.im * v.im;
                if (wgt == 0) continue;
                int b = (int)(or_hypot(or_freq(i, n, dx), fy) / fnyq * bins);
                if (b > bins) b = bins;
                hist[b] += wgt;
                tot += wgt;
            }
        }
    }
    free(a);
    if (tot <= 0) return 0;
    double run = 0.0;
    for (int b = 0; b <= bins; b++) {
        run += hist[b];
        if (run >= frac * tot) return (double)b / bins * fnyq;
    }
    return fnyq;
}

/* asmWrapWarn (optics/propagate.go:247-264) */
static void asm_wrap_warn(const or_cplx* ex, const or_cplx* ey, int n, double dx, double z,
                          const or_ctx* ctx, or_warnings* w, double carry) {
    if (!w || z == 0 || carry <= 0) return;
    if (fabs(z) * ctx->wavelength / (2 * dx) <= carry) return;
    double fsig = or_spectral_radius(ex, ey, n, dx, 0.90);
    if (fsig <= 0) return;
    double walk = fabs(z) * ctx->wavelength * fsig;
    if (walk > carry) warn_add(w, OR_WARN_ASM_ALIAS_WRAP, walk / carry);
}

/* ctx.transfer (optics/propagate.go:132-148) specialised per kernel kind */
static void transfer_asm(or_cplx* a, int n, double dx, const asm_tf* t) {
    or_fft2d(a, n, 0);
    for (int j = 0; j < n; j++) {
        double fy = or_freq(j, n, dx);
        for (int i = 0; i < n; i++) {
            size_t idx = (size_t)j * n + i;
            a[idx] = cmul(a[idx], asm_tf_eval(t, or_freq(i, n, dx), fy));
        }
    }
    or_fft2d(a, n, 1);
}

/* propASM (optics/propagate.go:325-332) */
static void prop_asm(or_cplx* ex, or_cplx* ey, int n, double dx, double z, const or_ctx* ctx,
                     or_warnings* w, or_stats* st) {
    int zero, reg;
    double alpha, limit;
    evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
    asm_wrap_warn(ex, ey, n, dx, z, ctx, w, (double)n * dx);
    double p_in = or_power(ex, ey, n, dx);
    asm_tf t = make_asm_tf(ctx->wavelength, z, zero, reg, alpha, limit);
    transfer_asm(ex, n, dx, &t);
    if (ey) transfer_asm(ey, n, dx, &t);
    double p_out = or_power(ex, ey, n, dx);
    asm_warn(w, z, p_in, p_out, zero, reg);
    if (st) { st->p_in = p_in; st->p_out = p_out; }
}

/* propASMPadCore (optics/propagate.go:346-387) */
static void prop_asm_pad_core(or_cplx* ex, or_cplx* ey, int n, double dx, double z,
                              const or_ctx* ctx, or_warnings* w, or_stats* st, double fx_off,
                              double fy_off) {
    int m = 2 * n;
    int zero, reg;
    double alpha, limit;
    evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
    asm_tf t = make_asm_tf(ctx->wavelength, z, zero, reg, alpha, limit);
    asm_wrap_warn(ex, ey, n, dx, z, ctx, w, (double)n * dx);
    double p_in = or_power(ex, ey, n, dx);
    size_t mm = (size_t)m * m;
    or_cplx* buf = (or_cplx*)malloc(sizeof(or_cplx) * mm);
    int off = n / 2;
    for (int comp = 0; comp < 2; comp++) {
        or_cplx* a = comp == 0 ? ex : ey;
        if (!a) continue;
        memset(buf, 0, sizeof(or_cplx) * mm);
        for (int j = 0; j < n; j++)
            memcpy(buf + (size_t)(j + off) * m + off, a + (size_t)j * n, sizeof(or_cplx) * (size_t)n);
        or_fft2d(buf, m, 0);
        for (int j = 0; j < m; j++) {
            double fy = or_freq(j, m, dx) + fy_off;
            for (int i = 0; i < m; i++) {
                size_t idx = (size_t)j * m + i;
                buf[idx] = cmul(buf[idx], asm_tf_eval(&t, or_freq(i, m, dx) + fx_off, fy));
            }
        }
        or_fft2d(buf, m, 1);
        for (int j = 0; j < n; j++)
            memcpy(a + (size_t)j * n, buf + (size_t)(j + off) * m + off, sizeof(or_cplx) * (size_t)n);
    }
    free(buf);
    double p_out = or_power(ex, ey, n, dx);
    asm_warn(w, z, p_in, p_out, zero, reg);
    if (st) { st->p_in = p_in; st->p_out = p_out; }
}

/* spectralCentroid (optics/propagate.go:391-416) */
void or_spectral_centroid(const or_cplx* ex, const or_cplx* ey, int n, double dx, double* fxo,
                          double* fyo) {
    double sx = 0, sy = 0, sw = 0;
    size_t nn = (size_t)n * n;
    or_cplx* a = (or_cplx*)malloc(sizeof(or_cplx) * nn);
    for (int comp = 0; comp < 2; comp++) {
        const or_cplx* c = comp == 0 ? ex : ey;
        if (!c) continue;
        memcpy(a, c, sizeof(or_cplx) * nn);
        or_fft2d(a, n, 0);
        for (int j = 0; j < n; j++) {
            double fy = or_freq(j, n, dx);
            for (int i = 0; i < n; i++) {
                double fx = or_freq(i, n, dx);
                or_cplx v = a[(size_t)j * n + i];
                double wgt = v.re * v.re + v.im * v.im;
                sw += wgt;
                sx += wgt * fx;
                sy += wgt * fy;
            }
        }
    }
    free(a);
    if (sw <= 0) { *fxo = 0; *fyo = 0; return; }
    *fxo = sx / sw;
    *fyo = sy / sw;
}

/* the carrier ramp closure of propASMShift/propASMShiftPad
 * (optics/propagate.go:437-444, 487-494) */
static void shift_ramp(or_cplx* a, int n, double dx, double fxc, double fyc, double sgn) {
    for (int j = 0; j < n; j++) {
        double y = coord(j, n, dx);
        for (int i = 0; i < n; i++) {
            size_t idx = (size_t)j * n + i;
            a[idx] = cmul(a[idx], cexpI(sgn * 2 * M_PI * (fxc * coord(i, n, dx) + fyc * y)));
        }
    }
}

/* propASMShift (optics/propagate.go:426-471) */
static void prop_asm_shift(or_cplx* ex, or_cplx* ey, int n, double dx, double z, const or_ctx* ctx,
                           or_warnings* w, or_stats* st) {
    double fxc, fyc;
    or_spectral_centroid(ex, ey, n, dx, &fxc, &fyc);
    if (st) { st->fx_c = fxc; st->fy_c = fyc; }
    double fbin = 1 / ((double)n * dx);
    if (fabs(fxc) < 0.5 * fbin && fabs(fyc) < 0.5 * fbin) {
        if (st) st->method_used = OR_ASM;
        prop_asm(ex, ey, n, dx, z, ctx, w, st);
        return;
    }
    int zero, reg;
    double alpha, limit;
    evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
    asm_tf t = make_asm_tf(ctx->wavelength, z, zero, reg, alpha, limit);
    asm_wrap_warn(ex, ey, n, dx, z, ctx, w, (double)n * dx);
    double p_in = or_power(ex, ey, n, dx);
    shift_ramp(ex, n, dx, fxc, fyc, -1);
    if (ey) shift_ramp(ey, n, dx, fxc, fyc, -1);
    for (int comp = 0; comp < 2; comp++) {
        or_cplx* a = comp == 0 ? ex : ey;
        if (!a) continue;
        or_fft2d(a, n, 0);
        for (int j = 0; j < n; j++) {
            double fy = or_freq(j, n, dx) + fyc;
            for (int i = 0; i < n; i++) {
                size_t idx = (size_t)j * n + i;
                a[idx] = cmul(a[idx], asm_tf_eval(&t, or_freq(i, n, dx) + fxc, fy));
            }
        }
        or_fft2d(a, n, 1);
    }
    shift_ramp(ex, n, dx, fxc, fyc, +1);
    if (ey) shift_ramp(ey, n, dx, fxc, fyc, +1);
    double p_out = or_power(ex, ey, n, dx);
    asm_warn(w, z, p_in, p_out, zero, reg);
    if (st) { st->p_in = p_in; st->p_out = p_out; st->method_used = OR_ASM_SHIFT; }
}

/* propASMShiftPad (optics/propagate.go:479-504) */
static void prop_asm_shift_pad(or_cplx* ex, or_cplx* ey, int n, double dx, double z,
                               const or_ctx* ctx, or_warnings* w, or_stats* st) {
    double fxc, fyc;
    or_spectral_centroid(ex, ey, n, dx, &fxc, &fyc);
    if (st) { st->fx_c = fxc; st->fy_c = fyc; }
    double fbin = 1 / ((double)n * dx);
    if (fabs(fxc) < 0.5 * fbin && fabs(fyc) < 0.5 * fbin) {
        if (st) st->method_used = OR_ASM_PAD;
        prop_asm_pad_core(ex, ey, n, dx, z, ctx, w, st, 0, 0);
        return;
    }
    shift_ramp(ex, n, dx, fxc, fyc, -1);
    if (ey) shift_ramp(ey, n, dx, fxc, fyc, -1);
    prop_asm_pad_core(ex, ey, n, dx, z, ctx, w, st, fxc, fyc);
    shift_ramp(ex, n, dx, fxc, fyc, +1);
    if (ey) shift_ramp(ey, n, dx, fxc, fyc, +1);
    if (st) st->method_used = OR_ASM_SHIFT_PAD;
}

/* propFresnelTF (optics/propagate.go:507-519) */
static void prop_fresnel_tf(or_cplx* ex, or_cplx* ey, int n, double dx, double z,
                            const or_ctx* ctx, or_warnings* w) {
    double wl = ctx->wavelength;
    double k = 2.0 * M_PI / wl;
    for (int comp = 0; comp < 2; comp++) {
        or_cplx* a = comp == 0 ? ex : ey;
        if (!a) continue;
        or_fft2d(a, n, 0);
        for (int j = 0; j < n; j++) {
            double fy = or_freq(j, n, dx);
            for (int i = 0; i < n; i++) {
                double fx = or_freq(i, n, dx);
                or_cplx h = cmul(cexpI(k * z), cexpI(-M_PI * wl * z * (fx * fx + fy * fy)));
                size_t idx = (size_t)j * n + i;
                a[idx] = cmul(a[idx], h);
            }
        }
        or_fft2d(a, n, 1);
    }
    if (fabs(z) > (double)n * dx * dx / wl) warn_add(w, OR_WARN_FRESNEL_TF_ALIAS, 0);
}

/* propFresnelIR (optics/propagate.go:533-590) */
static void prop_fresnel_ir(or_cplx* ex, or_cplx* ey, int n, double dx, double z,
                            const or_ctx* ctx, or_warnings* w) {
    double wl = ctx->wavelength;
    double k = 2.0 * M_PI / wl;
    int m = 2 * n;
    size_t mm = (size_t)m * m;
    or_cplx* h = (or_cplx*)malloc(sizeof(or_cplx) * mm);
    for (int j = 0; j < m; j++) {
        double dy = (double)j;
        if (j > m / 2) dy -= (double)m;
        double y = dy * dx;
        for (int i = 0; i < m; i++) {
            double dxi = (double)i;
            if (i > m / 2) dxi -= (double)m;
            double x = dxi * dx;
            h[(size_t)j * m + i] = cexpI(M_PI * (x * x + y * y) / (wl * z));
        }
    }
    or_fft2d(h, m, 0);
    or_cplx c0 = {0, -1 / (wl * z)};
    or_cplx pre = cmul(cmul(c0, cexpI(k * z)), cre(dx * dx));
    or_cplx* buf = (or_cplx*)malloc(sizeof(or_cplx) * mm);
    int off = n / 2;
    for (int comp = 0; comp < 2; comp++) {
        or_cplx* a = comp == 0 ? ex : ey;
        if (!a) continue;
        memset(buf, 0, sizeof(or_cplx) * mm);
        for (int j = 0; j < n; j++)
            memcpy(buf + (size_t)(j + off) * m + off, a + (size_t)j * n, sizeof(or_cplx) * (size_t)n);
        or_fft2d(buf, m, 0);
        for (size_t i = 0; i < mm; i++) buf[i] = cmul(buf[i], h[i]);
        or_fft2d(buf, m, 1);
        for (size_t i = 0; i < mm; i++) buf[i] = cmul(buf[i], pre);
        for (int j = 0; j < n; j++)
            memcpy(a + (size_t)j * n, buf + (size_t)(j + off) * m + off, sizeof(or_cplx) * (size_t)n);
    }
    free(buf);
    free(h);
    if (fabs(z) < (double)n * dx * dx / wl) warn_add(w, OR_WARN_FRESNEL_IR_ALIAS, 0);
}

/* Propagate (optics/propagate.go:65-102) for the methods on this path. */
/* Field.rmsRadius (optics/propagate.go:105-130), serial sums */
static double rms_radius(const or_cplx* ex, const or_cplx* ey, int n, double dx) {
    double p = 0, px = 0, py = 0;
    for (int j = 0; j < n; j++) {
        double y = coord(j, n, dx);
        for (int i = 0; i < n; i++) {
            size_t idx = (size_t)j * n + i;
            double w = ex[idx].re * ex[idx].re + ex[idx].im * ex[idx].im;
            if (ey) w += ey[idx].re * ey[idx].re + ey[idx].im * ey[idx].im;
            p += w;
            px += w * coord(i, n, dx);
            py += w * y;
        }
    }
    if (p <= 0) return 0;
    double cx = px / p, cy = py / p;
    double vr = 0;
    for (int j = 0; j < n; j++) {
        double y = coord(j, n, dx);
        for (int i = 0; i < n; i++) {
            size_t idx = (size_t)j * n + i;
            double w = ex[idx].re * ex[idx].re + ex[idx].im * ex[idx].im;
            if (ey) w += ey[idx].re * ey[idx].re + ey[idx].im * ey[idx].im;
            double x = coord(i, n, dx);
            vr += w * ((x - cx) * (x - cx) + (y - cy) * (y - cy));
        }
    }
    return 2 * sqrt(vr / p);
}

/* propFraunhofer (optics/propagate.go:603-662): FFT, dx^2, the centred-layout
 * checkerboard, fftshift by n/2, the quadratic phase and prefactor on the
 * output grid of spacing wl |z| / (n dx) (the caller's new DX), and the
 * near-field warning on the input's rms radius. */
static void prop_fraunhofer(or_cplx* ex, or_cplx* ey, int n, double dx, double z, const or_ctx* ctx,
                            or_warnings* w) {
    double wl = ctx->wavelength;
    double k = 2.0 * M_PI / wl;
    double dx_out = wl * fabs(z) / ((double)n * dx);
    double d_in = rms_radius(ex, ey, n, dx);
    size_t nn = (size_t)n * n;
    or_cplx* shift = (or_cplx*)calloc(nn, sizeof(or_cplx));
    or_cplx* comps[2] = {ex, ey};
    for (int c = 0; c < 2; c++) {
        or_cplx* a = comps[c];
        if (!a) continue;
        or_fft2d(a, n, 0);
        for (size_t i = 0; i < nn; i++) a[i] = cmul(a[i], cre(dx * dx));
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                if ((i + j) & 1) {
                    size_t idx = (size_t)j * n + i;
                    a[idx].re = -a[idx].re;
                    a[idx].im = -a[idx].im;
                }
        int half = n / 2;
        for (int j = 0; j < n; j++) {
            size_t row = (size_t)j * n;
            memcpy(shift + row, a + row + half, sizeof(or_cplx) * (size_t)half);
            memcpy(shift + row + half, a + row, sizeof(or_cplx) * (size_t)half);
        }
        for (int i = 0; i < n; i++)
            for (int j = 0; j < half; j++) {
                or_cplx t = shift[(size_t)j * n + i];
                shift[(size_t)j * n + i] = shift[(size_t)(j + half) * n + i];
                shift[(size_t)(j + half) * n + i] = t;
            }
        or_cplx c0 = cmul(cexpI(k * z), (or_cplx){0.0, -1 / (wl * z)});
        for (int j = 0; j < n; j++) {
            double y = ((double)j - (double)n / 2) * dx_out;
            for (int i = 0; i < n; i++) {
                double x = ((double)i - (double)n / 2) * dx_out;
                size_t idx = (size_t)j * n + i;
                shift[idx] = cmul(shift[idx], cmul(c0, cexpI(k * (x * x + y * y) / (2 * z))));
            }
        }
        memcpy(a, shift, sizeof(or_cplx) * nn);
    }
    free(shift);
    if (d_in > 0) {
        double nf = d_in * d_in / (wl * fabs(z));
        if (nf > 0.5) warn_add(w, OR_WARN_FRAUNHOFER_NEARFIELD, nf);
    }
}

int or_propagate(or_cplx* ex, or_cplx* ey, int n, double dx, double z, int method,
                 const or_ctx* ctx, or_warnings* w, or_stats* st) {
    if (st) { memset(st, 0, sizeof(*st)); st->method_used = method; }
    if (z == 0) return 0;
    if (!ctx) return 2;
    switch (method) {
    case OR_ASM: prop_asm(ex, ey, n, dx, z, ctx, w, st); break;
    case OR_ASM_PAD: prop_asm_pad_core(ex, ey, n, dx, z, ctx, w, st, 0, 0); break;
    case OR_ASM_SHIFT: prop_asm_shift(ex, ey, n, dx, z, ctx, w, st); break;
    case OR_ASM_SHIFT_PAD: prop_asm_shift_pad(ex, ey, n, dx, z, ctx, w, st); break;
    case OR_FRESNEL_TF: prop_fresnel_tf(ex, ey, n, dx, z, ctx, w); break;
    case OR_FRESNEL_IR: prop_fresnel_ir(ex, ey, n, dx, z, ctx, w); break;
    case OR_FRAUNHOFER:
        if (n & 1) return 3;  /* the reference's fftshift leaves a stale column for odd n */
        prop_fraunhofer(ex, ey, n, dx, z, ctx, w);
        break;
    default: return 2;
    }
    return 0;
}

/* PropagateUniaxial (optics/birefringent.go:19-61) */
static void uniaxial_apply(or_cplx* a, int n, double dx, double z, double k0, double no, double ne, int extra) {
    double az = fabs(z);
    or_fft2d(a, n, 0);
    for (int j = 0; j < n; j++) {
        double fy = or_freq(j, n, dx);
        for (int i = 0; i < n; i++) {
            double fx = or_freq(i, n, dx);
            double arg;
            if (extra) {
                double kx = 2 * M_PI * fx, ky = 2 * M_PI * fy;
                arg = k0 * k0 * ne * ne - (ne * ne / (no * no)) * kx * kx - ky * ky;
            } else {
                arg = k0 * k0 * no * no - (2 * M_PI * fx) * (2 * M_PI * fx) - (2 * M_PI * fy) * (2 * M_PI * fy);
            }
            or_cplx h = arg < 0 ? cre(exp(-az * sqrt(-arg))) : cexpI(z * sqrt(arg));
            size_t idx = (size_t)j * n + i;
            a[idx] = cmul(a[idx], h);
        }
    }
    or_fft2d(a, n, 1);
}
void or_propagate_uniaxial(or_cplx* ex, or_cplx* ey, int n, double dx, double z, double no, double ne, double wl) {
    double k0 = 2 * M_PI / wl;
    uniaxial_apply(ex, n, dx, z, k0, no, ne, 1);
    if (ey) uniaxial_apply(ey, n, dx, z, k0, no, ne, 0);
}


/* Go's complex division (runtime complex128div, Smith's algorithm), for the
 * divisor forms of PropagateVectorial: (kz, 0) and (0, kzi). */
static inline or_cplx cdiv(or_cplx n, or_cplx m) {
    or_cplx r;
    if (fabs(m.re) >= fabs(m.im)) {
        double ratio = m.im / m.re;
        double denom = m.re + ratio * m.im;
        r.re = (n.re + n.im * ratio) / denom;
        r.im = (n.im - n.re * ratio) / denom;
    } else {
        double ratio = m.re / m.im;
        double denom = m.im + ratio * m.re;
        r.re = (n.re * ratio + n.im) / denom;
        r.im = (n.im * ratio - n.re) / denom;
    }
    return r;
}

/* ---- Berreman 4x4 (optics/berreman.go) with Go complex128 semantics ---- */
typedef struct { or_cplx m[4][4]; } mat4;
static inline or_cplx cneg(or_cplx a) { or_cplx r = {-a.re, -a.im}; return r; }
/* buildBerreman (optics/berreman.go:90-112) */
static mat4 build_berreman(const or_cplx* e, double xi, double eta) {
#define E(i, j) e[3 * (i) + (j)]
    or_cplx r = cdiv(cre(1), E(2, 2));
    or_cplx x = cre(xi), y = cre(eta);
    mat4 D;
    D.m[0][0] = cmul(cmul(cneg(x), r), E(2, 0));
    D.m[0][1] = csub(cre(1), cmul(cmul(x, x), r));
    D.m[0][2] = cmul(cmul(cneg(x), r), E(2, 1));
    D.m[0][3] = cmul(cmul(cneg(x), r), y);
    D.m[1][0] = csub(csub(E(0, 0), cmul(y, y)), cmul(cmul(E(0, 2), r), E(2, 0)));
    D.m[1][1] = cmul(cmul(cneg(E(0, 2)), r), x);
    D.m[1][2] = csub(cadd(E(0, 1), cmul(y, x)), cmul(cmul(E(0, 2), r), E(2, 1)));
    D.m[1][3] = cmul(cmul(cneg(E(0, 2)), r), y);
    D.m[2][0] = cmul(cmul(cneg(y), r), E(2, 0));
    D.m[2][1] = cmul(cmul(cneg(y), r), x);
    D.m[2][2] = cmul(cmul(cneg(y), r), E(2, 1));
    D.m[2][3] = csub(cre(1), cmul(cmul(y, y), r));
    D.m[3][0] = csub(cadd(E(1, 0), cmul(x, y)), cmul(cmul(E(1, 2), r), E(2, 0)));
    D.m[3][1] = cmul(cmul(cneg(E(1, 2)), r), x);
    D.m[3][2] = csub(csub(E(1, 1), cmul(x, x)), cmul(cmul(E(1, 2), r), E(2, 1)));
    D.m[3][3] = cmul(cmul(cneg(E(1, 2)), r), y);
#undef E
    return D;
}
static or_cplx trace4(const mat4* A) { return cadd(cadd(cadd(A->m[0][0], A->m[1][1]), A->m[2][2]), A->m[3][3]); }
static mat4 mat_mul4(const mat4* A, const mat4* B) {
    mat4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            or_cplx s = cre(0);
            for (int k = 0; k < 4; k++) s = cadd(s, cmul(A->m[i][k], B->m[k][j]));
            o.m[i][j] = s;
        }
    return o;
}
/* addMat4(B, scaleMat4(I, c)) */
static mat4 add_scaled_identity(const mat4* B, or_cplx c) {
    mat4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o.m[i][j] = cadd(B->m[i][j], cmul(cre(i == j ? 1.0 : 0.0), c));
    return o;
}
/* charPoly4, Faddeev-LeVerrier (optics/berreman.go:127-143) */
static void char_poly4(const mat4* A, or_cplx* c0, or_cplx* c1, or_cplx* c2, or_cplx* c3) {
    mat4 B = *A, t;
    *c3 = cneg(trace4(&B));
    t = add_scaled_identity(&B, *c3);
    B = mat_mul4(A, &t);
    *c2 = cdiv(cneg(trace4(&B)), cre(2));
    t = add_scaled_identity(&B, *c2);
    B = mat_mul4(A, &t);
    *c1 = cdiv(cneg(trace4(&B)), cre(3));
    t = add_scaled_identity(&B, *c1);
    B = mat_mul4(A, &t);
    *c0 = cdiv(cneg(trace4(&B)), cre(4));
}
static or_cplx poly4(or_cplx z, or_cplx c0, or_cplx c1, or_cplx c2, or_cplx c3) {
    or_cplx t = cadd(cmul(cmul(cmul(z, z), z), z), cmul(cmul(cmul(c3, z), z), z));
    t = cadd(t, cmul(cmul(c2, z), z));
    t = cadd(t, cmul(c1, z));
    return cadd(t, c0);
}
/* roots4, Durand-Kerner (optics/berreman.go:184-210) */
static void roots4(or_cplx c0, or_cplx c1, or_cplx c2, or_cplx c3, or_cplx* z) {
    z[0] = (or_cplx){0.4, 0.9};
    z[1] = (or_cplx){0.4, -0.9};
    z[2] = (or_cplx){-0.4, 0.9};
    z[3] = (or_cplx){-0.4, -0.9};
    for (int iter = 0; iter < 300; iter++) {
        double max_d = 0;
        for (int i = 0; i < 4; i++) {
            or_cplx num = poly4(z[i], c0, c1, c2, c3);
            or_cplx den = cre(1);
            for (int j = 0; j < 4; j++)
                if (j != i) den = cmul(den, csub(z[i], z[j]));
            or_cplx dz = cdiv(num, den);
            z[i] = csub(z[i], dz);
            double d = or_hypot(dz.re, dz.im);
            if (d > max_d) max_d = d;
        }
        if (max_d < 1e-14) break;
    }
}
static or_cplx det3(or_cplx m[3][3]) {
    or_cplx a = cmul(m[0][0], csub(cmul(m[1][1], m[2][2]), cmul(m[1][2], m[2][1])));
    or_cplx b = cmul(m[0][1], csub(cmul(m[1][0], m[2][2]), cmul(m[1][2], m[2][0])));
    or_cplx c = cmul(m[0][2], csub(cmul(m[1][0], m[2][1]), cmul(m[1][1], m[2][0])));
    return cadd(csub(a, b), c);
}
/* nullVector4: the largest-norm adjugate column of A - qI (optics/berreman.go:224-272) */
static void null_vector4(const mat4* A, or_cplx q, or_cplx* best) {
    mat4 M = *A;
    for (int i = 0; i < 4; i++) M.m[i][i] = csub(M.m[i][i], q);
    double best_n = -1;
    for (int col = 0; col < 4; col++) {
        or_cplx v[4];
        for (int row = 0; row < 4; row++) {
            or_cplx m[3][3];
            int ri = 0;
            for (int r = 0; r < 4; r++) {
                if (r == row) continue;
                int ci = 0;
                for (int c = 0; c < 4; c++) {
                    if (c == col) continue;
                    m[ri][ci] = M.m[r][c];
                    ci++;
                }
                ri++;
            }
            or_cplx cof = det3(m);
            if ((row + col) % 2 == 1) cof = cneg(cof);
            v[row] = cof;
        }
        double nrm = 0;
        for (int k = 0; k < 4; k++) nrm += v[k].re * v[k].re + v[k].im * v[k].im;
        if (nrm > best_n) {
            best_n = nrm;
            for (int k = 0; k < 4; k++) best[k] = v[k];
        }
    }
}
static int forward_mode(or_cplx q) {
    if (q.im > 1e-12) return 1;
    if (q.im < -1e-12) return 0;
    return q.re > 0;
}
/* anisoMode (optics/berreman.go:38-86) */
static void aniso_mode(const or_cplx* eps, double xi, double eta, double k0, double z, or_cplx* Ex, or_cplx* Ey) {
    mat4 D = build_berreman(eps, xi, eta);
    or_cplx c0, c1, c2, c3, roots[4];
    char_poly4(&D, &c0, &c1, &c2, &c3);
    roots4(c0, c1, c2, c3, roots);
    or_cplx qs[2] = {{0, 0}, {0, 0}}, vs[2][4];
    memset(vs, 0, sizeof(vs));
    int cnt = 0, used[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++)
        if (forward_mode(roots[i]) && cnt < 2) {
            qs[cnt] = roots[i];
            null_vector4(&D, roots[i], vs[cnt]);
            used[i] = 1;
            cnt++;
        }
    while (cnt < 2) {
        double best = -1e300;
        int bi = -1;
        for (int i = 0; i < 4; i++)
            if (!used[i] && roots[i].re > best) {
                best = roots[i].re;
                bi = i;
            }
        if (bi < 0) break;
        qs[cnt] = roots[bi];
        null_vector4(&D, roots[bi], vs[cnt]);
        used[bi] = 1;
        cnt++;
    }
    or_cplx a = vs[0][0], b = vs[1][0], c = vs[0][2], d = vs[1][2];
    or_cplx det = csub(cmul(a, d), cmul(b, c));
    or_cplx m1 = cdiv(csub(cmul(d, *Ex), cmul(b, *Ey)), det);
    or_cplx m2 = cdiv(cadd(cmul(cneg(c), *Ex), cmul(a, *Ey)), det);
    or_cplx pf0 = cmul(cexpI(k0 * qs[0].re * z), cre(exp(-k0 * qs[0].im * z)));
    or_cplx pf1 = cmul(cexpI(k0 * qs[1].re * z), cre(exp(-k0 * qs[1].im * z)));
    *Ex = cadd(cmul(cmul(m1, pf0), a), cmul(cmul(m2, pf1), b));
    *Ey = cadd(cmul(cmul(m1, pf0), c), cmul(cmul(m2, pf1), d));
}
void or_propagate_anisotropic(or_cplx* ex, or_cplx* ey, int n, double dx, double z, const or_cplx* eps, double wl) {
    double k0 = 2 * M_PI / wl;
    or_fft2d(ex, n, 0);
    or_fft2d(ey, n, 0);
    for (int j = 0; j < n; j++) {
        double fy = or_freq(j, n, dx);
        double ky = 2 * M_PI * fy;
        for (int i = 0; i < n; i++) {
            double fx = or_freq(i, n, dx);
            double kx = 2 * M_PI * fx;
            size_t idx = (size_t)j * n + i;
            aniso_mode(eps, kx / k0, ky / k0, k0, z, &ex[idx], &ey[idx]);
        }
    }
    or_fft2d(ex, n, 1);
    or_fft2d(ey, n, 1);
}

/* PropagateVectorial (optics/vectorial.go:14-80) */
void or_propagate_vectorial(or_cplx* ex, or_cplx* ey, or_cplx* ez, int n, double dx, double z,
                            const or_ctx* ctx) {
    double wl = ctx->wavelength;
    double k = 2 * M_PI / wl;
    double az = fabs(z);
    int backward = z < 0;
    int zero, reg;
    double alpha, limit;
    evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
    size_t nn = (size_t)n * n;
    or_fft2d(ex, n, 0);
    or_fft2d(ey, n, 0);
    memset(ez, 0, nn * sizeof(or_cplx));
    for (int j = 0; j < n; j++) {
        double fy = or_freq(j, n, dx);
        double ky = 2 * M_PI * fy;
        for (int i = 0; i < n; i++) {
            double fx = or_freq(i, n, dx);
            double kx = 2 * M_PI * fx;
            double rho = wl * or_hypot(fx, fy);
            size_t idx = (size_t)j * n + i;
            or_cplx H;
            or_cplx s = cadd(cmul(cre(kx), ex[idx]), cmul(cre(ky), ey[idx]));
            or_cplx ns = {-s.re, -s.im};
            if (rho <= 1) {
                double kz = k * sqrt((1 - rho) * (1 + rho));
                H = cexpI(kz * z);
                if (kz > 0) ez[idx] = cdiv(ns, cre(kz));
            } else {
                double decay = k * az * sqrt((rho - 1) * (rho + 1));
                if (zero || (limit > 0 && decay > limit)) H = cre(0);
                else if (backward && reg) {
                    double a = exp(decay);
                    H = cre(a / (1 + alpha * alpha * a * a));
                } else if (backward) H = cre(exp(decay));
                else H = cre(exp(-decay));
                double kzi = k * sqrt((rho - 1) * (rho + 1));
                if (kzi > 0) ez[idx] = cdiv(ns, (or_cplx){0.0, kzi});
            }
            ex[idx] = cmul(ex[idx], H);
            ey[idx] = cmul(ey[idx], H);
            ez[idx] = cmul(ez[idx], H);
        }
    }
    or_fft2d(ex, n, 1);
    or_fft2d(ey, n, 1);
    or_fft2d(ez, n, 1);
}

/* Propagate1D (optics/propagate.go:718-738) */
int or_propagate1d(or_cplx* a, int n, double dx, double z, double wl, const or_ctx* ctx,
                   or_warnings* w) {
    int zero, reg;
    double alpha, limit;
    evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
    asm_tf t = make_asm_tf(wl, z, zero, reg, alpha, limit);
    double p_in = 0;
    for (int i = 0; i < n; i++) p_in += a[i].re * a[i].re + a[i].im * a[i].im;
    or_fft1d(a, n, 0);
    for (int i = 0; i < n; i++) a[i] = cmul(a[i], asm_tf1d_eval(&t, or_freq(i, n, dx)));
    or_fft1d(a, n, 1);
    double p_out = 0;
    for (int i = 0; i < n; i++) p_out += a[i].re * a[i].re + a[i].im * a[i].im;
    if (p_in > 0 && p_out < p_in * (1 - 1e-8) && z > 0) warn_add(w, OR_WARN_EVANESCENT_FILTERED, (p_in - p_out) / p_in);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Split-step BPM (optics/medium.go:22-65)                              */
/* ------------------------------------------------------------------ */
static void apply_medium_phase(or_cplx* ex, or_cplx* ey, int n, const or_cplx* index, double k,
                               double dz) {
    size_t nn = (size_t)n * n;
    for (size_t idx = 0; idx < nn; idx++) {
        or_cplx nv = index[idx];
        double ph = k * (nv.re - 1) * dz;
        or_cplx t = cmul(cre(exp(-k * nv.im * dz)), cexpI(ph));
        ex[idx] = cmul(ex[idx], t);
        if (ey) ey[idx] = cmul(ey[idx], t);
    }
}

int or_split_step(or_cplx* ex, or_cplx* ey, int n, double dx, double z, int steps,
                  const or_cplx* index, int n_planes, const or_ctx* ctx, or_warnings* w) {
    if (z == 0) return 0;
    if (steps <= 0) steps = 1;
    if (n_planes != 1 && n_planes != steps) return 2;
    double dz = z / (double)steps;
    double k = 2.0 * M_PI / ctx->wavelength;
    size_t nn = (size_t)n * n;
    for (int s = 0; s < steps; s++) {
        or_propagate(ex, ey, n, dx, dz / 2, OR_ASM, ctx, w, NULL);
        apply_medium_phase(ex, ey, n, index + (n_planes == 1 ? 0 : (size_t)s * nn), k, dz);
        or_propagate(ex, ey, n, dx, dz / 2, OR_ASM, ctx, w, NULL);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Polychromatic and volume (optics/polychromatic.go:18-43,             */
/* optics/volume.go:8-18)                                               */
/* ------------------------------------------------------------------ */
int or_polychromatic(const or_cplx* ex, const or_cplx* ey, int n, double dx, double z,
                     const double* wavelengths, const double* weights, int nw, int method,
                     const or_ctx* ctx, double* out, or_warnings* w) {
    if (nw <= 0) return 2;
    size_t nn = (size_t)n * n;
    memset(out, 0, sizeof(double) * nn);
    or_cplx* gx = (or_cplx*)malloc(sizeof(or_cplx) * nn);
    or_cplx* gy = ey ? (or_cplx*)malloc(sizeof(or_cplx) * nn) : NULL;
    for (int s = 0; s < nw; s++) {
        if (wavelengths[s] <= 0) { free(gx); free(gy); return 2; }
        memcpy(gx, ex, sizeof(or_cplx) * nn);
        if (gy) memcpy(gy, ey, sizeof(or_cplx) * nn);
        or_ctx c = *ctx;
        c.wavelength = wavelengths[s];
        or_propagate(gx, gy, n, dx, z, method, &c, w, NULL);
        for (size_t i = 0; i < nn; i++) out[i] += weights[s] * intensity_at(gx, gy, i);
    }
    free(gx);
    free(gy);
    return 0;
}

int or_volume(const or_cplx* ex, const or_cplx* ey, int n, double dx, const double* zs, int nz,
              int method, const or_ctx* ctx, or_cplx* out_ex, or_cplx* out_ey, or_warnings* w) {
    size_t nn = (size_t)n * n;
    for (int i = 0; i < nz; i++) {
        or_cplx* gx = out_ex + (size_t)i * nn;
        or_cplx* gy = (ey && out_ey) ? out_ey + (size_t)i * nn : NULL;
        memcpy(gx, ex, sizeof(or_cplx) * nn);
        if (gy) memcpy(gy, ey, sizeof(or_cplx) * nn);
        or_propagate(gx, gy, n, dx, zs[i], method, ctx, w, NULL);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Input constructors                                                   */
/* ------------------------------------------------------------------ */
/* BuildSource "gaussian" before NormalizePower (optics/source.go:113-125) */
void or_gaussian(or_cplx* ex, int n, double dx, double w0) {
    for (int j = 0; j < n; j++) {
        double dy = coord(j, n, dx) - 0;
        for (int i = 0; i < n; i++) {
            double ddx = coord(i, n, dx) - 0;
            double r2 = ddx * ddx + dy * dy;
            ex[(size_t)j * n + i] = cre(exp(-r2 / (w0 * w0)));
        }
    }
}

/* lensEl.Apply, no aperture, centred (optics/elements.go:203-220) */
void or_apply_lens(or_cplx* ex, or_cplx* ey, int n, double dx, double wl, double f) {
    double k = 2.0 * M_PI / wl;
    double c = -k / (2 * f);
    for (int j = 0; j < n; j++) {
        double dy = coord(j, n, dx) - 0;
        for (int i = 0; i < n; i++) {
            double ddx = coord(i, n, dx) - 0;
            double r2 = ddx * ddx + dy * dy;
            or_cplx t = cexpI(c * r2);
            size_t idx = (size_t)j * n + i;
            ex[idx] = cmul(ex[idx], t);
            if (ey) ey[idx] = cmul(ey[idx], t);
        }
    }
}

/* apertureEl.Apply with a hard-edged circle (optics/elements.go:288-310,
 * optics/shape.go:162-170, step at optics/elements.go:143-148) */
void or_apply_circle_aperture(or_cplx* ex, or_cplx* ey, int n, double dx, double radius) {
    for (int j = 0; j < n; j++) {
        double dy = coord(j, n, dx) - 0;
        for (int i = 0; i < n; i++) {
            double ddx = coord(i, n, dx) - 0;
            double d = radius - or_hypot(ddx, dy);
            or_cplx t = cre(d >= 0 ? 1.0 : 0.0);
            size_t idx = (size_t)j * n + i;
            ex[idx] = cmul(ex[idx], t);
            if (ey) ey[idx] = cmul(ey[idx], t);
        }
    }
}

/* lensEl.Apply with its aperture and offset (optics/elements.go:195-220) */
void or_apply_lens_at(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double wl, double f,
                      double aperture, double x0, double y0) {
    double k = 2.0 * M_PI / wl;
    double c = -k / (2 * f);
    double ap = aperture;
    for (int j = 0; j < n; j++) {
        double dy = coord(j, n, dx) - y0;
        for (int i = 0; i < n; i++) {
            double ddx = coord(i, n, dx) - x0;
            double r2 = ddx * ddx + dy * dy;
            size_t idx = (size_t)j * n + i;
            if (ap > 0 && r2 > ap * ap) {
                ex[idx] = cre(0);
                if (polarized) ey[idx] = cre(0);
                continue;
            }
            or_cplx t = cexpI(c * r2);
            ex[idx] = cmul(ex[idx], t);
            if (polarized) ey[idx] = cmul(ey[idx], t);
        }
    }
}

/* regularPolygonVertices, starVertices (optics/elements.go:316-338) */
void or_regular_polygon(int n, double radius, double phi0, double* verts) {
    for (int i = 0; i < n; i++) {
        double a = phi0 + 2 * M_PI * (double)i / (double)n;
        verts[2 * i] = radius * cos(a);
        verts[2 * i + 1] = radius * sin(a);
    }
}
void or_star(int n, double outer, double inner, double phi0, double* verts) {
    for (int i = 0; i < 2 * n; i++) {
        double a = phi0 + M_PI * (double)i / (double)n;
        double r = (i % 2 == 1) ? inner : outer;
        verts[2 * i] = r * cos(a);
        verts[2 * i + 1] = r * sin(a);
    }
}

/* pointInPolygon, distToSegment, polygonSignedDistance (optics/elements.go:342-395) */
static int point_in_polygon(const double* p, int n, double x, double y) {
    if (n < 3) return 0;
    int inside = 0;
    int j = n - 1;
    for (int i = 0; i < n; i++) {
        double xi = p[2 * i], yi = p[2 * i + 1], xj = p[2 * j], yj = p[2 * j + 1];
        if ((yi > y) != (yj > y)) {
            double xint = (xj - xi) * (y - yi) / (yj - yi) + xi;
            if (x < xint) inside = !inside;
        }
        j = i;
    }
    return inside;
}
static double dist_to_segment(double px, double py, double ax, double ay, double bx, double by) {
    double abx = bx - ax, aby = by - ay;
    double apx = px - ax, apy = py - ay;
    double t = (apx * abx + apy * aby) / (abx * abx + aby * aby);
    if (t < 0) t = 0;
    else if (t > 1) t = 1;
    return or_hypot(px - (ax + t * abx), py - (ay + t * aby));
}
static double polygon_signed_distance(const double* p, int n, double x, double y) {
    if (n < 3) return -1e300;
    double d = INFINITY;
    for (int i = 0; i < n; i++) {
        int k = (i + 1) % n;
        double dd = dist_to_segment(x, y, p[2 * i], p[2 * i + 1], p[2 * k], p[2 * k + 1]);
        if (dd < d) d = dd;
    }
    if (!point_in_polygon(p, n, x, y)) d = -d;
    return d;
}

/* shapeGeom.signedDistance (optics/shape.go:163-198) */
double or_shape_signed_distance(const or_shape* g, double dx, double dy) {
    double cr = cos(g->rotation), sr = sin(g->rotation);
    double u = cr * dx + sr * dy;
    double v = -sr * dx + cr * dy;
    switch (g->kind) {
    case 0: return g->radius - or_hypot(dx, dy);
    case 1: return fmin(g->width / 2 - fabs(u), g->height / 2 - fabs(v));
    case 2: return fmin(g->width / 2 - fabs(u), g->width / 2 - fabs(v));
    case 3: return fmin(g->width / 2 - fabs(u), g->height / 2 - fabs(v));
    case 4: return (1 - or_hypot(u / g->a, v / g->b)) * fmin(g->a, g->b);
    case 9: return polygon_signed_distance(g->verts, g->n_verts, u, v);
    case 6: {
        double r = or_hypot(dx, dy);
        return fmin(r - g->rin, g->rout - r);
    }
    case 7: {
        double d1 = fmin(g->height / 2 - fabs(v), g->width / 2 - fabs(u - g->sep / 2));
        double d2 = fmin(g->height / 2 - fabs(v), g->width / 2 - fabs(u + g->sep / 2));
        return fmax(d1, d2);
    }
    case 8: {
        double hw = g->width / 2, hl = g->length / 2;
        return fmax(fmin(hw - fabs(u), hl - fabs(v)), fmin(hl - fabs(u), hw - fabs(v)));
    }
    case 5: {
        double rn = pow(fabs(u) / g->a, g->order) + pow(fabs(v) / g->b, g->order);
        return (1 - rn) * fmin(g->a, g->b);
    }
    }
    return -1e300;
}

/* apertureEl.Apply (optics/elements.go:288-310): step or smoothStep
 * (optics/elements.go:143-153) of the signed distance */
void or_apply_aperture(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, const or_shape* g,
                       double x0, double y0) {
    double sig = g->edge_sigma;
    for (int j = 0; j < n; j++) {
        double dy = coord(j, n, dx) - y0;
        for (int i = 0; i < n; i++) {
            double ddx = coord(i, n, dx) - x0;
            double d = or_shape_signed_distance(g, ddx, dy);
            or_cplx t = sig > 0 ? cre(0.5 * (1 + erf(d / (M_SQRT2 * sig)))) : cre(d >= 0 ? 1.0 : 0.0);
            size_t idx = (size_t)j * n + i;
            ex[idx] = cmul(ex[idx], t);
            if (polarized) ey[idx] = cmul(ey[idx], t);
        }
    }
}

/* Field.ApplyJones (optics/field.go:109-131) */
void or_apply_jones(or_cplx* ex, or_cplx* ey, int* polarized, int n, or_cplx a, or_cplx b, or_cplx c,
                    or_cplx d) {
    size_t nn = (size_t)n * n;
    if (!*polarized) {
        if (b.re == 0 && b.im == 0 && c.re == 0 && c.im == 0 && d.re == a.re && d.im == a.im) {
            for (size_t i = 0; i < nn; i++) ex[i] = cmul(ex[i], a);
            return;
        }
        *polarized = 1;
        for (size_t i = 0; i < nn; i++) {
            or_cplx e = ex[i];
            ex[i] = cmul(a, e);
            ey[i] = cmul(c, e);
        }
        return;
    }
    for (size_t i = 0; i < nn; i++) {
        or_cplx e = ex[i], f = ey[i];
        ex[i] = cadd(cmul(a, e), cmul(b, f));
        ey[i] = cadd(cmul(c, e), cmul(d, f));
    }
}

/* Field.ScaleAmplitude (optics/field.go:134-143) */
void or_scale_amplitude(or_cplx* ex, or_cplx* ey, int polarized, int n, or_cplx s) {
    size_t nn = (size_t)n * n;
    for (size_t i = 0; i < nn; i++) ex[i] = cmul(ex[i], s);
    if (polarized)
        for (size_t i = 0; i < nn; i++) ey[i] = cmul(ey[i], s);
}

/* Field.ApplyBandlimit (optics/field.go:264-294) */
static void bandlimit_cut(or_cplx* a, int n, double dx, double f0, double sig) {
    or_fft2d(a, n, 0);
    for (int j = 0; j < n; j++) {
        double fy = or_freq(j, n, dx);
        for (int i = 0; i < n; i++) {
            double fx = or_freq(i, n, dx);
            double fr = or_hypot(fx, fy);
            if (fr > f0) {
                double d = (fr - f0) / sig;
                size_t idx = (size_t)j * n + i;
                a[idx] = cmul(a[idx], cre(exp(-d * d)));
            }
        }
    }
    or_fft2d(a, n, 1);
}
void or_bandlimit(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double fraction, double sigma) {
    if (fraction >= 1 || fraction <= 0) return;
    double fnyq = 1 / (2 * dx);
    double f0 = fraction * fnyq;
    double sig = sigma * fnyq;
    if (sig <= 0) sig = fnyq * 0.01;
    bandlimit_cut(ex, n, dx, f0, sig);
    if (polarized) bandlimit_cut(ey, n, dx, f0, sig);
}

/* ---- plane views (server/views.go, optics/polar.go) ---- */
static double norm2c(or_cplx z) { return z.re * z.re + z.im * z.im; }
static double wrap_pi(double a) {
    while (a > M_PI) a -= 2 * M_PI;
    while (a < -M_PI) a += 2 * M_PI;
    return a;
}
static double wrap_half_pi(double a) {
    while (a > M_PI / 2) a -= M_PI;
    while (a < -M_PI / 2) a += M_PI;
    return a;
}
typedef struct { double s0, s1, s2, s3, azimuth, ellipt, degree; } or_stokes;
/* StokesAt and derive (optics/polar.go:23-83) */
static or_stokes stokes_at(or_cplx ex, or_cplx ey) {
    or_stokes st = {0, 0, 0, 0, 0, 0, 0};
    double ax = or_hypot(ex.re, ex.im), ay = or_hypot(ey.re, ey.im);
    double px = atan2(ex.im, ex.re), py = atan2(ey.im, ey.re);
    double delta = wrap_pi(py - px);
    st.s0 = ax * ax + ay * ay;
    st.s1 = ax * ax - ay * ay;
    if (ax > 0 && ay > 0) {
        st.s2 = 2 * ax * ay * cos(delta);
        st.s3 = 2 * ax * ay * sin(delta);
    }
    if (st.s0 <= 0) return st;
    st.azimuth = wrap_half_pi(0.5 * atan2(st.s2, st.s1));
    double r = st.s3 / st.s0;
    if (r < -1) r = -1;
    if (r > 1) r = 1;
    st.ellipt = 0.5 * asin(r);
    st.degree = or_hypot(or_hypot(st.s1, st.s2), st.s3) / st.s0;
    return st;
}
static void put_f32(float* out, size_t i, double v, int nan) {
    if (nan) {
        uint32_t bits = 0x7FC00000u;
        memcpy(out + i, &bits, 4);
    } else {
        out[i] = (float)v;
    }
}
void or_plane_view(const or_cplx* ex, const or_cplx* ey, int n, int view, double mask_cut, float* out) {
    size_t nn = (size_t)n * n;
    or_cplx zero = {0, 0};
    if (view <= 3 || view == 6) {
        for (size_t i = 0; i < nn; i++) {
            double vx = norm2c(ex[i]), vy = ey ? norm2c(ey[i]) : 0;
            double v = view == 2 ? vx : view == 3 ? vy : vx + vy;
            put_f32(out, i, view == 1 ? sqrt(v) : v, 0);
        }
        return;
    }
    if (view == 4 || view == 5) {
        const or_cplx* src = view == 4 ? ex : ey;
        double mc = mask_cut > 0 ? mask_cut : 2e-3;
        double peak = 0;
        for (size_t i = 0; src && i < nn; i++)
            if (norm2c(src[i]) > peak) peak = norm2c(src[i]);
        double thr = mc * peak;
        for (size_t i = 0; i < nn; i++) {
            if (peak <= 0 || !src || norm2c(src[i]) < thr) put_f32(out, i, 0, 1);
            else put_f32(out, i, atan2(src[i].im, src[i].re), 0);
        }
        return;
    }
    double cp = mask_cut > 0 ? mask_cut : 1e-6;
    double max_s0 = 0;
    for (size_t i = 0; i < nn; i++) {
        double s0 = stokes_at(ex[i], ey ? ey[i] : zero).s0;
        if (s0 > max_s0) max_s0 = s0;
    }
    double cut = cp * max_s0;
    for (size_t i = 0; i < nn; i++) {
        or_stokes st = stokes_at(ex[i], ey ? ey[i] : zero);
        if (max_s0 > 0 && st.s0 < cut) {
            put_f32(out, i, 0, 1);
            continue;
        }
        double v = 0;
        switch (view) {
        case 7: v = st.azimuth; break;
        case 8: v = st.ellipt; break;
        case 9: v = st.s0 > 0 ? st.s1 / st.s0 : 0; break;
        case 10: v = st.s0 > 0 ? st.s2 / st.s0 : 0; break;
        case 11: v = st.s0 > 0 ? st.s3 / st.s0 : 0; break;
        case 12: v = st.degree; break;
        }
        put_f32(out, i, v, 0);
    }
}

/* hashNoise01 (optics/elements.go:710-721) */
double or_hash_noise01(int64_t x, int64_t y, int64_t seed) {
    uint64_t h = (uint64_t)seed * 0x9E3779B97F4A7C15ULL ^ (uint64_t)x * 0xC2B2AE3D27D4EB4FULL ^
                 (uint64_t)y * 0x165667B19E3779F9ULL;
    h ^= h >> 30;
    h *= 0xBF58476D1CE4E5B9ULL;
    h ^= h >> 27;
    h *= 0x94D049BB133111EBULL;
    h ^= h >> 31;
    return (double)(h >> 11) / (double)(1ULL << 53);
}

/* diffuserEl.Apply (optics/elements.go:722-783): hash noise on the lab
 * lattice, Gaussian smoothing in the frequency domain, rescale to the
 * requested standard deviation (serial sums), then amp * cexpI(phase). */
void or_apply_diffuser(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double sigma,
                       double corr, double amp, int64_t seed, double x0, double y0) {
    size_t nn = (size_t)n * n;
    double* phase = (double*)malloc(nn * sizeof(double));
    double ddx = dx;
    if (ddx <= 0) ddx = 1;
    for (int j = 0; j < n; j++) {
        int64_t gy = (int64_t)round((coord(j, n, dx) + y0) / ddx);
        for (int i = 0; i < n; i++) {
            int64_t gx = (int64_t)round((coord(i, n, dx) + x0) / ddx);
            phase[(size_t)j * n + i] = or_hash_noise01(gx, gy, seed) * 2 * M_PI - M_PI;
        }
    }
    if (corr > 0) {
        or_cplx* z = (or_cplx*)malloc(nn * sizeof(or_cplx));
        for (size_t i = 0; i < nn; i++) z[i] = cre(phase[i]);
        or_fft2d(z, n, 0);
        for (int j = 0; j < n; j++) {
            double fy = or_freq(j, n, dx);
            for (int i = 0; i < n; i++) {
                double fx = or_freq(i, n, dx);
                double f2 = fx * fx + fy * fy;
                size_t idx = (size_t)j * n + i;
                z[idx] = cmul(z[idx], cre(exp(-2 * M_PI * M_PI * corr * corr * f2)));
            }
        }
        or_fft2d(z, n, 1);
        for (size_t i = 0; i < nn; i++) phase[i] = z[i].re;
        free(z);
    }
    double m = 0, s2 = 0;
    for (size_t i = 0; i < nn; i++) {
        m += phase[i];
        s2 += phase[i] * phase[i];
    }
    m /= (double)nn;
    s2 = s2 / (double)nn - m * m;
    if (s2 > 0) {
        double sc = sigma / sqrt(s2);
        for (size_t i = 0; i < nn; i++) phase[i] = (phase[i] - m) * sc;
    }
    for (size_t i = 0; i < nn; i++) {
        or_cplx t = cmul(cre(amp), cexpI(phase[i]));
        ex[i] = cmul(ex[i], t);
        if (polarized) ey[i] = cmul(ey[i], t);
    }
    free(phase);
}

/* One i.i.d. uniform phase per pixel from the diffuser hash, lattice cell
 * (i - n/2, j - n/2): ex = exp(i * 2pi * u). */
void or_random_phase(or_cplx* ex, int n, int64_t seed) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double u = or_hash_noise01((int64_t)i - n / 2, (int64_t)j - n / 2, seed);
            ex[(size_t)j * n + i] = cexpI(2 * M_PI * u);
        }
}

/* Counter-based standard normal: two SplitMix64 draws, Box-Muller. */
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
double or_normal(uint64_t seed, uint64_t counter) {
    uint64_t h1 = splitmix64(seed ^ splitmix64(2 * counter));
    uint64_t h2 = splitmix64(seed ^ splitmix64(2 * counter + 1));
    double u1 = ((double)(h1 >> 11) + 0.5) / 9007199254740992.0;
    double u2 = (double)(h2 >> 11) / 9007199254740992.0;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* GenerateSchellRealizations, spectrum part (optics/coherence.go:29-53) */
void or_schell_sqrt_spectrum(or_cplx* s, int n, double dx, double sigma_c) {
    for (int j = 0; j < n; j++) {
        double dy = (j <= n / 2) ? (double)j * dx : (double)(j - n) * dx;
        for (int i = 0; i < n; i++) {
            double d = (i <= n / 2) ? (double)i * dx : (double)(i - n) * dx;
            double mu = exp(-(d * d + dy * dy) / (2 * sigma_c * sigma_c));
            s[(size_t)j * n + i] = cre(mu);
        }
    }
    or_fft2d(s, n, 0);
    size_t nn = (size_t)n * n;
    for (size_t i = 0; i < nn; i++) {
        double v = s[i].re;
        if (v < 0) v = 0;
        s[i] = cre(sqrt(v));
    }
}

/* GenerateSchellRealizations, one realisation: the envelope and the noise
 * (optics/coherence.go:54-78);
 * noise counter for pixel i of realisation r: (r*n*n + i)*2 (+1 for imag). */
void or_schell_realization(or_cplx* ex, const or_cplx* sqrt_s, int n, double dx, double w0,
                           uint64_t seed, uint64_t realization) {
    size_t nn = (size_t)n * n;
    const double inv_sqrt2 = 1 / M_SQRT2;
    for (size_t i = 0; i < nn; i++) {
        uint64_t c = (realization * (uint64_t)nn + i) * 2;
        or_cplx v = {or_normal(seed, c) * inv_sqrt2, or_normal(seed, c + 1) * inv_sqrt2};
        ex[i] = cmul(v, sqrt_s[i]);
    }
    or_fft2d(ex, n, 1);
    for (int j = 0; j < n; j++) {
        double y = ((double)j - (double)n / 2) * dx;
        for (int i = 0; i < n; i++) {
            double x = ((double)i - (double)n / 2) * dx;
            double I = exp(-2 * (x * x + y * y) / (w0 * w0));
            size_t idx = (size_t)j * n + i;
            ex[idx] = cmul(ex[idx], cre(sqrt(I)));
        }
    }
}

/* AverageIntensity (optics/coherence.go:85-100): out[i] accumulates
 * Intensity(i) of fields 0, 1, ..., m-1 in that order (ey NULL: scalar
 * fields), then is multiplied by the rounded reciprocal 1/m. Fields are m
 * contiguous n*n planes. m == 0 writes nothing (the reference returns nil). */
void or_average_intensity(const or_cplx* ex, const or_cplx* ey, int m, size_t nn, double* out) {
    if (m <= 0) return;
    for (size_t i = 0; i < nn; i++) out[i] = 0;
    for (int k = 0; k < m; k++) {
        const or_cplx* fx = ex + (size_t)k * nn;
        const or_cplx* fy = ey ? ey + (size_t)k * nn : NULL;
        for (size_t i = 0; i < nn; i++) out[i] += intensity_at(fx, fy, i);
    }
    double inv = 1 / (double)m;
    for (size_t i = 0; i < nn; i++) out[i] *= inv;
}

/* PropagatePartiallyCoherent (optics/coherence.go:105-116): m Schell
 * realisations (GenerateSchellRealizations, optics/coherence.go:22-28 for
 * the argument checks and m <= 0 -> 1), each propagated by Propagate, then
 * AverageIntensity. The realisations are generated and propagated one at a
 * time and accumulated in realisation order, which is the order
 * AverageIntensity sums them in. Returns 2 on invalid arguments. */
int or_propagate_partially_coherent(int n, double dx, double w0, double sigma_c, uint64_t seed, int m, double z,
                                    int method, const or_ctx* ctx, double* out, or_warnings* w) {
    if (!(w0 > 0) || !(sigma_c > 0) || n <= 0) return 2;
    if (m <= 0) m = 1;
    size_t nn = (size_t)n * n;
    or_cplx* sq = (or_cplx*)malloc(sizeof(or_cplx) * nn);
    or_cplx* f = (or_cplx*)malloc(sizeof(or_cplx) * nn);
    or_schell_sqrt_spectrum(sq, n, dx, sigma_c);
    for (size_t i = 0; i < nn; i++) out[i] = 0;
    int rc = 0;
    for (int k = 0; k < m; k++) {
        or_schell_realization(f, sq, n, dx, w0, seed, (uint64_t)k);
        if ((rc = or_propagate(f, NULL, n, dx, z, method, ctx, w, NULL))) break;
        for (size_t i = 0; i < nn; i++) out[i] += intensity_at(f, NULL, i);
    }
    double inv = 1 / (double)m;
    for (size_t i = 0; i < nn; i++) out[i] *= inv;
    free(sq);
    free(f);
    return rc;
}

/* ComputeStats (optics/metrics.go:31-91). */
void or_compute_stats(const or_cplx* ex, const or_cplx* ey, int n, double dx, double wl,
                      double strehl_aperture, double strehl_distance, or_plane_stats* s) {
    memset(s, 0, sizeof(*s));
    double p = 0, px = 0, py = 0;
    s->intensity_min = INFINITY;
    double ph_min = INFINITY, ph_max = -INFINITY;
    double peak = -1.0;
    for (int j = 0; j < n; j++) {
        double y = coord(j, n, dx);
        for (int i = 0; i < n; i++) {
            size_t idx = (size_t)j * n + i;
            double wv = intensity_at(ex, ey, idx);
            p += wv;
            px += wv * coord(i, n, dx);
            py += wv * y;
            if (wv < s->intensity_min) s->intensity_min = wv;
            if (wv > s->intensity_max) s->intensity_max = wv;
            if (wv > peak) peak = wv;
            double ph = atan2(ex[idx].im, ex[idx].re);
            if (ph < ph_min) ph_min = ph;
            if (ph > ph_max) ph_max = ph;
        }
    }
    s->power = p * dx * dx;
    s->peak = peak;
    s->phase_min = ph_min;
    s->phase_max = ph_max;
    if (p <= 0) return;
    s->centroid_x = px / p;
    s->centroid_y = py / p;
    double vx = 0, vy = 0;
    for (int j = 0; j < n; j++) {
        double y = coord(j, n, dx);
        for (int i = 0; i < n; i++) {
            double wv = intensity_at(ex, ey, (size_t)j * n + i);
            double ddx = coord(i, n, dx) - s->centroid_x;
            double ddy = y - s->centroid_y;
            vx += wv * ddx * ddx;
            vy += wv * ddy * ddy;
        }
    }
    s->rms_x = sqrt(vx / p);
    s->rms_y = sqrt(vy / p);
    if (strehl_aperture > 0 && strehl_distance > 0) {
        double ideal = s->power * M_PI * strehl_aperture * strehl_aperture /
                       (wl * wl * strehl_distance * strehl_distance);
        if (ideal > 0) s->strehl = s->peak / ideal;
    }
}

/* ------------------------------------------------------------------ */
/* Phase-mask loss and adjoint gradient (no reference counterpart).     */
/* Forward: U = A exp(i theta), A = F^-1 H F (propASM's transfer);      */
/* I = |U|^2, S = sum I, L = sum (I/S - T)^2.                            */
/* Backward: w = dL/dI = (g - sum g*I/S)/S with g = 2(I/S - T);         */
/* G = A^H (w U) = F^-1 conj(H) F (w U); dL/dtheta = -2 Im(u0 conj(G)). */
/* ------------------------------------------------------------------ */
double or_mask_loss_grad(const double* theta, const double* target, int n, double dx, double z,
                         const or_ctx* ctx, double* grad) {
    size_t nn = (size_t)n * n;
    or_cplx* u = (or_cplx*)malloc(sizeof(or_cplx) * nn);
    for (size_t i = 0; i < nn; i++) u[i] = cexpI(theta[i]);
    int zero, reg;
    double alpha, limit;
    evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
    asm_tf t = make_asm_tf(ctx->wavelength, z, zero, reg, alpha, limit);
    transfer_asm(u, n, dx, &t);
    double S = 0;
    for (size_t i = 0; i < nn; i++) S += u[i].re * u[i].re + u[i].im * u[i].im;
    double L = 0, gsum = 0;
    for (size_t i = 0; i < nn; i++) {
        double I = u[i].re * u[i].re + u[i].im * u[i].im;
        double d = I / S - target[i];
        L += d * d;
        gsum += 2 * d * (I / S);
    }
    if (grad) {
        for (size_t i = 0; i < nn; i++) {
            double I = u[i].re * u[i].re + u[i].im * u[i].im;
            double g = 2 * (I / S - target[i]);
            double wq = (g - gsum) / S;
            u[i].re *= wq;
            u[i].im *= wq;
        }
        /* A^H = F^-1 conj(H) F */
        or_fft2d(u, n, 0);
        for (int j = 0; j < n; j++) {
            double fy = or_freq(j, n, dx);
            for (int i = 0; i < n; i++) {
                size_t idx = (size_t)j * n + i;
                or_cplx h = asm_tf_eval(&t, or_freq(i, n, dx), fy);
                h.im = -h.im;
                u[idx] = cmul(u[idx], h);
            }
        }
        or_fft2d(u, n, 1);
        for (size_t i = 0; i < nn; i++) {
            or_cplx u0 = cexpI(theta[i]);
            /* Im(u0 * conj(G)) = u0.im*G.re - u0.re*G.im */
            grad[i] = -2 * (u0.im * u[i].re - u0.re * u[i].im);
        }
    }
    free(u);
    return L;
}

/* ------------------------------------------------------------------ */
/* Multi-plane phase design through a split-step medium (no reference   */
/* counterpart). Forward: u = exp(i theta); per slice s (PropagateSplit- */
/* Step, optics/medium.go:22-45): u = A u; checkpoint u; u = (u t_s)     */
/* exp(i phi_s) with t_s the medium screen (optics/medium.go:49-65);     */
/* u = A u; A = ASM over dz/2. Loss as or_mask_loss_grad. Backward:      */
/* lambda = dL/dI . U; per slice in reverse: lambda = A^H lambda;        */
/* dL/dphi_s = -2 Im(conj(lambda) v), v = (ck_s t_s) exp(i phi_s);       */
/* lambda = conj(t_s exp(i phi_s)) lambda; lambda = A^H lambda;          */
/* dL/dtheta = -2 Im(conj(lambda) u0).                                   */
/* ------------------------------------------------------------------ */
static void transfer_asm_conj(or_cplx* a, int n, double dx, const asm_tf* t) {
    or_fft2d(a, n, 0);
    for (int j = 0; j < n; j++) {
        double fy = or_freq(j, n, dx);
        for (int i = 0; i < n; i++) {
            size_t idx = (size_t)j * n + i;
            or_cplx h = asm_tf_eval(t, or_freq(i, n, dx), fy);
            h.im = -h.im;
            a[idx] = cmul(a[idx], h);
        }
    }
    or_fft2d(a, n, 1);
}

static or_cplx medium_screen(or_cplx nv, double k, double dz) {
    double ph = k * (nv.re - 1) * dz;
    return cmul(cre(exp(-k * nv.im * dz)), cexpI(ph));
}

double or_bpm_mask_loss_grad(const double* theta, const double* phi, const or_cplx* index, int n_planes,
                             const double* target, int n, double dx, double z, int steps, const or_ctx* ctx,
                             double* grad_theta, double* grad_phi) {
    size_t nn = (size_t)n * n;
    double dz = z / (double)steps;
    double k = 2.0 * M_PI / ctx->wavelength;
    int zero, reg;
    double alpha, limit;
    evanescent_policy(ctx, dz / 2, &zero, &reg, &alpha, &limit);
    asm_tf t = make_asm_tf(ctx->wavelength, dz / 2, zero, reg, alpha, limit);
    or_cplx* u = (or_cplx*)malloc(sizeof(or_cplx) * nn);
    or_cplx* ck = (or_cplx*)malloc(sizeof(or_cplx) * nn * (size_t)steps);
    for (size_t i = 0; i < nn; i++) u[i] = cexpI(theta[i]);
    for (int s = 0; s < steps; s++) {
        const or_cplx* ix = index + (n_planes == 1 ? 0 : (size_t)s * nn);
        transfer_asm(u, n, dx, &t);
        memcpy(ck + (size_t)s * nn, u, sizeof(or_cplx) * nn);
        for (size_t i = 0; i < nn; i++) {
            or_cplx v = cmul(u[i], medium_screen(ix[i], k, dz));
            u[i] = phi ? cmul(v, cexpI(phi[(size_t)s * nn + i])) : v;
        }
        transfer_asm(u, n, dx, &t);
    }
    double S = 0;
    for (size_t i = 0; i < nn; i++) S += u[i].re * u[i].re + u[i].im * u[i].im;
    double L = 0, gsum = 0;
    for (size_t i = 0; i < nn; i++) {
        double I = u[i].re * u[i].re + u[i].im * u[i].im;
        double d = I / S - target[i];
        L += d * d;
        gsum += 2 * d * (I / S);
    }
    if (grad_theta || grad_phi) {
        for (size_t i = 0; i < nn; i++) {
            double I = u[i].re * u[i].re + u[i].im * u[i].im;
            double wq = (2 * (I / S - target[i]) - gsum) / S;
            u[i].re *= wq;
            u[i].im *= wq;
        }
        for (int s = steps - 1; s >= 0; s--) {
            const or_cplx* ix = index + (n_planes == 1 ? 0 : (size_t)s * nn);
            transfer_asm_conj(u, n, dx, &t);
            for (size_t i = 0; i < nn; i++) {
                or_cplx ts = medium_screen(ix[i], k, dz);
                or_cplx e = phi ? cexpI(phi[(size_t)s * nn + i]) : cre(1.0);
                or_cplx v = cmul(cmul(ck[(size_t)s * nn + i], ts), e);
                if (grad_phi) grad_phi[(size_t)s * nn + i] = -2 * (u[i].re * v.im - u[i].im * v.re);
                or_cplx m = cmul(ts, e);
                m.im = -m.im;
                u[i] = cmul(u[i], m);
            }
            transfer_asm_conj(u, n, dx, &t);
        }
        if (grad_theta)
            for (size_t i = 0; i < nn; i++) {
                or_cplx u0 = cexpI(theta[i]);
                grad_theta[i] = -2 * (u[i].re * u0.im - u[i].im * u0.re);
            }
    }
    free(u);
    free(ck);
    return L;
}

/* ------------------------------------------------------------------ */
/* More train elements (optics/elements.go), each in the reference's     */
/* operation order; the field multiplies by t with Go's complex product. */
/* ------------------------------------------------------------------ */

/* nollNM, radialZernike, factorial (optics/elements.go:918-962) */
static const int or_noll[22][2] = {{0, 0}, {0, 0}, {1, 1}, {1, -1}, {2, 0}, {2, -2}, {2, 2}, {3, -1}, {3, 1},
                                   {3, -3}, {3, 3}, {4, 0}, {4, 2}, {4, -2}, {4, 4}, {4, -4}, {5, 1}, {5, -1},
                                   {5, 3}, {5, -3}, {5, 5}, {5, -5}};
static double or_fact(int n) {
    double f = 1;
    for (int i = 2; i <= n; i++) f *= (double)i;
    return f;
}
static double or_radial_zernike(int n, int m, double rho) {
    double s = 0;
    for (int k = 0; k <= (n - m) / 2; k++) {
        double num = or_fact(n - k);
        double den = or_fact(k) * or_fact((n + m) / 2 - k) * or_fact((n - m) / 2 - k);
        s += pow(-1, (double)k) * num / den * pow(rho, (double)(n - 2 * k));
    }
    return s;
}

int or_apply_element(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double wl, const or_element* e) {
    const double k = 2 * M_PI / wl;
    double c = 0, d = 0, cr = 1, sr = 0, thr = 0;
    switch (e->kind) {
        case OR_EL_CYLINDRICAL_LENS: /* optics/elements.go:234-270 */
            if (e->f == 0) return 2;
            c = -k / (2 * e->f);
            cr = cos(e->rotation);
            sr = sin(e->rotation);
            break;
        case OR_EL_APODIZER: /* optics/elements.go:431-456 */
            if (e->waist <= 0) return 2;
            break;
        case OR_EL_GRATING: /* optics/elements.go:464-531 */
            if (e->period <= 0 || e->duty < 0 || e->duty > 1 || e->sub < 0 || e->sub > 4) return 2;
            cr = cos(e->rotation);
            sr = sin(e->rotation);
            thr = cos(M_PI * e->duty);
            break;
        case OR_EL_AXICON: /* optics/elements.go:536-556 */
        case OR_EL_WEDGE:  /* optics/elements.go:600-622 */
            c = -k * (e->index - 1) * tan(e->alpha);
            cr = cos(e->rotation);
            sr = sin(e->rotation);
            break;
        case OR_EL_SPIRAL_PHASE: /* optics/elements.go:566-595 */
            if (e->f != 0) c = -k / (2 * e->f);
            break;
        case OR_EL_ZONE_PLATE: /* optics/elements.go:631-681 */
            if (e->f <= 0 || (e->sub != 0 && e->sub != 1)) return 2;
            break;
        case OR_EL_MIRROR: /* optics/elements.go:792-828 */
            if (e->reflectivity < 0 || e->reflectivity > 1) return 2;
            c = -k * e->curvature;
            d = -2 * k;
            break;
        case OR_EL_CONCAVE_MIRROR: /* optics/elements.go:839-899 */
        case OR_EL_CONVEX_MIRROR:
            if (e->radius <= 0 || e->reflectivity < 0 || e->reflectivity > 1) return 2;
            c = e->kind == OR_EL_CONVEX_MIRROR ? k / e->radius : -k / e->radius;
            break;
        case OR_EL_ZERNIKE: /* optics/elements.go:907-1008 */
            if (e->radius <= 0) return 2;
            break;
        case OR_EL_KERR: /* optics/elements.go:1172-1205 */
            if (e->length < 0) return 2;
            break;
        case OR_EL_SATURABLE_ABSORBER: /* optics/elements.go:1215-1242 */
            if (e->alpha0 < 0 || e->length < 0 || e->isat <= 0) return 2;
            break;
        default:
            return 2;
    }
    if (e->kind == OR_EL_KERR || e->kind == OR_EL_SATURABLE_ABSORBER) {
        const size_t nn = (size_t)n * n;
        for (size_t idx = 0; idx < nn; idx++) {
            /* Field.Intensity (optics/field.go:65-74) */
            double I = ex[idx].re * ex[idx].re + ex[idx].im * ex[idx].im;
            if (polarized) I += ey[idx].re * ey[idx].re + ey[idx].im * ey[idx].im;
            or_cplx t;
            if (e->kind == OR_EL_KERR) {
                if (I == 0) continue;
                t = cexpI(k * e->n2 * I * e->length);
                if (e->tpa != 0) t = cmul(t, cre(exp(-0.5 * e->tpa * I * e->length)));
            } else {
                t = cre(exp(-e->alpha0 * e->length / (2 * (1 + I / e->isat))));
            }
            ex[idx] = cmul(ex[idx], t);
            if (polarized) ey[idx] = cmul(ey[idx], t);
        }
        return 0;
    }
    for (int j = 0; j < n; j++) {
        const double y = coord(j, n, dx);
        for (int i = 0; i < n; i++) {
            const double x = coord(i, n, dx);
            const size_t idx = (size_t)j * n + i;
            or_cplx t = {1, 0};
            int zero = 0, skip = 0;
            switch (e->kind) {
                case OR_EL_CYLINDRICAL_LENS: {
                    const double ddy = y - e->y0, ddx = x - e->x0;
                    if (e->aperture > 0 && ddx * ddx + ddy * ddy > e->aperture * e->aperture) {
                        zero = 1;
                        break;
                    }
                    const double u = cr * ddx + sr * ddy;
                    t = cexpI(c * u * u);
                    break;
                }
                case OR_EL_APODIZER: {
                    const double ddy = y - e->y0, ddx = x - e->x0;
                    t = cre(e->amplitude * exp(-(ddx * ddx + ddy * ddy) / (e->waist * e->waist)));
                    break;
                }
                case OR_EL_GRATING: {
                    const double uu = cr * x + sr * y;
                    const double ph = 2 * M_PI * uu / e->period;
                    switch (e->sub) {
                        case 0: t = cre(0.5 * (1 + e->modulation * cos(ph + e->offset))); break;
                        case 1: t = cre(cos(ph + e->offset) >= thr ? 1.0 : 0.0); break;
                        case 2: t = cexpI(e->modulation / 2 * sin(ph + e->offset)); break;
                        case 3: t = cos(ph + e->offset) >= 0 ? cexpI(e->modulation) : cre(1); break;
                        default: {
                            double frac = fmod(uu / e->period, 1);
                            if (frac < 0) frac += 1;
                            t = cexpI(e->blaze_depth * frac);
                        }
                    }
                    break;
                }
                case OR_EL_AXICON: t = cexpI(c * or_hypot(x, y)); break;
                case OR_EL_SPIRAL_PHASE: {
                    const double ddy = y - e->y0, ddx = x - e->x0;
                    double ph = (double)e->sub * atan2(ddy, ddx);
                    if (c != 0) ph += c * (ddx * ddx + ddy * ddy);
                    t = cexpI(ph);
                    break;
                }
                case OR_EL_WEDGE: t = cexpI(c * (cr * x + sr * y)); break;
                case OR_EL_ZONE_PLATE: {
                    const double r2 = x * x + y * y;
                    if (r2 > e->radius * e->radius) {
                        skip = 1;
                        break;
                    }
                    const double r = sqrt(r2);
                    const int m = (int)floor(2 * (sqrt(e->f * e->f + r * r) - e->f) / wl + 0.5);
                    if (m % 2 != 0) t = e->sub == 0 ? cexpI(M_PI) : cre(0);
                    break;
                }
                case OR_EL_MIRROR:
                    t = cmul(cre(e->reflectivity), cexpI(c * (x * x + y * y) + d * (e->tilt_x * x + e->tilt_y * y)));
                    break;
                case OR_EL_CONCAVE_MIRROR:
                case OR_EL_CONVEX_MIRROR: {
                    const double ddy = y - e->y0, ddx = x - e->x0;
                    const double r2 = ddx * ddx + ddy * ddy;
                    if (e->aperture > 0 && r2 > e->aperture * e->aperture) {
                        zero = 1;
                        break;
                    }
                    t = cmul(cre(e->reflectivity), cexpI(c * r2));
                    break;
                }
                default: { /* Zernike */
                    const double rho = or_hypot(x, y) / e->radius;
                    if (rho > 1) {
                        skip = 1;
                        break;
                    }
                    const double theta = atan2(y, x);
                    double opd = 0;
                    for (int z = 1; z <= 21; z++) {
                        const double cz = e->zernike[z];
                        if (cz == 0) continue;
                        const int nn = or_noll[z][0], mm = or_noll[z][1];
                        const double rad = or_radial_zernike(nn, mm < 0 ? -mm : mm, rho);
                        const double ang = mm >= 0 ? cos((double)mm * theta) : sin((double)(-mm) * theta);
                        if (mm == 0) opd += cz * sqrt((double)(nn + 1)) * rad;
                        else opd += cz * sqrt(2 * (double)(nn + 1)) * rad * ang;
                    }
                    t = cexpI(2 * M_PI * opd);
                }
            }
            if (skip) continue;
            if (zero) {
                ex[idx] = cre(0);
                if (polarized) ey[idx] = cre(0);
                continue;
            }
            ex[idx] = cmul(ex[idx], t);
            if (polarized) ey[idx] = cmul(ey[idx], t);
        }
    }
    return 0;
}
