/*
 * wos_oracle.h — FP64 CPU restatement of the Tetsuhiro-WOS propagation path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the HIP library is compared
 * against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Every function restates the reference's Go code (file:line cited at each
 * definition in wos_oracle.c) in the same left-to-right IEEE operation order,
 * compiled with -ffp-contract=off so no multiply-add is fused: the kz*z phase
 * reaches ~5e5 rad, where one ulp of the argument is ~6e-11 rad.
 *
 * Complex arrays are interleaved complex128 (re, im), row-major idx = y*N + x,
 * exactly the memory layout of a Go []complex128.
 */
#ifndef WOS_ORACLE_H
#define WOS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } or_cplx;

/* Same field order and meaning as wos_ctx in include/wos_hip.h and as the Go
 * Context (optics/field.go:237-254) fields the path reads. */
typedef struct {
    double wavelength;
    int32_t evanescent_zero;      /* Evanescent == "zero" */
    int32_t backward_regularize;  /* BackwardRegularize */
    double evanescent_limit;      /* EvanescentLimit (nepers, 0 = off) */
    double tikhonov_alpha;        /* TikhonovAlpha (<= 0 -> 1e-3) */
} or_ctx;

/* Warning codes, same order as WOS_WARN_* in include/wos_hip.h. */
enum {
    OR_WARN_EVANESCENT_FILTERED = 0,
    OR_WARN_BACKWARD_REGULARIZED = 1,
    OR_WARN_BACKWARD_EVANESCENT = 2,
    OR_WARN_ASM_ALIAS_WRAP = 3,
    OR_WARN_FRESNEL_TF_ALIAS = 4,
    OR_WARN_FRESNEL_IR_ALIAS = 5,
    OR_WARN_FRAUNHOFER_NEARFIELD = 6,
    OR_WARN_COUNT = 8
};

/* Warnings.Add semantics (optics/field.go:322-338): count accumulates, value
 * is replaced when the new value is non-zero. */
typedef struct {
    int32_t count[OR_WARN_COUNT];
    double value[OR_WARN_COUNT];
} or_warnings;

/* Method codes, same as WOS_METHOD_* in include/wos_hip.h. */
enum {
    OR_ASM = 0, OR_ASM_PAD = 1, OR_ASM_SHIFT = 2, OR_ASM_SHIFT_PAD = 3,
    OR_FRESNEL_TF = 4, OR_FRESNEL_IR = 5, OR_FRAUNHOFER = 6
};

/* Per-call diagnostics of one Propagate (last call for loops). */
typedef struct {
    double p_in, p_out;     /* Field.Power before/after (0 if not computed) */
    double fx_c, fy_c;      /* spectral centroid (shift methods) */
    int32_t method_used;    /* after the asm_shift -> asm fallback */
    int32_t pad_;
} or_stats;

void or_set_threads(int p);            /* GOMAXPROCS stand-in for parFor */
int or_get_threads(void);

/* optics/fft.go */
void or_fft1d(or_cplx* a, int n, int inverse);          /* fft1DAnyCPU */
void or_fft1d_radix2(or_cplx* a, int n, int inverse);   /* fft1D (n pow2) */
void or_fft1d_bluestein(or_cplx* a, int n, int inverse);/* fft1DBluestein */
void or_fft2d(or_cplx* a, int n, int inverse);          /* fft2DCPU */

/* optics/field.go */
double or_freq(int i, int n, double dx);
double or_power(const or_cplx* ex, const or_cplx* ey, int n, double dx);
double or_hypot(double p, double q);                    /* Go math.Hypot formula */

/* optics/propagate.go: Propagate for the FFT methods. ey == NULL: scalar. */
int or_propagate(or_cplx* ex, or_cplx* ey, int n, double dx, double z, int method,
                 const or_ctx* ctx, or_warnings* w, or_stats* st);
double or_spectral_radius(const or_cplx* ex, const or_cplx* ey, int n, double dx, double frac);
void or_spectral_centroid(const or_cplx* ex, const or_cplx* ey, int n, double dx,
                          double* fx, double* fy);
int or_propagate1d(or_cplx* a, int n, double dx, double z, double wl,
                   const or_ctx* ctx, or_warnings* w);
/* PropagateUniaxial (optics/birefringent.go:19-61): Ex extraordinary, Ey
 * ordinary (ey NULL: scalar field) */
void or_propagate_uniaxial(or_cplx* ex, or_cplx* ey, int n, double dx, double z, double no, double ne, double wl);
/* PropagateAnisotropic (optics/berreman.go:13-36): Berreman 4x4 per spectral
 * bin for a relative permittivity eps (3x3 complex, row-major); both Ex and
 * Ey take part, as in the reference. */
void or_propagate_anisotropic(or_cplx* ex, or_cplx* ey, int n, double dx, double z, const or_cplx* eps, double wl);
/* PropagateVectorial (optics/vectorial.go:14-80): ex, ey in/out, ez out */
void or_propagate_vectorial(or_cplx* ex, or_cplx* ey, or_cplx* ez, int n, double dx, double z,
                            const or_ctx* ctx);

/* optics/medium.go: PropagateSplitStep with the index sampled on the grid.
 * index holds n_planes planes of n*n complex n(x,y); n_planes == 1 is a
 * z-invariant medium, n_planes == steps gives plane s to slice s (sampled at
 * zc = z*s/steps + dz/2 by the caller, as applyMediumPhase does). */
int or_split_step(or_cplx* ex, or_cplx* ey, int n, double dx, double z, int steps,
                  const or_cplx* index, int n_planes, const or_ctx* ctx, or_warnings* w);

/* optics/polychromatic.go, optics/volume.go */
int or_polychromatic(const or_cplx* ex, const or_cplx* ey, int n, double dx, double z,
                     const double* wavelengths, const double* weights, int nw,
                     int method, const or_ctx* ctx, double* out, or_warnings* w);
int or_volume(const or_cplx* ex, const or_cplx* ey, int n, double dx,
              const double* zs, int nz, int method, const or_ctx* ctx,
              or_cplx* out_ex, or_cplx* out_ey, or_warnings* w);

/* Input constructors (optics/source.go, optics/elements.go, optics/shape.go). */
void or_gaussian(or_cplx* ex, int n, double dx, double w0);       /* unit amplitude */
void or_apply_lens(or_cplx* ex, or_cplx* ey, int n, double dx, double wl, double f);
void or_apply_circle_aperture(or_cplx* ex, or_cplx* ey, int n, double dx, double radius);
double or_hash_noise01(int64_t x, int64_t y, int64_t seed);
void or_random_phase(or_cplx* ex, int n, int64_t seed);           /* exp(i 2pi u(i,j)) */

/* The element chain of a train (optics/simulator.go:397-500). Fields follow
 * Field (optics/field.go:9-40): ey is always allocated; *polarized says
 * whether it takes part. */
void or_apply_lens_at(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double wl, double f,
                      double aperture, double x0, double y0);             /* lensEl.Apply */
/* Aperture outline (optics/shape.go:28-48): the parsed shapeGeom. kind:
 * 0 circle, 1 slit, 2 square, 3 rectangle, 4 ellipse, 5 superellipse,
 * 6 ring, 7 double_slit, 8 cross, 9 polygon (triangle, polygon, star and
 * custom outlines: their vertices). */
typedef struct {
    int32_t kind, n_verts;
    double radius, width, height, a, b, order, rin, rout, inner, length, sep, rotation, edge_sigma;
    const double* verts;    /* n_verts (x, y) pairs */
} or_shape;
double or_shape_signed_distance(const or_shape* g, double dx, double dy);   /* shapeGeom.signedDistance */
void or_regular_polygon(int n, double radius, double phi0, double* verts);   /* regularPolygonVertices */
void or_star(int n, double outer, double inner, double phi0, double* verts); /* starVertices (2n points) */
void or_apply_aperture(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, const or_shape* g,
                       double x0, double y0);                             /* apertureEl.Apply */
void or_apply_diffuser(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double sigma,
                       double corr, double amp, int64_t seed, double x0, double y0);  /* diffuserEl.Apply */
void or_apply_jones(or_cplx* ex, or_cplx* ey, int* polarized, int n, or_cplx a, or_cplx b, or_cplx c,
                    or_cplx d);                                           /* Field.ApplyJones */
void or_scale_amplitude(or_cplx* ex, or_cplx* ey, int polarized, int n, or_cplx s);  /* Field.ScaleAmplitude */
void or_bandlimit(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double fraction,
                  double sigma);                                          /* Field.ApplyBandlimit */

/* servePlaneData's float32 views of one recorded field (server/server.go:581-
 * 660, server/views.go:131-244, optics/polar.go:23-96): 0 total, 1 amplitude,
 * 2 ex, 3 ey, 4 phase_x, 5 phase_y, 6 pol_intensity, 7 pol_azimuth,
 * 8 pol_ellip, 9 pol_s1, 10 pol_s2, 11 pol_s3, 12 pol_degree. ey NULL: scalar
 * field. Masked pixels are the float32 NaN the server writes (0x7FC00000). */
void or_plane_view(const or_cplx* ex, const or_cplx* ey, int n, int view, double mask_cut, float* out);

/* optics/coherence.go with a counter-based normal generator in place of
 * Go's math/rand stream (that stream is not reproducible without Go). */
double or_normal(uint64_t seed, uint64_t counter);
void or_schell_sqrt_spectrum(or_cplx* sqrt_s, int n, double dx, double sigma_c);
void or_schell_realization(or_cplx* ex, const or_cplx* sqrt_s, int n, double dx,
                           double w0, uint64_t seed, uint64_t realization);
void or_average_intensity(const or_cplx* ex, const or_cplx* ey, int m, size_t nn, double* out); /* AverageIntensity */
int or_propagate_partially_coherent(int n, double dx, double w0, double sigma_c, uint64_t seed, int m, double z,
                                    int method, const or_ctx* ctx, double* out,
                                    or_warnings* w);                    /* PropagatePartiallyCoherent */

/* optics/metrics.go:31-91: PlaneStats (field order of the Go struct). */
typedef struct {
    double power, peak, centroid_x, centroid_y, rms_x, rms_y, strehl;
    double intensity_min, intensity_max, phase_min, phase_max;
} or_plane_stats;
void or_compute_stats(const or_cplx* ex, const or_cplx* ey, int n, double dx, double wl,
                      double strehl_aperture, double strehl_distance, or_plane_stats* s);

/* Phase-mask design (no reference counterpart): forward ASM of exp(i theta),
 * loss L = sum (I/sum I - T)^2 and its gradient dL/dtheta by the adjoint
 * (conjugate transfer function). theta and target are n*n; grad may be NULL. */
double or_mask_loss_grad(const double* theta, const double* target, int n, double dx,
                         double z, const or_ctx* ctx, double* grad);

/* Multi-plane phase design through a split-step medium (no reference
 * counterpart): exp(i theta) through `steps` slices of PropagateSplitStep
 * whose screen s is multiplied by exp(i phi_s) (phi may be NULL: no design
 * phases); loss as or_mask_loss_grad; gradients by reversed slices with
 * conjugate operators and the checkpointed fields before each screen.
 * grad_theta (n*n) and grad_phi (steps*n*n) may be NULL. */
double or_bpm_mask_loss_grad(const double* theta, const double* phi, const or_cplx* index, int n_planes,
                             const double* target, int n, double dx, double z, int steps, const or_ctx* ctx,
                             double* grad_theta, double* grad_phi);

/* More train elements (optics/elements.go:156-178): the parsed parameters of
 * one element; kind OR_EL_*, sub = grating kind (0 amplitude_sin,
 * 1 amplitude_binary, 2 phase_sin, 3 phase_binary, 4 blazed), zone-plate
 * kind (0 phase, 1 amplitude) or spiral charge. */
#define OR_EL_CYLINDRICAL_LENS 0
#define OR_EL_APODIZER 1
#define OR_EL_GRATING 2
#define OR_EL_AXICON 3
#define OR_EL_SPIRAL_PHASE 4
#define OR_EL_WEDGE 5
#define OR_EL_ZONE_PLATE 6
#define OR_EL_MIRROR 7
#define OR_EL_CONCAVE_MIRROR 8
#define OR_EL_CONVEX_MIRROR 9
#define OR_EL_ZERNIKE 10
#define OR_EL_KERR 11
#define OR_EL_SATURABLE_ABSORBER 12
typedef struct {
    int32_t kind, sub;
    double f, aperture, x0, y0, rotation;
    double waist, amplitude;
    double alpha, index;
    double period, modulation, duty, blaze_depth, offset;
    double radius;
    double reflectivity, curvature, tilt_x, tilt_y;
    double zernike[22];
    double n2, tpa, length, alpha0, isat;
} or_element;
/* Apply one element at wavelength wl; returns 0, or 2 for parameters the
 * reference's constructor rejects. */
int or_apply_element(or_cplx* ex, or_cplx* ey, int polarized, int n, double dx, double wl, const or_element* e);

#ifdef __cplusplus
}
#endif
#endif
