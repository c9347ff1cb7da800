/*
 * wos_hip.h — C ABI of libwos_hip, the MI355X (gfx950) propagation engine
 * behind Tetsuhiro-WOS's fft2D/fft1DAny dispatch and its Propagate /
 * PropagateSplitStep / PropagatePolychromatic / Propagate3D operators.
 *
 * Conventions (all entry points):
 *  - Plain C types only; complex arrays are interleaved (re, im), row-major
 *    idx = y*N + x: complex128 = double[2], complex64 = float[2]. A Go
 *    []complex128 is passed as (*C.double)(unsafe.Pointer(&a[0])).
 *  - Host pointers are read and written synchronously; nothing is retained
 *    after return (cgo pointer rules). Device-resident batches use the opaque
 *    wos_batch handle.
 *  - Every call returns WOS_OK (0) or a WOS_ERR_* code; wos_hip_last_error()
 *    returns the calling thread's last message. No C++ exception crosses the
 *    ABI. The library never falls back to a CPU path: the Go caller keeps its
 *    own pure-Go fallback (optics/fft.go:136-141, optics/fft.go:210-215).
 *  - Forward transforms are unnormalised with exp(-2 pi i .); inverse
 *    transforms are scaled by 1/n (1-D) or 1/n^2 (2-D), as fft1DAny / fft2D.
 *  - Calls are thread-safe: one device mutex serialises work on the
 *    library's stream, as gpuExecMu does (optics/gpu.go:97-100).
 */
#ifndef WOS_HIP_H
#define WOS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define WOS_OK 0
#define WOS_ERR_NO_DEVICE 1   /* no HIP device / library not initialised */
#define WOS_ERR_INVALID 2     /* bad argument (size, method, pointer) */
#define WOS_ERR_HIP 3         /* a HIP runtime call failed */
#define WOS_ERR_UNSUPPORTED 4 /* size/method outside what the engine builds */
#define WOS_ERR_NOMEM 5       /* device allocation failed */

/* ---- precision of the bulk FFT passes ---- */
#define WOS_FP64 0 /* complex128 everywhere (parity path) */
#define WOS_FP32 1 /* complex64 bulk passes; FP64 twiddles, chirps, kz*z, sums */

/* ---- ints in a launch shape (autotune, set / get / load tuning) ---- */
#define WOS_SHAPE_LEN 11

/* ---- propagation methods (optics/propagate.go:40-50) ---- */
#define WOS_METHOD_ASM 0
#define WOS_METHOD_ASM_PAD 1
#define WOS_METHOD_ASM_SHIFT 2
#define WOS_METHOD_ASM_SHIFT_PAD 3
#define WOS_METHOD_FRESNEL_TF 4
#define WOS_METHOD_FRESNEL_IR 5
/* propFraunhofer (optics/propagate.go:603-662), even n: the output grid
 * spacing is wl |z| / (n dx), which the caller sets as the new Field.DX */
#define WOS_METHOD_FRAUNHOFER 6

/* ---- warning codes (optics/propagate.go:203-264, 515-518, 586-589) ---- */
#define WOS_WARN_EVANESCENT_FILTERED 0
#define WOS_WARN_BACKWARD_REGULARIZED 1
#define WOS_WARN_BACKWARD_EVANESCENT 2
#define WOS_WARN_ASM_ALIAS_WRAP 3
#define WOS_WARN_FRESNEL_TF_ALIAS 4
#define WOS_WARN_FRESNEL_IR_ALIAS 5
#define WOS_WARN_FRAUNHOFER_NEARFIELD 6
#define WOS_WARN_SLOTS 8

/* The Context fields the path reads (optics/field.go:237-254). */
typedef struct wos_ctx {
    double wavelength;           /* Wavelength, m */
    int32_t evanescent_zero;     /* Evanescent == "zero" */
    int32_t backward_regularize; /* BackwardRegularize */
    double evanescent_limit;     /* EvanescentLimit, nepers (0 = off) */
    double tikhonov_alpha;       /* TikhonovAlpha (<= 0 -> 1e-3) */
} wos_ctx;

/* Warnings.Add accumulation (optics/field.go:322-338): count[code] is the
 * number of Add calls, value[code] the last non-zero value. The Go shim
 * replays them into ctx.Warnings with the reference's messages. */
typedef struct wos_warnings {
    int32_t count[WOS_WARN_SLOTS];
    double value[WOS_WARN_SLOTS];
} wos_warnings;

/* Diagnostics of one Propagate call. */
typedef struct wos_prop_stats {
    double p_in, p_out;   /* Field.Power before / after (FP64, compensated) */
    double fx_c, fy_c;    /* spectral centroid (asm_shift*, else 0) */
    int32_t method_used;  /* method actually run (asm_shift* fall back) */
    int32_t reserved;
} wos_prop_stats;

/* PlaneStats of ComputeStats (optics/metrics.go:8-21), same field order. */
typedef struct wos_plane_stats {
    double power, peak, centroid_x, centroid_y, rms_x, rms_y, strehl;
    double intensity_min, intensity_max, phase_min, phase_max;
} wos_plane_stats;

/* ---- device / library ---- */
/* replaces wos_device_count (optics/gpu.go:28-33) */
int wos_hip_device_count(void);
/* replaces wos_device_name (optics/gpu.go:35-41); writes a NUL-terminated name */
int wos_hip_device_name(char* buf, int len);
/* creates the stream and twiddle cache on device 0; idempotent (gpuInit, optics/gpu.go:106-114) */
int wos_hip_init(void);
/* thread-local message of the last failing call on this thread */
const char* wos_hip_last_error(void);
/* library build string */
const char* wos_hip_version(void);
/* device time (HIP events) of the last split-step pipeline, ms */
double wos_hip_last_device_ms(void);
/* frees every cached plan/buffer (device memory returns to the runtime) */
int wos_hip_release_cache(void);
/* CheckGridMemory (optics/memory.go:74-81) for device memory: WOS_OK when
 * `units` resident propagations of n x n by `method` in `precision` fit the
 * budget (75% of the device's memory, memoryBudgetBytes' share above 8 GiB),
 * else WOS_ERR_NOMEM with a message like the reference's ("grid NxN would
 * need X of device memory but only Y is available ..."); need and budget
 * (nullable) receive the estimate and the budget in bytes. Batch and
 * pipeline creation apply the same check before they allocate.
 *
 * Failure injection (all entry points): the environment variable
 * WOS_HIP_FAIL, a comma-separated list of entry-point names or "all", makes
 * each listed entry point return WOS_ERR_HIP ("injected failure") before it
 * reads its arguments or touches the device, so a caller's fallback path can
 * be exercised without a broken GPU; with a listed wos_hip_device_count
 * (or "all") the library reports no device. */
int wos_hip_check_grid_memory(int n, int method, int precision, int units, uint64_t* need, uint64_t* budget);

/* ---- drop-in transforms: gpuFFT2D / gpuFFT1D (optics/gpu.go:275-304) ---- */
/* Lengths: powers of two up to 8192 run one workgroup per transform
 * (Stockham radix 16/8); other lengths whose Bluestein size (next power of
 * two >= 2n - 1) is at most 8192, i.e. n <= 4096, run mixed-radix Stockham
 * when n = 2^a 3^b 5^c 7^d and Bluestein otherwise. Beyond that, powers of
 * two up to 2^26 run a four-step factorisation N = N1 N2 over the same
 * kernels (column transforms, an exact twiddle, row transforms, a
 * transpose), and other lengths up to 2^25 run the reference's chirp-z
 * chain (optics/fft.go:159-204) on four-step transforms; fft2d runs such
 * lengths as row transforms, a transpose, row transforms, a transpose
 * back. All with the reference's unnormalised forward and 1/n inverse
 * (fft2d: 1/n per pass). Longer lengths, and grids beyond the device
 * memory budget, return WOS_ERR_UNSUPPORTED before reading the array (the
 * Go caller then keeps its CPU path, as optics/fft.go:136-141 does on any
 * GPU error). */
/* In-place 2-D transform of an n*n complex128 host array (fft2D, optics/fft.go:209-216). */
int wos_hip_fft2d_z(double* a, int n, int inverse);
/* In-place 1-D transform of a length-n complex128 host array (fft1DAny, optics/fft.go:135-142). */
int wos_hip_fft1d_z(double* a, int n, int inverse);
/* complex64 variants (no reference counterpart: the FP32 precision mode). */
int wos_hip_fft2d_c(float* a, int n, int inverse);
int wos_hip_fft1d_c(float* a, int n, int inverse);

/* ---- operators ---- */
/* Propagate (optics/propagate.go:65-102) for asm, asm_pad, asm_shift,
 * asm_shift_pad, fresnel_tf, fresnel_ir. ex/ey are n*n complex128 host arrays
 * updated in place; ey == NULL for a scalar field (Polarized == false).
 * precision selects the bulk-pass type (results are returned in complex128
 * either way). w and st may be NULL. z == 0 is a no-op. */
int wos_hip_propagate(double* ex, double* ey, int n, double dx, double z, int method,
                      const wos_ctx* ctx, int precision, wos_warnings* w, wos_prop_stats* st);

/* PropagateUniaxial (optics/birefringent.go:19-61): a uniaxial crystal of
 * thickness z, optic axis along x. Ex takes the extraordinary transfer
 * function, Ey (NULL: scalar field, left alone) the ordinary one, at
 * ctx->wavelength. no, ne <= 0 is an error, as in the reference. */
int wos_hip_propagate_uniaxial(double* ex, double* ey, int n, double dx, double z, double no, double ne,
                               const wos_ctx* ctx, int precision);

/* PropagateAnisotropic (optics/berreman.go:13-36): a homogeneous anisotropic
 * medium of thickness z >= 0 with relative permittivity eps (3x3 complex,
 * row-major, as 18 doubles re, im), by the Berreman 4x4 method per spectral
 * bin. Ex and Ey (both n*n complex128) always take part, as in the
 * reference; the biaxial element (optics/elements.go:1129-1155) passes
 * diag(nx^2, ny^2, nz^2). Degenerate modes (an isotropic eps) give NaN there,
 * as in the reference. */
int wos_hip_propagate_anisotropic(double* ex, double* ey, int n, double dx, double z, const double* eps,
                                  const wos_ctx* ctx, int precision);

/* PropagateVectorial (optics/vectorial.go:14-80), Propagate's "vectorial"
 * method: Ex and Ey (both complex128 n*n, Ey zeros for a field that had none)
 * propagate in place under exp(i kz z) and the evanescent policy; ez receives
 * the longitudinal component -(kx Ex + ky Ey)/kz of the propagated field.
 * Emits no warnings, as the reference. */
int wos_hip_propagate_vectorial(double* ex, double* ey, double* ez, int n, double dx, double z, const wos_ctx* ctx,
                                int precision);

/* Propagate1D (optics/propagate.go:718-738): 1-D ASM of a length-n complex128 array. */
int wos_hip_propagate1d(double* a, int n, double dx, double z, double wl, const wos_ctx* ctx,
                        int precision, wos_warnings* w);

/* PropagateSplitStep (optics/medium.go:22-45). The Go IndexFunc closure
 * cannot run on the device, so the caller samples it on the grid in the
 * reference's order: n_planes == 1 for a z-invariant medium n(x, y), or
 * n_planes == steps with plane s sampled at zc = z*s/steps + dz/2
 * (optics/medium.go:35,55). index is n_planes*n*n complex128.
 * On power-of-two grids the two ASM half steps between screens run as one
 * fused pass (H(dz/2)^2) and each half step's evanescent_filtered check is
 * taken from spectral power sums; other grids run the half steps one by one. */
int wos_hip_split_step(double* ex, double* ey, int n, double dx, double z, int steps,
                       const double* index, int n_planes, const wos_ctx* ctx, int precision,
                       wos_warnings* w);

/* Launch shape of the split-step slices on an n-point power-of-two grid:
 * column tile width (radix-16 columns: the default, or narrower from
 * n = 512; radix-8 columns: 2 or more), rows per row workgroup (0: default)
 * and the FFT stage schedules of the row and column passes (radix / col_radix:
 * 16, or 8 for 256..8192 points; 0 means 16); tile_w = 0 restores the
 * built-in choice. */
int wos_hip_split_step_set_shape(int precision, int n, int tile_w, int rows, int radix, int col_radix);

/* PropagatePolychromatic (optics/polychromatic.go:18-43): out (n*n float64)
 * = sum_s weight_s * |U_s|^2; the input field is not modified. */
int wos_hip_polychromatic(const double* ex, const double* ey, int n, double dx, double z,
                          const double* wavelengths, const double* weights, int n_samples,
                          int method, const wos_ctx* ctx, int precision, double* out,
                          wos_warnings* w);

/* Propagate3D (optics/volume.go:8-18): out_ex/out_ey hold nz planes of n*n
 * complex128 (out_ey may be NULL when ey is NULL). */
int wos_hip_volume(const double* ex, const double* ey, int n, double dx, const double* zs, int nz,
                   int method, const wos_ctx* ctx, int precision, double* out_ex, double* out_ey,
                   wos_warnings* w);

/* ---- partially coherent light (optics/coherence.go) ----
 * Go's math/rand NormFloat64 stream cannot be reproduced outside Go, so the
 * Gaussian-Schell realisations take their complex Gaussian noise from a
 * counter-based generator (SplitMix64 + Box-Muller; realisation r, pixel i
 * draws counters 2(r*n*n + i) and 2(r*n*n + i) + 1). Everything else follows
 * the reference: sqrt(max(0, Re FFT(mu))) in FP64, the inverse FFT, the
 * envelope sqrt(exp(-2 r^2 / w0^2)). A width or coherence <= 0 returns
 * WOS_ERR_INVALID ("schell: width and coherence must be > 0"); m <= 0 is
 * taken as 1, as the reference does. */

/* GenerateSchellRealizations (optics/coherence.go:22-81): m realisations of
 * n x n written to out as m contiguous complex128 planes (FP64 throughout). */
int wos_hip_schell_realizations(int n, double dx, double w0, double sigma_c, int64_t seed, int m, double* out);

/* AverageIntensity (optics/coherence.go:85-100) over m fields given as m
 * contiguous complex128 planes of n x n (ey NULL: scalar fields, else the Ey
 * planes in the same layout): out (n*n float64) sums Intensity(i) in field
 * order, then multiplies by the rounded reciprocal 1/m. m == 0 writes
 * nothing (the reference returns nil). */
int wos_hip_average_intensity(const double* ex, const double* ey, int m, int n, double* out);

/* PropagatePartiallyCoherent (optics/coherence.go:105-116): m Schell
 * realisations generated on the device, each propagated by `method`
 * (Propagate, any method, warnings accumulated as the reference's ctx
 * collects them), and their AverageIntensity in out (n*n float64). The
 * realisations never leave the device. */
int wos_hip_propagate_partially_coherent(int n, double dx, double w0, double sigma_c, int64_t seed, int m, double z,
                                         int method, const wos_ctx* ctx, int precision, double* out,
                                         wos_warnings* w);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): Propagate by asm or asm_pad (M = n or 2n a power of two
 * up to 8192) of a device-resident field, in place: `field` holds ncomp
 * (1 or 2) components of n x n, contiguous, in complex128 (WOS_FP64) or
 * complex64 (WOS_FP32). The kernels are those of the resident-field
 * Propagate (wos_hip_field_propagate gives the same values); every buffer
 * besides the field lives in the library's static per-device scratch of
 * 72 MiB (enough for one 2048^2 FP64 or two 2048^2 FP32 components); larger
 * calls return WOS_ERR_UNSUPPORTED with the size needed. No warnings or
 * power sums are produced. `stream` is a hipStream_t (NULL: the library's
 * stream); the call only enqueues work. Call wos_hip_init once beforehand.
 * Calls that overlap on one GPU share its scratch and race. */
int wos_hip_propagate_device_noalloc(void* field, int n, int ncomp, double dx, double z, int method,
                                     const wos_ctx* ctx, int precision, void* stream);

/* ---- device-resident batches (no reference counterpart: batching is new) ----
 * A batch holds n_units independent fields of n x n on the device in
 * `precision`, each propagated by `method` with its own wavelength and
 * distance: Jones components, polychromatic samples, volume planes or
 * realisations. One wos_hip_batch_run is one propagation of every unit. */
typedef struct wos_batch_s* wos_batch;

int wos_hip_batch_create(wos_batch* out, int n, double dx, int method, int precision, int n_units,
                         const double* wavelengths, const double* zs, const wos_ctx* ctx);
int wos_hip_batch_destroy(wos_batch b);
/* copy one unit from / to a complex128 host array of n*n */
int wos_hip_batch_upload(wos_batch b, int unit, const double* ex);
int wos_hip_batch_download(wos_batch b, int unit, double* ex);
/* fill every unit with a Gaussian-Schell realisation generated on the device
 * (optics/coherence.go:22-80 with a counter-based normal generator; unit u
 * is realisation u of `seed`) */
int wos_hip_batch_fill_schell(wos_batch b, double w0, double sigma_c, uint64_t seed);
/* Autotune the launch shape of this batch. Shapes are combinations of: the
 * column-tile width of the intermediate layout (tile_w); for the row passes
 * the FFT stage schedule (radix: 16 = radix-16 stages with 16 values per
 * thread; 8 = radix-8 stages with 8 values per thread and twice the
 * threads, built for power-of-two grids of 256..8192), the rows per
 * workgroup, the row groups of the pipelined kernels (0: one group, 2, 4)
 * and, radix-16, split LDS exchanges (real and imaginary parts in two
 * rounds through half the LDS; pipelined kernels of 2 or 4 rows then also
 * stage their twiddles in LDS), the forward and the inverse row pass each
 * taking its own rows, groups and split; for the column pass
 * its own stage schedule (col_radix: 16, 8, or 32 = radix-16 stages with 32
 * values per thread and split exchanges, 1024..4096 points, 4 or 8 columns
 * per workgroup), the columns per workgroup (col_w:
 * tile_w, tile_w/2 or tile_w/4) and, radix-16, split LDS exchanges
 * (col_split). For each tile_w the row and column shapes
 * are searched apart, each kernel timed over `iters` launches with HIP
 * events on the batch stream, and the tile_w with the smallest sum wins (a
 * captured graph is dropped). The launches propagate the batch's fields:
 * refill them afterwards. shape (may be NULL, else WOS_SHAPE_LEN ints)
 * receives {tile_w, col_w, rows, groups, split, radix, col_radix, col_split,
 * inv_rows, inv_groups, inv_split} (rows, groups, split: the forward row
 * pass; inv_*: the inverse row pass); best_ms
 * (may be NULL) the time of one whole run at that shape. On mixed-radix
 * grids (2^a 3^b 5^c 7^d sides) the choices are the layout tile width
 * (tile_w: 16, 8, 4, 2, 1 dividing the side) and the columns per column
 * workgroup (col_w: a built width dividing tile_w), each pair timed as a
 * whole run; Bluestein grids keep their fixed shape. */
int wos_hip_batch_autotune(wos_batch b, int iters, int* shape, double* best_ms);
/* Set the launch shape directly (e.g. from a cached autotune table);
 * radix / col_radix 0 mean 16; inv_rows < 0: the inverse row pass takes
 * rows, groups and split (inv_groups and inv_split are then ignored).
 * Mixed-radix grids take tile_w and col_w only (as autotune searches them)
 * and ignore the other fields. */
int wos_hip_batch_set_tuning(wos_batch b, int tile_w, int col_w, int rows, int groups, int split, int radix,
                             int col_radix, int col_split, int inv_rows, int inv_groups, int inv_split);
/* Persistent transfer (on != 0): each run is one launch of a persistent
 * kernel that takes the forward rows, columns and inverse rows of every unit
 * from per-XCD device work queues (plus a counter reset and a completion
 * check), instead of three launches. on == 1 queues each XCD's passes one
 * after another, on == 2 interleaves a round's three passes task by task. Built for asm batches of side 512,
 * 1024, 2048 (FP32) and 1024, 2048 (FP64) with its own fixed launch shape
 * (the tuning of wos_hip_batch_set_tuning / autotune applies to the
 * three-launch form only); requires workgroups placed on the 8 XCDs in turn
 * (checked once per process), else WOS_ERR_UNSUPPORTED. A captured graph is
 * dropped. wos_hip_batch_sync and wos_hip_batch_time report WOS_ERR_HIP if a
 * persistent run since enabling left work undone (a dependency wait gave
 * up), and the error stays until the mode is set again. on == 0 returns to
 * three launches. */
int wos_hip_batch_set_persistent(wos_batch b, int on);
/* capture one run into a hipGraph; later runs replay it */
int wos_hip_batch_capture(wos_batch b);
/* enqueue `iters` runs on the batch's stream (asynchronous) */
int wos_hip_batch_run(wos_batch b, int iters);
int wos_hip_batch_sync(wos_batch b);
/* device pointer of unit 0 (units are contiguous, n*n elements apart) */
void* wos_hip_batch_device_ptr(wos_batch b);
/* HIP-event timing on the batch's stream: ms_run = mean time of one run over
 * `iters` runs after `warmup`; kernel_ms[0..2] = mean duration of the row
 * forward, column (forward -> TF -> inverse) and row inverse launches, each
 * timed back to back over `iters` launches. kernel_ms may be NULL. */
int wos_hip_batch_time(wos_batch b, int warmup, int iters, double* ms_run, double* kernel_ms);

/* ---- heterogeneous pipeline (no reference counterpart: batching is new) --
 * Stages of batched asm / asm_pad propagations, each with its own grid size
 * and unit count, run back to back on one stream and captured into one
 * hipGraph. Stage inputs are kept (outputs go to separate buffers), so each
 * run recomputes the same outputs. A stage with src_units == 1 propagates one
 * source field for every unit (PropagatePolychromatic, optics/polychromatic.go:
 * 18-43, or Propagate3D, optics/volume.go:8-18); with src_units == units each
 * unit has its own source (independent Propagate calls). Units with z == 0
 * return their source (optics/propagate.go:66-68). The pipeline emits no
 * warnings. */
typedef struct wos_pipeline_s* wos_pipeline;

int wos_hip_pipeline_create(wos_pipeline* out, int precision);
int wos_hip_pipeline_destroy(wos_pipeline p);
/* add a stage of `units` propagations (unit u at wavelengths[u] over zs[u]);
 * *stage receives its index */
int wos_hip_pipeline_add(wos_pipeline p, int n, double dx, int method, int units, int src_units,
                         const double* wavelengths, const double* zs, const wos_ctx* ctx, int* stage);
/* the stage also writes sum_u weights[u] |U_u|^2 (n*n float64) in unit order,
 * the incoherent sum of PropagatePolychromatic */
int wos_hip_pipeline_set_intensity(wos_pipeline p, int stage, const double* weights);
/* the stage writes AverageIntensity of its outputs (optics/coherence.go:85-100):
 * sum_u |U_u|^2 in unit order times 1/units; with units Schell realisations
 * (wos_hip_pipeline_fill_schell) of one wavelength and distance the stage is
 * PropagatePartiallyCoherent (optics/coherence.go:105-116) for asm / asm_pad */
int wos_hip_pipeline_set_average(wos_pipeline p, int stage);
/* copy source src_index of a stage from a complex128 host array of n*n */
int wos_hip_pipeline_upload(wos_pipeline p, int stage, int src_index, const double* ex);
/* fill a stage's sources with Gaussian-Schell realisations first_realisation,
 * first_realisation+1, ... of `seed` generated on the device */
int wos_hip_pipeline_fill_schell(wos_pipeline p, int stage, double w0, double sigma_c, uint64_t seed,
                                 int first_realisation);
/* per-stage autotune as wos_hip_batch_autotune; stage_ms (may be NULL) gets
 * each stage's best time per run */
int wos_hip_pipeline_autotune(wos_pipeline p, int iters, double* stage_ms);
/* capture every stage into one hipGraph; later runs replay it */
int wos_hip_pipeline_capture(wos_pipeline p);
/* enqueue `iters` runs on the pipeline's stream (asynchronous) */
int wos_hip_pipeline_run(wos_pipeline p, int iters);
int wos_hip_pipeline_sync(wos_pipeline p);
/* output field of one unit of a stage as complex128 n*n */
int wos_hip_pipeline_download(wos_pipeline p, int stage, int unit, double* ex);
/* source field src_index of a stage as complex128 n*n */
int wos_hip_pipeline_download_source(wos_pipeline p, int stage, int src_index, double* ex);
/* intensity output of a stage (n*n float64) */
int wos_hip_pipeline_intensity(wos_pipeline p, int stage, double* out);
/* HIP-event timing: ms_run = mean time of one run; stage_ms[s] (may be NULL)
 * = mean time of stage s enqueued alone */
int wos_hip_pipeline_time(wos_pipeline p, int warmup, int iters, double* ms_run, double* stage_ms);

/* ---- plane statistics ----
 * ComputeStats (optics/metrics.go:31-91): power, peak, intensity range,
 * centroid, RMS widths, Strehl (when aperture and distance are > 0) and the
 * phase range of Ex. Sums are compensated FP64 (deterministic). */
int wos_hip_compute_stats(const double* ex, const double* ey, int n, double dx, double wl, double strehl_aperture,
                          double strehl_distance, wos_plane_stats* out);
/* the same for a device-resident field: a batch unit, a pipeline output */
int wos_hip_batch_stats(wos_batch b, int unit, double wl, double strehl_aperture, double strehl_distance,
                        wos_plane_stats* out);
int wos_hip_pipeline_stats(wos_pipeline p, int stage, int unit, double wl, double strehl_aperture,
                           double strehl_distance, wos_plane_stats* out);

/* ---- device-resident fields: the element chain of a train ----------------
 * runTrain (optics/simulator.go:397-500) applies elements and propagations to
 * one Field in turn. A wos_field keeps that Field (optics/field.go:9-40) in
 * HBM between the steps: Ex and Ey (Ey allocated also for a scalar field, as
 * Field does; `polarized` says whether it takes part), so a train makes no
 * host round trip between elements. Elements compute in FP64 in the
 * reference's operation order for both precisions (an FP32 field is rounded
 * once per element). All calls are synchronous. */
typedef struct wos_field_s* wos_field;

int wos_hip_field_create(wos_field* out, int n, double dx, int polarized, int precision);
int wos_hip_field_destroy(wos_field f);
/* Field.Clone (optics/field.go:42-52) */
int wos_hip_field_clone(wos_field f, wos_field* out);
/* complex128 host arrays of n*n; ey NULL: zeros */
int wos_hip_field_upload(wos_field f, const double* ex, const double* ey);
/* either pointer may be NULL */
int wos_hip_field_download(wos_field f, double* ex, double* ey);
/* Field.Polarized: 1 once a Jones element has made a scalar field a vector field */
int wos_hip_field_polarized(wos_field f);
/* device pointer of Ex (Ey follows n*n elements later) */
void* wos_hip_field_device_ptr(wos_field f);
/* Field.DX: the grid spacing, changed by a fraunhofer propagation */
double wos_hip_field_dx(wos_field f);
/* uniaxialEl.Apply = PropagateUniaxial (optics/elements.go:1102-1104) on the resident field */
int wos_hip_field_uniaxial(wos_field f, double z, double no, double ne, double wavelength);
/* biaxialEl.Apply = PropagateAnisotropic on the resident field (Ey takes
 * part also for a scalar field, whose Polarized flag is left as is) */
int wos_hip_field_anisotropic(wos_field f, double z, const double* eps, double wavelength);
/* Propagate (optics/propagate.go:65-102) on the resident field, Ey only if
 * polarized; warnings and stats as wos_hip_propagate */
int wos_hip_field_propagate(wos_field f, double z, int method, const wos_ctx* ctx, wos_warnings* w,
                            wos_prop_stats* st);
/* lensEl.Apply (optics/elements.go:195-220): phase (-k/(2f)) r^2 about
 * (x0, y0), zero outside `aperture` (radius, 0: none) */
int wos_hip_field_lens(wos_field f, double wavelength, double focal, double aperture, double x0, double y0);

/* Aperture outline: the parsed shapeGeom (optics/shape.go:28-48, 57-160).
 * Triangle, polygon, star and custom outlines are WOS_SHAPE_POLYGON with their
 * vertices (regularPolygonVertices / starVertices / parseVertices,
 * optics/elements.go:316-338, 397-425, formed by the caller). */
#define WOS_SHAPE_CIRCLE 0
#define WOS_SHAPE_SLIT 1
#define WOS_SHAPE_SQUARE 2
#define WOS_SHAPE_RECTANGLE 3
#define WOS_SHAPE_ELLIPSE 4
#define WOS_SHAPE_SUPERELLIPSE 5
#define WOS_SHAPE_RING 6
#define WOS_SHAPE_DOUBLE_SLIT 7
#define WOS_SHAPE_CROSS 8
#define WOS_SHAPE_POLYGON 9
typedef struct {
  int32_t kind, n_verts;
  double radius, width, height, a, b, order, rin, rout, inner, length, separation, rotation, edge_sigma;
  const double* verts; /* n_verts (x, y) pairs, host memory */
} wos_shape;
/* apertureEl.Apply (optics/elements.go:288-310): hard edge, or the erf
 * smooth step when edge_sigma > 0, of the signed distance about (x0, y0) */
int wos_hip_field_aperture(wos_field f, const wos_shape* shape, double x0, double y0);
/* diffuserEl.Apply (optics/elements.go:722-783): hashNoise01 phase on the lab
 * lattice anchored at (x0, y0), Gaussian correlation `correlation` (m, 0:
 * white), rescaled to standard deviation `sigma`, times `amplitude` */
int wos_hip_field_diffuser(wos_field f, double sigma, double correlation, double amplitude, int64_t seed, double x0,
                           double y0);
/* Field.ApplyJones (optics/field.go:109-131), abcd = {a.re, a.im, b.re, b.im,
 * c.re, c.im, d.re, d.im}; polarizer, retarder, rotator and custom_jones
 * (optics/elements.go:1018-1085) are this with their matrices */
int wos_hip_field_jones(wos_field f, const double* abcd);
/* More train elements (optics/elements.go:156-178), each applied in the
 * reference's per-pixel order: cylindrical_lens (elements.go:230-270),
 * apodizer (427-456), grating (458-531), axicon (534-556), spiral_phase
 * (560-595), wedge (598-622), zone_plate (625-681), mirror (787-828),
 * concave/convex mirror (831-899), zernike (902-1008), kerr (1168-1205) and
 * saturable_absorber (1211-1242), both on I = |Ex|^2 (+ |Ey|^2 when
 * polarized). `sub` is the grating
 * kind (0 amplitude_sin, 1 amplitude_binary, 2 phase_sin, 3 phase_binary,
 * 4 blazed), the zone-plate kind (0 phase, 1 amplitude) or the spiral charge;
 * the other fields carry the constructors' parameters (zernike: Noll
 * coefficients c1..c21 in waves at [1..21], normalisation radius in radius).
 * Parameters the reference's constructors reject return WOS_ERR_INVALID. */
#define WOS_EL_CYLINDRICAL_LENS 0
#define WOS_EL_APODIZER 1
#define WOS_EL_GRATING 2
#define WOS_EL_AXICON 3
#define WOS_EL_SPIRAL_PHASE 4
#define WOS_EL_WEDGE 5
#define WOS_EL_ZONE_PLATE 6
#define WOS_EL_MIRROR 7
#define WOS_EL_CONCAVE_MIRROR 8
#define WOS_EL_CONVEX_MIRROR 9
#define WOS_EL_ZERNIKE 10
#define WOS_EL_KERR 11
#define WOS_EL_SATURABLE_ABSORBER 12
typedef struct {
  int32_t kind, sub;
  double f, aperture, x0, y0, rotation;
  double waist, amplitude;
  double alpha, index;
  double period, modulation, duty, blaze_depth, offset;
  double radius;
  double reflectivity, curvature, tilt_x, tilt_y;
  double zernike[22];
  double n2, tpa, length, alpha0, isat; /* kerr, saturable_absorber */
} wos_element;
int wos_hip_field_element(wos_field f, const wos_element* element, double wavelength);
/* mediumEl.Apply (optics/elements.go:1105-1126): PropagateSplitStep through a
 * uniform index `index` + i `absorption` over `distance` in `steps` slices on
 * the resident field; warnings as wos_hip_split_step */
int wos_hip_field_medium(wos_field f, double distance, double index, double absorption, int steps,
                         const wos_ctx* ctx, wos_warnings* w);
/* Field.ScaleAmplitude (optics/field.go:134-143), the beamsplitter's arms */
int wos_hip_field_scale(wos_field f, double re, double im);
/* Field.ApplyBandlimit (optics/field.go:264-294); fraction outside (0, 1): no-op */
int wos_hip_field_bandlimit(wos_field f, double fraction, double sigma);
/* A propagate element with the context's band limit (optics/simulator.go:
 * 495-496): wos_hip_field_propagate, then wos_hip_field_bandlimit. For the
 * asm method on a power-of-two grid the two run as one transfer per
 * component (the roll-off applied between H and the inverse column FFT);
 * warnings and stats are the propagation's, with its output power taken
 * before the band limit. */
int wos_hip_field_propagate_bandlimited(wos_field f, double z, int method, const wos_ctx* ctx, double fraction,
                                        double sigma, wos_warnings* w, wos_prop_stats* st);
/* Plane views: the float32 little-endian planes servePlaneData streams
 * (server/server.go:581-660) for one recorded field, computed on the device
 * so a plane crosses PCIe as 4 bytes per pixel (server/views.go:131-244,
 * optics/polar.go:23-96). Masked pixels hold the float32 NaN 0x7FC00000.
 * mask_cut <= 0 takes the server's defaults (2e-3 of the component's peak
 * for phases, 1e-6 of the peak S0 for polarisation views). */
#define WOS_VIEW_TOTAL 0
#define WOS_VIEW_AMPLITUDE 1
#define WOS_VIEW_EX 2
#define WOS_VIEW_EY 3
#define WOS_VIEW_PHASE_X 4
#define WOS_VIEW_PHASE_Y 5
#define WOS_VIEW_POL_INTENSITY 6
#define WOS_VIEW_POL_AZIMUTH 7
#define WOS_VIEW_POL_ELLIP 8
#define WOS_VIEW_POL_S1 9
#define WOS_VIEW_POL_S2 10
#define WOS_VIEW_POL_S3 11
#define WOS_VIEW_POL_DEGREE 12
/* out: n*n floats (host memory) */
int wos_hip_field_view(wos_field f, int view, double mask_cut, float* out);
/* ComputeStats (optics/metrics.go:31-91) of the resident field */
int wos_hip_field_stats(wos_field f, double wl, double strehl_aperture, double strehl_distance,
                        wos_plane_stats* out);

/* ---- tuning tables (no reference counterpart) ----
 * A text file of launch shapes found by an earlier autotune, one entry a
 * line ('#' comments):
 *   n precision method units tile_w col_w rows groups split radix col_radix col_split
 *     [inv_rows inv_groups inv_split] [ms]
 * precision "fp64" or "fp32", method "asm" or "asm_pad", the shape fields as
 * wos_hip_batch_autotune reports them (an entry of eight shape fields and
 * a time has no inverse row shape of its own). A lookup takes the entries
 * of the grid's n, precision and method and the nearest unit count.
 * Loading applies the entry's shape (WOS_ERR_UNSUPPORTED: no entry). */
int wos_hip_batch_load_tuning(wos_batch b, const char* path, int* shape);
/* the launch shape of one pipeline stage (WOS_SHAPE_LEN ints, as above),
 * read or set */
int wos_hip_pipeline_get_tuning(wos_pipeline p, int stage, int* shape);
int wos_hip_pipeline_set_tuning(wos_pipeline p, int stage, const int* shape);
/* every stage with an entry takes its shape; *applied = how many did */
int wos_hip_pipeline_load_tuning(wos_pipeline p, const char* path, int* applied);

/* ---- phase-mask design operator (no reference counterpart) --------------
 * Forward/Backward pair behind a Go operator for phase-mask optimisation. For
 * each of `units` masks theta (n x n, radians): U = ASM_z(exp(i theta)) with
 * asmTF (optics/propagate.go:159-181) on the fft2D conventions
 * (optics/fft.go:209-242), I = |U|^2, S = sum I, and the loss
 * L = sum (I/S - T)^2 against one target T (n x n, real). The forward keeps U
 * on the device as the checkpoint of the backward, which returns
 * dL/dtheta = -2 Im(exp(i theta) conj(A^H (dL/dI . U))) with A^H the same
 * propagation under the conjugate transfer function. Sums are FP64 and
 * deterministic; precision selects the field arithmetic. */
typedef struct wos_mask_s* wos_mask;

int wos_hip_mask_create(wos_mask* out, int n, int units, double dx, double z, const wos_ctx* ctx, int precision,
                        const double* target);
int wos_hip_mask_destroy(wos_mask m);
/* theta: units*n*n doubles (unit-major, row-major) in host memory, or NULL
 * to use the masks already in device memory (the last upload, or what the
 * caller wrote at wos_hip_mask_device_theta); loss: units doubles (may be NULL) */
int wos_hip_mask_forward(wos_mask m, const double* theta, double* loss);
/* gradient of the last forward's losses: units*n*n doubles (may be NULL to
 * leave it on the device, see wos_hip_mask_device_grad); repeatable */
int wos_hip_mask_backward(wos_mask m, double* grad);
/* the forward's checkpoint U of one unit as complex128 n*n */
int wos_hip_mask_field(wos_mask m, int unit, double* out);
/* device pointer of the gradient (units*n*n doubles) */
void* wos_hip_mask_device_grad(wos_mask m);
/* device pointer of the masks theta (units*n*n doubles): an optimiser that
 * keeps its step on the GPU writes theta there and calls the forward with
 * theta = NULL, so no mask crosses PCIe */
void* wos_hip_mask_device_theta(wos_mask m);

/* ---- multi-plane phase design through a split-step medium (no reference
 * counterpart) ------------------------------------------------------------
 * exp(i theta) (n x n) goes through `steps` slices of PropagateSplitStep
 * (optics/medium.go:22-45) over z, with index planes sampled as for
 * wos_hip_split_step, and the screen of slice s multiplied by exp(i phi_s)
 * (design phase planes); the loss is the mask operator's. The forward keeps
 * the field before every screen on the device (steps checkpoints of n*n);
 * the backward runs the slices in reverse with the conjugate transfer
 * function and conjugate screens and returns dL/dtheta and dL/dphi_s. */
typedef struct wos_bpm_mask_s* wos_bpm_mask;

int wos_hip_bpm_mask_create(wos_bpm_mask* out, int n, double dx, double z, int steps, const double* index,
                            int n_planes, const wos_ctx* ctx, int precision, const double* target);
int wos_hip_bpm_mask_destroy(wos_bpm_mask m);
/* theta: n*n; phi: steps*n*n design phases or NULL (none); loss may be NULL */
int wos_hip_bpm_mask_forward(wos_bpm_mask m, const double* theta, const double* phi, double* loss);
/* gradients of the last forward's loss: grad_theta n*n, grad_phi steps*n*n (either may be NULL) */
int wos_hip_bpm_mask_backward(wos_bpm_mask m, double* grad_theta, double* grad_phi);

#ifdef __cplusplus
}
#endif
#endif /* WOS_HIP_H */
