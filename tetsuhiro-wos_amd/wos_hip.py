"""ctypes binding of libwos_hip's C ABI (include/wos_hip.h).

This is the host-side mirror of the reference's optics operators for the
propagation path: fft2d/fft1d stand for fft2D/fft1DAny (optics/fft.go:135,
209), propagate for Propagate (optics/propagate.go:65), split_step for
PropagateSplitStep (optics/medium.go:22; propagate_split_step takes its
IndexFunc), polychromatic for
PropagatePolychromatic (optics/polychromatic.go:18), volume for Propagate3D
(optics/volume.go:8) and propagate1d for Propagate1D
(optics/propagate.go:718). Errors raise WosError; there is no CPU fallback:
loading fails loudly when libwos_hip.so is absent.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwos_hip.so")

FP64, FP32 = 0, 1
METHODS = {"asm": 0, "asm_pad": 1, "asm_shift": 2, "asm_shift_pad": 3,
           "fresnel_tf": 4, "fresnel_ir": 5, "fraunhofer": 6}
WARN_CODES = ["evanescent_filtered", "backward_regularized", "backward_evanescent",
              "asm_alias_wrap", "fresnel_tf_alias", "fresnel_ir_alias", "fraunhofer_nearfield"]
EXPORTS = [
    "wos_hip_device_count", "wos_hip_device_name", "wos_hip_init", "wos_hip_last_error",
    "wos_hip_version", "wos_hip_release_cache", "wos_hip_fft2d_z", "wos_hip_fft1d_z",
    "wos_hip_fft2d_c", "wos_hip_fft1d_c", "wos_hip_propagate", "wos_hip_propagate1d",
    "wos_hip_split_step", "wos_hip_split_step_set_shape", "wos_hip_polychromatic", "wos_hip_volume",
    "wos_hip_batch_create", "wos_hip_batch_destroy", "wos_hip_batch_upload",
    "wos_hip_batch_download", "wos_hip_batch_fill_schell", "wos_hip_batch_capture",
    "wos_hip_batch_run", "wos_hip_batch_sync", "wos_hip_batch_device_ptr", "wos_hip_batch_time",
    "wos_hip_batch_autotune", "wos_hip_batch_set_tuning", "wos_hip_batch_set_persistent", "wos_hip_last_device_ms",
    "wos_hip_mask_create", "wos_hip_mask_destroy", "wos_hip_mask_forward", "wos_hip_mask_backward",
    "wos_hip_mask_field", "wos_hip_mask_device_grad", "wos_hip_mask_device_theta",
    "wos_hip_pipeline_create", "wos_hip_pipeline_destroy", "wos_hip_pipeline_add",
    "wos_hip_pipeline_set_intensity", "wos_hip_pipeline_upload", "wos_hip_pipeline_fill_schell",
    "wos_hip_pipeline_autotune", "wos_hip_pipeline_capture", "wos_hip_pipeline_run", "wos_hip_pipeline_sync",
    "wos_hip_pipeline_download", "wos_hip_pipeline_download_source", "wos_hip_pipeline_intensity",
    "wos_hip_pipeline_time", "wos_hip_compute_stats", "wos_hip_batch_stats", "wos_hip_pipeline_stats",
    "wos_hip_bpm_mask_create", "wos_hip_bpm_mask_destroy", "wos_hip_bpm_mask_forward", "wos_hip_bpm_mask_backward",
    "wos_hip_field_create", "wos_hip_field_destroy", "wos_hip_field_clone", "wos_hip_field_upload",
    "wos_hip_field_download", "wos_hip_field_polarized", "wos_hip_field_device_ptr", "wos_hip_field_propagate",
    "wos_hip_field_propagate_bandlimited", "wos_hip_field_element", "wos_hip_field_medium",
    "wos_hip_field_lens", "wos_hip_field_aperture", "wos_hip_field_diffuser", "wos_hip_field_jones",
    "wos_hip_field_scale", "wos_hip_field_bandlimit", "wos_hip_field_stats", "wos_hip_field_dx",
    "wos_hip_propagate_vectorial", "wos_hip_propagate_uniaxial", "wos_hip_field_uniaxial", "wos_hip_field_view",
    "wos_hip_propagate_anisotropic", "wos_hip_field_anisotropic", "wos_hip_schell_realizations",
    "wos_hip_average_intensity", "wos_hip_propagate_partially_coherent", "wos_hip_pipeline_set_average",
    "wos_hip_propagate_device_noalloc", "wos_hip_check_grid_memory", "wos_hip_batch_load_tuning",
    "wos_hip_pipeline_get_tuning", "wos_hip_pipeline_set_tuning", "wos_hip_pipeline_load_tuning",
]
# the tuning table committed beside the library (written by bench.py --write-tuning on an MI355X)
TUNING_TABLE = os.path.join(_HERE, "tuning_mi355x.txt")

VIEWS = ["total", "amplitude", "ex", "ey", "phase_x", "phase_y", "pol_intensity", "pol_azimuth", "pol_ellip",
         "pol_s1", "pol_s2", "pol_s3", "pol_degree"]

SHAPE_KINDS = {"circle": 0, "slit": 1, "square": 2, "rectangle": 3, "ellipse": 4, "superellipse": 5, "ring": 6,
               "double_slit": 7, "cross": 8, "polygon": 9}


ELEMENT_KINDS = ("cylindrical_lens", "apodizer", "grating", "axicon", "spiral_phase", "wedge", "zone_plate",
                 "mirror", "concave_mirror", "convex_mirror", "zernike", "kerr", "saturable_absorber")
GRATING_KINDS = ("amplitude_sin", "amplitude_binary", "phase_sin", "phase_binary", "blazed")


class Element(C.Structure):
    """wos_element: the parameters of one further train element."""
    _fields_ = [("kind", C.c_int32), ("sub", C.c_int32),
                ("f", C.c_double), ("aperture", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
                ("rotation", C.c_double), ("waist", C.c_double), ("amplitude", C.c_double),
                ("alpha", C.c_double), ("index", C.c_double), ("period", C.c_double),
                ("modulation", C.c_double), ("duty", C.c_double), ("blaze_depth", C.c_double),
                ("offset", C.c_double), ("radius", C.c_double), ("reflectivity", C.c_double),
                ("curvature", C.c_double), ("tilt_x", C.c_double), ("tilt_y", C.c_double),
                ("zernike", C.c_double * 22), ("n2", C.c_double), ("tpa", C.c_double), ("length", C.c_double),
                ("alpha0", C.c_double), ("isat", C.c_double)]


def make_element(element, **p):
    """A wos_element with the reference constructors' defaults
    (optics/elements.go): grating kind by name, zone-plate kind "phase" or
    "amplitude", spiral charge, Zernike coefficients c1..c21 in waves."""
    e = Element()
    e.kind = ELEMENT_KINDS.index(element)
    e.f = p.get("f", {"cylindrical_lens": 0.1, "zone_plate": 0.1}.get(element, 0.0))
    e.aperture, e.x0, e.y0 = p.get("aperture", 0.0), p.get("x", 0.0), p.get("y", 0.0)
    e.rotation = p.get("rotation", 0.0)
    e.waist, e.amplitude = p.get("waist", 0.001), p.get("amplitude", 1.0)
    e.alpha = p.get("alpha", 0.02 if element == "axicon" else 0.01)
    e.index = p.get("index", 1.5)
    e.period, e.modulation, e.duty = p.get("period", 1e-4), p.get("modulation", 1.0), p.get("duty", 0.5)
    e.blaze_depth, e.offset = p.get("blaze_depth", np.pi), p.get("offset", 0.0)
    e.radius = p.get("radius", {"zone_plate": 0.01, "zernike": 0.01}.get(element, 0.5))
    e.reflectivity, e.curvature = p.get("reflectivity", 1.0), p.get("curvature", 0.0)
    e.tilt_x, e.tilt_y = p.get("tilt_x", 0.0), p.get("tilt_y", 0.0)
    if element == "grating":
        e.sub = GRATING_KINDS.index(p.get("kind", "amplitude_sin"))
    elif element == "zone_plate":
        e.sub = ("phase", "amplitude").index(p.get("kind", "phase"))
    elif element == "spiral_phase":
        e.sub = int(p.get("charge", 1))
    for j in range(1, 22):
        e.zernike[j] = p.get(f"c{j}", 0.0)
    e.n2, e.tpa, e.length = p.get("n2", 1e-14), p.get("tpa", 0.0), p.get("length", 1e-3)
    e.alpha0, e.isat = p.get("alpha0", 1.0), p.get("isat", 1e6)
    return e


class Shape(C.Structure):
    """wos_shape: a parsed aperture outline (optics/shape.go:28-48)."""
    _fields_ = [("kind", C.c_int32), ("n_verts", C.c_int32)] + \
        [(k, C.c_double) for k in ("radius", "width", "height", "a", "b", "order", "rin", "rout", "inner",
                                   "length", "separation", "rotation", "edge_sigma")] + [("verts", C.c_void_p)]


def make_shape(kind, verts=None, **params):
    """Shape of `kind` ("circle" ... "cross", or "polygon" with an (m, 2)
    vertex array); unset parameters take the reference's defaults
    (optics/shape.go:51-55)."""
    d = {"radius": 1e-3, "width": 2e-3, "height": 2e-3, "a": 1e-3, "b": 2e-3, "order": 2, "rin": 5e-4,
         "rout": 1e-3, "inner": 5e-4, "length": 4e-3, "separation": 1e-3, "rotation": 0, "edge_sigma": 0}
    d.update(params)
    g = Shape(kind=SHAPE_KINDS[kind], n_verts=0, **d)
    if verts is not None:
        v = np.ascontiguousarray(verts, dtype=np.float64)
        g.n_verts = v.shape[0]
        g.verts = v.ctypes.data
        g._keep = v
    return g


# launch shape fields of a batch (wos_hip_batch_autotune / _set_tuning): rows,
# row_groups and split_lds are the forward row pass's, inv_* the inverse row
# pass's (inv_rows < 0 when set: as the forward pass)
TUNING_KEYS = ("tile_w", "col_w", "rows", "row_groups", "split_lds", "radix", "col_radix", "col_split_lds",
               "inv_rows", "inv_row_groups", "inv_split_lds")
SHAPE_LEN = len(TUNING_KEYS)


def _shape_array(shape):
    """A shape dict or sequence as the C ABI's WOS_SHAPE_LEN ints; missing
    inverse row fields mean "as the forward pass"."""
    if isinstance(shape, dict):
        vals = [shape.get(k, -1 if k == "inv_rows" else 0) for k in TUNING_KEYS]
    else:
        vals = list(shape) + [-1, 0, 0][len(shape) - 8:] if len(shape) < SHAPE_LEN else list(shape)
    return np.ascontiguousarray(vals[:SHAPE_LEN], dtype=np.int32)


class WosError(RuntimeError):
    pass


class PlaneStats(C.Structure):
    """wos_plane_stats: PlaneStats of ComputeStats."""
    _fields_ = [(k, C.c_double) for k in
                ("power", "peak", "centroid_x", "centroid_y", "rms_x", "rms_y", "strehl",
                 "intensity_min", "intensity_max", "phase_min", "phase_max")]


class Ctx(C.Structure):
    """wos_ctx: the Context fields of the path (optics/field.go:237-254)."""
    _fields_ = [("wavelength", C.c_double),
                ("evanescent_zero", C.c_int32),
                ("backward_regularize", C.c_int32),
                ("evanescent_limit", C.c_double),
                ("tikhonov_alpha", C.c_double)]


class Warnings(C.Structure):
    _fields_ = [("count", C.c_int32 * 8), ("value", C.c_double * 8)]

    def as_dict(self):
        return {code: (self.count[i], self.value[i])
                for i, code in enumerate(WARN_CODES) if self.count[i]}


class PropStats(C.Structure):
    _fields_ = [("p_in", C.c_double), ("p_out", C.c_double),
                ("fx_c", C.c_double), ("fy_c", C.c_double),
                ("method_used", C.c_int32), ("reserved", C.c_int32)]


def ctx_for(wl, evanescent="decay", limit=0.0, backward_regularize=False, alpha=0.0):
    return Ctx(wl, 1 if evanescent == "zero" else 0, 1 if backward_regularize else 0, limit, alpha)


_lib = None


def load(path=None):
    """Load libwos_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise WosError(f"libwos_hip.so not built at {path}; run __graft_entry__.build()")
    L = C.CDLL(path)
    P, D, I = C.c_void_p, C.c_double, C.c_int
    L.wos_hip_device_count.restype = I
    L.wos_hip_device_name.argtypes = [C.c_char_p, I]
    L.wos_hip_init.restype = I
    L.wos_hip_last_error.restype = C.c_char_p
    L.wos_hip_version.restype = C.c_char_p
    L.wos_hip_last_device_ms.restype = C.c_double
    for nm in ("wos_hip_fft2d_z", "wos_hip_fft1d_z", "wos_hip_fft2d_c", "wos_hip_fft1d_c"):
        getattr(L, nm).argtypes = [P, I, I]
        getattr(L, nm).restype = I
    L.wos_hip_propagate.argtypes = [P, P, I, D, D, I, C.POINTER(Ctx), I, C.POINTER(Warnings),
                                    C.POINTER(PropStats)]
    L.wos_hip_propagate1d.argtypes = [P, I, D, D, D, C.POINTER(Ctx), I, C.POINTER(Warnings)]
    L.wos_hip_split_step.argtypes = [P, P, I, D, D, I, P, I, C.POINTER(Ctx), I, C.POINTER(Warnings)]
    L.wos_hip_split_step_set_shape.argtypes = [I, I, I, I, I, I]
    L.wos_hip_polychromatic.argtypes = [P, P, I, D, D, P, P, I, I, C.POINTER(Ctx), I, P,
                                        C.POINTER(Warnings)]
    L.wos_hip_volume.argtypes = [P, P, I, D, P, I, I, C.POINTER(Ctx), I, P, P, C.POINTER(Warnings)]
    L.wos_hip_batch_create.argtypes = [C.POINTER(P), I, D, I, I, I, P, P, C.POINTER(Ctx)]
    L.wos_hip_batch_destroy.argtypes = [P]
    L.wos_hip_batch_upload.argtypes = [P, I, P]
    L.wos_hip_batch_download.argtypes = [P, I, P]
    L.wos_hip_batch_fill_schell.argtypes = [P, D, D, C.c_uint64]
    L.wos_hip_batch_capture.argtypes = [P]
    L.wos_hip_batch_run.argtypes = [P, I]
    L.wos_hip_batch_sync.argtypes = [P]
    L.wos_hip_batch_device_ptr.argtypes = [P]
    L.wos_hip_batch_device_ptr.restype = P
    L.wos_hip_batch_time.argtypes = [P, I, I, C.POINTER(D), P]
    L.wos_hip_batch_autotune.argtypes = [P, I, P, C.POINTER(D)]
    L.wos_hip_batch_set_tuning.argtypes = [P, I, I, I, I, I, I, I, I, I, I, I]
    L.wos_hip_batch_set_persistent.argtypes = [P, I]
    L.wos_hip_mask_create.argtypes = [C.POINTER(P), I, I, D, D, C.POINTER(Ctx), I, P]
    L.wos_hip_mask_destroy.argtypes = [P]
    L.wos_hip_mask_forward.argtypes = [P, P, P]
    L.wos_hip_mask_backward.argtypes = [P, P]
    L.wos_hip_mask_field.argtypes = [P, I, P]
    L.wos_hip_mask_device_grad.argtypes = [P]
    L.wos_hip_mask_device_grad.restype = P
    L.wos_hip_mask_device_theta.argtypes = [P]
    L.wos_hip_mask_device_theta.restype = P
    L.wos_hip_pipeline_create.argtypes = [C.POINTER(P), I]
    L.wos_hip_pipeline_destroy.argtypes = [P]
    L.wos_hip_pipeline_add.argtypes = [P, I, D, I, I, I, P, P, C.POINTER(Ctx), C.POINTER(I)]
    L.wos_hip_pipeline_set_intensity.argtypes = [P, I, P]
    L.wos_hip_pipeline_upload.argtypes = [P, I, I, P]
    L.wos_hip_pipeline_fill_schell.argtypes = [P, I, D, D, C.c_uint64, I]
    L.wos_hip_pipeline_autotune.argtypes = [P, I, P]
    L.wos_hip_pipeline_capture.argtypes = [P]
    L.wos_hip_pipeline_run.argtypes = [P, I]
    L.wos_hip_pipeline_sync.argtypes = [P]
    L.wos_hip_pipeline_download.argtypes = [P, I, I, P]
    L.wos_hip_pipeline_download_source.argtypes = [P, I, I, P]
    L.wos_hip_pipeline_intensity.argtypes = [P, I, P]
    L.wos_hip_pipeline_time.argtypes = [P, I, I, C.POINTER(D), P]
    L.wos_hip_compute_stats.argtypes = [P, P, I, D, D, D, D, C.POINTER(PlaneStats)]
    L.wos_hip_batch_stats.argtypes = [P, I, D, D, D, C.POINTER(PlaneStats)]
    L.wos_hip_pipeline_stats.argtypes = [P, I, I, D, D, D, C.POINTER(PlaneStats)]
    L.wos_hip_bpm_mask_create.argtypes = [C.POINTER(P), I, D, D, I, P, I, C.POINTER(Ctx), I, P]
    L.wos_hip_bpm_mask_destroy.argtypes = [P]
    L.wos_hip_bpm_mask_forward.argtypes = [P, P, P, P]
    L.wos_hip_bpm_mask_backward.argtypes = [P, P, P]
    L.wos_hip_field_create.argtypes = [P, I, D, I, I]
    L.wos_hip_field_destroy.argtypes = [P]
    L.wos_hip_field_clone.argtypes = [P, P]
    L.wos_hip_field_upload.argtypes = [P, P, P]
    L.wos_hip_field_download.argtypes = [P, P, P]
    L.wos_hip_field_polarized.argtypes = [P]
    L.wos_hip_field_device_ptr.argtypes = [P]
    L.wos_hip_field_device_ptr.restype = P
    L.wos_hip_field_propagate.argtypes = [P, D, I, C.POINTER(Ctx), C.POINTER(Warnings), C.POINTER(PropStats)]
    L.wos_hip_field_propagate_bandlimited.argtypes = [P, D, I, C.POINTER(Ctx), D, D, C.POINTER(Warnings),
                                                      C.POINTER(PropStats)]
    L.wos_hip_field_element.argtypes = [P, C.POINTER(Element), D]
    L.wos_hip_field_medium.argtypes = [P, D, D, D, I, C.POINTER(Ctx), C.POINTER(Warnings)]
    L.wos_hip_field_lens.argtypes = [P, D, D, D, D, D]
    L.wos_hip_field_aperture.argtypes = [P, C.POINTER(Shape), D, D]
    L.wos_hip_field_diffuser.argtypes = [P, D, D, D, C.c_int64, D, D]
    L.wos_hip_field_jones.argtypes = [P, P]
    L.wos_hip_field_scale.argtypes = [P, D, D]
    L.wos_hip_field_bandlimit.argtypes = [P, D, D]
    L.wos_hip_field_stats.argtypes = [P, D, D, D, C.POINTER(PlaneStats)]
    L.wos_hip_propagate_vectorial.argtypes = [P, P, P, I, D, D, C.POINTER(Ctx), I]
    L.wos_hip_propagate_uniaxial.argtypes = [P, P, I, D, D, D, D, C.POINTER(Ctx), I]
    L.wos_hip_field_uniaxial.argtypes = [P, D, D, D, D]
    L.wos_hip_field_view.argtypes = [P, I, D, P]
    L.wos_hip_propagate_anisotropic.argtypes = [P, P, I, D, D, P, C.POINTER(Ctx), I]
    L.wos_hip_field_anisotropic.argtypes = [P, D, P, D]
    L.wos_hip_field_dx.argtypes = [P]
    L.wos_hip_field_dx.restype = D
    L.wos_hip_schell_realizations.argtypes = [I, D, D, D, C.c_int64, I, P]
    L.wos_hip_average_intensity.argtypes = [P, P, I, I, P]
    L.wos_hip_propagate_partially_coherent.argtypes = [I, D, D, D, C.c_int64, I, D, I, C.POINTER(Ctx), I, P,
                                                       C.POINTER(Warnings)]
    L.wos_hip_pipeline_set_average.argtypes = [P, I]
    L.wos_hip_propagate_device_noalloc.argtypes = [P, I, I, D, D, I, C.POINTER(Ctx), I, P]
    L.wos_hip_check_grid_memory.argtypes = [I, I, I, I, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.wos_hip_batch_load_tuning.argtypes = [P, C.c_char_p, P]
    L.wos_hip_pipeline_get_tuning.argtypes = [P, I, P]
    L.wos_hip_pipeline_set_tuning.argtypes = [P, I, P]
    L.wos_hip_pipeline_load_tuning.argtypes = [P, C.c_char_p, C.POINTER(I)]
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = load().wos_hip_last_error().decode(errors="replace")
        raise WosError(f"{what} failed (rc={rc}): {msg}")


def _p(a):
    return None if a is None else a.ctypes.data


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def last_device_ms():
    """Device time of the last split-step pipeline (HIP events), ms."""
    return load().wos_hip_last_device_ms()


def device_count():
    return load().wos_hip_device_count()


def init():
    _check(load().wos_hip_init(), "wos_hip_init")


def device_name():
    buf = C.create_string_buffer(256)
    _check(load().wos_hip_device_name(buf, 256), "wos_hip_device_name")
    return buf.value.decode()


def fft2d(a, inverse=False, precision=FP64):
    """fft2D on a copy: forward unnormalised, inverse scaled by 1/n^2."""
    if precision == FP64:
        a = _c128(a).copy()
        _check(load().wos_hip_fft2d_z(_p(a), a.shape[0], int(inverse)), "fft2d_z")
    else:
        a = np.ascontiguousarray(a, dtype=np.complex64).copy()
        _check(load().wos_hip_fft2d_c(_p(a), a.shape[0], int(inverse)), "fft2d_c")
    return a


def fft1d(a, inverse=False, precision=FP64):
    """fft1DAny on a copy: inverse scaled by 1/n."""
    if precision == FP64:
        a = _c128(a).copy()
        _check(load().wos_hip_fft1d_z(_p(a), a.size, int(inverse)), "fft1d_z")
    else:
        a = np.ascontiguousarray(a, dtype=np.complex64).copy()
        _check(load().wos_hip_fft1d_c(_p(a), a.size, int(inverse)), "fft1d_c")
    return a


def propagate(ex, ey, dx, z, method, ctx, precision=FP64):
    """Propagate on copies; returns (ex, ey, warnings, stats)."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    ex = _c128(ex).copy()
    ey = None if ey is None else _c128(ey).copy()
    w, st = Warnings(), PropStats()
    _check(load().wos_hip_propagate(_p(ex), _p(ey), ex.shape[0], dx, z, m, C.byref(ctx), precision,
                                    C.byref(w), C.byref(st)), "propagate")
    return ex, ey, w, st


class Field:
    """Device-resident Field (wos_field): the element chain of a train runs
    on it without host round trips."""

    def __init__(self, n, dx, polarized=False, precision=FP64, _handle=None):
        self.n, self.dx, self.precision = n, dx, precision
        if _handle is not None:
            self.h = _handle
            return
        h = C.c_void_p()
        _check(load().wos_hip_field_create(C.byref(h), n, dx, int(polarized), precision), "field_create")
        self.h = h

    @classmethod
    def from_host(cls, ex, ey=None, dx=1e-6, polarized=None, precision=FP64):
        ex = _c128(ex)
        f = cls(ex.shape[0], dx, (ey is not None) if polarized is None else polarized, precision)
        f.upload(ex, ey)
        return f

    def close(self):
        if self.h:
            _check(load().wos_hip_field_destroy(self.h), "field_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def polarized(self):
        return bool(load().wos_hip_field_polarized(self.h))

    @property
    def grid_dx(self):
        """Field.DX on the device (a fraunhofer propagation changes it)."""
        return load().wos_hip_field_dx(self.h)

    def clone(self):
        h = C.c_void_p()
        _check(load().wos_hip_field_clone(self.h, C.byref(h)), "field_clone")
        return Field(self.n, self.dx, precision=self.precision, _handle=h)

    def upload(self, ex, ey=None):
        ex = _c128(ex)
        ey = None if ey is None else _c128(ey)
        _check(load().wos_hip_field_upload(self.h, _p(ex), _p(ey)), "field_upload")

    def download(self):
        ex = np.zeros((self.n, self.n), dtype=np.complex128)
        ey = np.zeros_like(ex)
        _check(load().wos_hip_field_download(self.h, _p(ex), _p(ey)), "field_download")
        return ex, ey

    def propagate(self, z, method, ctx, bandlimit=None):
        """Propagate; bandlimit=(fraction, sigma) applies ApplyBandlimit after
        it (one fused transfer for asm on a power-of-two grid)."""
        m = METHODS[method] if isinstance(method, str) else int(method)
        w, st = Warnings(), PropStats()
        if bandlimit is None:
            _check(load().wos_hip_field_propagate(self.h, z, m, C.byref(ctx), C.byref(w), C.byref(st)),
                   "field_propagate")
        else:
            _check(load().wos_hip_field_propagate_bandlimited(self.h, z, m, C.byref(ctx), float(bandlimit[0]),
                                                               float(bandlimit[1]), C.byref(w), C.byref(st)),
                   "field_propagate_bandlimited")
        return w, st

    def lens(self, wl, f, aperture=0.0, x0=0.0, y0=0.0):
        _check(load().wos_hip_field_lens(self.h, wl, f, aperture, x0, y0), "field_lens")

    def aperture(self, g, x0=0.0, y0=0.0):
        _check(load().wos_hip_field_aperture(self.h, C.byref(g), x0, y0), "field_aperture")

    def element(self, e, wl):
        """One of the further train elements (make_element)."""
        _check(load().wos_hip_field_element(self.h, C.byref(e), wl), "field_element")

    def medium(self, distance, index, absorption, steps, ctx):
        """mediumEl.Apply: split-step through a uniform index; returns the warnings."""
        w = Warnings()
        _check(load().wos_hip_field_medium(self.h, distance, index, absorption, steps, C.byref(ctx), C.byref(w)),
               "field_medium")
        return w

    def diffuser(self, sigma=np.pi, corr=2e-5, amp=1.0, seed=1, x0=0.0, y0=0.0):
        _check(load().wos_hip_field_diffuser(self.h, sigma, corr, amp, seed, x0, y0), "field_diffuser")

    def jones(self, a, b, c, d):
        m = np.array([complex(v) for v in (a, b, c, d)], dtype=np.complex128)
        _check(load().wos_hip_field_jones(self.h, _p(m)), "field_jones")

    def scale(self, s):
        s = complex(s)
        _check(load().wos_hip_field_scale(self.h, s.real, s.imag), "field_scale")

    def bandlimit(self, fraction, sigma):
        _check(load().wos_hip_field_bandlimit(self.h, fraction, sigma), "field_bandlimit")

    def uniaxial(self, z, no, ne, wl):
        _check(load().wos_hip_field_uniaxial(self.h, z, no, ne, wl), "field_uniaxial")

    def anisotropic(self, z, eps, wl):
        e = np.ascontiguousarray(eps, dtype=np.complex128)
        _check(load().wos_hip_field_anisotropic(self.h, z, e.ctypes.data, wl), "field_anisotropic")

    def view(self, name, mask_cut=0.0):
        """The server's float32 plane view `name` (VIEWS)."""
        out = np.zeros((self.n, self.n), dtype=np.float32)
        _check(load().wos_hip_field_view(self.h, VIEWS.index(name), mask_cut, out.ctypes.data), "field_view")
        return out

    def stats(self, wl, strehl_aperture=0.0, strehl_distance=0.0):
        st = PlaneStats()
        _check(load().wos_hip_field_stats(self.h, wl, strehl_aperture, strehl_distance, C.byref(st)), "field_stats")
        return st


def propagate_anisotropic(ex, ey, dx, z, eps, ctx, precision=FP64):
    """PropagateAnisotropic on copies; eps 3x3 complex; ey None: zeros."""
    ex = _c128(ex).copy()
    ey = np.zeros_like(ex) if ey is None else _c128(ey).copy()
    e = np.ascontiguousarray(eps, dtype=np.complex128)
    _check(load().wos_hip_propagate_anisotropic(_p(ex), _p(ey), ex.shape[0], dx, z, e.ctypes.data, C.byref(ctx),
                                                precision), "propagate_anisotropic")
    return ex, ey


def propagate_uniaxial(ex, ey, dx, z, no, ne, ctx, precision=FP64):
    """PropagateUniaxial on copies (ey None: scalar field)."""
    ex = _c128(ex).copy()
    ey = None if ey is None else _c128(ey).copy()
    _check(load().wos_hip_propagate_uniaxial(_p(ex), _p(ey), ex.shape[0], dx, z, no, ne, C.byref(ctx), precision),
           "propagate_uniaxial")
    return ex, ey


def propagate_vectorial(ex, ey, dx, z, ctx, precision=FP64):
    """PropagateVectorial on copies; ey None: zeros. Returns (ex, ey, ez)."""
    ex = _c128(ex).copy()
    ey = np.zeros_like(ex) if ey is None else _c128(ey).copy()
    ez = np.zeros_like(ex)
    _check(load().wos_hip_propagate_vectorial(_p(ex), _p(ey), _p(ez), ex.shape[0], dx, z, C.byref(ctx), precision),
           "propagate_vectorial")
    return ex, ey, ez


def compute_stats(ex, ey, dx, wl, strehl_aperture=0.0, strehl_distance=0.0):
    """ComputeStats of a host field on the device."""
    ex = _c128(ex)
    ey = None if ey is None else _c128(ey)
    st = PlaneStats()
    _check(load().wos_hip_compute_stats(_p(ex), _p(ey), ex.shape[0], dx, wl, strehl_aperture, strehl_distance,
                                        C.byref(st)), "compute_stats")
    return st


def propagate1d(a, dx, z, wl, ctx, precision=FP64):
    a = _c128(a).copy()
    w = Warnings()
    _check(load().wos_hip_propagate1d(_p(a), a.size, dx, z, wl, C.byref(ctx), precision, C.byref(w)),
           "propagate1d")
    return a, w


def split_step(ex, ey, dx, z, steps, index, ctx, precision=FP64):
    ex = _c128(ex).copy()
    ey = None if ey is None else _c128(ey).copy()
    index = _c128(index)
    planes = 1 if index.ndim == 2 else index.shape[0]
    w = Warnings()
    _check(load().wos_hip_split_step(_p(ex), _p(ey), ex.shape[0], dx, z, steps, _p(index), planes,
                                     C.byref(ctx), precision, C.byref(w)), "split_step")
    return ex, ey, w


def sample_index_planes(index_fn, n, dx, z, steps):
    """The IndexFunc of PropagateSplitStep (optics/medium.go:22-45) sampled
    as applyMediumPhase reads it (optics/medium.go:50-65): slice s at
    zc = z*s/steps + dz/2, pixel (i, j) at (X(i), Y(j)) with
    X(i) = (i - N/2) dx. index_fn(x, y, zc) takes broadcast x (1, n) and
    y (n, 1) arrays and returns the complex index. Planes are compared with
    the first as they are sampled; a z-invariant medium (every plane equal
    bit for bit) returns the one plane. Returns a (planes, n, n) complex128
    array with planes 1 or steps."""
    if steps <= 0:
        steps = 1
    dz = z / steps
    c = (np.arange(n) - n / 2) * dx
    x, y = c[None, :], c[:, None]

    def plane(s):
        zc = z * s / steps + dz / 2
        return np.ascontiguousarray(np.broadcast_to(index_fn(x, y, zc), (n, n)), dtype=np.complex128)

    first = plane(0)
    planes = None
    for s in range(1, steps):
        p = plane(s)
        if planes is None and not np.array_equal(p.view(np.float64), first.view(np.float64)):
            planes = np.empty((steps, n, n), dtype=np.complex128)
            planes[:s] = first
        if planes is not None:
            planes[s] = p
    return first[None] if planes is None else planes


def propagate_split_step(ex, ey, dx, z, index_fn, steps, ctx, precision=FP64):
    """PropagateSplitStep(f, z, n, steps, ctx) (optics/medium.go:22) with the
    reference's arguments: the IndexFunc is sampled on the host
    (sample_index_planes) and the planes go to wos_hip_split_step. Returns
    (ex, ey, warnings) on copies."""
    if z == 0:
        return _c128(ex).copy(), None if ey is None else _c128(ey).copy(), Warnings()
    steps = steps if steps > 0 else 1
    planes = sample_index_planes(index_fn, _c128(ex).shape[0], dx, z, steps)
    return split_step(ex, ey, dx, z, steps, planes[0] if planes.shape[0] == 1 else planes, ctx, precision)


def split_step_set_shape(precision, n, tile_w, rows=0, radix=0, col_radix=None):
    """Launch shape of the split-step slices (tile_w = 0: built-in choice);
    col_radix None: the row passes' radix."""
    _check(load().wos_hip_split_step_set_shape(precision, n, tile_w, rows, radix,
                                               radix if col_radix is None else col_radix), "split_step_set_shape")


def polychromatic(ex, ey, dx, z, wavelengths, weights, method, ctx, precision=FP64):
    m = METHODS[method] if isinstance(method, str) else int(method)
    ex = _c128(ex)
    ey = None if ey is None else _c128(ey)
    n = ex.shape[0]
    wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
    wt = np.ascontiguousarray(weights, dtype=np.float64)
    out = np.zeros((n, n), dtype=np.float64)
    w = Warnings()
    _check(load().wos_hip_polychromatic(_p(ex), _p(ey), n, dx, z, _p(wl), _p(wt), wl.size, m,
                                        C.byref(ctx), precision, _p(out), C.byref(w)), "polychromatic")
    return out, w


def volume(ex, ey, dx, zs, method, ctx, precision=FP64):
    m = METHODS[method] if isinstance(method, str) else int(method)
    ex = _c128(ex)
    ey = None if ey is None else _c128(ey)
    n = ex.shape[0]
    zs = np.ascontiguousarray(zs, dtype=np.float64)
    ox = np.zeros((zs.size, n, n), dtype=np.complex128)
    oy = None if ey is None else np.zeros((zs.size, n, n), dtype=np.complex128)
    w = Warnings()
    _check(load().wos_hip_volume(_p(ex), _p(ey), n, dx, _p(zs), zs.size, m, C.byref(ctx), precision,
                                 _p(ox), _p(oy), C.byref(w)), "volume")
    return ox, oy, w


def check_grid_memory(n, method="asm", precision=FP64, units=1):
    """CheckGridMemory on the device: (need, budget) in bytes; raises
    WosError (WOS_ERR_NOMEM) when the estimate exceeds the budget."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    need, budget = C.c_uint64(), C.c_uint64()
    _check(load().wos_hip_check_grid_memory(n, m, precision, units, C.byref(need), C.byref(budget)),
           "check_grid_memory")
    return need.value, budget.value


def propagate_device_noalloc(dev_ptr, n, ncomp, dx, z, method, ctx, precision=FP64, stream=None):
    """wos_hip_propagate_device_noalloc on device memory at dev_ptr (in place,
    enqueued on `stream`, None: the library's stream)."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    _check(load().wos_hip_propagate_device_noalloc(dev_ptr, n, ncomp, dx, z, m, C.byref(ctx), precision, stream),
           "propagate_device_noalloc")


def schell_realizations(n, dx, w0, sigma_c, seed, m):
    """GenerateSchellRealizations on the device: (m, n, n) complex128."""
    out = np.zeros((max(m, 1), n, n), dtype=np.complex128)
    _check(load().wos_hip_schell_realizations(n, dx, w0, sigma_c, seed, m, _p(out)), "schell_realizations")
    return out


def average_intensity(fields_x, fields_y=None):
    """AverageIntensity of m fields given as (m, n, n) arrays (None for m = 0)."""
    fx = _c128(fields_x)
    if fx.shape[0] == 0:
        return None
    fy = None if fields_y is None else _c128(fields_y)
    out = np.zeros(fx.shape[1:], dtype=np.float64)
    _check(load().wos_hip_average_intensity(_p(fx), _p(fy), fx.shape[0], fx.shape[1], _p(out)), "average_intensity")
    return out


def propagate_partially_coherent(n, dx, w0, sigma_c, seed, m, z, method, ctx, precision=FP64):
    """PropagatePartiallyCoherent on the device; returns (intensity, warnings)."""
    mcode = METHODS[method] if isinstance(method, str) else int(method)
    out = np.zeros((n, n), dtype=np.float64)
    w = Warnings()
    _check(load().wos_hip_propagate_partially_coherent(n, dx, w0, sigma_c, seed, m, z, mcode, C.byref(ctx),
                                                       precision, _p(out), C.byref(w)),
           "propagate_partially_coherent")
    return out, w


class Batch:
    """Device-resident batch of independent n x n propagations (wos_batch)."""

    def __init__(self, n, dx, wavelengths, zs, ctx, method="asm", precision=FP64):
        self.n = n
        self.precision = precision
        wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
        z = np.ascontiguousarray(zs, dtype=np.float64)
        assert wl.size == z.size
        self.units = wl.size
        m = METHODS[method] if isinstance(method, str) else int(method)
        h = C.c_void_p()
        _check(load().wos_hip_batch_create(C.byref(h), n, dx, m, precision, wl.size, _p(wl), _p(z),
                                           C.byref(ctx)), "batch_create")
        self.h = h

    def close(self):
        if self.h:
            _check(load().wos_hip_batch_destroy(self.h), "batch_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, unit, ex):
        ex = _c128(ex)
        _check(load().wos_hip_batch_upload(self.h, unit, _p(ex)), "batch_upload")

    def download(self, unit):
        out = np.zeros((self.n, self.n), dtype=np.complex128)
        _check(load().wos_hip_batch_download(self.h, unit, _p(out)), "batch_download")
        return out

    def fill_schell(self, w0, sigma_c, seed):
        _check(load().wos_hip_batch_fill_schell(self.h, w0, sigma_c, seed), "batch_fill_schell")

    def autotune(self, iters=5):
        """Pick the launch shape; returns ({tile_w, col_w, rows, row_groups,
        split_lds, radix, col_radix, col_split_lds, inv_rows, inv_row_groups,
        inv_split_lds}, ms per run)."""
        shape = np.zeros(SHAPE_LEN, dtype=np.int32)
        ms = C.c_double()
        _check(load().wos_hip_batch_autotune(self.h, iters, _p(shape), C.byref(ms)), "batch_autotune")
        return dict(zip(TUNING_KEYS, (int(x) for x in shape))), ms.value

    def set_tuning(self, tile_w, col_w=None, rows=0, groups=0, split=0, radix=0, col_radix=None, col_split=0,
                   inv_rows=-1, inv_groups=0, inv_split=0):
        """col_radix None: the row passes' radix; inv_rows < 0: the inverse
        row pass takes rows, groups and split."""
        _check(load().wos_hip_batch_set_tuning(self.h, tile_w, tile_w if col_w is None else col_w, rows, groups,
                                               split, radix, radix if col_radix is None else col_radix, col_split,
                                               inv_rows, inv_groups, inv_split),
               "batch_set_tuning")

    def load_tuning(self, path=None):
        """Apply the tuning table's shape for this batch; returns the shape
        dict, or None when the table has no entry for the grid."""
        shape = np.zeros(SHAPE_LEN, dtype=np.int32)
        rc = load().wos_hip_batch_load_tuning(self.h, (path or TUNING_TABLE).encode(), _p(shape))
        if rc == 4:
            return None
        _check(rc, "batch_load_tuning")
        return dict(zip(TUNING_KEYS, (int(x) for x in shape)))

    def set_persistent(self, on=True, interleave=False):
        """One persistent launch per run (per-XCD work queues) instead of
        three passes, the queued passes one after another or (interleave)
        task by task; raises WosError where it is not built."""
        _check(load().wos_hip_batch_set_persistent(self.h, (2 if interleave else 1) if on else 0),
               "batch_set_persistent")

    def capture(self):
        _check(load().wos_hip_batch_capture(self.h), "batch_capture")

    def run(self, iters=1):
        _check(load().wos_hip_batch_run(self.h, iters), "batch_run")

    def sync(self):
        _check(load().wos_hip_batch_sync(self.h), "batch_sync")

    def stats(self, unit, wl, strehl_aperture=0.0, strehl_distance=0.0):
        st = PlaneStats()
        _check(load().wos_hip_batch_stats(self.h, unit, wl, strehl_aperture, strehl_distance, C.byref(st)),
               "batch_stats")
        return st

    def time(self, warmup, iters, kernels=False):
        ms = C.c_double()
        k = np.zeros(3, dtype=np.float64)
        _check(load().wos_hip_batch_time(self.h, warmup, iters, C.byref(ms), _p(k) if kernels else None),
               "batch_time")
        return ms.value, (k if kernels else None)


class MaskOp:
    """Phase-mask design operator (wos_mask): forward loss of `units` masks
    and the adjoint gradient, the Forward/Backward pair a Go optimiser calls."""

    def __init__(self, n, units, dx, z, ctx, target, precision=FP64):
        self.n, self.units = n, units
        t = np.ascontiguousarray(target, dtype=np.float64)
        assert t.shape == (n, n)
        h = C.c_void_p()
        _check(load().wos_hip_mask_create(C.byref(h), n, units, dx, z, C.byref(ctx), precision, _p(t)),
               "mask_create")
        self.h = h

    def close(self):
        if self.h:
            _check(load().wos_hip_mask_destroy(self.h), "mask_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, theta):
        """theta: (units, n, n) radians -> losses (units,); None: the masks
        already in device memory"""
        th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float64).reshape(self.units, self.n, self.n)
        loss = np.zeros(self.units, dtype=np.float64)
        _check(load().wos_hip_mask_forward(self.h, _p(th), _p(loss)), "mask_forward")
        return loss

    def backward(self, to_host=True):
        """dL/dtheta of the last forward, (units, n, n); to_host False leaves
        it in device memory (wos_hip_mask_device_grad) and returns None"""
        g = np.zeros((self.units, self.n, self.n), dtype=np.float64) if to_host else None
        _check(load().wos_hip_mask_backward(self.h, _p(g)), "mask_backward")
        return g

    def field(self, unit):
        out = np.zeros((self.n, self.n), dtype=np.complex128)
        _check(load().wos_hip_mask_field(self.h, unit, _p(out)), "mask_field")
        return out


class Pipeline:
    """Heterogeneous device pipeline (wos_pipeline): stages of batched
    propagations of different grid sizes replayed as one hipGraph."""

    def __init__(self, precision=FP64):
        self.precision = precision
        self.stages = []
        h = C.c_void_p()
        _check(load().wos_hip_pipeline_create(C.byref(h), precision), "pipeline_create")
        self.h = h

    def close(self):
        if self.h:
            _check(load().wos_hip_pipeline_destroy(self.h), "pipeline_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, n, dx, wavelengths, zs, ctx, method="asm", src_units=None):
        wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
        z = np.ascontiguousarray(zs, dtype=np.float64)
        assert wl.size == z.size
        units = wl.size
        src = units if src_units is None else src_units
        m = METHODS[method] if isinstance(method, str) else int(method)
        st = C.c_int()
        _check(load().wos_hip_pipeline_add(self.h, n, dx, m, units, src, _p(wl), _p(z), C.byref(ctx),
                                           C.byref(st)), "pipeline_add")
        self.stages.append((n, units))
        return st.value

    def set_intensity(self, stage, weights):
        w = np.ascontiguousarray(weights, dtype=np.float64)
        _check(load().wos_hip_pipeline_set_intensity(self.h, stage, _p(w)), "pipeline_set_intensity")

    def set_average(self, stage):
        """AverageIntensity of the stage's outputs (sum of |U_u|^2, times 1/units)."""
        _check(load().wos_hip_pipeline_set_average(self.h, stage), "pipeline_set_average")

    def upload(self, stage, src_index, ex):
        ex = _c128(ex)
        _check(load().wos_hip_pipeline_upload(self.h, stage, src_index, _p(ex)), "pipeline_upload")

    def fill_schell(self, stage, w0, sigma_c, seed, first=0):
        _check(load().wos_hip_pipeline_fill_schell(self.h, stage, w0, sigma_c, seed, first), "pipeline_fill_schell")

    def autotune(self, iters=3):
        ms = np.zeros(len(self.stages), dtype=np.float64)
        _check(load().wos_hip_pipeline_autotune(self.h, iters, _p(ms)), "pipeline_autotune")
        return ms

    def get_tuning(self, stage):
        shape = np.zeros(SHAPE_LEN, dtype=np.int32)
        _check(load().wos_hip_pipeline_get_tuning(self.h, stage, _p(shape)), "pipeline_get_tuning")
        return dict(zip(TUNING_KEYS, (int(x) for x in shape)))

    def set_tuning(self, stage, shape):
        sh = _shape_array(shape)
        _check(load().wos_hip_pipeline_set_tuning(self.h, stage, _p(sh)), "pipeline_set_tuning")

    def load_tuning(self, path=None):
        """Stages with an entry in the tuning table take its shape; returns how many did."""
        n = C.c_int()
        _check(load().wos_hip_pipeline_load_tuning(self.h, (path or TUNING_TABLE).encode(), C.byref(n)),
               "pipeline_load_tuning")
        return n.value

    def capture(self):
        _check(load().wos_hip_pipeline_capture(self.h), "pipeline_capture")

    def run(self, iters=1):
        _check(load().wos_hip_pipeline_run(self.h, iters), "pipeline_run")

    def sync(self):
        _check(load().wos_hip_pipeline_sync(self.h), "pipeline_sync")

    def download(self, stage, unit):
        n = self.stages[stage][0]
        out = np.zeros((n, n), dtype=np.complex128)
        _check(load().wos_hip_pipeline_download(self.h, stage, unit, _p(out)), "pipeline_download")
        return out

    def source(self, stage, index):
        n = self.stages[stage][0]
        out = np.zeros((n, n), dtype=np.complex128)
        _check(load().wos_hip_pipeline_download_source(self.h, stage, index, _p(out)), "pipeline_download_source")
        return out

    def stats(self, stage, unit, wl, strehl_aperture=0.0, strehl_distance=0.0):
        st = PlaneStats()
        _check(load().wos_hip_pipeline_stats(self.h, stage, unit, wl, strehl_aperture, strehl_distance,
                                             C.byref(st)), "pipeline_stats")
        return st

    def intensity(self, stage):
        n = self.stages[stage][0]
        out = np.zeros((n, n), dtype=np.float64)
        _check(load().wos_hip_pipeline_intensity(self.h, stage, _p(out)), "pipeline_intensity")
        return out

    def time(self, warmup, iters, stages=False):
        ms = C.c_double()
        st = np.zeros(len(self.stages), dtype=np.float64)
        _check(load().wos_hip_pipeline_time(self.h, warmup, iters, C.byref(ms), _p(st) if stages else None),
               "pipeline_time")
        return ms.value, (st if stages else None)


class BpmMaskOp:
    """Multi-plane phase design through a split-step medium (wos_bpm_mask)."""

    def __init__(self, n, dx, z, steps, index, ctx, target, precision=FP64):
        self.n, self.steps = n, steps
        ix = _c128(index)
        planes = 1 if ix.ndim == 2 else ix.shape[0]
        t = np.ascontiguousarray(target, dtype=np.float64)
        h = C.c_void_p()
        _check(load().wos_hip_bpm_mask_create(C.byref(h), n, dx, z, steps, _p(ix), planes, C.byref(ctx), precision,
                                              _p(t)), "bpm_mask_create")
        self.h = h

    def close(self):
        if self.h:
            _check(load().wos_hip_bpm_mask_destroy(self.h), "bpm_mask_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, theta, phi=None):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ph = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        loss = C.c_double()
        _check(load().wos_hip_bpm_mask_forward(self.h, _p(th), _p(ph), C.byref(loss)), "bpm_mask_forward")
        return loss.value

    def backward(self):
        gt = np.zeros((self.n, self.n))
        gp = np.zeros((self.steps, self.n, self.n))
        _check(load().wos_hip_bpm_mask_backward(self.h, _p(gt), _p(gp)), "bpm_mask_backward")
        return gt, gp

