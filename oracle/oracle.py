"""ctypes view of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
Arrays are numpy complex128 (the memory layout of a Go []complex128),
row-major, index y*N + x.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ASM, ASM_PAD, ASM_SHIFT, ASM_SHIFT_PAD, FRESNEL_TF, FRESNEL_IR, FRAUNHOFER = range(7)
METHODS = {"asm": ASM, "asm_pad": ASM_PAD, "asm_shift": ASM_SHIFT,
           "asm_shift_pad": ASM_SHIFT_PAD, "fresnel_tf": FRESNEL_TF,
           "fresnel_ir": FRESNEL_IR, "fraunhofer": FRAUNHOFER}
WARN_CODES = ["evanescent_filtered", "backward_regularized", "backward_evanescent",
              "asm_alias_wrap", "fresnel_tf_alias", "fresnel_ir_alias", "fraunhofer_nearfield"]


class Ctx(C.Structure):
    """Context fields of the path (optics/field.go:237-254)."""
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


class Stats(C.Structure):
    _fields_ = [("p_in", C.c_double), ("p_out", C.c_double),
                ("fx_c", C.c_double), ("fy_c", C.c_double),
                ("method_used", C.c_int32), ("pad_", C.c_int32)]


class PlaneStats(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("power", "peak", "centroid_x", "centroid_y", "rms_x", "rms_y",
                 "strehl", "intensity_min", "intensity_max", "phase_min", "phase_max")]


def ctx_for(wl, evanescent="decay", limit=0.0, backward_regularize=False, alpha=0.0):
    """ctxFor (optics/propagate_test.go:8-10) plus the other policy fields."""
    return Ctx(wl, 1 if evanescent == "zero" else 0, 1 if backward_regularize else 0,
               limit, alpha)


_lib = None


class Cplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


SHAPE_KINDS = {"circle": 0, "slit": 1, "square": 2, "rectangle": 3, "ellipse": 4, "superellipse": 5, "ring": 6,
               "double_slit": 7, "cross": 8, "polygon": 9}


ELEMENT_KINDS = ("cylindrical_lens", "apodizer", "grating", "axicon", "spiral_phase", "wedge", "zone_plate",
                 "mirror", "concave_mirror", "convex_mirror", "zernike", "kerr", "saturable_absorber")
GRATING_KINDS = ("amplitude_sin", "amplitude_binary", "phase_sin", "phase_binary", "blazed")


class Element(C.Structure):
    """or_element: the parsed parameters of one train element
    (optics/elements.go:156-178)."""
    _fields_ = [("kind", C.c_int32), ("sub", C.c_int32),
                ("f", C.c_double), ("aperture", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
                ("rotation", C.c_double), ("waist", C.c_double), ("amplitude", C.c_double),
                ("alpha", C.c_double), ("index", C.c_double), ("period", C.c_double),
                ("modulation", C.c_double), ("duty", C.c_double), ("blaze_depth", C.c_double),
                ("offset", C.c_double), ("radius", C.c_double), ("reflectivity", C.c_double),
                ("curvature", C.c_double), ("tilt_x", C.c_double), ("tilt_y", C.c_double),
                ("zernike", C.c_double * 22), ("n2", C.c_double), ("tpa", C.c_double), ("length", C.c_double),
                ("alpha0", C.c_double), ("isat", C.c_double)]


def make_element(element, cls=Element, **p):
    """An element descriptor with the reference constructors' defaults
    (optics/elements.go): grating kind by name, zone-plate kind "phase" or
    "amplitude", spiral charge, Zernike coefficients c1..c21 in waves."""
    e = cls()
    e.kind = ELEMENT_KINDS.index(element)
    e.f = p.get("f", {"cylindrical_lens": 0.1, "zone_plate": 0.1}.get(element, 0.0))
    e.aperture, e.x0, e.y0 = p.get("aperture", 0.0), p.get("x", 0.0), p.get("y", 0.0)
    e.rotation = p.get("rotation", 0.0)
    e.waist, e.amplitude = p.get("waist", 0.001), p.get("amplitude", 1.0)
    e.alpha = p.get("alpha", 0.02 if element == "axicon" else 0.01)
    e.index = p.get("index", 1.5)
    e.period, e.modulation, e.duty = p.get("period", 1e-4), p.get("modulation", 1.0), p.get("duty", 0.5)
    e.blaze_depth, e.offset = p.get("blaze_depth", math.pi), p.get("offset", 0.0)
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
    """or_shape: a parsed aperture outline (optics/shape.go:28-48)."""
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


def regular_polygon(n, radius, phi0):
    v = np.zeros((n, 2))
    lib().or_regular_polygon(n, radius, phi0, v.ctypes.data)
    return v


def star(n, outer, inner, phi0):
    v = np.zeros((2 * n, 2))
    lib().or_star(n, outer, inner, phi0, v.ctypes.data)
    return v


def shape_signed_distance(g, dx, dy):
    return lib().or_shape_signed_distance(C.byref(g), dx, dy)


class OField:
    """Field (optics/field.go:9-40) on the host for the element chain: ey is
    always allocated; `polarized` says whether it takes part."""

    def __init__(self, ex, ey=None, dx=1e-6, polarized=None):
        self.ex = np.array(ex, dtype=np.complex128, order="C")
        self.ey = np.zeros_like(self.ex) if ey is None else np.array(ey, dtype=np.complex128, order="C")
        self.n = self.ex.shape[0]
        self.dx = dx
        self.polarized = (ey is not None) if polarized is None else bool(polarized)

    def clone(self):
        return OField(self.ex, self.ey, self.dx, self.polarized)

    def lens(self, wl, f, aperture=0.0, x0=0.0, y0=0.0):
        lib().or_apply_lens_at(_p(self.ex), _p(self.ey), int(self.polarized), self.n, self.dx, wl, f, aperture, x0, y0)

    def aperture(self, g, x0=0.0, y0=0.0):
        lib().or_apply_aperture(_p(self.ex), _p(self.ey), int(self.polarized), self.n, self.dx, C.byref(g), x0, y0)

    def element(self, e, wl):
        rc = lib().or_apply_element(_p(self.ex), _p(self.ey), int(self.polarized), self.n, self.dx, wl, C.byref(e))
        if rc:
            raise ValueError("element parameters rejected")

    def diffuser(self, sigma=np.pi, corr=2e-5, amp=1.0, seed=1, x0=0.0, y0=0.0):
        lib().or_apply_diffuser(_p(self.ex), _p(self.ey), int(self.polarized), self.n, self.dx, sigma, corr, amp,
                                seed, x0, y0)

    def jones(self, a, b, c, d):
        pol = C.c_int(int(self.polarized))
        cx = [Cplx(complex(v).real, complex(v).imag) for v in (a, b, c, d)]
        lib().or_apply_jones(_p(self.ex), _p(self.ey), C.byref(pol), self.n, *cx)
        self.polarized = bool(pol.value)

    def scale(self, s):
        s = complex(s)
        lib().or_scale_amplitude(_p(self.ex), _p(self.ey), int(self.polarized), self.n, Cplx(s.real, s.imag))

    def bandlimit(self, fraction, sigma):
        lib().or_bandlimit(_p(self.ex), _p(self.ey), int(self.polarized), self.n, self.dx, fraction, sigma)

    def anisotropic(self, z, eps, wl):
        e = np.ascontiguousarray(eps, dtype=np.complex128)
        lib().or_propagate_anisotropic(_p(self.ex), _p(self.ey), self.n, self.dx, z, e.ctypes.data, wl)

    def uniaxial(self, z, no, ne, wl):
        lib().or_propagate_uniaxial(_p(self.ex), _p(self.ey) if self.polarized else None, self.n, self.dx, z, no,
                                    ne, wl)

    def propagate(self, z, method, ctx):
        ex, ey, w, st = propagate(self.ex, self.ey if self.polarized else None, self.dx, z, method, ctx)
        self.ex = np.ascontiguousarray(ex)
        if self.polarized:
            self.ey = np.ascontiguousarray(ey)
        if method == "fraunhofer" and z != 0:  # the far-field grid (optics/propagate.go:661)
            self.dx = ctx.wavelength * abs(z) / (self.n * self.dx)
        return w, st


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        D, I = C.c_double, C.c_int
        L.or_fft1d.argtypes = [P, I, I]
        L.or_fft1d_radix2.argtypes = [P, I, I]
        L.or_fft1d_bluestein.argtypes = [P, I, I]
        L.or_fft2d.argtypes = [P, I, I]
        L.or_set_threads.argtypes = [I]
        L.or_get_threads.restype = I
        L.or_freq.argtypes = [I, I, D]
        L.or_freq.restype = D
        L.or_power.argtypes = [P, P, I, D]
        L.or_power.restype = D
        L.or_hypot.argtypes = [D, D]
        L.or_hypot.restype = D
        L.or_propagate.argtypes = [P, P, I, D, D, I, C.POINTER(Ctx), C.POINTER(Warnings),
                                   C.POINTER(Stats)]
        L.or_propagate.restype = I
        L.or_spectral_radius.argtypes = [P, P, I, D, D]
        L.or_spectral_radius.restype = D
        L.or_spectral_centroid.argtypes = [P, P, I, D, C.POINTER(D), C.POINTER(D)]
        L.or_propagate1d.argtypes = [P, I, D, D, D, C.POINTER(Ctx), C.POINTER(Warnings)]
        L.or_propagate1d.restype = I
        L.or_split_step.argtypes = [P, P, I, D, D, I, P, I, C.POINTER(Ctx), C.POINTER(Warnings)]
        L.or_split_step.restype = I
        L.or_polychromatic.argtypes = [P, P, I, D, D, P, P, I, I, C.POINTER(Ctx), P,
                                       C.POINTER(Warnings)]
        L.or_polychromatic.restype = I
        L.or_volume.argtypes = [P, P, I, D, P, I, I, C.POINTER(Ctx), P, P, C.POINTER(Warnings)]
        L.or_volume.restype = I
        L.or_gaussian.argtypes = [P, I, D, D]
        L.or_apply_lens.argtypes = [P, P, I, D, D, D]
        L.or_apply_circle_aperture.argtypes = [P, P, I, D, D]
        L.or_hash_noise01.argtypes = [C.c_int64, C.c_int64, C.c_int64]
        L.or_hash_noise01.restype = D
        L.or_random_phase.argtypes = [P, I, C.c_int64]
        L.or_normal.argtypes = [C.c_uint64, C.c_uint64]
        L.or_normal.restype = D
        L.or_schell_sqrt_spectrum.argtypes = [P, I, D, D]
        L.or_schell_realization.argtypes = [P, P, I, D, D, C.c_uint64, C.c_uint64]
        L.or_compute_stats.argtypes = [P, P, I, D, D, D, D, C.POINTER(PlaneStats)]
        L.or_average_intensity.argtypes = [P, P, I, C.c_size_t, P]
        L.or_propagate_partially_coherent.argtypes = [I, D, D, D, C.c_uint64, I, D, I, C.POINTER(Ctx), P,
                                                      C.POINTER(Warnings)]
        L.or_propagate_partially_coherent.restype = I
        L.or_apply_lens_at.argtypes = [P, P, I, I, D, D, D, D, D, D]
        L.or_apply_element.argtypes = [P, P, I, I, D, D, C.POINTER(Element)]
        L.or_apply_element.restype = I
        L.or_propagate_vectorial.argtypes = [P, P, P, I, D, D, C.POINTER(Ctx)]
        L.or_propagate_uniaxial.argtypes = [P, P, I, D, D, D, D, D]
        L.or_plane_view.argtypes = [P, P, I, I, D, P]
        L.or_propagate_anisotropic.argtypes = [P, P, I, D, D, P, D]
        L.or_shape_signed_distance.argtypes = [C.POINTER(Shape), D, D]
        L.or_shape_signed_distance.restype = D
        L.or_regular_polygon.argtypes = [I, D, D, P]
        L.or_star.argtypes = [I, D, D, D, P]
        L.or_apply_aperture.argtypes = [P, P, I, I, D, C.POINTER(Shape), D, D]
        L.or_apply_diffuser.argtypes = [P, P, I, I, D, D, D, D, C.c_int64, D, D]
        L.or_apply_jones.argtypes = [P, P, C.POINTER(I), I, Cplx, Cplx, Cplx, Cplx]
        L.or_scale_amplitude.argtypes = [P, P, I, I, Cplx]
        L.or_bandlimit.argtypes = [P, P, I, I, D, D, D]
        L.or_mask_loss_grad.argtypes = [P, P, I, D, D, C.POINTER(Ctx), P]
        L.or_mask_loss_grad.restype = D
        L.or_bpm_mask_loss_grad.argtypes = [P, P, P, I, P, I, D, D, I, C.POINTER(Ctx), P, P]
        L.or_bpm_mask_loss_grad.restype = D
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def set_threads(p):
    lib().or_set_threads(int(p))


def fft1d(a, inverse=False):
    a = _c128(a).copy()
    lib().or_fft1d(_p(a), a.size, int(inverse))
    return a


def fft1d_radix2(a, inverse=False):
    a = _c128(a).copy()
    lib().or_fft1d_radix2(_p(a), a.size, int(inverse))
    return a


def fft1d_bluestein(a, inverse=False):
    a = _c128(a).copy()
    lib().or_fft1d_bluestein(_p(a), a.size, int(inverse))
    return a


def fft2d(a, inverse=False):
    a = _c128(a).copy()
    n = a.shape[0]
    assert a.shape == (n, n)
    lib().or_fft2d(_p(a), n, int(inverse))
    return a


def power(ex, ey, dx):
    n = ex.shape[0]
    return lib().or_power(_p(_c128(ex)), _p(None if ey is None else _c128(ey)), n, dx)


def propagate(ex, ey, dx, z, method, ctx):
    """Propagate (optics/propagate.go:65) on copies; returns (ex, ey, warnings, stats)."""
    m = METHODS[method] if isinstance(method, str) else method
    ex = _c128(ex).copy()
    ey = None if ey is None else _c128(ey).copy()
    w, st = Warnings(), Stats()
    rc = lib().or_propagate(_p(ex), _p(ey), ex.shape[0], dx, z, m, C.byref(ctx),
                            C.byref(w), C.byref(st))
    if rc != 0:
        raise ValueError(f"or_propagate rc={rc}")
    return ex, ey, w, st


VIEWS = ["total", "amplitude", "ex", "ey", "phase_x", "phase_y", "pol_intensity", "pol_azimuth", "pol_ellip",
         "pol_s1", "pol_s2", "pol_s3", "pol_degree"]


def plane_view(ex, ey, view, mask_cut=0.0):
    """The server's float32 plane view of one field (ey None: scalar)."""
    ex = _c128(ex)
    ey = None if ey is None else _c128(ey)
    out = np.zeros(ex.shape, dtype=np.float32)
    lib().or_plane_view(_p(ex), _p(ey), ex.shape[0], VIEWS.index(view), mask_cut, out.ctypes.data)
    return out


def propagate_anisotropic(ex, ey, dx, z, eps, wl):
    """PropagateAnisotropic on copies; eps 3x3 complex; ey None: zeros."""
    ex = _c128(ex).copy()
    ey = np.zeros_like(ex) if ey is None else _c128(ey).copy()
    e = np.ascontiguousarray(eps, dtype=np.complex128)
    lib().or_propagate_anisotropic(_p(ex), _p(ey), ex.shape[0], dx, z, e.ctypes.data, wl)
    return ex, ey


def propagate_uniaxial(ex, ey, dx, z, no, ne, wl):
    """PropagateUniaxial on copies (ey None: scalar field)."""
    ex = _c128(ex).copy()
    ey = None if ey is None else _c128(ey).copy()
    lib().or_propagate_uniaxial(_p(ex), _p(ey), ex.shape[0], dx, z, no, ne, wl)
    return ex, ey


def propagate_vectorial(ex, ey, dx, z, ctx):
    """PropagateVectorial on copies; ey None: zeros. Returns (ex, ey, ez)."""
    ex = _c128(ex).copy()
    ey = np.zeros_like(ex) if ey is None else _c128(ey).copy()
    ez = np.zeros_like(ex)
    lib().or_propagate_vectorial(_p(ex), _p(ey), _p(ez), ex.shape[0], dx, z, C.byref(ctx))
    return ex, ey, ez


def propagate1d(a, dx, z, wl, ctx):
    a = _c128(a).copy()
    w = Warnings()
    lib().or_propagate1d(_p(a), a.size, dx, z, wl, C.byref(ctx), C.byref(w))
    return a, w


def split_step(ex, ey, dx, z, steps, index, ctx):
    """index: (n,n) complex (z-invariant) or (steps,n,n) per-slice planes."""
    ex = _c128(ex).copy()
    ey = None if ey is None else _c128(ey).copy()
    index = _c128(index)
    planes = 1 if index.ndim == 2 else index.shape[0]
    w = Warnings()
    rc = lib().or_split_step(_p(ex), _p(ey), ex.shape[0], dx, z, steps, _p(index), planes,
                             C.byref(ctx), C.byref(w))
    if rc != 0:
        raise ValueError(f"or_split_step rc={rc}")
    return ex, ey, w


def polychromatic(ex, ey, dx, z, wavelengths, weights, method, ctx):
    m = METHODS[method] if isinstance(method, str) else method
    ex = _c128(ex)
    ey = None if ey is None else _c128(ey)
    n = ex.shape[0]
    wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
    wt = np.ascontiguousarray(weights, dtype=np.float64)
    out = np.zeros((n, n), dtype=np.float64)
    w = Warnings()
    rc = lib().or_polychromatic(_p(ex), _p(ey), n, dx, z, _p(wl), _p(wt), wl.size, m,
                                C.byref(ctx), _p(out), C.byref(w))
    if rc != 0:
        raise ValueError(f"or_polychromatic rc={rc}")
    return out, w


def volume(ex, ey, dx, zs, method, ctx):
    m = METHODS[method] if isinstance(method, str) else method
    ex = _c128(ex)
    ey = None if ey is None else _c128(ey)
    n = ex.shape[0]
    zs = np.ascontiguousarray(zs, dtype=np.float64)
    ox = np.zeros((zs.size, n, n), dtype=np.complex128)
    oy = None if ey is None else np.zeros((zs.size, n, n), dtype=np.complex128)
    w = Warnings()
    lib().or_volume(_p(ex), _p(ey), n, dx, _p(zs), zs.size, m, C.byref(ctx), _p(ox), _p(oy),
                    C.byref(w))
    return ox, oy, w


def spectral_radius(ex, ey, dx, frac=0.9):
    n = ex.shape[0]
    return lib().or_spectral_radius(_p(_c128(ex)), _p(None if ey is None else _c128(ey)),
                                    n, dx, frac)


def spectral_centroid(ex, ey, dx):
    fx, fy = C.c_double(), C.c_double()
    lib().or_spectral_centroid(_p(_c128(ex)), _p(None if ey is None else _c128(ey)),
                               ex.shape[0], dx, C.byref(fx), C.byref(fy))
    return fx.value, fy.value


def gaussian(n, dx, w0):
    a = np.zeros((n, n), dtype=np.complex128)
    lib().or_gaussian(_p(a), n, dx, w0)
    return a


def apply_lens(ex, ey, dx, wl, f):
    lib().or_apply_lens(_p(ex), _p(ey), ex.shape[0], dx, wl, f)


def apply_circle_aperture(ex, ey, dx, radius):
    lib().or_apply_circle_aperture(_p(ex), _p(ey), ex.shape[0], dx, radius)


def hash_noise01(x, y, seed):
    return lib().or_hash_noise01(x, y, seed)


def random_phase(n, seed):
    a = np.zeros((n, n), dtype=np.complex128)
    lib().or_random_phase(_p(a), n, seed)
    return a


def schell_sqrt_spectrum(n, dx, sigma_c):
    s = np.zeros((n, n), dtype=np.complex128)
    lib().or_schell_sqrt_spectrum(_p(s), n, dx, sigma_c)
    return s


def schell_realization(sqrt_s, dx, w0, seed, r):
    n = sqrt_s.shape[0]
    a = np.zeros((n, n), dtype=np.complex128)
    lib().or_schell_realization(_p(a), _p(_c128(sqrt_s)), n, dx, w0, seed, r)
    return a


def average_intensity(fields_x, fields_y=None):
    """AverageIntensity over m fields given as (m, n, n) arrays."""
    fx = _c128(fields_x)
    fy = None if fields_y is None else _c128(fields_y)
    m, n = fx.shape[0], fx.shape[1]
    out = np.zeros((n, n), dtype=np.float64)
    lib().or_average_intensity(_p(fx), _p(fy), m, n * n, _p(out))
    return out


def propagate_partially_coherent(n, dx, w0, sigma_c, seed, m, z, method, ctx):
    """PropagatePartiallyCoherent with the counter-based Schell generator;
    returns (average intensity, warnings)."""
    mcode = METHODS[method] if isinstance(method, str) else int(method)
    out = np.zeros((n, n), dtype=np.float64)
    w = Warnings()
    rc = lib().or_propagate_partially_coherent(n, dx, w0, sigma_c, seed, m, z, mcode, C.byref(ctx), _p(out),
                                               C.byref(w))
    if rc:
        raise ValueError(f"or_propagate_partially_coherent: rc {rc}")
    return out, w


def compute_stats(ex, ey, dx, wl, ap=0.0, dist=0.0):
    s = PlaneStats()
    lib().or_compute_stats(_p(_c128(ex)), _p(None if ey is None else _c128(ey)), ex.shape[0],
                           dx, wl, ap, dist, C.byref(s))
    return s


def mask_loss_grad(theta, target, dx, z, ctx, want_grad=True):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    n = theta.shape[0]
    g = np.zeros_like(theta) if want_grad else None
    L = lib().or_mask_loss_grad(_p(theta), _p(target), n, dx, z, C.byref(ctx), _p(g))
    return L, g


def bpm_mask_loss_grad(theta, phi, index, target, dx, z, steps, ctx, want_grad=True):
    """Multi-plane phase design through a split-step medium; returns
    (loss, grad_theta, grad_phi)."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    n = theta.shape[0]
    index = _c128(index)
    planes = 1 if index.ndim == 2 else index.shape[0]
    ph = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    gt = np.zeros_like(theta) if want_grad else None
    gp = np.zeros((steps, n, n)) if want_grad else None
    L = lib().or_bpm_mask_loss_grad(_p(theta), _p(ph), _p(index), planes, _p(target), n, dx, z, steps,
                                    C.byref(ctx), _p(gt), _p(gp))
    return L, gt, gp


# ---- deterministic generators restated from the reference's tests ----

class SimpleRNG:
    """simpleRNG (optics/fft_test.go:62-68)."""

    def __init__(self, seed):
        self.seed = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.seed = (self.seed * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return float(self.seed >> 11) / 9007199254740992.0 - 0.5

    def complex_array(self, count):
        out = np.empty(count, dtype=np.complex128)
        for i in range(count):
            re = self.next()
            im = self.next()
            out[i] = complex(re, im)
        return out


def ref_field(n, seed):
    """refField (optics/gpu_test.go:12-22), vectorised."""
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    re = np.sin(seed * (i + 1) * 0.7) + 0.5 * np.cos(seed * (j + 1))
    im = np.cos(seed * (i - j) * 0.3) - 0.25 * np.sin(seed * (i * j + 1))
    return (re + 1j * im).astype(np.complex128)


def rel_l2(a, b):
    """relCmpL2 (optics/complex_field_test.go:15-26)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.sum(np.abs(b) ** 2)
    if den <= 0:
        return 0.0
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / den))
