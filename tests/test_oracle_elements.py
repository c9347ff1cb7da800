"""Pins of the oracle's element-chain restatements (lens, aperture outlines,
diffuser, Jones elements, amplitude scale, band limit) by closed forms,
cross-checks between independent outlines and physical invariants, so that
a wrong sign, offset or formula in one of them fails here."""
import math

import numpy as np
import pytest


def ones(n):
    return np.ones((n, n), dtype=np.complex128)


def mask(oracle, n, dx, g, x0=0.0, y0=0.0):
    f = oracle.OField(ones(n), dx=dx)
    f.aperture(g, x0, y0)
    return f.ex.real


def test_lens_inverse_offset_aperture_and_focus(oracle):
    n, dx, wl = 128, 2e-6, 633e-9
    rng = np.random.default_rng(1)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    f = oracle.OField(a, dx=dx)
    f.lens(wl, 0.05)
    f.lens(wl, -0.05)  # a lens of -f undoes it
    assert np.max(np.abs(f.ex - a)) <= 1e-13
    # an offset lens is the centred lens translated by whole pixels
    c = oracle.OField(ones(n), dx=dx)
    c.lens(wl, 0.02)
    o = oracle.OField(ones(n), dx=dx)
    o.lens(wl, 0.02, x0=3 * dx, y0=-2 * dx)
    # X(i) - 3 dx = X(i - 3), Y(j) + 2 dx = Y(j + 2)
    assert np.max(np.abs(o.ex[:-2, 3:] - c.ex[2:, :-3])) <= 1e-12
    # the aperture zeroes r > radius and leaves the modulus inside
    ap = oracle.OField(ones(n), dx=dx)
    ap.lens(wl, 0.02, aperture=40 * dx)
    x = (np.arange(n) - n / 2) * dx
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    out = r2 > (40 * dx) ** 2
    assert np.all(ap.ex[out] == 0) and np.allclose(np.abs(ap.ex[~out]), 1, atol=1e-15)
    # a converging lens focuses a Gaussian (w0) at z = f to w_f = wl f / (pi w0):
    # peak intensity grows by (w0 / w_f)^2
    n, dx, wl, w0, fl = 256, 2e-6, 0.5e-6, 100e-6, 0.01
    g = oracle.OField(oracle.gaussian(n, dx, w0), dx=dx)
    g.lens(wl, fl)
    g.propagate(fl, "asm", oracle.ctx_for(wl))
    gain = np.max(np.abs(g.ex) ** 2)
    want = (w0 / (wl * fl / (math.pi * w0))) ** 2
    assert abs(gain / want - 1) < 0.02, (gain, want)


def test_aperture_outlines_cross_checks(oracle):
    n, dx = 128, 1e-6
    R = 37.3e-6
    circ = mask(oracle, n, dx, oracle.make_shape("circle", radius=R))
    # the same disk as a ring without hole, an ellipse with a = b, a superellipse of order 2
    assert np.array_equal(circ, mask(oracle, n, dx, oracle.make_shape("ring", rin=0.0, rout=R)))
    assert np.array_equal(circ, mask(oracle, n, dx, oracle.make_shape("ellipse", a=R, b=R)))
    assert np.array_equal(circ, mask(oracle, n, dx, oracle.make_shape("superellipse", a=R, b=R, order=2)))
    # rectangle pixel count: centres at (i - n/2) dx, |x| <= w/2
    w, h = 41.3e-6, 17.7e-6
    rect = mask(oracle, n, dx, oracle.make_shape("rectangle", width=w, height=h))
    assert rect.sum() == (2 * math.floor(w / 2 / dx) + 1) * (2 * math.floor(h / 2 / dx) + 1)
    assert np.array_equal(rect, mask(oracle, n, dx, oracle.make_shape("slit", width=w, height=h)))
    assert np.array_equal(mask(oracle, n, dx, oracle.make_shape("square", width=w)),
                          mask(oracle, n, dx, oracle.make_shape("rectangle", width=w, height=w)))
    # a quarter turn swaps width and height
    assert np.array_equal(rect.T, mask(oracle, n, dx, oracle.make_shape("rectangle", width=w, height=h,
                                                                        rotation=math.pi / 2)))
    # the double slit is two rectangles at +-sep/2 (placed with the aperture offset)
    sep = 30.2e-6
    ds = mask(oracle, n, dx, oracle.make_shape("double_slit", width=w / 4, height=h, separation=sep))
    r1 = mask(oracle, n, dx, oracle.make_shape("rectangle", width=w / 4, height=h), x0=sep / 2)
    r2 = mask(oracle, n, dx, oracle.make_shape("rectangle", width=w / 4, height=h), x0=-sep / 2)
    assert np.array_equal(ds, np.maximum(r1, r2)) and ds.sum() == r1.sum() + r2.sum()
    # the cross is the union of a w x L and an L x w bar
    L = 70.3e-6
    cr = mask(oracle, n, dx, oracle.make_shape("cross", width=w / 3, length=L))
    b1 = mask(oracle, n, dx, oracle.make_shape("rectangle", width=w / 3, height=L))
    b2 = mask(oracle, n, dx, oracle.make_shape("rectangle", width=L, height=w / 3))
    assert np.array_equal(cr, np.maximum(b1, b2))
    # a regular 4-gon rotated by pi/4 is the square of half-width R/sqrt(2)
    sq_poly = mask(oracle, n, dx, oracle.make_shape("polygon", verts=oracle.regular_polygon(4, R, math.pi / 4)))
    assert np.array_equal(sq_poly, mask(oracle, n, dx, oracle.make_shape("square", width=R * math.sqrt(2))))
    # the polygon signed distance is the Euclidean distance to the nearest edge
    g = oracle.make_shape("polygon", verts=oracle.regular_polygon(4, R, math.pi / 4))
    hw = R / math.sqrt(2)
    assert abs(oracle.shape_signed_distance(g, 0.0, 0.0) - hw) < 1e-18
    assert abs(oracle.shape_signed_distance(g, hw + 3e-6, hw + 4e-6) + 5e-6) < 1e-18
    # a star lies between its inner and outer polygons
    st = mask(oracle, n, dx, oracle.make_shape("polygon", verts=oracle.star(5, R, R / 2, math.pi / 2)))
    inner = mask(oracle, n, dx, oracle.make_shape("circle", radius=R / 2 * math.cos(math.pi / 5) * 0.99))
    outer = circ
    assert np.all(st >= inner) and np.all(st <= outer) and inner.sum() < st.sum() < outer.sum()
    # the smooth edge: t(0) = 1/2 and t(d) + t(-d) = 1
    sig = 2e-6
    g = oracle.make_shape("circle", radius=R, edge_sigma=sig)
    t = mask(oracle, n, dx, g)
    x = (np.arange(n) - n / 2) * dx
    d = R - np.hypot(x[None, :], x[:, None])
    want = 0.5 * (1 + np.array([math.erf(v / (math.sqrt(2) * sig)) for v in d.ravel()]).reshape(d.shape))
    assert np.max(np.abs(t - want)) < 1e-14


# SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
# generators", OOPSLA 2014; Vigna's splitmix64.c): published outputs 1..5 of
# the generator started from state 0 and from state 1234567.
SPLITMIX64_FROM_0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC,
                     0x1B39896A51A8749B]
SPLITMIX64_FROM_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                           16408922859458223821]


def test_hash_noise01_is_splitmix64(oracle):
    """hashNoise01 (optics/elements.go:710-721) hashes
    h = seed*G ^ x*P2 ^ y*P5 (mod 2^64) with SplitMix64's output function,
    G = 0x9E3779B97F4A7C15 being SplitMix64's state increment. Output k of
    SplitMix64 from state s0 is that function of s0 + k*G, so inputs that
    make h equal s0 + k*G must reproduce the published outputs in the top 53
    bits (the value is (h >> 11) / 2^53, exact in FP64). The x and y terms
    are reached through the inverses of their odd multipliers, and the
    exclusive-or of the terms through seed and x together."""
    M = 1 << 64
    G, P2, P5 = 0x9E3779B97F4A7C15, 0xC2B2AE3D27D4EB4F, 0x165667B19E3779F9
    inv = lambda a: pow(a, -1, M)
    s64 = lambda v: v - M if v >= 1 << 63 else v

    def top53(x, y, seed):
        v = oracle.hash_noise01(s64(x % M), s64(y % M), s64(seed % M))
        assert 0.0 <= v < 1.0
        return int(v * 2.0 ** 53)

    for k, want in enumerate(SPLITMIX64_FROM_0, start=1):
        assert top53(0, 0, k) == want >> 11, k
        assert top53(k * G * inv(P2), 0, 0) == want >> 11, k
        assert top53(0, k * G * inv(P5), 0) == want >> 11, k
    for k, want in enumerate(SPLITMIX64_FROM_1234567, start=1):
        state = (1234567 + k * G) % M
        assert top53(0, 0, state * inv(G)) == want >> 11, k
        assert top53((state ^ (k * G % M)) * inv(P2), 0, k) == want >> 11, k
        assert top53(0, (state ^ (k * G % M)) * inv(P5), k) == want >> 11, k


def test_diffuser_statistics_anchoring_and_correlation(oracle):
    n, dx = 128, 1e-6
    rng = np.random.default_rng(2)
    a = np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n))) * (1 + rng.random((n, n)))

    def phase(corr, x0=0.0, sigma=0.3, seed=5):
        f = oracle.OField(ones(n), dx=dx)
        f.diffuser(sigma=sigma, corr=corr, amp=1.0, seed=seed, x0=x0)
        return np.angle(f.ex)

    # phase only at amp = 1, amplitude scales by amp
    f = oracle.OField(a, dx=dx)
    f.diffuser(sigma=1.0, corr=0.0, amp=1.0, seed=3)
    assert np.max(np.abs(np.abs(f.ex) - np.abs(a))) <= 1e-14
    f = oracle.OField(a, dx=dx)
    f.diffuser(sigma=1.0, corr=3e-6, amp=0.5, seed=3)
    assert np.max(np.abs(np.abs(f.ex) - 0.5 * np.abs(a))) <= 1e-14
    # the screen is rescaled to zero mean and the requested standard deviation
    for corr in (0.0, 4e-6):
        p = phase(corr)
        assert abs(p.mean()) < 1e-13 and abs(p.std() - 0.3) < 1e-12
    # white noise is uncorrelated between neighbours, the smoothed screen is not
    p0, p1 = phase(0.0), phase(5e-6)
    rho = lambda p: np.corrcoef(p[:, :-1].ravel(), p[:, 1:].ravel())[0, 1]
    assert abs(rho(p0)) < 0.05 and rho(p1) > 0.9
    # the screen belongs to the lab lattice: an element moved by +5 cells shows
    # the cells 5 columns further on, up to the window's own rescaling
    ps = phase(0.0, x0=5 * dx)
    A = np.vstack([p0[:, 5:].ravel(), np.ones(p0[:, 5:].size)]).T
    coef, res, _, _ = np.linalg.lstsq(A, ps[:, :-5].ravel(), rcond=None)
    assert abs(coef[0] - 1) < 0.05 and np.max(np.abs(A @ coef - ps[:, :-5].ravel())) < 1e-12


def jones_polarizer(angle, t=1.0):
    c, s = math.cos(angle), math.sin(angle)
    return t * c * c, t * s * c, t * s * c, t * s * s


def jones_retarder(d, th):
    c, s = math.cos(th), math.sin(th)
    e1, e2 = complex(math.cos(d / 2), math.sin(d / 2)), complex(math.cos(-d / 2), math.sin(-d / 2))
    a = (c * c) * e1 + (s * s) * e2
    b = (s * c) * (e1 - e2)
    dd = (s * s) * e1 + (c * c) * e2
    return a, b, b, dd


def jones_rotator(th):
    c, s = math.cos(th), math.sin(th)
    return c, -s, s, c


def test_jones_elements(oracle):
    n = 32
    rng = np.random.default_rng(4)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    # a scalar matrix keeps a scalar field scalar
    f = oracle.OField(a)
    f.jones(2j, 0, 0, 2j)
    assert not f.polarized and np.max(np.abs(f.ex - 2j * a)) < 1e-14
    # Malus: a polariser at theta passes cos^2 theta of x-polarised power
    for th in (0.3, 1.1):
        f = oracle.OField(a)
        f.jones(*jones_polarizer(th))
        assert f.polarized
        p = np.sum(np.abs(f.ex) ** 2 + np.abs(f.ey) ** 2) / np.sum(np.abs(a) ** 2)
        assert abs(p - math.cos(th) ** 2) < 1e-14
    # a rotator turns x polarisation by theta and keeps the power
    f = oracle.OField(a)
    f.jones(*jones_rotator(0.4))
    assert np.max(np.abs(f.ey - math.tan(0.4) * f.ex)) < 1e-13
    assert abs(np.sum(np.abs(f.ex) ** 2 + np.abs(f.ey) ** 2) / np.sum(np.abs(a) ** 2) - 1) < 1e-14
    # a quarter-wave plate at 45 degrees makes x polarisation circular
    f = oracle.OField(a)
    f.jones(*jones_retarder(math.pi / 2, math.pi / 4))
    assert np.max(np.abs(np.abs(f.ex) - np.abs(f.ey))) < 1e-13
    assert np.max(np.abs(np.abs(np.angle(f.ey / f.ex)) - math.pi / 2)) < 1e-12
    # on a vector field the full 2x2 product
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    f = oracle.OField(a, b)
    m = (1 + 2j, 0.5, -0.25j, 3)
    f.jones(*m)
    assert np.max(np.abs(f.ex - (m[0] * a + m[1] * b))) < 1e-13
    assert np.max(np.abs(f.ey - (m[2] * a + m[3] * b))) < 1e-13
    # amplitude scale on both components of a vector field, Ex only on a scalar one
    f = oracle.OField(a, b)
    f.scale(0.6j)
    assert np.max(np.abs(f.ex - 0.6j * a)) < 1e-14 and np.max(np.abs(f.ey - 0.6j * b)) < 1e-14
    f = oracle.OField(a)
    f.scale(2)
    assert np.array_equal(f.ey, np.zeros_like(a))


def test_bandlimit_closed_forms(oracle):
    n, dx = 128, 1e-6
    x = (np.arange(n) - n / 2) * dx
    # disabled outside (0, 1)
    rng = np.random.default_rng(6)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for fr in (1.0, 0.0, 1.5):
        f = oracle.OField(a, dx=dx)
        f.bandlimit(fr, 0.05)
        assert np.array_equal(f.ex, a)
    # a beam whose spectrum lies inside the cut-off passes unchanged
    # (w = 8 um: the window edge holds exp(-64) of the peak, so no wrap-around tail)
    g = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (8e-6) ** 2).astype(np.complex128)
    f = oracle.OField(g, dx=dx)
    f.bandlimit(0.9, 0.05)
    assert np.max(np.abs(f.ex - g)) < 1e-12
    # a single bin above the cut-off is scaled by exp(-((f - f0)/sigma)^2)
    fnyq = 1 / (2 * dx)
    for k in (58, 60):
        fx = k / (n * dx)
        w = np.exp(2j * np.pi * fx * x)[None, :] * np.ones((n, 1))
        f = oracle.OField(w, dx=dx)
        f.bandlimit(0.8, 0.05)
        d = (fx - 0.8 * fnyq) / (0.05 * fnyq)
        want = np.exp(-d * d) if fx > 0.8 * fnyq else 1.0
        assert np.max(np.abs(f.ex - want * w)) < 1e-12, k
    # a polarised field filters both components; a scalar one leaves Ey alone
    f = oracle.OField(a, 2 * a, dx=dx)
    f.bandlimit(0.5, 0.1)
    assert np.max(np.abs(f.ey - 2 * f.ex)) < 1e-12


def test_fraunhofer_gaussian_far_field(oracle):
    """propFraunhofer of a tilted Gaussian equals the closed-form far field
    U(X, Y) = e^{ikz} (-i/(wl z)) e^{ik(X^2+Y^2)/(2z)} pi w0^2
    exp(-pi^2 w0^2 ((X/(wl z) - f0)^2 + (Y/(wl z))^2)) on the grid of spacing
    wl z / (n dx), and warns only in the near field (Fresnel number > 0.5)."""
    n, dx, wl, w0, z = 256, 2e-6, 0.5e-6, 40e-6, 1.0
    x = (np.arange(n) - n / 2) * dx
    f0 = 3 / (n * dx)
    ex = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / w0 ** 2) * np.exp(2j * np.pi * f0 * x)[None, :]
    f = oracle.OField(ex, dx=dx)
    w, _ = f.propagate(z, "fraunhofer", oracle.ctx_for(wl))
    dxo = wl * z / (n * dx)
    assert abs(f.dx - dxo) <= 1e-18
    X = (np.arange(n) - n / 2) * dxo
    k = 2 * np.pi / wl
    fx, fy = X[None, :] / (wl * z), X[:, None] / (wl * z)
    want = (np.exp(1j * k * z) * (-1j / (wl * z)) * np.exp(1j * k * (X[None, :] ** 2 + X[:, None] ** 2) / (2 * z))
            * np.pi * w0 ** 2 * np.exp(-np.pi ** 2 * w0 ** 2 * ((fx - f0) ** 2 + fy ** 2)))
    err = np.sqrt(np.sum(np.abs(f.ex - want) ** 2) / np.sum(np.abs(want) ** 2))
    assert err < 1e-8, err
    assert not w.as_dict()
    # near field: F = d^2 / (wl z) with d = 2 rms radius = sqrt(2) w0 for this beam
    g = oracle.OField(ex, dx=dx)
    w, _ = g.propagate(1e-3, "fraunhofer", oracle.ctx_for(wl))
    cnt, val = w.as_dict()["fraunhofer_nearfield"]
    assert cnt == 1 and abs(val / (2 * w0 ** 2 / (wl * 1e-3)) - 1) < 1e-6
    # odd grids are refused (the reference's shift leaves a stale column there)
    with pytest.raises(ValueError):
        oracle.propagate(np.ones((9, 9)), None, dx, z, "fraunhofer", oracle.ctx_for(wl))


def test_vectorial_divergence_free_and_transverse(oracle):
    """PropagateVectorial: Ex and Ey propagate like the scalar ASM (up to the
    rounding order of the phase), the output satisfies kx Ex + ky Ey + kz Ez
    = 0 on every bin (kz imaginary on evanescent bins), and a tilted plane
    wave gets Ez = -(kx/kz) Ex."""
    n, dx, wl, z = 64, 0.2e-6, 0.5e-6, 3e-6
    rng = np.random.default_rng(9)
    x = (np.arange(n) - n / 2) * dx
    env = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (3e-6) ** 2)
    ex = env * np.exp(2j * np.pi * rng.random((n, n)))
    ey = 0.5 * env * np.exp(2j * np.pi * rng.random((n, n)))
    ctx = oracle.ctx_for(wl)
    vx, vy, vz = oracle.propagate_vectorial(ex, ey, dx, z, ctx)
    sx, sy, _, _ = oracle.propagate(ex, ey, dx, z, "asm", ctx)
    rel = lambda a, b: np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2))
    assert rel(vx, sx) < 1e-12 and rel(vy, sy) < 1e-12
    f = np.fft.fftfreq(n, dx)
    kx, ky = 2 * np.pi * f[None, :], 2 * np.pi * f[:, None]
    k = 2 * np.pi / wl
    rho2 = (wl * f[None, :]) ** 2 + (wl * f[:, None]) ** 2
    kz = np.where(rho2 <= 1, k * np.sqrt(np.abs(1 - rho2)) + 0j, 1j * k * np.sqrt(np.abs(rho2 - 1)))
    Fx, Fy, Fz = np.fft.fft2(vx), np.fft.fft2(vy), np.fft.fft2(vz)
    div = kx * Fx + ky * Fy + kz * Fz
    scale = np.max(np.abs(kx * Fx) + np.abs(ky * Fy))
    assert np.any(rho2 > 1) and np.max(np.abs(div)) / scale < 1e-12
    # ... evanescent bins included, where Ez carries the decayed field
    ev = rho2 > 1
    assert np.max(np.abs(div[ev])) / np.max(np.abs(kx * Fx)[ev]) < 1e-9
    # a tilted x-polarised plane wave (bin 5): Ez = -(kx / kz) Ex
    fx = 5 / (n * dx)
    w = np.exp(2j * np.pi * fx * x)[None, :] * np.ones((n, 1))
    px, py, pz = oracle.propagate_vectorial(w, None, dx, z, ctx)
    kxx = 2 * np.pi * fx
    kzz = k * np.sqrt(1 - (wl * fx) ** 2)
    assert np.max(np.abs(pz + (kxx / kzz) * px)) < 1e-12 and np.max(np.abs(py)) == 0


def test_uniaxial_closed_forms(oracle):
    """PropagateUniaxial: an isotropic crystal (no = ne = n) is the scalar ASM
    at wavelength wl/n; tilted plane waves pick up exp(i z kz) with the
    extraordinary kz = sqrt(k0^2 ne^2 - (ne/no)^2 kx^2) on Ex and the
    ordinary kz = sqrt(k0^2 no^2 - ky^2) on Ey; a scalar field leaves Ey."""
    n, dx, wl, z = 64, 1e-6, 633e-9, 30e-6
    rng = np.random.default_rng(12)
    x = (np.arange(n) - n / 2) * dx
    env = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (12e-6) ** 2)
    ex = env * np.exp(2j * np.pi * rng.random((n, n)))
    ey = env * np.exp(2j * np.pi * rng.random((n, n)))
    ux, uy = oracle.propagate_uniaxial(ex, ey, dx, z, 1.5, 1.5, wl)
    sx, sy, _, _ = oracle.propagate(ex, ey, dx, z, "asm", oracle.ctx_for(wl / 1.5))
    rel = lambda a, b: np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2))
    assert rel(ux, sx) < 1e-12 and rel(uy, sy) < 1e-12
    no, ne = 1.54, 1.55
    k0 = 2 * np.pi / wl
    fx, fy = 5 / (n * dx), 3 / (n * dx)
    px = np.exp(2j * np.pi * fx * x)[None, :] * np.ones((n, 1))
    py = np.exp(2j * np.pi * fy * x)[:, None] * np.ones((1, n))
    ox, oy = oracle.propagate_uniaxial(px, py, dx, z, no, ne, wl)
    kze = np.sqrt(k0 ** 2 * ne ** 2 - (ne / no) ** 2 * (2 * np.pi * fx) ** 2)
    kzo = np.sqrt(k0 ** 2 * no ** 2 - (2 * np.pi * fy) ** 2)
    assert np.max(np.abs(ox - np.exp(1j * z * kze) * px)) < 1e-11
    assert np.max(np.abs(oy - np.exp(1j * z * kzo) * py)) < 1e-11
    sx2, none = oracle.propagate_uniaxial(px, None, dx, z, no, ne, wl)
    assert none is None and np.array_equal(sx2, ox)


def test_plane_views_closed_forms(oracle):
    """The float32 plane views: intensity family and amplitude, masked
    wrapped phase, and the Stokes views for known polarisation states
    (linear at theta: s1 = cos 2theta, s2 = sin 2theta, azimuth theta; circular:
    s3 = +-1, ellipticity +-pi/4; degree 1 for a coherent field)."""
    n = 32
    x = (np.arange(n) - n / 2) / n
    env = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / 0.02)
    ph = np.exp(1j * 3 * x)[None, :] * np.ones((n, 1))
    th = 0.3
    ex, ey = env * np.cos(th) * ph, env * np.sin(th) * ph
    tot = oracle.plane_view(ex, ey, "total")
    want = (np.abs(ex) ** 2 + np.abs(ey) ** 2).astype(np.float32)
    assert np.max(np.abs(tot - want) / want.max()) < 1e-7
    assert np.allclose(oracle.plane_view(ex, ey, "amplitude") ** 2, tot, rtol=1e-6, atol=0)
    assert np.array_equal(oracle.plane_view(ex, None, "ey"), np.zeros((n, n), np.float32))
    assert np.array_equal(oracle.plane_view(ex, ey, "pol_intensity"), tot)
    # phase: masked (NaN, bits 0x7FC00000) below 2e-3 of the component's peak
    p = oracle.plane_view(ex, ey, "phase_x")
    keep = np.abs(ex) ** 2 >= 2e-3 * np.max(np.abs(ex) ** 2)
    assert np.all(np.isnan(p[~keep])) and np.all(p[~keep].view(np.uint32) == 0x7FC00000)
    assert np.max(np.abs(p[keep] - np.angle(ex[keep]))) < 1e-6
    assert np.all(np.isnan(oracle.plane_view(ex, None, "phase_y")))  # no Ey: no phase
    lit = np.abs(ex) ** 2 + np.abs(ey) ** 2 >= 1e-6 * np.max(np.abs(ex) ** 2 + np.abs(ey) ** 2)
    for view, v in (("pol_s1", np.cos(2 * th)), ("pol_s2", np.sin(2 * th)), ("pol_s3", 0.0),
                    ("pol_azimuth", th), ("pol_ellip", 0.0), ("pol_degree", 1.0)):
        got = oracle.plane_view(ex, ey, view)
        assert np.all(np.isnan(got[~lit])) and np.max(np.abs(got[lit] - v)) < 1e-6, view
    # circular light (Ey = +-i Ex): s3 = +-1, ellipticity +-pi/4
    for s in (1, -1):
        got3 = oracle.plane_view(env * ph, s * 1j * env * ph, "pol_s3")
        gote = oracle.plane_view(env * ph, s * 1j * env * ph, "pol_ellip")
        assert np.max(np.abs(got3[lit] - s)) < 1e-6 and np.max(np.abs(gote[lit] - s * np.pi / 4)) < 1e-6


def test_berreman_reductions(oracle):
    """PropagateAnisotropic: a uniaxial tensor diag(ne^2, no^2, no^2) on a
    field that varies along x only is PropagateUniaxial (exact in the x-z
    plane); a biaxial tensor at normal incidence gives exp(i k0 nx z) on Ex
    and exp(i k0 ny z) on Ey. (An isotropic tensor has degenerate modes: the
    adjugate null vectors vanish and the reference's 2x2 solve gives NaN, so
    it is no check here.)"""
    n, dx, wl, z = 32, 1e-6, 633e-9, 20e-6
    rng = np.random.default_rng(13)
    x = (np.arange(n) - n / 2) * dx
    rel = lambda a, b: np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2))
    no, ne = 1.658, 1.486
    line = np.exp(-(x / 6e-6) ** 2) * np.exp(2j * np.pi * rng.random(n))
    fx1, fy1 = line[None, :] * np.ones((n, 1)), 0.5 * line[None, :] * np.ones((n, 1))
    bx, by = oracle.propagate_anisotropic(fx1, fy1, dx, z, np.diag([ne ** 2, no ** 2, no ** 2]), wl)
    ux, uy = oracle.propagate_uniaxial(fx1, fy1, dx, z, no, ne, wl)
    assert rel(bx, ux) < 1e-10 and rel(by, uy) < 1e-10
    nx, ny, nz = 1.6, 1.5, 1.4
    k0 = 2 * np.pi / wl
    bx, by = oracle.propagate_anisotropic(np.ones((n, n)), 2 * np.ones((n, n)), dx, z,
                                          np.diag([nx ** 2, ny ** 2, nz ** 2]), wl)
    assert np.max(np.abs(bx - np.exp(1j * k0 * nx * z))) < 1e-12
    assert np.max(np.abs(by - 2 * np.exp(1j * k0 * ny * z))) < 1e-12


# ---------------------------------------------------------------- further elements

def _ones(oracle, n=96, dx=2e-6):
    return oracle.OField(np.ones((n, n), dtype=complex), dx=dx)


def test_element_relations_pin_the_restatement(oracle):
    """Relations the optics fixes: crossed cylindrical lenses (rotation 0 and
    pi/2) make the spherical lens; a mirror of curvature k is a lens of focal
    length 1/(2k); a concave mirror of radius R is a lens of f = R/2 and the
    convex one its conjugate; reflectivity scales the amplitude."""
    wl, n, dx = 633e-9, 96, 2e-6
    lens = _ones(oracle, n, dx)
    lens.lens(wl, 0.04)
    cyl = _ones(oracle, n, dx)
    cyl.element(oracle.make_element("cylindrical_lens", f=0.04), wl)
    cyl.element(oracle.make_element("cylindrical_lens", f=0.04, rotation=math.pi / 2), wl)
    assert np.max(np.abs(cyl.ex - lens.ex)) <= 1e-12
    mir = _ones(oracle, n, dx)
    mir.element(oracle.make_element("mirror", curvature=1 / (2 * 0.04), reflectivity=0.8), wl)
    assert np.max(np.abs(mir.ex - 0.8 * lens.ex)) <= 1e-12
    cc = _ones(oracle, n, dx)
    cc.element(oracle.make_element("concave_mirror", radius=0.08), wl)
    assert np.max(np.abs(cc.ex - lens.ex)) <= 1e-12
    cv = _ones(oracle, n, dx)
    cv.element(oracle.make_element("convex_mirror", radius=0.08, reflectivity=0.5), wl)
    assert np.max(np.abs(cv.ex - 0.5 * np.conj(lens.ex))) <= 1e-12


def test_element_closed_forms(oracle):
    """Tilted mirror: exp(-2ik(tx x + ty y)); wedge and axicon: linear phase
    -k (n-1) tan(alpha) along the wedge axis / the radius; spiral phase plate:
    charge times the azimuth; apodizer: amp exp(-r^2/w^2)."""
    wl, n, dx = 633e-9, 96, 2e-6
    k = 2 * math.pi / wl
    x = (np.arange(n) - n / 2) * dx
    X, Y = x[None, :], x[:, None]
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("mirror", tilt_x=3e-4, tilt_y=-2e-4), wl)
    assert np.max(np.abs(f.ex - np.exp(-2j * k * (3e-4 * X - 2e-4 * Y)))) <= 1e-12
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("wedge", alpha=0.004, index=1.45, rotation=0.7), wl)
    c = -k * 0.45 * math.tan(0.004)
    assert np.max(np.abs(f.ex - np.exp(1j * c * (math.cos(0.7) * X + math.sin(0.7) * Y)))) <= 1e-12
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("axicon", alpha=0.01, index=1.5), wl)
    c = -k * 0.5 * math.tan(0.01)
    assert np.max(np.abs(f.ex - np.exp(1j * c * np.hypot(X, Y)))) <= 1e-12
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("spiral_phase", charge=3, x=4e-6), wl)
    assert np.max(np.abs(f.ex - np.exp(3j * np.arctan2(Y + 0 * X, X - 4e-6)))) <= 1e-12
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("apodizer", waist=50e-6, amplitude=0.7), wl)
    assert np.max(np.abs(f.ex - 0.7 * np.exp(-(X ** 2 + Y ** 2) / 50e-6 ** 2))) <= 1e-15


def test_grating_kinds(oracle):
    """amplitude_sin: 0.5 (1 + m cos(2 pi u / period + offset)); phase_sin:
    exp(i m/2 sin(...)); blazed: exp(i depth frac(u / period)); the binary
    kinds take two values, open on a fraction of the period set by the duty
    cycle (amplitude) or by the sign of the cosine (phase)."""
    wl, n, dx = 633e-9, 128, 1e-6
    x = (np.arange(n) - n / 2) * dx
    X, Y = x[None, :], x[:, None]
    per, rot, off = 23.7e-6, 0.35, 0.2
    u = math.cos(rot) * X + math.sin(rot) * Y
    ph = 2 * math.pi * u / per + off
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("grating", kind="amplitude_sin", period=per, modulation=0.6, rotation=rot,
                                  offset=off), wl)
    assert np.max(np.abs(f.ex - 0.5 * (1 + 0.6 * np.cos(ph)))) <= 1e-12
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("grating", kind="phase_sin", period=per, modulation=1.3, rotation=rot,
                                  offset=off), wl)
    assert np.max(np.abs(f.ex - np.exp(1j * 0.65 * np.sin(ph)))) <= 1e-12
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("grating", kind="blazed", period=per, blaze_depth=2.1, rotation=rot), wl)
    assert np.max(np.abs(f.ex - np.exp(1j * 2.1 * np.mod(u / per, 1.0)))) <= 1e-12
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("grating", kind="amplitude_binary", period=per, duty=0.3, rotation=rot), wl)
    assert set(np.unique(f.ex.real)) <= {0.0, 1.0} and abs(np.mean(f.ex.real) - 0.3) < 0.02
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("grating", kind="phase_binary", period=per, modulation=0.9, rotation=rot), wl)
    vals = np.unique(np.round(f.ex, 12))
    assert len(vals) == 2 and abs(np.mean(np.abs(f.ex - 1) < 1e-12) - 0.5) < 0.03


def test_zone_plate_zone_radii(oracle):
    """Zones of the exact point-to-point zone plate, r_m^2 = m wl f + (m wl/2)^2:
    the reference rounds the zone index m = floor(2 (sqrt(f^2 + r^2) - f)/wl
    + 1/2) (optics/elements.go:662), so zone m spans the radii of m - 1/2 to
    m + 1/2 in that formula. The amplitude plate is clear in even zones and
    opaque in odd ones, the phase plate shifts odd zones by pi; nothing
    changes outside `radius`."""
    wl, n, dx, fz, R = 633e-9, 128, 1e-6, 2.13e-3, 55e-6
    x = (np.arange(n) - n / 2) * dx
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    edges = np.array([(m + 0.5) * wl * fz + ((m + 0.5) * wl / 2) ** 2 for m in range(12)])
    zone = np.searchsorted(edges, r2, side="right")
    near = np.min(np.abs(r2[..., None] - edges[None, None, :]), axis=-1) < 1e-12   # boundary pixels left out
    inside = (r2 <= R * R) & ~near
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("zone_plate", f=fz, radius=R, kind="amplitude"), wl)
    assert np.array_equal(f.ex.real[inside], (zone[inside] % 2 == 0).astype(float))
    assert np.all(f.ex[r2 > R * R] == 1)
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("zone_plate", f=fz, radius=R, kind="phase"), wl)
    assert np.max(np.abs(f.ex[inside] - np.where(zone[inside] % 2 == 0, 1, -1))) <= 1e-15


def test_zernike_defocus_and_coma(oracle):
    """Noll Z4 = sqrt(3) (2 rho^2 - 1) and Z7 = sqrt(8) (3 rho^3 - 2 rho) sin(theta)
    (coefficients in waves); outside the normalisation radius nothing changes."""
    wl, n, dx, R = 633e-9, 96, 2e-6, 70e-6
    x = (np.arange(n) - n / 2) * dx
    X, Y = x[None, :], x[:, None]
    rho, th = np.hypot(X, Y) / R, np.arctan2(Y, X)
    f = _ones(oracle, n, dx)
    f.element(oracle.make_element("zernike", radius=R, c4=0.3, c7=-0.2), wl)
    opd = 0.3 * math.sqrt(3) * (2 * rho ** 2 - 1) - 0.2 * math.sqrt(8) * (3 * rho ** 3 - 2 * rho) * np.sin(th)
    inside = rho <= 1
    assert np.max(np.abs(f.ex[inside] - np.exp(2j * math.pi * opd[inside]))) <= 1e-12
    assert np.all(f.ex[~inside] == 1)


def test_element_rejections(oracle):
    f = _ones(oracle, 16)
    for kind, p in (("apodizer", dict(waist=0)), ("grating", dict(period=0)), ("grating", dict(duty=1.5)),
                    ("zone_plate", dict(f=-1)), ("mirror", dict(reflectivity=1.5)), ("concave_mirror", dict(radius=0)),
                    ("zernike", dict(radius=0)), ("cylindrical_lens", dict(f=0))):
        with pytest.raises(ValueError):
            f.element(oracle.make_element(kind, **p), 633e-9)


def test_kerr_and_saturable_absorber(oracle):
    """Kerr without absorption only rotates the phase, by k n2 I L, and leaves
    zero-intensity pixels alone; two-photon absorption scales the amplitude
    by exp(-tpa I L / 2). The saturable absorber transmits exp(-alpha0 L / 2)
    in the amplitude at I = 0 and tends to 1 far above I_sat; I sums both
    components of a polarised field."""
    wl, n, dx = 633e-9, 32, 1e-6
    k = 2 * math.pi / wl
    rng = np.random.default_rng(3)
    ex = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    ex[0, :4] = 0
    ey = 0.5 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    ey[0, :4] = 0
    I = np.abs(ex) ** 2 + np.abs(ey) ** 2
    f = oracle.OField(ex.copy(), ey.copy(), dx=dx)
    f.element(oracle.make_element("kerr", n2=1e-7, length=2e-3), wl)
    assert np.allclose(np.abs(f.ex), np.abs(ex), rtol=1e-14, atol=0)
    assert np.max(np.abs(f.ex - ex * np.exp(1j * k * 1e-7 * I * 2e-3))) <= 1e-12 * np.max(np.abs(ex))
    assert np.array_equal(f.ex[0, :4], ex[0, :4])
    f = oracle.OField(ex.copy(), ey.copy(), dx=dx)
    f.element(oracle.make_element("kerr", n2=0.0, tpa=300.0, length=2e-3), wl)
    assert np.allclose(f.ey, ey * np.exp(-0.5 * 300.0 * I * 2e-3), rtol=1e-13, atol=0)
    f = oracle.OField(ex.copy(), ey.copy(), dx=dx)
    f.element(oracle.make_element("saturable_absorber", alpha0=500.0, isat=1e-12, length=2e-3), wl)
    bright = I > 1e-6
    assert np.allclose(f.ex[bright], ex[bright], rtol=1e-5)        # I >> I_sat: transparent
    assert np.allclose(f.ex[0, :4], 0)
    g = oracle.OField(np.zeros((n, n), dtype=complex) + 1e-9, dx=dx)
    g.element(oracle.make_element("saturable_absorber", alpha0=500.0, isat=1e6, length=2e-3), wl)
    assert np.allclose(g.ex, 1e-9 * math.exp(-500.0 * 2e-3 / 2), rtol=1e-12)   # I << I_sat

