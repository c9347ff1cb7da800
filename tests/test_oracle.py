"""Pins the FP64 CPU oracle before it is trusted as the checker.

Each test restates a known answer the reference's own tests hold (cited), or
a fixture computed by an independent library (tests/golden, numpy.fft), or a
closed form. They are chosen so plausible mistakes fail one of them: a wrong
sign of z or of the DFT exponent (plane-wave phase, tilted wave forward and
backward), transposed fx/fy (y-tilted wave), a dropped 1/n (round trips,
Parseval), a wrong evanescent form (decay constant), a wrong screen
(uniform / absorbing / stratified BPM), a wrong incoherent sum (polychromatic
uniform intensity) and a wrong gradient (central finite differences).
"""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "fft_golden.npz"))


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


# ---------------------------------------------------------------- FFT core

@pytest.mark.parametrize("n", [8, 16, 32, 64, 12, 27, 100])
def test_fft2d_matches_golden(oracle, golden, n):
    """fft2DCPU vs numpy on refField (optics/gpu_test.go:44-75 sizes)."""
    got = oracle.fft2d(oracle.ref_field(n, 1.3))
    tol = 1e-12 if (n & (n - 1)) == 0 else 1e-11
    assert rel_max(got, golden[f"fft2_refField_{n}"]) <= tol


@pytest.mark.parametrize("n", [1024, 4096, 1000, 2049])
def test_fft1d_matches_golden(oracle, golden, n):
    i = np.arange(n)
    got = oracle.fft1d(np.sin(i * 0.01) + 1j * np.cos(i * 0.017))
    assert rel_max(got, golden[f"fft1_signal_{n}"]) <= 1e-11


def test_bluestein_matches_naive_dft(oracle, golden):
    """TestBluesteinMatchesNaiveDFT (optics/asm_improve_test.go:23-38): 1e-9,
    and the brute-force DFT itself (optics/asm_improve_test.go:10-21)."""
    rng = oracle.SimpleRNG(99)
    for n in (100, 200, 480):
        a = rng.complex_array(n)
        k = np.arange(n)
        naive = np.exp(-2j * np.pi * np.outer(k, k) / n) @ a
        got = oracle.fft1d(a)
        assert np.max(np.abs(got - naive)) <= 1e-9
        assert np.max(np.abs(got - golden[f"fft1_rng99_{n}"])) <= 1e-9


def test_fft_round_trip_impulse_parseval(oracle):
    """TestFFTRoundTrip / TestFFTImpulse / TestFFTParseval (optics/fft_test.go:10-60)."""
    rng = oracle.SimpleRNG(42)
    a = rng.complex_array(64 * 64).reshape(64, 64)
    back = oracle.fft2d(oracle.fft2d(a), inverse=True)
    assert np.max(np.abs(back - a)) <= 1e-12
    imp = np.zeros((32, 32), complex)
    imp[0, 0] = 1
    assert np.max(np.abs(oracle.fft2d(imp) - 1)) <= 1e-12
    rng = oracle.SimpleRNG(7)
    a = rng.complex_array(128 * 128).reshape(128, 128)
    p1 = np.sum(np.abs(a) ** 2)
    p2 = np.sum(np.abs(oracle.fft2d(a)) ** 2)
    assert abs(p2 - p1 * 128 * 128) / (p1 * 128 * 128) <= 1e-12


def test_bluestein_round_trip_and_2d_non_pow2(oracle):
    """TestBluesteinRoundTrip, TestFFT2DNonPowerOfTwo (optics/asm_improve_test.go:40-83)."""
    rng = oracle.SimpleRNG(5)
    a = rng.complex_array(100)
    assert np.max(np.abs(oracle.fft1d(oracle.fft1d(a), inverse=True) - a)) <= 1e-12
    rng = oracle.SimpleRNG(11)
    b = rng.complex_array(100 * 100).reshape(100, 100)
    assert np.max(np.abs(oracle.fft2d(oracle.fft2d(b), inverse=True) - b)) <= 1e-11
    imp = np.zeros((100, 100), complex)
    imp[0, 0] = 1
    assert np.max(np.abs(oracle.fft2d(imp) - 1)) <= 1e-11


def test_hypot_and_freq(oracle):
    L = oracle.lib()
    assert L.or_hypot(3.0, 4.0) == 5.0
    assert L.or_hypot(0.0, 0.0) == 0.0
    assert L.or_hypot(-1e300, 1e300) == pytest.approx(math.sqrt(2) * 1e300)
    n, dx = 8, 0.5
    # freq (optics/field.go:297-303): the Nyquist bin i = n/2 is positive
    want = [i / (n * dx) if i <= n // 2 else (i - n) / (n * dx) for i in range(n)]
    assert [L.or_freq(i, n, dx) for i in range(n)] == want


# ---------------------------------------------------------------- propagators

def ctx(oracle, wl, **kw):
    return oracle.ctx_for(wl, **kw)


def test_plane_wave_phase(oracle):
    """TestPlaneWavePhase (optics/propagate_test.go:13-35)."""
    n, width, wl, z = 256, 0.01, 632.8e-9, 0.3
    ex = np.ones((n, n), complex)
    out, _, _, _ = oracle.propagate(ex, None, width / n, z, "asm", ctx(oracle, wl))
    c = out[n // 2, n // 2]
    got = math.remainder(math.atan2(c.imag, c.real) - 2 * math.pi / wl * z, 2 * math.pi)
    assert abs(got) <= 1e-6
    inten = np.abs(out) ** 2
    assert np.max(np.abs(inten - inten[0, 0])) <= 1e-12 * inten[0, 0]


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("sx,px", [(0.25, 30), (0.9, 45)])
def test_tilted_plane_wave_single_bin_exact(oracle, axis, sx, px):
    """TestTiltedPlaneWaveSingleBinExact (optics/complex_field_test.go:46-82),
    plus the y-tilted mirror case (catches a transposed fx/fy)."""
    wl, n = 632.8e-9, 256
    k = 2 * math.pi / wl
    dx = px / (sx / wl * n)
    z = 5e-4
    kz = k * math.sqrt(1 - sx * sx)
    x = (np.arange(n) - n / 2) * dx
    ramp = np.exp(1j * k * sx * x)
    ex = np.tile(ramp, (n, 1)) if axis == "x" else np.tile(ramp[:, None], (1, n))
    for method in ("asm", "asm_shift"):
        for zz in (z, -z):
            out, _, _, _ = oracle.propagate(ex, None, dx, zz, method, ctx(oracle, wl))
            want = ex * np.exp(1j * kz * zz)
            assert np.max(np.abs(out - want)) <= 1e-8, (method, zz)


def analytic_gaussian_fresnel(n, dx, wl, z, w0):
    """analyticGaussianFresnel (optics/complex_field_test.go:84-109)."""
    x = (np.arange(n) - n / 2) * dx
    beta = 1 / (w0 * w0) - 1j * math.pi / (wl * z)
    gamma = 2 * math.pi * x / (wl * z)
    G = np.sqrt(math.pi / beta) * np.exp(1j * math.pi * x * x / (wl * z)) * np.exp(-gamma ** 2 / (4 * beta))
    pre = np.exp(1j * 2 * math.pi / wl * z) / (1j * wl * z)
    return pre * np.outer(G, G)


def test_gaussian_matches_analytic_fresnel(oracle):
    """TestASMGaussianMatchesAnalyticComplex (optics/complex_field_test.go:114-138)."""
    wl, n, width, w0 = 632.8e-9, 256, 0.004, 5e-4
    dx = width / n
    ex = oracle.gaussian(n, dx, w0)
    for z in (0.005, 0.05, 0.3):
        want = analytic_gaussian_fresnel(n, dx, wl, z, w0)
        for m in ("asm", "asm_pad", "fresnel_tf"):
            out, _, _, _ = oracle.propagate(ex, None, dx, z, m, ctx(oracle, wl))
            assert oracle.rel_l2(out, want) <= 2e-3, (z, m)
    # fresnel_ir in its own validity region |z| >= N dx^2 / wl
    z = 0.3
    out, _, w, _ = oracle.propagate(ex, None, dx, z, "fresnel_ir", ctx(oracle, wl))
    assert oracle.rel_l2(out, analytic_gaussian_fresnel(n, dx, wl, z, w0)) <= 2e-3
    assert "fresnel_ir_alias" not in w.as_dict()


def test_variants_agree_with_asm(oracle):
    """TestASMPadShiftVariantsComplexAgree (optics/complex_field_test.go:222-246)."""
    n, width, wl, z = 256, 0.01, 632.8e-9, 0.1
    dx = width / n
    x = (np.arange(n) - n / 2) * dx
    ex = oracle.gaussian(n, dx, 1e-3) * np.exp(1j * 2 * math.pi / wl * math.sin(0.002) * x)[None, :]
    ref, _, _, _ = oracle.propagate(ex, None, dx, z, "asm", ctx(oracle, wl))
    for m in ("asm_pad", "asm_shift", "asm_shift_pad"):
        out, _, _, _ = oracle.propagate(ex, None, dx, z, m, ctx(oracle, wl))
        assert oracle.rel_l2(out, ref) <= 1e-6, m


def evanescent_field(n, dx, q):
    """evanescentField (optics/asm_improve_test.go:169-178)."""
    f0 = q / (n * dx)
    x = (np.arange(n) - n / 2) * dx
    return np.tile(np.exp(1j * 2 * math.pi * f0 * x), (n, 1))


def test_evanescent_decay_limit_and_backward(oracle):
    """TestEvanescentLimitTruncation, TestBackwardTikhonov (optics/asm_improve_test.go:182-253)."""
    n, dx, wl, q, z = 256, 0.25e-6, 1e-6, 96, 1e-6
    f0 = q / (n * dx)
    decay = (2 * math.pi / wl) * z * math.sqrt((wl * f0) ** 2 - 1)
    out, _, w, _ = oracle.propagate(evanescent_field(n, dx, q), None, dx, z, "asm", ctx(oracle, wl))
    assert abs(abs(out[0, 0]) - math.exp(-decay)) <= 1e-6
    assert "evanescent_filtered" in w.as_dict()
    out, _, _, _ = oracle.propagate(evanescent_field(n, dx, q), None, dx, z, "asm", ctx(oracle, wl, limit=1.0))
    assert oracle.power(out, None, dx) <= 1e-20
    # backward: default zeroes, Tikhonov keeps a finite positive value
    out, _, w, _ = oracle.propagate(evanescent_field(n, dx, q), None, dx, -z, "asm", ctx(oracle, wl))
    assert oracle.power(out, None, dx) <= 1e-20 and "backward_evanescent" in w.as_dict()
    out, _, w, _ = oracle.propagate(evanescent_field(n, dx, q), None, dx, -z, "asm",
                                    ctx(oracle, wl, backward_regularize=True, alpha=1e-3))
    assert np.isfinite(out).all() and oracle.power(out, None, dx) > 0
    assert "backward_regularized" in w.as_dict()
    # propagating content round-trips through the regularised inverse
    n2, width, wl2, z2 = 256, 0.01, 632.8e-9, 0.05
    g = oracle.gaussian(n2, width / n2, 1e-3)
    fwd, _, _, _ = oracle.propagate(g, None, width / n2, z2, "asm", ctx(oracle, wl2))
    back, _, _, _ = oracle.propagate(fwd, None, width / n2, -z2, "asm",
                                     ctx(oracle, wl2, backward_regularize=True, alpha=1e-3))
    assert oracle.rel_l2(back, g) <= 1e-6


def test_power_conservation(oracle):
    """TestPowerConservationASM (optics/propagate_test.go:95-112)."""
    n, width, wl = 512, 0.005, 532e-9
    dx = width / n
    g = oracle.gaussian(n, dx, 5e-4)
    p0 = oracle.power(g, None, dx)
    for z in (0.01, 0.1, 0.5, 1.5):
        out, _, _, st = oracle.propagate(g, None, dx, z, "asm", ctx(oracle, wl))
        assert abs(oracle.power(out, None, dx) - p0) / p0 <= 1e-6
        assert st.p_in == pytest.approx(p0, rel=1e-15)


def test_wrap_warning(oracle):
    """TestASMWrapWarning (optics/asm_wrap_test.go:14-83), restated on a
    plane wave through a sinusoidal phase grating (period 20 um, 1024 px over
    8 mm): the orders walk out of the window at z = 1 m, not at 10 mm."""
    n, width, wl = 1024, 0.008, 632.8e-9
    dx = width / n
    x = (np.arange(n) - n / 2) * dx
    ex = np.tile(np.exp(1j * 2.0 * np.sin(2 * math.pi * x / 2e-5)), (n, 1))
    _, _, w, _ = oracle.propagate(ex, None, dx, 1.0, "asm", ctx(oracle, wl))
    d = w.as_dict()
    assert "asm_alias_wrap" in d and d["asm_alias_wrap"][1] >= 2
    _, _, w, _ = oracle.propagate(ex, None, dx, 0.01, "asm", ctx(oracle, wl))
    assert "asm_alias_wrap" not in w.as_dict()


def test_propagate1d_plane_wave(oracle):
    """TestPropagate1DPlaneWave (optics/propagate1d_test.go:8-35)."""
    n, dx, wl, z = 256, 1e-5, 632.8e-9, 0.3
    out, _ = oracle.propagate1d(np.ones(n, complex), dx, z, wl, ctx(oracle, wl))
    got = math.remainder(math.atan2(out[n // 2].imag, out[n // 2].real) - 2 * math.pi / wl * z, 2 * math.pi)
    assert abs(got) <= 1e-6
    assert abs(np.sum(np.abs(out) ** 2) - n) / n <= 1e-9


# ---------------------------------------------------------------- BPM

def test_split_step_uniform_absorbing_stratified(oracle):
    """TestSplitStepUniformIndex / Absorption / Stratified (optics/medium_test.go:8-66)."""
    n, width, wl = 256, 0.01, 632.8e-9
    dx = width / n
    k0 = 2 * math.pi / wl
    ex = np.ones((n, n), complex)
    out, _, _ = oracle.split_step(ex, None, dx, 0.1, 20, np.full((n, n), 1.5 + 0j), ctx(oracle, wl))
    c = out[n // 2, n // 2]
    assert abs(math.remainder(math.atan2(c.imag, c.real) - k0 * 1.5 * 0.1, 2 * math.pi)) <= 1e-6
    assert abs(np.sum(np.abs(out) ** 2) - n * n) / (n * n) <= 1e-9
    wl2 = 1e-6
    k2 = 2 * math.pi / wl2
    out, _, _ = oracle.split_step(ex, None, dx, 1e-3, 10, np.full((n, n), 1 + 1e-4j), ctx(oracle, wl2))
    want = n * n * math.exp(-2 * k2 * 1e-4 * 1e-3)
    assert abs(np.sum(np.abs(out) ** 2) - want) / want <= 1e-3
    steps, z = 20, 0.1
    dz = z / steps
    planes = np.array([np.full((n, n), (1.5 if (z * s / steps + dz / 2) < 0.05 else 1.0) + 0j)
                       for s in range(steps)])
    out, _, _ = oracle.split_step(ex, None, dx, z, steps, planes, ctx(oracle, wl))
    c = out[n // 2, n // 2]
    assert abs(math.remainder(math.atan2(c.imag, c.real) - k0 * 0.125, 2 * math.pi)) <= 1e-6


# ---------------------------------------------------------------- batching operators

def test_polychromatic_uniform_and_single(oracle):
    """TestPolychromaticPlaneWave / SingleMatchesMono (optics/polychromatic_test.go:8-45)."""
    n, width = 256, 0.01
    dx = width / n
    ex = np.ones((n, n), complex)
    out, _ = oracle.polychromatic(ex, None, dx, 0.05, [632.8e-9, 532e-9], [1, 0.5], "asm",
                                  ctx(oracle, 632.8e-9))
    assert np.max(np.abs(out - 1.5)) <= 1e-9 * 1.5
    g = oracle.gaussian(n, dx, 1e-3)
    mono, _, _, _ = oracle.propagate(g, None, dx, 0.1, "asm", ctx(oracle, 632.8e-9))
    out, _ = oracle.polychromatic(g, None, dx, 0.1, [632.8e-9], [1.0], "asm", ctx(oracle, 632.8e-9))
    assert np.max(np.abs(out - np.abs(mono) ** 2)) <= 1e-9 * np.max(np.abs(mono) ** 2)


def test_volume_waist_law(oracle):
    """TestPropagate3DVolume (optics/volume_test.go:8-46)."""
    n, width, wl, w0 = 512, 0.01, 632.8e-9, 3e-4
    dx = width / n
    g = oracle.gaussian(n, dx, w0)
    zs = [0.05, 0.1, 0.2, 0.3]
    vol, _, _ = oracle.volume(g, None, dx, zs, "asm", ctx(oracle, wl))
    x = (np.arange(n) - n / 2) * dx
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    zr = math.pi * w0 * w0 / wl
    for k, z in enumerate(zs):
        I = np.abs(vol[k]) ** 2
        got = math.sqrt(2 * np.sum(I * r2) / np.sum(I))
        want = w0 * math.sqrt(1 + (z / zr) ** 2)
        assert abs(got - want) / want <= 0.01


def test_schell_coherence(oracle):
    """TestSchellSourceCoherence (optics/coherence_test.go:10-45), with the
    counter-based generator in place of math/rand."""
    n, dx, w0, sigc = 256, 1e-5, 1e-3, 2e-4
    sq = oracle.schell_sqrt_spectrum(n, dx, sigc)
    fields = [oracle.schell_realization(sq, dx, w0, 1, r) for r in range(128)]
    for d_px in (1, 2, 4, 8, 16):
        num = da = db = 0.0
        for f in fields:
            a, b = f[:, :n - d_px], f[:, d_px:]
            num += np.sum((a.conj() * b).real)
            da += np.sum(np.abs(a) ** 2)
            db += np.sum(np.abs(b) ** 2)
        got = num / math.sqrt(da * db)
        d = d_px * dx
        assert abs(got - math.exp(-d * d / (2 * sigc * sigc))) <= 0.03


def far_field_radius(I, n, dx_out):
    """farFieldRadius (optics/coherence_test.go:67-128): 1/e^2 radius from a
    weighted log-linear fit of the radial profile above 2% of its peak."""
    nb = 128
    dr = dx_out * n / 2 / nb
    c = (np.arange(n) - n / 2) * dx_out
    r = np.hypot(c[None, :], c[:, None])
    b = (r / dr).astype(int)
    keep = b < nb
    s = np.bincount(b[keep], weights=I[keep], minlength=nb)
    cnt = np.bincount(b[keep], minlength=nb)
    ok = cnt > 0
    v = np.where(ok, s / np.maximum(cnt, 1), 0.0)
    use = ok & (v >= 0.02 * v[ok].max())
    r2 = ((np.arange(nb) + 0.5) * dr) ** 2
    w, lv = cnt[use].astype(float), np.log(v[use])
    sw, sx, sy, sxx, sxy = w.sum(), (w * r2[use]).sum(), (w * lv).sum(), (w * r2[use] ** 2).sum(), \
        (w * r2[use] * lv).sum()
    slope = (sw * sxy - sx * sy) / (sw * sxx - sx * sx)
    return math.sqrt(-2 / slope)


def test_schell_far_field(oracle):
    """TestSchellFarField (optics/coherence_test.go:47-65):
    PropagatePartiallyCoherent of 160 realisations with the Fraunhofer
    method; the far-field 1/e^2 radius is (wl z / pi) sqrt(1/w0^2 +
    1/sigma_c^2) within 10%."""
    n, dx, w0, sigc, wl, z = 512, 1.25e-5, 1e-3, 3.75e-5, 632.8e-9, 1.0
    I, _ = oracle.propagate_partially_coherent(n, dx, w0, sigc, 2, 160, z, "fraunhofer", oracle.ctx_for(wl))
    got = far_field_radius(I, n, wl * z / (n * dx))
    want = (wl * z / math.pi) * math.sqrt(1 / w0 ** 2 + 1 / sigc ** 2)
    assert abs(got - want) / want <= 0.1, (got, want)


def test_average_intensity_closed_forms(oracle):
    """AverageIntensity (optics/coherence.go:85-100): m copies of one field
    average to its intensity (to the rounding of 1/m); two fields to
    (I1 + I2) / 2 exactly; doubling every field quadruples the result
    exactly; the Ey component counts for polarised fields."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    b = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    ia, ib = a.real ** 2 + a.imag ** 2, b.real ** 2 + b.imag ** 2
    for m in (1, 3, 7):
        got = oracle.average_intensity(np.stack([a] * m))
        assert np.max(np.abs(got - ia) / ia) <= 4e-16 * m
    assert np.array_equal(oracle.average_intensity(np.stack([a, b])), (ia + ib) / 2)
    assert np.array_equal(oracle.average_intensity(np.stack([2 * a, 2 * b, a])),
                          4 * oracle.average_intensity(np.stack([a, b, a / 2])))
    got = oracle.average_intensity(np.stack([a, a]), np.stack([b, b]))
    assert np.max(np.abs(got - (ia + ib)) / (ia + ib)) <= 4e-16


def test_partially_coherent_reduces_to_its_parts(oracle):
    """PropagatePartiallyCoherent (optics/coherence.go:105-116) with one
    realisation is the intensity of Propagate applied to that realisation;
    with z = 0 it is AverageIntensity of the realisations themselves; m <= 0
    is taken as 1; a width or coherence <= 0 is refused."""
    n, dx, w0, sc, seed, wl = 64, 4e-6, 60e-6, 15e-6, 9, 532e-9
    ctx = oracle.ctx_for(wl)
    sq = oracle.schell_sqrt_spectrum(n, dx, sc)
    r0 = oracle.schell_realization(sq, dx, w0, seed, 0)
    u, _, _, _ = oracle.propagate(r0, None, dx, 2e-3, "asm", ctx)
    got, _ = oracle.propagate_partially_coherent(n, dx, w0, sc, seed, 1, 2e-3, "asm", ctx)
    assert np.array_equal(got, u.real ** 2 + u.imag ** 2)
    got0, _ = oracle.propagate_partially_coherent(n, dx, w0, sc, seed, 0, 2e-3, "asm", ctx)
    assert np.array_equal(got0, got)
    fields = np.stack([oracle.schell_realization(sq, dx, w0, seed, k) for k in range(5)])
    got, _ = oracle.propagate_partially_coherent(n, dx, w0, sc, seed, 5, 0.0, "asm", ctx)
    assert np.array_equal(got, oracle.average_intensity(fields))
    for bad in ((0.0, sc), (w0, -1.0)):
        with pytest.raises(ValueError):
            oracle.propagate_partially_coherent(n, dx, bad[0], bad[1], seed, 2, 1e-3, "asm", ctx)


def test_normal_generator_moments(oracle):
    L = oracle.lib()
    v = np.array([L.or_normal(7, c) for c in range(20000)])
    assert abs(v.mean()) < 0.03 and abs(v.std() - 1) < 0.03


# ---------------------------------------------------------------- adjoint

def test_mask_adjoint_matches_central_differences(oracle):
    """Phase-mask gradient by the adjoint vs central finite differences of the
    FP64 forward (no reference counterpart)."""
    n, dx, wl, z = 64, 4e-6, 532e-9, 2e-3
    rng = np.random.default_rng(5)
    theta = rng.uniform(0, 2 * math.pi, (n, n))
    x = (np.arange(n) - n / 2) * dx
    T = np.exp(-((x[None, :] - 40e-6) ** 2 + x[:, None] ** 2) / (2 * (12e-6) ** 2))
    T /= T.sum()
    c = ctx(oracle, wl)
    L0, g = oracle.mask_loss_grad(theta, T, dx, z, c)
    h = 1e-5
    for (j, i) in [(3, 5), (31, 32), (60, 2), (17, 40)]:
        tp = theta.copy()
        tp[j, i] += h
        tm = theta.copy()
        tm[j, i] -= h
        fd = (oracle.mask_loss_grad(tp, T, dx, z, c, False)[0] - oracle.mask_loss_grad(tm, T, dx, z, c, False)[0]) / (2 * h)
        assert abs(fd - g[j, i]) <= 1e-4 * max(abs(fd), 1e-3 * np.max(np.abs(g)))


# ---------------------------------------------------------------- plane statistics

def test_compute_stats_closed_forms(oracle):
    """ComputeStats (optics/metrics.go:31-91) on an offset Gaussian with a
    constant phase: power pi w0^2/2, centroid (x0, y0), RMS widths w0/2 (the
    intensity exp(-2 r^2/w0^2) has sigma w0/2), peak 1, phase range = the
    phase, Strehl = peak / (P pi R^2 / (wl^2 z^2)); a second polarisation
    component adds its intensity."""
    n, dx, w0 = 256, 1e-6, 20e-6
    x0, y0, ph0 = 10e-6, -6e-6, 0.7
    x = (np.arange(n) - n / 2) * dx
    X, Y = x[None, :], x[:, None]
    ex = np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / w0 ** 2) * np.exp(1j * ph0)
    wl, ap, dist = 632.8e-9, 1e-3, 0.1
    s = oracle.compute_stats(ex, None, dx, wl, ap, dist)
    P = math.pi * w0 ** 2 / 2
    assert abs(s.power - P) <= 1e-12 * P
    assert abs(s.centroid_x - x0) <= 1e-15 and abs(s.centroid_y - y0) <= 1e-15
    assert abs(s.rms_x - w0 / 2) <= 1e-12 * w0 and abs(s.rms_y - w0 / 2) <= 1e-12 * w0
    assert abs(s.peak - 1.0) <= 1e-15 and s.intensity_max == s.peak
    imin = float(np.min(np.exp(-2 * ((X - x0) ** 2 + (Y - y0) ** 2) / w0 ** 2)))   # far corner
    assert abs(s.intensity_min - imin) <= 1e-12 * imin
    assert abs(s.phase_min - ph0) <= 1e-15 and abs(s.phase_max - ph0) <= 1e-15
    ideal = P * math.pi * ap ** 2 / (wl ** 2 * dist ** 2)
    assert abs(s.strehl - 1.0 / ideal) <= 1e-12 / ideal
    s2 = oracle.compute_stats(ex, 0.5 * ex, dx, wl)
    assert abs(s2.power - 1.25 * P) <= 1e-12 * P and s2.strehl == 0.0
    assert abs(s2.centroid_x - x0) <= 1e-15


def test_bpm_mask_adjoint_matches_central_differences(oracle):
    """Multi-plane phase design through a split-step medium: gradients with
    respect to the input phase and to each design phase plane vs central
    differences of the FP64 forward; with no design phases the forward is
    PropagateSplitStep of exp(i theta)."""
    n, dx, wl, steps, z = 32, 1e-6, 1.0e-6, 3, 60e-6
    rng = np.random.default_rng(9)
    theta = rng.uniform(0, 2 * math.pi, (n, n))
    phi = rng.uniform(0, 2 * math.pi, (steps, n, n))
    x = (np.arange(n) - n / 2) * dx
    r = np.hypot(x[None, :], x[:, None])
    index = np.array([np.where(r < (6 + 2 * s) * 1e-6, 1.5, 1.2) + 1e-4j for s in range(steps)])
    T = np.exp(-((x[None, :] - 3e-6) ** 2 + x[:, None] ** 2) / (2 * (3e-6) ** 2))
    T /= T.sum()
    c = ctx(oracle, wl)
    L0, gt, gp = oracle.bpm_mask_loss_grad(theta, phi, index, T, dx, z, steps, c)
    h = 1e-5

    def fd(mod):
        tp, pp = theta.copy(), phi.copy()
        tm, pm = theta.copy(), phi.copy()
        mod(tp, pp, h)
        mod(tm, pm, -h)
        return (oracle.bpm_mask_loss_grad(tp, pp, index, T, dx, z, steps, c, False)[0] -
                oracle.bpm_mask_loss_grad(tm, pm, index, T, dx, z, steps, c, False)[0]) / (2 * h)

    gmax = max(np.max(np.abs(gt)), np.max(np.abs(gp)))
    for (j, i) in [(3, 5), (16, 17), (30, 2)]:
        def mt(t, p, d, j=j, i=i):
            t[j, i] += d
        f = fd(mt)
        assert abs(f - gt[j, i]) <= 1e-4 * max(abs(f), 1e-3 * gmax)
    for (s, j, i) in [(0, 4, 9), (1, 15, 16), (2, 20, 11)]:
        def mp(t, p, d, s=s, j=j, i=i):
            p[s, j, i] += d
        f = fd(mp)
        assert abs(f - gp[s, j, i]) <= 1e-4 * max(abs(f), 1e-3 * gmax)
    # no design phases: the forward field is PropagateSplitStep's
    L1, _, _ = oracle.bpm_mask_loss_grad(theta, None, index, T, dx, z, steps, c, False)
    out, _, _ = oracle.split_step(np.exp(1j * theta), None, dx, z, steps, index, c)
    I = np.abs(out) ** 2
    assert abs(L1 - np.sum((I / I.sum() - T) ** 2)) <= 1e-12 * L1
