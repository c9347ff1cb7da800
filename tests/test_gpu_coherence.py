"""Partially coherent light on the device (optics/coherence.go:22-116) and
PropagateSplitStep called with its IndexFunc, against the FP64 oracle.

The Gaussian-Schell spectrum sqrt(max(0, Re FFT(mu))) takes square roots of
values at the FFT's rounding level in its tails, so two FFT implementations
give sources ~1e-8 apart there (rel-L2 ~1e-7 over the field). The
propagation and the averaging are therefore checked at the FP64 bar (1e-12)
on the device's own realisations, and the device realisations against the
oracle's at 1e-6.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def test_schell_realizations_match_oracle(wos, oracle):
    n, dx, w0, sc, seed, m = 128, 4e-6, 120e-6, 20e-6, 11, 4
    dev = wos.schell_realizations(n, dx, w0, sc, seed, m)
    sq = oracle.schell_sqrt_spectrum(n, dx, sc)
    for k in range(m):
        e = rel_l2(dev[k], oracle.schell_realization(sq, dx, w0, seed, k))
        print(f"schell realisation {k}: device vs oracle rel-L2 {e:.2e}", flush=True)
        assert e <= 1e-6
    # the batch generator draws the same realisations
    b = wos.Batch(n, dx, np.full(m, 532e-9), np.full(m, 1e-3), wos.ctx_for(532e-9), "asm", wos.FP64)
    b.fill_schell(w0, sc, seed)
    for k in range(m):
        assert np.array_equal(b.download(k), dev[k]), k
    b.close()
    assert wos.schell_realizations(n, dx, w0, sc, seed, 0).shape == (1, n, n)   # m <= 0 is one
    for bad in ((0.0, sc), (w0, -1e-6)):
        with pytest.raises(wos.WosError):
            wos.schell_realizations(n, dx, bad[0], bad[1], seed, 2)


def test_average_intensity_matches_oracle(wos, oracle):
    rng = np.random.default_rng(4)
    fx = rng.standard_normal((5, 96, 96)) + 1j * rng.standard_normal((5, 96, 96))
    fy = rng.standard_normal((5, 96, 96)) + 1j * rng.standard_normal((5, 96, 96))
    assert np.array_equal(wos.average_intensity(fx), oracle.average_intensity(fx))
    assert np.array_equal(wos.average_intensity(fx, fy), oracle.average_intensity(fx, fy))
    assert wos.average_intensity(fx[:0]) is None


@pytest.mark.parametrize("method", ["asm", "asm_pad", "fraunhofer", "fresnel_tf"])
def test_partially_coherent_matches_oracle(wos, oracle, method):
    """PropagatePartiallyCoherent: the device result equals AverageIntensity
    of the oracle's Propagate applied to the device's realisations (FP64
    1e-12, FP32 1e-5), the oracle's own PropagatePartiallyCoherent to the
    generator's 1e-6, and the warnings the oracle collects."""
    n, dx, w0, sc, seed, m, wl = 128, 4e-6, 100e-6, 25e-6, 3, 6, 632.8e-9
    z = 0.5 if method == "fraunhofer" else 5e-3
    real = wos.schell_realizations(n, dx, w0, sc, seed, m)
    props = [oracle.propagate(r, None, dx, z, method, oracle.ctx_for(wl))[0] for r in real]
    want = oracle.average_intensity(np.stack(props))
    got, wg = wos.propagate_partially_coherent(n, dx, w0, sc, seed, m, z, method, wos.ctx_for(wl))
    e = rel_l2(got, want)
    own, wo = oracle.propagate_partially_coherent(n, dx, w0, sc, seed, m, z, method, oracle.ctx_for(wl))
    e_own = rel_l2(got, own)
    print(f"partially coherent {method} FP64: vs oracle on device realisations {e:.2e}, "
          f"vs oracle's own realisations {e_own:.2e}", flush=True)
    assert e <= FP64_TOL
    assert e_own <= 1e-6
    assert wg.as_dict().keys() == wo.as_dict().keys()
    g32, _ = wos.propagate_partially_coherent(n, dx, w0, sc, seed, m, z, method, wos.ctx_for(wl), wos.FP32)
    e32 = rel_l2(g32, want)
    print(f"partially coherent {method} FP32 rel-L2 {e32:.2e}", flush=True)
    assert e32 <= FP32_TOL


def test_partially_coherent_pipeline_stage(wos, oracle):
    """m Schell realisations as one pipeline stage with the AverageIntensity
    epilogue, captured and replayed, equal the drop-in operator."""
    n, dx, w0, sc, seed, m, wl, z = 256, 4e-6, 200e-6, 40e-6, 8, 12, 532e-9, 10e-3
    for prec, tol in ((wos.FP64, 1e-13), (wos.FP32, FP32_TOL)):
        p = wos.Pipeline(prec)
        s = p.add(n, dx, np.full(m, wl), np.full(m, z), wos.ctx_for(wl), "asm")
        p.fill_schell(s, w0, sc, seed, 0)
        p.set_average(s)
        p.capture()
        p.run(1)
        p.sync()
        got = p.intensity(s)
        ref, _ = wos.propagate_partially_coherent(n, dx, w0, sc, seed, m, z, "asm", wos.ctx_for(wl), wos.FP64)
        e = rel_l2(got, ref)
        print(f"partially coherent pipeline stage prec={prec}: vs drop-in FP64 {e:.2e}", flush=True)
        assert e <= tol
        p.run(1)
        p.sync()
        assert np.array_equal(p.intensity(s), got)
        p.close()


def test_schell_far_field_on_device(wos):
    """TestSchellFarField (optics/coherence_test.go:47-65) through the device
    operator: 160 realisations, Fraunhofer, far-field radius within 10% of
    (wl z / pi) sqrt(1/w0^2 + 1/sigma_c^2)."""
    from test_oracle import far_field_radius
    n, dx, w0, sigc, wl, z = 512, 1.25e-5, 1e-3, 3.75e-5, 632.8e-9, 1.0
    I, _ = wos.propagate_partially_coherent(n, dx, w0, sigc, 2, 160, z, "fraunhofer", wos.ctx_for(wl))
    got = far_field_radius(I, n, wl * z / (n * dx))
    want = (wl * z / math.pi) * math.sqrt(1 / w0 ** 2 + 1 / sigc ** 2)
    print(f"schell far field radius {got:.4e} (closed form {want:.4e})", flush=True)
    assert abs(got - want) / want <= 0.1


def test_split_step_with_index_function(wos, oracle):
    """PropagateSplitStep with its IndexFunc: a z-invariant medium is sampled
    into one plane and gives the same field as the one plane passed
    directly, and so does the same plane repeated for every slice (the
    library collapses bit-equal planes); a stratified medium keeps its
    planes and matches the oracle."""
    from test_host_mirror import graded_index_fn, graded_index_planes
    n, dx, wl, steps = 128, 0.5e-6, 1.55e-6, 40
    ex = oracle.gaussian(n, dx, 10e-6)
    z = steps * 5e-6
    one, _, _ = wos.split_step(ex, None, dx, z, steps, graded_index_planes(n, dx), wos.ctx_for(wl))
    fn_out, _, _ = wos.propagate_split_step(ex, None, dx, z, graded_index_fn, steps, wos.ctx_for(wl))
    rep, _, _ = wos.split_step(ex, None, dx, z, steps, np.stack([graded_index_planes(n, dx)] * steps),
                               wos.ctx_for(wl))
    assert np.array_equal(fn_out, one) and np.array_equal(rep, one)
    zz, st = 2e-3, 8
    strat = lambda x, y, zc: np.where(zc < zz / 2, 1.5, 1.0) + 2e-6j + 0 * x * y
    g = oracle.gaussian(n, 2e-6, 40e-6)
    out, _, _ = wos.propagate_split_step(g, None, 2e-6, zz, strat, st, wos.ctx_for(632.8e-9))
    ref, _, _ = oracle.split_step(g, None, 2e-6, zz, st, wos.sample_index_planes(strat, n, 2e-6, zz, st),
                                  oracle.ctx_for(632.8e-9))
    assert rel_l2(out, ref) <= FP64_TOL
