"""Parity at the measured sizes: the 64 x 2048^2 headline batch at its
autotuned launch shape, the batched configuration's 2048^2 polychromatic and
1024^2 volume stages in FP64, and the split-step configuration's full 400
slices, each against the FP64 oracle (rel-L2 <= 1e-12 for FP64, <= 1e-5 for
FP32). The oracle runs on all host threads; each check keeps to a few
oracle propagations at these sizes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5
W0, SIGMA_C, SEED = 300e-6, 50e-6, 20240607


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def intensity(u):
    """Field.Intensity of a scalar field (optics/field.go:64-73): re^2 + im^2."""
    return u.real * u.real + u.imag * u.imag


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_headline_batch_units_match_oracle(wos, oracle, precision):
    """64 Gaussian-Schell fields of 2048^2 (dx 2 um) at 64 wavelengths over
    400-700 nm, asm over z = 20 mm: the batch is autotuned, captured and
    replayed, and units 0, 31 and 63 are compared with the oracle's
    Propagate of the same source; then again at the committed tuning
    table's shape, the one the benchmark runs."""
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    tol = FP64_TOL if prec == wos.FP64 else FP32_TOL
    n, dx, z, units = 2048, 2e-6, 20e-3, 64
    wls = np.linspace(400e-9, 700e-9, units)
    b = wos.Batch(n, dx, wls, np.full(units, z), wos.ctx_for(550e-9), "asm", prec)
    b.fill_schell(W0, SIGMA_C, SEED)
    picks = (0, 31, 63)
    srcs = {u: b.download(u) for u in picks}
    shape, ms = b.autotune(2)
    print(f"headline {precision}: autotuned shape {shape} ({ms:.3f} ms per run)", flush=True)
    b.fill_schell(W0, SIGMA_C, SEED)   # the tuning launches ran on the data
    for u in picks:
        assert np.array_equal(b.download(u), srcs[u]), u   # the generator is deterministic
    refs = {u: oracle.propagate(srcs[u], None, dx, z, "asm", oracle.ctx_for(wls[u]))[0] for u in picks}
    b.capture()
    b.run(1)
    b.sync()
    for u in picks:
        e = rel_l2(b.download(u), refs[u])
        print(f"headline {precision} unit {u} (wl {wls[u] * 1e9:.1f} nm): rel-L2 vs oracle {e:.3e}", flush=True)
        assert e <= tol, (u, e)
    # the committed tuning table's shape, which the benchmark runs
    table = b.load_tuning()
    b.fill_schell(W0, SIGMA_C, SEED)
    b.run(1)
    b.sync()
    for u in picks:
        e = rel_l2(b.download(u), refs[u])
        print(f"headline {precision} unit {u}, table shape {table}: rel-L2 vs oracle {e:.3e}", flush=True)
        assert e <= tol, (u, e)
    b.close()


def test_config4_fp64_stages_match_oracle(wos, oracle):
    """The batched configuration's shared-source stages in FP64 inside one
    captured graph: 64 wavelengths x 2048^2 with the polychromatic intensity
    epilogue, and 128 z-planes (0-20 mm) x 1024^2. Four wavelengths' fields
    and four planes are compared with the oracle's Propagate / Propagate3D
    on the stage's own source; the oracle's PropagatePolychromatic over the
    four wavelengths equals their share of the device fields; the device
    intensity equals the sum over all 64 downloaded fields in unit order."""
    ctx = wos.ctx_for(550e-9)
    p = wos.Pipeline(wos.FP64)
    wls = np.linspace(400e-9, 700e-9, 64)
    zs = np.linspace(0, 20e-3, 128)
    s_poly = p.add(2048, 2e-6, wls, np.full(64, 20e-3), ctx, "asm", src_units=1)
    p.set_intensity(s_poly, np.ones(64))
    p.fill_schell(s_poly, W0, SIGMA_C, SEED, 0)
    s_vol = p.add(1024, 2e-6, np.full(128, 532e-9), zs, ctx, "asm", src_units=1)
    p.fill_schell(s_vol, W0, SIGMA_C, SEED, 1)
    src_poly = p.source(s_poly, 0)
    src_vol = p.source(s_vol, 0)
    p.capture()
    p.run(1)
    p.sync()
    assert np.array_equal(p.source(s_poly, 0), src_poly)   # stages keep their inputs
    picks = (0, 21, 42, 63)
    acc = np.zeros((2048, 2048))
    part = np.zeros((2048, 2048))
    for u in range(64):
        f = p.download(s_poly, u)
        acc += intensity(f)
        if u in picks:
            part += intensity(f)
            ref, _, _, _ = oracle.propagate(src_poly, None, 2e-6, 20e-3, "asm", oracle.ctx_for(wls[u]))
            e = rel_l2(f, ref)
            print(f"config4 FP64 polychromatic unit {u}: rel-L2 vs oracle {e:.3e}", flush=True)
            assert e <= FP64_TOL, (u, e)
    dev = p.intensity(s_poly)
    e = rel_l2(dev, acc)
    print(f"config4 FP64 intensity epilogue vs the 64 fields summed in unit order: {e:.3e}", flush=True)
    assert e <= 1e-15
    ref_part, _ = oracle.polychromatic(src_poly, None, 2e-6, 20e-3, wls[list(picks)], np.ones(4), "asm",
                                       oracle.ctx_for(550e-9))
    e = rel_l2(part, ref_part)
    print(f"config4 FP64 four-wavelength intensity vs oracle PropagatePolychromatic: {e:.3e}", flush=True)
    assert e <= FP64_TOL
    vpicks = [0, 1, 64, 127]
    rv, _, _ = oracle.volume(src_vol, None, 2e-6, zs[vpicks], "asm", oracle.ctx_for(532e-9))
    for k, u in enumerate(vpicks):
        got = p.download(s_vol, u)
        if u == 0:
            assert np.array_equal(got, src_vol)   # z = 0: the source itself
        e = rel_l2(got, rv[k])
        print(f"config4 FP64 volume plane {u} (z {zs[u] * 1e3:.3f} mm): rel-L2 vs oracle {e:.3e}", flush=True)
        assert e <= FP64_TOL, (u, e)
    p.close()


def test_config3_full_400_slices_fp64_vs_oracle(wos, oracle):
    """The split-step configuration in full: 1024^2, dx 0.5 um, 1.55 um,
    400 slices of 5 um through the graded-index fibre with absorption,
    Gaussian w0 10 um; FP64 device path against the oracle's step-by-step
    PropagateSplitStep, and its warnings."""
    n, dx, wl, steps = 1024, 0.5e-6, 1.55e-6, 400
    x = (np.arange(n) - n / 2) * dx
    r = np.hypot(x[None, :], x[:, None])
    index = np.where(r < 100e-6, 1.45 - 0.01 * (r / 100e-6) ** 2, 1.44) + 1e-5j
    ex = oracle.gaussian(n, dx, 10e-6)
    z = steps * 5e-6
    out, _, wg = wos.split_step(ex, None, dx, z, steps, index, wos.ctx_for(wl))
    ref, _, wo = oracle.split_step(ex, None, dx, z, steps, index, oracle.ctx_for(wl))
    e = rel_l2(out, ref)
    ratio = np.sum(intensity(out)) / np.sum(intensity(ex))
    print(f"config3 400 slices FP64 rel-L2 vs oracle {e:.3e}, power ratio {ratio:.6f}", flush=True)
    assert e <= FP64_TOL
    assert wg.as_dict().keys() == wo.as_dict().keys()
