"""The persistent transfer: one launch per batch run in which workgroups take
the forward rows, columns and inverse rows of every unit from per-XCD device
work queues. The batched form of Propagate (optics/polychromatic.go:18-43
runs the same propagations one after another) must give what the
three-launch transfer gives at the same launch shape, bit for bit (the task
bodies are the same kernels' code; a task that read its input before the
producing pass finished would differ), and what the FP64 oracle gives
within the precision's tolerance. Unit counts that are not a multiple of 8
leave the XCD queues unequal; 3 units leave five queues empty.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5
W0, SIGMA_C, SEED = 300e-6, 50e-6, 20240607


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def persistent_shape(wos, precision):
    """The persistent kernel's launch shape, for the three-launch twin."""
    r = 2 if precision == wos.FP64 else 4
    return dict(tile_w=r, col_w=r, rows=r, groups=0, split=1, radix=16, col_radix=16, col_split=1)


def three_and_persistent(wos, n, dx, wls, z, precision, picks, runs=1, graph=False, interleave=False):
    """Sources, three-launch outputs and persistent outputs of `picks` after
    `runs` runs from the same seeded sources."""
    units = len(wls)
    b = wos.Batch(n, dx, wls, np.full(units, z), wos.ctx_for(550e-9), "asm", precision)
    b.set_tuning(**persistent_shape(wos, precision))
    b.fill_schell(W0, SIGMA_C, SEED)
    srcs = {u: b.download(u) for u in picks}
    b.run(runs)
    b.sync()
    three = {u: b.download(u) for u in picks}
    b.fill_schell(W0, SIGMA_C, SEED)
    b.set_persistent(True, interleave)
    if graph:
        b.capture()
        for _ in range(runs):   # replays with the host's own synchronisation between them
            b.run(1)
            b.sync()
    else:
        b.run(runs)
        b.sync()
    pers = {u: b.download(u) for u in picks}
    b.close()
    return srcs, three, pers


@pytest.mark.parametrize("precision,n,units,interleave", [("fp32", 512, 11, False), ("fp32", 512, 11, True),
                                                           ("fp32", 1024, 3, False), ("fp64", 1024, 11, False),
                                                           ("fp64", 1024, 11, True)])
def test_persistent_matches_three_launch_and_oracle(wos, oracle, precision, n, units, interleave):
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    tol = FP64_TOL if prec == wos.FP64 else FP32_TOL
    dx, z = 2e-6, 20e-3
    wls = np.linspace(400e-9, 700e-9, units)
    picks = tuple(range(units))
    print(f"persistent {precision} n={n} units={units} interleave={interleave}", flush=True)
    srcs, three, pers = three_and_persistent(wos, n, dx, wls, z, prec, picks, interleave=interleave)
    for u in picks:
        assert np.array_equal(pers[u], three[u]), (u, rel_l2(pers[u], three[u]))
    for u in (0, units - 1):
        ref, _, _, _ = oracle.propagate(srcs[u], None, dx, z, "asm", oracle.ctx_for(wls[u]))
        e = rel_l2(pers[u], ref)
        print(f"persistent {precision} n={n} unit {u}: rel-L2 vs oracle {e:.3e}", flush=True)
        assert e <= tol, (u, e)


@pytest.mark.parametrize("precision,interleave", [("fp32", False), ("fp32", True), ("fp64", False)])
def test_persistent_headline_batch(wos, oracle, precision, interleave):
    """64 x 2048^2 at dx 2 um, z 20 mm, 400-700 nm, captured into a graph and
    replayed twice (each replay propagates the previous output again):
    units of every XCD queue equal the three-launch runs, and unit 63 after
    one replay matches the oracle."""
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    tol = FP64_TOL if prec == wos.FP64 else FP32_TOL
    n, dx, z, units = 2048, 2e-6, 20e-3, 64
    wls = np.linspace(400e-9, 700e-9, units)
    picks = (0, 7, 8, 63)
    print(f"persistent headline {precision} interleave={interleave}: two graph replays", flush=True)
    _, three, pers = three_and_persistent(wos, n, dx, wls, z, prec, picks, runs=2, graph=True, interleave=interleave)
    for u in picks:
        assert np.array_equal(pers[u], three[u]), (u, rel_l2(pers[u], three[u]))
    print(f"persistent headline {precision}: one run against the oracle", flush=True)
    srcs, _, pers1 = three_and_persistent(wos, n, dx, wls, z, prec, (63,), interleave=interleave)
    ref, _, _, _ = oracle.propagate(srcs[63], None, dx, z, "asm", oracle.ctx_for(wls[63]))
    e = rel_l2(pers1[63], ref)
    print(f"persistent headline {precision} unit 63: rel-L2 vs oracle {e:.3e}", flush=True)
    assert e <= tol, e


def test_persistent_rejects_unbuilt_shapes(wos):
    """asm_pad batches and sides outside the built set keep the three-launch
    form and say why."""
    ctx = wos.ctx_for(550e-9)
    for n, method, prec in ((256, "asm_pad", wos.FP32), (256, "asm", wos.FP32), (512, "asm", wos.FP64)):
        b = wos.Batch(n, 2e-6, np.full(2, 550e-9), np.full(2, 1e-3), ctx, method, prec)
        with pytest.raises(wos.WosError, match="persistent"):
            b.set_persistent(True)
        b.run(1)   # still runs, in three launches
        b.sync()
        b.close()
