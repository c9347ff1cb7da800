"""Phase-mask design operator on the GPU: forward loss and hand-written
adjoint gradient (no reference counterpart). Checked against the oracle's
FP64 CPU adjoint (itself pinned to central differences in test_oracle.py)
and against central finite differences of the GPU's own FP64 forward."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def ring_target(n, dx, spots=16, radius=800e-6, sigma=12e-6):
    """16 Gaussian spots (sigma 12 um) on a ring of radius 800 um, sum 1."""
    x = (np.arange(n) - n / 2) * dx
    X, Y = x[None, :], x[:, None]
    t = np.zeros((n, n))
    for s in range(spots):
        a = 2 * math.pi * s / spots
        t += np.exp(-((X - radius * math.cos(a)) ** 2 + (Y - radius * math.sin(a)) ** 2) / (2 * sigma ** 2))
    return t / t.sum()


def masks(units, n, seed):
    """float32 masks, i.i.d. uniform [0, 2 pi)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 2 * math.pi, (units, n, n)).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n", [64, 128, 100])
def test_mask_operator_fp64_matches_cpu_adjoint(wos, oracle, n):
    """Loss and gradient of two masks vs the oracle's FP64 adjoint; n = 100
    runs the Bluestein path with the conjugate transfer function."""
    dx, wl, z = 4e-6, 532e-9, 2e-3
    target = ring_target(n, dx, radius=n * dx / 4, sigma=8e-6)
    theta = masks(2, n, 3)
    op = wos.MaskOp(n, 2, dx, z, wos.ctx_for(wl), target)
    loss = op.forward(theta)
    grad = op.backward()
    grad2 = op.backward()
    assert np.array_equal(grad, grad2)            # backward is repeatable on one checkpoint
    for u in range(2):
        L, g = oracle.mask_loss_grad(theta[u], target, dx, z, oracle.ctx_for(wl))
        eg = rel_l2(grad[u], g)
        print(f"n={n} unit {u}: loss {loss[u]:.6e} vs {L:.6e} (rel {abs(loss[u] - L) / L:.2e}), grad rel-L2 {eg:.2e}",
              flush=True)
        assert abs(loss[u] - L) <= 1e-12 * L
        assert eg <= 1e-10
    u_ref, _, _, _ = oracle.propagate(np.exp(1j * theta[1]), None, dx, z, "asm", oracle.ctx_for(wl))
    assert rel_l2(op.field(1), u_ref) <= 1e-12
    op.close()


def test_config5_mask_design_fd_and_cpu_adjoint(wos, oracle):
    """Config 5: 8 masks of 1024^2 (float32 uniform [0, 2 pi), seeded), plane
    wave at 532 nm, dx = 4 um, ASM over 20 mm; L2 between the normalised
    intensity and 16 Gaussian spots (sigma 12 um) on an 800 um ring.
    Gradient vs central differences of the FP64 forward on 32 seeded pixels
    (<= 1e-4 relative) and vs the FP64 CPU adjoint; FP32 error reported."""
    n, units, dx, wl, z = 1024, 8, 4e-6, 532e-9, 20e-3
    target = ring_target(n, dx)
    theta = masks(units, n, 20240611)
    ctx = wos.ctx_for(wl)
    op = wos.MaskOp(n, units, dx, z, ctx, target)
    loss = op.forward(theta)
    grad = op.backward()

    # CPU FP64 adjoint on two of the masks
    for u in (0, units - 1):
        L, g = oracle.mask_loss_grad(theta[u], target, dx, z, oracle.ctx_for(wl))
        eg = rel_l2(grad[u], g)
        print(f"config5 unit {u}: loss {loss[u]:.9e} vs CPU {L:.9e}; grad rel-L2 vs CPU adjoint {eg:.2e}", flush=True)
        assert abs(loss[u] - L) <= 1e-12 * L
        assert eg <= 1e-10

    # central differences: one pixel per mask per round (the masks are
    # independent units of one forward), 4 rounds = 32 pixels
    rng = np.random.default_rng(7)
    pix = rng.integers(0, n, size=(4, units, 2))
    h = 1e-3
    worst = 0.0
    gmax = np.max(np.abs(grad), axis=(1, 2))
    for r in range(4):
        tp, tm = theta.copy(), theta.copy()
        for u in range(units):
            j, i = pix[r, u]
            tp[u, j, i] += h
            tm[u, j, i] -= h
        lp, lm = op.forward(tp), op.forward(tm)
        for u in range(units):
            j, i = pix[r, u]
            fd = (lp[u] - lm[u]) / (2 * h)
            g = grad[u, j, i]
            scale = max(abs(fd), 1e-3 * gmax[u])
            err = abs(fd - g) / scale
            worst = max(worst, err)
            assert err <= 1e-4, (r, u, j, i, fd, g)
    print(f"config5 FD: 32 pixels, worst relative error {worst:.2e} (h = {h})", flush=True)

    # FP32 field arithmetic vs the FP64 path
    op32 = wos.MaskOp(n, units, dx, z, ctx, target, precision=wos.FP32)
    l32 = op32.forward(theta)
    g32 = op32.backward()
    el = float(np.max(np.abs(l32 - loss) / loss))
    eg = rel_l2(g32, grad)
    print(f"config5 FP32: loss max rel {el:.2e}, grad rel-L2 vs FP64 {eg:.2e}", flush=True)
    assert el <= 1e-5 and eg <= 1e-4
    op32.close()
    op.close()
