"""Phase-mask design operator on the GPU: forward loss and hand-written
adjoint gradient (no reference counterpart). Checked against the oracle's
FP64 CPU adjoint (itself pinned to central differences in test_oracle.py)
and against central finite differences of the GPU's own FP64 forward."""
import ctypes as C
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


@pytest.mark.parametrize("n", [64, 128, 100, 94])
def test_mask_operator_fp64_matches_cpu_adjoint(wos, oracle, n):
    """Loss and gradient of two masks vs the oracle's FP64 adjoint; n = 100
    runs the mixed-radix path and n = 94 = 2 47 the Bluestein path, both with
    the conjugate transfer function."""
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
    # device-resident masks: forward(None) reuses the uploaded theta; masks
    # written straight into device memory give the host-upload results
    assert np.array_equal(op.forward(None), loss)
    assert op.backward(to_host=False) is None
    assert np.array_equal(op.backward(), grad)
    theta2 = masks(2, n, 4)
    loss2, grad2 = op.forward(theta2), op.backward()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    op.forward(theta)
    assert hip.hipMemcpy(wos.load().wos_hip_mask_device_theta(op.h), theta2.ctypes.data, theta2.nbytes, 1) == 0
    assert np.array_equal(op.forward(None), loss2)
    assert np.array_equal(op.backward(), grad2)
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
    # independent units of one forward), 4 rounds = 32 pixels; the error is
    # relative to the difference quotient itself, |fd - g| / |fd| <= 1e-4 at
    # every pixel
    rng = np.random.default_rng(7)
    pix = rng.integers(0, n, size=(4, units, 2))
    h = 1e-3
    worst = 0.0
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
            err = abs(fd - g) / abs(fd)
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


# ---------------------------------------------------------------- multi-plane phase design through BPM

def bpm_case(n, dx, steps, seed):
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, 2 * math.pi, (n, n))
    phi = rng.uniform(0, 2 * math.pi, (steps, n, n))
    x = (np.arange(n) - n / 2) * dx
    r = np.hypot(x[None, :], x[:, None])
    index = np.array([np.where(r < n * dx * (0.1 + 0.05 * s), 1.5, 1.3) + 1e-4j for s in range(steps)])
    target = np.exp(-((x[None, :] - 4 * dx) ** 2 + x[:, None] ** 2) / (2 * (3 * dx) ** 2))
    return theta, phi, index, target / target.sum()


@pytest.mark.parametrize("n", [64, 100, 94])
def test_bpm_mask_operator_matches_cpu_adjoint(wos, oracle, n):
    """Loss and gradients (input phase and every design phase plane) vs the
    oracle's FP64 BPM adjoint; n = 100 runs the mixed-radix path, n = 94 the
    Bluestein path. FP32 error against FP64 reported."""
    dx, wl, steps, z = 1e-6, 1.0e-6, 4, 80e-6
    theta, phi, index, target = bpm_case(n, dx, steps, 2)
    L, gt, gp = oracle.bpm_mask_loss_grad(theta, phi, index, target, dx, z, steps, oracle.ctx_for(wl))
    op = wos.BpmMaskOp(n, dx, z, steps, index, wos.ctx_for(wl), target)
    lg = op.forward(theta, phi)
    g_t, g_p = op.backward()
    et, ep = rel_l2(g_t, gt), rel_l2(g_p, gp)
    print(f"bpm mask n={n}: loss {lg:.9e} vs {L:.9e}, grad theta {et:.2e}, grad phi {ep:.2e}", flush=True)
    assert abs(lg - L) <= 1e-12 * L and et <= 1e-10 and ep <= 1e-10
    op32 = wos.BpmMaskOp(n, dx, z, steps, index, wos.ctx_for(wl), target, precision=wos.FP32)
    l32 = op32.forward(theta, phi)
    t32, p32 = op32.backward()
    print(f"bpm mask n={n} FP32: loss {abs(l32 - L) / L:.2e}, grad theta {rel_l2(t32, gt):.2e}, "
          f"grad phi {rel_l2(p32, gp):.2e}", flush=True)
    assert abs(l32 - L) <= 1e-5 * L and rel_l2(t32, gt) <= 1e-4 and rel_l2(p32, gp) <= 1e-4
    # no design phases: the forward is PropagateSplitStep of exp(i theta)
    l0 = op.forward(theta)
    out, _, _ = oracle.split_step(np.exp(1j * theta), None, dx, z, steps, index, oracle.ctx_for(wl))
    I = np.abs(out) ** 2
    assert abs(l0 - np.sum((I / I.sum() - target) ** 2)) <= 1e-12 * l0
    op.close()
    op32.close()


def test_bpm_mask_fd_1024(wos, oracle):
    """1024^2 (config-3 grid, 20 slices of 5 um through a graded index) with
    design phase planes: GPU FP64 gradients vs central differences of the GPU
    forward on seeded pixels of theta and of three phase planes, and vs the
    oracle's adjoint."""
    n, dx, wl, steps = 1024, 0.5e-6, 1.55e-6, 20
    z = steps * 5e-6
    rng = np.random.default_rng(31)
    theta = rng.uniform(0, 0.5, (n, n))
    phi = rng.uniform(0, 0.5, (steps, n, n))
    x = (np.arange(n) - n / 2) * dx
    r = np.hypot(x[None, :], x[:, None])
    index = np.where(r < 100e-6, 1.45 - 0.01 * (r / 100e-6) ** 2, 1.44) + 1e-5j
    target = np.exp(-((x[None, :] - 8e-6) ** 2 + x[:, None] ** 2) / (2 * (4e-6) ** 2))
    target /= target.sum()
    op = wos.BpmMaskOp(n, dx, z, steps, index, wos.ctx_for(wl), target)
    L0 = op.forward(theta, phi)
    gt, gp = op.backward()
    Lc, ct, cp = oracle.bpm_mask_loss_grad(theta, phi, index, target, dx, z, steps, oracle.ctx_for(wl))
    print(f"bpm mask 1024: loss {L0:.9e} vs CPU {Lc:.9e}; grad theta {rel_l2(gt, ct):.2e}, grad phi {rel_l2(gp, cp):.2e}",
          flush=True)
    assert abs(L0 - Lc) <= 1e-12 * Lc and rel_l2(gt, ct) <= 1e-10 and rel_l2(gp, cp) <= 1e-10
    h = 1e-3
    worst = 0.0
    # pixels near the beam, where the gradient is not negligible
    picks = [("t", None, 512, 520), ("t", None, 500, 512), ("p", 0, 515, 505), ("p", 9, 512, 530), ("p", 19, 520, 512)]
    for kind, s, j, i in picks:
        tp, tm, pp, pm = theta.copy(), theta.copy(), phi.copy(), phi.copy()
        if kind == "t":
            tp[j, i] += h
            tm[j, i] -= h
            g = gt[j, i]
        else:
            pp[s, j, i] += h
            pm[s, j, i] -= h
            g = gp[s, j, i]
        fd = (op.forward(tp, pp) - op.forward(tm, pm)) / (2 * h)
        err = abs(fd - g) / abs(fd)
        worst = max(worst, err)
        assert err <= 1e-4, (kind, s, j, i, fd, g)
    print(f"bpm mask 1024 FD: worst relative error {worst:.2e}", flush=True)
    op.close()
