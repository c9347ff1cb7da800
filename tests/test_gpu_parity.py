"""GPU parity: libwos_hip (through its C ABI) against the FP64 oracle.

Tolerances: FP64 paths match the oracle to rel-L2 <= 1e-12 (the bar of the
FP64 GPU path); the drop-in FFT max error is also held to the reference's own
1e-9 GPU-vs-CPU bound (optics/gpu_test.go:62,153) tightened to 1e-12. FP32
paths are held to rel-L2 <= 1e-5 against the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def rel_l2(a, b):
    den = np.sum(np.abs(b) ** 2)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / den)) if den > 0 else 0.0


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


def test_device_is_gfx950(wos):
    name = wos.device_name()
    print("device:", name, flush=True)
    assert "gfx950" in name


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048])
def test_fft2d_pow2_matches_oracle(wos, oracle, n):
    """TestGPUFFT2DMatchesCPU (optics/gpu_test.go:44-75) on refField inputs."""
    src = oracle.ref_field(n, 1.3)
    cpu = oracle.fft2d(src)
    gpu = wos.fft2d(src)
    assert rel_max(gpu, cpu) <= FP64_TOL
    assert rel_l2(gpu, cpu) <= FP64_TOL
    back = wos.fft2d(gpu, inverse=True)
    assert np.max(np.abs(back - src)) <= 1e-12


@pytest.mark.parametrize("n", [64, 1024, 4096, 8192])
def test_fft1d_pow2_matches_oracle(wos, oracle, n):
    """TestGPUFFT1DMatchesCPU (optics/gpu_test.go:77-108)."""
    i = np.arange(n)
    src = np.sin(i * 0.01) + 1j * np.cos(i * 0.017)
    cpu = oracle.fft1d(src)
    gpu = wos.fft1d(src)
    assert rel_max(gpu, cpu) <= FP64_TOL
    back = wos.fft1d(gpu, inverse=True)
    assert np.max(np.abs(back - src)) <= 1e-12


@pytest.mark.parametrize("n", [64, 512, 2048])
def test_fft2d_fp32_error_vs_fp64_oracle(wos, oracle, n):
    src = oracle.ref_field(n, 0.7)
    cpu = oracle.fft2d(src)
    gpu = wos.fft2d(src, precision=wos.FP32)
    err = rel_l2(gpu.astype(np.complex128), cpu)
    print(f"n={n} FP32 fft2d rel-L2 {err:.3e}", flush=True)
    assert err <= FP32_TOL


def config1_input(oracle):
    """512^2, dx=2um, 632.8nm Gaussian w0=200um through a thin lens f=50mm."""
    n, dx, wl = 512, 2e-6, 632.8e-9
    ex = oracle.gaussian(n, dx, 200e-6)
    oracle.apply_lens(ex, None, dx, wl, 0.05)
    return ex, n, dx, wl


def test_config1_asm_fp64_parity(wos, oracle):
    ex, n, dx, wl = config1_input(oracle)
    ctx_o = oracle.ctx_for(wl)
    ctx_g = wos.ctx_for(wl)
    ref, _, wo, so = oracle.propagate(ex, None, dx, 0.05, "asm", ctx_o)
    out, _, wg, sg = wos.propagate(ex, None, dx, 0.05, "asm", ctx_g)
    err = rel_l2(out, ref)
    print(f"config1 asm FP64 rel-L2 {err:.3e}", flush=True)
    assert err <= FP64_TOL
    assert abs(sg.p_in - so.p_in) <= 1e-12 * so.p_in
    assert abs(sg.p_out - so.p_out) <= 1e-12 * so.p_out
    assert wg.as_dict().keys() == wo.as_dict().keys()


def test_config1_asm_fp32_error(wos, oracle):
    ex, n, dx, wl = config1_input(oracle)
    ref, _, _, _ = oracle.propagate(ex, None, dx, 0.05, "asm", oracle.ctx_for(wl))
    out, _, _, _ = wos.propagate(ex, None, dx, 0.05, "asm", wos.ctx_for(wl), precision=wos.FP32)
    err = rel_l2(out, ref)
    print(f"config1 asm FP32 rel-L2 {err:.3e}", flush=True)
    assert err <= FP32_TOL


def gaussian_tilted(oracle, n, width, wl, waist, theta):
    """buildGaussianTilted without NormalizePower (optics/asm_accuracy_test.go:10-16)."""
    dx = width / n
    ex = oracle.gaussian(n, dx, waist)
    k = 2 * np.pi / wl
    x = (np.arange(n) - n / 2) * dx
    ex = ex * np.exp(1j * k * np.sin(theta) * x)[None, :]
    return ex, dx


@pytest.mark.parametrize("method", ["asm", "asm_pad", "asm_shift", "asm_shift_pad", "fresnel_tf", "fresnel_ir"])
@pytest.mark.parametrize("jones", [False, True])
def test_methods_fp64_parity(wos, oracle, method, jones):
    n, wl = 256, 632.8e-9
    ex, dx = gaussian_tilted(oracle, n, 0.004, wl, 4e-4, 0.004)
    ey = 0.5 * ex[::-1, :].copy() if jones else None
    z = 0.05
    ref_x, ref_y, wo, so = oracle.propagate(ex, ey, dx, z, method, oracle.ctx_for(wl))
    out_x, out_y, wg, sg = wos.propagate(ex, ey, dx, z, method, wos.ctx_for(wl))
    err = rel_l2(out_x, ref_x)
    if jones:
        err = max(err, rel_l2(out_y, ref_y))
    print(f"{method} jones={jones} FP64 rel-L2 {err:.3e}", flush=True)
    assert err <= FP64_TOL
    assert sg.method_used == so.method_used
    assert wg.as_dict().keys() == wo.as_dict().keys()
    for code, (cnt, val) in wo.as_dict().items():
        assert wg.as_dict()[code][0] == cnt
        assert abs(wg.as_dict()[code][1] - val) <= 1e-9 * max(abs(val), 1e-300)


@pytest.mark.parametrize("method", ["asm_shift", "asm_shift_pad"])
@pytest.mark.parametrize("n", [256, 2048])
def test_shifted_methods_fp32_centroid(wos, oracle, method, n):
    """The FP32 path's spectral centroid is a parallel reduction (the FP64
    path keeps the reference's serial order): the carrier matches the
    oracle's to FP32 rounding and the shifted propagation to the FP32
    tolerance, on a grid where the centroid covers 4M (2048^2) pixels."""
    wl = 632.8e-9
    ex, dx = gaussian_tilted(oracle, n, n * 1.6e-5, wl, n * 1.6e-5 / 10, 0.004)
    ref, _, _, so = oracle.propagate(ex, None, dx, 0.05, method, oracle.ctx_for(wl))
    out, _, _, sg = wos.propagate(ex, None, dx, 0.05, method, wos.ctx_for(wl), precision=wos.FP32)
    err = rel_l2(out, ref)
    print(f"{method} n={n} FP32 rel-L2 {err:.3e}; carrier fx {sg.fx_c:.6e} (oracle {so.fx_c:.6e})", flush=True)
    assert sg.method_used == so.method_used
    assert abs(sg.fx_c - so.fx_c) <= 1e-6 * abs(so.fx_c) and abs(sg.fy_c - so.fy_c) <= 1e-6 * abs(so.fx_c)
    assert err <= FP32_TOL


@pytest.mark.parametrize("z", [-0.02, 0.02])
@pytest.mark.parametrize("reg", [False, True])
def test_evanescent_policies_parity(wos, oracle, z, reg):
    """Evanescent content (optics/asm_improve_test.go:169-253): forward decay,
    backward zeroing or Tikhonov."""
    n, dx, wl = 128, 0.25e-6, 1e-6
    rng = np.random.default_rng(3)
    ex = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    zz = z * 1e-4
    co = oracle.ctx_for(wl, backward_regularize=reg, alpha=1e-3)
    cg = wos.ctx_for(wl, backward_regularize=reg, alpha=1e-3)
    ref, _, wo, _ = oracle.propagate(ex, None, dx, zz, "asm", co)
    out, _, wg, _ = wos.propagate(ex, None, dx, zz, "asm", cg)
    assert rel_l2(out, ref) <= FP64_TOL
    assert wg.as_dict().keys() == wo.as_dict().keys()


EVAN_CASES = [(p, n, None, m) for p in ("fp32", "fp64") for n, m in ((256, "asm"), (2048, "asm"), (1000, "asm"),
                                                                    (1000, "asm_pad"))]
EVAN_CASES += [("fp32", 2048, (8, 8, 2, 2, 0, 16, 32, 1), "asm")]


@pytest.mark.parametrize("precision,n,shape,method", EVAN_CASES)
@pytest.mark.parametrize("z,reg", [(0.6e-6, False), (-0.6e-6, False), (-0.6e-6, True)])
def test_batch_evanescent_parity(wos, oracle, precision, n, shape, method, z, reg):
    """Batch column passes on grids with evanescent bins (dx = 0.3 um:
    rho reaches 2 and more in the corners), where the transfer function
    takes the decay, backward zeroing or Tikhonov factor
    (optics/propagate.go:168-180) and the mirrored half of each column
    takes its partner's values: the FP32 power-of-two passes evaluate the
    propagating factor everywhere and fix evanescent rows up afterwards.
    The second unit's wavelength is M dx / K with a Pythagorean K (100 at
    256 points, 300 at 1000, 600 for the 2000-point asm_pad transforms, 500
    at 2048), so bins such as (3K/5, 4K/5) lie on rho = 1, where the
    reference's own test picks the branch (zeroing: the whole value).
    Power-of-two and mixed-radix sides (1000 asm_pad: register-resident in
    FP64), 2048 points also under the 32-value column schedule; every unit
    matches the oracle's Propagate at its wavelength and distance to the
    precision's tolerance."""
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    M = 2 * n if method == "asm_pad" else n
    K = {256: 100, 1000: 300, 2000: 600, 2048: 500}[M]
    dx = 0.3e-6
    wls, zs = np.array([0.9e-6, M * dx / K]), np.array([z, 1.5 * z])
    rng = np.random.default_rng(n)
    srcs = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(2)]
    b = wos.Batch(n, dx, wls, zs, wos.ctx_for(1e-6, backward_regularize=reg, alpha=1e-3), method, prec)
    if shape is not None:
        b.set_tuning(*shape)
    for u, f in enumerate(srcs):
        b.upload(u, f)
    b.run(1)
    b.sync()
    for u in range(2):
        want, _, _, _ = oracle.propagate(srcs[u], None, dx, zs[u], method,
                                         oracle.ctx_for(wls[u], backward_regularize=reg, alpha=1e-3))
        e = rel_l2(b.download(u), want)
        print(f"evanescent {precision} {method} n={n} shape={shape} z={zs[u]:.2e} reg={reg} unit {u} "
              f"rel-L2 {e:.3e}", flush=True)
        assert e <= (FP64_TOL if prec == wos.FP64 else FP32_TOL), (n, method, shape, z, reg, u, e)
    b.close()


@pytest.mark.parametrize("n,method", [(256, "asm"), (256, "asm_shift"), (1000, "asm"), (1000, "asm_pad"),
                                      (1000, "asm_shift")])
@pytest.mark.parametrize("z,reg", [(0.6e-6, False), (-0.6e-6, False), (-0.6e-6, True)])
def test_propagate_fp32_evanescent_parity(wos, oracle, n, method, z, reg):
    """The FP32 drop-in Propagate with evanescent bins, at a wavelength
    M dx / K with a Pythagorean K (bins on rho = 1: optics/propagate.go:
    162-180 decides their branch), including the shifted-spectrum
    transfer (asm_shift: a carrier tilt moves the spectrum, the generic
    FP32 transfer function takes the reference's test within 1e-14 of the
    cut-off), against the oracle to the FP32 tolerance."""
    M = 2 * n if method == "asm_pad" else n
    dx = 0.3e-6
    wl = M * dx / {256: 100, 1000: 300, 2000: 600}[M]
    rng = np.random.default_rng(n + 1)
    x = (np.arange(n) - n / 2) * dx
    tilt = np.exp(2j * np.pi * 0.2 / wl * x)[None, :]   # a carrier for asm_shift's centroid
    src = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * tilt
    ctx_o = oracle.ctx_for(wl, backward_regularize=reg, alpha=1e-3)
    ctx_g = wos.ctx_for(wl, backward_regularize=reg, alpha=1e-3)
    want, _, _, so = oracle.propagate(src, None, dx, z, method, ctx_o)
    got, _, _, sg = wos.propagate(src, None, dx, z, method, ctx_g, precision=wos.FP32)
    e = rel_l2(got, want)
    print(f"drop-in FP32 {method} n={n} z={z:.1e} reg={reg} rel-L2 {e:.3e}", flush=True)
    assert sg.method_used == so.method_used
    assert e <= FP32_TOL, (n, method, z, reg, e)


def test_split_step_fp64_parity(wos, oracle):
    """PropagateSplitStep on a graded-index fibre (config-3 medium, smaller grid)."""
    n, dx, wl = 128, 0.5e-6, 1.55e-6
    x = (np.arange(n) - n / 2) * dx
    X, Y = np.meshgrid(x, x)
    r = np.hypot(X, Y)
    nre = np.where(r < 100e-6, 1.45 - 0.01 * (r / 100e-6) ** 2, 1.44)
    index = nre + 1j * 1e-5
    ex = oracle.gaussian(n, dx, 10e-6)
    ref, _, wo = oracle.split_step(ex, None, dx, 200e-6, 40, index, oracle.ctx_for(wl))
    out, _, wg = wos.split_step(ex, None, dx, 200e-6, 40, index, wos.ctx_for(wl))
    err = rel_l2(out, ref)
    print(f"split-step FP64 rel-L2 {err:.3e}", flush=True)
    assert err <= FP64_TOL
    assert wg.as_dict().keys() == wo.as_dict().keys()


def test_polychromatic_and_volume_fp64_parity(wos, oracle):
    n, wl = 128, 632.8e-9
    ex, dx = gaussian_tilted(oracle, n, 0.004, wl, 4e-4, 0.0)
    wls = [450e-9, 532e-9, 632.8e-9]
    wts = [0.5, 1.0, 0.75]
    ref, _ = oracle.polychromatic(ex, None, dx, 0.05, wls, wts, "asm", oracle.ctx_for(wl))
    out, _ = wos.polychromatic(ex, None, dx, 0.05, wls, wts, "asm", wos.ctx_for(wl))
    assert rel_l2(out, ref) <= FP64_TOL
    zs = [0.0, 0.01, 0.05, -0.02]
    rv, _, _ = oracle.volume(ex, None, dx, zs, "asm", oracle.ctx_for(wl))
    gv, _, _ = wos.volume(ex, None, dx, zs, "asm", wos.ctx_for(wl))
    for i in range(len(zs)):
        assert rel_l2(gv[i], rv[i]) <= FP64_TOL


@pytest.mark.parametrize("n", [256, 1024])
def test_propagate1d_fp64_parity(wos, oracle, n):
    dx, wl = 1e-5, 632.8e-9
    x = (np.arange(n) - n / 2) * dx
    a = np.exp(-(x / 3e-4) ** 2) * np.exp(1j * 0.3 * x / dx)
    ref, wo = oracle.propagate1d(a, dx, 0.3, wl, oracle.ctx_for(wl))
    out, wg = wos.propagate1d(a, dx, 0.3, wl, wos.ctx_for(wl))
    assert rel_l2(out, ref) <= FP64_TOL
    assert wg.as_dict().keys() == wo.as_dict().keys()


# ---------------------------------------------------------------- non-power-of-two (Bluestein)

@pytest.mark.parametrize("n", [12, 27, 100, 384, 1000, 11, 94, 1018])
def test_fft2d_bluestein_matches_oracle(wos, oracle, n):
    """TestGPUFFT2DMatchesCPU's non-power-of-two sizes (optics/gpu_test.go:54)
    against the oracle's chirp-z path (optics/fft.go:159-204): 12, 27, 100,
    384 and 1000 run the mixed-radix kernels, 11, 94 = 2 47 and 1018 = 2 509
    the device Bluestein path."""
    src = oracle.ref_field(n, 1.3)
    cpu = oracle.fft2d(src)
    gpu = wos.fft2d(src)
    err = rel_l2(gpu, cpu)
    print(f"n={n} bluestein fft2d rel-L2 {err:.3e} max {rel_max(gpu, cpu):.3e}", flush=True)
    assert err <= FP64_TOL and rel_max(gpu, cpu) <= FP64_TOL
    back = wos.fft2d(gpu, inverse=True)
    assert np.max(np.abs(back - src)) <= 1e-11


@pytest.mark.parametrize("n", [100, 1000, 2049])
def test_fft1d_bluestein_matches_oracle(wos, oracle, n):
    i = np.arange(n)
    src = np.sin(i * 0.01) + 1j * np.cos(i * 0.017)
    cpu = oracle.fft1d(src)
    gpu = wos.fft1d(src)
    assert rel_max(gpu, cpu) <= FP64_TOL
    assert np.max(np.abs(wos.fft1d(gpu, inverse=True) - src)) <= 1e-11


@pytest.mark.parametrize("method", ["asm", "asm_pad", "asm_shift", "asm_shift_pad", "fresnel_tf", "fresnel_ir"])
def test_methods_non_pow2_grid_fp64_parity(wos, oracle, method):
    """ASM family on an even non-power-of-two grid (TestASMNonPowerOfTwoGrid,
    optics/asm_improve_test.go:87-107)."""
    n, wl = 100, 632.8e-9
    ex, dx = gaussian_tilted(oracle, n, 0.004, wl, 4e-4, 0.003)
    ref, _, wo, so = oracle.propagate(ex, None, dx, 0.03, method, oracle.ctx_for(wl))
    out, _, wg, sg = wos.propagate(ex, None, dx, 0.03, method, wos.ctx_for(wl))
    err = rel_l2(out, ref)
    print(f"{method} n=100 FP64 rel-L2 {err:.3e}", flush=True)
    assert err <= FP64_TOL
    assert sg.method_used == so.method_used
    assert wg.as_dict().keys() == wo.as_dict().keys()


@pytest.mark.parametrize("n,method", [(384, "asm"), (384, "asm_pad"), (1536, "asm"), (400, "asm_shift"), (600, "asm_pad"),
                                      (1500, "asm"), (720, "asm"), (1920, "asm")])
def test_mixed_radix_grids_match_oracle(wos, oracle, n, method):
    """Grid sides 2^a 3^b 5^c run the mixed-radix kernels (radix 16/8/4/2
    stages, then 3 and 5) in place of Bluestein: 384 (16, 8, 3), its padded
    768 (16, 16, 3), 1536 (16, 16, 2, 3), 400 (16, 5, 5) with the shifted
    transfer function, 600 padded to 1200 (16, 3, 5, 5) and 1500
    (4, 3, 5, 5, 5), 720 (16, 3, 3, 5) and 1920 (16, 8, 3, 5), against the
    oracle's Bluestein path (the FP64 and FP32
    tolerances)."""
    wl = 632.8e-9
    ex, dx = gaussian_tilted(oracle, n, n * 4e-6, wl, n * 4e-6 / 8, 0.01)
    z = 0.01
    ref, _, wo, so = oracle.propagate(ex, None, dx, z, method, oracle.ctx_for(wl))
    for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
        out, _, wg, sg = wos.propagate(ex, None, dx, z, method, wos.ctx_for(wl), precision=prec)
        err = rel_l2(out, ref)
        print(f"mixed-radix n={n} {method} prec={prec} rel-L2 {err:.3e}", flush=True)
        assert err <= tol
        if prec == wos.FP64:  # FP32 output power moves ~1e-7: the 1e-8 power check is an FP64 property
            assert wg.as_dict().keys() == wo.as_dict().keys()
            # the oracle's Bluestein transform (the reference's algorithm) is itself
            # 1e-13 (384) to 4e-13 (1536) from numpy's FFT in rel-L2
            assert abs(sg.p_out - so.p_out) <= 1e-11 * so.p_out


@pytest.mark.parametrize("n,method", [(90, "asm"), (960, "asm"), (450, "asm_pad"), (1080, "asm"), (1152, "asm"),
                                      (480, "asm_shift"), (1350, "asm"), (700, "asm"), (350, "asm_pad"),
                                      (1470, "asm")])
def test_runtime_mixed_sides_match_oracle(wos, oracle, n, method):
    """2^a 3^b 5^c sides with no compiled plan run the runtime-length
    mixed-radix kernels (stage list as an argument, dynamic LDS): 90
    (2, 3, 3, 5), 960 (16, 4, 3, 5), 450 padded to 900 (4, 3, 3, 5, 5), 1080
    (8, 3, 3, 3, 5), 1152 (16, 8, 3, 3), 480 (16, 2, 3, 5) with the shifted
    transfer function, 1350 (2, 3, 3, 3, 5, 5), and with radix-7 stages 700
    (4, 5, 5, 7), 350 padded to 700 and 1470 (2, 3, 5, 7, 7), against the
    oracle's Bluestein path."""
    wl = 632.8e-9
    ex, dx = gaussian_tilted(oracle, n, n * 4e-6, wl, n * 4e-6 / 8, 0.01)
    z = 0.01
    ref, _, wo, so = oracle.propagate(ex, None, dx, z, method, oracle.ctx_for(wl))
    for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
        out, _, wg, sg = wos.propagate(ex, None, dx, z, method, wos.ctx_for(wl), precision=prec)
        err = rel_l2(out, ref)
        print(f"runtime mixed-radix n={n} {method} prec={prec} rel-L2 {err:.3e}", flush=True)
        assert err <= tol
        if prec == wos.FP64:
            assert wg.as_dict().keys() == wo.as_dict().keys()
            assert abs(sg.p_out - so.p_out) <= 1e-11 * so.p_out


@pytest.mark.parametrize("n", [90, 960, 1800, 3600, 4050, 49, 1400, 3430])
def test_runtime_mixed_fft_matches_numpy(wos, oracle, n):
    """The drop-in fft2D / fft1DAny on runtime-plan sides against numpy's FFT
    (FP64), with the round trip; 4050 (2, 3, 3, 3, 3, 5, 5) takes the
    largest workgroup, 512 threads; 49, 1400 and 3430 (2, 5, 7, 7, 7) run
    radix-7 stages."""
    src = oracle.ref_field(n, 1.3) if n <= 1920 else None
    if src is not None:
        want = np.fft.fft2(src)
        got = wos.fft2d(src)
        err = rel_l2(got, want)
        print(f"runtime mixed fft2d n={n} rel-L2 {err:.3e}", flush=True)
        assert err <= FP64_TOL
        assert np.max(np.abs(wos.fft2d(got, inverse=True) - src)) <= 1e-11
    i = np.arange(n)
    line = np.sin(i * 0.01) + 1j * np.cos(i * 0.017)
    want = np.fft.fft(line)
    got = wos.fft1d(line)
    err = rel_l2(got, want)
    print(f"runtime mixed fft1d n={n} rel-L2 {err:.3e}", flush=True)
    assert err <= FP64_TOL
    assert np.max(np.abs(wos.fft1d(got, inverse=True) - line)) <= 1e-12


def config2_input(oracle, seed=7):
    """1000^2, dx=1.5um, 532nm: x-polarised plane wave times an i.i.d. uniform
    random phase (diffuser hash, one value per pixel) inside a circular
    aperture of diameter 1.2 mm; Ey = 0 (Polarized)."""
    n, dx, wl = 1000, 1.5e-6, 532e-9
    ex = oracle.random_phase(n, seed)
    oracle.apply_circle_aperture(ex, None, dx, 0.6e-3)
    ey = np.zeros_like(ex)
    return ex, ey, n, dx, wl


def test_config2_asm_pad_jones_fp64_and_fp32(wos, oracle):
    ex, ey, n, dx, wl = config2_input(oracle)
    z = 10e-3
    rx, ry, wo, so = oracle.propagate(ex, ey, dx, z, "asm_pad", oracle.ctx_for(wl))
    gx, gy, wg, sg = wos.propagate(ex, ey, dx, z, "asm_pad", wos.ctx_for(wl))
    err = rel_l2(np.concatenate([gx, gy]), np.concatenate([rx, ry]))
    print(f"config2 asm_pad FP64 rel-L2 {err:.3e}", flush=True)
    assert err <= FP64_TOL
    assert wg.as_dict().keys() == wo.as_dict().keys()
    assert abs(sg.p_out - so.p_out) <= 1e-12 * so.p_out
    fx, fy, _, _ = wos.propagate(ex, ey, dx, z, "asm_pad", wos.ctx_for(wl), precision=wos.FP32)
    err32 = rel_l2(np.concatenate([fx, fy]), np.concatenate([rx, ry]))
    print(f"config2 asm_pad FP32 rel-L2 {err32:.3e}", flush=True)
    assert err32 <= FP32_TOL


# ---------------------------------------------------------------- split-step BPM

def graded_index(n, dx):
    """Config-3 medium: n = 1.45 - 0.01 (r/100um)^2 for r < 100um, else 1.44, + 1e-5 i."""
    x = (np.arange(n) - n / 2) * dx
    r = np.hypot(x[None, :], x[:, None])
    return np.where(r < 100e-6, 1.45 - 0.01 * (r / 100e-6) ** 2, 1.44) + 1e-5j


def test_split_step_fp32_error(wos, oracle):
    n, dx, wl = 128, 0.5e-6, 1.55e-6
    index = graded_index(n, dx)
    ex = oracle.gaussian(n, dx, 10e-6)
    ref, _, _ = oracle.split_step(ex, None, dx, 200e-6, 40, index, oracle.ctx_for(wl))
    out, _, _ = wos.split_step(ex, None, dx, 200e-6, 40, index, wos.ctx_for(wl), precision=wos.FP32)
    err = rel_l2(out, ref)
    print(f"split-step FP32 rel-L2 {err:.3e}", flush=True)
    assert err <= FP32_TOL


def test_split_step_z_dependent_planes_and_non_pow2(wos, oracle):
    """Stratified medium (TestSplitStepStratified, optics/medium_test.go:45-66)
    with one sampled plane per slice, on a power-of-two grid (fused pipeline)
    and on a non-power-of-two grid (half steps one by one)."""
    wl, z, steps = 632.8e-9, 2e-3, 8
    dz = z / steps
    for n in (128, 100):
        dx = 2e-6
        planes = np.array([np.full((n, n), (1.5 if (z * s / steps + dz / 2) < z / 2 else 1.0) + 2e-6j)
                           for s in range(steps)])
        ex = oracle.gaussian(n, dx, 40e-6)
        ref, _, wo = oracle.split_step(ex, None, dx, z, steps, planes, oracle.ctx_for(wl))
        out, _, wg = wos.split_step(ex, None, dx, z, steps, planes, wos.ctx_for(wl))
        err = rel_l2(out, ref)
        print(f"stratified n={n} FP64 rel-L2 {err:.3e}", flush=True)
        assert err <= FP64_TOL
        assert wg.as_dict().keys() == wo.as_dict().keys()


def test_split_step_evanescent_warning_parity(wos, oracle):
    """A field with sub-wavelength content loses power in every half step:
    the fused pipeline's spectral power checks must fire like the
    reference's spatial Power() pairs (optics/propagate.go:204-208)."""
    n, dx, wl = 64, 0.25e-6, 1e-6
    rng = np.random.default_rng(11)
    ex = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    index = np.full((n, n), 1.2 + 1e-4j)
    ref, _, wo = oracle.split_step(ex, None, dx, 2e-6, 4, index, oracle.ctx_for(wl))
    out, _, wg = wos.split_step(ex, None, dx, 2e-6, 4, index, wos.ctx_for(wl))
    assert rel_l2(out, ref) <= FP64_TOL
    do, dg = wo.as_dict(), wg.as_dict()
    assert do.keys() == dg.keys() and "evanescent_filtered" in do
    assert dg["evanescent_filtered"][0] == do["evanescent_filtered"][0] == 8
    assert abs(dg["evanescent_filtered"][1] - do["evanescent_filtered"][1]) <= 1e-9 * do["evanescent_filtered"][1]


def test_config3_bpm_short_vs_oracle(wos, oracle):
    """Config-3 grid and medium (1024^2, dx 0.5um, 1.55um, Gaussian w0 10um),
    20 of the 400 slices, FP64 vs the oracle's step-by-step BPM."""
    n, dx, wl, steps = 1024, 0.5e-6, 1.55e-6, 20
    index = graded_index(n, dx)
    ex = oracle.gaussian(n, dx, 10e-6)
    z = steps * 5e-6
    ref, _, _ = oracle.split_step(ex, None, dx, z, steps, index, oracle.ctx_for(wl))
    out, _, _ = wos.split_step(ex, None, dx, z, steps, index, wos.ctx_for(wl))
    err = rel_l2(out, ref)
    print(f"config3 20 steps FP64 rel-L2 {err:.3e}", flush=True)
    assert err <= FP64_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_split_step_radix8_schedules(wos, oracle, precision):
    """The split-step slices under the radix-8 schedule (the screen row pass
    and the H^2 column pass; FP32 with compensated radix-8 twiddles) for the
    row passes, the column pass or both: config-3 grid and medium, 20
    slices, against the oracle (the FP64 and FP32 tolerances) and the
    radix-16 run (rounding level)."""
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    tol = FP64_TOL if prec == wos.FP64 else FP32_TOL
    n, dx, wl, steps = 1024, 0.5e-6, 1.55e-6, 20
    index = graded_index(n, dx)
    ex = oracle.gaussian(n, dx, 10e-6)
    z = steps * 5e-6
    ref, _, wo = oracle.split_step(ex, None, dx, z, steps, index, oracle.ctx_for(wl))
    # (tile width, rows per row workgroup, row radix, column radix); (0, ...): built-in shape
    shapes = [(0, 0, 16, 16), (4, 0, 8, 8), (4, 2, 16, 8), (2, 4, 8, 16), (8, 1, 8, 8)]
    base = None
    try:
        for shape in shapes:
            wos.split_step_set_shape(prec, n, *shape)
            out, _, wg = wos.split_step(ex, None, dx, z, steps, index, wos.ctx_for(wl), precision=prec)
            e = rel_l2(out, ref)
            print(f"split-step {precision} shape {shape}: rel-L2 vs oracle {e:.3e}", flush=True)
            assert e <= tol, (shape, e)
            assert wg.as_dict().keys() == wo.as_dict().keys()
            if base is None:
                base = out
            else:
                d = rel_l2(out, base)
                assert d <= (1e-13 if prec == wos.FP64 else 2e-6), (shape, d)
    finally:
        wos.split_step_set_shape(prec, n, 0)


def test_config3_bpm_full_fp32_vs_fp64(wos, oracle):
    """Config 3 in full: 400 slices of dz = 5um. FP32 bulk path vs the FP64
    path (itself pinned to the oracle above) and the absorbed power against
    exp(-2 k0 Im(n) z) = 0.850316 (the light stays inside the periodic window)."""
    n, dx, wl, steps = 1024, 0.5e-6, 1.55e-6, 400
    index = graded_index(n, dx)
    ex = oracle.gaussian(n, dx, 10e-6)
    z = steps * 5e-6
    p0 = np.sum(np.abs(ex) ** 2)
    o64, _, w64 = wos.split_step(ex, None, dx, z, steps, index, wos.ctx_for(wl))
    t64 = wos.last_device_ms()
    o32, _, w32 = wos.split_step(ex, None, dx, z, steps, index, wos.ctx_for(wl), precision=wos.FP32)
    t32 = wos.last_device_ms()
    err = rel_l2(o32, o64)
    ratio = np.sum(np.abs(o64) ** 2) / p0
    want = np.exp(-2 * (2 * np.pi / wl) * 1e-5 * z)
    print(f"config3 400 steps: FP32 vs FP64 rel-L2 {err:.3e}; power ratio {ratio:.6f} (closed form {want:.6f}); "
          f"device ms FP64 {t64:.2f} ({steps / t64 * 1e3:.0f} steps/s) FP32 {t32:.2f} ({steps / t32 * 1e3:.0f} steps/s)",
          flush=True)
    assert err <= FP32_TOL
    assert abs(ratio - want) <= 1e-3
    assert w32.as_dict().keys() == w64.as_dict().keys()


# ---------------------------------------------------------------- launch shapes

@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_launch_shapes_bit_identical(wos, precision):
    """Layout tile width, column and row workgroup shapes change only the
    data movement: the one-group kernels give bit-identical results. The
    pipelined row kernels (row groups > 0, twiddles in registers or, with
    split exchanges, in LDS) and the split-exchange row kernel are other
    instantiations of the same arithmetic, where the compiler may
    fuse a complex product's multiply-adds the other way round: they must
    agree to rounding level."""
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    n = 1024
    b = wos.Batch(n, 2e-6, np.array([450e-9, 633e-9]), np.array([5e-3, 9e-3]), wos.ctx_for(550e-9), "asm", prec)
    # (layout tile width, columns per column workgroup, rows per row workgroup, row groups, split LDS
    #  [, the inverse row pass's rows, groups, split LDS])
    shapes = [(0, 0, 0, 0, 0), (1, 1, 1, 0, 0), (2, 2, 8, 0, 0), (4, 4, 4, 2, 0), (4, 2, 2, 4, 0), (8, 8, 1, 2, 0),
              (2, 1, 1, 4, 0), (8, 8, 8, 4, 0), (8, 4, 2, 2, 0), (8, 2, 4, 0, 0), (16, 4, 1, 4, 0), (4, 4, 4, 0, 1),
              (8, 4, 2, 0, 1), (2, 2, 8, 0, 1),
              (4, 4, 4, 4, 1), (4, 4, 4, 2, 1), (8, 4, 2, 4, 1), (8, 8, 2, 2, 1),   # pipelined, twiddles in LDS
              (4, 4, 4, 2, 1, 2, 2, 1), (4, 4, 4, 0, 1, 2, 0, 0), (8, 8, 1, 0, 0, 8, 4, 0),   # own inverse shapes
              (8, 8, 4, 2, 1, 2, 2, 1, 32), (4, 4, 4, 0, 1, -1, 0, 0, 32)]   # 32-value column schedule
    if prec == wos.FP32:
        shapes.append((16, 16, 4, 4, 0))   # 16-column workgroups of 1024-point columns fit LDS in FP32 only
    ref = None
    for tw, cw, rows, groups, split, *inv in shapes:
        col_radix, col_split = (inv[3], 1) if len(inv) > 3 else (None, 0)
        inv = inv[:3]
        if tw:
            b.set_tuning(tw, cw, rows, groups, split, 16, col_radix, col_split, *(inv or (-1, 0, 0)))
        b.fill_schell(100e-6, 30e-6, 3)
        b.run(1)
        b.sync()
        out = [b.download(u) for u in range(2)]
        if ref is None:
            ref = out
        else:
            for u in range(2):
                if groups == 0 and split == 0 and (not inv or inv[1:] == [0, 0]) and col_radix is None:
                    assert np.array_equal(out[u], ref[u]), (tw, cw, rows, groups, split, u)
                else:
                    e = rel_l2(out[u], ref[u])
                    assert e <= (1e-15 if prec == wos.FP64 else 1e-6), (tw, cw, rows, groups, split, inv, u, e)
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n,method", [(256, "asm"), (1024, "asm"), (512, "asm_pad"), (2048, "asm")])
def test_batch_radix8_schedule(wos, oracle, precision, n, method):
    """The radix-8 stage schedule (8 values per thread, radix-8 stages and a
    radix-4/2 remainder) is another factorisation of the same DFTs: under
    each launch shape it builds, for the row passes, the column pass or
    both, every unit of a batch run matches the radix-16 run to rounding
    and, up to 1024 points, the oracle's Propagate at the unit's wavelength
    and distance (the FP64 and FP32 tolerances). asm_pad covers the
    zero-padded embed window (M = 2n) of the row passes."""
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    M = 2 * n if method == "asm_pad" else n
    dx, wls, zs = 2e-6, np.array([450e-9, 633e-9]), np.array([5e-3, 9e-3])
    rng = np.random.default_rng(n)
    srcs = [oracle.gaussian(n, dx, n * dx / 6) * np.exp(1j * rng.uniform(-0.3, 0.3, (n, n))) for _ in range(2)]
    b = wos.Batch(n, dx, wls, zs, wos.ctx_for(550e-9), method, prec)

    def run(shape):
        b.set_tuning(*shape)
        for u, f in enumerate(srcs):
            b.upload(u, f)
        b.run(1)
        b.sync()
        return [b.download(u) for u in range(2)]

    ref16 = run((4, 4, 0, 0, 0, 16, 16))
    # (layout tile width, columns per column workgroup, rows per row workgroup, row groups)
    shapes8 = [(4, 4, 0, 0), (8, 4, 2, 2), (8, 4, 4, 4), (16, 4, 1, 0)]
    if M >= 1024:   # one- and two-column workgroups: 1024 points and up
        shapes8 += [(2, 1, 1, 2), (4, 2, 4, 4)]
    # the row and column passes take their schedules independently
    runs = [(tw, cw, rows, groups, 0, 8, 8) for tw, cw, rows, groups in shapes8]
    runs += [(4, 4, 0, 0, 1, 16, 8), (8, 4, 2, 2, 0, 8, 16), (4, 4, 0, 0, 1, 16, 16, 0), (4, 4, 0, 0, 1, 16, 16, 1),
             (8, 4, 2, 2, 0, 8, 16, 1)]
    outs = {}
    for shape in runs:
        out = outs[shape] = run(shape)
        for u in range(2):
            e = rel_l2(out[u], ref16[u])
            assert e <= (1e-14 if prec == wos.FP64 else 1e-6), (n, method, shape, u, e)
    # split column exchanges move the same values through half the LDS; as another
    # instantiation the compiler may contract a product's multiply-adds differently
    for u in range(2):
        e = rel_l2(outs[(4, 4, 0, 0, 1, 16, 16, 1)][u], outs[(4, 4, 0, 0, 1, 16, 16, 0)][u])
        assert e <= (1e-15 if prec == wos.FP64 else 1e-6), (u, e)
    if n <= 1024:
        out = outs[runs[0]]   # radix-8 row and column passes
        for u in range(2):
            want, _, _, _ = oracle.propagate(srcs[u], None, dx, zs[u], method, oracle.ctx_for(wls[u]))
            e = rel_l2(out[u], want)
            print(f"radix-8 {method} n={n} {precision} unit {u} rel-L2 vs oracle {e:.3e}", flush=True)
            assert e <= (FP64_TOL if prec == wos.FP64 else FP32_TOL), (n, method, u, e)
    if n == 256:
        tune, ms = b.autotune(1)
        assert tune["radix"] in (8, 16) and tune["col_radix"] in (8, 16) and ms > 0
        if tune["radix"] == 8:
            assert tune["split_lds"] == 0
    with pytest.raises(RuntimeError):
        b.set_tuning(4, 4, 0, 0, 1, 8, 16)   # split exchanges are radix-16 row kernels only
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n,method", [(1000, "asm"), (1080, "asm"), (500, "asm_pad"), (1000, "asm_pad"),
                                      (1536, "asm")])
def test_batch_mixed_radix_tile_widths(wos, oracle, precision, n, method):
    """On mixed-radix grids (1000 = 8 5 5 5 and 500 padded to 1000 on the
    compiled plan, 1080 = 8 27 5 on the runtime plan, and the sides with a
    factor 16, where 16-column layout tiles build: 1000 padded to 2000 =
    16 125 and 1536 = 16 96) every layout tile and column workgroup width
    the side builds (every power of two up to the tile) gives each unit of a
    batch run to rounding of the oracle's Propagate, shapes the side does
    not build are refused, and autotune settles on a built shape."""
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    tol = FP64_TOL if prec == wos.FP64 else FP32_TOL
    dx, wls, zs = 2e-6, np.array([450e-9, 633e-9]), np.array([5e-3, 9e-3])
    rng = np.random.default_rng(n)
    srcs = [oracle.gaussian(n, dx, n * dx / 6) * np.exp(1j * rng.uniform(-0.3, 0.3, (n, n))) for _ in range(2)]
    refs = [oracle.propagate(f, None, dx, z, method, oracle.ctx_for(w))[0] for f, w, z in zip(srcs, wls, zs)]
    b = wos.Batch(n, dx, wls, zs, wos.ctx_for(550e-9), method, prec)
    built = []
    for tw in (16, 8, 4, 2, 1):
        for cw in (c for c in (16, 8, 4, 2, 1) if c <= tw):
            try:
                b.set_tuning(tw, cw)
            except wos.WosError:
                continue
            built.append((tw, cw))
            for u, f in enumerate(srcs):
                b.upload(u, f)
            b.run(1)
            b.sync()
            for u in range(2):
                e = rel_l2(b.download(u), refs[u])
                print(f"mixed tile n={n} {method} prec={prec} tile {tw} columns {cw} unit {u} rel-L2 {e:.2e}",
                      flush=True)
                assert e <= tol
    assert len(built) >= 4 and any(tw != cw for tw, cw in built), built
    if (n if method == "asm" else 2 * n) % 16 == 0:
        assert any(tw == 16 for tw, _ in built), built
    for bad in ((4, 8), (3, 3), (32, 4)):   # columns wider than the tile, a tile not a power of two, too wide
        with pytest.raises(wos.WosError):
            b.set_tuning(*bad)
    shape, ms = b.autotune(2)
    print(f"mixed tile n={n} {method} prec={prec}: built {built}, autotune {shape['tile_w']}/{shape['col_w']} "
          f"({ms:.3f} ms)", flush=True)
    assert (shape["tile_w"], shape["col_w"]) in built
    b.close()
