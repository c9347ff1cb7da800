"""Edge cases of the C ABI on the GPU: the smallest and largest grids, the
largest Bluestein length, z = 0, and the error returns for invalid and
unsupported arguments (nothing is computed for them)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def test_one_point_transforms_are_identity(wos):
    a = np.array([[3.0 - 2.0j]])
    for inv in (False, True):
        assert np.array_equal(wos.fft2d(a, inverse=inv), a)
        assert np.array_equal(wos.fft1d(a.ravel(), inverse=inv), a.ravel())


def test_largest_power_of_two_grid(wos, oracle):
    """8192^2 complex128 (1 GiB): forward and inverse against the oracle."""
    n = 8192
    rng = np.random.default_rng(8)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 0.5
    g = wos.fft2d(a)
    r = oracle.fft2d(a)
    e = rel_l2(g, r)
    print(f"8192^2 fft2d FP64 rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-12
    back = wos.fft2d(g, inverse=True)
    assert rel_l2(back, a) <= 1e-12


@pytest.mark.parametrize("n", [1000, 3000])
def test_host_transfers_round_trip(wos, n):
    """Host <-> device copies go through a ring of eight 8 MiB pinned chunks:
    n = 1000 (16 MB, one full chunk and a partial one) and n = 3000 (144 MB,
    18 chunks, every slot reused) come back bit-exact in FP64, and equal to
    the float32 rounding in FP32."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for prec in (wos.FP64, wos.FP32):
        b = wos.Batch(n, 1e-6, [633e-9, 532e-9], [1e-3, 1e-3], wos.ctx_for(633e-9), "asm", prec)
        b.upload(1, a)
        b.upload(0, -a)
        got1, got0 = b.download(1), b.download(0)
        want = a if prec == wos.FP64 else a.astype(np.complex64).astype(np.complex128)
        assert np.array_equal(got1, want) and np.array_equal(got0, -want), (n, prec)
        b.close()


def test_largest_bluestein_length(wos):
    """L = 4095 (internal 8192-point transforms) in 1-D against numpy."""
    n = 4095
    rng = np.random.default_rng(4)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(wos.fft1d(a), np.fft.fft(a)) <= 1e-12
    assert rel_l2(wos.fft1d(a, inverse=True), np.fft.ifft(a)) <= 1e-12


def test_zero_distance_is_a_no_op(wos):
    rng = np.random.default_rng(5)
    ex = rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))
    for m in ("asm", "asm_pad", "fresnel_tf", "fresnel_ir"):
        out, _, w, _ = wos.propagate(ex, None, 1e-6, 0.0, m, wos.ctx_for(633e-9))
        assert np.array_equal(out, ex) and not w.as_dict()


def test_invalid_and_unsupported_arguments(wos):
    lib = wos.load()
    a = np.zeros(16, dtype=np.complex128)
    ctx = wos.ctx_for(633e-9)
    # invalid: null data, non-positive sizes, bad method, bad precision, bad dx
    assert lib.wos_hip_fft2d_z(None, 4, 0) == 2
    assert lib.wos_hip_fft2d_z(a.ctypes.data, 0, 0) == 2
    assert lib.wos_hip_fft1d_z(a.ctypes.data, -3, 0) == 2
    b = np.ones((4, 4), dtype=np.complex128)
    assert lib.wos_hip_propagate(b.ctypes.data, None, 4, 1e-6, 1e-3, 99, C.byref(ctx), 0, None, None) in (2, 4)
    assert lib.wos_hip_propagate(b.ctypes.data, None, 4, 1e-6, 1e-3, 0, C.byref(ctx), 7, None, None) == 2
    assert lib.wos_hip_propagate(b.ctypes.data, None, 4, 0.0, 1e-3, 0, C.byref(ctx), 0, None, None) == 2
    assert np.array_equal(b, np.ones((4, 4)))                    # nothing written
    # unsupported: beyond the largest transform (powers of two up to 2^26,
    # other lengths whose Bluestein length is at most 2^26); refused before
    # the buffer is read
    big = np.zeros(1, dtype=np.complex128)
    assert lib.wos_hip_fft2d_z(big.ctypes.data, 1 << 27, 0) == 4
    assert lib.wos_hip_fft1d_z(big.ctypes.data, 1 << 27, 0) == 4
    assert lib.wos_hip_fft1d_z(big.ctypes.data, (1 << 25) + 1, 0) == 4
    assert b"not supported" in lib.wos_hip_last_error()
    # a handle-based call with a bad shape reports instead of launching
    bt = wos.Batch(256, 1e-6, [633e-9], [1e-3], ctx, "asm", wos.FP32)
    with pytest.raises(wos.WosError):
        bt.set_tuning(3, 3, 1, 0)
    with pytest.raises(wos.WosError):
        bt.set_tuning(4, 4, 3, 0)
    bt.close()


def test_fresnel_ir_spectrum_cache(wos, oracle):
    """fresnel_ir keeps the impulse-response spectrum of the last (precision,
    grid, dx, wl, z): repeated calls reuse it, a change of z or precision
    rebuilds it, and every result equals the first (bit for bit) and the
    oracle (1e-12)."""
    n, dx, wl = 128, 4e-6, 633e-9
    rng = np.random.default_rng(9)
    ex = np.exp(1j * rng.uniform(0, 0.5, (n, n))) * np.exp(-((np.arange(n) - n / 2) ** 2)[None, :] / 400.0)
    ctx = wos.ctx_for(wl)
    first = {}
    for z in (0.01, 0.02, 0.01, 0.01):
        for prec in (wos.FP64, wos.FP32, wos.FP64):
            out, _, _, _ = wos.propagate(ex, None, dx, z, "fresnel_ir", ctx, precision=prec)
            key = (z, prec)
            if key in first:
                assert np.array_equal(out, first[key]), key
            else:
                first[key] = out
    for z in (0.01, 0.02):
        ref, _, _, _ = oracle.propagate(ex, None, dx, z, "fresnel_ir", oracle.ctx_for(wl))
        assert rel_l2(first[(z, wos.FP64)], ref) <= 1e-12
        assert rel_l2(first[(z, wos.FP32)], ref) <= 1e-5


def test_concurrent_callers_match_serial_runs(wos):
    """The C ABI serialises every call on a per-device lock: drop-in
    propagations, 2-D FFTs and a resident-field element chain issued from
    six host threads at once give the bit-exact results of serial runs."""
    import threading
    rng = np.random.default_rng(12)
    fields = [rng.standard_normal((128, 128)) + 1j * rng.standard_normal((128, 128)) for _ in range(6)]
    ctx = wos.ctx_for(633e-9)

    def work(a, prec):
        out, _, _, _ = wos.propagate(a, None, 2e-6, 3e-3 if prec == wos.FP64 else 5e-3, "asm_pad", ctx, precision=prec)
        spec = wos.fft2d(a)
        f = wos.Field.from_host(a, dx=2e-6, precision=prec)
        f.lens(633e-9, 0.05)
        f.element(wos.make_element("grating", kind="phase_sin", period=31e-6, modulation=0.8), 633e-9)
        f.propagate(2e-3, "asm", ctx, bandlimit=(0.9, 0.05))
        got = f.download()[0]
        f.close()
        return out, spec, got

    precs = [wos.FP64, wos.FP32] * 3
    serial = [work(a, p) for a, p in zip(fields, precs)]
    results = [None] * 6
    errors = []

    def run(i):
        try:
            for _ in range(3):
                results[i] = work(fields[i], precs[i])
        except Exception as exc:  # reported below
            errors.append(exc)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(6):
        for a, b in zip(results[i], serial[i]):
            assert np.array_equal(a, b), i


def test_grid_memory_budget_refuses_before_allocating(wos):
    """CheckGridMemory on the device (optics/memory.go:74-81): the headline
    batch fits the budget (75% of the device's memory); a batch or pipeline
    stage far beyond it is refused with the reference's kind of message
    before anything is allocated, and the library keeps working."""
    need, budget = wos.check_grid_memory(2048, "asm", wos.FP64, 64)
    print(f"64 x 2048^2 FP64 needs {need / 2**30:.1f} GiB of a {budget / 2**30:.1f} GiB budget", flush=True)
    assert 0 < need < budget
    with pytest.raises(wos.WosError, match="would need .* of device memory"):
        wos.check_grid_memory(2048, "asm", wos.FP64, 5000)
    with pytest.raises(wos.WosError, match="would need"):
        wos.Batch(2048, 2e-6, np.full(5000, 532e-9), np.full(5000, 1e-3), wos.ctx_for(532e-9), "asm", wos.FP64)
    p = wos.Pipeline(wos.FP32)
    with pytest.raises(wos.WosError, match="would need"):
        p.add(4096, 1e-6, np.full(3000, 532e-9), np.full(3000, 1e-3), wos.ctx_for(532e-9), "asm")
    s = p.add(64, 1e-6, np.full(2, 532e-9), np.full(2, 1e-3), wos.ctx_for(532e-9), "asm")
    p.run(1)
    p.sync()
    p.close()
    b = wos.Batch(64, 1e-6, [532e-9], [1e-3], wos.ctx_for(532e-9), "asm", wos.FP64)
    b.run(1)
    b.sync()
    b.close()


def test_failure_injection_on_the_device(wos, oracle, monkeypatch):
    """With the GPU present, a listed entry point fails before it touches the
    caller's array (the fallback runs on the original data), and without the
    variable the same call runs on the device again."""
    from test_abi import fft2d_with_fallback
    a = oracle.ref_field(64, 0.9)
    monkeypatch.setenv("WOS_HIP_FAIL", "wos_hip_fft2d_z")
    got, rc = fft2d_with_fallback(wos, oracle, a)
    assert rc == 3
    want = oracle.fft2d(a)
    assert np.array_equal(got, want)
    assert wos.device_count() > 0
    monkeypatch.delenv("WOS_HIP_FAIL")
    got, rc = fft2d_with_fallback(wos, oracle, a)
    assert rc == 0
    assert float(np.sqrt(np.sum(np.abs(got - want) ** 2) / np.sum(np.abs(want) ** 2))) <= 1e-12
