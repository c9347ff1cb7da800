"""The propagation that allocates nothing (wos_hip_propagate_device_noalloc)
against the resident-field Propagate it mirrors (wos_hip_field_propagate):
two calls in a row, the second on a smaller field, so the second must not
see what the first left in the shared scratch; oversize and unsupported
calls are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def field_of(n, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(n) - n / 2) / n
    env = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / 0.05)
    return env * np.exp(2j * np.pi * rng.random((n, n)))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_matches_field_propagate(wos, precision):
    prec = wos.FP64 if precision == "fp64" else wos.FP32
    tol = 1e-12 if prec == wos.FP64 else 1e-5
    # (n, polarized, method, dx, wavelength, z): the second call is smaller
    calls = [(2048 if prec == wos.FP32 else 1024, True, "asm", 2e-6, 532e-9, 20e-3),
             (512, False, "asm_pad", 1.5e-6, 633e-9, 5e-3),
             (256, True, "asm", 1e-6, 450e-9, -2e-3)]
    for n, pol, method, dx, wl, z in calls:
        ex, ey = field_of(n, n), (field_of(n, n + 1) if pol else None)
        ref = wos.Field.from_host(ex, ey, dx=dx, polarized=pol, precision=prec)
        ref.propagate(z, method, wos.ctx_for(wl))
        want = ref.download()
        ref.close()
        f = wos.Field.from_host(ex, ey, dx=dx, polarized=pol, precision=prec)
        wos.propagate_device_noalloc(wos.load().wos_hip_field_device_ptr(f.h), n, 2 if pol else 1, dx, z, method,
                                     wos.ctx_for(wl), prec)
        got = f.download()
        f.close()
        comps = 2 if pol else 1
        e = max(rel_l2(got[c], want[c]) for c in range(comps))
        same = all(np.array_equal(got[c], want[c]) for c in range(comps))
        print(f"noalloc {precision} n={n} {method} comps={comps}: rel-L2 vs field propagate {e:.2e}, "
              f"bit-equal {same}", flush=True)
        assert e <= tol
        assert same


def test_noalloc_refuses_what_it_cannot_hold(wos):
    f = wos.Field(64, 1e-6, precision=wos.FP64)
    p = wos.load().wos_hip_field_device_ptr(f.h)
    ctx = wos.ctx_for(532e-9)
    for args, msg in [((p, 4096, 1, 1e-6, 1e-3, "asm", ctx, wos.FP64), "scratch"),      # 4096^2 FP64: too large
                      ((p, 2048, 2, 1e-6, 1e-3, "asm", ctx, wos.FP64), "scratch"),      # two 2048^2 FP64 components
                      ((p, 100, 1, 1e-6, 1e-3, "asm", ctx, wos.FP64), "power of two"),
                      ((p, 64, 1, 1e-6, 1e-3, "fresnel_tf", ctx, wos.FP64), "asm"),
                      ((p, 64, 3, 1e-6, 1e-3, "asm", ctx, wos.FP64), "bad arguments")]:
        with pytest.raises(wos.WosError, match=msg):
            wos.propagate_device_noalloc(*args)
    wos.propagate_device_noalloc(p, 64, 1, 1e-6, 0.0, "asm", ctx, wos.FP64)   # z = 0: nothing to do
    f.close()
