"""The drop-in fft1DAny / fft2D (optics/fft.go:135-153, 209-242) on lengths
beyond one workgroup's transform: powers of two above 8192 (four-step
factorisation over the engine's own kernels, up to 2^26) and other lengths
whose Bluestein length exceeds 8192 (the chirp-z chain on four-step
transforms).

Bars: FP64 rel-L2 <= 1e-12 against the oracle (which follows the
reference's fft1DBluestein and fft2DCPU) and against numpy for powers of two.
For Bluestein lengths numpy is held to a bound set by the reference's own
chirp: its phase pi k^2 / L is rounded in double, an error up to
ulp(pi L) / 2 per element that numpy's transform does not carry, so that
bound is 4 ulp(pi L). FP32 <= 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def chirp_bound(L):
    return max(1e-12, 4 * float(np.spacing(np.pi * L)))


def tone(N, f):
    """exp(2 pi i f k / N) with the phase reduced exactly in integers."""
    k = np.arange(N, dtype=np.int64)
    return np.exp(2j * np.pi * ((f * k) % N).astype(np.float64) / N)


@pytest.mark.parametrize("N", [16384, 32768, 1 << 20])
def test_long_power_of_two_1d(wos, oracle, N):
    rng = np.random.default_rng(N)
    a = rand_c(rng, N)
    g = wos.fft1d(a)
    e_np, e_or = rel_l2(g, np.fft.fft(a)), rel_l2(g, oracle.fft1d(a))
    gi = wos.fft1d(a, inverse=True)
    e_inv = rel_l2(gi, np.fft.ifft(a))
    g32 = wos.fft1d(a, precision=wos.FP32)
    e32 = rel_l2(g32, np.fft.fft(a))
    print(f"N={N}: fwd vs numpy {e_np:.2e} vs oracle {e_or:.2e}, inverse {e_inv:.2e}, FP32 {e32:.2e}", flush=True)
    assert e_np <= 1e-12 and e_or <= 1e-12 and e_inv <= 1e-12
    assert e32 <= 1e-5


@pytest.mark.parametrize("L", [4097, 5003, 6000, 9973, 100003])
def test_long_bluestein_1d(wos, oracle, L):
    rng = np.random.default_rng(L)
    a = rand_c(rng, L)
    for inv in (False, True):
        g = wos.fft1d(a, inverse=inv)
        e_or = rel_l2(g, oracle.fft1d(a, inverse=inv))
        e_np = rel_l2(g, np.fft.ifft(a) if inv else np.fft.fft(a))
        print(f"L={L} inverse={inv}: vs oracle {e_or:.2e}, vs numpy {e_np:.2e} "
              f"(bound {chirp_bound(L):.1e})", flush=True)
        assert e_or <= 1e-12
        assert e_np <= chirp_bound(L)
    g32 = wos.fft1d(a, precision=wos.FP32)
    assert rel_l2(g32, np.fft.fft(a)) <= 1e-5


def test_largest_power_of_two_1d(wos):
    """N = 2^26 (1 GiB of complex128; N1 = N2 = 8192): a tone lands on its
    bin, and the inverse brings the input back."""
    N, f = 1 << 26, 12345677
    x = tone(N, f)
    g = wos.fft1d(x)
    want = np.zeros(N, dtype=np.complex128)
    want[f] = N
    e = rel_l2(g, want)
    back = wos.fft1d(g, inverse=True)
    e_back = rel_l2(back, x)
    print(f"2^26 tone: {e:.2e}, round trip {e_back:.2e}", flush=True)
    assert e <= 1e-12 and e_back <= 1e-12


def test_largest_bluestein_length_1d(wos):
    """L = 2^25 - 1, Bluestein length 2^26: a tone lands on its bin (within
    the chirp bound) and the inverse brings it back."""
    L, f = (1 << 25) - 1, 777777
    x = tone(L, f)
    g = wos.fft1d(x)
    want = np.zeros(L, dtype=np.complex128)
    want[f] = L
    e = rel_l2(g, want)
    back = wos.fft1d(g, inverse=True)
    e_back = rel_l2(back, x)
    print(f"L=2^25-1 tone: {e:.2e}, round trip {e_back:.2e} (bound {chirp_bound(L):.1e})", flush=True)
    assert e <= chirp_bound(L) and e_back <= 2 * chirp_bound(L)


def test_long_bluestein_2d_against_the_oracle(wos, oracle):
    """4097^2 (Bluestein length 16384 on both passes) in both precisions."""
    n = 4097
    rng = np.random.default_rng(4097)
    a = rand_c(rng, (n, n))
    r = oracle.fft2d(a)
    g = wos.fft2d(a)
    e = rel_l2(g, r)
    ri = oracle.fft2d(a, inverse=True)
    ei = rel_l2(wos.fft2d(a, inverse=True), ri)
    e32 = rel_l2(wos.fft2d(a, precision=wos.FP32), r)
    print(f"4097^2: fwd {e:.2e} inverse {ei:.2e} FP32 {e32:.2e}", flush=True)
    assert e <= 1e-12 and ei <= 1e-12 and e32 <= 1e-5


def test_long_power_of_two_2d(wos):
    """16384^2 complex128 (4 GiB): a separable field u(y) v(x) transforms to
    U(fy) V(fx) exactly, and the inverse brings it back. A wrong index
    order, twiddle or transpose breaks the separable product."""
    n = 16384
    rng = np.random.default_rng(16384)
    u, v = rand_c(rng, n), rand_c(rng, n)
    a = np.multiply.outer(u, v)
    g = wos.fft2d(a)
    U, V = np.fft.fft(u), np.fft.fft(v)
    num = den = 0.0
    for r0 in range(0, n, 1024):
        w = np.multiply.outer(U[r0:r0 + 1024], V)
        num += float(np.sum(np.abs(g[r0:r0 + 1024] - w) ** 2))
        den += float(np.sum(np.abs(w) ** 2))
    e = float(np.sqrt(num / den))
    del a
    back = wos.fft2d(g, inverse=True)
    num = den = 0.0
    for r0 in range(0, n, 1024):
        w = np.multiply.outer(u[r0:r0 + 1024], v)
        num += float(np.sum(np.abs(back[r0:r0 + 1024] - w) ** 2))
        den += float(np.sum(np.abs(w) ** 2))
    e_back = float(np.sqrt(num / den))
    print(f"16384^2 separable: {e:.2e}, round trip {e_back:.2e}", flush=True)
    assert e <= 1e-12 and e_back <= 1e-12
