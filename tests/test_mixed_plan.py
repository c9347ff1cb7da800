"""CPU check of the mixed-radix Stockham design (wos_mixed.hpp) for the grid
sides the kernels build: the plan (radix-16 stages, one leftover power of
two, then radix 3 and 5), the per-stage butterfly maps (loads x[j + r N/RAD],
stores y[(j - k) RAD + k + r NS], k = j mod NS, each output written once),
the twiddle table layout ([k][r-1] per stage, conjugated for the inverse)
and the radix-3/5 butterfly formulas, emulated in numpy and compared with
numpy's FFT."""
import numpy as np
import pytest

SIZES = [100, 200, 300, 384, 400, 500, 600, 640, 720, 768, 800, 1000, 1200, 1280, 1500, 1536, 1920, 2000, 2160, 2400,
         2560, 3000, 3072, 3840, 4000]


def plan(n):
    a = b = c = 0
    m = n
    while m % 2 == 0:
        m //= 2
        a += 1
    while m % 3 == 0:
        m //= 3
        b += 1
    while m % 5 == 0:
        m //= 5
        c += 1
    d = 0
    while m % 7 == 0:
        m //= 7
        d += 1
    assert m == 1
    return [16] * (a // 4) + ([1 << (a % 4)] if a % 4 else []) + [3] * b + [5] * c + [7] * d


def twiddles(n):
    rad, tab, ns = plan(n), [], None
    ns = rad[0]
    for s in range(1, len(rad)):
        L = ns * rad[s]
        for k in range(ns):
            for r in range(1, rad[s]):
                tab.append(np.exp(-2j * np.pi * ((r * k) % L) / L))
        ns = L
    return np.array(tab)


def mul_i(d, a):
    return -1j * a if d < 0 else 1j * a


def butterfly(R, d, v):
    if R == 3:
        s3 = np.sin(2 * np.pi / 3)
        a, b = v[1] + v[2], v[1] - v[2]
        t, u = v[0] - 0.5 * a, mul_i(d, s3 * b)
        return [v[0] + a, t + u, t - u]
    if R == 5:
        c1, c2 = np.cos(2 * np.pi / 5), np.cos(4 * np.pi / 5)
        s1, s2 = np.sin(2 * np.pi / 5), np.sin(4 * np.pi / 5)
        a1, b1, a2, b2 = v[1] + v[4], v[1] - v[4], v[2] + v[3], v[2] - v[3]
        t1, t2 = v[0] + c1 * a1 + c2 * a2, v[0] + c2 * a1 + c1 * a2
        u1, u2 = mul_i(d, s1 * b1 + s2 * b2), mul_i(d, s2 * b1 - s1 * b2)
        return [v[0] + a1 + a2, t1 + u1, t2 + u2, t2 - u2, t1 - u1]
    if R == 7:  # pairs a_k = v_k + v_(7-k), b_k = v_k - v_(7-k), as dft7 in the kernel
        a = [v[k] + v[7 - k] for k in (1, 2, 3)]
        b = [v[k] - v[7 - k] for k in (1, 2, 3)]
        out = [v[0] + sum(a)] + [None] * 6
        for m in (1, 2, 3):
            t = v[0] + sum(np.cos(2 * np.pi * m * k / 7) * a[k - 1] for k in (1, 2, 3))
            u = mul_i(d, sum(np.sin(2 * np.pi * m * k / 7) * b[k - 1] for k in (1, 2, 3)))
            out[m], out[7 - m] = t + u, t - u
        return out
    k = np.arange(R)
    return list(np.exp(d * 2j * np.pi * np.outer(k, k) / R) @ np.array(v))


def transform(x, d):
    n = len(x)
    rad, tw = plan(n), twiddles(n)
    data, ns, off = x.astype(complex).copy(), 1, 0
    for s, R in enumerate(rad):
        nb = n // R
        out = np.empty_like(data)
        written = np.zeros(n, bool)
        for j in range(nb):
            k = j % ns
            v = [data[j + r * nb] for r in range(R)]
            if ns > 1:
                for r in range(1, R):
                    w = tw[off + k * (R - 1) + r - 1]
                    v[r] = v[r] * (w if d < 0 else np.conj(w))
            v = butterfly(R, d, v)
            o = (j - k) * R + k
            for r in range(R):
                assert not written[o + r * ns]
                written[o + r * ns] = True
                out[o + r * ns] = v[r]
        assert written.all()
        data = out
        if s >= 1:
            off += ns * (R - 1)
        ns *= R
    assert off == len(tw)
    return data


@pytest.mark.parametrize("n", SIZES)
def test_mixed_radix_plan_matches_numpy(n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    want = np.fft.fft(x)
    assert np.linalg.norm(transform(x, -1) - want) <= 1e-13 * np.linalg.norm(want)
    want = np.fft.ifft(x) * n
    assert np.linalg.norm(transform(x, +1) - want) <= 1e-13 * np.linalg.norm(want)


# Sides without a compiled plan, and sides with factors of 7, run the
# runtime-length kernels (k_rows_mixed_rt / k_cols_mixed_rt): same plan,
# radix-7 stages last, same butterfly maps and tables.
RT_SIZES = [6, 90, 450, 480, 900, 960, 1080, 1152, 1800, 2700, 3240, 3600, 3888, 4050,
            7, 49, 700, 1400, 2100, 2940, 3430]


def smooth_sides(limit):
    out = []
    for n in range(2, limit + 1):
        m = n
        for f in (2, 3, 5, 7):
            while m % f == 0:
                m //= f
        if m == 1 and n & (n - 1):
            out.append(n)
    return out


def bmax(r):
    return 1 if r >= 16 else 4 if r == 2 else 3 if r == 3 else 2


def rt_threads(n):
    """Thread count mixed_rt_plan picks: the smallest multiple of 64 with
    ceil((n / RAD) / T) <= bmax(RAD) for every stage."""
    t = 64
    for r in plan(n):
        t = max(t, -(-(n // r) // bmax(r)))
    return -(-t // 64) * 64


@pytest.mark.parametrize("n", smooth_sides(4096))
def test_runtime_plan_gives_every_butterfly_one_thread(n):
    t = rt_threads(n)
    assert t <= 512 and len(plan(n)) <= 21
    assert (n + n // 16 + 1) * 16 <= 160 * 1024  # FP64 transform in LDS
    for r in plan(n):
        nb = n // r
        hits = np.zeros(nb, int)
        for b in range(bmax(r)):
            j = np.arange(t) + b * t
            hits[j[j < nb]] += 1
        assert (hits == 1).all()


@pytest.mark.parametrize("n", RT_SIZES)
def test_runtime_sides_match_numpy(n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    want = np.fft.fft(x)
    assert np.linalg.norm(transform(x, -1) - want) <= 1e-13 * np.linalg.norm(want)


# ------------------------------------------------- register-resident plan (wos_mixreg.hpp)

def _dft5(v, d):
    c1, c2, s1, s2 = np.cos(2 * np.pi / 5), np.cos(4 * np.pi / 5), np.sin(2 * np.pi / 5), np.sin(4 * np.pi / 5)
    a1, b1, a2, b2 = v[1] + v[4], v[1] - v[4], v[2] + v[3], v[2] - v[3]
    t1, t2 = v[0] + c1 * a1 + c2 * a2, v[0] + c2 * a1 + c1 * a2
    u1, u2 = d * 1j * (s1 * b1 + s2 * b2), d * 1j * (s2 * b1 - s1 * b2)
    return np.array([v[0] + a1 + a2, t1 + u1, t2 + u2, t2 - u2, t1 - u1])


def _dft10(v, d):
    """2 x 5: DFT-5 over n1 of x[2 n1 + n2], twiddle W10^(n2 k1), DFT-2; X[k1 + 5 k2]."""
    a, b = _dft5(v[0::2], d), _dft5(v[1::2], d) * np.exp(d * 2j * np.pi * np.arange(5) / 10)
    return np.concatenate([a + b, a - b])


def _dft_any(v, d):
    n = len(v)
    if n == 10:
        return _dft10(v, d)
    if n == 5:
        return _dft5(v, d)
    return np.array([np.sum(v * np.exp(d * 2j * np.pi * np.arange(n) * k / n)) for k in range(n)])


def _reg_tables(fwd):
    """Forward stages 1..S-1, then the inverse (reversed) plan's stages 1..S-1:
    entry [k][r-1] = exp(-2 pi i r k / (NS RAD)) (MRPlan<N>::tw_off)."""
    tab, offs = [], {}
    for dname, rads in (("f", fwd), ("i", fwd[::-1])):
        ns = rads[0]
        for s in range(1, len(rads)):
            offs[(dname, s)] = len(tab)
            L = ns * rads[s]
            tab += [np.exp(-2j * np.pi * ((r * k) % L) / L) for k in range(ns) for r in range(1, rads[s])]
            ns = L
    return np.array(tab), offs


def _reg_run(regs, n, fwd, v, d, tab, offs):
    """Run the stages on per-thread registers regs[t] (slot b RAD + r = input
    x[j + r n/RAD], j = t + b T), exchanging through an LDS array checked to
    be written exactly once per exchange."""
    T = n // v
    rads = fwd if d < 0 else fwd[::-1]
    ns = 1
    for s, rad in enumerate(rads):
        assert v % rad == 0
        for t in range(T):
            for b in range(v // rad):
                j = t + b * T
                k = j % ns
                blk = regs[t][b * rad:(b + 1) * rad].copy()
                if ns > 1:
                    w = tab[offs[("f" if d < 0 else "i", s)] + k * (rad - 1) + np.arange(rad - 1)]
                    blk[1:] *= w if d < 0 else np.conj(w)
                regs[t][b * rad:(b + 1) * rad] = _dft_any(blk, d)
        if s + 1 < len(rads):
            lds = np.full(n, np.nan + 0j)
            for t in range(T):
                for b in range(v // rad):
                    j = t + b * T
                    k = j % ns
                    for r in range(rad):
                        p = (j - k) * rad + k + r * ns
                        assert np.isnan(lds[p])
                        lds[p] = regs[t][b * rad + r]
            assert not np.isnan(lds).any()
            rn = rads[s + 1]
            for t in range(T):
                regs[t] = np.array([lds[t + b * T + r * (n // rn)] for b in range(v // rn) for r in range(rn)])
        ns *= rad
    return regs


@pytest.mark.parametrize("n,fwd,v", [(2000, [10, 10, 10, 2], 10),
                                    (1536, [8, 8, 8, 3], 24), (1000, [10, 10, 10], 10)])
def test_register_resident_plans(n, fwd, v):
    """The register-resident mixed-radix designs (MRDef in wos_mixreg.hpp):
    2000 points as [10, 10, 10, 2] on 200 threads of 10 values, 1536 as
    [8, 8, 8, 3] on 64 threads of 24, 1000 as [10, 10, 10] on 100 threads of
    10; the inverse plan reversed so the forward output map is the inverse
    input map (forward -> transfer function -> inverse stays in registers),
    the twiddle layout, and the mirror map of the transfer function (row
    N - p of thread t's row p = t + T m lies in thread T - t at
    m' = V - 1 - m, thread 0 at m' = V - m)."""
    T = n // v
    tab, offs = _reg_tables(fwd)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r0 = fwd[0]
    regs = [np.array([x[t + b * T + r * (n // r0)] for b in range(v // r0) for r in range(r0)]) for t in range(T)]
    regs = _reg_run(regs, n, fwd, v, -1, tab, offs)
    rl = fwd[-1]
    y = np.zeros(n, complex)
    for t in range(T):
        for b in range(v // rl):
            for r in range(rl):
                y[t + b * T + r * (n // rl)] = regs[t][b * rl + r]
    want = np.fft.fft(x)
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()
    # slot (m % NB) rl + m // NB of thread t holds row t + T m
    nb = v // rl
    for t in range(T):
        for m in range(v):
            assert t + (m % nb) * T + (m // nb) * (n // rl) == t + T * m
            pm = (n - (t + T * m)) % n
            assert pm == (T * ((v - m) % v) if t == 0 else (T - t) + T * (v - 1 - m))
    regs = _reg_run(regs, n, fwd, v, +1, tab, offs)   # fused: the inverse starts from the forward's registers
    r0i = fwd[::-1][-1]
    z = np.zeros(n, complex)
    for t in range(T):
        for b in range(v // r0i):
            for r in range(r0i):
                z[t + b * T + r * (n // r0i)] = regs[t][b * r0i + r]
    assert np.abs(z / n - x).max() <= 1e-13
