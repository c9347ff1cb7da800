"""Device-resident element chain (wos_field) against the oracle's
restatements: every element on its own, then whole trains with
propagations, a beamsplitter's cloned arm and the band limit, in FP64
(<= 1e-12) and FP32 (<= 1e-5)."""
import math

import numpy as np
import pytest

from test_oracle_elements import jones_polarizer, jones_retarder, jones_rotator

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def rel_l2(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def field(n, seed, polarized=False):
    rng = np.random.default_rng(seed)
    x = (np.arange(n) - n / 2) / n
    env = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / 0.1)
    ex = env * np.exp(2j * np.pi * rng.random((n, n)))
    ey = env * np.exp(2j * np.pi * rng.random((n, n))) * 0.5 if polarized else None
    return ex, ey


def compare(wos, oracle, n, dx, polarized, op, seed=1, tol64=FP64_TOL):
    """Apply op(f) to a device field (FP64 and FP32) and to the oracle field."""
    ex, ey = field(n, seed, polarized)
    o = oracle.OField(ex, ey, dx=dx)
    op(o, oracle)
    ro = np.concatenate([o.ex, o.ey]) if o.polarized else o.ex
    errs = {}
    for prec in (wos.FP64, wos.FP32):
        f = wos.Field.from_host(ex, ey, dx=dx, precision=prec)
        op(f, wos)
        gx, gy = f.download()
        assert f.polarized == o.polarized
        got = np.concatenate([gx, gy]) if f.polarized else gx
        if not f.polarized:
            assert np.array_equal(gy, np.zeros_like(gy)) or ey is not None
        errs[prec] = rel_l2(got, ro)
        f.close()
    assert errs[wos.FP64] <= tol64 and errs[wos.FP32] <= FP32_TOL, errs
    return errs


@pytest.mark.parametrize("polarized", [False, True])
def test_lens(wos, oracle, polarized):
    e = compare(wos, oracle, 128, 2e-6, polarized, lambda f, m: f.lens(633e-9, 0.03, aperture=100e-6, x0=6e-6,
                                                                        y0=-4e-6))
    print(f"lens polarized={polarized}: FP64 {e[wos.FP64]:.2e} FP32 {e[wos.FP32]:.2e}", flush=True)


SHAPES = [
    ("circle", dict(radius=70e-6), None),
    ("slit", dict(width=30.5e-6, height=150e-6, rotation=0.3), None),
    ("square", dict(width=90.5e-6, rotation=-0.2), None),
    ("rectangle", dict(width=120.3e-6, height=50.7e-6), None),
    ("ellipse", dict(a=80e-6, b=40e-6, rotation=0.7), None),
    ("superellipse", dict(a=80e-6, b=60e-6, order=4.5), None),
    ("ring", dict(rin=30e-6, rout=90e-6), None),
    ("double_slit", dict(width=10.5e-6, height=100e-6, separation=40e-6), None),
    ("cross", dict(width=20.5e-6, length=150e-6, rotation=0.4), None),
    ("polygon", dict(rotation=0.1), ("polygon", 6, 90e-6)),
    ("polygon", dict(), ("star", 5, 100e-6)),
    ("circle", dict(radius=70e-6, edge_sigma=5e-6), None),
]


@pytest.mark.parametrize("idx", range(len(SHAPES)))
def test_aperture_outlines(wos, oracle, idx):
    kind, params, verts = SHAPES[idx]
    v = None
    if verts is not None:
        v = (oracle.regular_polygon(verts[1], verts[2], math.pi / verts[1]) if verts[0] == "polygon"
             else oracle.star(verts[1], verts[2], verts[2] / 2.5, math.pi / 2))

    e = compare(wos, oracle, 128, 2e-6, idx % 2 == 1,
                lambda f, m: f.aperture(m.make_shape(kind, verts=v, **params), x0=5e-6, y0=-3e-6))
    print(f"aperture {kind} {params}: FP64 {e[wos.FP64]:.2e} FP32 {e[wos.FP32]:.2e}", flush=True)
    assert e[wos.FP64] <= 1e-14  # the masks agree pixel for pixel


@pytest.mark.parametrize("n,corr", [(128, 0.0), (128, 6e-6), (100, 6e-6)])
def test_diffuser(wos, oracle, n, corr):
    e = compare(wos, oracle, n, 2e-6, n == 128, lambda f, m: f.diffuser(sigma=1.3, corr=corr, amp=0.8, seed=11,
                                                                        x0=8e-6, y0=2e-6))
    print(f"diffuser n={n} corr={corr}: FP64 {e[wos.FP64]:.2e} FP32 {e[wos.FP32]:.2e}", flush=True)


@pytest.mark.parametrize("polarized", [False, True])
@pytest.mark.parametrize("which", ["polarizer", "retarder", "rotator", "scalar", "custom"])
def test_jones(wos, oracle, polarized, which):
    m = {"polarizer": jones_polarizer(0.6, 0.9), "retarder": jones_retarder(math.pi / 2, math.pi / 4),
         "rotator": jones_rotator(0.35), "scalar": (0.5 - 0.5j, 0, 0, 0.5 - 0.5j),
         "custom": (1 + 2j, 0.5, -0.25j, 3)}[which]
    e = compare(wos, oracle, 64, 1e-6, polarized, lambda f, mod: f.jones(*m))
    assert e[wos.FP64] <= 1e-15


@pytest.mark.parametrize("n", [128, 100])
def test_scale_and_bandlimit(wos, oracle, n):
    compare(wos, oracle, n, 1e-6, True, lambda f, m: f.scale(0.3 + 0.4j))
    for pol in (False, True):
        e = compare(wos, oracle, n, 1e-6, pol, lambda f, m: f.bandlimit(0.7, 0.05))
        print(f"bandlimit n={n} polarized={pol}: FP64 {e[wos.FP64]:.2e} FP32 {e[wos.FP32]:.2e}", flush=True)
    compare(wos, oracle, n, 1e-6, False, lambda f, m: f.bandlimit(1.0, 0.05))


def run_train(f, m, ctx, bl):
    """A train with the element kinds of runTrain: lens, propagation (each
    followed by the band limit), aperture, diffuser, a polariser that makes
    the field a vector field, a retarder, and propagations after it."""
    f.lens(633e-9, 0.05, aperture=220e-6)
    f.propagate(0.01, "asm", ctx)
    f.bandlimit(*bl)
    f.aperture(m.make_shape("circle", radius=150e-6, edge_sigma=4e-6))
    f.diffuser(sigma=0.8, corr=6e-6, amp=1.0, seed=3)
    f.propagate(0.004, "asm_pad", ctx)
    f.bandlimit(*bl)
    f.jones(*jones_polarizer(0.5))
    f.jones(*jones_retarder(math.pi / 2, 0.2))
    f.propagate(0.006, "asm", ctx)
    f.bandlimit(*bl)


@pytest.mark.parametrize("n", [256, 200])
def test_train_matches_oracle(wos, oracle, n):
    dx = 2e-6
    ex, _ = field(n, 7)
    bl = (0.9, 0.05)
    o = oracle.OField(ex, dx=dx)
    run_train(o, oracle, oracle.ctx_for(633e-9), bl)
    # the beamsplitter's reflected arm: a clone scaled by i sqrt(r) exp(i phase)
    arm_o = o.clone()
    arm_o.scale(1j * math.sqrt(0.3) * complex(math.cos(0.2), math.sin(0.2)))
    o.scale(math.sqrt(0.7))
    arm_o.propagate(0.003, "asm", oracle.ctx_for(633e-9))
    ref = np.concatenate([o.ex, o.ey, arm_o.ex, arm_o.ey])
    so = oracle.compute_stats(o.ex, o.ey, dx, 633e-9)
    for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
        f = wos.Field.from_host(ex, dx=dx, precision=prec)
        run_train(f, wos, wos.ctx_for(633e-9), bl)
        arm = f.clone()
        arm.scale(1j * math.sqrt(0.3) * complex(math.cos(0.2), math.sin(0.2)))
        f.scale(math.sqrt(0.7))
        arm.propagate(0.003, "asm", wos.ctx_for(633e-9))
        assert f.polarized and arm.polarized
        got = np.concatenate([*f.download(), *arm.download()])
        e = rel_l2(got, ref)
        sg = f.stats(633e-9)
        print(f"train n={n} prec={prec}: rel-L2 {e:.2e}, power {sg.power:.9e} vs {so.power:.9e}", flush=True)
        assert e <= tol
        assert abs(sg.power - so.power) <= (1e-12 if prec == wos.FP64 else 1e-5) * so.power
        f.close()
        arm.close()


@pytest.mark.parametrize("n,polarized", [(256, False), (256, True), (200, True)])
def test_propagate_bandlimited_matches_oracle(wos, oracle, n, polarized):
    """A propagate element with the context's band limit (Propagate, then
    ApplyBandlimit): one fused transfer per component on power-of-two grids,
    the two steps on others, against the oracle's two steps. The warnings
    and stats are the propagation's: dx = 0.25 um keeps evanescent bands on
    the grid, so evanescent_filtered fires, its value and p_out taken from
    the power before the band limit (the fused pass sums |H U|^2 over the
    spectrum)."""
    dx, wl, z, bl = 0.25e-6, 633e-9, 20e-6, (0.8, 0.05)
    ex, ey = field(n, 31, polarized)
    o = oracle.OField(ex, ey, dx=dx)
    wo, so = o.propagate(z, "asm", oracle.ctx_for(wl))
    o.bandlimit(*bl)
    ref = np.concatenate([o.ex, o.ey]) if polarized else o.ex
    assert "evanescent_filtered" in wo.as_dict()
    for prec, tol, ptol in ((wos.FP64, FP64_TOL, 1e-12), (wos.FP32, FP32_TOL, 1e-5)):
        f = wos.Field.from_host(ex, ey, dx=dx, precision=prec)
        wg, sg = f.propagate(z, "asm", wos.ctx_for(wl), bandlimit=bl)
        got = np.concatenate(f.download()) if polarized else f.download()[0]
        e = rel_l2(got, ref)
        print(f"propagate+bandlimit n={n} polarized={polarized} prec={prec}: rel-L2 {e:.2e}, "
              f"p_out {sg.p_out:.12e} vs {so.p_out:.12e}", flush=True)
        assert e <= tol
        assert abs(sg.p_in - so.p_in) <= ptol * so.p_in and abs(sg.p_out - so.p_out) <= ptol * so.p_out
        assert wg.as_dict().keys() == wo.as_dict().keys()
        for code, (cnt, val) in wo.as_dict().items():
            assert wg.as_dict()[code][0] == cnt
            if prec == wos.FP64:
                assert abs(wg.as_dict()[code][1] - val) <= 1e-9 * max(abs(val), 1e-300), code
        f.close()


@pytest.mark.parametrize("polarized", [False, True])
def test_fraunhofer_matches_oracle(wos, oracle, polarized):
    """Propagate with method fraunhofer: the drop-in call and the resident
    field against the oracle (far field at 0.5 m, then a near-field case that
    raises fraunhofer_nearfield); the field's grid spacing becomes wl z/(n dx)
    and a following asm propagation runs on it."""
    n, dx, wl = 256, 2e-6, 633e-9
    ex, ey = field(n, 21, polarized)
    for z in (0.5, 2e-3):
        rx, ry, wo, _ = oracle.propagate(ex, ey, dx, z, "fraunhofer", oracle.ctx_for(wl))
        ref = np.concatenate([rx, ry]) if polarized else rx
        for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
            gx, gy, wg, _ = wos.propagate(ex, ey, dx, z, "fraunhofer", wos.ctx_for(wl), precision=prec)
            got = np.concatenate([gx, gy]) if polarized else gx
            e = rel_l2(got, ref)
            print(f"fraunhofer z={z} polarized={polarized} prec={prec}: rel-L2 {e:.2e}, warnings {wg.as_dict()}",
                  flush=True)
            assert e <= tol
            assert wg.as_dict().keys() == wo.as_dict().keys()
            for key, (c, v) in wo.as_dict().items():
                assert wg.as_dict()[key][0] == c and abs(wg.as_dict()[key][1] - v) <= 1e-9 * abs(v)
    o = oracle.OField(ex, ey, dx=dx)
    o.propagate(0.5, "fraunhofer", oracle.ctx_for(wl))
    o.propagate(0.02, "asm", oracle.ctx_for(wl))
    f = wos.Field.from_host(ex, ey, dx=dx)
    f.propagate(0.5, "fraunhofer", wos.ctx_for(wl))
    assert abs(f.grid_dx - o.dx) <= 1e-18
    f.propagate(0.02, "asm", wos.ctx_for(wl))
    gx, gy = f.download()
    got = np.concatenate([gx, gy]) if polarized else gx
    ref = np.concatenate([o.ex, o.ey]) if polarized else o.ex
    assert rel_l2(got, ref) <= FP64_TOL
    f.close()
    with pytest.raises(wos.WosError):
        wos.propagate(np.ones((9, 9)), None, dx, 0.5, "fraunhofer", wos.ctx_for(wl))


@pytest.mark.parametrize("n,dx,wl,z", [(128, 0.2e-6, 0.5e-6, 4e-6), (94, 1e-6, 633e-9, -20e-6), (256, 2e-6, 633e-9, 0.02)])
def test_vectorial_matches_oracle(wos, oracle, n, dx, wl, z):
    """PropagateVectorial (Ex, Ey in place, Ez out) against the oracle, with
    evanescent bins (dx < wl/2), a backward step and a Bluestein grid
    (94 = 2 47)."""
    ex, ey = field(n, 31, True)
    ctx_o, ctx_g = oracle.ctx_for(wl), wos.ctx_for(wl)
    ref = np.concatenate(oracle.propagate_vectorial(ex, ey, dx, z, ctx_o))
    for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
        got = np.concatenate(wos.propagate_vectorial(ex, ey, dx, z, ctx_g, precision=prec))
        e = rel_l2(got, ref)
        print(f"vectorial n={n} z={z} prec={prec}: rel-L2 {e:.2e}", flush=True)
        assert e <= tol


@pytest.mark.parametrize("polarized", [False, True])
@pytest.mark.parametrize("n,dx,z", [(128, 0.3e-6, 12e-6), (94, 1e-6, -40e-6)])
def test_uniaxial_matches_oracle(wos, oracle, polarized, n, dx, z):
    """PropagateUniaxial (calcite-like no 1.658, ne 1.486) as a drop-in call
    and as a resident-field element, against the oracle; evanescent bins at
    dx 0.3 um, a backward step and a Bluestein grid (94 = 2 47)."""
    wl, no, ne = 633e-9, 1.658, 1.486
    ex, ey = field(n, 41, polarized)
    rx, ry = oracle.propagate_uniaxial(ex, ey, dx, z, no, ne, wl)
    ref = np.concatenate([rx, ry]) if polarized else rx
    for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
        gx, gy = wos.propagate_uniaxial(ex, ey, dx, z, no, ne, wos.ctx_for(wl), precision=prec)
        got = np.concatenate([gx, gy]) if polarized else gx
        e = rel_l2(got, ref)
        print(f"uniaxial n={n} z={z} polarized={polarized} prec={prec}: rel-L2 {e:.2e}", flush=True)
        assert e <= tol
    e = compare(wos, oracle, n, dx, polarized, lambda f, m: f.uniaxial(z, no, ne, wl), seed=41)
    with pytest.raises(wos.WosError):
        wos.propagate_uniaxial(ex, ey, dx, z, 0.0, ne, wos.ctx_for(wl))


@pytest.mark.parametrize("polarized", [False, True])
def test_plane_views_match_oracle(wos, oracle, polarized):
    """Every float32 plane view of a resident field against the oracle on
    the same field values: NaN masks identical (bit pattern 0x7FC00000),
    values within one float32 rounding."""
    n = 96
    ex, ey = field(n, 51, polarized)
    for prec in (wos.FP64, wos.FP32):
        f = wos.Field.from_host(ex, ey, dx=1e-6, precision=prec)
        hx, hy = f.download()  # the values the device holds
        for view in wos.VIEWS:
            for mc in (0.0, 0.05):
                got = f.view(view, mc)
                ref = oracle.plane_view(hx, hy if polarized else None, view, mc)
                nan = np.isnan(ref)
                assert np.array_equal(np.isnan(got), nan), (view, mc, prec)
                assert np.all(got[nan].view(np.uint32) == 0x7FC00000)
                d = np.abs(got[~nan].astype(np.float64) - ref[~nan].astype(np.float64))
                scale = np.maximum(np.abs(ref[~nan]).astype(np.float64), 1e-30)
                assert np.all(d <= 2.5e-7 * scale + 1e-12), (view, mc, prec, float(np.max(d / scale)))
        f.close()


def rotated_biaxial(nx, ny, nz, angle, loss=0.0):
    """diag(nx^2, ny^2, nz^2) rotated about z by `angle`, with absorption."""
    c, s = math.cos(angle), math.sin(angle)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    return R @ np.diag([nx ** 2 + 1j * loss, ny ** 2, nz ** 2]) @ R.T


@pytest.mark.parametrize("polarized", [False, True])
@pytest.mark.parametrize("n,eps", [(64, "biaxial"), (64, "rotated"), (46, "rotated")])
def test_anisotropic_matches_oracle(wos, oracle, polarized, n, eps):
    """PropagateAnisotropic (Berreman 4x4 per bin) as a drop-in call and as a
    resident-field element against the oracle: the biaxial element's diagonal
    tensor, and a rotated absorbing tensor with off-diagonal terms; a
    Bluestein grid (n = 46 = 2 23)."""
    dx, wl, z = 0.5e-6, 633e-9, 8e-6
    e = np.diag([1.6 ** 2, 1.5 ** 2, 1.4 ** 2]).astype(complex) if eps == "biaxial" else \
        rotated_biaxial(1.62, 1.5, 1.45, 0.4, 2e-3)
    ex, ey = field(n, 61, polarized)
    rx, ry = oracle.propagate_anisotropic(ex, ey, dx, z, e, wl)
    ref = np.concatenate([rx, ry])
    for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
        gx, gy = wos.propagate_anisotropic(ex, ey, dx, z, e, wos.ctx_for(wl), precision=prec)
        err = rel_l2(np.concatenate([gx, gy]), ref)
        print(f"anisotropic n={n} {eps} polarized={polarized} prec={prec}: rel-L2 {err:.2e}", flush=True)
        assert err <= tol
    # the element on a resident field: Ey takes part, Polarized stays as it was
    o = oracle.OField(ex, ey, dx=dx)
    o.anisotropic(z, e, wl)
    f = wos.Field.from_host(ex, ey, dx=dx)
    f.anisotropic(z, e, wl)
    assert f.polarized == polarized
    gx, gy = f.download()
    assert rel_l2(np.concatenate([gx, gy]), np.concatenate([o.ex, o.ey])) <= FP64_TOL
    f.close()


ELEMENTS = [
    ("cylindrical_lens", dict(f=0.03, aperture=100e-6, x=5e-6, y=-3e-6, rotation=0.3)),
    ("apodizer", dict(waist=80e-6, amplitude=0.8, x=-6e-6)),
    ("grating", dict(kind="amplitude_sin", period=37.3e-6, modulation=0.7, rotation=0.4, offset=0.2)),
    ("grating", dict(kind="amplitude_binary", period=37.3e-6, duty=0.37, rotation=0.4, offset=0.2)),
    ("grating", dict(kind="phase_sin", period=37.3e-6, modulation=1.3, rotation=0.4)),
    ("grating", dict(kind="phase_binary", period=41.1e-6, modulation=2.2, offset=0.1)),
    ("grating", dict(kind="blazed", period=29.9e-6, blaze_depth=2.1, rotation=-0.7)),
    ("axicon", dict(alpha=0.01, index=1.45)),
    ("spiral_phase", dict(charge=3, f=0.05, x=3e-6)),
    ("wedge", dict(alpha=0.005, rotation=1.1)),
    ("zone_plate", dict(f=2.13e-3, radius=110e-6, kind="phase")),
    ("zone_plate", dict(f=2.13e-3, radius=110e-6, kind="amplitude")),
    ("mirror", dict(reflectivity=0.9, curvature=3.3, tilt_x=2e-4, tilt_y=-1e-4)),
    ("concave_mirror", dict(radius=0.07, aperture=110e-6, x=4e-6)),
    ("convex_mirror", dict(radius=0.09, reflectivity=0.8)),
    ("zernike", dict(radius=100e-6, c4=0.3, c7=-0.2, c12=0.1, c21=0.05)),
    ("kerr", dict(kind="tpa", n2=1e-7, tpa=300.0, length=2e-3)),
    ("kerr", dict(n2=-2e-7, length=1e-3)),
    ("saturable_absorber", dict(alpha0=500.0, isat=0.3, length=2e-3)),
]


@pytest.mark.parametrize("polarized", [False, True])
@pytest.mark.parametrize("name,params", ELEMENTS, ids=[f"{n}-{p.get('kind', '')}" for n, p in ELEMENTS])
def test_further_elements_match_oracle(wos, oracle, name, params, polarized):
    """The further train elements (wos_hip_field_element) on a resident field
    against the oracle's restatements: FP64 and FP32 (the per-pixel
    transmission is formed in FP64 in both, so threshold elements open the
    same pixels)."""
    e = compare(wos, oracle, 128, 2e-6, polarized, lambda f, m: f.element(m.make_element(name, **params), 633e-9))
    print(f"{name} {params.get('kind', '')} polarized={polarized}: FP64 {e[wos.FP64]:.2e} FP32 {e[wos.FP32]:.2e}",
          flush=True)


def test_further_elements_reject_bad_parameters(wos):
    f = wos.Field.from_host(np.ones((16, 16), dtype=complex), dx=1e-6)
    for name, p in (("apodizer", dict(waist=0)), ("grating", dict(period=-1)), ("zone_plate", dict(f=0)),
                    ("mirror", dict(reflectivity=2)), ("zernike", dict(radius=0)), ("cylindrical_lens", dict(f=0)),
                    ("kerr", dict(length=-1)), ("saturable_absorber", dict(isat=0))):
        with pytest.raises(RuntimeError):
            f.element(wos.make_element(name, **p), 633e-9)
    f.close()


@pytest.mark.parametrize("n,polarized", [(128, True), (100, False)])
def test_medium_element_matches_oracle(wos, oracle, n, polarized):
    """The medium element (PropagateSplitStep through a uniform index n + i k)
    on a resident field against the oracle's split-step BPM: the merged device
    pipeline on a power-of-two grid, the half steps one by one on 100."""
    dx, wl, dist, steps = 1e-6, 633e-9, 60e-6, 6
    ex, ey = field(n, 41, polarized)
    ref_x, ref_y, wo = oracle.split_step(ex, ey, dx, dist, steps, np.full((n, n), 1.45 + 2e-4j), oracle.ctx_for(wl))
    ref = np.concatenate([ref_x, ref_y]) if polarized else ref_x
    for prec, tol in ((wos.FP64, FP64_TOL), (wos.FP32, FP32_TOL)):
        f = wos.Field.from_host(ex, ey, dx=dx, precision=prec)
        wg = f.medium(dist, 1.45, 2e-4, steps, wos.ctx_for(wl))
        got = np.concatenate(f.download()) if polarized else f.download()[0]
        e = rel_l2(got, ref)
        print(f"medium n={n} polarized={polarized} prec={prec}: rel-L2 {e:.2e}", flush=True)
        assert e <= tol
        if prec == wos.FP64:
            assert wg.as_dict().keys() == wo.as_dict().keys()
        f.close()

