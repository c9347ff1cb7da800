"""Headline benchmark: 2048^2 ASM propagations/s on one MI355X, FP32 vs FP64.

Workload (the polychromatic set of the batched configuration): 64 fields of
2048 x 2048, dx = 2 um, each a seeded Gaussian-Schell realisation (w0 = 300
um, sigma_c = 50 um) generated on the device, each propagated by the angular
spectrum method over z = 20 mm at its own wavelength (64 wavelengths evenly
spaced over 400-700 nm). One step is one complete propagation of all 64
fields (forward 2-D FFT -> transfer function -> inverse 2-D FFT), device
resident, replayed from a captured hipGraph. `value` is the FP32 bulk path;
the FP64 path runs the same batch beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tetsuhiro-wos_amd"))

N = 2048
DX = 2e-6
Z = 20e-3
UNITS = 64
W0, SIGMA_C, SEED = 300e-6, 50e-6, 20240607
HBM_PEAK_GBS = 8000.0
# MI355X VALU issue ceiling: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction
VALU_WAVE_INSTR_PER_S = 614.4e9
# HBM bytes per full-batch launch (64 x 2048^2), rocprofv3 PMC: 2 x FETCH_SIZE + WRITE_SIZE
# (calibrated on a 1 GiB copy), from profiles/asm_2048x64_hbm_traffic.md, measured at the launch shapes
# PMC_SHAPE (tile_w, col_w, rows, row_groups, split_lds, radix, col_radix, col_split_lds, inv_rows,
# inv_row_groups, inv_split_lds); a run whose autotuned shape differs reports no traffic figure
PMC_HBM_BYTES = {
    "fp32": {"row_fft_fwd": (1048872 * 2 + 2097201) * 1024, "col_fft_tf_ifft": (1048974 * 2 + 2097747) * 1024,
             "row_fft_inv": (1048772 * 2 + 2097158) * 1024},
    "fp64": {"row_fft_fwd": (2097380 * 2 + 4195109) * 1024, "col_fft_tf_ifft": (2100070 * 2 + 4275576) * 1024,
             "row_fft_inv": (2097794 * 2 + 4195266) * 1024},
}
# VALU wave-instructions per full-batch launch (SQ_INSTS_VALU), from profiles/asm_2048x64_sq_counters.md
PMC_VALU = {"fp32": {"row_fft_fwd": 1608 * 131072, "col_fft_tf_ifft": 2831 * 131072, "row_fft_inv": 1783 * 131072}}
PMC_SHAPE = {"fp32": (8, 8, 2, 2, 0, 16, 32, 1, 2, 2, 1), "fp64": (4, 4, 2, 0, 1, 16, 16, 1, 2, 2, 1)}
METRIC = "2048² ASM props/s: FP32 vs FP64 at rel-L2≤1e-5; FFT HBM GB/s % of peak"


def rel_l2(a, b):
    den = np.sum(np.abs(b) ** 2)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / den))


def tune_batch(b, tune, retune, iters=5):
    """The launch shape of a batch: the committed tuning table's entry unless
    retune (or no entry), else the autotuner. Returns (shape, ms, source)."""
    if not tune:
        return None, None, "default"
    if not retune:
        shape = b.load_tuning()
        if shape is not None:
            return shape, b.time(1, 5)[0], "table"
    shape, ms = b.autotune(iters)
    return shape, ms, "autotune"


def run_gpu(wos, precision, steps, warmup, shape=None, retune=False):
    wls = np.linspace(400e-9, 700e-9, UNITS)
    zs = np.full(UNITS, Z)
    b = wos.Batch(N, DX, wls, zs, wos.ctx_for(550e-9), "asm", precision)
    b.fill_schell(W0, SIGMA_C, SEED)
    inp0 = b.download(0)
    if shape:   # a fixed launch shape: tile_w, col_w, rows, row groups, split[, radix[, col_radix[, col_split
        #         [, inv_rows, inv_groups, inv_split]]]] (no inverse fields: as the forward row pass)
        shape = list(shape) + [16] * (6 - len(shape))
        shape = shape + [shape[5]] * (7 - len(shape))
        shape = shape + [0] * (8 - len(shape))
        if len(shape) < 11:
            shape = shape[:8] + shape[2:5]
        b.set_tuning(*shape)
        tune = (dict(zip(wos.TUNING_KEYS, shape)), b.time(1, 5)[0])
        source = "fixed"
    else:
        s_, ms_, source = tune_batch(b, True, retune)
        tune = (s_, ms_)
    b.fill_schell(W0, SIGMA_C, SEED)   # the tuning launches ran on the data
    b.capture()
    b.run(1)
    b.sync()
    out = {"unit0_in": inp0, "unit0": b.download(0), "unit_last": b.download(UNITS - 1),
           "wl0": wls[0], "tune": dict(tune[0], ms_per_run=round(tune[1], 4), source=source)}
    b.run(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.run(steps)
    b.sync()
    t1 = time.perf_counter()
    out["ms_per_step"] = (t1 - t0) * 1e3 / steps
    ms_ev, kms = b.time(2, max(steps, 10), kernels=True)
    out["ms_per_step_events"] = ms_ev
    out["kernel_ms"] = kms
    b.close()
    return out


def bench_persistent(wos, steps):
    """The headline batch with the persistent transfer (one launch per step,
    per-XCD work queues; the queued passes one after another, and
    interleaved task by task), captured and replayed like the headline run;
    units 0 and 63 compared with the three-launch run at the persistent
    kernel's shape (bit-identical by construction)."""
    out = {}
    wls = np.linspace(400e-9, 700e-9, UNITS)
    for prec, name in ((wos.FP32, "fp32"), (wos.FP64, "fp64")):
        b = wos.Batch(N, DX, wls, np.full(UNITS, Z), wos.ctx_for(550e-9), "asm", prec)
        r = 2 if prec == wos.FP64 else 4
        b.set_tuning(r, r, r, 0, 1, 16, 16, 1)
        b.fill_schell(W0, SIGMA_C, SEED)
        b.run(1)
        b.sync()
        three = [b.download(0), b.download(UNITS - 1)]
        res = {}
        for il in (False, True):
            b.fill_schell(W0, SIGMA_C, SEED)
            b.set_persistent(True, il)
            b.capture()
            b.run(1)
            b.sync()
            same = bool(np.array_equal(b.download(0), three[0]) and np.array_equal(b.download(UNITS - 1), three[1]))
            b.run(2)
            b.sync()
            t0 = time.perf_counter()
            b.run(steps)
            b.sync()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            res["interleaved" if il else "pass_after_pass"] = {
                "ms_per_step": round(ms, 4), "value": round(UNITS / (ms * 1e-3), 1),
                "equals_three_launch": same}
        b.close()
        out[name] = res
    out["unit"] = "props/s"
    return out


def pcie_inclusive(wos, inp, wl, precision, reps=5):
    """Drop-in rate: one 2048^2 Propagate (asm) per call from and to host
    complex128 memory (upload, device pipeline with power sums, download)."""
    import ctypes
    ctx = wos.ctx_for(wl)
    lib = wos.load()
    a = np.array(inp, dtype=np.complex128, order="C")   # the caller's slice, propagated in place
    w, st = wos.Warnings(), wos.PropStats()

    def call():
        rc = lib.wos_hip_propagate(a.ctypes.data, None, N, DX, Z, wos.METHODS["asm"], ctypes.byref(ctx), precision,
                                   ctypes.byref(w), ctypes.byref(st))
        if rc:
            raise RuntimeError(lib.wos_hip_last_error().decode())

    call()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    return reps / (time.perf_counter() - t0)


def fft2d_dropin_ms(wos, inp, reps=10):
    """Drop-in gpuFFT2D stand-in: one 2048^2 complex128 2-D FFT per call on a
    host buffer (upload, transform, download), alternating forward and
    inverse so the values stay bounded."""
    import ctypes
    lib = wos.load()
    a = np.array(inp, dtype=np.complex128, order="C")
    for d in (0, 1):
        if lib.wos_hip_fft2d_z(a.ctypes.data, N, d):
            raise RuntimeError(lib.wos_hip_last_error().decode())
    t0 = time.perf_counter()
    for i in range(reps):
        if lib.wos_hip_fft2d_z(a.ctypes.data, N, i & 1):
            raise RuntimeError(lib.wos_hip_last_error().decode())
    return (time.perf_counter() - t0) * 1e3 / reps


def bench_bpm(wos):
    """Split-step BPM: 1024^2, dx 0.5 um, 1.55 um, 400 slices of 5 um through
    the graded-index medium, Gaussian w0 10 um; device time per slice."""
    n, dx, wl, steps = 1024, 0.5e-6, 1.55e-6, 400
    x = (np.arange(n) - n / 2) * dx
    r = np.hypot(x[None, :], x[:, None])
    index = np.where(r < 100e-6, 1.45 - 0.01 * (r / 100e-6) ** 2, 1.44) + 1e-5j
    ex = np.exp(-(r / 10e-6) ** 2).astype(np.complex128)
    out = {}
    res = {}
    for name, prec in (("fp64", wos.FP64), ("fp32", wos.FP32)):
        # warm-up: tables, buffers and clocks settle before the timed run
        wos.split_step(ex, None, dx, 20 * 5e-6, 20, index, wos.ctx_for(wl), precision=prec)
        o, _, _ = wos.split_step(ex, None, dx, steps * 5e-6, steps, index, wos.ctx_for(wl), precision=prec)
        ms = wos.last_device_ms()
        res[name] = o
        out[name] = {"steps_per_s": round(steps / (ms * 1e-3), 1), "device_ms": round(ms, 3)}
    out["rel_l2_fp32_vs_fp64"] = rel_l2(res["fp32"], res["fp64"])
    out["power_ratio_fp64"] = float(np.sum(np.abs(res["fp64"]) ** 2) / np.sum(np.abs(ex) ** 2))
    # a Strang step is two ASM half steps: 2 x 8 n^2 S algorithmic bytes
    for name, S in (("fp64", 16), ("fp32", 8)):
        out[name]["frac_algorithmic"] = round(out[name]["steps_per_s"] * 2 * 8 * n * n * S / (HBM_PEAK_GBS * 1e9), 4)
    out["frac_note"] = ("algorithmic bytes (two ASM half steps per slice) over time against 8 TB/s; the merged "
                        "half steps move a quarter of them and the 8/16 MiB field stays in the last-level cache, "
                        "so this is not an HBM measurement and can exceed 1")
    check20 = {}
    for name, prec in (("fp64", wos.FP64), ("fp32", wos.FP32)):
        check20[name], _, _ = wos.split_step(ex, None, dx, 20 * 5e-6, 20, index, wos.ctx_for(wl), precision=prec)
    return out, check20


def config4_pipeline(wos, prec):
    """The batched configuration in one pipeline: 64 x 2048^2 polychromatic
    intensity, 128 z-planes at 1024^2, 32 mixed fields (384^2, 1000^2,
    1536^2); Gaussian-Schell inputs."""
    p = wos.Pipeline(prec)
    ctx = wos.ctx_for(550e-9)
    s = p.add(2048, 2e-6, np.linspace(400e-9, 700e-9, 64), np.full(64, 20e-3), ctx, "asm", src_units=1)
    p.set_intensity(s, np.ones(64))
    p.fill_schell(s, W0, SIGMA_C, SEED, 0)
    s = p.add(1024, 2e-6, np.full(128, 532e-9), np.linspace(0, 20e-3, 128), ctx, "asm", src_units=1)
    p.fill_schell(s, W0, SIGMA_C, SEED, 1)
    for k, n in enumerate((384, 1000, 1536)):
        cnt = len(range(k, 32, 3))
        s = p.add(n, 4.096e-3 / n, np.full(cnt, 532e-9), np.full(cnt, 20e-3), ctx, "asm")
        p.fill_schell(s, W0, SIGMA_C, SEED, 2 + 100 * k)
    return p


def bench_pipeline(wos, steps, tune=True, retune=False):
    """One hipGraph: 64 x 2048^2 polychromatic intensity, 128 z-planes at
    1024^2, 32 mixed fields (384^2, 1000^2, 1536^2); Gaussian-Schell inputs."""
    out, res = {}, {}
    for name, prec in (("fp32", wos.FP32), ("fp64", wos.FP64)):
        p = config4_pipeline(wos, prec)
        source = "default"
        if tune:
            source = "table"
            if retune or p.load_tuning() < 5:
                p.autotune(3)
                source = "autotune"
        p.capture()
        ms, st = p.time(2, max(steps, 10), stages=True)
        p.run(1)
        p.sync()
        res[name] = (p.intensity(0), p.download(3, 10))
        S = 8 if prec == wos.FP32 else 16
        B = 8 * S * (64 * 2048 ** 2 + 128 * 1024 ** 2 + 11 * 384 ** 2 + 11 * 1000 ** 2 + 10 * 1536 ** 2)
        out[name] = {"ms_per_graph_run": round(ms, 4), "props_per_s": round(224 / (ms * 1e-3), 1),
                     "stage_ms": [round(x, 4) for x in st],
                     "hbm_frac_algorithmic": round(B / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     "tuning": source}
        p.close()
    out["hbm_note"] = ("algorithmic bytes 8 M^2 S per propagation; the shared-source stages compute the "
                       "source's forward spectrum once and move about half of them")
    out["rel_l2_fp32_vs_fp64_intensity"] = rel_l2(res["fp32"][0], res["fp64"][0])
    out["rel_l2_fp32_vs_fp64_1000sq_field"] = rel_l2(res["fp32"][1], res["fp64"][1])
    return out


def bench_adjoint(wos):
    """Phase-mask design: 8 masks of 1024^2, forward loss + adjoint gradient."""
    n, units, dx, wl, z = 1024, 8, 4e-6, 532e-9, 20e-3
    x = (np.arange(n) - n / 2) * dx
    t = np.zeros((n, n))
    for k in range(16):
        a = 2 * np.pi * k / 16
        t += np.exp(-((x[None, :] - 800e-6 * np.cos(a)) ** 2 + (x[:, None] - 800e-6 * np.sin(a)) ** 2) / (2 * 12e-6 ** 2))
    t /= t.sum()
    theta = np.random.default_rng(1).uniform(0, 2 * np.pi, (units, n, n)).astype(np.float32).astype(np.float64)
    out = {}
    keep = {"theta0": theta[0], "grad0": {}, "loss0": {}}
    for name, prec in (("fp32", wos.FP32), ("fp64", wos.FP64)):
        op = wos.MaskOp(n, units, dx, z, wos.ctx_for(wl), t, precision=prec)
        keep["loss0"][name] = float(op.forward(theta)[0])
        keep["grad0"][name] = op.backward()[0]
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            op.forward(theta)
            op.backward()
        el = (time.perf_counter() - t0) / reps
        # device-resident step: theta stays in device memory, the gradient too
        op.forward(None)
        op.backward(to_host=False)
        t0 = time.perf_counter()
        for _ in range(reps):
            op.forward(None)
            op.backward(to_host=False)
        eld = (time.perf_counter() - t0) / reps
        out[name] = {"ms_forward_backward_8_masks": round(el * 1e3, 2), "note": "host theta in, host gradient out",
                     "ms_device_resident": round(eld * 1e3, 3)}
        op.close()
    out["grad_rel_l2_fp32_vs_fp64"] = rel_l2(keep["grad0"]["fp32"], keep["grad0"]["fp64"])
    return out, keep


def config12_inputs():
    """Inputs of configurations 1 and 2 in closed form.
    1: 512^2, dx 2 um, 632.8 nm: unit-amplitude Gaussian (w0 200 um) through
       a thin lens f = 50 mm, phase (-k/(2f)) r^2; asm over 50 mm.
    2: 1000^2, dx 1.5 um, 532 nm: x-polarised plane wave times a seeded i.i.d.
       uniform [0, 2 pi) phase per pixel inside a circle of diameter 1.2 mm,
       Ey = 0 (Polarized, both components propagate); asm_pad over 10 mm."""
    n, dx, wl = 512, 2e-6, 632.8e-9
    x = (np.arange(n) - n / 2) * dx
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    k = 2 * np.pi / wl
    c1 = np.exp(-r2 / (200e-6) ** 2) * np.exp(1j * ((-k / (2 * 0.05)) * r2))
    n2, dx2 = 1000, 1.5e-6
    x2 = (np.arange(n2) - n2 / 2) * dx2
    inside = np.hypot(x2[None, :], x2[:, None]) <= 0.6e-3
    phase = np.random.default_rng(7).uniform(0, 2 * np.pi, (n2, n2))
    c2 = np.where(inside, np.exp(1j * phase), 0)
    return {"config1": (c1, None, n, dx, wl, 0.05, "asm"),
            "config2": (c2, np.zeros_like(c2), n2, dx2, 532e-9, 10e-3, "asm_pad")}


def bench_config12(wos, tune=True, retune=False):
    """Configurations 1 and 2 as device-resident batches (one unit per field
    component; timed as stream launches, and as hipGraph replays) and as
    drop-in Propagate calls on host buffers. Algorithmic bytes per component: 8 M^2 S (M = 2n for
    asm_pad)."""
    out, fields = {}, {}
    for name, (ex, ey, n, dx, wl, z, method) in config12_inputs().items():
        comps = [ex] if ey is None else [ex, ey]
        M = 2 * n if method == "asm_pad" else n
        res = {"grid": n, "method": method, "components": len(comps)}
        for pname, prec in (("fp64", wos.FP64), ("fp32", wos.FP32)):
            b = wos.Batch(n, dx, [wl] * len(comps), [z] * len(comps), wos.ctx_for(wl), method, prec)
            _, _, source = tune_batch(b, tune, retune, 3)
            for u, f in enumerate(comps):
                b.upload(u, f)
            # the checked fields first: timing propagates the batch in place
            b.run(1)
            b.sync()
            fields[(name, pname)] = [b.download(u) for u in range(len(comps))]
            # stream launches back to back (a 512^2 propagation takes less GPU
            # time than one graph launch takes to submit), then graph replays
            ms = sorted(b.time(3, 30)[0] for _ in range(3))[1]   # median of three timings
            b.capture()
            ms_graph = sorted(b.time(3, 30)[0] for _ in range(3))[1]
            b.close()
            S = 16 if prec == wos.FP64 else 8
            t0 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                wos.propagate(ex, ey, dx, z, method, wos.ctx_for(wl), precision=prec)
            # bytes the three passes move by construction: asm 6 n^2 S per component; asm_pad
            # never stores the zero rows of the 2n grid: n^2 in + 2n^2 out (rows), 2n^2 each way
            # (columns), 2n^2 in + n^2 out (inverse rows) = 10 n^2 S
            moved = len(comps) * (10 if method == "asm_pad" else 6) * n * n * S
            res[pname] = {"props_per_s": round(len(comps) / (ms * 1e-3), 1), "device_ms": round(ms, 4),
                          "graph_replay_ms": round(ms_graph, 4),
                          "hbm_frac_algorithmic": round(len(comps) * 8 * M * M * S / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                          "hbm_frac_moved": round(moved / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                          "dropin_ms_per_call": round((time.perf_counter() - t0) * 1e3 / reps, 2),
                          "tuning": source}
        a32, a64 = np.concatenate(fields[(name, "fp32")]), np.concatenate(fields[(name, "fp64")])
        res["rel_l2_fp32_vs_fp64"] = rel_l2(a32, a64)
        res["hbm_note"] = ("hbm_frac_algorithmic prices 8 M^2 S per component (M the transform side); "
                           "hbm_frac_moved the bytes the kernels move (6 n^2 S asm, 10 n^2 S for the pruned asm_pad); "
                           "a 512^2 field stays in the last-level cache, so neither is an HBM measurement there")
        out[name] = res
    return out, fields


def propagate_bl(f, z, ctx, bl=(0.9, 0.05)):
    """A propagate element with the band limit: the device field runs it as
    one call (one fused transfer for asm), the oracle as Propagate then
    ApplyBandlimit."""
    if type(f).__name__ == "OField":
        f.propagate(z, "asm", ctx)
        f.bandlimit(*bl)
    else:
        f.propagate(z, "asm", ctx, bandlimit=bl)


def train_steps(f, m, ctx, shape):
    """A train of runTrain's element kinds on one field: lens, propagation
    (each followed by the band limit 0.9/0.05), soft circular aperture,
    correlated diffuser, polariser (the field becomes a vector field),
    quarter-wave plate, two more propagations."""
    c, s = np.cos(0.5), np.sin(0.5)
    pol = (c * c, s * c, s * c, s * s)
    e1, e2 = np.exp(1j * np.pi / 4), np.exp(-1j * np.pi / 4)
    c2, s2 = np.cos(0.2), np.sin(0.2)
    ret = ((c2 * c2) * e1 + (s2 * s2) * e2, (s2 * c2) * (e1 - e2), (s2 * c2) * (e1 - e2), (s2 * s2) * e1 + (c2 * c2) * e2)
    f.lens(633e-9, 0.2, aperture=1.8e-3)
    propagate_bl(f, 0.02, ctx)
    f.aperture(shape)
    f.diffuser(sigma=0.8, corr=10e-6, amp=1.0, seed=3)
    propagate_bl(f, 0.01, ctx)
    f.jones(*pol)
    f.jones(*ret)
    propagate_bl(f, 0.01, ctx)


def bench_train(wos, n=2048, dx=2e-6, reps=5):
    """The element train above on a device-resident 2048^2 field (no host
    round trip between elements), upload and download outside the timing."""
    x = (np.arange(n) - n / 2) * dx
    g = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (0.8e-3) ** 2).astype(np.complex128)
    out, res = {}, {}
    for name, prec in (("fp64", wos.FP64), ("fp32", wos.FP32)):
        shape = wos.make_shape("circle", radius=1.5e-3, edge_sigma=5e-6)
        f = wos.Field.from_host(g, dx=dx, precision=prec)
        train_steps(f, wos, wos.ctx_for(633e-9), shape)
        res[name] = f.download()
        ts = []
        for _ in range(reps):
            f.upload(g)
            t0 = time.perf_counter()
            train_steps(f, wos, wos.ctx_for(633e-9), shape)
            ts.append(time.perf_counter() - t0)
        f.close()
        out[name] = {"ms_per_train": round(float(np.median(ts)) * 1e3, 3)}
    out["elements_per_train"] = 11
    out["rel_l2_fp32_vs_fp64"] = rel_l2(np.concatenate(res["fp32"]), np.concatenate(res["fp64"]))
    return out, g, res["fp64"]


def cpu_timed(fn, reps, warm=1):
    """Median wall time of `reps` calls after `warm` untimed calls."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_model():
    """The host CPU's model string as /proc/cpuinfo gives it."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Host threads the CPU baseline runs on, and what they are taken from:
    every CPU this process may run on (sched_getaffinity), capped by
    OMP_NUM_THREADS when the environment sets it (a job's CPU share on a
    shared host); nproc is reported beside it."""
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = nproc
    threads, source = usable, "sched_getaffinity"
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and 0 < int(omp) < threads:
        threads, source = int(omp), "OMP_NUM_THREADS (the job's CPU share)"
    return {"nproc": nproc, "affinity_cpus": usable, "threads": threads, "threads_from": source,
            "cpu_model": cpu_model()}


def cpu_configs(cfg12, fields, bpm, bpm_check20, adj):
    """Restated reference CPU path (the FP64 oracle) on each configuration:
    the checker for the GPU outputs and the CPU timing beside them (median
    of the runs after one warm-up; the host threads of host_threads(), and 8
    threads for configuration 1 as in the reference's published 8-core
    figures)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    out = host_threads()
    threads = out["threads"]
    for name, (ex, ey, n, dx, wl, z, method) in config12_inputs().items():
        O.set_threads(threads)
        ctx = O.ctx_for(wl)
        ref = O.propagate(ex, ey, dx, z, method, ctx)
        refc = np.concatenate([ref[0]] if ey is None else [ref[0], ref[1]])
        t = cpu_timed(lambda: O.propagate(ex, ey, dx, z, method, ctx), 5)
        r = {"cpu_ms_per_propagate": round(t * 1e3, 1),
             "rel_l2_fp64_vs_oracle": rel_l2(np.concatenate(fields[(name, "fp64")]), refc),
             "rel_l2_fp32_vs_oracle": rel_l2(np.concatenate(fields[(name, "fp32")]), refc)}
        if name == "config1" and threads > 8:
            O.set_threads(8)
            r["cpu_ms_per_propagate_8_threads"] = round(cpu_timed(lambda: O.propagate(ex, ey, dx, z, method, ctx), 5) * 1e3, 1)
            O.set_threads(threads)
        r["gpu_speedup_fp64"] = round(t * cfg12[name]["fp64"]["props_per_s"] / cfg12[name]["components"], 1)
        out[name] = r
    # configuration 3: a 5-slice sample of the 400-slice run
    n, dx, wl = 1024, 0.5e-6, 1.55e-6
    x = (np.arange(n) - n / 2) * dx
    rr = np.hypot(x[None, :], x[:, None])
    index = np.where(rr < 100e-6, 1.45 - 0.01 * (rr / 100e-6) ** 2, 1.44) + 1e-5j
    ex = np.exp(-(rr / 10e-6) ** 2).astype(np.complex128)
    ctx = O.ctx_for(wl)
    t = cpu_timed(lambda: O.split_step(ex, None, dx, 5 * 5e-6, 5, index, ctx), 5)
    ref, _, _ = O.split_step(ex, None, dx, 20 * 5e-6, 20, index, ctx)
    out["config3"] = {"cpu_steps_per_s": round(5 / t, 2), "sample": "5 slices of 5 um",
                      "gpu_speedup_fp64": round(bpm["fp64"]["steps_per_s"] * t / 5, 1),
                      "rel_l2_fp64_vs_oracle_20_slices": rel_l2(bpm_check20["fp64"], ref),
                      "rel_l2_fp32_vs_oracle_20_slices": rel_l2(bpm_check20["fp32"], ref)}
    # configuration 4: one unit of each grid, weighted by the unit counts of the graph
    ctx = O.ctx_for(532e-9)
    per = {}
    for n, dx in ((2048, 2e-6), (1024, 2e-6), (384, 4.096e-3 / 384), (1000, 4.096e-3 / 1000), (1536, 4.096e-3 / 1536)):
        f = np.exp(-((np.arange(n) - n / 2) * dx / W0) ** 2)
        f = (f[None, :] * f[:, None]).astype(np.complex128)
        per[n] = cpu_timed(lambda: O.propagate(f, None, dx, 20e-3, "asm", ctx), 5 if n <= 1024 else 3)
    total = 64 * per[2048] + 128 * per[1024] + 11 * per[384] + 11 * per[1000] + 10 * per[1536]
    out["config4"] = {"cpu_s_per_224_props": round(total, 2), "cpu_props_per_s": round(224 / total, 2),
                      "sample": "one propagation per grid size (median of 5, 3 at 1536^2 and 2048^2), "
                                "weighted 64/128/11/11/10"}
    # configuration 5: one mask's forward + adjoint, times 8
    n, dx, wl, z = 1024, 4e-6, 532e-9, 20e-3
    xx = (np.arange(n) - n / 2) * dx
    tg = np.zeros((n, n))
    for k in range(16):
        a = 2 * np.pi * k / 16
        tg += np.exp(-((xx[None, :] - 800e-6 * np.cos(a)) ** 2 + (xx[:, None] - 800e-6 * np.sin(a)) ** 2) / (2 * 12e-6 ** 2))
    tg /= tg.sum()
    ctx = O.ctx_for(wl)
    loss0, grad0 = O.mask_loss_grad(adj["theta0"], tg, dx, z, ctx)
    t = cpu_timed(lambda: O.mask_loss_grad(adj["theta0"], tg, dx, z, ctx), 5)
    out["config5"] = {"cpu_ms_forward_backward_8_masks": round(8 * t * 1e3, 1), "sample": "one mask, times 8",
                      "grad_rel_l2_fp64_vs_oracle": rel_l2(adj["grad0"]["fp64"], grad0),
                      "grad_rel_l2_fp32_vs_oracle": rel_l2(adj["grad0"]["fp32"], grad0),
                      "loss_rel_fp64_vs_oracle": abs(adj["loss0"]["fp64"] - loss0) / loss0}
    return out


def cpu_train(g, out64, dx=2e-6):
    """The element train on the oracle (the checker and the CPU time beside
    the device-resident train)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_threads(host_threads()["threads"])
    shape = O.make_shape("circle", radius=1.5e-3, edge_sigma=5e-6)
    t0 = time.perf_counter()
    f = O.OField(g, dx=dx)
    train_steps(f, O, O.ctx_for(633e-9), shape)
    t = time.perf_counter() - t0
    return {"cpu_ms_per_train": round(t * 1e3, 1),
            "rel_l2_fp64_vs_oracle": rel_l2(np.concatenate(out64), np.concatenate([f.ex, f.ey]))}


def cpu_baseline(inp, wl, reps=5):
    """Restated reference CPU path (the FP64 oracle: radix-2 FFT rows and
    columns over the host threads, serial TF loop and Power sums, as
    optics/fft.go:91-242 and optics/propagate.go:132-148) on one 2048^2 unit
    of the same batch: the median of `reps` propagations after one warm-up,
    on the host threads of host_threads() and on 8 threads (the reference's
    published figures are for 8 cores, README.md:83-89)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ht = host_threads()
    threads = ht["threads"]
    ctx = O.ctx_for(wl)
    O.set_threads(threads)
    out, _, _, _ = O.propagate(inp, None, DX, Z, "asm", ctx)   # the warm-up, kept as the checker's output
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.propagate(inp, None, DX, Z, "asm", ctx)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    O.set_threads(8)
    med8 = cpu_timed(lambda: O.propagate(inp, None, DX, Z, "asm", ctx), reps)
    O.set_threads(threads)
    return out, dict({"value": round(1 / med, 3), "unit": "props/s", "cores": threads, "kind": "port",
                      "sample": f"median of {reps} FP64 ASM propagations of one 2048x2048 complex128 unit "
                                f"(unit 0 of the batch) after 1 warm-up, {threads} threads "
                                f"(runs {', '.join(f'{x:.3f}' for x in ts)} s)",
                      "s_per_prop": round(med, 4),
                      "eight_threads": {"s_per_prop": round(med8, 4), "props_per_s": round(1 / med8, 3),
                                        "reference_8_cores_s_per_prop": 0.7,
                                        "reference_source": "README.md:83-89 (~0.7 s at 2048^2, 8 cores)"}},
                     **ht)


def write_tuning(wos, path):
    """Autotune the batches and pipeline stages the benchmark runs and write
    their shapes as a tuning table (the format wos_hip_batch_load_tuning
    reads)."""
    lines = [f"# launch shapes of libwos_hip on {wos.device_name()}, written by bench.py --write-tuning",
             "# n precision method units tile_w col_w rows groups split radix col_radix col_split inv_rows inv_groups"
             " inv_split ms"]
    pname = {wos.FP32: "fp32", wos.FP64: "fp64"}

    def add(n, prec, method, units, shape, ms):
        vals = " ".join(str(int(shape[k])) for k in wos.TUNING_KEYS)
        lines.append(f"{n} {pname[prec]} {method} {units} {vals} {ms:.4f}")

    for n, dx, method, units in ((N, DX, "asm", UNITS), (512, 2e-6, "asm", 1), (1000, 1.5e-6, "asm_pad", 2)):
        for prec in (wos.FP32, wos.FP64):
            b = wos.Batch(n, dx, np.linspace(500e-9, 600e-9, units), np.full(units, Z), wos.ctx_for(550e-9),
                          method, prec)
            b.fill_schell(W0, SIGMA_C, SEED)
            shape, ms = b.autotune(5)
            add(n, prec, method, units, shape, ms)
            b.close()
    for prec in (wos.FP32, wos.FP64):
        p = config4_pipeline(wos, prec)
        ms = p.autotune(3)
        for i, (n, units) in enumerate(p.stages):
            add(n, prec, "asm", units, p.get_tuning(i), ms[i])
        p.close()
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the headline batch (no other workloads)")
    ap.add_argument("--no-tune", action="store_true",
                    help="extra workloads at their default launch shapes (profiling: no autotuner launches)")
    ap.add_argument("--shape32", default=None, help="fixed FP32 launch shape tile_w,col_w,rows,groups,split"
                    "[,radix[,col_radix[,col_split[,inv_rows,inv_groups,inv_split]]]]")
    ap.add_argument("--shape64", default=None, help="fixed FP64 launch shape (as --shape32)")
    ap.add_argument("--only", default=None, choices=["config12", "bpm", "pipeline", "adjoint", "train", "persistent"],
                    help="run one of the other workloads alone and print its figures (profiling)")
    ap.add_argument("--retune", action="store_true", help="autotune even where the tuning table has a shape")
    ap.add_argument("--write-tuning", default=None, metavar="PATH",
                    help="autotune the benchmark's batches and pipeline stages, write the tuning table, exit")
    args = ap.parse_args()
    shape32 = [int(x) for x in args.shape32.split(",")] if args.shape32 else None
    shape64 = [int(x) for x in args.shape64.split(",")] if args.shape64 else None
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the units of a batch are independent; several GPUs would each take a share of them)", file=sys.stderr)
        sys.exit(2)

    import wos_hip as wos
    wos.init()
    if args.write_tuning:
        lines = write_tuning(wos, args.write_tuning)
        print(json.dumps({"tuning_table": args.write_tuning, "entries": len(lines) - 2}), flush=True)
        return
    if args.only:
        tune = not args.no_tune
        run = {"config12": lambda: bench_config12(wos, tune, args.retune)[0], "bpm": lambda: bench_bpm(wos)[0],
               "pipeline": lambda: bench_pipeline(wos, args.steps, tune, args.retune),
               "adjoint": lambda: bench_adjoint(wos)[0], "train": lambda: bench_train(wos)[0],
               "persistent": lambda: bench_persistent(wos, args.steps)}[args.only]
        print(json.dumps({"only": args.only, "result": run()}), flush=True)
        return
    r32 = run_gpu(wos, wos.FP32, args.steps, args.warmup, shape32, args.retune)
    r64 = run_gpu(wos, wos.FP64, args.steps, args.warmup, shape64, args.retune)

    props32 = UNITS / (r32["ms_per_step"] * 1e-3)
    props64 = UNITS / (r64["ms_per_step"] * 1e-3)
    err32 = max(rel_l2(r32["unit0"], r64["unit0"]), rel_l2(r32["unit_last"], r64["unit_last"]))

    # Per-kernel roofline of the FP32 pipeline. Algorithmic bytes of a launch:
    # one read + one write of the 64 fields (each pass must stream the batch
    # through HBM once); the fused column kernel does the work of two of the
    # four passes of an unfused forward + inverse 2-D FFT with one of them.
    names = ["row_fft_fwd", "col_fft_tf_ifft", "row_fft_inv"]
    k = r32["kernel_ms"]
    dom = int(np.argmax(k))
    bytes_launch = 2.0 * UNITS * N * N * 8
    per_kernel = []
    for i, nm in enumerate(names):
        gbs = bytes_launch / (k[i] * 1e-3) / 1e9
        per_kernel.append({"kernel": nm, "ms": round(k[i], 4), "achieved_GBs": round(gbs, 1),
                           "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": PMC_HBM_BYTES["fp32"][nm],
                           "valu_issue_frac": round(PMC_VALU["fp32"][nm] / VALU_WAVE_INSTR_PER_S / (k[i] * 1e-3), 3)})
    achieved = per_kernel[dom]["achieved_GBs"]
    tuned32 = tuple(r32["tune"][k] for k in wos.TUNING_KEYS)
    traffic_ok = tuned32 == PMC_SHAPE["fp32"]
    if not traffic_ok:   # the PMC constants belong to another launch shape
        for pk in per_kernel:
            pk["traffic"] = None
    # whole step: algorithmic bytes 8*M^2*S per propagation (two 2-D FFTs x
    # two out-of-cache passes x read + write)
    step_gbs = props32 * 8.0 * N * N * 8 / 1e9

    line = {
        "metric": METRIC,
        "value": round(props32, 1),
        "unit": "props/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex; fp64 twiddles, kz*z phase and sums)",
        "data": "synthetic: seeded Gaussian-Schell fields generated on the device",
        "config": {"workload": "64 x 2048^2 ASM propagations per step (64 wavelengths 400-700 nm, "
                               "dx=2um, z=20mm), device resident, hipGraph replay",
                   "grid": N, "units_per_step": UNITS, "method": "asm"},
        "fp64": {"value": round(props64, 1), "unit": "props/s", "ms_per_step": round(r64["ms_per_step"], 4),
                 "kernel_ms": [round(x, 4) for x in r64["kernel_ms"]], "autotune": r64["tune"]},
        "autotune": r32["tune"],
        "fp32_over_fp64": round(props32 / props64, 3),
        "rel_l2_fp32_vs_fp64": err32,
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": PMC_HBM_BYTES["fp32"][names[dom]] if traffic_ok else None,
                     "traffic_shape": list(PMC_SHAPE["fp32"]),
                     "traffic_note": "PMC FETCH_SIZE x 2 + WRITE_SIZE per launch at traffic_shape"
                                     + ("" if traffic_ok else f"; this run's autotuned shape {list(tuned32)} differs, "
                                                              "so no traffic figure is given"),
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "kernels": per_kernel,
                     "step_algorithmic_GBs": round(step_gbs, 1),
                     "step_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                     # the three executed passes each read and write the batch once
                     "step_moved_GBs": round(props32 * 6.0 * N * N * 8 / 1e9, 1),
                     "step_moved_frac": round(props32 * 6.0 * N * N * 8 / 1e9 / HBM_PEAK_GBS, 4),
                     "step_note": "step_frac prices 8 N^2 S per propagation (four passes of an unfused forward + "
                                  "inverse 2-D FFT); step_moved_frac the 6 N^2 S the three executed passes move"},
    }
    if args.no_extra:
        print(json.dumps(line), flush=True)
        return
    line["pcie_inclusive"] = {
        "unit": "props/s", "workload": "one 2048^2 asm Propagate per call, complex128 host buffers in and out",
        "fp64": round(pcie_inclusive(wos, r64["unit0_in"], r64["wl0"], wos.FP64), 1),
        "fp32": round(pcie_inclusive(wos, r64["unit0_in"], r64["wl0"], wos.FP32), 1),
        "fft2d_z_2048_ms": round(fft2d_dropin_ms(wos, r64["unit0_in"]), 3),
        "fft2d_note": "one complex128 2-D FFT of a host buffer per call, transfers included; the reference "
                      "publishes 45 ms for this call through cuFFT on a Tesla T4 (README.md:113)"}
    line["persistent_one_launch"] = bench_persistent(wos, args.steps)
    cfg12, fields = bench_config12(wos, not args.no_tune, args.retune)
    line["config1_512_asm"] = cfg12["config1"]
    line["config2_1000_asm_pad_jones"] = cfg12["config2"]
    line["bpm_1024_400_slices"], bpm20 = bench_bpm(wos)
    line["pipeline_one_graph_224_props"] = bench_pipeline(wos, args.steps, not args.no_tune, args.retune)
    line["mask_adjoint"], adj = bench_adjoint(wos)
    line["element_train_2048"], train_in, train_out64 = bench_train(wos)
    if not args.no_cpu:
        ref, cb = cpu_baseline(r64["unit0_in"], r64["wl0"])
        line["cpu_baseline"] = cb
        line["rel_l2_fp64_vs_oracle"] = rel_l2(r64["unit0"], ref)
        line["rel_l2_fp32_vs_oracle"] = rel_l2(r32["unit0"], ref)
        line["cpu_configs"] = cpu_configs(cfg12, fields, line["bpm_1024_400_slices"], bpm20, adj)
        line["cpu_configs"]["element_train_2048"] = cpu_train(train_in, train_out64)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
