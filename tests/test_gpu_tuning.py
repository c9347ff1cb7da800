"""Tuning tables: launch shapes from a file instead of an autotune at every
start. The lookup, the refusals and the committed table itself."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_tuning_table_lookup(wos, tmp_path):
    t = tmp_path / "tune.txt"
    t.write_text("# test table\n"
                 "1024 fp32 asm 4 8 4 2 0 0 16 16 0 0.1\n"
                 "1024 fp32 asm 64 4 4 4 0 1 16 16 0\n"
                 "1024 fp32 asm 16 4 4 4 2 1 16 16 1 2 2 1 0.5   # its own inverse row shape\n"
                 "1024 fp64 asm 4 4 4 2 0 1 16 16 1   # FP64 entry\n"
                 "1000 fp32 asm_pad 2 8 2 1 0 0 16 16 0\n")
    ctx = wos.ctx_for(532e-9)
    b = wos.Batch(1024, 2e-6, np.full(8, 532e-9), np.full(8, 1e-3), ctx, "asm", wos.FP32)
    assert b.load_tuning(str(t)) == dict(tile_w=8, col_w=4, rows=2, row_groups=0, split_lds=0, radix=16,
                                         col_radix=16, col_split_lds=0, inv_rows=2, inv_row_groups=0,
                                         inv_split_lds=0)   # 8 units: nearer to 4 than to 64; inverse as forward
    b.close()
    b = wos.Batch(1024, 2e-6, np.full(16, 532e-9), np.full(16, 1e-3), ctx, "asm", wos.FP32)
    assert b.load_tuning(str(t)) == dict(tile_w=4, col_w=4, rows=4, row_groups=2, split_lds=1, radix=16,
                                         col_radix=16, col_split_lds=1, inv_rows=2, inv_row_groups=2,
                                         inv_split_lds=1)
    b.close()
    b = wos.Batch(1024, 2e-6, np.full(40, 532e-9), np.full(40, 1e-3), ctx, "asm", wos.FP32)
    assert b.load_tuning(str(t))["rows"] == 4                                # 40 units: the 64-unit entry
    b.close()
    b = wos.Batch(1024, 2e-6, np.full(2, 532e-9), np.full(2, 1e-3), ctx, "asm", wos.FP64)
    assert b.load_tuning(str(t))["col_split_lds"] == 1
    b.close()
    b = wos.Batch(512, 2e-6, [532e-9], [1e-3], ctx, "asm", wos.FP32)
    assert b.load_tuning(str(t)) is None                                     # no entry for 512
    b.close()
    b = wos.Batch(1000, 1.5e-6, np.full(2, 532e-9), np.full(2, 1e-3), ctx, "asm_pad", wos.FP32)
    assert b.load_tuning(str(t))["tile_w"] == 8                              # a mixed-radix grid: tile widths
    b.close()
    bad = tmp_path / "bad.txt"
    bad.write_text("1024 fp16 asm 4 8 4 2 0 0 16 16 0\n")
    b = wos.Batch(1024, 2e-6, [532e-9], [1e-3], ctx, "asm", wos.FP32)
    with pytest.raises(wos.WosError, match="bad entry on line 1"):
        b.load_tuning(str(bad))
    bad.write_text("1024 fp32 asm 4 3 3 2 0 0 16 16 0\n")                   # a shape the kernels do not build
    with pytest.raises(wos.WosError, match="not built"):
        b.load_tuning(str(bad))
    bad.write_text("1024 fp32 asm 4 4 4 4 0 1 16 16 0 3 0 0\n")             # an inverse row shape not built
    with pytest.raises(wos.WosError, match="not built"):
        b.load_tuning(str(bad))
    bad.write_text("1024 fp32 asm 4 4 4 4 0 1 16 16 0 4 0\n")               # two of the three inverse fields
    with pytest.raises(wos.WosError, match="bad entry on line 1"):
        b.load_tuning(str(bad))
    with pytest.raises(wos.WosError, match="cannot open"):
        b.load_tuning(str(tmp_path / "missing.txt"))
    b.close()


def test_pipeline_tuning_table_replays_match(wos, tmp_path):
    """Stages take the table's shapes; the replay then equals an eager run
    at those shapes, and get_tuning reports what was applied."""
    t = tmp_path / "tune.txt"
    t.write_text("256 fp32 asm 3 8 8 2 0 0 16 16 0\n384 fp32 asm 2 8 2 1 0 0 16 16 0\n")
    ctx = wos.ctx_for(532e-9)
    p = wos.Pipeline(wos.FP32)
    s0 = p.add(256, 2e-6, np.array([450e-9, 550e-9, 650e-9]), np.full(3, 2e-3), ctx, "asm", src_units=1)
    s1 = p.add(384, 2e-6, np.array([532e-9, 600e-9]), np.full(2, 3e-3), ctx, "asm")
    s2 = p.add(128, 2e-6, np.array([532e-9]), np.full(1, 1e-3), ctx, "asm")
    rng = np.random.default_rng(3)
    for s, n, k in ((s0, 256, 1), (s1, 384, 2), (s2, 128, 1)):
        for u in range(k):
            p.upload(s, u, rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    assert p.load_tuning(str(t)) == 2
    assert p.get_tuning(s0)["tile_w"] == 8 and p.get_tuning(s0)["rows"] == 2
    assert p.get_tuning(s1)["col_w"] == 2
    p.run(1)
    p.sync()
    eager = [p.download(s0, u) for u in range(3)] + [p.download(s1, u) for u in range(2)]
    p.capture()
    p.run(1)
    p.sync()
    again = [p.download(s0, u) for u in range(3)] + [p.download(s1, u) for u in range(2)]
    assert all(np.array_equal(a, b) for a, b in zip(eager, again))
    p.set_tuning(s2, dict(tile_w=4, col_w=4, rows=0, row_groups=0, split_lds=0, radix=16, col_radix=16,
                          col_split_lds=0))
    assert p.get_tuning(s2)["tile_w"] == 4
    p.close()


def test_committed_tuning_table_applies(wos):
    """The table committed beside the library parses and its headline entry
    applies to the benchmark's batch in both precisions."""
    assert os.path.exists(wos.TUNING_TABLE)
    for prec in (wos.FP32, wos.FP64):
        b = wos.Batch(2048, 2e-6, np.linspace(400e-9, 700e-9, 64), np.full(64, 20e-3), wos.ctx_for(550e-9), "asm",
                      prec)
        shape = b.load_tuning()
        print(f"committed tuning table, 64 x 2048^2 precision {prec}: {shape}", flush=True)
        assert shape is not None and shape["tile_w"] > 0
        b.close()
