"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/wos_hip.h declares, the binding's structures match the C
layout, and the error path reports a missing device instead of computing."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wos_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wos_hip_\w+)\s*\(", text)))


def test_header_declares_and_library_exports_everything():
    import wos_hip
    lib = wos_hip.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(wos_hip.EXPORTS) == names


def test_library_symbols_are_c_linkage():
    import wos_hip
    out = subprocess.run(["nm", "-D", "--defined-only", wos_hip.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (wos_hip_\w+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layouts_match_header(tmp_path):
    """Compile a probe against the header with gcc and compare offsets."""
    import wos_hip
    src = tmp_path / "probe.c"
    src.write_text(r'''
#include <stddef.h>
#include <stdio.h>
#include "wos_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(wos_ctx), offsetof(wos_ctx, evanescent_zero),
         offsetof(wos_ctx, backward_regularize), offsetof(wos_ctx, evanescent_limit), offsetof(wos_ctx, tikhonov_alpha));
  printf("%zu %zu\n", sizeof(wos_warnings), offsetof(wos_warnings, value));
  printf("%zu %zu\n", sizeof(wos_prop_stats), offsetof(wos_prop_stats, method_used));
  printf("%d %d %d %d %d %d\n", WOS_METHOD_ASM, WOS_METHOD_ASM_PAD, WOS_METHOD_ASM_SHIFT,
         WOS_METHOD_ASM_SHIFT_PAD, WOS_METHOD_FRESNEL_TF, WOS_METHOD_FRESNEL_IR);
  printf("%d %d\n", WOS_FP64, WOS_FP32);
  printf("%zu %zu %zu %zu %zu\n", sizeof(wos_shape), offsetof(wos_shape, radius), offsetof(wos_shape, separation),
         offsetof(wos_shape, edge_sigma), offsetof(wos_shape, verts));
  printf("%d %d %d %d %d %d %d %d %d %d\n", WOS_SHAPE_CIRCLE, WOS_SHAPE_SLIT, WOS_SHAPE_SQUARE, WOS_SHAPE_RECTANGLE,
         WOS_SHAPE_ELLIPSE, WOS_SHAPE_SUPERELLIPSE, WOS_SHAPE_RING, WOS_SHAPE_DOUBLE_SLIT, WOS_SHAPE_CROSS,
         WOS_SHAPE_POLYGON);
  return 0;
}
''')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    ctx = [int(x) for x in lines[0].split()]
    assert ctx == [C.sizeof(wos_hip.Ctx), wos_hip.Ctx.evanescent_zero.offset,
                   wos_hip.Ctx.backward_regularize.offset, wos_hip.Ctx.evanescent_limit.offset,
                   wos_hip.Ctx.tikhonov_alpha.offset]
    assert [int(x) for x in lines[1].split()] == [C.sizeof(wos_hip.Warnings), wos_hip.Warnings.value.offset]
    assert [int(x) for x in lines[2].split()] == [C.sizeof(wos_hip.PropStats), wos_hip.PropStats.method_used.offset]
    assert [int(x) for x in lines[3].split()] == [wos_hip.METHODS[m] for m in
                                                  ("asm", "asm_pad", "asm_shift", "asm_shift_pad",
                                                   "fresnel_tf", "fresnel_ir")]
    assert [int(x) for x in lines[4].split()] == [wos_hip.FP64, wos_hip.FP32]
    S = wos_hip.Shape
    assert [int(x) for x in lines[5].split()] == [C.sizeof(S), S.radius.offset, S.separation.offset,
                                                  S.edge_sigma.offset, S.verts.offset]
    assert [int(x) for x in lines[6].split()] == [wos_hip.SHAPE_KINDS[k] for k in
                                                  ("circle", "slit", "square", "rectangle", "ellipse",
                                                   "superellipse", "ring", "double_slit", "cross", "polygon")]


def test_oracle_structs_match_library_structs():
    import oracle as O
    import wos_hip
    assert C.sizeof(O.Ctx) == C.sizeof(wos_hip.Ctx)
    assert [f[0] for f in O.Ctx._fields_] == [f[0] for f in wos_hip.Ctx._fields_]
    assert C.sizeof(O.Warnings) == C.sizeof(wos_hip.Warnings)
    assert O.WARN_CODES == wos_hip.WARN_CODES
    assert O.METHODS == wos_hip.METHODS
    assert C.sizeof(O.Shape) == C.sizeof(wos_hip.Shape)
    assert [f[0] for f in O.Shape._fields_] == [f[0] for f in wos_hip.Shape._fields_]
    assert O.SHAPE_KINDS == wos_hip.SHAPE_KINDS


def test_no_device_reports_error_without_computing():
    """Without a visible GPU every compute entry point returns
    WOS_ERR_NO_DEVICE (the Go shim then falls back to its pure-Go path);
    nothing is computed on the host."""
    import numpy as np
    import wos_hip
    lib = wos_hip.load()
    if lib.wos_hip_device_count() > 0:
        pytest.skip("a GPU is visible; the no-device path is not reachable")
    a = np.zeros((8, 8), dtype=np.complex128)
    rc = lib.wos_hip_fft2d_z(a.ctypes.data, 8, 0)
    assert rc == 1
    assert b"device" in lib.wos_hip_last_error()
    with pytest.raises(wos_hip.WosError):
        wos_hip.propagate(a, None, 1e-6, 1e-3, "asm", wos_hip.ctx_for(632.8e-9))
    with pytest.raises(wos_hip.WosError):
        wos_hip.Field(8, 1e-6)


def fft2d_with_fallback(wos_hip, oracle, a):
    """fft2D's dispatch (optics/fft.go:209-216): the GPU call, and the CPU
    path on any error, run on the caller's untouched array."""
    b = np.array(a, dtype=np.complex128, order="C")
    rc = wos_hip.load().wos_hip_fft2d_z(b.ctypes.data, b.shape[0], 0)
    if rc != 0:
        assert np.array_equal(b, a)   # a failed call leaves the slice as it was
        return oracle.fft2d(b), rc
    return b, rc


def test_failure_injection_drives_the_fallback(monkeypatch):
    """WOS_HIP_FAIL makes the listed entry points fail before they touch
    their arguments, on any machine: the caller's CPU path then runs on the
    unmodified data; unlisted entry points behave as without it; "all" (or a
    listed wos_hip_device_count) reports no device (GPUAvailable false)."""
    import oracle as O
    import wos_hip
    lib = wos_hip.load()
    a = O.ref_field(16, 1.3)
    monkeypatch.setenv("WOS_HIP_FAIL", "wos_hip_propagate,wos_hip_fft2d_z")
    got, rc = fft2d_with_fallback(wos_hip, O, a)
    assert rc == 3 and b"injected failure" in lib.wos_hip_last_error()
    assert np.allclose(got, np.fft.fft2(a), rtol=0, atol=1e-11)
    x = a.copy()
    with pytest.raises(wos_hip.WosError, match="injected"):
        wos_hip.propagate(x, None, 1e-6, 1e-3, "asm", wos_hip.ctx_for(632.8e-9))
    assert np.array_equal(x, a)
    count = lib.wos_hip_device_count()
    rc1 = lib.wos_hip_fft1d_z(x.ctypes.data, 16, 0)   # not listed
    assert rc1 != 3 or b"injected" not in lib.wos_hip_last_error()
    assert rc1 == (0 if count > 0 else 1)
    monkeypatch.setenv("WOS_HIP_FAIL", "all")
    assert lib.wos_hip_device_count() == 0
    assert lib.wos_hip_fft1d_z(x.ctypes.data, 16, 0) == 3
    monkeypatch.setenv("WOS_HIP_FAIL", "wos_hip_fft2d")   # a prefix names no entry point
    assert lib.wos_hip_fft2d_z(a.copy().ctypes.data, 16, 0) != 3 or b"injected" not in lib.wos_hip_last_error()
