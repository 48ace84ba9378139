"""The C-ABI library loads and exports every symbol include/clkmi.h declares,
and its host-side validation answers with OpenCL codes -- no GPU needed."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "clkmi.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int32_t|const char\*)\s+(clk_\w+)\s*\(", text, re.M)))


def test_header_declares_the_five_kernels():
    names = declared_functions()
    for k in ("gaussian_blur_separable_3d", "mean_separable_3d", "minimum_separable_3d",
              "maximum_separable_3d", "median_box_3d"):
        assert "clk_" + k in names


def test_library_exports_every_declared_symbol():
    from clkmi import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(_lib.EXPORTED)


def test_dtype_sizes_and_version():
    from clkmi import _lib
    L = _lib.lib()
    assert [L.clk_dtype_size(i) for i in range(6)] == [1, 1, 2, 2, 4, 0]
    assert L.clk_version() >= 100


def test_host_validation_without_device():
    import clkmi
    from clkmi import _lib
    L = _lib.lib()
    img = _lib.ClkImage(0x1000, 3, 0, 8, 8, 8)
    bad_type = _lib.ClkImage(0x100000, 7, 0, 8, 8, 8)
    rc = L.clk_mean_separable_3d(None, ctypes.byref(img), None, 0, 3, 0.0)
    assert rc == -38
    rc = L.clk_mean_separable_3d(None, ctypes.byref(bad_type), ctypes.byref(img), 0, 3, 0.0)
    assert rc == -30
    assert b"pixel type" in L.clk_last_error_string()
    rc = L.clk_maximum_separable_3d(None, ctypes.byref(img), ctypes.byref(img), 0, 3, 0.0)
    assert rc == -8
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.execute("sobel_3d", {"dst": clkmi.Image(0x1000, 1, 2, 2, 2)})
    assert e.value.code == -46


def test_unbuilt_library_fails_loudly(tmp_path, monkeypatch):
    from clkmi import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="not built"):
        _lib.lib()


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of clk_image, clk_arg and clk_filter_job have the C
    sizes and offsets (compiled with gcc from include/clkmi.h)."""
    import subprocess
    from clkmi import _lib
    src = tmp_path / "layout.c"
    src.write_text('''#include <stdio.h>
#include <stddef.h>
#include "clkmi.h"
int main(void) {
  printf("%zu %zu %zu %zu\\n", sizeof(clk_image), sizeof(clk_arg), offsetof(clk_arg, image), sizeof(clk_filter_job));
  printf("%zu %zu %zu %zu\\n", offsetof(clk_filter_job, dst), offsetof(clk_filter_job, src),
         offsetof(clk_filter_job, sigma), offsetof(clk_filter_job, radius));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    want = [ctypes.sizeof(_lib.ClkImage), ctypes.sizeof(_lib.ClkArg), _lib.ClkArg.image.offset,
            ctypes.sizeof(_lib.ClkFilterJob), _lib.ClkFilterJob.dst.offset, _lib.ClkFilterJob.src.offset,
            _lib.ClkFilterJob.sigma.offset, _lib.ClkFilterJob.radius.offset]
    assert [int(v) for v in out] == want


def test_tuning_knobs_validate_values():
    """Launch-shape knobs accept their documented ranges only (host only)."""
    import clkmi
    for key, good, bad in [("autotune", 1, 2), ("strided_order", 3, 4), ("ystatic_cols", 4, 3),
                           ("xreg_maxc", 16, 17), ("strided_rounds", 64, 0), ("batch_seg", 8, 7),
                           ("xreg_units", 4, 3), ("x_tasks", 2, 3), ("sphere_direct", 2, 3), ("yring", 2, 3),
                           ("yring_deep", 1, 2)]:
        old = clkmi.get_tuning(key)
        try:
            clkmi.set_tuning(key, good)
            assert clkmi.get_tuning(key) == good
            with pytest.raises(clkmi.ClkError) as e:
                clkmi.set_tuning(key, bad)
            assert e.value.code == -30
        finally:
            clkmi.set_tuning(key, old)
    with pytest.raises(clkmi.ClkError):
        clkmi.set_tuning("no_such_knob", 1)


def test_autotune_cache_round_trip(tmp_path):
    """Saved launch shapes load back (host only; no tuning runs here)."""
    import clkmi
    src = tmp_path / "shapes.txt"
    src.write_text("# comment\n0 3 4 1 25 1024 1024 512 : strided_rounds=16 strided_order=1 ystatic_cols=2\n"
                   "2 3 3 0 7 1024 1024 512 : xreg_units=4\n"
                   "0 4 4 2 13 1024 1024 512 : strided_rounds=8 strided_order=2 ystatic_cols=0 yring=2 "
                   "yring_deep=1\n")
    clkmi.autotune_load(src)
    out = tmp_path / "out.txt"
    clkmi.autotune_save(out)
    lines = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert "0 3 4 1 25 1024 1024 512 : strided_rounds=16 strided_order=1 ystatic_cols=2" in lines
    assert "2 3 3 0 7 1024 1024 512 : xreg_units=4" in lines
    assert ("0 4 4 2 13 1024 1024 512 : strided_rounds=8 strided_order=2 ystatic_cols=0 yring=2 yring_deep=1"
            in lines)
    bad = tmp_path / "bad.txt"
    bad.write_text("0 3 4 1 25 1024 1024 512 : no_such_knob=1\n")
    with pytest.raises(clkmi.ClkError) as e:
        clkmi.autotune_load(bad)
    assert e.value.code == -30
    with pytest.raises(clkmi.ClkError):
        clkmi.autotune_load(tmp_path / "missing.txt")


def test_noalloc_validation_without_device():
    """clk_filter_3d_noalloc checks its job on the host: NULL, bad kind,
    overlap, and temporaries beyond its fixed array (checked before any
    device call)."""
    from clkmi import _lib
    L = _lib.lib()
    assert L.clk_filter_3d_noalloc(None, None) == -30
    j = _lib.ClkFilterJob()
    j.kind = 0
    j.src = _lib.ClkImage(0x100000000, 3, 0, 256, 256, 257)
    j.dst = _lib.ClkImage(0x200000000, 4, 0, 256, 256, 257)
    for k in range(3):
        j.sigma[k] = 1.0
    assert L.clk_filter_3d_noalloc(None, ctypes.byref(j)) == -30
    assert b"fixed array" in L.clk_last_error_string()
    j.kind = 9
    assert L.clk_filter_3d_noalloc(None, ctypes.byref(j)) == -30
    j.kind = 2
    j.dst = j.src
    assert L.clk_filter_3d_noalloc(None, ctypes.byref(j)) == -8
