"""The C ABI: the library loads, exports every declared entry point, and the
Python structs match the C layouts.  No GPU call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "footile_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:ft_status|void|const char \*)\s*(ft_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def built_lib():
    from footile_amd import _lib, build
    if not os.path.exists(_lib.library_path()):
        build.build()
    return ctypes.CDLL(_lib.library_path())


def test_header_declares_the_plotter_surface():
    names = declared_functions()
    for n in ["ft_plotter_new", "ft_plotter_fill", "ft_plotter_stroke", "ft_plotter_set_tolerance",
              "ft_plotter_set_transform", "ft_plotter_set_join", "ft_plotter_read", "ft_fig_fill",
              "ft_batch_new", "ft_batch_run"]:
        assert n in names


def test_library_exports_every_declared_symbol(built_lib):
    missing = [n for n in declared_functions() if not hasattr(built_lib, n)]
    assert not missing, missing


def test_struct_layouts_match_c(tmp_path):
    from footile_amd._lib import FillDescC
    from footile_amd.diff import _Layout
    from footile_amd.path import OP_DTYPE
    src = tmp_path / "layout.c"
    src.write_text(f'''#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ft_path_op), sizeof(ft_fill_desc),
         offsetof(ft_fill_desc, dst), offsetof(ft_fill_desc, fmt), offsetof(ft_fill_desc, clr),
         offsetof(ft_fill_desc, transform), offsetof(ft_fill_desc, tolerance),
         offsetof(ft_fill_desc, pen_width), offsetof(ft_fill_desc, fresh), sizeof(ft_diff_layout),
         offsetof(ft_diff_layout, path_image), offsetof(ft_diff_layout, path_cubics),
         offsetof(ft_diff_layout, flags));
  return 0;
}}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [OP_DTYPE.itemsize, ctypes.sizeof(FillDescC), FillDescC.dst.offset, FillDescC.fmt.offset,
            FillDescC.clr.offset, FillDescC.transform.offset, FillDescC.tolerance.offset,
            FillDescC.pen_width.offset, FillDescC.fresh.offset, ctypes.sizeof(_Layout), _Layout.path_image.offset,
            _Layout.path_cubics.offset, _Layout.flags.offset]
    assert got == want


def test_missing_library_fails_loudly(monkeypatch):
    from footile_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "_LIB_FILE", "/nonexistent/libfootile_hip.so")
    with pytest.raises(_lib.FootileError):
        _lib.lib()


def test_ops_array_layout():
    from footile_amd.path import CUBIC, ops_array
    a = ops_array([(CUBIC, (1, 2, 3, 4, 5, 6))])
    raw = a.tobytes()
    assert len(raw) == 28
    assert np.frombuffer(raw[:4], "<u4")[0] == CUBIC
    assert list(np.frombuffer(raw[4:], "<f4")) == [1, 2, 3, 4, 5, 6]
