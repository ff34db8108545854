"""The C-ABI boundary: the library loads without a GPU, exports every symbol
include/*.h declares, and its taco_tensor_t is byte-identical to taco's."""
import ctypes
import os
import re
import subprocess

import pytest

from taco_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "taco_hip.h"),
           os.path.join(ROOT, "include", "taco", "taco_tensor_t.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{)]*\)\s*;", text):
            name = m.group(1)
            if name not in ("if", "while", "return", "sizeof"):
                names.add(name)
    return names


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    declared = declared_functions()
    assert "taco_hip_spmv_compute" in declared and "init_taco_tensor_t" in declared
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert declared <= set(_lib.EXPORTED), declared - set(_lib.EXPORTED)


def test_no_gpu_calls_fail_cleanly_here():
    L = _lib.lib()
    n = L.taco_hip_device_count()
    assert n >= 0
    flags = L.taco_hip_get_default_compiler_flags().decode()
    assert "--offload-arch=gfx" in flags and "-packed-fp32-ops" in flags


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "%s"
int main(void) {
  printf("%%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu\n", sizeof(taco_tensor_t),
         offsetof(taco_tensor_t, order), offsetof(taco_tensor_t, dimensions),
         offsetof(taco_tensor_t, csize), offsetof(taco_tensor_t, mode_ordering),
         offsetof(taco_tensor_t, mode_types), offsetof(taco_tensor_t, indices),
         offsetof(taco_tensor_t, vals), offsetof(taco_tensor_t, fill_value),
         offsetof(taco_tensor_t, vals_size));
  return 0;
}
"""


def layout_of(header, tmp_path, tag):
    src = tmp_path / f"layout_{tag}.c"
    exe = tmp_path / f"layout_{tag}"
    src.write_text(LAYOUT_C % header)
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True,
                          text=True).stdout.split()


def test_descriptor_layout_matches_ctypes(tmp_path):
    ours = [int(v) for v in layout_of(HEADERS[1], tmp_path, "ours")]
    T = _lib.TacoTensorT
    assert ours == [ctypes.sizeof(T), T.order.offset, T.dimensions.offset,
                    T.csize.offset, T.mode_ordering.offset, T.mode_types.offset,
                    T.indices.offset, T.vals.offset, T.fill_value.offset,
                    T.vals_size.offset]


REF_HEADER = "/root/reference/include/taco/taco_tensor_t.h"


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference tree absent")
def test_descriptor_layout_matches_taco(tmp_path):
    assert layout_of(HEADERS[1], tmp_path, "ours") == layout_of(REF_HEADER, tmp_path, "ref")


def test_init_deinit_descriptor():
    L = _lib.lib()
    dims = (ctypes.c_int32 * 2)(7, 9)
    order = (ctypes.c_int32 * 2)(0, 1)
    types = (ctypes.c_int * 2)(0, 1)
    fill = ctypes.c_double(0.5)
    t = L.init_taco_tensor_t(2, 64, dims, order, types, ctypes.byref(fill))
    assert t.contents.order == 2 and t.contents.csize == 64
    assert t.contents.dimensions[1] == 9 and t.contents.mode_types[1] == 1
    got = ctypes.cast(t.contents.fill_value, ctypes.POINTER(ctypes.c_double))[0]
    assert got == 0.5
    L.taco_hip_set_codegen_enabled(1)
    assert L.taco_hip_should_use_codegen() == 1
    L.taco_hip_set_codegen_enabled(0)
    assert L.taco_hip_should_use_codegen() == 0
    L.deinit_taco_tensor_t(t)


def test_unified_alloc_zero_is_null():
    assert _lib.lib().taco_hip_unified_alloc(0) is None


def test_product_does_not_import_the_oracle():
    """The shipped package never reaches into oracle/ (it is the checker)."""
    pkg = os.path.join(ROOT, "taco_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")) or f == "Makefile":
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text and "taco_ref" not in text, f
