"""The C ABI boundary without a device: libspai_hip.so loads and exports every
entry point that include/*.h declares, the structs keep the reference's
layout (include/methods_c_api.h:7-46), the headers compile as C and link
against the library, and the Cython mirror imports.  No compute calls."""
import ctypes
import re
import shutil
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "parallellocallyoptimalapproximatesparsematrixinverses_amd"
LIB = PKG / "libspai_hip.so"
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#[^\n]*", "", text)
        # drop struct / typedef bodies, keep prototypes
        text = re.sub(r"struct\s+\w+\s*\{.*?\};", "", text, flags=re.S)
        for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
            names.add(m.group(2))
    return names


def test_headers_declare_entry_points():
    names = declared_functions()
    for required in ("spai_global_cg", "spai_global_mr", "spai_global_lomr", "inner_outer_mr", "inner_outer_lomr",
                     "free_mem", "spai_run_ex", "spai_pcg", "spai_loss_forward", "spai_loss_backward"):
        assert required in names


@pytest.mark.skipif(not LIB.exists(), reason="libspai_hip.so is not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout():
    from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib
    assert ctypes.sizeof(_lib.MatrixData) == 48
    assert ctypes.sizeof(_lib.MethodParams) == 64
    assert ctypes.sizeof(_lib.SingleIterationData) == 32
    assert ctypes.sizeof(_lib.Result) == 88
    assert ctypes.sizeof(_lib.SpaiExtParams) == 64
    assert [f for f, _ in _lib.MethodParams._fields_] == [
        "max_iterations", "tolerance", "max_density", "enable_dropping", "num_threads", "use_initial_m",
        "use_preconditioner", "inner_iterations"]


C_PROGRAM = r"""
#include <stddef.h>
#include <stdio.h>
#include "methods_c_api.h"
#include "spai_ext.h"
int main(void) {
  void* fns[] = {(void*)&spai_global_cg, (void*)&spai_global_mr, (void*)&spai_global_lomr,
                 (void*)&inner_outer_mr, (void*)&inner_outer_lomr, (void*)&free_mem, (void*)&spai_run_ex,
                 (void*)&spai_pcg, (void*)&spai_matvec, (void*)&spai_session_create,
                 (void*)&spai_loss_plan_create, (void*)&spai_loss_backward};
  int ok = 1;
  for (size_t i = 0; i < sizeof(fns) / sizeof(fns[0]); ++i) ok &= fns[i] != NULL;
  printf("%d %zu %zu %zu %zu %zu\n", ok, sizeof(MatrixData), sizeof(MethodParams), sizeof(SingleIterationData),
         sizeof(Result), offsetof(Result, error_message));
  return 0;
}
"""


@pytest.mark.skipif(not LIB.exists() or shutil.which("gcc") is None, reason="needs gcc and the built library")
def test_headers_compile_as_c_and_link(tmp_path, spawn):
    src = tmp_path / "abi.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "abi"
    rc, _, err = spawn(["gcc", "-std=c99", "-Wall", "-Werror", "-I", ROOT / "include", src, "-o", exe,
                        "-L", PKG, "-lspai_hip", f"-Wl,-rpath,{PKG}"])
    assert rc == 0, err
    rc, out, err = spawn([exe])
    assert rc == 0, err
    assert out.split() == ["1", "48", "64", "32", "88", "80"]


def test_cython_mirror_imports():
    from parallellocallyoptimalapproximatesparsematrixinverses_amd import methods_cython
    from parallellocallyoptimalapproximatesparsematrixinverses_amd.methods_cython import backward_error
    assert hasattr(methods_cython.global_spai, "cg")
    assert hasattr(methods_cython.global_spai, "mr")
    assert hasattr(methods_cython.global_spai, "lomr")
    assert hasattr(methods_cython.inner_outer, "mr")
    assert hasattr(methods_cython.inner_outer, "lomr")
    for name in ("pcg", "matvec", "parallel_threads"):
        assert callable(getattr(backward_error, name))
