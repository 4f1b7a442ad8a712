"""The C++ interface include/methods.hpp (reference include/methods.hpp:11-155)
through host programs built by tests/cpp/Makefile against libspai_hip.so:

  sparse_matrix_test  the Eigen-free SparseMatrix against dense brute force
                      (no device needed);
  methods_api_test    the reference's tests/test.cpp ported to this interface:
                      all five builders, batch-count independence, absolute
                      stop, dropping budget, density stop, reported residual,
                      and the reference's validation messages as exceptions.
"""
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "cpp"
LIB = HERE.parent.parent / "parallellocallyoptimalapproximatesparsematrixinverses_amd" / "libspai_hip.so"


def _run(spawn, name, timeout):
    """Builds the program if needed and runs it through the session's
    launcher process (this process never spawns children)."""
    exe = HERE / "_bin" / name
    if not exe.exists():
        if not LIB.exists():
            pytest.skip("libspai_hip.so is not built")
        rc, out, err = spawn(["make", "-s", "-C", HERE, f"_bin/{name}"], timeout=300)
        assert rc == 0, err[-2000:]
    print(f"case: {name}", flush=True)
    rc, out, err = spawn([exe], timeout=timeout)
    print(out[-4000:])
    assert rc == 0, out[-2000:] + err[-2000:]
    assert "all checks passed" in out


def test_sparse_matrix_host_class(spawn):
    _run(spawn, "sparse_matrix_test", 60)


@pytest.mark.gpu
def test_methods_api_port_of_reference_test(spawn):
    _run(spawn, "methods_api_test", 110)
