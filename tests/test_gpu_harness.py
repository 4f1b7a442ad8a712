"""Device harness (scripts/utils/backward_error.py mirror) against the oracle:
right-hand sides bit-identical to the reference recipe, PCG histories and
status strings as backward_error.pcg gives them."""
import numpy as np
import pytest

from parallellocallyoptimalapproximatesparsematrixinverses_amd import harness, problems

pytestmark = pytest.mark.gpu


def row_dot_matvec(A, x):
    """Sequential row dots in column order (methods_cython/backward_error.pyx:16-24)."""
    A = A.tocsr()
    y = np.zeros(A.shape[0])
    for i in range(A.shape[0]):
        s = 0.0
        for k in range(A.indptr[i], A.indptr[i + 1]):
            s += A.data[k] * x[A.indices[k]]
        y[i] = s
    return y


def test_random_rhs_matches_reference_recipe():
    A = problems.diffusion3d(6, seed=2)
    got = list(harness.random_rhs_realizations(A, seed=7, realizations=3))
    rng = np.random.default_rng(7)
    for b in got:
        np.testing.assert_array_equal(b, row_dot_matvec(A, rng.random(A.shape[0])))
    np.testing.assert_array_equal(harness.random_rhs(A, seed=7), got[0])


def test_prepared_solver_against_oracle(oracle_lib):
    A = problems.diffusion3d(10, seed=1)
    Pr = problems.diag_inverse(A)
    M = oracle_lib.run("lomr", A, Pr, None, 8, 0.0, 1.0, False, 1, fixed_pattern=True)["M"]
    system = harness.PreparedPCGSystem(A)
    solver = system.prepare_preconditioner(M)
    b = harness.random_rhs(A)
    for limit, expect in ((harness.DEFAULT_MAX_ITERATIONS, None), (3, "iteration limit reached"), (0, "converged")):
        res = solver.calculate(limit, b=b)
        ho, so, _ = oracle_lib.pcg(A.tocsr(), M.tocsr(), b, limit or A.shape[0], harness.TOLERANCE)
        assert res.status == harness.STATUS_STRINGS[so]
        if expect:
            assert res.status == expect
        assert abs(len(res.history) - len(ho)) <= 1
        k = min(len(res.history), len(ho)) - 1
        np.testing.assert_allclose(res.history[:k], ho[:k], rtol=1e-9)
        assert res.converged == (so == 1)
    # b = None draws the default right-hand side (seed 1)
    res = solver.calculate()
    assert res.solver == "PCG" and res.converged
    solver.close()


def test_calculate_backward_error_many_rhs(oracle_lib):
    A = problems.poisson2d(20)
    M = problems.diag_inverse(A)
    system = harness.PreparedPCGSystem(A)
    solver = system.prepare_preconditioner(M)
    for b in harness.random_rhs_realizations(A, seed=3, realizations=4):
        res = solver.calculate(200, b=b)
        ho, so, _ = oracle_lib.pcg(A.tocsr(), M.tocsr(), b, 200, harness.TOLERANCE)
        assert res.converged and so == 1
        assert abs(len(res.history) - len(ho)) <= 1
    solver.close()
    one = harness.calculate_backward_error(A, M, 200, b=b)
    assert one.converged


def test_torch_after_library_use(gpu_lib, spawn):
    """The package loads torch before libspai_hip.so, so torch still sees the
    device after the library has run (one HIP runtime in the process).  The
    child is started by the session's launcher process, not by this GPU test
    process (tests/spawner.py)."""
    import sys
    code = (
        "import numpy as np\n"
        "from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib, problems\n"
        "A = problems.poisson2d(6).tocsr()\n"
        "assert np.allclose(_lib.matvec(A, np.ones(A.shape[0])), A @ np.ones(A.shape[0]))\n"
        "import torch\n"
        "assert torch.cuda.is_available()\n"
        "print('ok')\n")
    from pathlib import Path
    root = Path(__file__).resolve().parent.parent
    print("case: child process, library then torch", flush=True)
    rc, out, err = spawn([sys.executable, "-c", code], cwd=root, timeout=240)
    assert rc == 0 and out.strip().endswith("ok"), err[-2000:]
