"""Device runs on the reference's other seven sample matrices against the
convergence tables the reference committed for them
(scripts/results/results_sample/dropping/max_iter_*/max_density_0.030/<name>/*-convergence.csv,
excerpted into tests/golden/).

The reference ran every method variant with 256 column batches, M0 = I,
Pr = diag(A)^-1, tolerance 1e-9 and dropping at density 0.03
(scripts/common.py:225-263, 380-409).  The device runs the same inputs with
batch-ordered reductions (the trace products summed in the reference's
column-batch order), so each ||R_i||_F must equal the table's value exactly:
the tables print 17 significant digits, which round-trip a double.

Row counts are per matrix and method family, so that each run takes a few
seconds: the inner-outer products M_outer R have near-dense columns once M
reaches its density cap (about a million expansion terms per column on
Poisson32k), and bundle1's and the random matrices' patterns fill quickly.
The CPU oracle is held to the same tables in tests/test_oracle_golden.py.
"""
import numpy as np
import pytest

import golden_data as gd

pytestmark = pytest.mark.gpu

# name -> (iterations of the Global methods, iterations of the inner-outer ones)
ROWS = {
    "Poisson32k": (12, 3),
    "wathen100": (12, 3),
    "4bw100eigs20k": (30, 30),
    "4bw100eigs20k2": (30, 30),
    "bundle1": (3, 2),
    "rand20k": (8, 3),
    "rand20k2": (6, 2),
}


@pytest.mark.parametrize("name", list(ROWS))
@pytest.mark.parametrize("variant", range(7))
def test_device_reproduces_sample_history(gpu_lib, name, variant):
    method, inner = gd.VARIANTS[variant]
    iters = ROWS[name][1 if method.startswith("io") else 0]
    print(f"case: {name} dropping d0.030 {method}/{inner} {iters} iterations", flush=True)
    A = gd.matrix(name)
    M0, Pr = gd.harness_inputs(A)
    g = gpu_lib.run(method, A, M0, Pr, iters, 1e-9, 0.03, True, 256, inner_iterations=inner, batch_ordered=True)
    ref = gd.convergence(name, "dropping", "0.030")[:, variant]
    h = g["history"]["residual_norm"]
    assert len(h) == iters + 1
    assert not np.isnan(ref[:len(h)]).any()
    np.testing.assert_array_equal(h, ref[:len(h)])
