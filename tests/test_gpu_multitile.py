"""Fixed-pattern LOMR and MR at sizes where the fused pass walks several tiles
per wavefront, against the oracle (src/global_spai/locally_minimal_residual.cpp:54-118,
minimal_residual.cpp:31-93; PCG methods_cython/backward_error.pyx:69-144).

The fused per-iteration pass gives each wavefront tiles of 64/GL consecutive
columns (GL lanes per column) and walks them with the grid the device holds
resident at once.  A wavefront with more than one tile loads the next tile's
bounds and first step chunk while it finishes the current one; that path only
runs above (resident wavefronts x 64/GL) columns -- at most 65,536 columns
for GL = 8 on 256 CUs (8 wavefronts per SIMD).  The cases:

  diffusion3d(42)   n = 74,088,  7-point, GL = 8   (above the threshold at any occupancy)
  stencil27(24)     n = 13,824, 27-point, GL = 32
  diffusion3d(64)   n = 262,144, the benchmark workload (6-8 tiles per wave)

Each test first asserts, from the session's launch shape, that the pass it
runs really has more tiles than resident wavefronts.  Bars: FP64 M within
1e-12 relative Frobenius and histories within 1e-10 of the oracle; FP32 bulk
within 1e-4 of the FP64 oracle with equal iteration counts; PCG iteration
counts within +-1 of the oracle's PCG on the oracle's M.
"""
import numpy as np
import pytest

from helpers import assert_same_result, env_set, rel_fro
from parallellocallyoptimalapproximatesparsematrixinverses_amd import problems

pytestmark = pytest.mark.gpu

# the oracle's batch count; fixed-pattern results do not depend on it beyond
# the summation order of its trace products
_THREADS = 16

_CASES = {
    "diffusion42": lambda: problems.diffusion3d(42, seed=1),
    "stencil27_24": lambda: problems.stencil27(24, seed=4),
    "diffusion64": lambda: problems.diffusion3d(64, seed=1),
}
_cache = {}


def _problem(name):
    if name not in _cache:
        A = _CASES[name]()
        _cache[name] = (A, problems.diag_inverse(A))
    return _cache[name]


_oracle_runs = {}


def _oracle(oracle_lib, name, method, iters):
    key = (name, method, iters)
    if key not in _oracle_runs:
        A, Pr = _problem(name)
        _oracle_runs[key] = oracle_lib.run(method, A, Pr, None, iters, 0.0, 1.0, False, _THREADS,
                                           fixed_pattern=True)
    return _oracle_runs[key]


def _session_build(gpu_lib, name, method, iters, precision, layout="csr"):
    """Session and build; layout "csr" keeps the general fused pass (the
    7-point cases would otherwise take the diagonal pass for LOMR)."""
    A, Pr = _problem(name)
    with env_set("SPAI_FIXED_LAYOUT", layout):
        s = gpu_lib.Session(method, A, M0=Pr, max_iterations=iters, precision=precision)
    info = s.info()
    if layout == "csr":
        assert info["fused"] == 1, info
        waves = info["grid_blocks"] * info["waves_per_block"]
        assert info["tiles"] > waves, f"{name}: {info['tiles']} tiles on {waves} resident wavefronts"
    else:
        assert info["fused"] == 2 and info["diagonals_s2"] == 25, info
    it = s.build()
    return s, it, s.result(), info


@pytest.mark.parametrize("name", ["diffusion42", "stencil27_24", "diffusion64"])
@pytest.mark.parametrize("method", ["mr", "lomr"])
def test_multitile_fp64_matches_oracle(gpu_lib, oracle_lib, name, method):
    iters = 6
    print(f"case: multitile fp64 {method} {name}", flush=True)
    s, it, g, info = _session_build(gpu_lib, name, method, iters, "fp64")
    s.close()
    o = _oracle(oracle_lib, name, method, iters)
    assert it == o["iterations"] == iters
    assert_same_result(g, o, hist_rtol=1e-10, m_rtol=1e-12)


@pytest.mark.parametrize("name", ["diffusion42", "stencil27_24", "diffusion64"])
@pytest.mark.parametrize("method", ["mr", "lomr"])
def test_multitile_fp32_against_fp64_oracle(gpu_lib, oracle_lib, name, method):
    iters = 6
    print(f"case: multitile fp32 {method} {name}", flush=True)
    s, it, g, info = _session_build(gpu_lib, name, method, iters, "fp32")
    s.close()
    o = _oracle(oracle_lib, name, method, iters)
    assert it == o["iterations"] == iters
    np.testing.assert_array_equal(g["M"].indices, o["M"].indices)
    assert rel_fro(g["M"], o["M"]) <= 1e-4
    np.testing.assert_allclose(g["history"]["residual_norm"], o["history"]["residual_norm"], rtol=1e-4)


@pytest.mark.parametrize("layout", ["auto", "csr"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config2_full_lomr_and_pcg_against_oracle(gpu_lib, oracle_lib, precision, layout):
    """The benchmark step itself: 64^3 heterogeneous diffusion, 30 LOMR
    iterations on pattern(A) from M0 = diag(A)^-1, then PCG to 1e-8 with
    x_true ~ N(0,1) (seed 2), all on one session as the benchmark runs it;
    the diagonal pass (the default here) and the general fused pass."""
    name = "diffusion64"
    A, Pr = _problem(name)
    print(f"case: config2 full {precision} {layout}", flush=True)
    s, it, g, info = _session_build(gpu_lib, name, "lomr", 30, precision, layout)
    o = _oracle(oracle_lib, name, "lomr", 30)
    assert it == o["iterations"] == 30
    if precision == "fp64":
        assert_same_result(g, o, hist_rtol=1e-10, m_rtol=1e-12)
    else:
        assert rel_fro(g["M"], o["M"]) <= 1e-4
        np.testing.assert_allclose(g["history"]["residual_norm"], o["history"]["residual_norm"], rtol=1e-4)
    b = A @ np.random.default_rng(2).standard_normal(A.shape[0])
    s.set_rhs(b)
    st, cnt = s.pcg(5000, 1e-8)
    hist, x = s.pcg_output(cnt)
    s.close()
    ho, so, _ = oracle_lib.pcg(A.tocsr(), o["M"].tocsr(), b, 5000, 1e-8, threads=_THREADS)
    assert st == so == 1
    assert abs(cnt - len(ho)) <= 1, (cnt, len(ho))
    n = min(cnt, len(ho)) - 1
    dev = float(np.max(np.abs(hist[:n] / ho[:n] - 1.0)))
    print(f"pcg iterations device {cnt} oracle {len(ho)}, history max relative deviation {dev:.2e}", flush=True)
    assert dev <= (1e-3 if precision == "fp32" else 1e-6)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-8
