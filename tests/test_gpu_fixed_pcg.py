"""Fixed-pattern builders (FP64 and FP32-bulk) and the device PCG against the oracle."""
import numpy as np
import pytest
from scipy import sparse

from helpers import assert_same_result, rel_fro
from parallellocallyoptimalapproximatesparsematrixinverses_amd import problems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["cg", "mr", "lomr"])
def test_fixed_fp64_poisson(gpu_lib, oracle_lib, method):
    """BASELINE config 1 shape: 2D Poisson 32^2, M0 = diag(A)^-1, M on pattern(A)."""
    A = problems.poisson2d(32)
    Pr = problems.diag_inverse(A)
    print(f"case: fixed fp64 {method} poisson32", flush=True)
    g = gpu_lib.run(method, A, Pr, None, 20, 0.0, 1.0, False, 1, fixed_pattern=True)
    o = oracle_lib.run(method, A, Pr, None, 20, 0.0, 1.0, False, 1, fixed_pattern=True)
    assert_same_result(g, o, hist_rtol=1e-10, m_rtol=1e-12)


@pytest.mark.parametrize("method", ["cg", "mr", "lomr"])
def test_fixed_fp64_diffusion(gpu_lib, oracle_lib, method):
    A = problems.diffusion3d(10, seed=1)
    Pr = problems.diag_inverse(A)
    print(f"case: fixed fp64 {method} diffusion10", flush=True)
    g = gpu_lib.run(method, A, Pr, None, 15, 0.0, 1.0, False, 1, fixed_pattern=True)
    o = oracle_lib.run(method, A, Pr, None, 15, 0.0, 1.0, False, 1, fixed_pattern=True)
    assert_same_result(g, o, hist_rtol=1e-10, m_rtol=1e-12)


@pytest.mark.parametrize("method", ["io_mr", "io_lomr"])
@pytest.mark.parametrize("inner", [2, 5])
def test_fixed_inner_outer_fp64(gpu_lib, oracle_lib, method, inner):
    """Inner-outer on the fixed pattern (BASELINE config 4 form, reduced size):
    anisotropic 2D diffusion, M0 = diag(A)^-1."""
    A = problems.anisotropic2d(24)
    M0 = problems.diag_inverse(A)
    print(f"case: fixed fp64 {method} inner={inner}", flush=True)
    g = gpu_lib.run(method, A, M0, None, 12, 0.0, 1.0, False, 1, inner_iterations=inner, fixed_pattern=True)
    o = oracle_lib.run(method, A, M0, None, 12, 0.0, 1.0, False, 1, inner_iterations=inner, fixed_pattern=True)
    assert_same_result(g, o, hist_rtol=1e-10, m_rtol=1e-11)


SQRT_FLT_EPS = float(np.sqrt(np.finfo(np.float32).eps))


@pytest.mark.parametrize("method", ["io_mr", "io_lomr"])
@pytest.mark.parametrize("size,iters,inner", [(24, 12, 5), (24, 20, 2), (64, 20, 5)])
def test_fixed_inner_outer_fp32(gpu_lib, oracle_lib, method, size, iters, inner):
    """FP32-bulk inner-outer against the FP64 oracle.  Inner-outer LOMR in FP32
    solves each column's 2x2 Gram with the rank threshold sqrt(FLT_EPSILON)
    in place of the reference's DBL_EPSILON (FP32 rounding of AZ and AQ
    makes smaller pivots noise; tests/test_oracle_ops.py pins what that rule
    computes).  The oracle run with the same threshold (rank_eps) is the FP64
    restatement of that algorithm: the FP32 result must equal it to FP32
    rounding.  Against the reference's rule it differs by the columns solved
    as rank 1 (M by a few percent, ||I - AM||_F by about 1e-3), and its
    residual still decreases monotonically."""
    A = problems.anisotropic2d(size)
    M0 = problems.diag_inverse(A)
    print(f"case: fixed fp32 {method} aniso{size} {iters} iterations inner={inner}", flush=True)
    g = gpu_lib.run(method, A, M0, None, iters, 0.0, 1.0, False, 1, inner_iterations=inner, fixed_pattern=True,
                    precision="fp32")
    o = oracle_lib.run(method, A, M0, None, iters, 0.0, 1.0, False, 1, inner_iterations=inner, fixed_pattern=True)
    assert g["iterations"] == o["iterations"] == iters
    h32, h64 = g["history"]["residual_norm"], o["history"]["residual_norm"]
    if method == "io_mr":
        assert rel_fro(g["M"], o["M"]) < 1e-6
        np.testing.assert_allclose(h32, h64, rtol=1e-6)
    else:
        o32 = oracle_lib.run(method, A, M0, None, iters, 0.0, 1.0, False, 1, inner_iterations=inner,
                             fixed_pattern=True, rank_eps=SQRT_FLT_EPS)
        assert rel_fro(g["M"], o32["M"]) < 1e-6
        np.testing.assert_allclose(h32, o32["history"]["residual_norm"], rtol=1e-6)
        # against the reference's rank rule: the first LOMR inner step (a
        # plain MR step) agrees, later ones differ by the rank-1 columns
        np.testing.assert_allclose(h32[:2], h64[:2], rtol=1e-6)
        np.testing.assert_allclose(h32, h64, rtol=2e-3)
        assert rel_fro(g["M"], o["M"]) < 5e-2
        assert np.all(np.diff(h32) <= 1e-6 * h32[:-1])


@pytest.mark.parametrize("method", ["cg", "mr", "lomr"])
def test_fixed_fp32_against_fp64_oracle(gpu_lib, oracle_lib, method):
    A = problems.diffusion3d(10, seed=1)
    Pr = problems.diag_inverse(A)
    print(f"case: fixed fp32 {method} diffusion10", flush=True)
    g = gpu_lib.run(method, A, Pr, None, 15, 0.0, 1.0, False, 1, fixed_pattern=True, precision="fp32")
    o = oracle_lib.run(method, A, Pr, None, 15, 0.0, 1.0, False, 1, fixed_pattern=True)
    assert g["iterations"] == o["iterations"]
    np.testing.assert_array_equal(g["M"].indices, o["M"].indices)
    # FP32 bulk: ||M32 - M64||_F / ||M64||_F at single-precision level
    assert rel_fro(g["M"], o["M"]) < 1e-4
    np.testing.assert_allclose(g["history"]["residual_norm"], o["history"]["residual_norm"], rtol=1e-4)


def _pcg_pair(gpu_lib, oracle_lib, A, M, b, limit, tol, m_fp32=False):
    hg, sg, xg = gpu_lib.pcg(A, M, b, limit, tol, m_fp32=m_fp32)
    ho, so, xo = oracle_lib.pcg(A, M, b, limit, tol, threads=1)
    return (hg, sg, xg), (ho, so, xo)


def test_pcg_poisson(gpu_lib, oracle_lib):
    A = problems.poisson2d(32)
    Pr = problems.diag_inverse(A)
    M = oracle_lib.run("mr", A, Pr, None, 20, 0.0, 1.0, False, 1, fixed_pattern=True)["M"]
    x_true = np.random.default_rng(0).uniform(-1, 1, A.shape[0])
    b = A @ x_true
    print("case: pcg poisson32", flush=True)
    (hg, sg, xg), (ho, so, xo) = _pcg_pair(gpu_lib, oracle_lib, A.tocsr(), M.tocsr(), b, 1000, 1e-8)
    assert sg == so == 1
    assert abs(len(hg) - len(ho)) <= 1
    n = min(len(hg), len(ho)) - 1
    np.testing.assert_allclose(hg[:n], ho[:n], rtol=1e-6)
    assert hg[-1] <= 1e-8
    np.testing.assert_allclose(xg, x_true, rtol=0, atol=1e-5)


def test_pcg_limit_and_identity_preconditioner(gpu_lib, oracle_lib):
    A = problems.diffusion3d(8, seed=2).tocsr()
    I = sparse.identity(A.shape[0], format="csr")
    b = A @ np.random.default_rng(5).standard_normal(A.shape[0])
    print("case: pcg limit", flush=True)
    (hg, sg, _), (ho, so, _) = _pcg_pair(gpu_lib, oracle_lib, A, I, b, 10, 1e-12)
    assert sg == so == 0
    assert len(hg) == len(ho) == 10
    np.testing.assert_allclose(hg, ho, rtol=1e-9)


def test_pcg_fp32_preconditioner(gpu_lib, oracle_lib):
    A = problems.diffusion3d(12, seed=1)
    Pr = problems.diag_inverse(A)
    M = oracle_lib.run("lomr", A, Pr, None, 10, 0.0, 1.0, False, 1, fixed_pattern=True)["M"]
    b = A @ np.random.default_rng(2).standard_normal(A.shape[0])
    print("case: pcg fp32 M", flush=True)
    (hg, sg, _), (ho, so, _) = _pcg_pair(gpu_lib, oracle_lib, A.tocsr(), M.tocsr(), b, 2000, 1e-8, m_fp32=True)
    assert sg == so == 1
    assert abs(len(hg) - len(ho)) <= 2


def test_matvec(gpu_lib):
    A = problems.diffusion3d(9, seed=3).tocsr()
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    y = gpu_lib.matvec(A, x)
    np.testing.assert_allclose(y, A @ x, rtol=1e-13, atol=1e-13)


def test_session_build_and_solve(gpu_lib, oracle_lib):
    A = problems.diffusion3d(16, seed=1)
    Pr = problems.diag_inverse(A)
    b = A @ np.random.default_rng(2).standard_normal(A.shape[0])
    print("case: session lomr diffusion16", flush=True)
    s = gpu_lib.Session("lomr", A, M0=Pr, max_iterations=12)
    it = s.build()
    res = s.result()
    o = oracle_lib.run("lomr", A, Pr, None, 12, 0.0, 1.0, False, 1, fixed_pattern=True)
    assert it == o["iterations"] == 12
    assert rel_fro(res["M"], o["M"]) < 1e-12
    s.set_rhs(b)
    st, cnt = s.pcg(5000, 1e-8)
    hist, x = s.pcg_output(cnt)
    ho, so, _ = oracle_lib.pcg(A.tocsr(), o["M"].tocsr(), b, 5000, 1e-8)
    assert st == so == 1
    assert abs(cnt - len(ho)) <= 1
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-8
    # rebuild reuses the session: M's values reach the PCG through the
    # transposition kept from the first build, so the solve repeats exactly
    assert s.build() == 12
    st2, cnt2 = s.pcg(5000, 1e-8)
    hist2, x2 = s.pcg_output(cnt2)
    assert st2 == st and cnt2 == cnt
    np.testing.assert_array_equal(hist2, hist)
    np.testing.assert_array_equal(x2, x)
    s.close()


def test_session_fp32_rebuild_repeats_solve(gpu_lib):
    """FP32-bulk session: the PCG after a rebuild (values moved by the kept
    transposition, FP32 copy included) repeats the first solve exactly, and
    matches a fresh session's."""
    A = problems.diffusion3d(16, seed=3)
    Pr = problems.diag_inverse(A)
    b = A @ np.random.default_rng(4).standard_normal(A.shape[0])
    print("case: session fp32 rebuild diffusion16", flush=True)
    runs = []
    for builds in (1, 2):
        s = gpu_lib.Session("lomr", A, M0=Pr, max_iterations=9, precision="fp32")
        for _ in range(builds):
            assert s.build() == 9
        s.set_rhs(b)
        st, cnt = s.pcg(5000, 1e-8)
        runs.append((st, cnt) + s.pcg_output(cnt))
        s.close()
    (st1, c1, h1, x1), (st2, c2, h2, x2) = runs
    assert st1 == st2 == 1 and c1 == c2
    np.testing.assert_array_equal(h1, h2)
    np.testing.assert_array_equal(x1, x2)


def test_session_kernel_timing(gpu_lib):
    """HIP-event timing of the fused LOMR/MR pass: one average per build over
    the iterations that ran; CG builds report none."""
    A = problems.diffusion3d(12, seed=1)
    Pr = problems.diag_inverse(A)
    s = gpu_lib.Session("lomr", A, M0=Pr, max_iterations=7)
    s.build()
    assert s.build_kernel_us() == (-1.0, 0)
    s.set_kernel_timing(True)
    assert s.build() == 7
    us, launches = s.build_kernel_us()
    assert launches == 7 and 0.0 < us < 1e5
    s.set_kernel_timing(False)
    s.build()
    assert s.build_kernel_us()[1] == 0
    s.close()
    m = gpu_lib.Session("cg", A, M0=Pr, max_iterations=3)
    m.set_kernel_timing(True)
    m.build()
    assert m.build_kernel_us() == (-1.0, 0)
    m.close()
    m = gpu_lib.Session("mr", A, M0=Pr, max_iterations=3)
    m.set_kernel_timing(True)
    assert m.build() == 3
    assert m.build_kernel_us()[1] == 3
    m.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_session_build_graph_replays(gpu_lib, oracle_lib, precision):
    """Untimed builds replay one recorded graph: every replay, with eager
    (timed) builds and solves in between, gives the same M and history as the
    eager build, and the first matches the oracle."""
    A = problems.diffusion3d(10, seed=2)
    Pr = problems.diag_inverse(A)
    b = A @ np.random.default_rng(3).standard_normal(A.shape[0])
    s = gpu_lib.Session("lomr", A, M0=Pr, max_iterations=9, precision=precision)
    s.set_rhs(b)
    s.set_kernel_timing(True)
    assert s.build() == 9
    eager = s.result()
    s.set_kernel_timing(False)
    o = oracle_lib.run("lomr", A, Pr, None, 9, 0.0, 1.0, False, 1, fixed_pattern=True)
    assert rel_fro(eager["M"], o["M"]) < (1e-12 if precision == "fp64" else 1e-4)
    for k in range(4):
        assert s.build() == 9
        r = s.result()
        np.testing.assert_array_equal(r["M"].data, eager["M"].data)
        np.testing.assert_array_equal(r["history"]["residual_norm"], eager["history"]["residual_norm"])
        st, cnt = s.pcg(2000, 1e-8)
        assert st == 1
        if k == 1:
            s.set_kernel_timing(True)
            s.build()
            s.set_kernel_timing(False)
    s.close()


@pytest.mark.parametrize("n,per_row", [(3000, 3), (800, 10), (1500, 6)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fixed_lomr_wide_patterns(gpu_lib, oracle_lib, n, per_row, precision):
    """Fixed-pattern LOMR on B^T B + sigma I patterns: 32 lanes per column
    (3000/3), columns of A longer than 64 with tiles beyond the register-held
    entries (800/10), and S2 columns over 1024 entries, which take the
    unfused passes (1500/6)."""
    A, _, _ = problems.random_gram(n, per_row=per_row, seed=3)
    Pr = problems.diag_inverse(A)
    print(f"case: lomr wide n={n} per_row={per_row} {precision}", flush=True)
    g = gpu_lib.run("lomr", A, Pr, None, 8, 0.0, 1.0, False, 1, fixed_pattern=True, precision=precision)
    o = oracle_lib.run("lomr", A, Pr, None, 8, 0.0, 1.0, False, 1, fixed_pattern=True)
    assert g["iterations"] == o["iterations"] == 8
    if precision == "fp64":
        assert_same_result(g, o, hist_rtol=1e-10, m_rtol=1e-12)
    else:
        assert rel_fro(g["M"], o["M"]) < 1e-4
        np.testing.assert_allclose(g["history"]["residual_norm"], o["history"]["residual_norm"], rtol=1e-4)


def _nine_point(m):
    """2D 9-point stencil (kron of 1D tridiagonals plus a diagonal shift): 9
    entries per interior column of A and of S, so a column has more steps
    than product 2 takes from registers."""
    T = sparse.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    E = sparse.diags([1.0, 1.0, 1.0], [-1, 0, 1], shape=(m, m))
    A = (sparse.kron(T, E) + sparse.kron(E, T) + 0.5 * sparse.identity(m * m)).tocsc()
    A.sort_indices()
    return A


@pytest.mark.parametrize("method", ["mr", "lomr"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fixed_nine_point_product2_paths(gpu_lib, oracle_lib, method, precision):
    """Columns of 9 S entries: FP32 takes product 2's terms from the LDS stash,
    FP64 reloads them (the stash does not fit its LDS budget).  This stencil is
    indefinite, and its LOMR step sizes are sensitive to the summation order
    of the trace products (the device's differs from the oracle's): M differs
    by 8e-12 relative from iteration 4 on and the difference does not grow, so
    the FP64 LOMR bar on M is 1e-10 here (histories stay within 1e-12)."""
    A = _nine_point(40)
    Pr = problems.diag_inverse(A)
    print(f"case: fixed {precision} {method} nine-point 40^2", flush=True)
    g = gpu_lib.run(method, A, Pr, None, 12, 0.0, 1.0, False, 1, fixed_pattern=True, precision=precision)
    o = oracle_lib.run(method, A, Pr, None, 12, 0.0, 1.0, False, 1, fixed_pattern=True)
    assert g["iterations"] == o["iterations"] == 12
    if precision == "fp64":
        assert_same_result(g, o, hist_rtol=1e-12, m_rtol=1e-10 if method == "lomr" else 1e-12)
    else:
        assert rel_fro(g["M"], o["M"]) < 1e-4
        np.testing.assert_allclose(g["history"]["residual_norm"], o["history"]["residual_norm"], rtol=1e-4)


@pytest.mark.parametrize("m_fp32", [False, True])
def test_prepared_pcg_smaller_limit_after_failed_refreshes(gpu_lib, oracle_lib, m_fp32):
    """One prepared system, two solves: the first to a limit of 2000, the
    second to 300.  The tolerance 1e-17 lies below the attainable accuracy, so
    the recursive residual keeps crossing it while the refreshed true residual
    stays above it, and the solve goes on with beta = 0 after each refresh
    (backward_error.pyx:101-133).  Both solves must run to their own limit
    (status 0, history of exactly `limit` entries) as the oracle's do; the
    second solve must not stop early because the first ran longer."""
    from parallellocallyoptimalapproximatesparsematrixinverses_amd import harness
    A = problems.diffusion3d(12, seed=1)
    Pr = problems.diag_inverse(A)
    M = oracle_lib.run("lomr", A, Pr, None, 10, 0.0, 1.0, False, 1, fixed_pattern=True)["M"].tocsr()
    b = A @ np.random.default_rng(2).standard_normal(A.shape[0])
    system = harness._DeviceSystem(A.tocsr(), M, m_fp32=m_fp32)
    try:
        for limit in (2000, 300, 2000):
            print(f"case: prepared pcg limit {limit} m_fp32={m_fp32}", flush=True)
            hg, sg, _ = system.solve(b, limit, 1e-17)
            ho, so, _ = oracle_lib.pcg(A.tocsr(), M, b, limit, 1e-17)
            assert sg == so == 0
            assert len(hg) == len(ho) == limit
            np.testing.assert_allclose(hg[:30], ho[:30], rtol=1e-4 if m_fp32 else 1e-8)
            assert np.all(hg[100:] < 1e-14)
    finally:
        system.close()
