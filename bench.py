"""Benchmark: Global SPAI LOMR + PCG time-to-1e-8 on the 64^3 heterogeneous-diffusion problem.

Workload (one "step" = one full pass, operands already resident in HBM):
  A: 3D 7-point variable-coefficient diffusion on 64^3 (n=262,144, nnz=1,810,432),
     face coefficients = harmonic mean of exp(g), g ~ N(0,1), seed 1 (synthetic).
  build: Global SPAI LOMR, M on pattern(A), M0 = diag(A)^-1, 30 iterations.
  solve: PCG with b = A x_true, x_true ~ N(0,1) seed 2, x0 = 0, tolerance 1e-8.
The timed form is FP32-bulk (M, directions, residual and products in FP32;
step sizes, trace products, ||R||_F and the whole PCG recurrence in FP64); the
plain FP64 form is timed the same way and reported beside it.

roofline: the dominant kernel of the FP32-bulk step, the fused LOMR pass (one
iteration), timed with HIP events on the session's stream over one extra build
after the timed steps; algorithmic bytes per launch are stated in main().
roofline_pcg_spmv: the PCG's A-SpMV, timed by a standalone replay.
spgemm: the reference-semantics two-phase SpGEMM C = A A (the product A M
  with M on A's pattern), GB/s of algorithmic bytes and GFLOP/s.
cold: the same step from nothing on the device, once per form: session
  creation (upload of A and M0, S/S2 symbolic phase, slot plan, PCG operands),
  right-hand side upload, the first build (graph recording included), the
  solve (a fresh iteration schedule) and the download of M and x; median of
  five.  `value` stays the warm step (operands resident).
achievable_bandwidth: streaming read / copy / triad kernels of this library
  on 2 GiB arrays; every roofline object also states its fraction of the best
  of copy and triad.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SPAI+PCG time-to-1e-8 (s), FP32-bulk vs FP64 on GPU; SpGEMM HBM GB/s"
HBM_PEAK_GBS = 8000.0
ITERATIONS = 30
TOL = 1e-8
PCG_LIMIT = 20000


def problem():
    from parallellocallyoptimalapproximatesparsematrixinverses_amd import problems

    A = problems.diffusion3d(64, seed=1)
    Pr = problems.diag_inverse(A)
    x_true = np.random.default_rng(2).standard_normal(A.shape[0])
    return A, Pr, A @ x_true


def time_form(lib, A, M0, b, precision, steps, warmup):
    s = lib.Session("lomr", A, M0=M0, max_iterations=ITERATIONS, precision=precision)
    s.set_rhs(b)
    out = {}
    for _ in range(warmup):
        s.build()
        s.pcg(PCG_LIMIT, TOL)
    t_build = t_pcg = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        it = s.build()
        c = time.perf_counter()
        st, cnt = s.pcg(PCG_LIMIT, TOL)
        d = time.perf_counter()
        t_build += c - a
        t_pcg += d - c
    total = time.perf_counter() - t0
    hist, x = s.pcg_output(cnt)
    out.update(seconds=total / steps, build_s=t_build / steps, pcg_s=t_pcg / steps, spai_iterations=it,
               pcg_iterations=cnt - 1, status=st,
               true_rel_residual=float(np.linalg.norm(b - A @ x) / np.linalg.norm(b)))
    out["M"] = s.result()["M"]
    out["spmv_a_us"] = s.spmv_a_us()
    out["pcg_layouts"] = dict(zip(("A", "M"), s.pcg_layouts()))
    out["launch_shape"] = s.info()
    # kernel timing: the same build once more with HIP events recorded on the
    # session's stream around every per-iteration pass (outside the timed
    # steps, whose pacing the events would change)
    s.set_kernel_timing(True)
    s.build()
    us, launches = s.build_kernel_us()
    out["lomr_pass_us"] = us if us > 0 else None
    out["lomr_passes_timed"] = launches
    s.close()
    return out


def cold_step(lib, A, M0, b, precision):
    """One step from nothing on the device, through the C ABI: create the
    session (uploads, symbolic phase, plan), upload b, build, solve, and
    download M (C-ABI Result arrays) and x.  A and M0 are canonical CSC with
    int64 indices, which the calls read in place, as a C caller hands over its
    arrays.  Returns (seconds, pcg iterations)."""
    t = time.perf_counter()
    s = lib.Session("lomr", A, M0=M0, max_iterations=ITERATIONS, precision=precision)
    s.set_rhs(b)
    s.build()
    st, cnt = s.pcg(PCG_LIMIT, TOL)
    hist, x = s.pcg_output(cnt)
    s.download_result()
    s.close()
    return time.perf_counter() - t, cnt - 1


def fused_pass_bytes(n, nnz_a, nnz_s, nnz_s2, sv, method):
    """Algorithmic bytes of one fused iteration pass (int32 indices and
    pointers, s_v bytes per value): A once ((n+1) 4 + nnz(A) (4 + s_v)), the
    patterns of S and S2 once ((n+1) 4 + nnz 4 each); LOMR reads and writes R
    on S2 and M, Z, P on S (2 nnz(S2) s_v + 6 nnz(S) s_v); MR writes R
    (recomputed as I - A M) and reads and writes M and Z (nnz(S2) s_v +
    4 nnz(S) s_v)."""
    base = (n + 1) * 4 + nnz_a * (4 + sv) + (n + 1) * 4 + nnz_s * 4 + (n + 1) * 4 + nnz_s2 * 4
    if method == "lomr":
        return base + 2 * nnz_s2 * sv + 6 * nnz_s * sv
    return base + nnz_s2 * sv + 4 * nnz_s * sv


def dia_pass_bytes(n, nd_a, nd_s, sv, method="lomr"):
    """Algorithmic bytes of one LOMR or MR iteration on the diagonal pass (s_v
    bytes per value, every slot of the layout counted): A's values once (nd_a
    n s_v), S's column masks (4 n), M, Z and (LOMR) P read and written (6 or
    4 nd_s n s_v) and Pr once (n s_v); R' is formed from M' in registers and
    never stored."""
    arrays = 6 if method == "lomr" else 4
    return nd_a * n * sv + 4 * n + arrays * nd_s * n * sv + n * sv


def lomr_pass_bytes(n, nnz_a, info, sv, method):
    """Algorithmic bytes of the per-iteration pass the session's build runs
    (info = Session.info()): the diagonal pass or the general fused pass."""
    if info["fused"] == 2:
        return dia_pass_bytes(n, info["diagonals_a"], info["diagonals_s"], sv, method)
    return fused_pass_bytes(n, nnz_a, info["nnz_s"], info["nnz_s2"], sv, method)


def lomr_pass_kernel(info, method, precision):
    """Name of the per-iteration pass kernel of a session's build."""
    if info["fused"] == 2:
        shape = {3: "1-D Laplacian", 5: "2-D Laplacian", 7: "3-D Laplacian", 9: "2-D box"}.get(info["diagonals_s"], "?")
        return (f"DiaPass<{precision}, {shape}> (one {method.upper()} iteration on the diagonal layout: "
                f"{info['diagonals_s']} diagonals in A and S, R' = I - A M' formed in registers)")
    return "FusedPass<%s, %s> (one fused %s iteration)" % (precision, method.upper(), method.upper())


def n_diagonals(A):
    """Distinct diagonals (col - row) holding entries of A."""
    c = A.tocoo()
    return int(np.unique(c.col.astype(np.int64) - c.row).size)


def matrix_bytes(n, nnz, v_bytes, nd=None):
    """Bytes of one pass over a matrix in the PCG's layout: diagonals (nd
    diagonals of n values and a 4-byte row mask) or CSR/SELL (int32 offsets
    and indices)."""
    if nd is not None:
        return nd * n * v_bytes + 4 * n
    return (n + 1) * 4 + nnz * (4 + v_bytes)


def pcg_iteration_bytes(n, nnz_a, nnz_m, m_bytes, nd_a=None, nd_m=None):
    """Algorithmic bytes of one PCG iteration: one pass over A (FP64 values)
    and one over M (m_bytes values), in the layouts the solve reads (nd_a /
    nd_m diagonals, or CSR when None), and the FP64 vectors -- p = z + beta p
    with A p reads z, p and writes p, Ap; the update reads x, p, r, Ap and
    writes x, r, z -- 11 n doubles."""
    return matrix_bytes(n, nnz_a, 8, nd_a) + matrix_bytes(n, nnz_m, m_bytes, nd_m) + 11 * 8 * n


def pcg_layout_diagonals(A, layouts):
    """(nd_a, nd_m) for pcg_iteration_bytes from a session's PCG layouts; M
    lies on pattern(A) | I here, which is pattern(A) for these matrices."""
    nd = n_diagonals(A)
    return (nd if layouts[0] == "dia" else None), (nd if layouts[1] == "dia" else None)


def roof(bytes_, seconds, achievable=None):
    """GB/s of algorithmic bytes and the fractions of HBM peak (and of the
    in-repo measured streaming bandwidth, when given)."""
    gbs = bytes_ / seconds / 1e9
    out = {"algorithmic_bytes": int(bytes_), "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if achievable:
        out["frac_of_achievable"] = round(gbs / achievable, 4)
    return out


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Host threads this process may use, capped at 16 (the CPU share of one
    GPU on the measurement hosts)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(A, M0, b, threads, full_runs=3):
    """Oracle (C++/OpenMP restatement of the reference) on the host cores.
    At `threads` threads: the whole workload (FP64 fixed-pattern LOMR, 30
    iterations, then the PCG on its M), median of `full_runs` runs.  At one
    thread, a bounded sample: LOMR at 2 and 6 iterations (slope extrapolated
    to 30) plus the full PCG."""
    from oracle import oracle

    runs = []
    for _ in range(full_runs):
        t = time.perf_counter()
        r = oracle.run("lomr", A, M0, None, ITERATIONS, 0.0, 1.0, False, threads, fixed_pattern=True,
                       omp_threads=threads)
        oracle.pcg(A.tocsr(), r["M"].tocsr(), b, PCG_LIMIT, TOL, threads=threads)
        runs.append(time.perf_counter() - t)
    t = time.perf_counter()
    oracle.run("lomr", A, M0, None, 2, 0.0, 1.0, False, 1, fixed_pattern=True, omp_threads=1)
    t2 = time.perf_counter() - t
    t = time.perf_counter()
    r6 = oracle.run("lomr", A, M0, None, 6, 0.0, 1.0, False, 1, fixed_pattern=True, omp_threads=1)
    t6 = time.perf_counter() - t
    t = time.perf_counter()
    oracle.pcg(A.tocsr(), r6["M"].tocsr(), b, PCG_LIMIT, TOL, threads=1)
    tp = time.perf_counter() - t
    one = t2 + (t6 - t2) / 4.0 * (ITERATIONS - 2) + tp
    return {"value": round(float(np.median(runs)), 4), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"the whole workload (oracle FP64 LOMR {ITERATIONS} iterations + PCG to {TOL:g}), median of "
                      f"{full_runs} runs on {threads} OpenMP threads",
            "runs_s": [round(x, 4) for x in runs], "cpu_model": cpu_model(),
            "one_thread": {"value": round(one, 4), "unit": "s", "cores": 1,
                           "sample": f"LOMR at 2 and 6 iterations ({t2:.2f} s, {t6:.2f} s; slope extrapolated to "
                                     f"{ITERATIONS}) + full PCG ({tp:.2f} s)"}}


def calibrate_copy(mib):
    """Plain streaming kernel of known size: reads and writes mib MiB once per
    launch.  Its FETCH_SIZE / WRITE_SIZE counters give the factor between the
    counters and HBM bytes."""
    import torch

    x = torch.ones(mib * 2 ** 17, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "torch.mul fp64", "bytes_read": x.numel() * 8, "bytes_written": y.numel() * 8,
                      "launches": 5}))


def spgemm_only(repeats):
    """Only the standalone SpGEMM of the headline line (C = A A on the 64^3
    diffusion matrix, symbolic and numeric phases, operands resident):
    `repeats` products after one warm-up, for kernel traces of that product
    alone."""
    from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib

    A, _, _ = problem()
    _lib.spgemm(A, A, repeats=1)
    C, s = _lib.spgemm(A, A, repeats=repeats)
    print(json.dumps({"spgemm": "C = A A, diffusion3d 64^3", "repeats": repeats, "seconds_per_product": s,
                      "nnz_C": int(C.nnz)}))


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def _fused_kernel_roof(s, method, n, nnz_a, sv, achievable):
    """The session's per-iteration fused pass: average HIP-event time over
    one more build, and its algorithmic bytes (lomr_pass_bytes)."""
    info = s.info()
    s.set_kernel_timing(True)
    s.build()
    us, launches = s.build_kernel_us()
    s.set_kernel_timing(False)
    if not info["fused"] or us <= 0:
        return None
    out = {"kernel": lomr_pass_kernel(info, method, "float" if sv == 4 else "double"), "avg_launch_us": round(us, 3),
           "launches_timed": launches}
    out.update(roof(lomr_pass_bytes(n, nnz_a, info, sv, method), us * 1e-6, achievable))
    return out


def run_configs(with_cpu, reps=5, only=None):
    """One JSON line per configuration besides the headline one: device times
    of the build and the solves (operands resident; medians of `reps` runs
    after one warm-up), the dominant kernels' GB/s of algorithmic bytes with
    their fractions of HBM peak and of the measured streaming bandwidth, and,
    with with_cpu, the oracle on a bounded sample."""
    from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib, problems
    from parallellocallyoptimalapproximatesparsematrixinverses_amd.autograd import ResidualLoss

    threads = host_threads()
    bw = _lib.stream_bandwidth(2 << 30, 20)
    achievable = max(bw["copy"], bw["triad"])
    print(json.dumps({"achievable_bandwidth": {"copy_GBs": round(bw["copy"], 1), "triad_GBs": round(bw["triad"], 1),
                                               "read_GBs": round(bw["read"], 1),
                                               "arrays": "FP64, 2 GiB each, median of 20 launches"}}), flush=True)
    want = set(only) if only else {1, 3, 4, 5}

    if 1 in want:
        # config 1: 2D Poisson 32^2, MR on pattern(A), 20 iterations, PCG to 1e-8
        A = problems.poisson2d(32)
        M0 = problems.diag_inverse(A)
        b = A @ np.random.default_rng(0).uniform(-1, 1, A.shape[0])
        s = _lib.Session("mr", A, M0=M0, max_iterations=20)
        s.set_rhs(b)
        s.build()
        s.pcg(1000, 1e-8)
        tb = _median_time(s.build, reps)
        tp = _median_time(lambda: s.pcg(1000, 1e-8), reps)
        st, cnt = s.pcg(1000, 1e-8)
        n = A.shape[0]
        fused = _fused_kernel_roof(s, "mr", n, A.nnz, 8, achievable)
        nnz_m = s.info()["nnz_s"]
        layouts = s.pcg_layouts()
        s.close()
        pcg_roof = {"kernel": f"PCG iteration (two kernels, A and M in layouts {layouts}; solve time / iterations)",
                    "avg_iteration_us": round(tp / max(cnt - 1, 1) * 1e6, 3)}
        pcg_roof.update(roof(pcg_iteration_bytes(n, A.nnz, nnz_m, 8, *pcg_layout_diagonals(A, layouts)),
                             tp / max(cnt - 1, 1), achievable))
        line = {"config": "poisson2d 32^2, MR fixed pattern 20 its + PCG 1e-8", "build_s": tb, "pcg_s": tp,
                "pcg_iterations": cnt - 1, "status": st, "roofline_build_pass": fused, "roofline_pcg": pcg_roof,
                "note": "n = 1,024: every kernel is launch-bound, far from any bandwidth bound"}
        if with_cpu:
            from oracle import oracle
            t = time.perf_counter()
            r = oracle.run("mr", A, M0, None, 20, 0.0, 1.0, False, threads, fixed_pattern=True, omp_threads=threads)
            oracle.pcg(A.tocsr(), r["M"].tocsr(), b, 1000, 1e-8, threads=threads)
            line["cpu_baseline"] = {"value": time.perf_counter() - t, "unit": "s", "cores": threads, "kind": "port",
                                    "sample": "full workload"}
        print(json.dumps(line), flush=True)

    if 3 in want:
        # config 3: A = B^T B + sigma I, n = 100,000, CG with per-column top-16 dropping, 25 iterations
        A, _, _ = problems.random_gram(100_000, seed=3, power_iterations=50)
        Pr = problems.diag_inverse(A)
        r = _lib.run("cg", A, None, Pr, 25, 0.0, 0.5, True, 256, column_top_k=16)
        dev = []
        for _ in range(3):
            r = _lib.run("cg", A, None, Pr, 25, 0.0, 0.5, True, 256, column_top_k=16)
            dev.append(r["device_seconds"])
        # the dynamic build's dominant kernel: the two-phase SpGEMM, here A M
        # with the final M (FP64; 3 (n+1) 4 + 12 (nnz(A) + nnz(M) + nnz(C)))
        n = A.shape[0]
        C, sg_s = _lib.spgemm(A, r["M"], repeats=10)
        sg = {"kernel": "SpGemm C = A M (final M; symbolic + numeric phases), FP64", "seconds": round(sg_s, 7),
              "nnz_C": int(C.nnz)}
        sg.update(roof(3 * (n + 1) * 4 + 12 * (A.nnz + r["M"].nnz + C.nnz), sg_s, achievable))
        line = {"config": "random gram n=100000 (nnz %d), CG top-16 dropping 25 its" % A.nnz,
                "build_s_including_transfers": float(np.median(dev)), "nnz_M": int(r["M"].nnz), "spgemm": sg}
        if with_cpu:
            from oracle import oracle
            t = time.perf_counter()
            oracle.run("cg", A, None, Pr, 2, 0.0, 0.5, True, 256, column_top_k=16, omp_threads=threads)
            t2 = time.perf_counter() - t
            line["cpu_baseline"] = {"value": t2 / 2 * 25, "unit": "s", "cores": threads, "kind": "port",
                                    "sample": "2 of 25 iterations (%.2f s), scaled to 25" % t2}
        print(json.dumps(line), flush=True)

    if 4 in want:
        # config 4: 1024^2 anisotropic, inner-outer LOMR 5 x 20 FP32 bulk, 16 right-hand sides
        A = problems.anisotropic2d(1024)
        M0 = problems.diag_inverse(A)
        s = _lib.Session("io_lomr", A, M0=M0, max_iterations=100, precision="fp32", inner_iterations=5)
        s.build()
        tb = _median_time(s.build, 3)
        counts = []
        t = time.perf_counter()
        for seed in range(16):
            bb = A @ np.random.default_rng(seed).standard_normal(A.shape[0])
            s.set_rhs(bb)
            st, cnt = s.pcg(50000, 1e-8)
            counts.append(cnt - 1)
        tp = time.perf_counter() - t
        n = A.shape[0]
        nnz_m = s.info()["nnz_s"]
        spmv_us = s.spmv_a_us()
        layouts = s.pcg_layouts()
        s.close()
        nd_a, nd_m = pcg_layout_diagonals(A, layouts)
        per_it = tp / sum(counts)
        pcg_roof = {"kernel": f"PCG iteration (two kernels, FP32 M, A and M in layouts {layouts}; 16 solves' "
                              "time / their iterations, host set-up of each right-hand side included)",
                    "avg_iteration_us": round(per_it * 1e6, 3)}
        pcg_roof.update(roof(pcg_iteration_bytes(n, A.nnz, nnz_m, 4, nd_a, nd_m), per_it, achievable))
        spmv_roof = {"kernel": f"SpMV with A + <p, Ap> (A in layout {layouts[0]}, FP64), standalone replay",
                     "avg_launch_us": round(spmv_us, 3)}
        spmv_roof.update(roof(matrix_bytes(n, A.nnz, 8, nd_a) + 16 * n, spmv_us * 1e-6, achievable))
        line = {"config": "anisotropic2d 1024^2, inner-outer LOMR 5x20 FP32 bulk + 16 RHS PCG 1e-8", "build_s": tb,
                "pcg_16_rhs_s": tp, "pcg_iterations": counts, "roofline_pcg": pcg_roof, "roofline_spmv_a": spmv_roof}
        if with_cpu:
            from oracle import oracle
            t = time.perf_counter()
            oracle.run("io_lomr", A, M0, None, 5, 0.0, 1.0, False, threads, inner_iterations=5, fixed_pattern=True,
                       omp_threads=threads)
            t5 = time.perf_counter() - t
            line["cpu_baseline"] = {"value": t5 * 20, "unit": "s", "cores": threads, "kind": "port",
                                    "sample": "one outer iteration (5 inner, FP64, %.2f s) scaled to 20; solves not "
                                              "included" % t5}
        print(json.dumps(line), flush=True)

    if 5 in want:
        # config 5: 27-point 128^3, MR FP32 bulk 10 iterations; loss and gradients in FP32 per 65,536 columns
        import torch

        A = problems.stencil27(128, seed=4)
        M0 = problems.diag_inverse(A)
        s = _lib.Session("mr", A, M0=M0, max_iterations=10, precision="fp32")
        s.build()
        tb = _median_time(s.build, 3)
        n = A.shape[0]
        fused = _fused_kernel_roof(s, "mr", n, A.nnz, 4, achievable)
        M = s.result()["M"].tocsc()
        s.close()
        M.sort_indices()
        plan = ResidualLoss(A, M, torch.float32, block_cols=65536)
        m = torch.tensor(M.data, dtype=torch.float32, device="cuda")
        a = torch.tensor(A.data, dtype=torch.float32, device="cuda")
        plan.forward_value(m, a)
        tf = _median_time(lambda: plan.forward_value(m, a), 3)
        tg = _median_time(lambda: plan.gradients(m, a), 3)
        info = plan.info()
        plan.close()
        # loss algorithmic bytes (FP32 values, int32 indices): forward reads A
        # and M (patterns and values) and the pattern of R (S2); backward also
        # writes and reads back each R block and writes df/dM and df/dA
        nnz_r, nnz_m = info["nnz_r"], info["nnz_m"]
        fwd = (n + 1) * 4 + A.nnz * 8 + (n + 1) * 4 + nnz_m * 8 + (n + 1) * 4 + nnz_r * 4
        bwd = fwd + 2 * nnz_r * 4 + nnz_m * 4 + A.nnz * 4
        lf = {"kernel": "loss forward (LossColumns per 65,536-column block)"}
        lf.update(roof(fwd, tf, achievable))
        lb = {"kernel": "loss backward (LossColumns with df/dM, then LossGradA per block)"}
        lb.update(roof(bwd, tg, achievable))
        line = {"config": "stencil27 128^3 (n=2097152, nnz %d), MR FP32 bulk 10 its; loss fwd/bwd FP32, 65536-column "
                          "blocks" % A.nnz, "build_s": tb, "loss_forward_s": tf, "loss_backward_s": tg,
                "nnz_R": nnz_r, "blocks": info["blocks"], "roofline_build_pass": fused, "roofline_loss_forward": lf,
                "roofline_loss_backward": lb}
        if with_cpu:
            from oracle import oracle
            A64 = problems.stencil27(64, seed=4)
            M064 = problems.diag_inverse(A64)
            t = time.perf_counter()
            oracle.run("mr", A64, M064, None, 1, 0.0, 1.0, False, threads, fixed_pattern=True, omp_threads=threads)
            t1 = time.perf_counter() - t
            line["cpu_baseline"] = {"value": t1 * 8 * 10, "unit": "s", "cores": threads, "kind": "port",
                                    "sample": "FP64 MR, 1 iteration on 64^3 (%.2f s), scaled by 8 (size) and 10 "
                                              "(iterations)" % t1}
        print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 launches of an FP64 y = 1.0 * x kernel over MIB MiB (counter calibration)")
    ap.add_argument("--spgemm", type=int, default=0, metavar="REPEATS",
                    help="only run REPEATS standalone SpGEMM products C = A A (for kernel traces)")
    ap.add_argument("--configs", action="store_true",
                    help="instead of the headline line, one line per other configuration (1, 3, 4, 5)")
    ap.add_argument("--only", type=str, default="",
                    help="with --configs: comma-separated configuration numbers to run (default all)")
    args = ap.parse_args()
    if args.configs:
        only = [int(x) for x in args.only.split(",") if x.strip()]
        run_configs(not args.no_cpu_baseline, only=only)
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.spgemm:
        spgemm_only(args.spgemm)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib

    A, M0, b = problem()
    n, nnz = A.shape[0], A.nnz
    fp64 = time_form(_lib, A, M0, b, "fp64", args.steps, args.warmup)
    fp32 = time_form(_lib, A, M0, b, "fp32", args.steps, args.warmup)
    cold = {}
    for form in ("fp64", "fp32"):
        runs = [cold_step(_lib, A, M0, b, form) for _ in range(5)]
        cold[form] = {"seconds": round(float(np.median([r[0] for r in runs])), 6),
                      "runs_s": [round(r[0], 6) for r in runs], "pcg_iterations": runs[-1][1]}
    bw = _lib.stream_bandwidth(2 << 30, 20)
    achievable = max(bw["copy"], bw["triad"])
    dM = float(np.linalg.norm((fp32["M"] - fp64["M"]).data) / np.linalg.norm(fp64["M"].data))

    # roofline of the dominant kernel of the FP32-bulk step: the LOMR pass
    # (one iteration; the diagonal pass on this stencil), algorithmic bytes of
    # lomr_pass_bytes with s_v = 4
    shape = fp32["launch_shape"]
    pass_bytes = lomr_pass_bytes(n, nnz, shape, 4, "lomr")
    # the same iteration in the general fused pass's algorithmic bytes (R
    # stored on S2, CSC patterns), for comparison with that pass
    fused_equiv_bytes = fused_pass_bytes(n, nnz, shape["nnz_s"], shape["nnz_s2"], 4, "lomr")
    pass_kernel = lomr_pass_kernel(shape, "lomr", "float")
    us = fp32["lomr_pass_us"]
    achieved = pass_bytes / (us * 1e-6) / 1e9
    traffic = None
    prof = ROOT / "profiles" / ("lomr_dia_diffusion64_traffic.json" if shape["fused"] == 2
                                else "lomr_fused_diffusion64_traffic.json")
    if prof.exists():
        traffic = json.loads(prof.read_text()).get("fp32_bytes_per_launch")
    # the PCG's A-SpMV + <p, Ap>: on the diagonal layout the solve reads,
    # nd*n*8 (values) + 4n (row masks) + 2*8*n (p read, Ap written) bytes per
    # launch; on CSR/SELL (n+1)*4 + nnz*(4+8) + 2*8*n
    csr_spmv_bytes = (n + 1) * 4 + nnz * 12 + 16 * n
    nd_a = n_diagonals(A)
    spmv_layout = fp32["pcg_layouts"]["A"]
    spmv_bytes = matrix_bytes(n, nnz, 8, nd_a if spmv_layout == "dia" else None) + 16 * n
    spmv_kernel = {"dia": f"PcgSpmvADia (SpMV with A + <p, Ap> on the {nd_a}-diagonal layout the solve reads)",
                   "sell": "PcgSpmvASell (SpMV with A + <p, Ap> on the sliced-ELLPACK copy the solve reads)",
                   "csr": "PcgSpmvA (SpMV with A + <p, Ap> on CSR)"}[spmv_layout]
    spmv_us = fp32["spmv_a_us"]
    # one PCG iteration of the FP32-bulk solve (K1 + K2, FP32 M) from the
    # solve's time over its iterations (the per-solve start, refresh, final
    # residual and read-back included)
    lay = (fp32["pcg_layouts"]["A"], fp32["pcg_layouts"]["M"])
    it_s = fp32["pcg_s"] / max(fp32["pcg_iterations"], 1)
    pcg_it_roof = {"kernel": f"PCG iteration (PcgDirSpmvA + PcgUpdateSpmvM, FP32 M, layouts {lay}), "
                             "solve time / iterations", "bound": "hbm", "avg_iteration_us": round(it_s * 1e6, 3)}
    pcg_it_roof.update(roof(pcg_iteration_bytes(n, nnz, nnz, 4, *pcg_layout_diagonals(A, lay)), it_s,
                            max(bw["copy"], bw["triad"])))
    # the FP32-over-FP64 step ratio with every kernel at the HBM roofline:
    # algorithmic bytes of the two steps (LOMR passes + PCG iterations)
    nd_pair = pcg_layout_diagonals(A, lay)
    step_bytes = {}
    for form, res, sv in (("fp32", fp32, 4), ("fp64", fp64, 8)):
        step_bytes[form] = ((res["spai_iterations"] + 1) * lomr_pass_bytes(n, nnz, res["launch_shape"], sv, "lomr")
                            + res["pcg_iterations"] * pcg_iteration_bytes(n, nnz, nnz, sv, *nd_pair))
    ratio_bound = step_bytes["fp64"] / step_bytes["fp32"]
    spmv_traffic = None
    prof2 = ROOT / "profiles" / f"pcg_spmv_a_{spmv_layout}_diffusion64_traffic.json"
    if prof2.exists():
        spmv_traffic = json.loads(prof2.read_text()).get("bytes_per_launch")

    # the reference-semantics SpGEMM (symbolic + numeric phases, FP64): C = A A,
    # i.e. A M with M on A's pattern; algorithmic bytes 3 (n+1) 4 + 12 (nnz(A)
    # + nnz(A) + nnz(C)), operands resident, mean of 10 products
    C, spgemm_s = _lib.spgemm(A, A, repeats=10)
    spgemm_bytes = 3 * (n + 1) * 4 + 12 * (nnz + nnz + C.nnz)
    spgemm_flops = 2 * int(np.diff(A.indptr)[A.indices].sum())

    threads = host_threads()
    cpu = None if args.no_cpu_baseline else cpu_baseline(A, M0, b, threads)

    line = {
        "metric": METRIC,
        "value": round(fp32["seconds"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp32["seconds"] * 1e3, 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 bulk + f64 reductions",
        "data": "synthetic",
        "config": {"workload": "diffusion3d 64^3 (n=262144, nnz=1810432), Global SPAI LOMR on pattern(A), "
                               "30 iterations, M0=diag(A)^-1, PCG to 1e-8, x_true~N(0,1) seed 2",
                   "n": n, "nnz": nnz},
        "fp32_bulk": {k: v for k, v in fp32.items() if k not in ("M",)},
        "fp64": {k: v for k, v in fp64.items() if k not in ("M",)},
        "speedup_fp32_vs_fp64": round(fp64["seconds"] / fp32["seconds"], 4),
        "cold": {"fp32_bulk": cold["fp32"], "fp64": cold["fp64"],
                 "speedup_fp32_vs_fp64": round(cold["fp64"]["seconds"] / cold["fp32"]["seconds"], 4)},
        "speedup_target": {"target": 1.5, "met_warm": fp64["seconds"] / fp32["seconds"] >= 1.5,
                           "met_cold": cold["fp64"]["seconds"] / cold["fp32"]["seconds"] >= 1.5,
                           "bound_at_roofline": round(ratio_bound, 3),
                           "note": "the PCG recurrence is FP64 in both forms (only M is FP32); with every kernel "
                                   "at the HBM roofline the step ratio is the ratio of the two steps' algorithmic "
                                   "bytes (LOMR passes + PCG iterations), bound_at_roofline"},
        "extra_pcg_iterations_fp32": fp32["pcg_iterations"] - fp64["pcg_iterations"],
        "rel_fro_M32_vs_M64": dM,
        "roofline": {"kernel": pass_kernel, "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "frac_of_achievable": round(achieved / achievable, 4),
                     "algorithmic_bytes_per_launch": pass_bytes, "avg_launch_us": round(us, 3),
                     "fused_pass_equivalent_GBs": round(fused_equiv_bytes / (us * 1e-6) / 1e9, 1),
                     "launches_timed": fp32["lomr_passes_timed"]},
        "roofline_pcg_spmv": {"kernel": spmv_kernel + ", standalone replay", "bound": "hbm",
                              "achieved": round(spmv_bytes / (spmv_us * 1e-6) / 1e9, 1), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(spmv_bytes / (spmv_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                              "frac_of_achievable": round(spmv_bytes / (spmv_us * 1e-6) / 1e9 / achievable, 4),
                              "traffic": spmv_traffic, "algorithmic_bytes_per_launch": spmv_bytes,
                              "csr_equivalent_GBs": round(csr_spmv_bytes / (spmv_us * 1e-6) / 1e9, 1),
                              "avg_launch_us": round(spmv_us, 3)},
        "roofline_pcg_iteration": pcg_it_roof,
        "spgemm": {"kernel": "SpGemm C = A A (every column short: ExpansionSizes, MergeCount, scan, MergeNumeric), FP64",
                   "seconds": round(spgemm_s, 7), "nnz_C": int(C.nnz), "algorithmic_bytes": spgemm_bytes,
                   "achieved_GBs": round(spgemm_bytes / spgemm_s / 1e9, 1),
                   "frac": round(spgemm_bytes / spgemm_s / 1e9 / HBM_PEAK_GBS, 4),
                   "frac_of_achievable": round(spgemm_bytes / spgemm_s / 1e9 / achievable, 4),
                   "GFLOPs": round(spgemm_flops / spgemm_s / 1e9, 1)},
        "achievable_bandwidth": {"read_GBs": round(bw["read"], 1), "copy_GBs": round(bw["copy"], 1),
                                 "triad_GBs": round(bw["triad"], 1), "achievable_GBs": round(achievable, 1),
                                 "arrays": "FP64, 2 GiB each, median of 20 launches"},
        "cpu_baseline": cpu,
        "device": _lib.device_name(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
