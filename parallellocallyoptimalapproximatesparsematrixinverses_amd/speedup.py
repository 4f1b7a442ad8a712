"""Thread-speedup study in the reference's format, with the device beside the
host-thread columns.

Mirrors scripts/methods_thread_speedup.py:
  * THREAD_COUNTS, TIMING_COLUMNS     the thread counts and the timing CSV
                                      columns (:17-20)
  * time_device                       _get_average_running_time /
                                      _run_method_for_threads (:23-93): every
                                      method variant at every p, `tries` timed
                                      end-to-end calls each (uploads and
                                      downloads included, as the reference
                                      times its Cython call)
  * speedup_cell                      _build_speedup_table_row (:96-105):
                                      t, S_p = t_1 / t_p, E_p = S_p / p
  * read_timing_csv, write_timing_csv _write_timing_csv (:131-139)
  * summary_rows, write_summary       the speedup summary table (:142-176),
                                      written as Markdown, with a device
                                      column added

On the device, `num_threads` is the number of column batches the builder
emulates (the reference's batch decomposition), which fixes the per-batch
dropping thresholds and, with batch-ordered reductions, the summation order;
it does not change the device's parallelism.  The host-thread columns come
from a timing CSV of the reference (its committed results, or a new run of
methods_thread_speedup.py).

  python -m parallellocallyoptimalapproximatesparsematrixinverses_amd.speedup \\
      --matrix A.mtx --cpu-timings all_thread_speedup_timings.csv --tries 3 \\
      --max-iterations 25 --max-density 0.03 --out results/
"""
import argparse
import csv
import time
from collections import defaultdict
from pathlib import Path

import numpy as np

from . import _lib, datasets

THREAD_COUNTS = (1, 2, 4, 8, 16, 32, 64)
TIMING_COLUMNS = ("method", "num_threads", "repetition", "runtime_seconds")


def time_device(A, max_iterations, max_density, enable_dropping, tries, tolerance=1e-9,
                thread_counts=THREAD_COUNTS, variants=None, batch_ordered=False):
    """Timing rows (TIMING_COLUMNS) of every variant at every batch count,
    `tries` repetitions each, in the reference's loop order."""
    if tries <= 0:
        raise ValueError("tries must be positive")
    A = datasets.load_csc(A)
    M0, Pr = datasets.harness_inputs(A)
    rows = []
    for label, method, inner in (variants or datasets.VARIANTS):
        for p in thread_counts:
            for rep in range(tries):
                start = time.perf_counter()
                _lib.run(method, A, M0, Pr, max_iterations, tolerance, max_density, enable_dropping, p,
                         inner_iterations=inner, batch_ordered=batch_ordered)
                rows.append({"method": label, "num_threads": p, "repetition": rep + 1,
                             "runtime_seconds": time.perf_counter() - start})
    return rows


def read_timing_csv(path):
    with open(path, newline="", encoding="utf-8") as fh:
        return [{"method": r["method"], "num_threads": int(r["num_threads"]), "repetition": int(r["repetition"]),
                 "runtime_seconds": float(r["runtime_seconds"])} for r in csv.DictReader(fh)]


def write_timing_csv(path, rows):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", newline="", encoding="utf-8") as fh:
        writer = csv.DictWriter(fh, fieldnames=TIMING_COLUMNS)
        writer.writeheader()
        writer.writerows(rows)
    return path


def averages(rows):
    """{method: {p: mean runtime over the repetitions}}, methods in first-seen order."""
    acc = defaultdict(lambda: defaultdict(list))
    for r in rows:
        acc[r["method"]][r["num_threads"]].append(r["runtime_seconds"])
    return {m: {p: float(np.mean(v)) for p, v in by_p.items()} for m, by_p in acc.items()}


def speedup_cell(t_p, t_1, p):
    """(t, S_p, E_p) as the reference's table cell computes them; S_p is NaN
    when either time is not positive."""
    s = t_1 / t_p if t_1 > 0.0 and t_p > 0.0 else float("nan")
    return t_p, s, s / p


def summary_rows(cpu_rows, device_rows, thread_counts=THREAD_COUNTS):
    """Per method: the host cells at every p (from cpu_rows) and the device's
    mean time over its batch counts with its speedup over the host at p = 1
    and at its fastest p."""
    cpu = averages(cpu_rows)
    dev = averages(device_rows)
    out = []
    for method in list(cpu) + [m for m in dev if m not in cpu]:
        row = {"method": method}
        c = cpu.get(method, {})
        t1 = c.get(1, float("nan"))
        for p in thread_counts:
            row[f"p_{p}"] = speedup_cell(c[p], t1, p) if p in c else None
        if method in dev:
            t_dev = float(np.mean(list(dev[method].values())))
            best = min(c.values()) if c else float("nan")
            row["device"] = (t_dev, t1 / t_dev, best / t_dev)  # NaN without host timings
        else:
            row["device"] = None
        out.append(row)
    return out


def write_summary(path, rows, device_name="GPU", thread_counts=THREAD_COUNTS):
    """Markdown table: one row per method, a cell per p (t, S_p, E_p) and the
    device cell (t, speedup over p = 1, speedup over the fastest p)."""
    head = ["Method"] + [f"p = {p}" for p in thread_counts] + [device_name]
    lines = ["| " + " | ".join(head) + " |", "|" + "---|" * len(head)]
    for r in rows:
        cells = [r["method"]]
        for p in thread_counts:
            c = r[f"p_{p}"]
            cells.append("" if c is None else f"t={c[0]:.3f}s S_p={c[1]:.2f} E_p={c[2]:.2f}")
        d = r["device"]
        cells.append("" if d is None else f"t={d[0]:.4f}s vs p=1 {d[1]:.1f}x, vs best p {d[2]:.1f}x")
        lines.append("| " + " | ".join(cells) + " |")
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text("\n".join(lines) + "\n", encoding="utf-8")
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--matrix", required=True, help="MatrixMarket file, or .npz of CSC arrays")
    ap.add_argument("--cpu-timings", help="timing CSV of the reference (methods_thread_speedup.py)")
    ap.add_argument("--tries", "-N", type=int, required=True)
    ap.add_argument("--max-iterations", type=int, default=100)
    ap.add_argument("--tolerance", type=float, default=1e-9)
    ap.add_argument("--max-density", type=float, default=0.03)
    ap.add_argument("--disable-dropping", dest="enable_dropping", action="store_false", default=True)
    ap.add_argument("--out", type=Path, default=Path("."))
    args = ap.parse_args(argv)
    if str(args.matrix).endswith(".npz"):
        z = np.load(args.matrix, allow_pickle=False)  # scipy CSC arrays: indptr, indices, data, shape
        from scipy import sparse
        A = sparse.csc_matrix((z["data"], z["indices"], z["indptr"]), shape=tuple(z["shape"]))
    else:
        A = datasets.load_matrix_market(args.matrix)
    dev = time_device(A, args.max_iterations, args.max_density, args.enable_dropping, args.tries, args.tolerance)
    write_timing_csv(args.out / "all_thread_speedup_timings_device.csv", dev)
    cpu = read_timing_csv(args.cpu_timings) if args.cpu_timings else []
    write_summary(args.out / "all_speedup_summary.md", summary_rows(cpu, dev))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
