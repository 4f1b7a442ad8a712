"""Writes the golden fixtures of tests/golden/ from the reference's committed
data files (run in a container where the reference tree is mounted).

  * <name>.npz: the sample matrix (scripts/sample_datasets/<name>-dataset.mtx),
    canonicalised as scripts/common.py:396-400 does (CSC, duplicates summed,
    zeros eliminated, indices sorted).  The larger matrices (LOWER) are stored
    as their lower triangle with int32 indices and a "lower" flag; they are
    exactly symmetric, and tests/golden_data.py rebuilds the full matrix as
    L + strict_lower(L)^T, which adds no two values.
  * <name>-<mode>-d<density>-convergence.csv: the first rows of the reference's
    per-iteration ||R_i||_F table for the seven method variants
    (scripts/results/results_sample/<mode>/max_iter_*/max_density_*/<name>/*-convergence.csv).
  * <name>-<mode>-d<density>-{construction_status,backward_error,solver_status}.csv
    (PIPELINES only): the reference's end-of-build statuses, the PCG
    relative-residual histories with its preconditioners (150-entry limit) and
    the solver statuses, copied whole.
  * <name>-<mode>-i<iterations>-d<density>-thread_speedup_timings.csv
    (SPEEDUP only): the reference's CPU timings of every method variant at
    p = 1..64 threads (scripts/results/results_parallelization_speedup/...),
    copied whole.
"""
import csv
import glob
import sys
from pathlib import Path

import numpy as np
from scipy import sparse
from scipy.io import mmread

REF = Path("/root/reference/scripts")
OUT = Path(__file__).resolve().parent
ROWS = 41

DATASETS = ["tri100eigs4k", "bcsstk21", "msc04515"]
LOWER = ["Poisson32k", "wathen100", "4bw100eigs20k", "4bw100eigs20k2", "bundle1", "rand20k", "rand20k2"]
PIPELINES = [("tri100eigs4k", "no_dropping", "1.000"), ("Poisson32k", "dropping", "0.030")]
SPEEDUP = [("tri100eigs4k", "dropping", "25", "0.030")]


def main():
    for name in DATASETS + LOWER:
        A = mmread(REF / "sample_datasets" / f"{name}-dataset.mtx").tocsc().astype(np.float64)
        A.sum_duplicates()
        A.eliminate_zeros()
        A.sort_indices()
        if name in LOWER:
            if abs(A - A.T).max() != 0.0:
                raise ValueError(f"{name} is not exactly symmetric")
            L = sparse.tril(A).tocsc()
            L.sort_indices()
            np.savez_compressed(OUT / f"{name}.npz", indptr=L.indptr.astype(np.int32),
                                indices=L.indices.astype(np.int32), data=L.data, shape=np.array(L.shape),
                                lower=np.array(1))
        else:
            np.savez_compressed(OUT / f"{name}.npz", indptr=A.indptr.astype(np.int64),
                                indices=A.indices.astype(np.int64), data=A.data, shape=np.array(A.shape))
        for f in sorted(glob.glob(str(REF / f"results/results_sample/*/max_iter_*/max_density_*/{name}/*-convergence.csv"))):
            if "no_dropping" in f and "max_density_0.030" in f:
                continue  # written by an earlier revision (see tests/test_oracle_golden.py)
            p = Path(f)
            mode = p.parts[-5]
            dens = p.parts[-3].split("_")[-1]
            with open(f) as fh:
                rows = list(csv.reader(fh))
            with open(OUT / f"{name}-{mode}-d{dens}-convergence.csv", "w", newline="") as fh:
                csv.writer(fh).writerows(rows[:ROWS + 1])
    for name, mode, dens in PIPELINES:
        for f in sorted(glob.glob(str(REF / f"results/results_sample/{mode}/max_iter_*/max_density_{dens}/{name}/*.csv"))):
            kind = Path(f).stem.split("-")[-1]
            if kind in ("construction_status", "backward_error", "solver_status"):
                with open(f) as fh:
                    rows = list(csv.reader(fh))
                with open(OUT / f"{name}-{mode}-d{dens}-{kind}.csv", "w", newline="") as fh:
                    csv.writer(fh).writerows(rows)
    for name, mode, its, dens in SPEEDUP:
        src = REF / f"results/results_parallelization_speedup/{mode}/max_iter_{its}/max_density_{dens}/{name}"
        with open(src / "all_thread_speedup_timings.csv") as fh:
            rows = list(csv.reader(fh))
        with open(OUT / f"{name}-{mode}-i{its}-d{dens}-thread_speedup_timings.csv", "w", newline="") as fh:
            csv.writer(fh).writerows(rows)
    return 0


if __name__ == "__main__":
    sys.exit(main())
