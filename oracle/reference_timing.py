"""Times the reference's own CPU path (test infrastructure / CPU baseline).

TEST INFRASTRUCTURE ONLY: runs oracle/_ref/gw_ref_blas, the reference's
run_g0w0 with its BlasLapackBackend (src/linalg/linalg_blas_lapack.cpp,
zgetrf/zgetrs/zgemm from the LAPACK provider the Makefile linked), built by
oracle/Makefile from the reference sources.  The product never calls it;
bench.py's cpu_baseline leg does.

Schedule: the reference CLI's --frequency-parallel serial --kernel-parallel
openmp with OMP_NUM_THREADS = MKL_NUM_THREADS = the cores used.  The
reference recomputes W_c(w') for every external frequency w_n
(src/gw/sigma_c_driver.cpp:132-169), so a run at F frequencies does F^2
identical (w_n, w') units, each a complex LU + solve against the identity
(n x n) and the quadratic forms of every (state, k) panel.  A run at a small
F' therefore gives the full run exactly in operation count:

    t(F) = t_setup + (F / F')^2 * t_units(F'),

where t_units is the time of its F'^2 units (build Pi0 + invert epsilon +
Sigma_c accumulation, the reference's own GwTimings) and t_setup the V_ph
gather.  "hoisted" reports what the same units cost when each w' is done
once (F units), the schedule the device path runs.

Usage: python oracle/reference_timing.py [--configs nmo24 nmo120 nmo200]
prints one JSON line per (config, F').
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GW_REF_BLAS = os.path.join(HERE, "_ref", "gw_ref_blas")


def available() -> bool:
    return os.path.exists(GW_REF_BLAS)


def write_factor_input(path: str, mo, fermi: float, b, vxc, nocc: int, nfreq: int, npar: int,
                       state: int = -1, threads: int = 1) -> None:
    """oracle/ref_driver.cpp input with the ERI given by its factor B."""
    nmo = len(mo)
    with open(path, "wb") as fh:
        np.array([nmo, nocc, nfreq, npar, state, threads, b.shape[1]], dtype="<i8").tofile(fh)
        np.array([fermi, 0.0], dtype="<f8").tofile(fh)
        np.ascontiguousarray(mo, dtype="<f8").tofile(fh)
        np.ascontiguousarray(vxc, dtype="<f8").tofile(fh)
        np.ascontiguousarray(b, dtype="<f8").tofile(fh)


def run(nmo: int, nocc: int, seed: int, nfreq: int, threads: int, npar: int = 16,
        timeout: float | None = None) -> dict:
    """One run of the reference (all states) on the seeded synthetic molecule."""
    sys.path.insert(0, REPO)
    from minigw_amd.synthetic import synthetic_factor
    mo, fermi, b, vxc = synthetic_factor(nmo, nocc, seed)
    with tempfile.TemporaryDirectory() as td:
        ip, op = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_factor_input(ip, mo, fermi, b, vxc, nocc, nfreq, npar)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), MKL_NUM_THREADS=str(threads),
                   MKL_THREADING_LAYER="GNU")
        t0 = time.perf_counter()
        res = subprocess.run([GW_REF_BLAS, ip, op, "blas-lapack", "kernel-openmp"], env=env,
                             capture_output=True, text=True, timeout=timeout)
        wall = time.perf_counter() - t0
        if res.returncode != 0:
            raise RuntimeError(f"gw_ref_blas failed: {res.stderr[-500:]}")
    line = next(l for l in res.stdout.splitlines() if l.startswith("timings "))
    vals = dict(re.findall(r"(\w+) ([-+0-9.eE]+)", line))
    t = {k: float(v) for k, v in vals.items()}
    units = t["build_pi0"] + t["invert_epsilon"] + t["sigma_c"]
    return {"nmo": nmo, "nocc": nocc, "nph": nocc * (nmo - nocc), "nfreq_run": nfreq,
            "threads": threads, "process_wall_s": wall, "sigma_c_wall_s": t["sigma_c_wall"],
            "setup_s": t["screening_setup"], "units_s": units, "unit_s": units / nfreq ** 2}


def extrapolate(r: dict, nfreq: int) -> dict:
    """Full-run Sigma_c time at nfreq, reference schedule (F^2 units) and hoisted (F)."""
    ref_s = r["setup_s"] + r["unit_s"] * nfreq ** 2
    hoisted_s = r["setup_s"] + r["unit_s"] * nfreq
    return {"nfreq": nfreq, "reference_schedule_s": ref_s, "hoisted_schedule_s": hoisted_s,
            "reference_qp_solves_per_s": r["nmo"] / ref_s,
            "hoisted_qp_solves_per_s": r["nmo"] / hoisted_s}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", nargs="*", default=["nmo24", "nmo120", "nmo200"])
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    args = ap.parse_args()
    sys.path.insert(0, REPO)
    from minigw_amd.synthetic import CONFIGS
    plan = {"nmo24": [200], "nmo120": [4, 8], "nmo200": [1, 4]}
    for name in args.configs:
        nmo, nocc, seed = CONFIGS[name]
        for fp in plan[name]:
            r = run(nmo, nocc, seed, fp, args.threads)
            r.update(config=name, extrapolated=extrapolate(r, 200), exact=fp == 200)
            print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
