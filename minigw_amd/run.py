"""G0W0 on the hip backend from a gw_input file, writing miniGW's output files.

    python -m minigw_amd.run --input-dir DIR [--output-dir DIR] [--freq-points 200]
        [--pade-params 16] [--state 5 | --all-states] [--eta 0.0]
        [--linalg-backend hip|hip-int8|hip-int8x4|hip-fp16|hip-cd]
        [--precision fp64|mixed|mixed_int8|mixed_int8x4|fp64_cd] [--cd-tolerance 1e-12]
        [--tasks-per-gpu 4] [--print-memory-footprint]
        [--frequency-parallel ...] [--kernel-parallel ...] [--contraction-panel-size N]
        [--scalapack-ranks-per-group N]

The options, defaults and outputs follow miniGW's CLI (src/cli.cpp:33-104,
include/cli.hpp:10-22; src/main.cpp:38-133).  The backend name selects the
form as in the reference's own hip backend ("hip" FP64, "hip-int8" five
7-bit digits, "hip-int8x4" four 8-bit digits, "hip-fp16" the FP16 split,
"hip-cd" the decomposed ERI); --precision, when given, overrides it.  The
reference's scheduling options are accepted and checked as it checks them:
the frequencies run batched on one device whatever --frequency-parallel /
--kernel-parallel say, every column is contracted at once whatever
--contraction-panel-size says, and --tasks-per-gpu maps local rank 0 to its
device (set_cuda_device_for_local_rank).  Input: DIR/gw_input.h5 (or the raw form
DIR/gw_input.raw), E_c_before_Pade.out and E_c.out for the selected state
(the first state with --all-states), and gw.out ending in the
"--- G0W0 Calculation Summary ---" table that scripts/compare_reference.py
reads.  Everything between the input and the files runs on the GPU.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np


# backend name -> form, as the reference-side hip backend names them
BACKENDS = {"hip": "fp64", "hip-int8": "mixed_int8", "hip-int8x4": "mixed_int8x4",
            "hip-fp16": "mixed", "hip-cd": "fp64_cd"}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="minigw_amd.run")
    ap.add_argument("--input-dir", default=".")
    ap.add_argument("--output-dir", default="./gw_output")
    ap.add_argument("--freq-points", type=int, default=200)
    ap.add_argument("--pade-params", type=int, default=16)
    ap.add_argument("--state", type=int, default=5)
    ap.add_argument("--all-states", action="store_true")
    ap.add_argument("--eta", type=float, default=0.0)
    ap.add_argument("--precision", default=None,
                    choices=["fp64", "mixed", "mixed_int8", "mixed_int8x4", "fp64_cd"])
    ap.add_argument("--cd-tolerance", type=float, default=1e-12,
                    help="fp64_cd: ERI decomposition stops at this fraction of the first pivot")
    ap.add_argument("--linalg-backend", default="hip", choices=list(BACKENDS))
    ap.add_argument("--frequency-parallel", default="auto")
    ap.add_argument("--kernel-parallel", default="auto")
    ap.add_argument("--contraction-panel-size", type=int, default=32)
    ap.add_argument("--scalapack-ranks-per-group", type=int, default=4)
    ap.add_argument("--tasks-per-gpu", type=int, default=4)
    ap.add_argument("--print-memory-footprint", action="store_true")
    args = ap.parse_args(argv)
    if args.freq_points <= 0:
        raise SystemExit("--freq-points must be positive")
    if args.pade_params <= 0:
        raise SystemExit("--pade-params must be positive")
    for name in ("contraction_panel_size", "scalapack_ranks_per_group", "tasks_per_gpu"):
        if getattr(args, name) <= 0:
            raise SystemExit(f"--{name.replace('_', '-')} must be positive")
    if args.frequency_parallel not in ("auto", "serial", "mpi", "openmp"):
        raise SystemExit("--frequency-parallel must be one of: auto, serial, mpi, openmp")
    if args.kernel_parallel not in ("auto", "serial", "openmp"):
        raise SystemExit("--kernel-parallel must be one of: auto, serial, openmp")
    precision = args.precision or BACKENDS[args.linalg_backend]
    if not args.all_states and args.state == 0:
        raise SystemExit("--state must be 1-based and positive")

    from . import gwio, hip

    path = os.path.join(args.input_dir, "gw_input.h5")
    if not os.path.exists(path):
        path = os.path.join(args.input_dir, "gw_input.raw")
    os.makedirs(args.output_dir, exist_ok=True)
    log = open(os.path.join(args.output_dir, "gw.out"), "w")

    def out(line=""):
        print(line)
        log.write(line + "\n")

    t0 = time.perf_counter()
    inp = gwio.read_gw_input(path)
    nmo = inp.nmo
    out(f"Loading data from: {args.input_dir}")
    out(f"  - nmo = {nmo}, nocc = {inp.nocc}, particle-hole states = {inp.nph}")
    states = list(range(nmo)) if args.all_states else [args.state - 1]
    if not all(0 <= s < nmo for s in states):
        raise SystemExit("Selected state is out of range")
    prec = {"fp64": hip.PREC_FP64, "mixed": hip.PREC_MIXED, "mixed_int8": hip.PREC_MIXED_INT8,
            "mixed_int8x4": hip.PREC_MIXED_INT8_4, "fp64_cd": hip.PREC_FP64_CD}[precision]
    if args.print_memory_footprint:
        out(f"  - device memory estimate: "
            f"{hip.estimate_bytes(nmo, inp.nocc, args.freq_points, precision=prec) / 2**30:.3f} GiB")
    with hip.HipContext(hip.device_for_local_rank(0, args.tasks_per_gpu)) as ctx:
        ctx.set_cd_tolerance(args.cd_tolerance)
        t1 = time.perf_counter()
        res = hip.run_g0w0(inp, args.freq_points, args.pade_params, states=states, eta=args.eta,
                           precision=prec, ctx=ctx)
        t2 = time.perf_counter()
        row = res.sigma_c[0]                     # the output state: selected, or state 0
        gwio.write_sigma_before_pade(os.path.join(args.output_dir, "E_c_before_Pade.out"), row)
        vals = gwio.pade_real_axis(row, res.omegas, args.pade_params, inp.fermi_energy, ctx=ctx)
        gwio.write_sigma_after_pade(os.path.join(args.output_dir, "E_c.out"), vals,
                                    inp.fermi_energy)
        t3 = time.perf_counter()
    out(f"read input:      {t1 - t0:8.2f} s")
    out(f"G0W0 on device:  {t2 - t1:8.2f} s  ({precision}, {args.freq_points} frequencies)")
    out(f"pade + files:    {t3 - t2:8.2f} s")
    out("")
    out("--- G0W0 Calculation Summary ---")
    table = gwio.format_energy_table(inp.mo_energy, res.sigma_x_diag, np.diag(inp.vxc), res.qp,
                                     states)
    print(table, end="")
    log.write(table)
    out("")
    out("--- End of G0W0 Calculation (hip backend) ---")
    log.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
