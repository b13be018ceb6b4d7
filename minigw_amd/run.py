"""G0W0 on the hip backend from a gw_input file, writing miniGW's output files.

    python -m minigw_amd.run --input-dir DIR [--output-dir DIR] [--freq-points 200]
        [--pade-params 16] [--state 5 | --all-states] [--eta 0.0]
        [--precision fp64|mixed|mixed_int8|fp64_cd] [--cd-tolerance 1e-12]

The options, defaults and outputs follow miniGW's CLI (src/cli.cpp,
include/cli.hpp:10-22; src/main.cpp:38-133): DIR/gw_input.h5 (or the raw form
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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="minigw_amd.run")
    ap.add_argument("--input-dir", default=".")
    ap.add_argument("--output-dir", default="./gw_output")
    ap.add_argument("--freq-points", type=int, default=200)
    ap.add_argument("--pade-params", type=int, default=16)
    ap.add_argument("--state", type=int, default=5)
    ap.add_argument("--all-states", action="store_true")
    ap.add_argument("--eta", type=float, default=0.0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "mixed_int8", "fp64_cd"])
    ap.add_argument("--cd-tolerance", type=float, default=1e-12,
                    help="fp64_cd: ERI decomposition stops at this fraction of the first pivot")
    ap.add_argument("--linalg-backend", default="hip", choices=["hip"])
    args = ap.parse_args(argv)
    if args.freq_points <= 0:
        raise SystemExit("--freq-points must be positive")
    if args.pade_params <= 0:
        raise SystemExit("--pade-params must be positive")
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
    prec = {"fp64": hip.PREC_FP64, "mixed": hip.PREC_MIXED,
            "mixed_int8": hip.PREC_MIXED_INT8, "fp64_cd": hip.PREC_FP64_CD}[args.precision]
    with hip.HipContext(hip.device_for_local_rank(0, 1)) as ctx:
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
    out(f"G0W0 on device:  {t2 - t1:8.2f} s  ({args.precision}, {args.freq_points} frequencies)")
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
