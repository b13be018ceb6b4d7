"""G0W0 QP-solve throughput of the hip backend on one MI355X.

A step is one molecule's G0W0 on the device: all imaginary frequencies of the
transformed Gauss-Legendre grid, every state's Sigma_c(i w), Pade fit and QP
fixed point, on the ERI already resident in HBM.  value = QP solves per second
(one QP solve = one state's QP energy) of the mixed-precision form; the plain
FP64 form runs beside it as the error reference and for the mixed/FP64 ratio, and
the decomposed-ERI FP64 form (pivoted Cholesky of the pair matrix on the
device, screening in the rank-r pair space) is reported as
"decomposed_fp64" with its own accuracy against the plain form.

Prints one JSON line.  Default workload: the synthetic nmo=200, nocc=40
molecule (nph=6400, ERI 12.8 GB), 200 frequencies, 16 Pade parameters.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "G0W0 QP solves/sec + p50/p99 latency, mixed vs FP64 HIP, QP err ≤1 meV"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def contraction_flops(n: int, m: int, precision: int) -> float:
    """Algorithmic flops of one frequency's contraction Z = (L^-1 S) X, L^-1
    lower triangular: 2 * (n(n+1)/2) * m for the m contracted columns (the
    unordered (p,k) pairs when q(p,k) = q(k,p)); three products in the FP16
    split form, fifteen int8 slice products in the int8 form."""
    f = float(n) * (n + 1) * m
    return {1: 3.0, 2: 15.0}.get(precision, 1.0) * f


HA_MEV = 27.211386245988 * 1e3

# HBM bytes per contraction launch (FETCH_SIZE + WRITE_SIZE, each scaled by
# what the counter reads for a 2 GiB copy), averaged over the molecule's
# launches, from two rocprofv3 --pmc runs of this bench with that form alone;
# the commands and per-dispatch figures are in profiles/kernels_nmo200_*.md
# (int8, fp64, fp16split, fp64_cd).
TRAFFIC = {("mixed_int8", "nmo200", 200): (456.06e9 + 0.39e9),
           ("fp64", "nmo200", 200): (1422.64e9 + 0.39e9),
           ("mixed", "nmo200", 200): (137.12e9 + 0.20e9),
           ("fp64_cd", "nmo200", 200): (64.87e9 + 0.16e9)}


def accuracy(hip, ctx, inp, mix, f64, npar: int, trials: int = 3) -> dict:
    """Mixed against FP64 on one molecule: Sigma_c(i w) error, QP error, and
    how well the FP64 QP energies themselves are determined.  The Pade/QP step
    is re-run on the device on the FP64 Sigma_c with relative noise at the
    mixed error level; a state is 'resolvable' when that moves its QP energy by
    less than 1 meV (otherwise no FP64-vs-mixed comparison below 1 meV can be
    expected of that state)."""
    ds = np.abs(mix.sigma_c - f64.sigma_c)
    scale = np.abs(f64.sigma_c).max()
    rel = float(ds.max() / scale)
    dq = np.abs(mix.qp - f64.qp) * HA_MEV
    rng = np.random.default_rng(11)
    vxc = np.diag(inp.vxc)

    def spread(eps):
        sp = np.zeros(inp.nmo)
        for _ in range(trials):
            noise = rng.standard_normal(f64.sigma_c.shape) + 1j * rng.standard_normal(f64.sigma_c.shape)
            qp, _ = hip.pade_qp(f64.sigma_c + eps * scale * noise, f64.omegas, f64.sigma_x_diag,
                                inp.mo_energy, vxc, inp.fermi_energy, npar, ctx=ctx)
            sp = np.maximum(sp, np.nan_to_num(np.abs(qp - f64.qp), nan=np.inf) * HA_MEV)
        return sp

    sp_mix = spread(max(rel, 1e-16))
    sp_tiny = spread(1e-14)
    ok = sp_mix < 1.0
    return {
        "sigma_c": {"max_abs_mev": float(ds.max() * HA_MEV),
                    "median_state_max_abs_mev": float(np.median(ds.max(axis=1)) * HA_MEV),
                    "max_rel": rel, "max_abs_sigma_ha": float(scale)},
        "qp": {"max_mev": float(np.nanmax(dq)), "median_mev": float(np.nanmedian(dq)),
               "within_1meV": int(np.sum(dq <= 1.0)), "states": int(inp.nmo),
               "fp64_qp_unconverged": int(np.sum(f64.qp_iterations >= 200))},
        "qp_conditioning": {
            "spread_at_1e-14_rel_noise_median_mev": float(np.median(sp_tiny)),
            "spread_at_mixed_err_median_mev": float(np.median(sp_mix)),
            "resolvable_states": int(ok.sum()),
            "resolvable_within_1meV": int(np.sum(dq[ok] <= 1.0)),
            "resolvable_max_err_mev": float(dq[ok].max()) if ok.any() else None},
    }


def h2o_check(hip, ctx, npar: int, prec: int) -> dict:
    """The same report on H2O (nmo=18, the reference's regression input, kept
    as a fixture in tests/golden), whose QP energies are well conditioned."""
    from minigw_amd.synthetic import GwInput
    d = np.load(os.path.join(REPO, "tests", "golden", "h2o_input.npz"))
    inp = GwInput(mo_energy=d["mo_energy"], eri=d["eri"], vxc=d["vxc"], nocc=int(d["nocc"]),
                  fermi_energy=float(d["fermi_energy"]))
    ctx.upload_eri(inp.eri, inp.nocc)
    f64 = hip.run_g0w0(inp, 200, npar, ctx=ctx, upload=False, precision=hip.PREC_FP64)
    mix = hip.run_g0w0(inp, 200, npar, ctx=ctx, upload=False, precision=prec)
    return accuracy(hip, ctx, inp, mix, f64, npar)


def cpu_baseline(inp, nfreq: int, sample_freqs: int) -> dict:
    """Oracle (numpy + LAPACK/BLAS, all host cores) on `sample_freqs` of the
    molecule's imaginary frequencies, scaled to QP solves/s for all nfreq."""
    from threadpoolctl import threadpool_info
    from oracle import oracle as orc

    ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
    omegas, _ = orc.transformed_legendre_grid(nfreq)
    pick = np.linspace(0, nfreq - 1, sample_freqs).round().astype(int)
    t0 = time.perf_counter()
    orc.quadratic_forms_all(inp.eri, ph, omegas[pick], list(range(inp.nmo)), 0.0)
    dt = time.perf_counter() - t0
    per_freq = dt / sample_freqs
    threads = max([p.get("num_threads", 1) for p in threadpool_info()] or [1])
    return {"value": inp.nmo / (per_freq * nfreq), "unit": "QP solves/s", "cores": int(threads),
            "kind": "port",
            "sample": f"{sample_freqs} of {nfreq} frequencies (W_c by getrf/getrs + W_c X gemm, "
                      f"all {inp.nmo} states), {dt:.1f} s, scaled x{nfreq / sample_freqs:.0f}; "
                      f"Sigma_c/Pade/QP excluded (negligible)"}


def _factor_job(nmo: int, nocc: int, seed: int):
    from minigw_amd.synthetic import synthetic_factor
    mo, fermi, b, vxc = synthetic_factor(nmo, nocc, seed)
    return dict(nmo=nmo, nocc=nocc, mo=mo, fermi=fermi, b=b, vxc=vxc)


def run_batch(hip, count: int, workers: int, nfreq: int, npar: int, prec: int) -> dict:
    """Uneven-nmo batch: `count` molecules (nmo log-uniform in [24, 120],
    nocc = round(nmo / 5)), all submitted at once to the request loop."""
    from minigw_amd.synthetic import batch_molecules
    mols = [_factor_job(*m) for m in batch_molecules(count)]
    with hip.HipServer(0, workers=workers, hbm_budget=200 << 30, batch_freqs=64) as srv:
        # warm-up: compile-free, but first-touch allocations and graph capture
        t = srv.submit(factor=mols[0]["b"], mo_energy=mols[0]["mo"], vxc=mols[0]["vxc"],
                       nocc=mols[0]["nocc"], fermi=mols[0]["fermi"], nfreq=nfreq, npar=npar,
                       precision=prec)
        srv.wait(t)
        t0 = time.perf_counter()
        tickets = [srv.submit(factor=m["b"], mo_energy=m["mo"], vxc=m["vxc"], nocc=m["nocc"],
                              fermi=m["fermi"], nfreq=nfreq, npar=npar, precision=prec)
                   for m in mols]
        lat = [srv.wait(t)[1] for t in tickets]
        wall = time.perf_counter() - t0
        st = srv.stats()
    solves = sum(m["nmo"] for m in mols)
    return {"molecules": count, "nmo_range": [min(m["nmo"] for m in mols), max(m["nmo"] for m in mols)],
            "qp_solves_per_s": round(solves / wall, 2), "molecules_per_s": round(count / wall, 3),
            "wall_s": round(wall, 3), "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1), "workers": workers,
            "peak_concurrent_jobs": st["peak_running"]}


def run_serve(hip, jobs: int, rate: float, workers: int, npar: int, prec: int) -> dict:
    """Poisson arrivals at `rate` jobs/s; nmo log-uniform in [24, 200],
    nocc = round(nmo / 5), 100-200 frequencies; latency = arrival to result."""
    import math
    from minigw_amd.synthetic import SplitMix64, SERVE_SEED
    rng = SplitMix64(SERVE_SEED)
    u = rng.uniform(3 * jobs)
    specs = []
    t_arr = 0.0
    for i in range(jobs):
        t_arr += -math.log(1.0 - u[3 * i]) / rate
        nmo = int(round(math.exp(math.log(24) + u[3 * i + 1] * (math.log(200) - math.log(24)))))
        nfreq = 100 + int(u[3 * i + 2] * 101)
        specs.append((t_arr, nmo, max(1, int(round(nmo / 5))), nfreq, SERVE_SEED + 1 + i))
    data = [_factor_job(nmo, nocc, seed) for _, nmo, nocc, _, seed in specs]
    with hip.HipServer(0, workers=workers, hbm_budget=200 << 30, batch_freqs=32) as srv:
        t0 = time.perf_counter()
        tickets = []
        for (ta, nmo, nocc, nfreq, _), d in zip(specs, data):
            delay = ta - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(factor=d["b"], mo_energy=d["mo"], vxc=d["vxc"], nocc=nocc,
                                      fermi=d["fermi"], nfreq=nfreq, npar=npar, precision=prec))
        lat = [srv.wait(t)[1] for t in tickets]
        wall = time.perf_counter() - t0
        st = srv.stats()
    solves = sum(s[1] for s in specs)
    return {"jobs": jobs, "rate_per_s": rate, "qp_solves_per_s": round(solves / wall, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1), "wall_s": round(wall, 2),
            "workers": workers, "peak_concurrent_jobs": st["peak_running"],
            "nmo_range": [min(s[1] for s in specs), max(s[1] for s in specs)]}


FORMS = {
    # name: (precision code, contraction kernel, products per element, vendor peak,
    #        measured-peak key, dtype label)
    "mixed": (1, "k_contract_split<FP16>", 3, 2.5e15, "fp16_mfma_32x32", "fp16x3-split+fp64"),
    "mixed_int8": (2, "k_contract_i8", 15, 5.0e15, "int8_mfma_32x32", "int8x5-slice+fp64"),
    "fp64": (0, "k_contract_f64", 1, 78.6e12, "fp64_mfma", "fp64"),
    # FP64 with the ERI decomposed on the device (pair-space screening, rank r)
    "fp64_cd": (3, "k_contract_f64", 1, 78.6e12, "fp64_mfma", "fp64"),
}


RANDOM_PEAK = {"mixed_int8": "int8_mfma_32x32_random", "fp64": "fp64_mfma_random",
               "fp64_cd": "fp64_mfma_random"}


def roofline(form: str, ph: dict, n: int, m: int, nfreq: int, config: str, peaks: dict) -> dict:
    """Contraction kernel: algorithmic flops (or int8 ops) per launch over its
    average launch time (HIP events around the kernel on the context's stream)."""
    code, kernel, products, peak, key, _ = FORMS[form]
    launches = max(ph["contract_launches"], 1.0)
    t_launch = ph["contract"] / launches * 1e-3
    cols = ph.get("contract_columns", m) or m      # unordered (p,k) pairs when the ERI is symmetric
    # launches are nb frequencies each except the last: average work per launch;
    # the decomposed form contracts over the rank r of the ERI's pair matrix
    k = int(ph["cd_rank"]) if form == "fp64_cd" else n
    work = contraction_flops(k, cols, 0) * products * nfreq / launches
    achieved = work / t_launch / 1e12
    return {"bound": "mfma", "kernel": kernel, "achieved": round(achieved, 2),
            "peak": peak / 1e12, "unit": "TOP/s" if form == "mixed_int8" else "TFLOP/s",
            "frac": round(achieved * 1e12 / peak, 4),
            "traffic": TRAFFIC.get((form, config, nfreq)),
            "launch_ms": round(t_launch * 1e3, 3), "work_per_launch": work,
            "measured_peak": round(peaks[key] / 1e12, 1),
            # the same MFMA on random operands (the rate real data allows: the
            # chip's MFMA throughput depends on the operand values)
            "measured_peak_random_operands": (round(peaks[RANDOM_PEAK[form]] / 1e12, 1)
                                              if form in RANDOM_PEAK else None)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="nmo200", choices=["nmo24", "nmo120", "nmo200"])
    ap.add_argument("--nfreq", type=int, default=200)
    ap.add_argument("--npar", type=int, default=16)
    ap.add_argument("--precision", default="mixed_int8", choices=list(FORMS))
    ap.add_argument("--alt-precision", default="mixed", choices=list(FORMS) + ["none"],
                    help="second mixed form, timed and checked against FP64 beside the headline one")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cd-steps", type=int, default=3,
                    help="steps of the decomposed-ERI FP64 form reported beside (0: skip)")
    ap.add_argument("--cpu-sample-freqs", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-molecules", type=int, default=32)
    ap.add_argument("--serve-jobs", type=int, default=12)
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[1.0])
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(frequency sharding across GPUs is a design note only)"}))
        return 2

    from minigw_amd import hip
    from minigw_amd.synthetic import CONFIGS, synthetic_input

    nmo, nocc, seed = CONFIGS[args.config]
    t0 = time.perf_counter()
    inp = synthetic_input(nmo, nocc, seed)
    log(f"bench: generated {args.config} nmo={nmo} nocc={nocc} in {time.perf_counter() - t0:.1f} s")
    prec = FORMS[args.precision][0]
    alt = None if args.alt_precision in ("none", args.precision, "fp64") else args.alt_precision
    n, m = inp.nph, nmo * nmo

    with hip.HipContext(0) as ctx:
        t0 = time.perf_counter()
        ctx.upload_eri(inp.eri, inp.nocc)
        log(f"bench: ERI resident ({inp.eri.nbytes / 1e9:.2f} GB) in {time.perf_counter() - t0:.1f} s")

        peaks = ctx.measure_peaks()
        log("bench: measured peaks " + ", ".join(f"{k} {v / 1e12:.1f}T" for k, v in peaks.items()))

        def step(p):
            return hip.run_g0w0(inp, args.nfreq, args.npar, precision=p, ctx=ctx, upload=False)

        if args.autotune:
            tb, tms = hip.autotune(inp, args.nfreq, args.npar, prec, ctx=ctx)
            log(f"bench: autotuned frequency batch {tb} ({tms:.2f} ms per frequency)")
        for i in range(args.warmup):
            step(prec)
            log(f"bench: warmup {i + 1}/{args.warmup} done")
        lat, phases, res = [], [], None
        t_start = time.perf_counter()
        for i in range(args.steps):
            t1 = time.perf_counter()
            res = step(prec)
            lat.append(time.perf_counter() - t1)
            phases.append(ctx.last_timings())
            log(f"bench: step {i + 1}/{args.steps} {lat[-1] * 1e3:.1f} ms")
        elapsed = time.perf_counter() - t_start

        def timed(p, steps, what):
            r, ts = None, []
            for i in range(steps):
                t1 = time.perf_counter()
                r = step(p)
                ts.append(time.perf_counter() - t1)
                log(f"bench: {what} step {i + 1}/{steps} {ts[-1] * 1e3:.1f} ms")
            return r, ts, ctx.last_timings()

        # plain FP64 form: error reference and ratio; the other mixed form beside it
        f64 = alt_res = None
        if prec != 0 and args.fp64_steps > 0:
            f64, f64_lat, f64_phase = timed(0, args.fp64_steps, "fp64")
            if alt:
                step(FORMS[alt][0])
                alt_res, alt_lat, alt_phase = timed(FORMS[alt][0], 1, alt)
        cd_res = None
        if prec != 3 and args.cd_steps > 0 and f64 is not None:
            step(3)
            cd_res, cd_lat, cd_phase = timed(3, args.cd_steps, "fp64_cd")
        report = {}
        if f64 is not None:
            log("bench: accuracy report ...")
            report["accuracy_vs_fp64"] = accuracy(hip, ctx, inp, res, f64, args.npar)
            if alt_res is not None:
                report["alt_accuracy"] = accuracy(hip, ctx, inp, alt_res, f64, args.npar)
            report["accuracy_vs_fp64_h2o"] = h2o_check(hip, ctx, args.npar, prec)
            if alt_res is not None:
                report["alt_accuracy_h2o"] = h2o_check(hip, ctx, args.npar, FORMS[alt][0])
            if cd_res is not None:
                report["cd_accuracy"] = accuracy(hip, ctx, inp, cd_res, f64, args.npar)
                report["cd_accuracy_h2o"] = h2o_check(hip, ctx, args.npar, 3)

    ms = elapsed / args.steps * 1e3
    value = nmo * args.steps / elapsed
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "QP solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": FORMS[args.precision][5],
        "data": "synthetic (seeded SplitMix64 generator)",
        "config": {"workload": f"{args.config}: synthetic nmo={nmo} nocc={nocc} nph={n}, "
                               f"{args.nfreq} GL freqs, {args.npar} Pade, all states, "
                               f"{args.precision}", "nmo": nmo, "nocc": nocc, "nph": n,
                   "nfreq": args.nfreq, "npar": args.npar, "precision": args.precision},
        "p50_ms": round(float(np.percentile(np.array(lat) * 1e3, 50)), 2),
        "p99_ms": round(float(np.percentile(np.array(lat) * 1e3, 99)), 2),
        "phases_ms": {k: round(v, 3) for k, v in ph.items()},
        "roofline": roofline(args.precision, ph, n, m, args.nfreq, args.config, peaks),
        "measured_peaks_tflops_or_tbps": {k: round(v / 1e12, 2) for k, v in peaks.items()},
    }
    if f64 is not None:
        fv = nmo * len(f64_lat) / sum(f64_lat)
        out["fp64"] = {"value": round(fv, 3), "ms_per_step": round(float(np.mean(f64_lat)) * 1e3, 2),
                       "phases_ms": {k: round(v, 3) for k, v in f64_phase.items()},
                       "roofline": roofline("fp64", f64_phase, n, m, args.nfreq, args.config, peaks)}
        out["mixed_over_fp64"] = round(value / fv, 3)
        if alt_res is not None:
            av = nmo * len(alt_lat) / sum(alt_lat)
            out["alt"] = {"precision": alt, "value": round(av, 3), "over_fp64": round(av / fv, 3),
                          "ms_per_step": round(float(np.mean(alt_lat)) * 1e3, 2),
                          "phases_ms": {k: round(v, 3) for k, v in alt_phase.items()},
                          "roofline": roofline(alt, alt_phase, n, m, args.nfreq, args.config, peaks)}
        if cd_res is not None:
            cv = nmo * len(cd_lat) / sum(cd_lat)
            out["decomposed_fp64"] = {
                "precision": "fp64_cd", "value": round(cv, 3), "over_fp64": round(cv / fv, 3),
                "ms_per_step": round(float(np.mean(cd_lat)) * 1e3, 2),
                "p50_ms": round(float(np.percentile(np.array(cd_lat) * 1e3, 50)), 2),
                "rank": int(cd_phase["cd_rank"]),
                "phases_ms": {k: round(v, 3) for k, v in cd_phase.items()},
                "roofline": roofline("fp64_cd", cd_phase, n, m, args.nfreq, args.config, peaks)}
        out.update(report)
    cd_too = prec != 3 and args.cd_steps > 0          # the same workloads in the decomposed form
    if args.batch_molecules > 0:
        log("bench: uneven-nmo batch ...")
        out["batch"] = run_batch(hip, args.batch_molecules, args.workers, args.nfreq, args.npar, prec)
        if cd_too:
            out["batch_decomposed"] = run_batch(hip, args.batch_molecules, args.workers, args.nfreq,
                                                args.npar, 3)
    if args.serve_jobs > 0:
        out["serve"] = []
        for rate in args.serve_rates:
            log(f"bench: serving at {rate} jobs/s ...")
            out["serve"].append(run_serve(hip, args.serve_jobs, rate, args.workers, args.npar, prec))
        if cd_too:
            out["serve_decomposed"] = [run_serve(hip, args.serve_jobs, rate, args.workers, args.npar, 3)
                                       for rate in args.serve_rates]
    if not args.no_cpu:
        log("bench: cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(inp, args.nfreq, args.cpu_sample_freqs)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
