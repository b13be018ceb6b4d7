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


def contraction_flops(n: int, m: int) -> float:
    """Algorithmic flops of one frequency's contraction q_j = -||(L^-1 S) x_j||^2,
    L^-1 lower triangular: 2 * (n(n+1)/2) * m for the m contracted columns
    (the unordered (p,k) pairs when q(p,k) = q(k,p)).  One product whatever
    the arithmetic: the split forms' extra slice products are not useful work."""
    return float(n) * (n + 1) * m


def contraction_bytes(n: int, m: int, nb: float, form: str) -> float:
    """Algorithmic HBM bytes of one contraction launch over nb frequencies: the
    operands once each in the form's representation (the triangle of M per
    frequency, the panel X once per launch) and the per-row-tile partial sums
    written (128-row tiles)."""
    per = {"mixed_int8": 5.0, "mixed_int8x4": 4.0, "mixed": 4.0, "fp64": 8.0, "fp64_cd": 8.0}[form]
    return per * (n * (n + 1) / 2 * nb + n * m) + 8.0 * (n / 128) * m * nb


HA_MEV = 27.211386245988 * 1e3

# HBM bytes per contraction launch (FETCH_SIZE + WRITE_SIZE, each scaled by
# what the counter reads for a 2 GiB copy), averaged over the molecule's
# launches, from two rocprofv3 --pmc runs of this bench with that form alone;
# the commands and per-dispatch figures are in profiles/kernels_nmo200_*.md
# (int8x4, int8, fp64, fp16split, fp64_cd).
TRAFFIC = {("mixed_int8x4", "nmo200", 200): (794.78e9 + 0.78e9),
           ("mixed_int8", "nmo200", 200): (985.16e9 + 0.78e9),
           ("fp64", "nmo200", 200): (2844.54e9 + 0.78e9),
           ("mixed", "nmo200", 200): (266.04e9 + 0.39e9),
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


def host_threads() -> int:
    """Host threads the CPU baseline uses: the job's OpenMP share when set, else nproc."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_port(inp, nfreq: int, sample_freqs: int, threads: int) -> dict:
    """The oracle (numpy + LAPACK/BLAS, `threads` host threads) on `sample_freqs`
    of the molecule's imaginary frequencies in the hoisted schedule (each w'
    once), scaled to QP solves/s for all nfreq."""
    from threadpoolctl import threadpool_limits
    from oracle import oracle as orc

    ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
    omegas, _ = orc.transformed_legendre_grid(nfreq)
    pick = np.linspace(0, nfreq - 1, sample_freqs).round().astype(int)
    with threadpool_limits(threads):
        t0 = time.perf_counter()
        orc.quadratic_forms_all(inp.eri, ph, omegas[pick], list(range(inp.nmo)), 0.0)
        dt = time.perf_counter() - t0
    per_freq = dt / sample_freqs
    return {"value": inp.nmo / (per_freq * nfreq), "unit": "QP solves/s", "cores": threads,
            "kind": "port", "schedule": "hoisted (each w' once)",
            "sample": f"{sample_freqs} of {nfreq} frequencies (W_c by getrf/getrs + W_c X gemm, "
                      f"all {inp.nmo} states), {dt:.1f} s, scaled x{nfreq / sample_freqs:.0f}; "
                      f"Sigma_c/Pade/QP excluded (negligible)"}


def cpu_baseline(config: str, inp, nfreq: int, sample_freqs: int) -> dict:
    """The reference's own CPU path on the host cores: oracle/_ref/gw_ref_blas
    (the reference's run_g0w0 with its BLAS/LAPACK backend, frequencies serial,
    OpenMP kernel loops and LAPACK threads = the cores used) on the same
    molecule at 1 frequency, i.e. one of the F^2 identical (w_n, w') units its
    schedule repeats, extrapolated to F = nfreq exactly in operation count
    (oracle/reference_timing.py).  value = the reference's own F^2 schedule;
    the same units done once per w' (the hoisted schedule the device runs)
    and the numpy port are reported beside it.  Falls back to the port when
    the reference binary or its LAPACK provider is missing."""
    from minigw_amd.synthetic import CONFIGS
    from oracle import reference_timing as rt

    threads = host_threads()
    port = cpu_port(inp, nfreq, sample_freqs, threads)
    if not rt.available():
        port["note"] = "reference binary oracle/_ref/gw_ref_blas not built: the port instead"
        return port
    nmo, nocc, seed = CONFIGS[config]
    try:
        r = rt.run(nmo, nocc, seed, 1, threads, timeout=600)
    except Exception as e:                      # noqa: BLE001 (reported, not hidden)
        port["note"] = f"reference binary failed ({e}); the port instead"
        return port
    ex = rt.extrapolate(r, nfreq)
    return {"value": ex["reference_qp_solves_per_s"], "unit": "QP solves/s", "cores": threads,
            "kind": "reference", "schedule": "reference (W_c recomputed for every (w_n, w'): F^2 units)",
            "sample": f"1 of the {nfreq}^2 (w_n, w') units of the reference's BLAS/LAPACK backend "
                      f"on this molecule ({r['unit_s']:.1f} s: complex zgetrf + zgetrs against I, "
                      f"zgemm panels of all {nmo} states), + V_ph setup {r['setup_s']:.2f} s, "
                      f"extrapolated to {nfreq} frequencies ({ex['reference_schedule_s']:.3g} s)",
            "unit_s": r["unit_s"],
            "hoisted_value": ex["hoisted_qp_solves_per_s"],
            "hoisted_note": "the reference's units done once per w' (F units), the device's schedule",
            "port": port}


def storage_check(hip, inp, nfreq: int, npar: int, prec: int, res, resident_upload_s: float) -> dict:
    """The headline molecule with its ERI in the streamed storage (the ERI
    crosses to the device once in slabs of rows p; only the (pk|ia) block and
    (pi|pi) stay): bytes held, upload time, one molecule's time after a
    warm-up, and its result against the resident storage's, bit for bit."""
    with hip.HipContext(0) as c:
        c.set_eri_storage(hip.ERI_STREAMED)
        t0 = time.perf_counter()
        c.upload_eri(inp.eri, inp.nocc)
        upload_s = time.perf_counter() - t0
        storage, held = c.eri_storage()
        hip.run_g0w0(inp, nfreq, npar, precision=prec, ctx=c, upload=False)
        t1 = time.perf_counter()
        r = hip.run_g0w0(inp, nfreq, npar, precision=prec, ctx=c, upload=False)
        ms = (time.perf_counter() - t1) * 1e3
    same = bool(np.array_equal(r.sigma_c, res.sigma_c) and np.array_equal(r.qp, res.qp, equal_nan=True)
                and np.array_equal(r.qp_iterations, res.qp_iterations))
    return {"storage": hip.ERI_STORAGE_NAMES[storage], "resident_eri_bytes": int(inp.eri.nbytes),
            "streamed_eri_bytes": int(held),
            "upload_s": {"resident": round(resident_upload_s, 3), "streamed": round(upload_s, 3)},
            "ms_per_molecule_streamed": round(ms, 2), "bitwise_equal_to_resident": same}


def _factor_job(nmo: int, nocc: int, seed: int):
    from minigw_amd.synthetic import synthetic_factor
    mo, fermi, b, vxc = synthetic_factor(nmo, nocc, seed)
    return dict(nmo=nmo, nocc=nocc, mo=mo, fermi=fermi, b=b, vxc=vxc)


def run_batch(hip, count: int, workers: int, nfreq: int, npar: int, prec: int,
              grouping: int = 16, reps: int = 3) -> dict:
    """Uneven-nmo batch: `count` molecules (nmo log-uniform in [24, 120],
    nocc = round(nmo / 5)), all submitted at once to the request loop, whose
    workers group molecules of one padded order (up to `grouping` per group;
    1: every molecule alone on its worker).  The batch runs `reps` times, each
    on a new server after one warm-up job (batches run back to back on one
    server have taken longer each time on some boxes); the figures are those
    of the repetition with the median wall time, and every repetition's wall
    time is listed."""
    from minigw_amd.synthetic import batch_molecules
    mols = [_factor_job(*m) for m in batch_molecules(count)]
    runs = []
    for _ in range(reps):
        with hip.HipServer(0, workers=workers, hbm_budget=200 << 30, batch_freqs=64) as srv:
            srv.set_grouping(grouping)
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
            runs.append((time.perf_counter() - t0, lat))
            st = srv.stats()
    wall, lat = sorted(runs, key=lambda r: r[0])[len(runs) // 2]
    solves = sum(m["nmo"] for m in mols)
    return {"molecules": count, "nmo_range": [min(m["nmo"] for m in mols), max(m["nmo"] for m in mols)],
            "qp_solves_per_s": round(solves / wall, 2), "molecules_per_s": round(count / wall, 3),
            "wall_s": round(wall, 3), "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1), "workers": workers,
            "grouping": grouping, "peak_concurrent_jobs": st["peak_running"],
            "repetitions": reps, "wall_s_each": [round(r[0], 3) for r in runs]}


def run_serve(hip, jobs: int, rates, workers: int, npar: int, prec: int) -> list:
    """Poisson job streams (minigw_amd.synthetic.serve_jobs: nmo log-uniform
    in [24, 200], nocc = round(nmo / 5), 100-200 frequencies) at each rate of
    `rates` jobs/s, the same jobs at every rate; latency = arrival to result.
    Mixed jobs run under the loop's error judge with FP64 fallback."""
    from minigw_amd.synthetic import serve_jobs
    base = serve_jobs(jobs, 1.0)
    data = [_factor_job(nmo, nocc, seed) for _, nmo, nocc, _, seed in base]
    out = []
    for rate in rates:
        log(f"bench: serving {jobs} jobs at {rate} jobs/s ...")
        with hip.HipServer(0, workers=workers, hbm_budget=200 << 30, batch_freqs=32) as srv:
            t0 = time.perf_counter()
            tickets = []
            for (ta, nmo, nocc, nfreq, _), d in zip(base, data):
                delay = ta / rate - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                tickets.append(srv.submit(factor=d["b"], mo_energy=d["mo"], vxc=d["vxc"], nocc=nocc,
                                          fermi=d["fermi"], nfreq=nfreq, npar=npar, precision=prec))
            done = [srv.wait(t) for t in tickets]
            wall = time.perf_counter() - t0
            st = srv.stats()
        lat = [d[1] for d in done]
        jerr = [d[0].timings["judge_err"] for d in done]
        solves = sum(s[1] for s in base)
        out.append({"jobs": jobs, "rate_per_s": rate, "qp_solves_per_s": round(solves / wall, 2),
                    "jobs_per_s": round(jobs / wall, 3),
                    "p50_ms": round(float(np.percentile(lat, 50)), 1),
                    "p99_ms": round(float(np.percentile(lat, 99)), 1), "wall_s": round(wall, 2),
                    "workers": workers, "peak_concurrent_jobs": st["peak_running"],
                    "nmo_range": [min(s[1] for s in base), max(s[1] for s in base)],
                    "judged_jobs": int(sum(e >= 0 for e in jerr)),
                    "judge_max_rel_err": max(jerr), "fp64_fallbacks":
                        int(sum(d[0].timings["fell_back"] != 0 for d in done))})
    return out


FORMS = {
    # name: (precision code, contraction kernel, products per element, vendor peak,
    #        measured-peak key, dtype label)
    "mixed": (1, "k_contract_split<FP16>", 3, 2.5e15, "fp16_mfma_32x32", "fp16x3-split+fp64"),
    "mixed_int8": (2, "k_contract_i8<5, 7>", 15, 5.0e15, "int8_mfma_32x32", "int8x5-slice+fp64"),
    # four balanced 8-bit digits per operand: 10 digit products
    "mixed_int8x4": (4, "k_contract_i8<4, 8>", 10, 5.0e15, "int8_mfma_32x32", "int8x4-digit+fp64"),
    "fp64": (0, "k_contract_f64", 1, 78.6e12, "fp64_mfma", "fp64"),
    # FP64 with the ERI decomposed on the device (pair-space screening, rank r)
    "fp64_cd": (3, "k_contract_f64", 1, 78.6e12, "fp64_mfma", "fp64"),
}


RANDOM_PEAK = {"mixed_int8": "int8_mfma_32x32_random", "mixed_int8x4": "int8_mfma_32x32_random",
               "fp64": "fp64_mfma_random",
               "fp64_cd": "fp64_mfma_random"}


def roofline(form: str, ph: dict, n: int, m: int, nfreq: int, config: str, peaks: dict) -> dict:
    """Contraction kernel: algorithmic flops per launch (one product,
    n(n+1) m per frequency on the executed columns) over its average launch
    time (HIP events around the kernel on the context's stream), against the
    vendor dense peak of the instruction the form issues (int8 5.0 POP/s,
    FP16 2.5 PFLOP/s, FP64 78.6 TFLOP/s).  mfma_pipe_frac counts every slice
    product the form issues (15 int8, 3 FP16): how busy the MFMA pipe is, not
    useful work.  traffic_over_algorithmic: measured HBM bytes per launch over
    the operands' bytes read once."""
    code, kernel, products, peak, key, _ = FORMS[form]
    launches = max(ph["contract_launches"], 1.0)
    t_launch = ph["contract"] / launches * 1e-3
    cols = ph.get("contract_columns", m) or m      # unordered (p,k) pairs when the ERI is symmetric
    # launches are nb frequencies each except the last: average work per launch;
    # the decomposed form contracts over the rank r of the ERI's pair matrix
    k = int(ph["cd_rank"]) if form == "fp64_cd" else n
    nb = nfreq / launches
    work = contraction_flops(k, cols) * nb
    achieved = work / t_launch
    traffic = TRAFFIC.get((form, config, nfreq))
    alg_bytes = contraction_bytes(k, cols, nb, form)
    return {"bound": "mfma", "kernel": kernel, "achieved": round(achieved / 1e12, 2),
            "peak": peak / 1e12, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": traffic, "algorithmic_bytes": alg_bytes,
            "traffic_over_algorithmic": round(traffic / alg_bytes, 1) if traffic else None,
            "mfma_pipe_frac": round(achieved * products / peak, 4),
            "products_per_element": products,
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
    ap.add_argument("--precision", default="mixed_int8x4", choices=list(FORMS))
    ap.add_argument("--alt-precision", default="mixed_int8", choices=list(FORMS) + ["none"],
                    help="second mixed form, timed and checked against FP64 beside the headline one")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cd-steps", type=int, default=3,
                    help="steps of the decomposed-ERI FP64 form reported beside (0: skip)")
    ap.add_argument("--cpu-sample-freqs", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-molecules", type=int, default=32)
    ap.add_argument("--serve-jobs", type=int, default=50)
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[1.0, 3.0, 10.0])
    ap.add_argument("--decomposed-workloads", action="store_true",
                    help="also run the batch and serving workloads in the decomposed-ERI form")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-storage-check", action="store_true",
                    help="skip the headline molecule in the streamed ERI storage")
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
        ctx.set_eri_storage(hip.ERI_RESIDENT)
        ctx.upload_eri(inp.eri, inp.nocc)
        resident_upload_s = time.perf_counter() - t0
        log(f"bench: ERI resident ({inp.eri.nbytes / 1e9:.2f} GB) in {resident_upload_s:.1f} s")

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
    if not args.no_storage_check:
        log("bench: streamed ERI storage ...")
        out["eri_storage"] = storage_check(hip, inp, args.nfreq, args.npar, prec, res,
                                           resident_upload_s)
    cd_too = prec != 3 and args.decomposed_workloads  # the same workloads in the decomposed form
    if args.batch_molecules > 0:
        log("bench: uneven-nmo batch ...")
        out["batch"] = run_batch(hip, args.batch_molecules, args.workers, args.nfreq, args.npar, prec)
        out["batch_ungrouped"] = run_batch(hip, args.batch_molecules, args.workers, args.nfreq,
                                           args.npar, prec, grouping=1)
        if cd_too:
            out["batch_decomposed"] = run_batch(hip, args.batch_molecules, args.workers, args.nfreq,
                                                args.npar, 3)
    if args.serve_jobs > 0:
        out["serve"] = run_serve(hip, args.serve_jobs, args.serve_rates, args.workers, args.npar, prec)
        if cd_too:
            out["serve_decomposed"] = run_serve(hip, args.serve_jobs, args.serve_rates, args.workers,
                                                args.npar, 3)
    if not args.no_cpu:
        log("bench: cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.config, inp, args.nfreq, args.cpu_sample_freqs)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
