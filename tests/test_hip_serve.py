"""Request loop and device-side ERI construction (needs an MI355X)."""
import time

import numpy as np
import pytest

from minigw_amd.synthetic import GwInput, eri_from_factor, synthetic_factor, synthetic_input, with_gap
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hip():
    from minigw_amd import hip as h
    return h


def test_eri_from_factor_on_device(hip):
    say("case: ERI built on the device from its factor, nmo=14")
    nmo, nocc = 14, 4
    mo, fermi, b, vxc = synthetic_factor(nmo, nocc, 77)
    host = eri_from_factor(b, nmo)
    with hip.HipContext(0) as ctx:
        ctx.upload_eri_factor(b, nmo, nocc)
        dev = ctx.download_eri(nmo)
        assert np.abs(dev - host).max() <= 1e-14 * np.abs(host).max()
        for perm in [(1, 0, 2, 3), (2, 3, 0, 1), (0, 1, 3, 2)]:
            assert np.array_equal(dev, dev.transpose(perm))
        inp = GwInput(mo_energy=mo, eri=host, vxc=vxc, nocc=nocc, fermi_energy=fermi)
        res = hip.run_g0w0(inp, 32, 8, ctx=ctx, upload=False)
        assert res.timings["contract_columns"] == nmo * (nmo + 1) // 2   # symmetric path taken
    want = orc.run_g0w0(inp, 32, 8)
    assert np.abs(res.sigma_c - want.sigma_c).max() < 1e-11 * np.abs(want.sigma_c).max()


def _jobs():
    out = []
    for i, (nmo, nocc) in enumerate([(12, 3), (20, 4), (28, 6), (16, 4), (36, 7), (24, 5)]):
        mo, fermi, b, vxc = synthetic_factor(nmo, nocc, 1000 + i)
        out.append(dict(nmo=nmo, nocc=nocc, mo=mo, fermi=fermi, b=b, vxc=vxc,
                        precision=(0, 1, 3, 2, 0, 3)[i]))
    return out


def test_serve_alone_equals_mix(hip):
    """Each job's result in a mix of concurrent jobs equals the same job alone."""
    say("case: 6 jobs (nmo 12..36, FP64, mixed, int8 and decomposed) on 4 workers vs alone")
    jobs = _jobs()
    with hip.HipServer(0, workers=4, hbm_budget=64 << 30, batch_freqs=16) as srv:
        tickets = [srv.submit(factor=j["b"], mo_energy=j["mo"], vxc=j["vxc"], nocc=j["nocc"],
                              fermi=j["fermi"], nfreq=40, npar=12, precision=j["precision"])
                   for j in jobs]
        mixed = [srv.wait(t) for t in tickets]
        st = srv.stats()
    say(f"   server stats {st}")
    assert st["completed"] == len(jobs) and st["running"] == 0 and st["peak_running"] >= 2
    with hip.HipContext(0) as ctx:
        for j, (res, lat, run) in zip(jobs, mixed):
            ctx.upload_eri_factor(j["b"], j["nmo"], j["nocc"])
            inp = GwInput(mo_energy=j["mo"], eri=None, vxc=j["vxc"], nocc=j["nocc"],
                          fermi_energy=j["fermi"])
            alone = hip.run_g0w0(inp, 40, 12, ctx=ctx, upload=False, precision=j["precision"])
            assert np.array_equal(alone.sigma_c, res.sigma_c), j["nmo"]
            assert np.array_equal(alone.qp, res.qp, equal_nan=True)
            assert np.array_equal(alone.qp_iterations, res.qp_iterations)
            assert lat >= run > 0


def test_serve_admission_and_rejection(hip):
    say("case: admission by HBM estimate")
    jobs = _jobs()[:3]
    est = max(hip.estimate_bytes(j["nmo"], j["nocc"], 40, 0, 16, j["precision"]) for j in jobs)
    budget = int(est * 1.5)
    with hip.HipServer(0, workers=3, hbm_budget=budget, batch_freqs=16) as srv:
        tickets = [srv.submit(factor=j["b"], mo_energy=j["mo"], vxc=j["vxc"], nocc=j["nocc"],
                              fermi=j["fermi"], nfreq=40, npar=12, precision=j["precision"])
                   for j in jobs]
        for t in tickets:
            srv.wait(t)
        assert srv.stats()["peak_running"] >= 1
        big = synthetic_input(64, 12, 5)
        assert hip.estimate_bytes(64, 12, 40, 0, 16, hip.PREC_FP64) > budget
        t = srv.submit(big, nfreq=40, npar=12)
        with pytest.raises(hip.HipError) as e:
            srv.wait(t)
        assert e.value.code == hip.MGW_ERR_MEMORY


def test_graph_replays_match_eager(hip):
    """Frequency batches captured as hipGraphs and replayed (within a molecule
    and across repeated molecules, with eager work on another context in
    between) give exactly the eager results."""
    say("case: hipGraph replay vs eager, nmo=30, 150 frequencies (3 batches)")
    inp = synthetic_input(30, 6, 41)
    other = synthetic_input(18, 4, 42)
    with hip.HipContext(0) as ctx, hip.HipContext(0) as side:
        ctx.upload_eri(inp.eri, inp.nocc)
        side.upload_eri(other.eri, other.nocc)
        for prec in (hip.PREC_FP64, hip.PREC_MIXED, hip.PREC_MIXED_INT8, hip.PREC_MIXED_INT8_4,
                     hip.PREC_FP64_CD):
            ctx.set_graphs(False)
            eager = hip.run_g0w0(inp, 150, 16, ctx=ctx, upload=False, precision=prec)
            side_ref = hip.run_g0w0(other, 40, 8, ctx=side, upload=False, precision=prec)
            ctx.set_graphs(True)
            for rep in range(3):
                res = hip.run_g0w0(inp, 150, 16, ctx=ctx, upload=False, precision=prec)
                assert res.timings["graph_batches"] > 0
                assert np.array_equal(res.sigma_c, eager.sigma_c), (prec, rep)
                assert np.array_equal(res.qp, eager.qp, equal_nan=True)
                s2 = hip.run_g0w0(other, 40, 8, ctx=side, upload=False, precision=prec)
                assert np.array_equal(s2.sigma_c, side_ref.sigma_c)


def test_graph_cache_tells_molecules_of_one_shape_apart(hip):
    """Two molecules with the same nmo and a different nocc pad to the same
    orders (n = 144 and 161, both padded to 256) and reuse the same buffers on
    one context: the second must not replay the first's captured batches,
    which carry n as a kernel argument."""
    say("case: graphs on one context, nmo=30 nocc=6 then nocc=7 (same padded shapes)")
    first, second = synthetic_input(30, 6, 45), synthetic_input(30, 7, 46)
    for prec in (hip.PREC_FP64, hip.PREC_MIXED_INT8, hip.PREC_MIXED_INT8_4, hip.PREC_FP64_CD):
        with hip.HipContext(0) as ctx:
            ctx.upload_eri(second.eri, second.nocc)
            eager = hip.run_g0w0(second, 100, 16, ctx=ctx, upload=False, precision=prec)
            ctx.set_graphs(True)
            ctx.upload_eri(first.eri, first.nocc)
            hip.run_g0w0(first, 100, 16, ctx=ctx, upload=False, precision=prec)
            ctx.upload_eri(second.eri, second.nocc)
            res = hip.run_g0w0(second, 100, 16, ctx=ctx, upload=False, precision=prec)
            assert res.timings["graph_batches"] > 0
            assert np.array_equal(res.sigma_c, eager.sigma_c), prec
            assert np.array_equal(res.qp, eager.qp, equal_nan=True), prec


def test_autotune_picks_a_batch_and_keeps_results(hip):
    say("case: frequency-batch and unit-shape autotuner, nmo=30")
    inp = synthetic_input(30, 6, 43)
    with hip.HipContext(0) as ctx:
        ctx.upload_eri(inp.eri, inp.nocc)
        before = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False)
        nb, ms = hip.autotune(inp, 100, 16, ctx=ctx)
        assert nb in (8, 16, 32, 64) and ms > 0
        assert hip.autotune_result(inp, 100, 16) == (nb, 0)      # FP64: no unit shape
        after = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False)
        batches = -(-100 // nb)                                    # equal batches of the cap
        assert after.timings["freq_batch"] == -(-100 // batches)
        assert np.array_equal(after.sigma_c, before.sigma_c)
        # mixed form: the unit shape changes the tile order only, not the sums
        p = hip.PREC_MIXED_INT8
        before = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=p)
        nb, ms = hip.autotune(inp, 100, 16, p, ctx=ctx)
        tb, uf = hip.autotune_result(inp, 100, 16, p)
        assert tb == nb and uf in (2, 4, 8, 16)
        after = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=p)
        assert np.array_equal(after.sigma_c, before.sigma_c)
        # decomposed form: pair-space batches up to 256 frequencies, no unit shape
        p = hip.PREC_FP64_CD
        before = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=p)
        nb, ms = hip.autotune(inp, 100, 16, p, ctx=ctx)
        assert nb in (16, 32, 64, 128, 256) and ms > 0
        assert hip.autotune_result(inp, 100, 16, p) == (nb, 0)
        after = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=p)
        batches = -(-100 // nb)                                    # equal batches of the cap
        assert after.timings["freq_batch"] == -(-100 // batches)
        assert np.array_equal(after.sigma_c, before.sigma_c)


def test_split_factorization_matches_one_stream(hip, monkeypatch):
    """The factorization split over two (and four) streams, forced on a small
    molecule, gives bit for bit the one-stream results; with graphs on, the
    batches that split run eagerly (a graph captured across the side streams
    keeps the memory of the buffers it referenced after they are freed), and
    device memory in use returns to its level after the context's buffers are
    released."""
    say("case: factorization on 1 / 2 / 4 streams, nmo=30, eager; graphs off for split batches")
    inp = synthetic_input(30, 6, 44)
    base = {}
    monkeypatch.setenv("MGW_FACTOR_STREAMS", "1")
    with hip.HipContext(0) as ctx:
        ctx.upload_eri(inp.eri, inp.nocc)
        for prec in (hip.PREC_FP64, hip.PREC_MIXED_INT8, hip.PREC_MIXED_INT8_4):
            base[prec] = hip.run_g0w0(inp, 150, 16, ctx=ctx, upload=False, precision=prec)
    monkeypatch.setenv("MGW_FACTOR_SPLIT_ORDER", "0")
    for streams in ("2", "4"):
        monkeypatch.setenv("MGW_FACTOR_STREAMS", streams)
        with hip.HipContext(0) as ctx:
            ctx.upload_eri(inp.eri, inp.nocc)
            for prec in (hip.PREC_FP64, hip.PREC_MIXED_INT8, hip.PREC_MIXED_INT8_4):
                ctx.set_graphs(False)
                eager = hip.run_g0w0(inp, 150, 16, ctx=ctx, upload=False, precision=prec)
                assert np.array_equal(eager.sigma_c, base[prec].sigma_c), (streams, prec)
                ctx.set_graphs(True)
                for rep in range(3):
                    res = hip.run_g0w0(inp, 150, 16, ctx=ctx, upload=False, precision=prec)
                    assert res.timings["graph_batches"] == 0
                    assert np.array_equal(res.sigma_c, base[prec].sigma_c), (streams, prec, rep)
                    assert np.array_equal(res.qp, base[prec].qp, equal_nan=True)


@pytest.mark.parametrize("form", ["int8", "int8x4"])
def test_error_judge_passes_and_forced_fallback_gives_fp64(hip, form):
    """The judge of an int8 form measures its q error on sampled frequencies
    (~1e-9 / ~1e-8, below the 1e-7 / 1e-6 defaults) and leaves the result
    alone; a tolerance no mixed form can meet forces the FP64 fallback, whose
    result is bit for bit the FP64 form's.  Without fallback the failed judge
    is only reported."""
    say(f"case: error judge and FP64 fallback, nmo=30, {form} form")
    prec = hip.PREC_MIXED_INT8 if form == "int8" else hip.PREC_MIXED_INT8_4
    default_tol = 1e-7 if form == "int8" else 1e-6
    inp = synthetic_input(30, 6, 47)
    with hip.HipContext(0) as ctx:
        ctx.upload_eri(inp.eri, inp.nocc)
        plain = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=prec)
        f64 = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=hip.PREC_FP64)
        assert ctx.last_judge() == {"max_rel_err": -1.0, "samples": 0, "fell_back": 0}
        ctx.set_judge(3)
        judged = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=prec)
        j = ctx.last_judge()
        say(f"   judge: {j}")
        assert j["samples"] == 3 and 0.0 < j["max_rel_err"] < default_tol and j["fell_back"] == 0
        assert np.array_equal(judged.sigma_c, plain.sigma_c)
        ctx.set_judge(3, tol=1e-30)
        fb = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=prec)
        assert ctx.last_judge()["fell_back"] == 1
        assert np.array_equal(fb.sigma_c, f64.sigma_c) and np.array_equal(fb.qp, f64.qp, equal_nan=True)
        ctx.set_judge(3, tol=1e-30, fallback=False)
        kept = hip.run_g0w0(inp, 100, 16, ctx=ctx, upload=False, precision=prec)
        j = ctx.last_judge()
        assert j["fell_back"] == 0 and j["max_rel_err"] > 1e-30
        assert np.array_equal(kept.sigma_c, plain.sigma_c)
        with pytest.raises(hip.HipError) as e:
            ctx.set_judge(65)
        assert e.value.code == hip.MGW_ERR_ARG


@pytest.mark.parametrize("form", ["int8", "int8x4"])
def test_small_gap_molecule_mixed_forms_stand(hip, form):
    """An ill-conditioned molecule: the nmo=30 synthetic molecule with its
    HOMO-LUMO gap shrunk to 1e-4 Ha (every virtual level moved down by the
    same amount), so that I + S V S reaches cond ~1/gap at small omega.  The
    FP64 form still matches the oracle to 2e-14; the int8 forms' Sigma_c error grows
    (measured 3.0e-8 / 5.1e-9 relative, against 5.2e-10 / 7.7e-11 at the
    generator's gap) while the q error that the judge measures stays near
    2e-8 / 3e-9, below its default tolerance: the result stands and no
    fallback fires."""
    prec = hip.PREC_MIXED_INT8 if form == "int8" else hip.PREC_MIXED_INT8_4
    sigma_bound = 3e-8 if form == "int8" else 2e-7
    inp = with_gap(synthetic_input(30, 6, 47), 1e-4)
    say(f"case: HOMO-LUMO gap 1e-4 Ha, nmo=30, FP64 and {form} forms")
    with hip.HipContext(0) as ctx:
        ctx.upload_eri(inp.eri, inp.nocc)
        f64 = hip.run_g0w0(inp, 200, 16, ctx=ctx, upload=False, precision=hip.PREC_FP64)
        ctx.set_judge(3)
        mix = hip.run_g0w0(inp, 200, 16, ctx=ctx, upload=False, precision=prec)
        j = ctx.last_judge()
    want = orc.run_g0w0(inp, 200, 16)
    scale = np.abs(want.sigma_c).max()
    err64 = np.abs(f64.sigma_c - want.sigma_c).max() / scale
    errmix = np.abs(mix.sigma_c - f64.sigma_c).max() / scale
    say(f"   FP64 vs oracle {err64:.2e}, {form} vs FP64 {errmix:.2e} (bound {sigma_bound:.0e}), "
        f"judge {j}")
    assert err64 < 2e-14
    assert errmix < sigma_bound
    assert j["samples"] == 3 and j["fell_back"] == 0
    assert 0.0 < j["max_rel_err"] < (1e-7 if form == "int8" else 1e-6) / 5


def test_cancel_flag_stops_between_batches(hip):
    """A stop flag set before the call (set from the host, the kernels all run
    to completion) ends mgw_hip_g0w0 with MGW_ERR_STATE; cleared, the same
    context runs the job normally."""
    import ctypes
    say("case: cancel flag, nmo=30, 8 frequencies per batch")
    inp = synthetic_input(30, 6, 48)
    flag = ctypes.c_int32(1)
    with hip.HipContext(0) as ctx:
        ctx.upload_eri(inp.eri, inp.nocc)
        ctx.set_batch_limit(8)
        ref = hip.run_g0w0(inp, 64, 16, ctx=ctx, upload=False)
        assert ref.timings["freq_batch"] == 8
        ctx.set_cancel_flag(flag)
        with pytest.raises(hip.HipError) as e:
            hip.run_g0w0(inp, 64, 16, ctx=ctx, upload=False)
        assert e.value.code == hip.MGW_ERR_STATE and "cancelled" in str(e.value)
        flag.value = 0
        again = hip.run_g0w0(inp, 64, 16, ctx=ctx, upload=False)
        assert np.array_equal(again.sigma_c, ref.sigma_c)
        ctx.set_cancel_flag(None)


def test_serve_faults_cancel_and_fallback(hip):
    """Failure handling of the request loop, each condition forced from the
    host: a job whose device budget is cut (fault-injection hook) fails with
    MGW_ERR_MEMORY and releases its reservation, the next job runs; a pending
    job is cancelled at once and a running one at its next frequency batch
    (MGW_ERR_STATE), without disturbing the job beside it; a judge
    tolerance no mixed form can meet makes the loop rerun a job in FP64."""
    say("case: request loop faults: budget cut, cancellation, judge fallback")
    small = _jobs()[:2]
    kw = dict(nfreq=40, npar=12)
    with hip.HipServer(0, workers=1, hbm_budget=64 << 30, batch_freqs=8) as srv:
        srv.set_budget_scale(0.01)
        j = small[0]
        t = srv.submit(factor=j["b"], mo_energy=j["mo"], vxc=j["vxc"], nocc=j["nocc"], fermi=j["fermi"], **kw)
        with pytest.raises(hip.HipError) as e:
            srv.wait(t)
        assert e.value.code == hip.MGW_ERR_MEMORY
        assert srv.stats()["reserved_bytes"] == 0
        srv.set_budget_scale(1.0)
        ok, _, _ = srv.wait(srv.submit(factor=j["b"], mo_energy=j["mo"], vxc=j["vxc"], nocc=j["nocc"],
                                       fermi=j["fermi"], **kw))
        assert np.all(np.isfinite(ok.sigma_c))
        # one worker: a long job runs while a second waits; cancel the waiting one,
        # then the running one
        mo, fermi, b, vxc = synthetic_factor(120, 24, 49)
        t_long = srv.submit(factor=b, mo_energy=mo, vxc=vxc, nocc=24, fermi=fermi, nfreq=200,
                            npar=16)
        t_wait = srv.submit(factor=j["b"], mo_energy=j["mo"], vxc=j["vxc"], nocc=j["nocc"],
                            fermi=j["fermi"], **kw)
        srv.cancel(t_wait)
        with pytest.raises(hip.HipError) as e:
            srv.wait(t_wait)
        assert e.value.code == hip.MGW_ERR_STATE and "before the job started" in str(e.value)
        while srv.stats()["running"] == 0:
            time.sleep(0.001)
        srv.cancel(t_long)
        with pytest.raises(hip.HipError) as e:
            srv.wait(t_long)
        assert e.value.code == hip.MGW_ERR_STATE and "cancelled" in str(e.value)
        # judge fallback through the loop
        srv.set_judge(2, tol=1e-30)
        t = srv.submit(factor=j["b"], mo_energy=j["mo"], vxc=j["vxc"], nocc=j["nocc"],
                       fermi=j["fermi"], precision=hip.PREC_MIXED_INT8, **kw)
        fb, _, _ = srv.wait(t)
        assert fb.timings["fell_back"] == 1
        st = srv.stats()
        assert st["running"] == 0 and st["reserved_bytes"] == 0
    with hip.HipContext(0) as ctx:
        ctx.upload_eri_factor(j["b"], j["nmo"], j["nocc"])
        inp = GwInput(mo_energy=j["mo"], eri=None, vxc=j["vxc"], nocc=j["nocc"], fermi_energy=j["fermi"])
        f64 = hip.run_g0w0(inp, 40, 12, ctx=ctx, upload=False)
    assert np.array_equal(fb.sigma_c, f64.sigma_c)


def test_split_batches_leave_no_device_memory_held(hip):
    """A molecule whose order splits the factorization over two streams
    (nph = 4096), run with graphs on: once the context's buffers are released,
    the device memory held by live allocations returns to its level before
    the job (memory the pool keeps for reuse is not held by any; a graph
    captured across the side streams used to keep the memory of the buffers
    it referenced).  A small job first reserves the runtime's per-queue
    scratch (k_pade_qp uses 3 KB per lane), which stays."""
    import ctypes
    rt = ctypes.CDLL("libamdhip64.so")

    def used_gib():
        """Device memory held by live allocations: in use on the device, less
        what the default memory pool keeps unused for reuse (the contexts
        allocate from it; hipMemPoolAttrReservedMemCurrent = 5, UsedMemCurrent = 7)."""
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        assert rt.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
        pool, res, use = ctypes.c_void_p(), ctypes.c_uint64(0), ctypes.c_uint64(0)
        idle = 0
        if rt.hipDeviceGetDefaultMemPool(ctypes.byref(pool), 0) == 0:
            assert rt.hipMemPoolGetAttribute(pool, 5, ctypes.byref(res)) == 0
            assert rt.hipMemPoolGetAttribute(pool, 7, ctypes.byref(use)) == 0
            idle = max(0, res.value - use.value)
        return (total.value - free.value - idle) / 2**30

    say("case: nmo=160 nocc=32 (nph 4096, split factorization) with graphs, memory after release")
    small = synthetic_input(18, 4, 51)
    mo, fermi, b, vxc = synthetic_factor(160, 32, 52)
    big = GwInput(mo_energy=mo, eri=None, vxc=vxc, nocc=32, fermi_energy=fermi)
    with hip.HipContext(0) as ctx:
        ctx.set_graphs(True)
        ctx.set_batch_limit(8)
        hip.run_g0w0(small, 40, 12, ctx=ctx, precision=hip.PREC_MIXED_INT8_4)
        ctx.release_buffers()
        before = used_gib()
        ctx.upload_eri_factor(b, 160, 32)
        for rep in range(2):
            res = hip.run_g0w0(big, 40, 12, ctx=ctx, upload=False, precision=hip.PREC_MIXED_INT8_4)
            assert res.timings["graph_batches"] == 0 and np.all(np.isfinite(res.sigma_c))
        peak = used_gib()
        ctx.release_buffers()
        after = used_gib()
    say(f"   device memory in use: {before:.2f} GiB before, {peak:.2f} with the job, {after:.2f} after release")
    assert peak > before + 5.0
    assert after < before + 0.25
