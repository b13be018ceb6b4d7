"""Molecules grouped through mgw_hip_g0w0_group (needs an MI355X).

Molecules of one padded particle-hole order share every frequency-batch
launch (dielectric, factorization, inverse, digit slicing); each result must
be bit for bit the molecule's own mgw_hip_g0w0 result, in every form, with
stacked batches that split molecules, with captured graphs replayed, and
next to molecules that run alone inside the call (another order, eta != 0).
Per-job failures (an invalid job, a cancelled job) leave the other jobs'
results intact."""
import ctypes

import numpy as np
import pytest

from minigw_amd.synthetic import synthetic_input

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hip():
    from minigw_amd import hip as h
    return h


# (nmo, nocc, seed): the first four pad to nph 256, the last to 512
MOLS = [(24, 5, 101), (20, 4, 102), (26, 5, 103), (22, 6, 104), (40, 12, 105)]


def _molecules():
    return [synthetic_input(nmo, nocc, seed) for nmo, nocc, seed in MOLS]


def _contexts(hip, inps):
    ctxs = [hip.HipContext(0) for _ in inps]
    for c, inp in zip(ctxs, inps):
        c.upload_eri(inp.eri, inp.nocc)
    return ctxs


def _same(a, b):
    return (np.array_equal(a.sigma_c, b.sigma_c) and np.array_equal(a.qp, b.qp, equal_nan=True)
            and np.array_equal(a.qp_iterations, b.qp_iterations)
            and np.array_equal(a.sigma_x_diag, b.sigma_x_diag))


@pytest.mark.parametrize("form", ["PREC_FP64", "PREC_MIXED_INT8_4", "PREC_MIXED_INT8", "PREC_MIXED"])
def test_group_equals_alone(hip, form):
    prec = getattr(hip, form)
    inps = _molecules()
    nfreq = [120, 80, 100, 60, 90]
    npar = [16, 12, 16, 8, 16]
    say(f"case: group of {len(inps)} molecules (nph 256 x4, 512 x1), {form}, nfreq {nfreq}")
    ctxs = _contexts(hip, inps)
    try:
        res, status = hip.run_g0w0_group(inps, ctxs, nfreq, npar, prec)
        assert list(status) == [hip.MGW_OK] * len(inps)
        for c in ctxs[1:]:
            assert c.last_timings()["freq_batch"] > 0
        # the four molecules of order 256 shared their batches: one stacked batch
        assert int(res[0].timings["freq_batch"]) == sum(nfreq[:4])
        for i, (inp, c) in enumerate(zip(inps, ctxs)):
            alone = hip.run_g0w0(inp, nfreq[i], npar[i], ctx=c, upload=False, precision=prec)
            assert _same(res[i], alone), (form, i)
    finally:
        for c in ctxs:
            c.close()


def test_group_with_split_stacks_and_graphs(hip):
    """A batch limit on the lead below one molecule's frequencies splits the
    stack across molecules; graphs captured on the first call replay on the
    second; both equal the molecules' own runs."""
    prec = hip.PREC_MIXED_INT8_4
    inps = _molecules()[:4]
    say("case: group of 4 (nph 256), int8x4, stacked batches of 70 frequencies, graphs")
    ctxs = _contexts(hip, inps)
    try:
        ctxs[0].set_batch_limit(70)
        ctxs[0].set_graphs(True)
        first, st1 = hip.run_g0w0_group(inps, ctxs, 100, 16, prec)
        second, st2 = hip.run_g0w0_group(inps, ctxs, 100, 16, prec)
        assert list(st1) == list(st2) == [hip.MGW_OK] * 4
        assert int(first[0].timings["freq_batch"]) <= 70
        assert second[0].timings["graph_batches"] > 0
        ctxs[0].set_graphs(False)
        ctxs[0].set_batch_limit(0)
        for i, (inp, c) in enumerate(zip(inps, ctxs)):
            alone = hip.run_g0w0(inp, 100, 16, ctx=c, upload=False, precision=prec)
            assert _same(first[i], alone) and _same(second[i], alone), i
    finally:
        for c in ctxs:
            c.close()


def test_group_mixes_alone_forms_and_reports_per_job(hip):
    """eta != 0 and the decomposed form run alone inside the call; an invalid
    job (a state out of range) and a job cancelled through its flag end with
    their own status while the others complete unchanged."""
    inps = _molecules()[:4]
    say("case: group with eta != 0, decomposed form, an invalid and a cancelled job")
    ctxs = _contexts(hip, inps)
    flag = ctypes.c_int32(0)
    try:
        prec = [hip.PREC_FP64, hip.PREC_FP64_CD, hip.PREC_FP64, hip.PREC_FP64]
        eta = [0.0, 0.0, 0.0, 0.05]
        res, status = hip.run_g0w0_group(inps, ctxs, 60, 12, prec, eta)
        assert list(status) == [hip.MGW_OK] * 4
        for i, (inp, c) in enumerate(zip(inps, ctxs)):
            alone = hip.run_g0w0(inp, 60, 12, ctx=c, upload=False, precision=prec[i], eta=eta[i])
            assert _same(res[i], alone), i
        # job 1 invalid, job 2 cancelled before the call
        states = [None, [0, inps[1].nmo + 3], None, None]
        ctxs[2].set_cancel_flag(flag)
        flag.value = 1
        res, status = hip.run_g0w0_group(inps, ctxs, 60, 12, hip.PREC_MIXED_INT8_4, 0.0, states)
        say(f"   status {list(status)}")
        assert status[0] == hip.MGW_OK and status[3] == hip.MGW_OK
        assert status[1] == hip.MGW_ERR_ARG and status[2] == hip.MGW_ERR_STATE
        assert res[1] is None and res[2] is None
        ctxs[2].set_cancel_flag(None)
        for i in (0, 3):
            alone = hip.run_g0w0(inps[i], 60, 12, ctx=ctxs[i], upload=False,
                                 precision=hip.PREC_MIXED_INT8_4)
            assert _same(res[i], alone), i
        with pytest.raises(hip.HipError) as e:
            hip.run_g0w0_group(inps[:2], [ctxs[0], ctxs[0]], 60, 12)
        assert e.value.code == hip.MGW_ERR_ARG
    finally:
        for c in ctxs:
            c.close()


def test_group_judges_each_member(hip):
    """Grouped mixed jobs under the error judge: the judges' sampled
    frequencies of all members run as one stacked FP64 batch; a member whose
    judge tolerance no form can meet reruns in FP64 (its result bit for bit
    the FP64 form's) while the others keep their int8 results, each equal to
    its run alone under the same judge."""
    prec = hip.PREC_MIXED_INT8_4
    inps = _molecules()[:4]
    say("case: group of 4 (nph 256), int8x4, judge on every member, member 2 forced to fall back")
    ctxs = _contexts(hip, inps)
    try:
        for i, c in enumerate(ctxs):
            c.set_judge(2, tol=1e-30 if i == 2 else 0.0)
        res, status = hip.run_g0w0_group(inps, ctxs, 80, 12, prec)
        assert list(status) == [hip.MGW_OK] * 4
        judges = [c.last_judge() for c in ctxs]
        say(f"   judges: {judges}")
        for i, j in enumerate(judges):
            assert j["samples"] == 2 and 0.0 < j["max_rel_err"] < 1e-6
            assert j["fell_back"] == (1 if i == 2 else 0)
        for i, (inp, c) in enumerate(zip(inps, ctxs)):
            alone = hip.run_g0w0(inp, 80, 12, ctx=c, upload=False, precision=prec)
            assert c.last_judge()["fell_back"] == (1 if i == 2 else 0)
            assert _same(res[i], alone), i
        f64 = hip.run_g0w0(inps[2], 80, 12, ctx=ctxs[2], upload=False, precision=hip.PREC_FP64)
        assert _same(res[2], f64)
    finally:
        for c in ctxs:
            c.close()
