"""ERI storage of a context (needs an MI355X).

The streamed storage passes the ERI through the device in slabs of whole rows
p and keeps only the particle-hole block (pk|ia) and the exchange diagonal
(pi|pi) (the counterpart of the reference's StreamingPanel mode,
src/workspace/device_pq_ph_panel.cu:168-190,266-282).  Every result must be
the resident storage's bit for bit: G0W0 in every form from a host ERI and
from its factor, with one slab and with many (slab starts not aligned to the
64-row tiles of the factor product), and the panel view (V_ph, W_c, quadratic
forms, device panels).  What needs the whole tensor (the decomposed form,
downloading the ERI) is MGW_ERR_UNSUPPORTED.  ERI_AUTO follows the reference's
rule (resident while the tensor takes under a quarter of the free memory,
here the context's budget); the request loop starts jobs streamed under
ERI_AUTO when they do not fit resident, with the results of their resident
runs."""
import numpy as np
import pytest

from minigw_amd.synthetic import GwInput, synthetic_factor, synthetic_input
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hip():
    from minigw_amd import hip as h
    return h


def _ctx(hip, monkeypatch, storage, slab_rows=None, nmo=None):
    """A context taking `storage`; slab_rows rows p per streamed slab."""
    if slab_rows is not None:
        monkeypatch.setenv("MGW_ERI_SLAB_BYTES", str(slab_rows * nmo ** 3 * 8))
    else:
        monkeypatch.delenv("MGW_ERI_SLAB_BYTES", raising=False)
    c = hip.HipContext(0)
    c.set_eri_storage(storage)
    return c


def _same(a, b):
    return (np.array_equal(a.sigma_c, b.sigma_c) and np.array_equal(a.qp, b.qp, equal_nan=True)
            and np.array_equal(a.qp_iterations, b.qp_iterations)
            and np.array_equal(a.sigma_x_diag, b.sigma_x_diag)
            and (a.q is None or np.array_equal(a.q, b.q)))


FORMS = ["PREC_FP64", "PREC_MIXED_INT8_4", "PREC_MIXED_INT8", "PREC_MIXED"]


@pytest.mark.parametrize("form", FORMS)
def test_streamed_host_eri_equals_resident(hip, monkeypatch, form):
    prec = getattr(hip, form)
    nmo, nocc = 26, 5
    inp = synthetic_input(nmo, nocc, 301)
    say(f"case: host ERI nmo={nmo}, {form}, resident vs streamed (1 slab, slabs of 3 rows)")
    res = _ctx(hip, monkeypatch, hip.ERI_RESIDENT)
    try:
        res.upload_eri(inp.eri, nocc)
        st, nbytes = res.eri_storage()
        assert st == hip.ERI_RESIDENT and nbytes == nmo ** 4 * 8
        want = hip.run_g0w0(inp, 60, 12, ctx=res, upload=False, precision=prec, want_q=True)
    finally:
        res.close()
    for rows in (None, 3):
        c = _ctx(hip, monkeypatch, hip.ERI_STREAMED, rows, nmo)
        try:
            c.upload_eri(inp.eri, nocc)
            st, nbytes = c.eri_storage()
            held = (nmo * nmo * nocc * (nmo - nocc) + nmo * nocc) * 8
            assert st == hip.ERI_STREAMED and nbytes == held
            got = hip.run_g0w0(inp, 60, 12, ctx=c, upload=False, precision=prec, want_q=True)
            assert _same(got, want), (form, rows)
        finally:
            c.close()


@pytest.mark.parametrize("form", ["PREC_FP64", "PREC_MIXED_INT8_4"])
def test_streamed_factor_equals_resident(hip, monkeypatch, form):
    """The row-slab product of the factor gives the mirrored full product's
    values: slabs of 2 rows p start at row 2 p nmo of B, off the 64-row tiles."""
    prec = getattr(hip, form)
    nmo, nocc = 30, 6
    mo, fermi, b, vxc = synthetic_factor(nmo, nocc, 302)
    inp = GwInput(mo_energy=mo, eri=None, vxc=vxc, nocc=nocc, fermi_energy=fermi)
    say(f"case: factor nmo={nmo}, {form}, resident vs streamed slabs of 2 rows")
    runs = []
    for storage, rows in ((hip.ERI_RESIDENT, None), (hip.ERI_STREAMED, 2)):
        c = _ctx(hip, monkeypatch, storage, rows, nmo)
        try:
            c.upload_eri_factor(b, nmo, nocc)
            assert c.eri_storage()[0] == storage
            runs.append(hip.run_g0w0(inp, 48, 12, ctx=c, upload=False, precision=prec,
                                     want_q=True))
        finally:
            c.close()
    assert _same(runs[1], runs[0]), form


def test_streamed_panel_view_equals_resident(hip, monkeypatch):
    """DevicePqPhPanelView's operations in the streamed storage: V_ph built on
    the device, W_c from it, fused quadratic forms and device panels read on a
    second context, bit for bit the resident ones (and near the oracle)."""
    nmo, nocc = 24, 5
    inp = synthetic_input(nmo, nocc, 11)
    ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
    v = orc.v_ph_matrix(inp.eri, ph)
    pi0 = orc.pi0_diag_real(0.8, ph.delta)
    wc_ref = orc.compute_w_c(v, pi0)
    panels = [(0, 0, 24), (23, 20, 4), (7, 3, 1), (5, 0, 0)]
    out = {}
    for mode in (hip.ERI_RESIDENT, hip.ERI_STREAMED):
        say(f"case: panel view nmo={nmo}, storage {hip.ERI_STORAGE_NAMES[mode]}")
        ctx = _ctx(hip, monkeypatch, hip.ERI_AUTO, 5, nmo)
        src = _ctx(hip, monkeypatch, hip.ERI_AUTO, 5, nmo)
        try:
            ws = hip.HipScreeningWorkspace(v, ctx)
            view = hip.HipPqPhPanelView(inp.eri, nocc, ctx, mode)
            assert view.storage_mode() == mode
            assert view.storage_mode_name() == hip.ERI_STORAGE_NAMES[mode]
            view.build_v_ph()
            ws.compute_w_c(pi0)
            wc = ws.w_c()
            assert np.abs(wc - wc_ref).max() <= 1e-12 * np.abs(wc_ref).max()
            fused = [view.quadratic_forms(p, k0, w) for p, k0, w in panels]
            other = hip.HipPqPhPanelView(inp.eri, nocc, src, mode)
            dev = [ws.quadratic_forms_panel(other.fill_panel(p, k0, w)) for p, k0, w in panels]
            for (p, k0, w), q, qd in zip(panels, fused, dev):
                want = orc.quadratic_forms_panel(wc_ref, orc.pq_panel(inp.eri, ph, p, k0, w))
                assert np.allclose(q.real, want, rtol=1e-11, atol=1e-13), (mode, p, k0, w)
                assert np.array_equal(q, qd), (mode, p, k0, w)
            out[mode] = (wc, fused)
        finally:
            src.close()
            ctx.close()
    wr, qr = out[hip.ERI_RESIDENT]
    ws_, qs = out[hip.ERI_STREAMED]
    assert np.array_equal(wr, ws_)
    assert all(np.array_equal(a, b) for a, b in zip(qr, qs))


def test_streamed_storage_limits_and_auto(hip, monkeypatch):
    """The decomposed form and downloading the ERI need the whole tensor;
    ERI_AUTO is resident while the tensor takes under a quarter of the memory
    left to the context (its budget), streamed beyond; an unknown storage is
    an argument error."""
    nmo, nocc = 24, 5
    inp = synthetic_input(nmo, nocc, 11)
    full = nmo ** 4 * 8
    say("case: streamed storage limits, AUTO against the context budget")
    c = _ctx(hip, monkeypatch, hip.ERI_STREAMED)
    try:
        with pytest.raises(hip.HipError) as e:
            c.eri_storage()
        assert e.value.code == hip.MGW_ERR_STATE
        c.upload_eri(inp.eri, nocc)
        with pytest.raises(hip.HipError) as e:
            c.download_eri(nmo)
        assert e.value.code == hip.MGW_ERR_UNSUPPORTED
        with pytest.raises(hip.HipError) as e:
            hip.run_g0w0(inp, 20, 8, ctx=c, upload=False, precision=hip.PREC_FP64_CD)
        assert e.value.code == hip.MGW_ERR_UNSUPPORTED
        with pytest.raises(hip.HipError) as e:
            c.set_eri_storage(3)
        assert e.value.code == hip.MGW_ERR_ARG
        c.set_eri_storage(hip.ERI_AUTO)
        c.set_memory_budget(3 * full)                 # a quarter of it is below the tensor
        c.upload_eri(inp.eri, nocc)
        assert c.eri_storage()[0] == hip.ERI_STREAMED
        c.set_memory_budget(5 * full)
        c.upload_eri(inp.eri, nocc)
        assert c.eri_storage()[0] == hip.ERI_RESIDENT
        assert np.array_equal(c.download_eri(nmo), inp.eri)
        c.set_memory_budget(0)
    finally:
        c.close()
    # the streamed footprint: held block + diagonal + one slab (here the whole
    # tensor by default, one row p with MGW_ERI_SLAB_BYTES at one row)
    held = (nmo * nmo * nocc * (nmo - nocc) + nmo * nocc) * 8
    assert hip.eri_bytes(nmo, nocc, hip.ERI_RESIDENT) == full
    assert hip.eri_bytes(nmo, nocc, hip.ERI_STREAMED) == held + full
    monkeypatch.setenv("MGW_ERI_SLAB_BYTES", str(nmo ** 3 * 8))
    assert hip.eri_bytes(nmo, nocc, hip.ERI_STREAMED) == held + nmo ** 3 * 8
    assert hip.eri_bytes(nmo, nocc, hip.ERI_AUTO) == 0


def _serve_estimate(hip, d, storage):
    """The request loop's reservation for a factor job of 40 frequencies in
    batches of 8 (FP64): its estimate plus the factor on the device, with the
    storage's ERI bytes."""
    nmo, nocc, kdim = d["nmo"], d["nocc"], d["b"].shape[1]
    est = hip.estimate_bytes(nmo, nocc, 40, 0, 8, hip.PREC_FP64)
    est += ((nmo * nmo + 63) // 64 * 64 + 64) * ((kdim + 15) // 16 * 16) * 8
    return est - hip.eri_bytes(nmo, nocc, hip.ERI_RESIDENT) + hip.eri_bytes(nmo, nocc, storage)


def test_request_loop_streams_jobs_that_do_not_fit_resident(hip, monkeypatch):
    """Under ERI_AUTO (slabs of one row p) a job whose resident estimate
    exceeds the whole budget but whose streamed one fits runs streamed instead
    of failing; with a budget that holds one resident and one streamed job,
    jobs run side by side in both storages.  Every result equals the job's
    resident run alone; the default storage still refuses the large job."""
    monkeypatch.setenv("MGW_ERI_SLAB_BYTES", "1")        # one row p per slab
    data = []
    for nmo, nocc, seed in [(40, 8, 401), (40, 8, 402), (40, 8, 403)]:
        mo, fermi, b, vxc = synthetic_factor(nmo, nocc, seed)
        data.append(dict(nmo=nmo, nocc=nocc, mo=mo, fermi=fermi, b=b, vxc=vxc))
    res_b = _serve_estimate(hip, data[0], hip.ERI_RESIDENT)
    str_b = _serve_estimate(hip, data[0], hip.ERI_STREAMED)
    assert str_b < res_b
    say(f"case: request loop, ERI_AUTO, nmo=40: resident {res_b} B, streamed {str_b} B")

    def submit(srv, d):
        return srv.submit(factor=d["b"], mo_energy=d["mo"], vxc=d["vxc"], nocc=d["nocc"],
                          fermi=d["fermi"], nfreq=40, npar=8)

    got = []
    with hip.HipServer(0, workers=1, hbm_budget=(res_b + str_b) // 2, batch_freqs=8) as srv:
        with pytest.raises(hip.HipError) as e:          # resident by default: too large
            srv.wait(submit(srv, data[0]))
        assert e.value.code == hip.MGW_ERR_MEMORY
        srv.set_eri_storage(hip.ERI_AUTO)
        g = srv.wait(submit(srv, data[0]))[0]
        assert g.timings["eri_storage"] == hip.ERI_STREAMED
        got.append(g)
    with hip.HipServer(0, workers=2, hbm_budget=res_b + str_b + 4096, batch_freqs=8) as srv:
        srv.set_grouping(1)
        srv.set_eri_storage(hip.ERI_AUTO)
        tickets = [submit(srv, d) for d in data]
        runs = [srv.wait(t)[0] for t in tickets]
    storages = [r.timings["eri_storage"] for r in runs]
    say(f"   side by side: storages {storages}")
    assert set(storages) <= {hip.ERI_RESIDENT, hip.ERI_STREAMED}
    for d, g in zip(data + data[:1], runs + got):
        with hip.HipContext(0) as c:
            c.set_eri_storage(hip.ERI_RESIDENT)
            c.upload_eri_factor(d["b"], d["nmo"], d["nocc"])
            inp = GwInput(mo_energy=d["mo"], eri=None, vxc=d["vxc"], nocc=d["nocc"],
                          fermi_energy=d["fermi"])
            alone = hip.run_g0w0(inp, 40, 8, ctx=c, upload=False)
        assert np.array_equal(g.sigma_c, alone.sigma_c)
        assert np.array_equal(g.qp, alone.qp, equal_nan=True)
