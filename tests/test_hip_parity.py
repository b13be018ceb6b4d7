"""Parity of the HIP path (through the C ABI) with the oracle and the
reference's own outputs.  Needs an MI355X."""
import math
import os

import numpy as np
import pytest

from minigw_amd.synthetic import GwInput, synthetic_input, with_gap, CONFIGS
from oracle import oracle as orc
from qp_bounds import SIGMA_REL, fp64_qp_violations, mixed_qp_violations

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hip():
    from minigw_amd import hip as h
    return h


@pytest.fixture(scope="module")
def ctx(hip):
    c = hip.HipContext(0)
    yield c
    c.close()


def h2o_input():
    d = np.load(os.path.join(GOLD, "h2o_input.npz"))
    return GwInput(mo_energy=d["mo_energy"], eri=d["eri"], vxc=d["vxc"], nocc=int(d["nocc"]),
                   fermi_energy=float(d["fermi_energy"]))


def golden_input(name, r):
    """The input of a reference fixture case."""
    if name.startswith("h2o"):
        return h2o_input()
    if name == "syn_nmo24":
        return synthetic_input(*CONFIGS["nmo24"])
    if name in ("syn_gap", "syn_gap_eta"):
        return with_gap(synthetic_input(30, 6, 47), 1e-4)
    return GwInput(mo_energy=r["mo_energy"], eri=r["eri"], vxc=r["vxc"], nocc=int(r["nocc"]),
                   fermi_energy=float(r["fermi_energy"]))


# ------------------------------------------------------------ MFMA layout
def test_mfma_f64_tile_layout(ctx):
    say("case: f64 MFMA tile C = A B^T, 64x64")
    rng = np.random.default_rng(0)
    a = rng.standard_normal((64, 64))
    b = rng.standard_normal((64, 64))
    c = ctx.selftest_tile(a, b)
    out = os.environ.get("MGW_PROBE_DIR")      # optional dump of the probe for inspection
    if out and os.path.isdir(out):
        np.savez(os.path.join(out, "tile_selftest.npz"), a=a, b=b, c=c)
    assert np.allclose(c, a @ b.T, rtol=1e-13, atol=1e-12)


# ------------------------------------------------------------ workspace boundary
@pytest.mark.parametrize("nmo,nocc,seed", [(10, 3, 7), (24, 5, 20240601), (40, 9, 3)])
def test_compute_wc_matches_oracle(hip, ctx, nmo, nocc, seed):
    say(f"case: compute_w_c nmo={nmo} nocc={nocc}")
    inp = synthetic_input(nmo, nocc, seed)
    ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
    v = orc.v_ph_matrix(inp.eri, ph)
    ws = hip.HipScreeningWorkspace(v, ctx)
    assert ws.nph == ph.size
    for w in [1e-12, 0.37, 25.0]:
        pi0 = orc.pi0_diag(w, ph.delta, 0.0)
        ws.compute_w_c(pi0)
        want = orc.compute_w_c(v, pi0.real)
        got = ws.w_c()
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-12, err
        rng = np.random.default_rng(1)
        x = rng.standard_normal((ph.size, 37))
        q = ws.quadratic_forms_panel(x)
        qw = orc.quadratic_forms_panel(want, x)
        assert np.allclose(q.real, qw, rtol=1e-11, atol=1e-13)
        assert np.all(q.imag == 0)


@pytest.mark.parametrize("eta", [0.01, 0.2])
def test_compute_wc_complex_matches_oracle(hip, ctx, eta):
    """eta != 0: complex Pi0, W_c = eps^-1 diag(Pi0) against the oracle's
    complex getrf/getrs, and the complex quadratic forms (no conjugate)."""
    say(f"case: compute_w_c complex, eta={eta}, nmo=24 and 40")
    for nmo, nocc, seed in [(24, 5, 20240601), (40, 9, 3)]:
        inp = synthetic_input(nmo, nocc, seed)
        ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
        v = orc.v_ph_matrix(inp.eri, ph)
        ws = hip.HipScreeningWorkspace(v, ctx)
        for w in [1e-12, 0.37, 25.0]:
            pi0 = orc.pi0_diag(w, ph.delta, eta)
            ws.compute_w_c(pi0)
            want = orc.compute_w_c(v, pi0)
            got = ws.w_c()
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-12, (nmo, w, err)
            rng = np.random.default_rng(2)
            x = rng.standard_normal((ph.size, 70))
            q = ws.quadratic_forms_panel(x)
            qw = orc.quadratic_forms_panel(want, x)
            assert np.abs(q - qw).max() <= 1e-11 * np.abs(qw).max(), (nmo, w)
        view = hip.HipPqPhPanelView(inp.eri, inp.nocc, ctx)
        pi0 = orc.pi0_diag(0.8, ph.delta, eta)
        ws.compute_w_c(pi0)
        wc = orc.compute_w_c(v, pi0)
        for p, k0, width in [(0, 0, nmo), (3, 2, 5)]:
            q = view.quadratic_forms(p, k0, width)
            want = orc.quadratic_forms_panel(wc, orc.pq_panel(inp.eri, ph, p, k0, width))
            assert np.abs(q - want).max() <= 1e-11 * np.abs(want).max(), (p, k0, width)


def test_eri_panels_match_oracle(hip, ctx):
    say("case: ERI-resident panels nmo=24")
    inp = synthetic_input(24, 5, 11)
    ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
    ws = hip.HipScreeningWorkspace(orc.v_ph_matrix(inp.eri, ph), ctx)
    view = hip.HipPqPhPanelView(inp.eri, inp.nocc, ctx)
    view.build_v_ph()                     # V_ph rebuilt on the device from the ERI
    pi0 = orc.pi0_diag_real(0.8, ph.delta)
    ws.compute_w_c(pi0)
    wc = orc.compute_w_c(orc.v_ph_matrix(inp.eri, ph), pi0)
    for p, k0, width in [(0, 0, 24), (4, 0, 24), (23, 20, 4), (7, 3, 1)]:
        q = view.quadratic_forms(p, k0, width)
        want = orc.quadratic_forms_panel(wc, orc.pq_panel(inp.eri, ph, p, k0, width))
        assert np.allclose(q.real, want, rtol=1e-11, atol=1e-13), (p, k0, width)
    with pytest.raises(hip.HipError) as e:
        view.quadratic_forms(24, 0, 1)
    assert "p index out of range" in str(e.value)
    with pytest.raises(hip.HipError):
        view.quadratic_forms(0, 20, 5)


def test_device_panels_across_contexts(hip, ctx):
    """fill_panel on the panel source's context, quadratic_forms_panel on the
    workspace's: the reference's two device objects, agreeing with the fused
    entry point and with the oracle."""
    say("case: device panels, panel source and workspace on two contexts, nmo=24")
    inp = synthetic_input(24, 5, 11)
    ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
    v = orc.v_ph_matrix(inp.eri, ph)
    ws = hip.HipScreeningWorkspace(v, ctx)
    pi0 = orc.pi0_diag_real(0.8, ph.delta)
    ws.compute_w_c(pi0)
    wc = orc.compute_w_c(v, pi0)
    src_ctx = hip.HipContext(0)
    try:
        view = hip.HipPqPhPanelView(inp.eri, inp.nocc, src_ctx)
        fused = hip.HipPqPhPanelView(inp.eri, inp.nocc, ctx)
        for p, k0, width in [(0, 0, 24), (23, 20, 4), (7, 3, 1), (5, 0, 0)]:
            panel = view.fill_panel(p, k0, width)
            assert panel.ld >= ph.size and panel.ld % 256 == 0 and panel.cols == width
            q = ws.quadratic_forms_panel(panel)
            want = orc.quadratic_forms_panel(wc, orc.pq_panel(inp.eri, ph, p, k0, width))
            assert np.allclose(q.real, want, rtol=1e-11, atol=1e-13), (p, k0, width)
            assert np.array_equal(q, fused.quadratic_forms(p, k0, width)), (p, k0, width)
        with pytest.raises(hip.HipError) as e:
            view.fill_panel(24, 0, 1)
        assert "p index out of range" in str(e.value)
        host = np.zeros(ph.size)
        bad = hip.HipDevicePanel(host.ctypes.data, ph.size, ph.size, 1)
        with pytest.raises(hip.HipError) as e:
            ws.quadratic_forms_panel(bad)             # host memory is refused, not read
        assert e.value.code == hip.MGW_ERR_ARG
    finally:
        src_ctx.close()


def test_boundary_errors(hip, ctx):
    say("case: boundary error paths")
    with pytest.raises(hip.HipError) as e:
        hip.HipScreeningWorkspace(np.zeros((3, 4)), ctx)
    v = np.eye(5) * 0.3
    ws = hip.HipScreeningWorkspace(v, ctx)
    with pytest.raises(hip.HipError) as e:
        ws.compute_w_c(np.full(4, -0.5))
    assert e.value.code == hip.MGW_ERR_ARG
    p0 = -0.5 + 0.01j                          # n = 1 closed form per diagonal entry
    ws.compute_w_c(np.full(5, p0))
    assert np.allclose(ws.w_c(), np.eye(5) * (p0 / (1 - 0.3 * p0)), atol=1e-15)
    ws.compute_w_c(np.full(5, -0.5))
    assert np.allclose(ws.w_c(), np.eye(5) * (-0.5 / (1 + 0.5 * 0.3)), atol=1e-15)


# ------------------------------------------------------------ full chain, FP64
def test_h2o_state5_fixture(hip, ctx):
    say("case: H2O state 5, 200 GL, 16 Pade, FP64")
    inp = h2o_input()
    fx = np.load(os.path.join(GOLD, "h2o_serial_fixtures.npz"))
    res = hip.run_g0w0(inp, 200, 16, states=[4], ctx=ctx)
    sc = res.sigma_c[0]
    ref = fx["e_c_before_pade"]
    for x, y in zip(sc, ref):  # compare_reference.py component check at 1e-8
        assert math.isclose(x.real, y.real, rel_tol=1e-8, abs_tol=1e-8)
        assert math.isclose(x.imag, y.imag, rel_tol=1e-8, abs_tol=1e-8)
    assert np.abs(sc - ref).max() < 1e-11
    row = fx["summary_row_orbital5_ev"]
    assert math.isclose(res.qp[4] * orc.K_HARTREE_TO_EV, row[4], rel_tol=1e-5, abs_tol=1e-5)
    assert np.isnan(res.qp[0]) and res.qp_iterations[4] > 0


@pytest.mark.parametrize("name", ["h2o_all", "syn_nmo24", "syn_small", "syn_odd", "h2o_eta",
                                  "syn_small_eta", "syn_gap", "syn_gap_eta"])
def test_all_states_vs_reference_path(hip, ctx, name):
    """FP64 device path against the reference's own outputs: Sigma_c(i w) to
    2e-14 of max |Sigma_c| from both reference backends; QP energies within
    1e-8 Ha of both, or, on states the reference's own backends show to be
    ill-conditioned, within the bound their amplification gives at that
    Sigma_c precision (tests/qp_bounds.py).  syn_gap(_eta) is the
    ill-conditioned case (eta = 0 and 1e-3): the nmo=30 molecule with a
    1e-4 Ha HOMO-LUMO gap, where the
    quadrature's d^2 + w'^2 cancels to a few digits and only the host's
    rounding of every step reproduces the reference."""
    r = dict(np.load(os.path.join(GOLD, f"ref_{name}.npz")))
    rb = dict(np.load(os.path.join(GOLD, f"ref_{name}_blas.npz")))
    say(f"case: all states vs reference path, {name}")
    inp = golden_input(name, r)
    nf, npar = int(r["nfreq"]), int(r["npar"])
    res = hip.run_g0w0(inp, nf, npar, eta=float(r["eta"]), ctx=ctx)
    assert np.array_equal(res.omegas, r["omega"])
    assert np.allclose(res.sigma_x_diag, r["sigma_x_diag"], rtol=1e-14, atol=1e-15)
    scale = np.abs(r["sigma_c"]).max()
    dsig = np.abs(res.sigma_c - r["sigma_c"]).max() / scale
    dsig_b = np.abs(res.sigma_c - rb["sigma_c"]).max() / scale
    say(f"   rel |dSigma_c| = {dsig:.2e} (ReferenceBackend), {dsig_b:.2e} (BLAS/LAPACK)")
    assert dsig <= SIGMA_REL["fp64"] and dsig_b <= SIGMA_REL["fp64"]
    bad, summary = fp64_qp_violations(res.qp, name)
    say(f"   {summary}")
    assert not bad, (bad, summary)


def test_pade_qp_kernel_exact_on_same_sigma(hip, ctx):
    """Pade + QP on the device equals the oracle's bit for bit when Sigma_c agrees."""
    say("case: device Pade/QP vs oracle on the device's own Sigma_c, nmo=24")
    inp = synthetic_input(*CONFIGS["nmo24"])
    res = hip.run_g0w0(inp, 64, 16, ctx=ctx)
    qp, it = orc.solve_qp(inp.mo_energy, inp.fermi_energy, res.sigma_x_diag, np.diag(inp.vxc),
                          res.sigma_c, res.omegas, list(range(inp.nmo)), 16)
    assert np.array_equal(it, res.qp_iterations)
    assert np.array_equal(qp, res.qp)


def test_pade_qp_entry_point_matches_oracle(hip, ctx):
    """mgw_hip_pade_qp on a host Sigma_c (oracle's, perturbed, a state
    subset) equals the oracle's solve_qp bit for bit."""
    say("case: mgw_hip_pade_qp vs oracle solve_qp, nmo=24, states 0,4,5,6,23")
    inp = synthetic_input(*CONFIGS["nmo24"])
    want = orc.run_g0w0(inp, 48, 12)
    rng = np.random.default_rng(3)
    states = [0, 4, 5, 6, 23]
    sig = want.sigma_c[states] * (1 + 1e-3 * rng.standard_normal((len(states), 48)))
    qp_o, it_o = orc.solve_qp(inp.mo_energy, inp.fermi_energy, want.sigma_x_diag,
                              np.diag(inp.vxc), sig, want.omegas, states, 12)
    qp_d, it_d = hip.pade_qp(sig, want.omegas, want.sigma_x_diag, inp.mo_energy, np.diag(inp.vxc),
                             inp.fermi_energy, 12, states=states, ctx=ctx)
    assert np.array_equal(qp_d, qp_o, equal_nan=True)
    assert np.array_equal(it_d, it_o)
    assert np.isnan(qp_d[1]) and it_d[1] == 0
    with pytest.raises(hip.HipError) as e:
        hip.pade_qp(sig, want.omegas, want.sigma_x_diag, inp.mo_energy, np.diag(inp.vxc),
                    inp.fermi_energy, 0, states=states, ctx=ctx)
    assert e.value.code == hip.MGW_ERR_ARG and "npar must be positive" in str(e.value)


@pytest.mark.parametrize("nfreq,npar", [(96, 80), (96, 200), (40, 64), (40, 65), (24, 1)])
def test_pade_qp_any_order_matches_oracle(hip, ctx, nfreq, npar):
    """Any Pade order the reference accepts (it fits min(npar, nfreq) points,
    src/gw/pade.cpp:25-48): beyond 64 points the device keeps the fraction's
    tables in global memory instead of the thread's registers; QP energies
    and iteration counts equal the oracle's bit for bit."""
    say(f"case: mgw_hip_pade_qp, nfreq={nfreq}, npar={npar}, nmo=24")
    inp = synthetic_input(*CONFIGS["nmo24"])
    want = orc.run_g0w0(inp, nfreq, 16)
    rng = np.random.default_rng(nfreq + npar)
    sig = want.sigma_c * (1 + 1e-3 * rng.standard_normal(want.sigma_c.shape))
    states = list(range(inp.nmo))
    qp_o, it_o = orc.solve_qp(inp.mo_energy, inp.fermi_energy, want.sigma_x_diag,
                              np.diag(inp.vxc), sig, want.omegas, states, npar)
    qp_d, it_d = hip.pade_qp(sig, want.omegas, want.sigma_x_diag, inp.mo_energy, np.diag(inp.vxc),
                             inp.fermi_energy, npar, ctx=ctx)
    assert np.array_equal(qp_d, qp_o, equal_nan=True)
    assert np.array_equal(it_d, it_o)


@pytest.mark.parametrize("nmo,nocc,nfreq,eta", [(60, 12, 12, 0.0), (90, 18, 6, 0.0),
                                                 (60, 12, 8, 0.03)])
def test_multiblock_vs_oracle(hip, ctx, nmo, nocc, nfreq, eta):
    say(f"case: multi-block factorization nmo={nmo} nocc={nocc} nfreq={nfreq} eta={eta}")
    inp = synthetic_input(nmo, nocc, 99)
    res = hip.run_g0w0(inp, nfreq, 6, eta=eta, ctx=ctx)
    want = orc.run_g0w0(inp, nfreq, 6, eta=eta)
    d = np.abs(res.sigma_c - want.sigma_c).max() / np.abs(want.sigma_c).max()
    say(f"   rel |dSigma_c| = {d:.3e}")
    assert d < 1e-11


@pytest.mark.parametrize("nmo,nocc", [(2, 1), (9, 1), (9, 8), (7, 3)])
def test_extreme_occupations_vs_oracle(hip, ctx, nmo, nocc):
    """One particle-hole pair (nmo = 2), one occupied or one virtual orbital:
    nph of 1 to 12, all inside one zero-padded 256-wide tile.  FP64 to the
    oracle at 1e-11, the decomposed-ERI form at 1e-10, the int8 forms within
    1e-9 (5 x 7 bits) and 1e-8 (4 x 8 bits) of FP64; QP energies of the FP64
    form equal the oracle's fixed point
    on the same Sigma_c."""
    say(f"case: extreme occupation nmo={nmo} nocc={nocc}")
    inp = synthetic_input(nmo, nocc, 5)
    want = orc.run_g0w0(inp, 32, 8)
    scale = np.abs(want.sigma_c).max()
    res = hip.run_g0w0(inp, 32, 8, ctx=ctx)
    d = np.abs(res.sigma_c - want.sigma_c).max() / scale
    cd = hip.run_g0w0(inp, 32, 8, ctx=ctx, upload=False, precision=hip.PREC_FP64_CD)
    dcd = np.abs(cd.sigma_c - want.sigma_c).max() / scale
    mix = hip.run_g0w0(inp, 32, 8, ctx=ctx, upload=False, precision=hip.PREC_MIXED_INT8)
    dmix = np.abs(mix.sigma_c - res.sigma_c).max() / scale
    mix4 = hip.run_g0w0(inp, 32, 8, ctx=ctx, upload=False, precision=hip.PREC_MIXED_INT8_4)
    dmix4 = np.abs(mix4.sigma_c - res.sigma_c).max() / scale
    say(f"   rel |dSigma_c|: FP64 {d:.2e}, decomposed {dcd:.2e}, int8 vs FP64 {dmix:.2e}, "
        f"int8x4 vs FP64 {dmix4:.2e}")
    assert d < 1e-11 and dcd < 1e-10 and dmix < SIGMA_REL["int8"] and dmix4 < SIGMA_REL["int8x4"]
    qp, _ = orc.solve_qp(inp.mo_energy, inp.fermi_energy, res.sigma_x_diag, np.diag(inp.vxc),
                         res.sigma_c, res.omegas, list(range(nmo)), 8)
    assert np.allclose(res.qp, qp, rtol=0, atol=1e-12, equal_nan=True)


def test_subset_states_and_errors(hip, ctx):
    say("case: state subset + job errors")
    inp = synthetic_input(16, 4, 5)
    full = hip.run_g0w0(inp, 24, 8, ctx=ctx)
    sub = hip.run_g0w0(inp, 24, 8, states=[2, 9, 15], ctx=ctx, upload=False)
    assert np.allclose(sub.sigma_c, full.sigma_c[[2, 9, 15]], rtol=1e-13, atol=1e-15)
    assert np.isnan(sub.qp[0]) and not np.isnan(sub.qp[9])
    with pytest.raises(hip.HipError) as e:
        hip.run_g0w0(inp, 24, 8, eta=0.01, ctx=ctx, upload=False, precision=hip.PREC_MIXED)
    assert e.value.code == hip.MGW_ERR_UNSUPPORTED
    with pytest.raises(hip.HipError) as e:
        hip.run_g0w0(inp, 24, 8, states=[3, 3], ctx=ctx, upload=False)
    assert e.value.code == hip.MGW_ERR_ARG
    with pytest.raises(hip.HipError) as e:
        hip.run_g0w0(inp, 24, 0, ctx=ctx, upload=False)
    assert e.value.code == hip.MGW_ERR_ARG and "npar must be positive" in str(e.value)
    # a Pade order beyond the points fits all of them, as the reference does
    more = hip.run_g0w0(inp, 24, 65, ctx=ctx, upload=False)
    same = hip.run_g0w0(inp, 24, 24, ctx=ctx, upload=False)
    assert np.array_equal(more.qp, same.qp, equal_nan=True)
    big = hip.run_g0w0(inp, 96, 80, ctx=ctx, upload=False)
    qp_o, it_o = orc.solve_qp(inp.mo_energy, inp.fermi_energy, big.sigma_x_diag, np.diag(inp.vxc),
                              big.sigma_c, big.omegas, list(range(inp.nmo)), 80)
    assert np.array_equal(big.qp, qp_o, equal_nan=True) and np.array_equal(big.qp_iterations, it_o)


# ------------------------------------------------------------ mixed precision
MIXED_CASES = {
    # case: (input factory, nfreq, npar); fixture cases come with the reference's outputs
    "h2o_all": None, "syn_nmo24": None, "syn_small": None, "syn_odd": None,
    "syn60": (lambda: synthetic_input(60, 12, 31), 48, 16),
    "syn90": (lambda: synthetic_input(90, 18, 31), 48, 16),
}


def _mixed_case(case):
    """(input, nfreq, npar, reference fixture or None)."""
    if MIXED_CASES[case] is None:
        r = dict(np.load(os.path.join(GOLD, f"ref_{case}.npz")))
        return golden_input(case, r), int(r["nfreq"]), int(r["npar"]), r
    fac, nf, npar = MIXED_CASES[case]
    return fac(), nf, npar, None


def _check_mixed(hip, ctx, case, prec, form, sigma_tol):
    inp, nf, npar, r = _mixed_case(case)
    f64 = hip.run_g0w0(inp, nf, npar, ctx=ctx)
    mix = hip.run_g0w0(inp, nf, npar, ctx=ctx, upload=False, precision=prec)
    scale = np.abs(f64.sigma_c).max()
    d = np.abs(mix.sigma_c - f64.sigma_c).max() / scale
    msg = f"   rel |dSigma_c| vs FP64 {d:.2e}"
    if r is not None:                       # against the reference's own output too
        dr = np.abs(mix.sigma_c - r["sigma_c"]).max() / np.abs(r["sigma_c"]).max()
        msg += f", vs the reference {dr:.2e}"
        assert dr < sigma_tol
    bad, summary = mixed_qp_violations(mix.qp, f64.qp, form, case, r is not None)
    say(msg + f"; {summary}")
    assert d < sigma_tol
    assert not bad, (bad, summary)


@pytest.mark.parametrize("case", ["h2o_all", "syn_nmo24", "syn60"])
def test_mixed_vs_fp64(hip, ctx, case):
    """FP16 hi/lo split form against the FP64 form (and the reference's
    output): Sigma_c(i w) to 1e-6 relative; QP energies within 1 meV except
    the pinned states of tests/golden/qp_mev_exceptions.json."""
    say(f"case: FP16-split mixed vs FP64 {case}")
    _check_mixed(hip, ctx, case, hip.PREC_MIXED, "fp16split", 1e-6)


@pytest.mark.parametrize("case", list(MIXED_CASES))
@pytest.mark.parametrize("form", ["int8", "int8x4"])
def test_mixed_int8_vs_fp64(hip, ctx, case, form):
    """Both int8 forms against the FP64 form (and the reference's output):
    Sigma_c(i w) to SIGMA_REL of the form (five 7-bit digits truncate at
    ~2^-35 of a row's largest element: 1e-9; four 8-bit digits at ~2^-32:
    1e-8; the products are exact); QP energies within 1 meV except the pinned
    states of tests/golden/qp_mev_exceptions.json (none on H2O)."""
    say(f"case: {form} mixed vs FP64 {case}")
    prec = hip.PREC_MIXED_INT8 if form == "int8" else hip.PREC_MIXED_INT8_4
    _check_mixed(hip, ctx, case, prec, form, SIGMA_REL[form])


# ------------------------------------------------------------ decomposed-ERI form
def _golden_case(name):
    r = dict(np.load(os.path.join(GOLD, f"ref_{name}.npz")))
    if name == "h2o_all":
        inp = h2o_input()
    elif name == "syn_nmo24":
        inp = synthetic_input(*CONFIGS["nmo24"])
    elif name == "syn_gap":
        inp = with_gap(synthetic_input(30, 6, 47), 1e-4)
    else:
        inp = GwInput(mo_energy=r["mo_energy"], eri=r["eri"], vxc=r["vxc"], nocc=int(r["nocc"]),
                      fermi_energy=float(r["fermi_energy"]))
    return r, inp


@pytest.mark.parametrize("name", ["h2o_all", "syn_nmo24", "syn_small", "syn_odd", "syn_gap"])
def test_fp64_cd_vs_reference_path(hip, ctx, name):
    """Decomposed-ERI form (device pivoted Cholesky of the pair matrix, pair-space
    screening) against the reference path's all-state outputs: Sigma_c to
    1e-10 absolute; on the ill-conditioned syn_gap case (max |Sigma_c| 2472
    Ha) to 5e-11 of max |Sigma_c| (measured 8.3e-12, 2e-8 Ha): its q is the
    difference of two norms, ||L_C^-1 l||^2 - ||l||^2 (Woodbury in pair
    space), which cancels more as S^2 grows toward 4 / gap, while the plain
    form stays at 2e-15.  The QP energies are those the oracle's Pade + QP step
    gives on the same Sigma_c (the QP step of these molecules is too
    ill-conditioned for a tolerance on the QP energies alone: 1e-12 relative
    noise on Sigma_c moves some by 0.1 Ha).  The synthetic ERIs have rank
    <= 3 nmo, which the decomposition finds."""
    r, inp = _golden_case(name)
    say(f"case: FP64 decomposed-ERI form vs reference path, {name}")
    nf, npar = int(r["nfreq"]), int(r["npar"])
    res = hip.run_g0w0(inp, nf, npar, ctx=ctx, precision=hip.PREC_FP64_CD)
    rank = int(ctx.last_timings()["cd_rank"])
    dsig = np.abs(res.sigma_c - r["sigma_c"]).max()
    say(f"   rank {rank}; max |dSigma_c| = {dsig:.3e}")
    npairs = inp.nmo * (inp.nmo + 1) // 2
    assert 0 < rank <= (npairs if name == "h2o_all" else min(npairs, 3 * inp.nmo))
    assert dsig < (5e-11 * np.abs(r["sigma_c"]).max() if name == "syn_gap" else 1e-10)
    states = list(range(inp.nmo))
    qp_orc, _ = orc.solve_qp(inp.mo_energy, inp.fermi_energy, res.sigma_x_diag, np.diag(inp.vxc),
                             res.sigma_c, res.omegas, states, npar)
    dq = np.abs(res.qp - r["qp"])
    say(f"   QP vs reference: {np.sum(dq < 1e-8)}/{inp.nmo} within 1e-8 Ha (max {dq.max():.2e}); "
        f"vs the oracle's Pade/QP on this Sigma_c: max {np.nanmax(np.abs(res.qp - qp_orc)):.2e}")
    assert np.allclose(res.qp, qp_orc, rtol=0, atol=1e-12, equal_nan=True)


def test_fp64_cd_h2o_state5_fixture(hip, ctx):
    """The reference's h2o_serial E_c_before_Pade.out through the decomposed form,
    with compare_reference.py's 1e-8 component check."""
    say("case: H2O state 5 fixture, decomposed-ERI form")
    inp = h2o_input()
    fx = np.load(os.path.join(GOLD, "h2o_serial_fixtures.npz"))
    res = hip.run_g0w0(inp, 200, 16, states=[4], ctx=ctx, precision=hip.PREC_FP64_CD)
    sc, ref = res.sigma_c[0], fx["e_c_before_pade"]
    for x, y in zip(sc, ref):
        assert math.isclose(x.real, y.real, rel_tol=1e-8, abs_tol=1e-8)
        assert math.isclose(x.imag, y.imag, rel_tol=1e-8, abs_tol=1e-8)
    say(f"   max |dSigma_c| = {np.abs(sc - ref).max():.3e}")
    assert np.abs(sc - ref).max() < 1e-10
    row = fx["summary_row_orbital5_ev"]
    assert math.isclose(res.qp[4] * orc.K_HARTREE_TO_EV, row[4], rel_tol=1e-5, abs_tol=1e-5)


def test_fp64_cd_grows_factor_and_matches_fp64(hip, ctx):
    """A pair matrix of rank 300 (> the decomposition's first capacity of 256
    columns at nmo = 40) exercises the factor's growth; the result matches
    the n x n FP64 form."""
    from minigw_amd.synthetic import eri_from_factor, synthetic_factor
    say("case: decomposed form, rank 300 at nmo 40 (factor growth) vs FP64 form")
    nmo, nocc, k = 40, 8, 300
    mo, fermi, _, vxc = synthetic_factor(nmo, nocc, 11)
    rng = np.random.default_rng(4)
    b = rng.normal(0, 0.05, (nmo, nmo, k))
    b = 0.5 * (b + b.transpose(1, 0, 2))
    inp = GwInput(mo_energy=mo, eri=eri_from_factor(b.reshape(nmo * nmo, k), nmo), vxc=vxc,
                  nocc=nocc, fermi_energy=fermi)
    f64 = hip.run_g0w0(inp, 24, 16, ctx=ctx)
    cdr = hip.run_g0w0(inp, 24, 16, ctx=ctx, upload=False, precision=hip.PREC_FP64_CD)
    rank = int(ctx.last_timings()["cd_rank"])
    d = np.abs(cdr.sigma_c - f64.sigma_c).max()
    say(f"   rank {rank}; max |dSigma_c| vs FP64 form = {d:.3e}")
    assert rank == k
    assert d < 1e-11


def test_fp64_cd_tolerance_and_errors(hip, ctx):
    """A loose stopping tolerance lowers the rank; eta != 0, an ERI without
    pair symmetry and a tolerance outside [0, 1) are refused."""
    say("case: decomposed form: tolerance, unsupported inputs, argument errors")
    inp = synthetic_input(16, 4, 5)
    tight = hip.run_g0w0(inp, 16, 8, ctx=ctx, precision=hip.PREC_FP64_CD)
    r_tight = int(ctx.last_timings()["cd_rank"])
    try:
        ctx.set_cd_tolerance(0.2)
        loose = hip.run_g0w0(inp, 16, 8, ctx=ctx, upload=False, precision=hip.PREC_FP64_CD)
        r_loose = int(ctx.last_timings()["cd_rank"])
    finally:
        ctx.set_cd_tolerance(1e-12)
    say(f"   rank {r_tight} at 1e-12, {r_loose} at 0.2")
    assert r_loose < r_tight
    assert np.all(np.isfinite(loose.sigma_c))
    assert np.abs(loose.sigma_c - tight.sigma_c).max() > 1e-9
    with pytest.raises(hip.HipError) as e:
        ctx.set_cd_tolerance(1.0)
    assert e.value.code == hip.MGW_ERR_ARG
    with pytest.raises(hip.HipError) as e:
        hip.run_g0w0(inp, 16, 8, eta=0.01, ctx=ctx, upload=False, precision=hip.PREC_FP64_CD)
    assert e.value.code == hip.MGW_ERR_UNSUPPORTED
    bad = GwInput(mo_energy=inp.mo_energy, eri=inp.eri.copy(), vxc=inp.vxc, nocc=inp.nocc,
                  fermi_energy=inp.fermi_energy)
    bad.eri[0, 0, 1, 2] += 1e-3          # (00|12) != (12|00); (pk|ia) block untouched
    with pytest.raises(hip.HipError) as e:
        hip.run_g0w0(bad, 16, 8, ctx=ctx, precision=hip.PREC_FP64_CD)
    assert e.value.code == hip.MGW_ERR_UNSUPPORTED
    ok = hip.run_g0w0(inp, 16, 8, ctx=ctx, precision=hip.PREC_FP64_CD)
    assert np.abs(ok.sigma_c - tight.sigma_c).max() == 0.0


# ------------------------------------------------------------ not positive definite
def _indefinite_input(nmo, nocc, seed, alpha):
    """The seeded molecule with a rank-1 pair term subtracted from its ERI,
    eri[p,q,r,s] -= alpha U[p,q] U[r,s] (U symmetric; 8-fold symmetry kept):
    V_ph gets a negative eigenvalue, so I + S V S is indefinite at the low
    imaginary frequencies (S^2 = 4 D / (w^2 + D^2) large) and positive definite
    at the high ones."""
    inp = synthetic_input(nmo, nocc, seed)
    rng = np.random.default_rng(seed)
    u = rng.normal(0.0, 1.0, (nmo, nmo))
    u = 0.5 * (u + u.T)
    eri = inp.eri - alpha * np.einsum("pq,rs->pqrs", u, u)
    return GwInput(mo_energy=inp.mo_energy, eri=np.ascontiguousarray(eri), vxc=inp.vxc,
                   nocc=inp.nocc, fermi_energy=inp.fermi_energy)


@pytest.mark.parametrize("nmo,nocc", [(30, 6), (48, 10)])
def test_indefinite_dielectric_takes_pivoted_lu(hip, ctx, nmo, nocc):
    """V_ph with a negative eigenvalue: where I + S V S is not positive
    definite the FP64 path screens by LU with partial pivoting (the reference's
    getrf/getrs) instead of stopping; W_c and q against the oracle's LU, the
    whole chain against the oracle, and the int8 form, whose factorization
    breaks down there, rerun in FP64 by its judge's fallback."""
    say(f"case: indefinite dielectric matrix, nmo={nmo} nocc={nocc}")
    inp = _indefinite_input(nmo, nocc, 7, 0.05)
    ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
    v = orc.v_ph_matrix(inp.eri, ph)
    assert np.linalg.eigvalsh(0.5 * (v + v.T)).min() < -0.5
    ws = hip.HipScreeningWorkspace(v, ctx)
    view = hip.HipPqPhPanelView(inp.eri, inp.nocc, ctx)
    for w in (1e-3, 0.5, 40.0):
        pi0 = orc.pi0_diag_real(w, ph.delta)
        s = np.sqrt(-pi0)
        spd = np.linalg.eigvalsh(np.eye(ph.size) + s[:, None] * v * s[None, :]).min() > 0
        ws.compute_w_c(pi0)
        want = orc.compute_w_c(v, pi0)
        err = np.abs(ws.w_c().real - want).max() / np.abs(want).max()
        q = view.quadratic_forms(2, 0, nmo)
        qw = orc.quadratic_forms_panel(want, orc.pq_panel(inp.eri, ph, 2, 0, nmo))
        eq = np.abs(q.real - qw).max() / np.abs(qw).max()
        say(f"   w={w}: positive definite {spd}, W_c rel err {err:.2e}, q rel err {eq:.2e}")
        assert err < 1e-11 and eq < 1e-11
    res = hip.run_g0w0(inp, 40, 12, ctx=ctx)
    lu = int(res.timings["lu_freqs"])
    want = orc.run_g0w0(inp, 40, 12)
    d = np.abs(res.sigma_c - want.sigma_c).max() / np.abs(want.sigma_c).max()
    say(f"   G0W0: {lu} of 40 frequencies by pivoted LU, rel |dSigma_c| {d:.2e}")
    assert 0 < lu < 40 and d < 1e-11
    # the same with frequency batches smaller than the run of failing
    # frequencies: breakdowns of every batch are found and screened by LU
    nb = max(1, lu // 3)
    ctx.set_batch_limit(nb)
    try:
        split = hip.run_g0w0(inp, 40, 12, ctx=ctx, upload=False)
    finally:
        ctx.set_batch_limit(0)
    say(f"   batches of {nb}: {int(split.timings['lu_freqs'])} frequencies by pivoted LU")
    assert int(split.timings["freq_batch"]) <= nb and int(split.timings["lu_freqs"]) == lu
    assert np.array_equal(split.sigma_c, res.sigma_c)
    with pytest.raises(hip.HipError) as e:
        hip.run_g0w0(inp, 40, 12, ctx=ctx, upload=False, precision=hip.PREC_MIXED_INT8)
    assert e.value.code == hip.MGW_ERR_NUMERIC
    ctx.set_judge(2)
    try:
        fb = hip.run_g0w0(inp, 40, 12, ctx=ctx, upload=False, precision=hip.PREC_MIXED_INT8)
        assert ctx.last_judge()["fell_back"] == 2
    finally:
        ctx.set_judge(0)
    assert np.array_equal(fb.sigma_c, res.sigma_c)
