"""The reference's own `gw` executable with the hip backend linked in
(oracle/_ref/gw_hip, built from the reference sources and the integration
guide by oracle/build_integration.py) on the H2O regression input: the
reference's main, backend factory, driver, QP solver and output writers, with
Sigma_c from this library.  Its output files are compared with the reference's
h2o_serial outputs by the regression test's own checks.  Needs an MI355X."""
import math
import os
import subprocess

import numpy as np
import pytest

from minigw_amd.synthetic import GwInput

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")
EXE = os.path.join(REPO, "oracle", "_ref", "gw_hip")
HA_TO_EV = 27.211386245988

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/gw_hip not built "
                               "(python oracle/build_integration.py, needs the reference tree)")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def h2o_dir(tmp_path_factory):
    from minigw_amd import gwio
    d = np.load(os.path.join(GOLD, "h2o_input.npz"))
    inp = GwInput(mo_energy=d["mo_energy"], eri=d["eri"], vxc=d["vxc"], nocc=int(d["nocc"]),
                  fermi_energy=float(d["fermi_energy"]))
    path = tmp_path_factory.mktemp("h2o")
    gwio.write_gw_input_raw(str(path / "gw_input.h5"), inp)   # main reads gw_input.h5
    return path


def run_gw(h2o_dir, out, form, *extra):
    out.mkdir()
    res = subprocess.run([EXE, "--input-dir", str(h2o_dir), "--output-dir", str(out),
                          "--linalg-backend", form, "--state", "5"] + list(extra),
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr
    return res.stdout


def read_before_pade(out):
    vals = []
    for line in (out / "E_c_before_Pade.out").read_text().splitlines():
        re_, im_ = line.strip().strip("()").split(",")
        vals.append(complex(float(re_), float(im_)))
    return np.array(vals)


def summary_row(out, state=5):
    text = (out / "gw.out").read_text().splitlines()
    start = text.index("--- G0W0 Calculation Summary ---")
    for line in text[start + 1:]:
        if line.strip().startswith("--- End"):
            break
        parts = line.split()
        if len(parts) == 6 and parts[0] == str(state):
            return [float(x) for x in parts[1:]]
    return None


@needs_exe
@pytest.mark.parametrize("form", ["hip", "hip-int8", "hip-int8x4"])
def test_gw_executable_h2o_against_regression_outputs(h2o_dir, tmp_path, form):
    """E_c_before_Pade.out within the regression test's 1e-8 component check
    (FP64 also within 1e-11 of it), E_c.out within the spread of the
    reference's own backends (2e-6 Ha) in FP64 and within 1 meV in the int8
    forms, the summary row within scripts/compare_reference.py's 1e-5 check.
    The real-axis continuation is ill-conditioned: white noise of 8e-11 of
    max |Sigma_c| on the FP64 Sigma_c moves it by up to 0.1 Ha near -110 eV
    and by 2e-2 Ha inside |E - E_F| <= 1 Ha (oracle.pade_real_axis), so no
    bound on E_c.out follows from a form's Sigma_c precision; the forms' errors
    are smooth across frequencies, which keeps the curves far closer.  The
    four-digit form (hip-int8x4) passes the reference's own regression checks
    (the E_c_before_Pade.out components and the summary row) and its E_c.out
    is held at the order measured on the device (3.3e-4 Ha over the curve,
    4.9e-5 Ha within |E - E_F| <= 1 Ha): below 1e-3 Ha and 1e-4 Ha."""
    say(f"case: reference gw executable, --linalg-backend {form}, H2O state 5, 200 frequencies")
    fx = np.load(os.path.join(GOLD, "h2o_serial_fixtures.npz"))
    out = tmp_path / "out"
    stdout = run_gw(h2o_dir, out, form)
    assert "Using the hip device path on device 0" in stdout
    if form.startswith("hip-int8"):
        assert "hip judge: largest relative error of q" in stdout
    got = read_before_pade(out)
    want = fx["e_c_before_pade"]
    assert len(got) == 200
    for a, b in zip(got, want):
        assert math.isclose(a.real, b.real, rel_tol=1e-8, abs_tol=1e-8)
        assert math.isclose(a.imag, b.imag, rel_tol=1e-8, abs_tol=1e-8)
    if form == "hip":
        assert np.abs(got - want).max() < 1e-11
    ec = np.loadtxt(out / "E_c.out")
    assert ec.shape == (19836, 3) and np.array_equal(ec[:, 0], fx["e_c"][:, 0])
    tol = {"hip": 2e-6, "hip-int8": 1e-3 / HA_TO_EV, "hip-int8x4": 1e-3}[form]
    dev = np.abs(ec[:, 1:] - fx["e_c"][:, 1:]).max(axis=1)
    say(f"   E_c.out: largest deviation {dev.max():.2e} Ha at {ec[dev.argmax(), 0]:.2f} eV")
    assert dev.max() < tol
    if form == "hip-int8x4":
        fermi_ev = float(np.load(os.path.join(GOLD, "h2o_input.npz"))["fermi_energy"]) * HA_TO_EV
        window = np.abs(ec[:, 0] - fermi_ev) <= HA_TO_EV
        say(f"   window |E - E_F| <= 1 Ha: largest deviation {dev[window].max():.2e} Ha")
        assert dev[window].max() < 1e-4
    row = summary_row(out)
    assert row is not None
    for a, b in zip(row, fx["summary_row_orbital5_ev"]):
        assert math.isclose(a, b, rel_tol=1e-5, abs_tol=1e-5), (row, fx["summary_row_orbital5_ev"])


@needs_exe
def test_gw_executable_per_frequency_workspace_matches_hoisted(h2o_dir, tmp_path):
    """--linalg-backend hip-workspace runs the reference's per-frequency loop on
    HipScreeningWorkspace / HipPqPhPanelView (one factorization per (w_n, w'),
    panels filled on one context and contracted on another); it must agree
    with the hoisted schedule to FP64 rounding."""
    say("case: reference gw executable, hip-workspace vs hip, H2O state 5, 40 frequencies")
    a = run_gw(h2o_dir, tmp_path / "ws", "hip-workspace", "--freq-points", "40")
    assert "per-frequency workspace" in a
    run_gw(h2o_dir, tmp_path / "ho", "hip", "--freq-points", "40")
    ws = read_before_pade(tmp_path / "ws")
    ho = read_before_pade(tmp_path / "ho")
    assert len(ws) == 40
    assert np.abs(ws - ho).max() <= 1e-12 * np.abs(ho).max()
