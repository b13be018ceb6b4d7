"""gw_input readers and the output writers (include/mgw_io.h) on the CPU.

Byte compatibility is checked against the reference's own regression files:
their values are held in tests/golden/h2o_serial_fixtures.npz (parsed from
the 17-digit text, so they print back to the same text), and against the
files themselves where the reference tree is present."""
import math
import os

import numpy as np
import pytest

from minigw_amd.synthetic import GwInput, synthetic_input

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = "/root/reference/regression_tests/h2o_serial"


@pytest.fixture(scope="module")
def gwio():
    from minigw_amd import gwio as g
    return g


def h2o_input():
    d = np.load(os.path.join(GOLD, "h2o_input.npz"))
    return GwInput(mo_energy=d["mo_energy"], eri=d["eri"], vxc=d["vxc"], nocc=int(d["nocc"]),
                   fermi_energy=float(d["fermi_energy"]))


def test_raw_round_trip(gwio, tmp_path):
    inp = h2o_input()
    p = str(tmp_path / "gw_input.raw")
    gwio.write_gw_input_raw(p, inp)
    h = gwio.probe_gw_input(p)
    assert h == {"nmo": 18, "nocc": 5, "fermi_energy": inp.fermi_energy, "format": gwio.FORMAT_RAW}
    back = gwio.read_gw_input(p)
    assert np.array_equal(back.mo_energy, inp.mo_energy)
    assert np.array_equal(back.eri, inp.eri)
    assert np.array_equal(back.vxc, inp.vxc)
    assert back.nocc == inp.nocc and back.fermi_energy == inp.fermi_energy


def test_input_errors(gwio, tmp_path):
    from minigw_amd import hip
    bad = tmp_path / "bad.raw"
    bad.write_bytes(b"not a gw input file")
    with pytest.raises(hip.HipError) as e:
        gwio.read_gw_input(str(bad))
    assert e.value.code == hip.MGW_ERR_IO
    with pytest.raises(hip.HipError) as e:
        gwio.read_gw_input(str(tmp_path / "missing.h5"))
    assert e.value.code == hip.MGW_ERR_IO
    inp = synthetic_input(6, 2, 1)
    p = str(tmp_path / "t.raw")
    gwio.write_gw_input_raw(p, inp)
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-8])                     # truncated
    with pytest.raises(hip.HipError) as e:
        gwio.read_gw_input(p)
    assert e.value.code == hip.MGW_ERR_IO
    inp.nocc = 6                                       # nocc must be in [1, nmo-1]
    with pytest.raises(hip.HipError) as e:
        gwio.write_gw_input_raw(p, inp)
    assert e.value.code == hip.MGW_ERR_ARG and "nocc must be in [1, nmo-1]" in str(e.value)


def test_hdf5_reader_on_reference_input(gwio):
    path = os.path.join(REF, "gw_input.h5")
    if not os.path.exists(path):
        pytest.skip("reference tree absent")
    from minigw_amd import hip
    try:
        got = gwio.read_gw_input(path)
    except hip.HipError as e:
        if e.code == hip.MGW_ERR_UNSUPPORTED:
            pytest.skip(f"no libhdf5: {e}")
        raise
    want = h2o_input()
    assert np.array_equal(got.mo_energy, want.mo_energy)
    assert np.array_equal(got.eri, want.eri)
    assert np.array_equal(got.vxc, want.vxc)
    assert got.nocc == want.nocc and got.fermi_energy == want.fermi_energy


def test_sigma_before_pade_bytes(gwio, tmp_path):
    fx = np.load(os.path.join(GOLD, "h2o_serial_fixtures.npz"))
    p = tmp_path / "E_c_before_Pade.out"
    gwio.write_sigma_before_pade(str(p), fx["e_c_before_pade"])
    text = p.read_text()
    want = "".join(f"({z.real:.17g},{z.imag:.17g})\n" for z in fx["e_c_before_pade"])
    assert text == want
    ref = os.path.join(REF, "E_c_before_Pade.out")
    if os.path.exists(ref):
        assert text == open(ref).read()


def test_sigma_after_pade_and_grid_bytes(gwio, tmp_path):
    """The E_c.out writer's text from the reference's own values; the energy
    column is recomputed from x_start, interval and E_F as src/io.cpp does."""
    fx = np.load(os.path.join(GOLD, "h2o_serial_fixtures.npz"))
    ec = fx["e_c"]
    inp = h2o_input()
    p = tmp_path / "E_c.out"
    gwio.write_sigma_after_pade(str(p), ec[:, 1] + 1j * ec[:, 2], inp.fermi_energy)
    text = p.read_text()
    want = "".join(f"{a:.17g} {b:.17g} {c:.17g}\n" for a, b, c in ec)
    assert text == want
    ref = os.path.join(REF, "E_c.out")
    if os.path.exists(ref):
        assert text == open(ref).read()
    from minigw_amd import hip
    om, w = hip.transformed_legendre_grid(200)
    g = tmp_path / "grid.out"
    gwio.write_frequency_grid(str(g), om, w)
    assert g.read_text() == "".join(f"{a:.17g} {b:.17g}\n" for a, b in zip(om, w))


def test_energy_table_format(gwio):
    """Header and rule lines as the reference prints them; rows parse with the
    comparison script's rule (six fields, first the 1-based orbital)."""
    inp = h2o_input()
    rng = np.random.default_rng(0)
    sx = -rng.random(18)
    qp = inp.mo_energy + 0.01
    table = gwio.format_energy_table(inp.mo_energy, sx, np.diag(inp.vxc), qp, [4, 0, 17])
    lines = table.splitlines()
    assert lines[0] == ("Orbital   KS Energy      Sigma_x        Sigma_c        Vxc            "
                        "QP Energy      ")
    assert lines[1] == ("-------   ---------      -------        -------        -------        "
                        "---------      ")
    K = 27.211386245988
    for line, p in zip(lines[2:], [4, 0, 17]):
        sc = qp[p] + inp.vxc[p, p] - inp.mo_energy[p] - sx[p]
        vals = [inp.mo_energy[p] * K, sx[p] * K, sc * K, inp.vxc[p, p] * K, qp[p] * K]
        assert line == f"{p + 1:<10d}" + "".join(f"{v:<15.6g}" for v in vals)
        parts = line.split()
        assert len(parts) == 6 and parts[0] == str(p + 1)
        assert all(math.isclose(float(a), b, rel_tol=1e-5) for a, b in zip(parts[1:], vals))


@pytest.mark.parametrize("argv,message", [
    (["--freq-points", "0"], "--freq-points must be positive"),
    (["--pade-params", "0"], "--pade-params must be positive"),
    (["--contraction-panel-size", "0"], "--contraction-panel-size must be positive"),
    (["--tasks-per-gpu", "0"], "--tasks-per-gpu must be positive"),
    (["--scalapack-ranks-per-group", "0"], "--scalapack-ranks-per-group must be positive"),
    (["--frequency-parallel", "threads"], "--frequency-parallel must be one of"),
    (["--kernel-parallel", "mpi"], "--kernel-parallel must be one of"),
])
def test_runner_checks_options_as_the_reference_cli(tmp_path, argv, message):
    """The runner refuses the option values the reference CLI refuses, with
    its messages (src/cli.cpp:75-100), before it touches a device."""
    from minigw_amd import run
    with pytest.raises(SystemExit) as e:
        run.main(["--input-dir", str(tmp_path), "--output-dir", str(tmp_path / "o")] + argv)
    assert message in str(e.value)


def test_runner_backend_names():
    """Backend names select the forms as the reference-side hip backend does."""
    from minigw_amd import run
    assert run.BACKENDS == {"hip": "fp64", "hip-int8": "mixed_int8", "hip-int8x4": "mixed_int8x4",
                            "hip-fp16": "mixed", "hip-cd": "fp64_cd"}
