"""CPU checks of the C ABI library: it loads, exports every declared symbol,
and its host-only entry points (frequency grid, memory estimate) agree with
the oracle.  No device work is started here."""
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    inc = os.path.join(REPO, "include")
    names = set()
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, h)).read(), flags=re.S)
            names |= set(re.findall(r"\b(mgw_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from minigw_amd import hip
    names = declared_symbols()
    assert len(names) >= 40
    assert "mgw_input_read" in names and "mgw_hip_pade_real_axis" in names
    missing = [n for n in names if not hasattr(hip.lib, n)]
    assert not missing, missing


def test_version_string():
    from minigw_amd import hip
    assert b"gfx950" in hip.lib.mgw_hip_version()


@pytest.mark.parametrize("nfreq", [1, 7, 48, 200])
def test_grid_matches_oracle(nfreq):
    from minigw_amd import hip
    from oracle import oracle as orc
    om, w = hip.transformed_legendre_grid(nfreq)
    oom, ow = orc.transformed_legendre_grid(nfreq)
    assert np.array_equal(om, oom)
    assert np.array_equal(w, ow)


def test_grid_matches_reference_fixture(golden_dir):
    from minigw_amd import hip
    r = np.load(os.path.join(golden_dir, "ref_h2o_all.npz"))
    om, w = hip.transformed_legendre_grid(200)
    assert np.array_equal(om, r["omega"]) and np.array_equal(w, r["weight"])


def test_grid_rejects_bad_size():
    from minigw_amd import hip
    with pytest.raises(hip.HipError) as e:
        hip.transformed_legendre_grid(0)
    assert e.value.code == hip.MGW_ERR_ARG
    assert "num_freq_points must be positive" in str(e.value)


def test_estimate_bytes_scales_like_eri_plus_batch():
    from minigw_amd import hip
    nmo200 = hip.estimate_bytes(200, 40, 200, 0, 16, hip.PREC_FP64)
    eri = 200 ** 4 * 8
    assert eri < nmo200 < eri + 16 * 6400 ** 2 * 8 * 1.2 + 2e9
    assert hip.estimate_bytes(200, 40, 200, 0, 32, hip.PREC_FP64) > nmo200
    assert hip.estimate_bytes(200, 40, 200, 0, 16, hip.PREC_MIXED) > nmo200
    assert hip.estimate_bytes(1, 0, 10) == 0
