"""gw_input files and the G0W0 output files over the C ABI of include/mgw_io.h.

read_gw_input mirrors read_gw_input_hdf5 (src/hdf5_input.cpp:192-212): HDF5
through the system libhdf5 loaded at run time, or the raw fallback written by
write_gw_input_raw.  The writers produce E_c_before_Pade.out, E_c.out, the
frequency grid and the summary table byte for byte as src/io.cpp:79-126 does;
E_c.out's real-axis Pade values come from the device (mgw_hip_pade_real_axis).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip
from .synthetic import GwInput

_lib = hip.lib
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_sig = {
    "mgw_input_probe": (C.c_int, [C.c_char_p, _ip, _ip, _dp, _ip]),
    "mgw_input_read": (C.c_int, [C.c_char_p, C.c_int32, _dp, _dp, _dp]),
    "mgw_input_write_raw": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_double, _dp, _dp, _dp]),
    "mgw_write_frequency_grid": (C.c_int, [C.c_char_p, _dp, _dp, C.c_int32]),
    "mgw_write_sigma_before_pade": (C.c_int, [C.c_char_p, _dp, _dp, C.c_int32]),
    "mgw_write_sigma_after_pade": (C.c_int, [C.c_char_p, C.c_double, C.c_double, C.c_double,
                                             C.c_int32, _dp, _dp]),
    "mgw_format_energy_table": (C.c_int, [C.c_int32, _dp, _dp, _dp, _dp, _ip, C.c_int32,
                                          C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "mgw_io_last_error": (C.c_char_p, []),
    "mgw_hip_pade_real_axis": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int32, C.c_int32, C.c_double,
                                         C.c_double, C.c_double, C.c_int32, _dp, _dp]),
}
for _name, (_res, _args) in _sig.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

FORMAT_HDF5, FORMAT_RAW = 1, 2
E_C_POINTS, E_C_X_START, E_C_INTERVAL = 19836, -6.5, 0.00065     # include/io.hpp:29-31, main.cpp:118


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(_dp)


def _call(name: str, *args):
    rc = getattr(_lib, name)(*args)
    if rc != hip.MGW_OK:
        raise hip.HipError(name, rc, (_lib.mgw_io_last_error() or b"").decode())


def probe_gw_input(path: str) -> dict:
    nmo, nocc, fmt = C.c_int32(), C.c_int32(), C.c_int32()
    fermi = C.c_double()
    _call("mgw_input_probe", path.encode(), C.byref(nmo), C.byref(nocc), C.byref(fermi), C.byref(fmt))
    return {"nmo": nmo.value, "nocc": nocc.value, "fermi_energy": fermi.value, "format": fmt.value}


def read_gw_input(path: str) -> GwInput:
    h = probe_gw_input(path)
    n = h["nmo"]
    mo, eri, vxc = np.zeros(n), np.zeros((n, n, n, n)), np.zeros((n, n))
    _call("mgw_input_read", path.encode(), n, _p(mo), _p(eri), _p(vxc))
    return GwInput(mo_energy=mo, eri=eri, vxc=vxc, nocc=h["nocc"], fermi_energy=h["fermi_energy"])


def write_gw_input_raw(path: str, inp: GwInput) -> None:
    n = len(inp.mo_energy)
    _call("mgw_input_write_raw", path.encode(), n, int(inp.nocc), float(inp.fermi_energy),
          _p(_f64(inp.mo_energy)), _p(_f64(inp.eri)), _p(_f64(inp.vxc)))


def write_frequency_grid(path: str, omegas, weights) -> None:
    o, w = _f64(omegas), _f64(weights)
    _call("mgw_write_frequency_grid", path.encode(), _p(o), _p(w), len(o))


def write_sigma_before_pade(path: str, sigma_row) -> None:
    s = np.asarray(sigma_row)
    re, im = _f64(s.real), _f64(np.imag(s))
    _call("mgw_write_sigma_before_pade", path.encode(), _p(re), _p(im), len(re))


def pade_real_axis(sigma_row, omegas, npar: int, fermi: float, npoints: int = E_C_POINTS,
                   x_start: float = E_C_X_START, interval: float = E_C_INTERVAL,
                   ctx: hip.HipContext | None = None) -> np.ndarray:
    """Device Thiele Pade of one state's Sigma_c(i w) at x_start + i * interval - fermi."""
    ctx = ctx or hip.HipContext()
    s = np.asarray(sigma_row)
    re, im, om = _f64(s.real), _f64(np.imag(s)), _f64(omegas)
    ore, oim = np.zeros(npoints), np.zeros(npoints)
    ctx.call("mgw_hip_pade_real_axis", _p(re), _p(im), _p(om), len(om), int(npar), float(fermi),
             float(x_start), float(interval), int(npoints), _p(ore), _p(oim))
    return ore + 1j * oim


def write_sigma_after_pade(path: str, values, fermi: float, x_start: float = E_C_X_START,
                           interval: float = E_C_INTERVAL) -> None:
    v = np.asarray(values)
    re, im = _f64(v.real), _f64(np.imag(v))
    _call("mgw_write_sigma_after_pade", path.encode(), float(x_start), float(interval), float(fermi),
          len(re), _p(re), _p(im))


def format_energy_table(mo_energy, sigma_x_diag, vxc_diag, qp, states) -> str:
    st = np.ascontiguousarray(states, dtype=np.int32)
    args = (len(mo_energy), _p(_f64(mo_energy)), _p(_f64(sigma_x_diag)), _p(_f64(vxc_diag)),
            _p(_f64(qp)), st.ctypes.data_as(_ip), len(st))
    need = C.c_size_t()
    _call("mgw_format_energy_table", *args, None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value + 1)
    _call("mgw_format_energy_table", *args, buf, need.value + 1, C.byref(need))
    return buf.value.decode()
