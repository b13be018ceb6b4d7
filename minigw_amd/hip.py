"""Python mirror of the `hip` backend's boundary, over the C ABI (include/mgw_hip.h).

Names and argument meaning follow the reference's device screening interface
so tests read like the reference's own:

* ``HipScreeningWorkspace(v_ph)`` -- DeviceScreeningWorkspace
  (include/workspace/device_screening_workspace.hpp:27-59): ``compute_w_c(pi0)``,
  ``quadratic_forms_panel(x_panel)``, ``nph``, ``estimated_device_bytes``.
* ``HipPqPhPanelView(eri, nocc, ctx, mode)`` -- DevicePqPhPanelView
  (include/workspace/device_pq_ph_panel.hpp:67-99) with its storage mode
  (``ERI_AUTO`` / ``ERI_RESIDENT`` / ``ERI_STREAMED``, the reference's Auto /
  FullEriResident / StreamingPanel), whose panels feed
  ``quadratic_forms(p, k0, width)``.
* ``run_g0w0(inp, ...)`` -- run_g0w0 (src/gw/gw.cpp:125-191) with the hoisted,
  frequency-batched device path.

Errors raise ``HipError("mgw_hip_<fn> failed: <message>")`` like the
reference's check_* helpers (src/workspace/device_screening_workspace.cu:25-43).
The library is loaded from minigw_amd/lib/libmgw_hip.so; if it is missing the
import fails -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmgw_hip.so")

MGW_OK, MGW_ERR_ARG, MGW_ERR_HIP, MGW_ERR_NUMERIC, MGW_ERR_UNSUPPORTED, MGW_ERR_MEMORY, \
    MGW_ERR_STATE, MGW_ERR_IO = range(8)
PREC_FP64, PREC_MIXED, PREC_MIXED_INT8, PREC_FP64_CD, PREC_MIXED_INT8_4 = 0, 1, 2, 3, 4
ERI_AUTO, ERI_RESIDENT, ERI_STREAMED = 0, 1, 2
ERI_STORAGE_NAMES = {ERI_AUTO: "auto", ERI_RESIDENT: "full-eri-resident", ERI_STREAMED: "streaming-panel"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class HipError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed: {msg}")
        self.code = code


class G0W0Job(C.Structure):
    _fields_ = [("nmo", C.c_int32), ("nocc", C.c_int32), ("nfreq", C.c_int32),
                ("npar", C.c_int32), ("nstates", C.c_int32), ("states", _ip),
                ("mo_energy", _dp), ("vxc_diag", _dp), ("fermi_energy", C.c_double),
                ("eta", C.c_double), ("precision", C.c_int32)]


class G0W0Out(C.Structure):
    _fields_ = [("omega", _dp), ("weight", _dp), ("sigma_x_diag", _dp), ("sigma_c_re", _dp),
                ("sigma_c_im", _dp), ("qp", _dp), ("qp_iterations", _ip), ("q", _dp)]


class ServeJob(C.Structure):
    _fields_ = [("job", G0W0Job), ("eri", _dp), ("eri_factor", _dp), ("kdim", C.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"HIP library not built: {LIB_PATH} (run minigw_amd/build.py)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "mgw_hip_version": (C.c_char_p, []),
        "mgw_hip_last_error": (C.c_char_p, [vp]),
        "mgw_hip_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "mgw_hip_device_for_local_rank": (C.c_int, [C.c_size_t, C.c_size_t, C.POINTER(C.c_int)]),
        "mgw_hip_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "mgw_hip_destroy": (C.c_int, [vp]),
        "mgw_hip_device_bytes": (C.c_int, [vp, C.POINTER(C.c_size_t)]),
        "mgw_hip_set_memory_budget": (C.c_int, [vp, C.c_size_t]),
        "mgw_hip_upload_vph": (C.c_int, [vp, _dp, C.c_int32]),
        "mgw_hip_compute_wc": (C.c_int, [vp, _dp, _dp, C.c_int32]),
        "mgw_hip_download_wc": (C.c_int, [vp, _dp, _dp, C.c_int32]),
        "mgw_hip_quadratic_forms_host": (C.c_int, [vp, _dp, C.c_int32, C.c_int32, _dp, _dp]),
        "mgw_hip_upload_eri": (C.c_int, [vp, _dp, C.c_int32, C.c_int32]),
        "mgw_hip_upload_eri_factor": (C.c_int, [vp, _dp, C.c_int32, C.c_int32, C.c_int32]),
        "mgw_hip_download_eri": (C.c_int, [vp, _dp]),
        "mgw_hip_set_eri_storage": (C.c_int, [vp, C.c_int32]),
        "mgw_hip_eri_storage": (C.c_int, [vp, _ip, C.POINTER(C.c_size_t)]),
        "mgw_hip_eri_bytes": (C.c_size_t, [C.c_int32] * 3),
        "mgw_hip_release_buffers": (C.c_int, [vp]),
        "mgw_hip_build_vph_from_eri": (C.c_int, [vp]),
        "mgw_hip_quadratic_forms": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, _dp, _dp]),
        "mgw_hip_fill_panel": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp),
                                         C.POINTER(C.c_int32)]),
        "mgw_hip_quadratic_forms_device": (C.c_int, [vp, vp, C.c_int32, C.c_int32, C.c_int32, _dp,
                                                     _dp]),
        "mgw_hip_g0w0": (C.c_int, [vp, C.POINTER(G0W0Job), C.POINTER(G0W0Out)]),
        "mgw_hip_g0w0_group": (C.c_int, [C.POINTER(vp), C.POINTER(G0W0Job), C.POINTER(G0W0Out),
                                         C.c_int32, _ip]),
        "mgw_hip_pade_qp": (C.c_int, [vp, C.c_int32, C.c_int32, _ip, C.c_int32, C.c_int32, _dp,
                                      _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, _ip]),
        "mgw_hip_last_timings": (C.c_int, [vp, _dp, C.c_int32]),
        "mgw_hip_set_graphs": (C.c_int, [vp, C.c_int32]),
        "mgw_hip_set_cd_tolerance": (C.c_int, [vp, C.c_double]),
        "mgw_hip_set_judge": (C.c_int, [vp, C.c_int32, C.c_double, C.c_int32]),
        "mgw_hip_last_judge": (C.c_int, [vp, _dp, _ip, _ip]),
        "mgw_hip_autotune": (C.c_int, [vp, C.POINTER(G0W0Job), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_double)]),
        "mgw_hip_autotune_result": (C.c_int, [C.POINTER(G0W0Job), C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int32)]),
        "mgw_hip_selftest_tile": (C.c_int, [vp, _dp, _dp, _dp]),
        "mgw_hip_probe_mfma": (C.c_int, [vp, C.c_int32, C.c_int32, vp, vp, vp]),
        "mgw_hip_measure_peaks": (C.c_int, [vp, _dp]),
        "mgw_grid_transformed_legendre": (C.c_int, [C.c_int32, _dp, _dp]),
        "mgw_hip_estimate_bytes": (C.c_size_t, [C.c_int32] * 6),
        "mgw_hip_server_create": (C.c_int, [C.c_int, C.c_int32, C.c_size_t, C.c_int32, C.POINTER(vp)]),
        "mgw_hip_server_destroy": (C.c_int, [vp]),
        "mgw_hip_server_submit": (C.c_int, [vp, C.POINTER(ServeJob), C.POINTER(G0W0Out),
                                            C.POINTER(C.c_int64)]),
        "mgw_hip_server_wait": (C.c_int, [vp, C.c_int64, C.c_double, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "mgw_hip_server_job_error": (C.c_char_p, [vp, C.c_int64]),
        "mgw_hip_server_set_judge": (C.c_int, [vp, C.c_int32, C.c_double, C.c_int32]),
        "mgw_hip_server_set_budget_scale": (C.c_int, [vp, C.c_double]),
        "mgw_hip_server_set_grouping": (C.c_int, [vp, C.c_int32]),
        "mgw_hip_server_set_eri_storage": (C.c_int, [vp, C.c_int32]),
        "mgw_hip_server_job_eri_storage": (C.c_int, [vp, C.c_int64, _ip]),
        "mgw_hip_server_cancel": (C.c_int, [vp, C.c_int64]),
        "mgw_hip_server_job_judge": (C.c_int, [vp, C.c_int64, _dp, _ip]),
        "mgw_hip_server_release": (C.c_int, [vp, C.c_int64]),
        "mgw_hip_set_cancel_flag": (C.c_int, [vp, _ip]),
        "mgw_hip_set_batch_limit": (C.c_int, [vp, C.c_int32]),
        "mgw_hip_server_stats": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                           C.POINTER(C.c_int32), C.POINTER(C.c_size_t)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _ptr(a: np.ndarray | None, kind=_dp):
    return None if a is None else a.ctypes.data_as(kind)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(fn: str, rc: int, ctx=None):
    if rc != MGW_OK:
        msg = lib.mgw_hip_last_error(ctx)
        raise HipError(fn, rc, msg.decode() if msg else "")


def transformed_legendre_grid(nfreq: int):
    om = np.zeros(nfreq)
    w = np.zeros(nfreq)
    _check("mgw_grid_transformed_legendre", lib.mgw_grid_transformed_legendre(nfreq, _ptr(om), _ptr(w)))
    return om, w


def estimate_bytes(nmo: int, nocc: int, nfreq: int, nstates: int = 0, nbatch: int = 0,
                   precision: int = PREC_FP64) -> int:
    return int(lib.mgw_hip_estimate_bytes(nmo, nocc, nfreq, nstates, nbatch, precision))


def eri_bytes(nmo: int, nocc: int, storage: int) -> int:
    """Device bytes an ERI of this size holds in a storage (ERI_RESIDENT or
    ERI_STREAMED, the latter with its upload slab)."""
    return lib.mgw_hip_eri_bytes(nmo, nocc, storage)


def device_count() -> int:
    """Visible HIP devices (0 where the runtime finds none)."""
    n = C.c_int(0)
    return n.value if lib.mgw_hip_device_count(C.byref(n)) == MGW_OK else 0


def device_for_local_rank(local_rank: int, tasks_per_gpu: int) -> int:
    d = C.c_int(0)
    _check("mgw_hip_device_for_local_rank", lib.mgw_hip_device_for_local_rank(local_rank, tasks_per_gpu, C.byref(d)))
    return d.value


class HipContext:
    """One device context (one HIP stream); not thread-safe."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check("mgw_hip_create", lib.mgw_hip_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib.mgw_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def call(self, name: str, *args):
        _check(name, getattr(lib, name)(self.h, *args), self.h)

    @property
    def device_bytes(self) -> int:
        b = C.c_size_t(0)
        self.call("mgw_hip_device_bytes", C.byref(b))
        return b.value

    def set_memory_budget(self, nbytes: int):
        self.call("mgw_hip_set_memory_budget", nbytes)

    def set_eri_storage(self, storage: int):
        """Storage of the ERI for the uploads that follow (ERI_AUTO by default)."""
        self.call("mgw_hip_set_eri_storage", int(storage))

    def eri_storage(self) -> tuple:
        """(storage the uploaded ERI took, device bytes it holds)."""
        st, nb = C.c_int32(0), C.c_size_t(0)
        self.call("mgw_hip_eri_storage", C.byref(st), C.byref(nb))
        return st.value, nb.value

    def upload_eri(self, eri: np.ndarray, nocc: int):
        e = _f64(eri)
        nmo = e.shape[0]
        assert e.shape == (nmo, nmo, nmo, nmo)
        self.call("mgw_hip_upload_eri", _ptr(e), nmo, nocc)

    def upload_eri_factor(self, b: np.ndarray, nmo: int, nocc: int):
        """ERI built on the device from B (nmo^2 x K): eri[pq][rs] = sum_K B[pq,K] B[rs,K]."""
        bb = _f64(b)
        assert bb.shape[0] == nmo * nmo
        self.call("mgw_hip_upload_eri_factor", _ptr(bb), nmo, nocc, bb.shape[1])

    def download_eri(self, nmo: int) -> np.ndarray:
        out = np.zeros((nmo, nmo, nmo, nmo))
        self.call("mgw_hip_download_eri", _ptr(out))
        return out

    def set_graphs(self, enable: bool):
        """Capture/replay each frequency batch as a hipGraph (no per-phase timings then)."""
        self.call("mgw_hip_set_graphs", 1 if enable else 0)

    def set_cd_tolerance(self, tol: float):
        """Stopping tolerance of the PREC_FP64_CD pair-matrix decomposition
        (relative to its first pivot; default 1e-12)."""
        self.call("mgw_hip_set_cd_tolerance", float(tol))

    def set_cancel_flag(self, flag):
        """flag: a ctypes c_int32 polled between frequency batches (None removes it)."""
        self._cancel = flag
        self.call("mgw_hip_set_cancel_flag", None if flag is None else C.pointer(flag))

    def set_batch_limit(self, max_freqs: int):
        self.call("mgw_hip_set_batch_limit", int(max_freqs))

    def set_judge(self, nsample: int, tol: float = 0.0, fallback: bool = True):
        """Error judge of the mixed forms: nsample frequencies recomputed in FP64
        (0 = off); with fallback, a failed judge or a factorization breakdown
        reruns the job in FP64."""
        self.call("mgw_hip_set_judge", int(nsample), float(tol), 1 if fallback else 0)

    def last_judge(self) -> dict:
        e, n, f = C.c_double(0), C.c_int32(0), C.c_int32(0)
        self.call("mgw_hip_last_judge", C.byref(e), C.byref(n), C.byref(f))
        return {"max_rel_err": e.value, "samples": n.value, "fell_back": f.value}

    def release_buffers(self):
        self.call("mgw_hip_release_buffers")

    def last_timings(self) -> dict:
        t = np.zeros(15)
        self.call("mgw_hip_last_timings", _ptr(t), 15)
        keys = ["build_v", "dielectric", "factorize", "contract", "sigma_c", "pade_qp", "total",
                "freq_batch", "contract_launches", "split_reduce", "contract_columns",
                "graph_batches", "cd_rank", "cd_wall_ms", "lu_freqs"]
        return dict(zip(keys, t.tolist()))

    def probe_mfma(self, kind: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        """Raw single-MFMA results: kind 0 f64 (a, b: [t, 64]); 1 bf16 / 2 f16 16x16x32,
        3 bf16 / 4 f16 32x32x16 (a, b: [t, 64, 8] uint16); 5 i8 32x32x32
        (a, b: [t, 64, 16] int8, result int32)."""
        if kind == 0:
            a, b = _f64(a), _f64(b)
            d = np.zeros((a.shape[0], 64, 4))
        elif kind == 5:
            a = np.ascontiguousarray(a, dtype=np.int8)
            b = np.ascontiguousarray(b, dtype=np.int8)
            d = np.zeros((a.shape[0], 64, 16), dtype=np.int32)
        else:
            a = np.ascontiguousarray(a, dtype=np.uint16)
            b = np.ascontiguousarray(b, dtype=np.uint16)
            d = np.zeros((a.shape[0], 64, 4 if kind <= 2 else 16), dtype=np.float32)
        self.call("mgw_hip_probe_mfma", kind, a.shape[0], a.ctypes.data, b.ctypes.data, d.ctypes.data)
        return d

    def measure_peaks(self) -> dict:
        """Measured MFMA rates (flop/s or op/s; constant and random operands) and HBM
        copy bytes/s of this device."""
        o = np.zeros(10)
        self.call("mgw_hip_measure_peaks", _ptr(o))
        return {"fp64_mfma": o[0], "fp16_mfma": o[1], "bf16_mfma": o[2], "hbm_copy": o[3],
                "fp16_mfma_32x32": o[4], "bf16_mfma_32x32": o[5], "int8_mfma_32x32": o[6],
                "int8_mfma_16x16": o[7], "int8_mfma_32x32_random": o[8],
                "fp64_mfma_random": o[9]}

    def selftest_tile(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a, b = _f64(a), _f64(b)
        c = np.zeros((64, 64))
        self.call("mgw_hip_selftest_tile", _ptr(a), _ptr(b), _ptr(c))
        return c


class HipScreeningWorkspace:
    """DeviceScreeningWorkspace over the hip backend."""

    def __init__(self, v_ph: np.ndarray, ctx: HipContext | None = None):
        v = _f64(v_ph)
        if v.ndim != 2 or v.shape[0] != v.shape[1]:
            raise HipError("mgw_hip_upload_vph", MGW_ERR_ARG, "V_ph must be square")
        self.ctx = ctx or HipContext()
        self._n = v.shape[0]
        self.ctx.call("mgw_hip_upload_vph", _ptr(v), self._n)

    @property
    def nph(self) -> int:
        return self._n

    @property
    def estimated_device_bytes(self) -> int:
        return self.ctx.device_bytes

    def compute_w_c(self, pi0: np.ndarray):
        p = np.asarray(pi0)
        re = _f64(p.real)
        im = _f64(p.imag) if np.iscomplexobj(p) else None
        self.ctx.call("mgw_hip_compute_wc", _ptr(re), _ptr(im), re.shape[0])

    def w_c(self) -> np.ndarray:
        """W_c (n x n, complex128; imaginary part zero at eta = 0)."""
        re = np.zeros((self._n, self._n))
        im = np.zeros((self._n, self._n))
        self.ctx.call("mgw_hip_download_wc", _ptr(re), _ptr(im), self._n)
        return re + 1j * im

    def quadratic_forms_panel(self, x_panel) -> np.ndarray:
        """q for a host panel (n_ph x nvec, row-major) or a HipDevicePanel."""
        if isinstance(x_panel, HipDevicePanel):
            q, qi = np.zeros(x_panel.cols), np.zeros(x_panel.cols)
            self.ctx.call("mgw_hip_quadratic_forms_device", x_panel.ptr, x_panel.rows, x_panel.ld,
                          x_panel.cols, _ptr(q), _ptr(qi))
            return q + 1j * qi
        x = _f64(x_panel)
        n, nvec = x.shape
        q, qi = np.zeros(nvec), np.zeros(nvec)
        self.ctx.call("mgw_hip_quadratic_forms_host", _ptr(x), n, nvec, _ptr(q), _ptr(qi))
        return q + 1j * qi


@dataclass
class HipDevicePanel:
    """DeviceComplexPanelView: a device panel (real FP64, column-major, column
    c at ptr + c * ld) owned by the HipPqPhPanelView that filled it, valid until
    its next fill_panel."""
    ptr: int
    rows: int
    ld: int
    cols: int


class HipPqPhPanelView:
    """DevicePqPhPanelView sharing a workspace context; `mode` as the
    reference's DevicePqPhPanelStorageMode (ERI_AUTO, ERI_RESIDENT,
    ERI_STREAMED)."""

    def __init__(self, eri: np.ndarray, nocc: int, ctx: HipContext, mode: int = ERI_AUTO):
        self.ctx = ctx
        self.nmo = eri.shape[0]
        self.nocc = nocc
        ctx.set_eri_storage(mode)
        ctx.upload_eri(eri, nocc)

    @property
    def nph(self) -> int:
        return self.nocc * (self.nmo - self.nocc)

    def storage_mode(self) -> int:
        return self.ctx.eri_storage()[0]

    def storage_mode_name(self) -> str:
        return ERI_STORAGE_NAMES[self.storage_mode()]

    def estimated_device_bytes(self) -> int:
        return self.ctx.eri_storage()[1]

    def build_v_ph(self):
        self.ctx.call("mgw_hip_build_vph_from_eri")

    def quadratic_forms(self, p: int, k0: int, width: int) -> np.ndarray:
        q, qi = np.zeros(max(width, 0)), np.zeros(max(width, 0))
        self.ctx.call("mgw_hip_quadratic_forms", p, k0, width, _ptr(q), _ptr(qi))
        return q + 1j * qi

    def fill_panel(self, p: int, k0: int, width: int) -> HipDevicePanel:
        """X(ia, kk) = (p, k0 + kk | i a) in device memory of this view's context."""
        ptr, ld = C.c_void_p(), C.c_int32()
        self.ctx.call("mgw_hip_fill_panel", p, k0, width, C.byref(ptr), C.byref(ld))
        return HipDevicePanel(ptr.value or 0, self.nph, ld.value, width)


@dataclass
class HipG0W0Result:
    omegas: np.ndarray
    weights: np.ndarray
    states: list
    sigma_x_diag: np.ndarray
    sigma_c: np.ndarray        # (nstates, nfreq) complex
    qp: np.ndarray             # (nmo,), NaN for unselected states
    qp_iterations: np.ndarray
    timings: dict
    q: np.ndarray | None = None   # (nfreq, nstates * nmo) quadratic forms, if asked for


def run_g0w0(inp, nfreq: int = 200, npar: int = 16, states=None, eta: float = 0.0,
             precision: int = PREC_FP64, ctx: HipContext | None = None,
             upload: bool = True, want_q: bool = False) -> HipG0W0Result:
    """G0W0 for one molecule on the device (ERI uploaded unless upload=False);
    want_q also returns the quadratic forms q[f, s * nmo + k]."""
    ctx = ctx or HipContext()
    nmo = inp.nmo
    if upload:
        ctx.upload_eri(inp.eri, inp.nocc)
    st = None if states is None else np.ascontiguousarray(states, dtype=np.int32)
    ns = nmo if st is None else len(st)
    mo = _f64(inp.mo_energy)
    vxc = _f64(np.diag(inp.vxc))
    job = G0W0Job(nmo, inp.nocc, nfreq, npar, 0 if st is None else ns, _ptr(st, _ip),
                  _ptr(mo), _ptr(vxc), float(inp.fermi_energy), float(eta), precision)
    om, w, sx = np.zeros(nfreq), np.zeros(nfreq), np.zeros(nmo)
    sre, sim = np.zeros((ns, nfreq)), np.zeros((ns, nfreq))
    qp, it = np.zeros(nmo), np.zeros(nmo, dtype=np.int32)
    q = np.zeros((nfreq, ns * nmo)) if want_q else None
    out = G0W0Out(_ptr(om), _ptr(w), _ptr(sx), _ptr(sre), _ptr(sim), _ptr(qp), _ptr(it, _ip),
                  _ptr(q))
    ctx.call("mgw_hip_g0w0", C.byref(job), C.byref(out))
    return HipG0W0Result(om, w, list(range(nmo)) if st is None else st.tolist(), sx,
                         sre + 1j * sim, qp, it, ctx.last_timings(), q)


def _job_and_out(inp, nfreq, npar, states, eta, precision, want_q):
    """The C job and output records of one molecule, and the arrays they point to."""
    nmo = inp.nmo
    st = None if states is None else np.ascontiguousarray(states, dtype=np.int32)
    ns = nmo if st is None else len(st)
    mo = _f64(inp.mo_energy)
    vxc = _f64(np.diag(inp.vxc))
    job = G0W0Job(nmo, inp.nocc, nfreq, npar, 0 if st is None else ns, _ptr(st, _ip),
                  _ptr(mo), _ptr(vxc), float(inp.fermi_energy), float(eta), precision)
    arr = dict(st=st, mo=mo, vxc=vxc, om=np.zeros(nfreq), w=np.zeros(nfreq), sx=np.zeros(nmo),
               sre=np.zeros((ns, nfreq)), sim=np.zeros((ns, nfreq)), qp=np.zeros(nmo),
               it=np.zeros(nmo, dtype=np.int32), q=np.zeros((nfreq, ns * nmo)) if want_q else None)
    out = G0W0Out(_ptr(arr["om"]), _ptr(arr["w"]), _ptr(arr["sx"]), _ptr(arr["sre"]),
                  _ptr(arr["sim"]), _ptr(arr["qp"]), _ptr(arr["it"], _ip), _ptr(arr["q"]))
    return job, out, arr


def _result(arr, nmo, timings) -> "HipG0W0Result":
    st = arr["st"]
    return HipG0W0Result(arr["om"], arr["w"], list(range(nmo)) if st is None else st.tolist(),
                         arr["sx"], arr["sre"] + 1j * arr["sim"], arr["qp"], arr["it"], timings,
                         arr["q"])


def run_g0w0_group(inps, ctxs, nfreq=200, npar=16, precision: int = PREC_FP64, eta=0.0,
                   states=None, want_q: bool = False):
    """Several molecules through mgw_hip_g0w0_group: inps[i] on the ERI already
    resident in ctxs[i] (ctxs[0] leads: its stream and batch buffers run the
    shared frequency batches).  nfreq, npar, precision, eta and states may be
    one value for all or a list per molecule.  Returns (results, status):
    results[i] is None where status[i] != MGW_OK."""
    k = len(inps)
    if len(ctxs) != k or k == 0:
        raise ValueError("one context per molecule")

    def per(v):
        return list(v) if isinstance(v, (list, tuple)) else [v] * k
    nf, npr, prec, et = per(nfreq), per(npar), per(precision), per(eta)
    per_mol = (isinstance(states, list) and len(states) == k
               and all(v is None or isinstance(v, (list, tuple, np.ndarray)) for v in states))
    sts = list(states) if per_mol else [states] * k
    jobs = (G0W0Job * k)()
    outs = (G0W0Out * k)()
    keep = []
    for i, inp in enumerate(inps):
        jobs[i], outs[i], arr = _job_and_out(inp, nf[i], npr[i], sts[i], et[i], prec[i], want_q)
        keep.append(arr)
    handles = (C.c_void_p * k)(*[c.h for c in ctxs])
    status = np.zeros(k, dtype=np.int32)
    rc = lib.mgw_hip_g0w0_group(handles, jobs, outs, k, _ptr(status, _ip))
    if rc != MGW_OK and not np.any(status == MGW_OK):
        _check("mgw_hip_g0w0_group", rc, ctxs[0].h)
    res = [(_result(keep[i], inps[i].nmo, ctxs[i].last_timings()) if status[i] == MGW_OK else None)
           for i in range(k)]
    return res, status


def pade_qp(sigma_c, omegas, sigma_x_diag, mo_energy, vxc_diag, fermi: float, npar: int = 16,
            states=None, ctx: HipContext | None = None):
    """Device Pade fit + QP fixed point on a given Sigma_c(i w) (nstates x
    nfreq complex); mirrors solve_qp_energies (src/gw/qp_solver.cpp:10-47).
    Returns (qp[nmo], iterations[nmo])."""
    ctx = ctx or HipContext()
    sig = np.asarray(sigma_c)
    mo = _f64(mo_energy)
    nmo = len(mo)
    st = None if states is None else np.ascontiguousarray(states, dtype=np.int32)
    ns = nmo if st is None else len(st)
    if sig.shape != (ns, len(omegas)):
        raise ValueError(f"sigma_c shape {sig.shape} != ({ns}, {len(omegas)})")
    sre, sim = _f64(sig.real), _f64(sig.imag)
    qp, it = np.zeros(nmo), np.zeros(nmo, dtype=np.int32)
    ctx.call("mgw_hip_pade_qp", nmo, 0 if st is None else ns, _ptr(st, _ip), len(omegas), npar,
             _ptr(_f64(omegas)), _ptr(sre), _ptr(sim), _ptr(_f64(sigma_x_diag)), _ptr(mo),
             _ptr(_f64(vxc_diag)), float(fermi), _ptr(qp), _ptr(it, _ip))
    return qp, it


class HipServer:
    """Continuous-batching request loop (worker threads, one stream each,
    HBM-budget admission) over mgw_hip_server_*."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget: int = 200 << 30,
                 batch_freqs: int = 32):
        h = C.c_void_p()
        _check("mgw_hip_server_create",
               lib.mgw_hip_server_create(device, workers, hbm_budget, batch_freqs, C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            lib.mgw_hip_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, inp=None, *, factor=None, mo_energy=None, vxc=None, nocc=None, fermi=None,
               nfreq: int = 200, npar: int = 16, states=None, precision: int = PREC_FP64) -> int:
        """Queue one molecule: either `inp` (GwInput with .eri) or its `factor`
        B (nmo^2 x K) with mo_energy, vxc, nocc, fermi."""
        if inp is not None:
            mo, vx, nocc, fermi = inp.mo_energy, inp.vxc, inp.nocc, inp.fermi_energy
            eri, fac, kdim = _f64(inp.eri), None, 0
        else:
            mo, vx = mo_energy, vxc
            eri, fac = None, _f64(factor)
            kdim = fac.shape[1]
        mo = _f64(mo)
        nmo = mo.shape[0]
        vxd = _f64(np.diag(vx))
        st = None if states is None else np.ascontiguousarray(states, dtype=np.int32)
        ns = nmo if st is None else len(st)
        res = {"omega": np.zeros(nfreq), "weight": np.zeros(nfreq), "sx": np.zeros(nmo),
               "sre": np.zeros((ns, nfreq)), "sim": np.zeros((ns, nfreq)), "qp": np.zeros(nmo),
               "it": np.zeros(nmo, dtype=np.int32)}
        job = G0W0Job(nmo, int(nocc), nfreq, npar, 0 if st is None else ns, _ptr(st, _ip),
                      _ptr(mo), _ptr(vxd), float(fermi), 0.0, precision)
        sj = ServeJob(job, _ptr(eri), _ptr(fac), kdim)
        out = G0W0Out(_ptr(res["omega"]), _ptr(res["weight"]), _ptr(res["sx"]), _ptr(res["sre"]),
                      _ptr(res["sim"]), _ptr(res["qp"]), _ptr(res["it"], _ip))
        t = C.c_int64(0)
        _check("mgw_hip_server_submit", lib.mgw_hip_server_submit(self.h, C.byref(sj), C.byref(out), C.byref(t)))
        self._keep[t.value] = (res, eri, fac, st, nmo)
        return t.value

    def set_judge(self, nsample: int, tol: float = 0.0, fallback: bool = True):
        """Error judge of mixed jobs submitted from now on (0 samples: off)."""
        _check("mgw_hip_server_set_judge",
               lib.mgw_hip_server_set_judge(self.h, int(nsample), float(tol), 1 if fallback else 0))

    def set_grouping(self, max_jobs: int):
        """Most jobs a worker runs together through mgw_hip_g0w0_group (1: off)."""
        _check("mgw_hip_server_set_grouping", lib.mgw_hip_server_set_grouping(self.h, int(max_jobs)))

    def set_eri_storage(self, storage: int):
        """ERI storage of jobs submitted from now on: ERI_RESIDENT (default),
        ERI_STREAMED, or ERI_AUTO (resident when it fits the budget left, else
        streamed)."""
        _check("mgw_hip_server_set_eri_storage", lib.mgw_hip_server_set_eri_storage(self.h, int(storage)))

    def set_budget_scale(self, scale: float):
        """Fault injection: per-job device budget = scale x estimate."""
        _check("mgw_hip_server_set_budget_scale", lib.mgw_hip_server_set_budget_scale(self.h, scale))

    def cancel(self, ticket: int):
        _check("mgw_hip_server_cancel", lib.mgw_hip_server_cancel(self.h, ticket))

    def wait(self, ticket: int, timeout: float = -1.0):
        """Returns (HipG0W0Result, latency_ms, run_ms); raises HipError on failure.
        The result's timings hold the judge outcome ("judge_err", "fell_back") and
        the storage the job's ERI took ("eri_storage").
        The job's record is released once it has been reported."""
        st, lat, run = C.c_int32(0), C.c_double(0), C.c_double(0)
        _check("mgw_hip_server_wait", lib.mgw_hip_server_wait(self.h, ticket, timeout, C.byref(st),
                                                               C.byref(lat), C.byref(run)))
        if st.value == -1:
            raise TimeoutError(f"job {ticket} not finished")
        res, _, _, stt, nmo = self._keep.pop(ticket)
        je, fb, es = C.c_double(-1.0), C.c_int32(0), C.c_int32(0)
        lib.mgw_hip_server_job_judge(self.h, ticket, C.byref(je), C.byref(fb))
        lib.mgw_hip_server_job_eri_storage(self.h, ticket, C.byref(es))
        msg = lib.mgw_hip_server_job_error(self.h, ticket)
        lib.mgw_hip_server_release(self.h, ticket)
        if st.value != MGW_OK:
            raise HipError("mgw_hip_server_wait", st.value, msg.decode() if msg else "")
        out = HipG0W0Result(res["omega"], res["weight"],
                            list(range(nmo)) if stt is None else stt.tolist(), res["sx"],
                            res["sre"] + 1j * res["sim"], res["qp"], res["it"],
                            {"judge_err": je.value, "fell_back": fb.value,
                             "eri_storage": es.value})
        return out, lat.value, run.value

    def stats(self) -> dict:
        c, r, p, b = C.c_int64(0), C.c_int32(0), C.c_int32(0), C.c_size_t(0)
        _check("mgw_hip_server_stats", lib.mgw_hip_server_stats(self.h, C.byref(c), C.byref(r),
                                                                 C.byref(p), C.byref(b)))
        return {"completed": c.value, "running": r.value, "peak_running": p.value,
                "reserved_bytes": b.value}


def autotune(inp, nfreq: int = 200, npar: int = 16, precision: int = PREC_FP64,
             ctx: HipContext | None = None):
    """Pick the frequency batch size (and, for the mixed forms, the contraction's
    unit shape) for this molecule's shape on the resident ERI; returns (batch,
    ms per frequency).  Later run_g0w0 calls use them."""
    ctx = ctx or HipContext()
    mo = _f64(inp.mo_energy)
    vxc = _f64(np.diag(inp.vxc))
    job = G0W0Job(inp.nmo, inp.nocc, nfreq, npar, 0, None, _ptr(mo), _ptr(vxc),
                  float(inp.fermi_energy), 0.0, precision)
    b, t = C.c_int32(0), C.c_double(0)
    ctx.call("mgw_hip_autotune", C.byref(job), C.byref(b), C.byref(t))
    return b.value, t.value


def autotune_result(inp, nfreq: int = 200, npar: int = 16, precision: int = PREC_FP64):
    """(batch, unit frequencies) the autotuner keeps for this shape (0 = none)."""
    mo = _f64(inp.mo_energy)
    vxc = _f64(np.diag(inp.vxc))
    job = G0W0Job(inp.nmo, inp.nocc, nfreq, npar, 0, None, _ptr(mo), _ptr(vxc),
                  float(inp.fermi_energy), 0.0, precision)
    b, u = C.c_int32(0), C.c_int32(0)
    _check("mgw_hip_autotune_result", lib.mgw_hip_autotune_result(C.byref(job), C.byref(b), C.byref(u)))
    return b.value, u.value
