"""mgw_hip_quadratic_forms_noalloc (needs an MI355X): the allocation-free
variant of mgw_hip_quadratic_forms, through the C ABI, against the existing
entry point on two calls in a row (the second with a smaller panel), with the
ERI resident and streamed, its limits, and the device bytes of the context
unchanged by it."""
import ctypes as C

import numpy as np
import pytest

from minigw_amd.synthetic import synthetic_input
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hip():
    from minigw_amd import hip as h
    h.lib.mgw_hip_quadratic_forms_noalloc.restype = C.c_int
    h.lib.mgw_hip_quadratic_forms_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                                       C.c_void_p]
    return h


@pytest.fixture(scope="module")
def hiprt():
    """The HIP runtime the library uses, for the caller's own device buffers."""
    rt = C.CDLL("libamdhip64.so")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


def _noalloc(hip, ctx, p, k0, width, rt):
    """The entry point on a device buffer of the caller's (filled with NaN first)."""
    n = max(width, 1)
    host = np.full(n, np.nan)
    dev = C.c_void_p()
    assert rt.hipMalloc(C.byref(dev), n * 8) == 0
    try:
        assert rt.hipMemcpy(dev, host.ctypes.data, n * 8, 1) == 0          # host to device
        rc = hip.lib.mgw_hip_quadratic_forms_noalloc(ctx.h, p, k0, width, dev)
        if rc != hip.MGW_OK:
            raise hip.HipError("mgw_hip_quadratic_forms_noalloc", rc,
                               hip.lib.mgw_hip_last_error(ctx.h).decode())
        assert rt.hipMemcpy(host.ctypes.data, dev, n * 8, 2) == 0          # device to host
    finally:
        rt.hipFree(dev)
    return host[:width]


@pytest.mark.parametrize("storage", ["ERI_RESIDENT", "ERI_STREAMED"])
def test_noalloc_quadratic_forms_match_the_allocating_entry_point(hip, hiprt, storage):
    for nmo, nocc, seed in [(24, 5, 20240601), (40, 9, 3)]:
        say(f"case: no-alloc quadratic forms, nmo={nmo} nocc={nocc}, {storage}")
        inp = synthetic_input(nmo, nocc, seed)
        ph = orc.particle_hole_basis(inp.mo_energy, inp.nocc)
        with hip.HipContext(0) as ctx:
            view = hip.HipPqPhPanelView(inp.eri, inp.nocc, ctx, getattr(hip, storage))
            view.build_v_ph()
            ws = hip.HipScreeningWorkspace.__new__(hip.HipScreeningWorkspace)
            ws.ctx, ws._n = ctx, ph.size
            with pytest.raises(hip.HipError) as e:
                _noalloc(hip, ctx, 0, 0, nmo, hiprt)
            assert e.value.code == hip.MGW_ERR_STATE
            ws.compute_w_c(orc.pi0_diag_real(0.8, ph.delta))
            # two calls in a row, the second with a smaller panel, against the
            # allocating entry point (run first: it grows the context's buffers,
            # the allocation-free one must not)
            panels = [(3, 0, nmo), (nmo - 1, 4, 7)]
            want = [view.quadratic_forms(p, k0, width) for p, k0, width in panels]
            held = ctx.device_bytes
            for (p, k0, width), w in zip(panels, want):
                got = _noalloc(hip, ctx, p, k0, width, hiprt)
                say(f"   p={p} k0={k0} width={width}: max |dq| {np.abs(got - w.real).max():.1e}")
                assert np.array_equal(got, w.real)
            assert ctx.device_bytes == held
            assert _noalloc(hip, ctx, 0, 2, 0, hiprt).size == 0
            with pytest.raises(hip.HipError) as e:
                _noalloc(hip, ctx, nmo, 0, 1, hiprt)
            assert e.value.code == hip.MGW_ERR_ARG
            ws.compute_w_c(orc.pi0_diag(0.8, ph.delta, 0.05))
            with pytest.raises(hip.HipError) as e:
                _noalloc(hip, ctx, 0, 0, nmo, hiprt)
            assert e.value.code == hip.MGW_ERR_UNSUPPORTED


def test_noalloc_rejects_panels_beyond_its_scratch(hip, hiprt):
    say("case: no-alloc limits, nmo=260 (width 260 > 256)")
    nmo, nocc = 260, 2
    from minigw_amd.synthetic import synthetic_factor
    mo, fermi, b, vxc = synthetic_factor(nmo, nocc, 9)
    with hip.HipContext(0) as ctx:
        ctx.upload_eri_factor(b, nmo, nocc)
        ctx.call("mgw_hip_build_vph_from_eri")
        ph = orc.particle_hole_basis(mo, nocc)
        ws = hip.HipScreeningWorkspace.__new__(hip.HipScreeningWorkspace)
        ws.ctx, ws._n = ctx, ph.size
        ws.compute_w_c(orc.pi0_diag_real(0.8, ph.delta))
        with pytest.raises(hip.HipError) as e:
            _noalloc(hip, ctx, 0, 0, nmo, hiprt)
        assert e.value.code == hip.MGW_ERR_ARG and "scratch limit" in str(e.value)
        view = hip.HipPqPhPanelView.__new__(hip.HipPqPhPanelView)
        view.ctx = ctx
        want = view.quadratic_forms(1, 4, 256)
        held = ctx.device_bytes
        assert np.array_equal(_noalloc(hip, ctx, 1, 4, 256, hiprt), want.real)
        assert ctx.device_bytes == held
