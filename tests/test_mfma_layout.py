"""Lane layout of the MFMA instructions the kernels use, checked on the device
with one-hot inputs (each probe is a single in-bounds wave).

The kernels assume (wave64):
  f64 16x16x4:  a = A[lane % 16][lane / 16], b = B[lane / 16][lane % 16],
                d[r] = D[lane / 16 + 4 * r][lane % 16]
  16x16x32 (bf16/f16): a[e] = A[lane % 16][k(lane / 16, e)],
                b[e] = B[k(lane / 16, e)][lane % 16] with the same k map,
                d[r] = D[4 * (lane / 16) + r][lane % 16]
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from minigw_amd import hip
    c = hip.HipContext(0)
    yield c
    c.close()


def _dump(name, **arrs):
    out = os.environ.get("MGW_PROBE_DIR")      # optional dump of the probes for inspection
    if out and os.path.isdir(out):
        np.savez_compressed(os.path.join(out, name), **arrs)


def test_f64_layout(ctx):
    a = np.zeros((64 * 64, 64))
    b = np.zeros((64 * 64, 64))
    for la in range(64):
        for lb in range(64):
            a[la * 64 + lb, la] = 1.0
            b[la * 64 + lb, lb] = 1.0
    d = ctx.probe_mfma(0, a, b)
    _dump("probe_f64.npz", d=d)
    found = {}
    for la in range(64):
        for lb in range(64):
            nz = np.argwhere(d[la * 64 + lb] != 0)
            if len(nz):
                found[(la, lb)] = [tuple(x) for x in nz]
    print("f64 pairs producing output:", len(found), "e.g.", list(found.items())[:6], flush=True)
    for (la, lb), pos in found.items():
        assert la // 16 == lb // 16                       # same k slot
        i, j = la % 16, lb % 16
        assert pos == [(j + 16 * (i % 4), i // 4)], ((la, lb), pos)
    assert len(found) == 64 * 16


def _bits(x, kind):
    x = np.asarray(x, dtype=np.float32)
    if kind == 2:
        return x.astype(np.float16).view(np.uint16)
    return (x.view(np.uint32) >> 16).astype(np.uint16)   # exact for small integers


@pytest.mark.parametrize("kind", [1, 2])
def test_16bit_layout(ctx, kind):
    # rows: A one-hot at (lane, e), B all ones -> row i filled in every column
    n = 64 * 8
    a = np.zeros((n, 64, 8), dtype=np.float32)
    for la in range(64):
        for e in range(8):
            a[la * 8 + e, la, e] = 1.0
    # B value at (lane, e) encodes its k slot: 1 + e + 8 * (lane // 16)
    bv = np.zeros((64, 8), dtype=np.float32)
    for lb in range(64):
        for e in range(8):
            bv[lb, e] = 1 + e + 8 * (lb // 16)
    b = np.broadcast_to(bv, (n, 64, 8))
    d = ctx.probe_mfma(kind, _bits(a, kind), _bits(b, kind))
    _dump(f"probe_16_{kind}.npz", d=d)
    for la in range(64):
        for e in range(8):
            out = d[la * 8 + e]
            nz = np.argwhere(out != 0)
            i = la % 16
            want = sorted((j + 16 * (i // 4), i % 4) for j in range(16))
            assert sorted(tuple(x) for x in nz) == want, (kind, la, e, nz[:5])
            # the paired B slot carries the same k: value 1 + e + 8 * (la // 16)
            vals = set(float(out[l, r]) for l, r in nz)
            assert vals == {1.0 + e + 8 * (la // 16)}, (kind, la, e, vals)


@pytest.mark.parametrize("kind", [3, 4])
def test_32x32_layout(ctx, kind):
    """32x32x16: a[e] = A[lane % 32][k(lane / 32, e)], b likewise with columns,
    d[r] = D[8 (r / 4) + 4 (lane / 32) + r % 4][lane % 32]."""
    bits_kind = 1 if kind == 3 else 2
    n = 64 * 8
    a = np.zeros((n, 64, 8), dtype=np.float32)
    for la in range(64):
        for e in range(8):
            a[la * 8 + e, la, e] = 1.0
    bv = np.zeros((64, 8), dtype=np.float32)
    for lb in range(64):
        for e in range(8):
            bv[lb, e] = 1 + e + 8 * (lb // 32)
    b = np.broadcast_to(bv, (n, 64, 8))
    d = ctx.probe_mfma(kind, _bits(a, bits_kind), _bits(b, bits_kind))
    _dump(f"probe_32_{kind}.npz", d=d)
    for la in range(64):
        for e in range(8):
            out = d[la * 8 + e]
            nz = sorted(tuple(x) for x in np.argwhere(out != 0))
            i = la % 32
            want = sorted((j + 32 * ((i % 8) // 4), 4 * (i // 8) + i % 4) for j in range(32))
            assert nz == want, (kind, la, e, nz[:4])
            vals = set(float(out[l, r]) for l, r in nz)
            assert vals == {1.0 + e + 8 * (la // 32)}, (kind, la, e, vals)


def test_i8_32x32x32_layout(ctx):
    """v_mfma_i32_32x32x32_i8: a byte e of lane l = A[l % 32][k(l / 32, e)],
    b likewise with columns and the same k map, d[r] as the 32x32x16 forms."""
    n = 64 * 16
    a = np.zeros((n, 64, 16), dtype=np.int8)
    for la in range(64):
        for e in range(16):
            a[la * 16 + e, la, e] = 1
    bv = np.zeros((64, 16), dtype=np.int8)
    for lb in range(64):
        for e in range(16):
            bv[lb, e] = 1 + e + 16 * (lb // 32)
    b = np.broadcast_to(bv, (n, 64, 16))
    d = ctx.probe_mfma(5, a, b)
    _dump("probe_i8.npz", d=d)
    for la in range(64):
        for e in range(16):
            out = d[la * 16 + e]
            nz = sorted(tuple(x) for x in np.argwhere(out != 0))
            i = la % 32
            want = sorted((j + 32 * ((i % 8) // 4), 4 * (i // 8) + i % 4) for j in range(32))
            assert nz == want, (la, e, nz[:4])
            vals = set(int(out[l, r]) for l, r in nz)
            assert vals == {1 + e + 16 * (la // 32)}, (la, e, vals)
