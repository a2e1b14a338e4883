"""vp9r_mc_batch_planes_noalloc, the motion compensation entry point that
allocates nothing (its descriptors and planes staged in the library's fixed
device array): its output equals vp9r_mc_batch_planes' on the same inputs,
on two calls in a row, the second with fewer and smaller blocks (so stale
contents of the shared array from the first call would show), at 8 and 10
bits; sizes beyond its limit are refused before any device work."""
import ctypes

import numpy as np
import pytest

import rivet_vp9_amd as vp

pytestmark = pytest.mark.gpu


def batch(rng, n, geo, sizes):
    descs = []
    out_stride = 64 * 8
    for i in range(n):
        w, h = sizes[i % len(sizes)]
        a_k, b_k = int(rng.integers(0, len(geo))), int(rng.integers(0, len(geo)))
        pw, ph, _ = geo[a_k]
        d = vp.McDesc()
        d.x, d.y = int(rng.integers(-20 * 16, (pw + 10) * 16)), int(rng.integers(-20 * 16, (ph + 10) * 16))
        d.x2, d.y2 = d.x + int(rng.integers(-64, 64)), d.y + int(rng.integers(-64, 64))
        d.xs, d.ys = (16, 16) if i % 3 else (int(rng.integers(8, 33)), int(rng.integers(8, 33)))
        d.w, d.h, d.filter = w, h, int(rng.integers(0, 4))
        d.ref_idx, d.ref2_idx = a_k, (b_k if i % 4 == 1 else -1)
        d.out_off = (i // 8) * 64 * out_stride + (i % 8) * 64
        descs.append(d)
    return (vp.McDesc * n)(*descs), out_stride, (n // 8 + 1) * 64


@pytest.mark.parametrize("bd", [8, 10])
def test_noalloc_equals_planes_on_two_calls(bd):
    g = vp.Gpu(0)
    lib = vp.library()
    rng = np.random.default_rng(1100 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    max_v = (1 << bd) - 1
    geo = [(150, 90, 160), (76, 45, 96), (300, 180, 320)]
    planes, keep = (vp.RefPlane * 3)(), []
    for k, (pw, ph, stride) in enumerate(geo):
        d = g.upload(rng.integers(0, max_v + 1, size=(ph + 4) * stride).astype(dt))
        keep.append(d)
        planes[k].data, planes[k].stride, planes[k].last_x, planes[k].last_y = d.ptr.value, stride, pw - 1, ph - 1
    calls = [(96, [(64, 64), (32, 32), (16, 8), (8, 16), (64, 32), (4, 4)]), (20, [(8, 8), (4, 4), (16, 16)])]
    for n, sizes in calls:
        descs, out_stride, rows = batch(rng, n, geo, sizes)
        outs = []
        for fn in (lib.vp9r_mc_batch_planes, lib.vp9r_mc_batch_planes_noalloc):
            d_out = g.upload(np.zeros(rows * out_stride, dtype=dt))
            st = fn(g.handle, n, ctypes.cast(descs, ctypes.c_void_p), ctypes.cast(planes, ctypes.c_void_p), 3,
                    d_out.ptr, out_stride, bd)
            assert st == 0, vp.library().vp9r_last_error()
            outs.append(d_out.read(dt, rows * out_stride))
        assert np.array_equal(outs[0], outs[1]), f"{n} blocks at {bd} bits"


def test_noalloc_refuses_oversized_batches():
    g = vp.Gpu(0)
    lib = vp.library()
    planes = (vp.RefPlane * 1)()
    d = g.upload(np.zeros(64 * 64, dtype=np.uint8))
    planes[0].data, planes[0].stride, planes[0].last_x, planes[0].last_y = d.ptr.value, 64, 63, 63
    descs = (vp.McDesc * 65537)()
    st = lib.vp9r_mc_batch_planes_noalloc(g.handle, 65537, ctypes.cast(descs, ctypes.c_void_p),
                                          ctypes.cast(planes, ctypes.c_void_p), 1, d.ptr, 64, 8)
    assert st == vp.ERR_INVALID_INPUT
    assert b"65536" in vp.library().vp9r_last_error()
