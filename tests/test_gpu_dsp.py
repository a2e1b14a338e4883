"""GPU parity of the per-kernel seam (the src/dsp signatures) against the CPU
oracle on the same seeded inputs, through the C ABI: inverse transforms
(itx.rs:701-749) at every size, type and bit depth, conforming and not;
intra prediction (intra.rs:19-175) for every mode; motion compensation
(inter.rs:289-395, pixel.rs:22-50) inside, across and far outside the
reference's edges, every filter, unscaled and scaled, single and compound;
the loop filter's edge passes (lf.rs:178-259).  Bit-exact."""
import ctypes

import numpy as np
import pytest

import oracle_lib
import rivet_vp9_amd as vp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    return vp.Gpu(0)


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rand_block(rng, n, mag):
    n0 = 1 << n
    b = np.zeros(n0 * n0, dtype=np.int64)
    kind = rng.integers(0, 4)
    if kind == 0:
        b[0] = rng.integers(-mag, mag + 1)
    elif kind == 1:
        for _ in range(rng.integers(1, 6)):
            b[rng.integers(0, n0 * n0)] = rng.integers(-mag, mag + 1)
    elif kind == 2:
        rows = rng.integers(1, n0 + 1)
        b[:rows * n0] = rng.integers(-mag, mag + 1, size=rows * n0)
    else:
        b[:] = rng.integers(-mag, mag + 1, size=n0 * n0)
    return b.astype(np.int32)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_itx_host_matches_oracle(gpu, bd):
    rng = np.random.default_rng(100 + bd)
    lib = vp.library()
    olib = oracle_lib.lib()
    max_v = (1 << bd) - 1
    for n in range(2, 6):
        n0 = 1 << n
        for tx_type in range(4):
            for it in range(24):
                lossless = 1 if (n == 2 and tx_type == 0 and it % 4 == 3) else 0
                mag = [64, 1 << 12, (1 << (7 + bd)) - 1, 2 ** 31 - 1][it % 4]
                b = rand_block(rng, n, mag)
                eob = 1 if not b[1:].any() else 2
                stride = n0 + int(rng.integers(0, 5))
                pred = rng.integers(0, max_v + 1, size=(n0, stride)).astype(np.uint16)
                want = pred.copy()
                olib.orc_itx_add(ptr(b), n, tx_type, lossless, eob, bd, ptr(want), stride)
                got = pred.copy()
                st = lib.vp9r_itx_add_host(gpu.handle, ptr(b), n, tx_type, lossless, eob, bd, ptr(got), stride)
                assert st == 0
                assert np.array_equal(got[:, :n0], want[:, :n0]), f"n={n} type={tx_type} bd={bd} it={it}"


def test_itx_batch_matches_oracle(gpu):
    rng = np.random.default_rng(7)
    lib = vp.library()
    olib = oracle_lib.lib()
    for bd in (8, 10):
        count = 600
        stride = 64 * 24
        tx = rng.integers(0, 4, size=count).astype(np.uint8)
        tx_type = np.where(tx == 3, 0, rng.integers(0, 4, size=count)).astype(np.uint8)
        coefs, offs, eobs, dst_off = [], [], [], []
        off = 0
        for i in range(count):
            n = 2 + int(tx[i])
            b = rand_block(rng, n, [64, 2000, 30000][i % 3] >> (n - 2))
            coefs.append(b)
            offs.append(off)
            off += b.size
            eobs.append(1 if not b[1:].any() else 2)
            # block i at column (i % 24) * 64, row (i // 24) * 32
            dst_off.append((i // 24) * 32 * stride + (i % 24) * 64)
        rows = ((count + 23) // 24) * 32
        max_v = (1 << bd) - 1
        plane = rng.integers(0, max_v + 1, size=(rows, stride)).astype(np.uint16)
        want = plane.copy()
        for i in range(count):
            n = 2 + int(tx[i])
            r0, c0 = divmod(dst_off[i], stride)
            sub = np.ascontiguousarray(want[r0:r0 + (1 << n), c0:c0 + (1 << n)])
            olib.orc_itx_add(ptr(coefs[i]), n, int(tx_type[i]), 0, eobs[i], bd, ptr(sub), 1 << n)
            want[r0:r0 + (1 << n), c0:c0 + (1 << n)] = sub
        dt = np.uint8 if bd == 8 else np.uint16
        d_plane = gpu.upload(plane.astype(dt))
        d_coefs = gpu.upload(np.concatenate(coefs))
        d_offs = gpu.upload(np.array(offs, dtype=np.uint32))
        d_types = gpu.upload(tx_type)
        d_eobs = gpu.upload(np.array(eobs, dtype=np.uint16))
        d_dst = gpu.upload(np.array(dst_off, dtype=np.uint64))
        st = lib.vp9r_itx_add_batch(gpu.handle, count, d_coefs.ptr, d_offs.ptr, ptr(tx), d_types.ptr, d_eobs.ptr,
                                    d_dst.ptr, d_plane.ptr, stride, bd, 0)
        assert st == 0
        got = d_plane.read(dt, rows * stride).reshape(rows, stride).astype(np.uint16)
        assert np.array_equal(got, want), f"bd={bd}"


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_intra_host_matches_oracle(gpu, bd):
    rng = np.random.default_rng(200 + bd)
    lib = vp.library()
    olib = oracle_lib.lib()
    max_v = (1 << bd) - 1
    for log2 in range(2, 6):
        size = 1 << log2
        for mode in range(10):
            for hl in (0, 1):
                for ha in (0, 1):
                    above = rng.integers(0, max_v + 1, size=2 * size + 1).astype(np.int32)
                    left = rng.integers(0, max_v + 1, size=size).astype(np.int32)
                    want = np.zeros((size, size + 3), dtype=np.uint16)
                    got = np.zeros((size, size + 3), dtype=np.uint16)
                    olib.orc_intra_predict(mode, log2, ptr(above), ptr(left), hl, ha, bd, ptr(want), size + 3)
                    st = lib.vp9r_intra_predict_host(gpu.handle, mode, log2, ptr(above), ptr(left), hl, ha, bd,
                                                     ptr(got), size + 3)
                    assert st == 0
                    assert np.array_equal(got[:, :size], want[:, :size]), f"mode={mode} size={size} bd={bd}"


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_mc_batch_matches_oracle(gpu, bd):
    rng = np.random.default_rng(300 + bd)
    lib = vp.library()
    olib = oracle_lib.lib()
    max_v = (1 << bd) - 1
    pw, ph = 150, 90
    stride = pw + 10
    nref = 2
    refs = []
    for _ in range(nref):
        r = rng.integers(0, max_v + 1, size=(ph, stride)).astype(np.uint16)
        r.ravel()[::97] = max_v
        refs.append(r)
    dt = np.uint8 if bd == 8 else np.uint16
    d_refs = [gpu.upload(r.astype(dt)) for r in refs]
    ref_ptrs = (ctypes.c_void_p * nref)(*[d.ptr.value for d in d_refs])
    sizes = [(4, 4), (4, 8), (8, 4), (8, 8), (16, 8), (8, 16), (16, 16), (32, 16), (16, 32), (32, 32),
             (64, 32), (32, 64), (64, 64)]
    descs = []
    out_stride = 64 * 8
    expected = []
    for i in range(96):
        w, h = sizes[i % len(sizes)]
        f = int(rng.integers(0, 4))
        kind = i % 3
        if kind == 0:
            x = int(rng.integers(0, (pw - w - 8) * 16))
        elif kind == 1:
            x = int(rng.integers(-80 * 16, (pw + 10) * 16))
        else:
            x = int(rng.integers(-3 * 16, 4 * 16))
        y = int(rng.integers(-40 * 16, (ph + 10) * 16)) if i % 2 else int(rng.integers(3 * 16, (ph - h - 5) * 16))
        xs, ys = (16, 16) if i % 4 != 3 else (int(rng.integers(1, 33)), int(rng.integers(1, 33)))
        comp = i % 5 == 4
        x2, y2 = x + int(rng.integers(-40, 40)), y + int(rng.integers(-40, 40))
        d = vp.McDesc()
        d.x, d.y, d.x2, d.y2, d.xs, d.ys, d.w, d.h, d.filter = x, y, x2, y2, xs, ys, w, h, f
        d.ref_idx, d.ref2_idx = 0, (1 if comp else -1)
        d.out_off = (i // 8) * 64 * out_stride + (i % 8) * 64
        descs.append(d)
        a = np.zeros((h, w), dtype=np.uint16)
        olib.orc_inter_predict(ptr(refs[0]), stride, pw - 1, ph - 1, x, y, xs, ys, w, h, f, bd, ptr(a), w)
        if comp:
            b = np.zeros((h, w), dtype=np.uint16)
            olib.orc_inter_predict(ptr(refs[1]), stride, pw - 1, ph - 1, x2, y2, xs, ys, w, h, f, bd, ptr(b), w)
            c = np.zeros((h, w), dtype=np.uint16)
            olib.orc_avg(ptr(a), ptr(b), w, h, ptr(c), w)
            a = c
        expected.append(a)
    rows = (len(descs) // 8 + 1) * 64
    d_out = gpu.upload(np.zeros(rows * out_stride, dtype=dt))
    arr = (vp.McDesc * len(descs))(*descs)
    st = lib.vp9r_mc_batch(gpu.handle, len(descs), ctypes.cast(arr, ctypes.c_void_p), ref_ptrs, nref, stride,
                           pw - 1, ph - 1, d_out.ptr, out_stride, bd)
    assert st == 0
    out = d_out.read(dt, rows * out_stride).reshape(rows, out_stride).astype(np.uint16)
    for i, d in enumerate(descs):
        r0, c0 = divmod(d.out_off, out_stride)
        got = out[r0:r0 + d.h, c0:c0 + d.w]
        assert np.array_equal(got, expected[i]), f"desc {i}: {d.w}x{d.h} at ({d.x},{d.y}) step ({d.xs},{d.ys}) filter {d.filter}"


def picture(rng, w, h, max_v, kind):
    base = int(rng.integers(0, max_v + 1))
    y, x = np.mgrid[0:h, 0:w]
    if kind == 0:
        v = base + ((x // 8 + y // 8) % 2) * rng.integers(0, 3, size=(h, w))
    elif kind == 1:
        v = base + rng.integers(-3, 4, size=(h, w)) * max(max_v // 255, 1)
    elif kind == 2:
        v = rng.integers(0, max_v + 1, size=(h, w))
    elif kind == 3:
        v = np.where((x // 4 + y // 4) % 2 == 0, 0, max_v)
    else:
        v = base + ((x + y) % 5) * rng.integers(0, 5, size=(h, w))
    return np.clip(v, 0, max_v).astype(np.uint16)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_lf_edges_matches_oracle(gpu, bd):
    rng = np.random.default_rng(400 + bd)
    lib = vp.library()
    olib = oracle_lib.lib()
    max_v = (1 << bd) - 1
    w = h = 96
    dt = np.uint8 if bd == 8 else np.uint16
    for it in range(40):
        pic = picture(rng, w, h, max_v, it % 5)
        n_passes = 4
        tables = np.zeros((n_passes, 1024), dtype=np.uint8)
        x0s, y0s, verts, nes, nrs, firsts = [], [], [], [], [], []
        want = pic.copy()
        for p in range(n_passes):
            e = oracle_lib.Edges()
            fs = rng.integers(-1, 3, size=(16, 16))
            for ed in range(16):
                fs[ed][(fs[ed] == 2)] = 2 if ed % 4 == 0 else 1
            l = rng.integers(1, 64, size=(16, 16))
            limit = np.clip(l >> rng.integers(0, 3, size=(16, 16)), 1, 9)
            blimit = 2 * (l + 2) + limit
            thresh = l >> 4
            for ed in range(16):
                for r in range(16):
                    e.fs[ed][r] = int(fs[ed][r])
                    e.limit[ed][r] = int(limit[ed][r])
                    e.blimit[ed][r] = int(blimit[ed][r])
                    e.thresh[ed][r] = int(thresh[ed][r])
            tables[p] = np.frombuffer(bytes(e), dtype=np.uint8)
            vertical = (p % 2) == 0
            n_runs = 8 if it % 3 == 0 else 16
            n_edges = 8 if it % 4 == 0 else 16
            x0, y0 = 16, 16
            # edges from `first` on (lf.rs:178-190), as the reference's
            # callers resume a pass
            first = 0 if it % 2 == 0 else int(rng.integers(0, n_edges))
            olib.orc_lf_filter_edges(ptr(want), w, x0, y0, int(vertical), first, n_edges, n_runs, ctypes.byref(e),
                                     bd)
            x0s.append(x0), y0s.append(y0), verts.append(int(vertical)), nes.append(n_edges), nrs.append(n_runs)
            firsts.append(first)
        d = gpu.upload(pic.astype(dt))
        st = lib.vp9r_lf_edges_batch(gpu.handle, n_passes, ptr(tables), ptr(np.array(x0s, np.uint16)),
                                     ptr(np.array(y0s, np.uint16)), ptr(np.array(verts, np.uint8)),
                                     ptr(np.array(firsts, np.uint8)), ptr(np.array(nes, np.uint8)),
                                     ptr(np.array(nrs, np.uint8)), d.ptr, w, h, bd)
        assert st == 0
        got = d.read(dt, w * h).reshape(h, w).astype(np.uint16)
        assert np.array_equal(got, want), f"it={it} bd={bd}"


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_mc_batch_per_reference_planes(gpu, bd):
    """inter::predict with each reference its own RefPlane (inter.rs:21-27):
    three planes of different strides and visible sizes (as references of
    another frame size are), single and compound across them."""
    rng = np.random.default_rng(700 + bd)
    lib = vp.library()
    olib = oracle_lib.lib()
    max_v = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    geo = [(150, 90, 160), (76, 45, 96), (300, 180, 320)]  # visible w, h, stride
    refs, planes, keep = [], (vp.RefPlane * 3)(), []
    for k, (pw, ph, stride) in enumerate(geo):
        r = rng.integers(0, max_v + 1, size=(ph + 4, stride)).astype(np.uint16)
        refs.append(r)
        d = gpu.upload(r.astype(dt))
        keep.append(d)
        planes[k].data, planes[k].stride, planes[k].last_x, planes[k].last_y = d.ptr.value, stride, pw - 1, ph - 1
    out_stride = 64 * 8
    descs, expected = [], []
    for i in range(72):
        w, h = [(8, 8), (16, 8), (8, 16), (16, 16), (32, 32), (64, 64), (4, 4), (32, 16)][i % 8]
        a_k, b_k = int(rng.integers(0, 3)), int(rng.integers(0, 3))
        pw, ph, stride = geo[a_k]
        x = int(rng.integers(-20 * 16, (pw + 10) * 16))
        y = int(rng.integers(-20 * 16, (ph + 10) * 16))
        xs, ys = (16, 16) if i % 3 else (int(rng.integers(8, 33)), int(rng.integers(8, 33)))
        f = int(rng.integers(0, 4))
        comp = i % 4 == 1
        x2, y2 = x + int(rng.integers(-64, 64)), y + int(rng.integers(-64, 64))
        d = vp.McDesc()
        d.x, d.y, d.x2, d.y2, d.xs, d.ys, d.w, d.h, d.filter = x, y, x2, y2, xs, ys, w, h, f
        d.ref_idx, d.ref2_idx = a_k, (b_k if comp else -1)
        d.out_off = (i // 8) * 64 * out_stride + (i % 8) * 64
        descs.append(d)
        a = np.zeros((h, w), dtype=np.uint16)
        olib.orc_inter_predict(ptr(refs[a_k]), stride, pw - 1, ph - 1, x, y, xs, ys, w, h, f, bd, ptr(a), w)
        if comp:
            bw, bh, bs = geo[b_k]
            b = np.zeros((h, w), dtype=np.uint16)
            olib.orc_inter_predict(ptr(refs[b_k]), bs, bw - 1, bh - 1, x2, y2, xs, ys, w, h, f, bd, ptr(b), w)
            c = np.zeros((h, w), dtype=np.uint16)
            olib.orc_avg(ptr(a), ptr(b), w, h, ptr(c), w)
            a = c
        expected.append(a)
    rows = (len(descs) // 8 + 1) * 64
    d_out = gpu.upload(np.zeros(rows * out_stride, dtype=dt))
    arr = (vp.McDesc * len(descs))(*descs)
    assert lib.vp9r_mc_batch_planes(gpu.handle, len(descs), ctypes.cast(arr, ctypes.c_void_p),
                                    ctypes.cast(planes, ctypes.c_void_p), 3, d_out.ptr, out_stride, bd) == 0
    out = d_out.read(dt, rows * out_stride).reshape(rows, out_stride).astype(np.uint16)
    for i, d in enumerate(descs):
        r0, c0 = divmod(d.out_off, out_stride)
        assert np.array_equal(out[r0:r0 + d.h, c0:c0 + d.w], expected[i]), f"desc {i}"


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_avg_batch_matches_oracle(gpu, bd):
    """pixel::avg (pixel.rs:22-50) as a batched launch."""
    rng = np.random.default_rng(800 + bd)
    lib = vp.library()
    olib = oracle_lib.lib()
    max_v = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    stride = 256
    n = 40
    descs = (vp.AvgDesc * n)()
    a_all, b_all, want = [], [], np.zeros((n * 64, stride), dtype=np.uint16)
    off = 0
    for i in range(n):
        w, h = [(4, 4), (8, 8), (16, 8), (64, 64), (32, 64), (8, 32), (64, 16), (4, 8)][i % 8]
        a = rng.integers(0, max_v + 1, size=(h, w)).astype(np.uint16)
        b = rng.integers(0, max_v + 1, size=(h, w)).astype(np.uint16)
        descs[i].a_off = descs[i].b_off = off
        descs[i].dst_off = i * 64 * stride + (i % 3) * 64
        descs[i].w, descs[i].h = w, h
        a_all.append(a.ravel()), b_all.append(b.ravel())
        off += w * h
        c = np.zeros((h, w), dtype=np.uint16)
        olib.orc_avg(ptr(a), ptr(b), w, h, ptr(c), w)
        r0, c0 = divmod(descs[i].dst_off, stride)
        want[r0:r0 + h, c0:c0 + w] = c
    da = gpu.upload(np.concatenate(a_all).astype(dt))
    db = gpu.upload(np.concatenate(b_all).astype(dt))
    dd = gpu.upload(np.zeros(n * 64 * stride, dtype=dt))
    assert lib.vp9r_avg_batch(gpu.handle, n, ctypes.cast(descs, ctypes.c_void_p), da.ptr, db.ptr, dd.ptr, stride,
                              bd) == 0
    got = dd.read(dt, n * 64 * stride).reshape(n * 64, stride).astype(np.uint16)
    assert np.array_equal(got, want)


def test_add_const_batch_matches_oracle(gpu):
    """pixel::add_const (pixel.rs:9-16): the clamp of v to +-65536 and of
    the result to 0..255, at every block size."""
    rng = np.random.default_rng(900)
    lib = vp.library()
    olib = oracle_lib.lib()
    stride = 128
    plane = rng.integers(0, 256, size=(64 * 8, stride)).astype(np.uint16)
    want = plane.copy()
    n = 32
    descs = (vp.AddConstDesc * n)()
    for i in range(n):
        size = [4, 8, 16, 32][i % 4]
        v = int([rng.integers(-300, 300), rng.integers(-10, 10), 70000, -70000][i % 4 if i % 8 < 4 else 1])
        r0, c0 = (i // 4) * 64, (i % 4) * 32  # disjoint blocks
        descs[i].dst_off, descs[i].v, descs[i].n = r0 * stride + c0, v, size
        blk = np.ascontiguousarray(want[r0:r0 + size, c0:c0 + size])
        olib.orc_add_const(ptr(blk), size, size, v)
        want[r0:r0 + size, c0:c0 + size] = blk
    dd = gpu.upload(plane.astype(np.uint8))
    assert lib.vp9r_add_const_batch(gpu.handle, n, ctypes.cast(descs, ctypes.c_void_p), dd.ptr, stride) == 0
    got = dd.read(np.uint8, plane.size).reshape(plane.shape).astype(np.uint16)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_intra_predict_batch_matches_oracle(gpu, bd):
    """intra::predict (intra.rs:19-175) as a batched device launch: every
    mode, size and edge availability in one call."""
    rng = np.random.default_rng(1000 + bd)
    lib = vp.library()
    olib = oracle_lib.lib()
    max_v = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    stride = 256
    cases = [(m, l, hl, ha) for m in range(10) for l in (2, 3, 4, 5) for hl, ha in ((1, 1), (0, 1), (1, 0), (0, 0))]
    descs = (vp.IntraDesc * len(cases))()
    edges, want = [], np.zeros((len(cases) // 8 * 32 + 32, stride), dtype=np.uint16)
    eoff = 0
    for i, (m, l, hl, ha) in enumerate(cases):
        size = 1 << l
        above = rng.integers(0, max_v + 1, size=2 * size + 1).astype(np.int32)
        left = rng.integers(0, max_v + 1, size=size).astype(np.int32)
        r0, c0 = (i // 8) * 32, (i % 8) * 32
        descs[i].dst_off, descs[i].edge_off = r0 * stride + c0, eoff
        descs[i].mode, descs[i].log2, descs[i].have_left, descs[i].have_above = m, l, hl, ha
        edges += [above, left]
        eoff += 3 * size + 1
        blk = np.zeros((size, size), dtype=np.uint16)
        olib.orc_intra_predict(m, l, ptr(above), ptr(left), hl, ha, bd, ptr(blk), size)
        want[r0:r0 + size, c0:c0 + size] = blk
    de = gpu.upload(np.concatenate(edges))
    dd = gpu.upload(np.zeros(want.size, dtype=dt))
    assert lib.vp9r_intra_predict_batch(gpu.handle, len(cases), ctypes.cast(descs, ctypes.c_void_p), de.ptr, dd.ptr,
                                        stride, bd) == 0
    got = dd.read(dt, want.size).reshape(want.shape).astype(np.uint16)
    assert np.array_equal(got, want)


def test_context_outlives_its_frames_and_buffers():
    """Frame stores and device buffers keep their context alive: releasing
    the context first, then the objects made from it, is safe, and they
    still work in between."""
    import gc
    lib = vp.library()
    g = vp.Gpu(0)
    f = g.frame(64, 64, 8, 1, 1)
    b = g.upload(np.arange(256, dtype=np.uint8))
    handle = ctypes.c_void_p(g.handle.value)
    g._g = None  # the wrapper forgets it; free the caller's reference now
    lib.vp9r_gpu_free(handle)
    assert np.array_equal(b.read(np.uint8, 256), np.arange(256, dtype=np.uint8))
    s, r = f.geometry(0)
    f.upload_plane(0, np.full((r, s), 7, dtype=np.uint16))
    assert (f.download_plane(0) == 7).all()
    assert len(f.packed()) == vp.library().vp9r_packed_size(64, 64, 8, 1, 1)
    del f, b
    gc.collect()


def test_seam_time_needs_a_timed_call():
    g = vp.Gpu(0)
    with pytest.raises(vp.InvalidInputError):
        g.seam_ms()
