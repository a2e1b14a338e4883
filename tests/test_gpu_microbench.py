"""GPU parity of the kernel microbenchmark's batches (tests/microbench.py)
against the oracle, every block, through the C ABI: 2M inverse-transform
blocks (vp9r_itx_add_batch, itx.rs:701-749) and 1M motion-compensation
blocks (vp9r_mc_batch, inter.rs:289-395 and pixel.rs:22-50) from seed 3 at
8 bits, and a smaller 10-bit batch of each.  Bit-exact."""
import os

import numpy as np
import pytest

import microbench as mb
import rivet_vp9_amd as vp

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def gpu():
    return vp.Gpu(0)


def run_itx(gpu, s):
    dt = np.uint8 if s.bit_depth == 8 else np.uint16
    d_plane = gpu.upload(s.plane.astype(dt))
    bufs = [gpu.upload(a) for a in (s.coefs, s.coef_off, s.tx_type, s.eob, s.dst_off)]
    st = vp.library().vp9r_itx_add_batch(gpu.handle, s.count, bufs[0].ptr, bufs[1].ptr, mb.ptr(s.tx),
                                         bufs[2].ptr, bufs[3].ptr, bufs[4].ptr, d_plane.ptr, mb.PLANE_STRIDE,
                                         s.bit_depth, 0)
    assert st == 0, vp.library().vp9r_last_error()
    assert gpu.seam_ms() > 0
    return d_plane.read(dt, s.rows * mb.PLANE_STRIDE).reshape(s.rows, mb.PLANE_STRIDE)


def run_mc(gpu, m):
    import ctypes
    dt = np.uint8 if m.bit_depth == 8 else np.uint16
    d_refs = [gpu.upload(r.astype(dt)) for r in m.refs]
    ref_ptrs = (ctypes.c_void_p * mb.N_REFS)(*[d.ptr.value for d in d_refs])
    d_out = gpu.upload(np.zeros(m.rows * mb.PLANE_STRIDE, dtype=dt))
    st = vp.library().vp9r_mc_batch(gpu.handle, m.count, mb.ptr(m.desc), ref_ptrs, mb.N_REFS, mb.REF_W,
                                    mb.REF_W - 1, mb.REF_H - 1, d_out.ptr, mb.PLANE_STRIDE, m.bit_depth)
    assert st == 0, vp.library().vp9r_last_error()
    assert gpu.seam_ms() > 0
    return d_out.read(dt, m.rows * mb.PLANE_STRIDE).reshape(m.rows, mb.PLANE_STRIDE)


def first_diff(got, want):
    r, c = np.argwhere(got.astype(np.uint16) != want)[0]
    return f"first difference at row {r} col {c}: {got[r, c]} != {want[r, c]}"


@pytest.mark.parametrize("n,bd", [(2_000_000, 8), (100_000, 10)])
def test_itx_batch_every_block(gpu, n, bd):
    print(f"itx batch: {n} blocks, seed 3, {bd}-bit", flush=True)
    s = mb.ItxSet(n, 3, bd)
    got = run_itx(gpu, s)
    want = mb.oracle_itx(s, THREADS)
    assert np.array_equal(got, want), first_diff(got, want)


@pytest.mark.parametrize("n,bd", [(1_000_000, 8), (50_000, 10)])
def test_mc_batch_every_block(gpu, n, bd):
    print(f"mc batch: {n} blocks, seed 3, {bd}-bit", flush=True)
    m = mb.McSet(n, 3, bd)
    got = run_mc(gpu, m)
    want = mb.oracle_mc(m, THREADS)
    assert np.array_equal(got, want), first_diff(got, want)


@pytest.mark.parametrize("persistent", [0, vp.SHAPE_PERSISTENT])
@pytest.mark.parametrize("reg", [0, vp.SHAPE_REG])
@pytest.mark.parametrize("waves", [1, 2, 4])
@pytest.mark.parametrize("bd", [8, 10])
def test_itx_every_tuned_shape(bd, waves, reg, persistent):
    """Each launch shape the autotuner may choose (1, 2 or 4 waves per
    transform workgroup; transposes through LDS or in registers, which
    32x32 lacks; one workgroup per blocks or persistent workgroups) gives
    the oracle's output at every size."""
    g = vp.Gpu(0)
    small = waves | reg | persistent
    g.set_tuning([small] * 3 + [waves | persistent], [small] * 3 + [waves | persistent], 32, 32)
    s = mb.ItxSet(60_000, 30 + waves + reg + persistent, bd)
    got = run_itx(g, s)
    want = mb.oracle_itx(s, THREADS)
    assert np.array_equal(got, want), first_diff(got, want)


def test_itx_register_shape_refused_at_32x32():
    g = vp.Gpu(0)
    with pytest.raises(vp.InvalidInputError):
        g.set_tuning([1, 1, 1, 1 | vp.SHAPE_REG], [1] * 4, 32, 32)


@pytest.mark.parametrize("rows", [16, 32])
@pytest.mark.parametrize("bd", [8, 10])
def test_mc_every_tuned_shape(bd, rows):
    """Tiles of 16 and of 32 rows give the oracle's output."""
    g = vp.Gpu(0)
    g.set_tuning([1] * 4, [1] * 4, rows, rows)
    m = mb.McSet(30_000, 40 + rows, bd)
    got = run_mc(g, m)
    want = mb.oracle_mc(m, THREADS)
    assert np.array_equal(got, want), first_diff(got, want)


def test_autotune_records_its_choices():
    g = vp.Gpu(0)
    assert g.tuning()["tuned"] is False
    choice, times = g.autotune()
    print(choice, [round(t, 1) for t in times], flush=True)
    assert choice["tuned"] is True
    shapes = choice["itx_shape_8bit"] + choice["itx_shape_hbd"]
    assert all((c & 15) in (1, 2, 4) for c in shapes)
    assert not (choice["itx_shape_8bit"][3] & vp.SHAPE_REG or choice["itx_shape_hbd"][3] & vp.SHAPE_REG)
    assert choice["mc_tile_rows_8bit"] in (16, 32) and choice["mc_tile_rows_hbd"] in (16, 32)
    assert len(times) == vp.Gpu.TUNE_TIMES
    # every shape timed except register transposes at 32x32
    for c in range(2):
        for k in range(4):
            for cand in range(12):
                t = times[(c * 4 + k) * 12 + cand]
                assert (t < 0) == (k == 3 and (cand // 3) % 2 == 1), (c, k, cand, t)
    assert all(t > 0 for t in times[96:])
    # the tuned context still gives the oracle's output
    s = mb.ItxSet(20_000, 5, 8)
    assert np.array_equal(run_itx(g, s), mb.oracle_itx(s, THREADS))
