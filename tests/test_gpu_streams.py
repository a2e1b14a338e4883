"""GPU parity on generated streams: features the committed streams do not
reach (several tile columns, scaled references, 4:2:2 / 4:4:0 / 4:4:4 at
every bit depth, the 1080p benchmark workload), every frame compared with
the CPU oracle after reconstruction and after the loop filter, with intra
work as leveled blocks (the default) and as (superblock, plane) groups."""
import os

import pytest

import gpu_lockstep
import oracle_lib
import rivet_vp9_amd as vp
import streamgen

pytestmark = pytest.mark.gpu

CASES = {
    "sinusoid-352x288": dict(width=352, height=288, frames=6, seed=1, base_q_idx=120, content=0),
    "rects-640x360-2tiles": dict(width=640, height=360, frames=5, seed=2, tile_cols_log2=1, content=1,
                                 intra_percent=10),
    "resize-200x120": dict(width=200, height=120, frames=8, seed=3, resize_every=3),
    "10bit-420": dict(width=160, height=96, frames=4, seed=5, bit_depth=10),
    "12bit-444": dict(width=160, height=96, frames=4, seed=5, bit_depth=12, ss_x=0, ss_y=0),
    "8bit-422": dict(width=160, height=96, frames=4, seed=6, ss_x=1, ss_y=0),
    "10bit-440": dict(width=160, height=96, frames=4, seed=7, bit_depth=10, ss_x=0, ss_y=1),
    "lossless": dict(width=96, height=64, frames=3, seed=8, base_q_idx=0),
    # loop filter sharpness > 0 (limits shifted and capped, lf.rs / 8.8.1) and
    # the extreme levels
    "sharpness5-level40": dict(width=352, height=288, frames=3, seed=9, content=0, lf_level=40, lf_sharpness=5),
    "sharpness7-level63-10bit": dict(width=320, height=192, frames=3, seed=10, bit_depth=10, lf_level=63,
                                     lf_sharpness=7),
    "sharpness3-level1-444": dict(width=160, height=96, frames=3, seed=11, ss_x=0, ss_y=0, lf_level=1,
                                  lf_sharpness=3),
    "level0-filter-off": dict(width=160, height=96, frames=3, seed=12, lf_level=0),
    "1080p-workload": dict(streamgen.CONFIG2, frames=3),
    # segmentation: per-segment quantisers (ALT_Q) and filter levels (ALT_L),
    # absolute and delta, forced reference and skip, written, temporally
    # predicted and kept maps (recon.rs:219-230, loopfilter.rs:15-47)
    "segmentation-352x288": dict(width=352, height=288, frames=6, seed=61, content=1, intra_percent=10,
                                 segmentation=1),
    "segmentation-640x360-2tiles-10bit": dict(width=640, height=360, frames=4, seed=62, tile_cols_log2=1,
                                              bit_depth=10, content=1, intra_percent=10, segmentation=1),
    "segmentation-12bit-444": dict(width=160, height=96, frames=4, seed=63, bit_depth=12, ss_x=0, ss_y=0,
                                   segmentation=1),
    # the 4K 10-bit workload, resized every other frame: 2:1 and 1:2 scaled references
    "4k-10bit-scaled": dict(streamgen.CONFIG4, frames=4, resize_every=2),
}


@pytest.fixture(scope="module")
def gpu():
    return vp.Gpu(0)


@pytest.fixture(scope="module")
def gpu_one_launch():
    old = os.environ.get("VP9R_WAVEFRONT")
    os.environ["VP9R_WAVEFRONT"] = "one"
    try:
        g = vp.Gpu(0)
    finally:
        if old is None:
            del os.environ["VP9R_WAVEFRONT"]
        else:
            os.environ["VP9R_WAVEFRONT"] = old
    return g


@pytest.mark.parametrize("name", list(CASES))
def test_stream_lockstep(gpu, name):
    print(f"stream {name}", flush=True)
    frames = streamgen.generate(**CASES[name])
    n, err = gpu_lockstep.run_frames(frames, gpu)
    assert err is None, f"{name}: {n} frames equal, then {err}"
    assert n == len(frames)


@pytest.mark.skip(reason="the one-launch wavefront kernels did not finish on the GPU (device-side waits "
                         "across workgroups); the schedule was removed and VP9R_WAVEFRONT=one is an error")
@pytest.mark.parametrize("name", ["rects-640x360-2tiles", "resize-200x120", "12bit-444"])
def test_stream_lockstep_one_launch(gpu_one_launch, name):
    print(f"stream {name} (one-launch wavefronts)", flush=True)
    frames = streamgen.generate(**CASES[name])
    n, err = gpu_lockstep.run_frames(frames, gpu_one_launch)
    assert err is None, f"{name}: {n} frames equal, then {err}"


def test_one_launch_schedule_is_an_error():
    """The removed one-launch schedule is refused, not silently replaced."""
    old = os.environ.get("VP9R_WAVEFRONT")
    os.environ["VP9R_WAVEFRONT"] = "one"
    try:
        with pytest.raises(vp.UnsupportedError):
            vp.Gpu(0)
    finally:
        if old is None:
            del os.environ["VP9R_WAVEFRONT"]
        else:
            os.environ["VP9R_WAVEFRONT"] = old


@pytest.fixture(scope="module")
def gpu_groups():
    old = os.environ.get("VP9R_INTRA")
    os.environ["VP9R_INTRA"] = "groups"
    try:
        g = vp.Gpu(0)
    finally:
        if old is None:
            del os.environ["VP9R_INTRA"]
        else:
            os.environ["VP9R_INTRA"] = old
    return g


@pytest.mark.parametrize("name", ["1080p-workload", "12bit-444", "8bit-422"])
def test_stream_lockstep_intra_groups(gpu_groups, name):
    """Intra as (superblock, plane) groups (VP9R_INTRA=groups)."""
    print(f"stream {name} (intra groups)", flush=True)
    frames = streamgen.generate(**CASES[name])
    n, err = gpu_lockstep.run_frames(frames, gpu_groups)
    assert err is None, f"{name}: {n} frames equal, then {err}"


@pytest.fixture(scope="module")
def gpu_lf_levels():
    old = os.environ.get("VP9R_LF")
    os.environ["VP9R_LF"] = "levels"
    try:
        g = vp.Gpu(0)
    finally:
        if old is None:
            del os.environ["VP9R_LF"]
        else:
            os.environ["VP9R_LF"] = old
    return g


@pytest.mark.parametrize("name", ["1080p-workload", "12bit-444", "8bit-422"])
def test_stream_lockstep_lf_levels(gpu_lf_levels, name):
    """One loop filter level per launch throughout (VP9R_LF=levels; by
    default a single frame's last levels go two per launch)."""
    print(f"stream {name} (loop filter levels)", flush=True)
    frames = streamgen.generate(**CASES[name])
    n, err = gpu_lockstep.run_frames(frames, gpu_lf_levels)
    assert err is None, f"{name}: {n} frames equal, then {err}"


def test_prepared_run_equals_reconstruct(gpu):
    """vp9r_gpu_prepare + vp9r_gpu_run (the benchmark's path) gives the same
    frames as vp9r_gpu_reconstruct + vp9r_gpu_loop_filter, run twice."""
    frames = streamgen.generate(**dict(CASES["rects-640x360-2tiles"], frames=4))
    p = vp.Parser()
    recs = []
    for f in frames:
        rec = p.parse(f)
        recs.append(gpu.prepare(rec))
    ref = oracle_lib.CpuDecoder()
    want = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
    for attempt in range(2):
        slots = [None] * 8
        got = []
        for pr in recs:
            h = pr.hdr
            cur = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
            refs = [None] * 3 if h.frame_is_intra else [slots[h.ref_frame_idx[k]] for k in range(3)]
            gpu.run(pr, refs, cur)
            for k in range(8):
                if (h.refresh_frame_flags >> k) & 1:
                    slots[k] = cur
            if h.show_frame:
                got.append(oracle_lib.md5(cur.packed()))
        gpu.synchronize()
        assert got == want, f"pass {attempt}"


def test_async_download_matches(gpu):
    """vp9r_gpu_download_packed_async into pinned memory equals the
    synchronous download, with the frame store rewritten afterwards."""
    frames = streamgen.generate(**dict(CASES["sinusoid-352x288"], frames=3))
    p = vp.Parser()
    preps = [gpu.prepare(p.parse(f)) for f in frames]
    h = preps[0].hdr
    size = vp.library().vp9r_packed_size(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
    host = vp.HostBuffer(size)
    a = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
    b = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
    gpu.run(preps[0], [None] * 3, a)
    want = a.packed()
    a.packed_async(host)
    # frame 1 predicts from a into b; then a is rewritten by frame 0 again
    gpu.run(preps[1], [a, a, a], b)
    gpu.run(preps[0], [None] * 3, a)
    a.wait()
    assert host.bytes(size) == want
    gpu.synchronize()


def test_reused_frame_stores_match(gpu):
    """Frame stores reused across frames (and filled with noise first) give
    the oracle's frames: reconstruction must not depend on what a store
    held before.  1080p 4:2:0 has an odd number of 8x8 rows, whose last
    chroma 4x4 edge reads past the 8x8-aligned area."""
    import numpy as np
    frames = streamgen.generate(**dict(streamgen.CONFIG2, frames=6))
    ref = oracle_lib.CpuDecoder()
    want = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
    p = vp.Parser()
    preps = [gpu.prepare(p.parse(f)) for f in frames]
    h = preps[0].hdr
    rng = np.random.default_rng(9)
    pool = [gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y) for _ in range(3)]
    for f in pool:
        for pl in range(3):
            s, r = f.geometry(pl)
            f.upload_plane(pl, rng.integers(0, 256, (r, s)))
    slots = [None] * 8
    got = []
    for pr in preps:
        hh = pr.hdr
        busy = {id(x) for x in slots if x is not None}
        cur = next(f for f in pool if id(f) not in busy)
        refs = [None] * 3 if hh.frame_is_intra else [slots[hh.ref_frame_idx[k]] for k in range(3)]
        gpu.run(pr, refs, cur)
        for k in range(8):
            if (hh.refresh_frame_flags >> k) & 1:
                slots[k] = cur
        if hh.show_frame:
            got.append(oracle_lib.md5(cur.packed()))
    gpu.synchronize()
    assert got == want


def test_decoder_api_on_the_1080p_workload():
    """The decoder API (its own frame pool) on the first frames of the
    1080p workload equals the oracle frame for frame."""
    frames = streamgen.generate(**dict(streamgen.CONFIG2, frames=10))
    ref = oracle_lib.CpuDecoder()
    d = vp.Decoder(0)
    for i, f in enumerate(frames):
        want = [oracle_lib.md5(o.packed()) for o in ref.decode(f)]
        got = [oracle_lib.md5(o.packed()) for o in d.decode_all(f)]
        assert got == want, f"frame {i}"
