"""Cross-stream batches (vp9r_gpu_run_batch): independent frames of several
streams reconstructed and loop filtered by the same launches.  Each stream's
frames must equal the oracle's, as when the stream runs alone: sizes,
bit depths (8-bit and high-bit-depth frames share a batch), subsampling,
scaled references and lossless frames mixed, streams of unequal length."""
import pytest

import oracle_lib
import rivet_vp9_amd as vp
import streamgen

pytestmark = pytest.mark.gpu

MIX = [
    dict(width=352, height=288, frames=5, seed=11, base_q_idx=120, content=0),
    dict(width=640, height=360, frames=4, seed=12, tile_cols_log2=1, content=1, intra_percent=10),
    dict(width=200, height=120, frames=6, seed=13, resize_every=2),
    dict(width=160, height=96, frames=3, seed=14, bit_depth=10),
    dict(width=160, height=96, frames=4, seed=15, bit_depth=12, ss_x=0, ss_y=0),
    dict(width=160, height=96, frames=4, seed=16, ss_x=1, ss_y=0),
    dict(width=96, height=64, frames=2, seed=17, base_q_idx=0),
]


@pytest.fixture(scope="module")
def gpu():
    return vp.Gpu(0)


class StreamState:
    def __init__(self, gpu, frames):
        self.gpu = gpu
        p = vp.Parser()
        self.preps = [gpu.prepare(p.parse(f)) for f in frames]
        ref = oracle_lib.CpuDecoder()
        self.want = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
        self.slots = [None] * 8
        self.got = []
        self.next = 0

    def job(self):
        """The next frame's (prepared, refs, cur), or None at the end."""
        if self.next >= len(self.preps):
            return None
        pr = self.preps[self.next]
        h = pr.hdr
        cur = self.gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        refs = [None] * 3 if h.frame_is_intra else [self.slots[h.ref_frame_idx[k]] for k in range(3)]
        return pr, refs, cur

    def done(self, job):
        pr, _, cur = job
        h = pr.hdr
        for k in range(8):
            if (h.refresh_frame_flags >> k) & 1:
                self.slots[k] = cur
        if h.show_frame:
            self.got.append(oracle_lib.md5(cur.packed()))
        self.next += 1


def test_mixed_streams_batched_equal_the_oracle(gpu):
    streams = []
    for i, cfg in enumerate(MIX):
        print(f"generating stream {i}: {cfg}", flush=True)
        streams.append(StreamState(gpu, streamgen.generate(**cfg)))
    step = 0
    while True:
        jobs = [(s, s.job()) for s in streams]
        jobs = [(s, j) for s, j in jobs if j is not None]
        if not jobs:
            break
        print(f"batch {step}: {len(jobs)} frames", flush=True)
        gpu.run_batch([j for _, j in jobs])
        gpu.synchronize()
        for s, j in jobs:
            s.done(j)
        step += 1
    for i, s in enumerate(streams):
        assert s.got == s.want, f"stream {i} ({MIX[i]})"


def test_batch_of_one_and_repeated_batches(gpu):
    """The same prepared frames batched twice in a row (slot ring reuse), and
    as batches of one, give the same frames."""
    frames = streamgen.generate(**MIX[0])
    for attempt in range(2):
        s = StreamState(gpu, frames)
        while (j := s.job()) is not None:
            gpu.run_batch([j])
            s.done(j)
        gpu.synchronize()
        assert s.got == s.want, f"pass {attempt}"


def test_batch_rejects_dependent_frames(gpu):
    """A frame that references another frame of the same batch, or two frames
    writing one store, are refused before anything is queued."""
    frames = streamgen.generate(**dict(MIX[0], frames=2))
    p = vp.Parser()
    p0, p1 = [gpu.prepare(p.parse(f)) for f in frames]
    h = p0.hdr
    a = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
    b = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
    assert not p1.hdr.frame_is_intra
    with pytest.raises(vp.InvalidInputError):
        gpu.run_batch([(p0, [None] * 3, a), (p1, [a, a, a], b)])
    with pytest.raises(vp.InvalidInputError):
        gpu.run_batch([(p0, [None] * 3, a), (p0, [None] * 3, a)])
    gpu.run_batch([(p0, [None] * 3, a)])
    gpu.synchronize()


def test_filter_batch_equals_the_oracle(gpu):
    """vp9r_gpu_filter_batch: frames reconstructed with the filter off (their
    records at level 0), then filtered together with their own levels and
    sharpness (intra and inter frames, 8 and 10 bits, 4:4:4), equal the
    oracle's decode; refused while stage timing is on."""
    cases = [dict(width=352, height=288, frames=2, seed=31, lf_level=40, lf_sharpness=5),
             dict(width=200, height=120, frames=2, seed=32, lf_level=63, lf_sharpness=7, bit_depth=10),
             dict(width=160, height=96, frames=2, seed=33, lf_level=20, lf_sharpness=0, ss_x=0, ss_y=0)]
    jobs_lf, want = [], []
    keep = []
    for cfg in cases:
        frames = streamgen.generate(**cfg)
        ref = oracle_lib.CpuDecoder()
        md5s = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
        # frame 0 (intra) decoded alone; frame 1 (inter) predicts from it
        p = vp.Parser()
        rec0 = p.parse(frames[0])
        h = rec0.hdr
        f0 = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        gpu.run(gpu.prepare(rec0), [None] * 3, f0)
        rec1 = p.parse(frames[1])
        assert not rec1.hdr.frame_is_intra
        p_lf = gpu.prepare(rec1)
        rec1.hdr.lf_level = 0
        p_off = gpu.prepare(rec1)
        cur = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        gpu.run(p_off, [f0] * 3, cur)
        jobs_lf.append((p_lf, cur))
        want.append(md5s[1])
        keep += [f0, p_off]
    gpu.synchronize()
    ms = gpu.filter_batch(jobs_lf)
    assert ms > 0
    assert [oracle_lib.md5(c.packed()) for _, c in jobs_lf] == want
    gpu.set_timing(True)
    try:
        with pytest.raises(vp.InvalidInputError):
            gpu.filter_batch(jobs_lf)
    finally:
        gpu.set_timing(False)


def test_filter_batch_large_equals_the_oracle(gpu):
    """A batch of 2040 superblocks or more (four 1080p frames) takes the
    loop filter decisions one workgroup per superblock: the same frames as
    the oracle's decode."""
    frames = streamgen.generate(width=1920, height=1080, frames=2, seed=34, lf_level=36, lf_sharpness=3)
    ref = oracle_lib.CpuDecoder()
    md5s = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
    p = vp.Parser()
    rec0 = p.parse(frames[0])
    h = rec0.hdr
    assert h.sb64_rows * h.sb64_cols * 4 >= 2040
    f0 = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
    gpu.run(gpu.prepare(rec0), [None] * 3, f0)
    rec1 = p.parse(frames[1])
    p_lf = gpu.prepare(rec1)
    rec1.hdr.lf_level = 0
    p_off = gpu.prepare(rec1)
    jobs = []
    for _ in range(4):
        cur = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        gpu.run(p_off, [f0] * 3, cur)
        jobs.append((p_lf, cur))
    gpu.synchronize()
    gpu.filter_batch(jobs)
    assert [oracle_lib.md5(c.packed()) for _, c in jobs] == [md5s[1]] * 4
