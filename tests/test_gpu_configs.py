"""The benchmark configurations run exactly, on the GPU, against the oracle.

Config 1: the 352x288 8-bit 4:2:0 sinusoid stream, 30 frames from seed 1
(base_q_idx 120, golden refresh every 8 frames, one tile column, no
high-precision motion vectors), every frame compared with the oracle after
reconstruction and after the loop filter, and the decoder API's per-frame
MD5s equal to the oracle's (the reference's per-frame MD5 check,
tests/common/mod.rs:123-182, examples/vp9_bench.rs:112-135).

Config 5: the seed-5 serving mix, all 32 streams with their sizes (640x360,
1280x720, 1920x1080) and formats (8-bit 4:2:0, 10-bit 4:2:0, 12-bit 4:4:4),
each cut to its first 6 frames, served together through vp9r_server_*;
every frame equals the oracle's decode of its stream alone, and one stream
of each format served alone equals it too."""
import pytest

import gpu_lockstep
import oracle_lib
import rivet_vp9_amd as vp
import streamgen
from test_gpu_server import fmt, serve

pytestmark = pytest.mark.gpu

CONFIG5_FRAMES = 6


def oracle_md5s(frames):
    ref = oracle_lib.CpuDecoder()
    return [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]


def test_config1_stream_lockstep_and_decoder_md5s():
    frames = streamgen.generate(**streamgen.CONFIG1)
    assert len(frames) == 30
    n, err = gpu_lockstep.run_frames(frames, vp.Gpu(0))
    assert err is None, err
    assert n == 30
    want = oracle_md5s(frames)
    dec = vp.Decoder(0)
    got = [oracle_lib.md5(fr.packed()) for f in frames for fr in dec.decode_all(f)]
    assert got == want
    assert dec.frames_decoded() == 30


@pytest.fixture(scope="module")
def config5():
    items = []
    for _, cfg in streamgen.config5(5):
        cfg = dict(cfg, frames=CONFIG5_FRAMES)
        frames = streamgen.generate(**cfg)
        items.append((cfg, frames, oracle_md5s(frames)))
    return items


def test_config5_mix_has_every_size_and_format(config5):
    sizes = {(c["width"], c["height"]) for c, _, _ in config5}
    formats = {(c["bit_depth"], c["ss_x"], c["ss_y"]) for c, _, _ in config5}
    assert len(config5) == 32
    assert sizes == {(640, 360), (1280, 720), (1920, 1080)}
    assert formats == {(8, 1, 1), (10, 1, 1), (12, 0, 0)}


def test_config5_all_32_streams_together(config5):
    srv = vp.Server(0, parse_threads=8)
    got, status, lat = serve(srv, [(cfg, frames) for cfg, frames, _ in config5], timeout_s=300)
    st = srv.stats()
    srv.close_server()
    print(f"32 streams: {st}; latency max {max(lat):.2f} ms", flush=True)
    assert status == [vp.OK] * 32
    for i, (cfg, _, want) in enumerate(config5):
        assert got[i] == want, f"stream {i} {fmt(cfg)}"
    assert st["max_batch"] > 1


def test_config5_streams_alone(config5):
    picked = {}
    for cfg, frames, want in config5:
        picked.setdefault((cfg["bit_depth"], cfg["ss_x"]), (cfg, frames, want))
    for cfg, frames, want in picked.values():
        print(f"alone: {fmt(cfg)}", flush=True)
        srv = vp.Server(0)
        got, status, _ = serve(srv, [(cfg, frames)])
        srv.close_server()
        assert status == [vp.OK] and got[0] == want
