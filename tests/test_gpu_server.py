"""The multi-stream server (vp9r_server_*): every stream's frames equal the
oracle's, decoded alone and in a mix of sizes, bit depths and subsamplings;
admission under a small HBM budget (streams wait, then run); a stream that
fails ends alone with its error while the others finish."""
import pytest

import oracle_lib
import rivet_vp9_amd as vp
import streamgen

pytestmark = pytest.mark.gpu

MIX = [
    dict(width=352, height=288, frames=6, seed=21, base_q_idx=120, content=0),
    dict(width=640, height=360, frames=5, seed=22, tile_cols_log2=1, content=1, intra_percent=10),
    dict(width=200, height=120, frames=7, seed=23, resize_every=3),
    dict(width=160, height=96, frames=4, seed=24, bit_depth=10),
    dict(width=160, height=96, frames=4, seed=25, bit_depth=12, ss_x=0, ss_y=0),
    dict(width=160, height=96, frames=4, seed=26, ss_x=1, ss_y=0),
    dict(width=96, height=64, frames=3, seed=27, base_q_idx=0),
]

_streams = {}


def stream(i):
    """(frames, oracle MD5s) of MIX[i], generated once."""
    if i not in _streams:
        frames = streamgen.generate(**MIX[i])
        ref = oracle_lib.CpuDecoder()
        want = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
        _streams[i] = (frames, want)
    return _streams[i]


def fmt(cfg):
    return (cfg["width"], cfg["height"], cfg.get("bit_depth", 8), cfg.get("ss_x", 1), cfg.get("ss_y", 1))


def serve(server, items, timeout_s=60):
    """Opens a stream per (cfg, frames), submits every packet, closes, and
    collects MD5s per stream until every stream has ended: returns
    ({stream: [md5]}, {stream: status}, latencies)."""
    import hashlib
    import time
    ids = []
    for cfg, frames in items:
        sid = server.open(*fmt(cfg))
        ids.append(sid)
    for sid, (cfg, frames) in zip(ids, items):
        for f in frames:
            server.submit(sid, f)
        server.close(sid)
    got = {sid: [] for sid in ids}
    status = {}
    lat = []
    deadline = time.time() + timeout_s
    while len(status) < len(ids):
        assert time.time() < deadline, f"server did not finish: {server.stats()}"
        fr = server.poll(1000)
        if fr is None:
            continue
        if fr.eos:
            status[fr.stream] = fr.status
            continue
        assert fr.index == len(got[fr.stream])
        got[fr.stream].append(hashlib.md5(fr.view()).hexdigest())
        lat.append(fr.latency_ms)
        fr.release()
    return [got[s] for s in ids], [status[s] for s in ids], lat


def test_each_stream_alone_equals_the_oracle():
    for i, cfg in enumerate(MIX):
        print(f"stream {i} alone: {cfg}", flush=True)
        frames, want = stream(i)
        srv = vp.Server(0)
        got, status, _ = serve(srv, [(cfg, frames)])
        srv.close_server()
        assert status == [vp.OK]
        assert got[0] == want, f"stream {i}"


def test_mixed_streams_equal_the_oracle():
    items = [(cfg, stream(i)[0]) for i, cfg in enumerate(MIX)]
    srv = vp.Server(0, parse_threads=4)
    got, status, lat = serve(srv, items)
    st = srv.stats()
    srv.close_server()
    print(f"stats {st}; latency max {max(lat):.2f} ms", flush=True)
    assert status == [vp.OK] * len(MIX)
    for i in range(len(MIX)):
        assert got[i] == stream(i)[1], f"stream {i}"
    assert st["frames_shown"] == sum(len(stream(i)[1]) for i in range(len(MIX)))
    assert st["max_batch"] > 1  # frames of several streams shared launches


def test_admission_under_a_small_budget():
    """A budget for two 640x360 footprints: streams wait for admission and
    all finish, with the same frames."""
    big = vp.Server.footprint(640, 360)
    items = [(cfg, stream(i)[0]) for i, cfg in enumerate(MIX)]
    srv = vp.Server(0, hbm_budget=2 * big, parse_threads=2)
    ids = [srv.open(*fmt(cfg)) for cfg, _ in items]
    st0 = srv.stats()
    assert st0["streams_waiting"] > 0
    for sid, (cfg, frames) in zip(ids, items):
        for f in frames:
            srv.submit(sid, f)
        srv.close(sid)
    import hashlib
    got = {sid: [] for sid in ids}
    ended = {}
    while len(ended) < len(ids):
        fr = srv.poll(20000)
        assert fr is not None, f"stalled: {srv.stats()}"
        if fr.eos:
            ended[fr.stream] = fr.status
            continue
        got[fr.stream].append(hashlib.md5(fr.view()).hexdigest())
        fr.release()
        assert srv.stats()["store_bytes"] <= 2 * big
    srv.close_server()
    for i, sid in enumerate(ids):
        assert ended[sid] == vp.OK
        assert got[sid] == stream(i)[1], f"stream {i}"
    with pytest.raises(vp.InvalidInputError):
        vp.Server(0, hbm_budget=big // 2).open(640, 360)


def test_a_failing_stream_ends_alone():
    """A truncated packet ends its stream with an error; a frame larger than
    the declared size ends its stream with InvalidInput; the others finish
    with their frames."""
    f0, want0 = stream(0)
    f1, want1 = stream(1)
    f3, _ = stream(3)
    broken = list(f3[:2]) + [f3[2][:7]] + list(f3[3:])
    items = [(MIX[0], f0), (MIX[3], broken), (dict(MIX[1], width=320, height=180), f1)]
    srv = vp.Server(0)
    got, status, _ = serve(srv, items)
    srv.close_server()
    assert status[0] == vp.OK and got[0] == want0
    assert status[1] != vp.OK
    assert status[2] == vp.ERR_INVALID_INPUT
