#!/usr/bin/env python3
"""Single-stream GPU reconstruction benchmark: 1080p decoded frames/s.

Workload: the 1920x1080 8-bit 4:2:0 profile-0 stream of 60 frames from seed 2
(64 moving textured rectangles over a panning background, 4 tile columns,
base_q_idx 100, NEWMV / NEARESTMV / intra), made by the committed stream
generator (tests/streamgen).  Every frame is parsed on the CPU and its record
(blocks, transform blocks, coefficients, mode info) plus work lists are
copied to HBM before the timed region; a step is the GPU reconstruction and
loop filter of one frame (vp9r_gpu_run) from those resident inputs, with the
reference slots chained exactly as the decoder chains them.  The stream is
replayed from its key frame once it ends.

Also printed (in the same JSON line): the roofline of the dominant stage and
per-stage times, measured with HIP events on the stream the kernels run on
over a second timed region of the same frames (the events cost the stream a
few microseconds each, so the headline region has none; "stage_pass_fps" is
that region's rate); the PCIe-inclusive rate (every frame also copied to
pinned host memory on the copy stream); synchronous per-frame latency
p50/p99; and the CPU baseline: the oracle's reconstruction and loop filter
of the same records on all host cores (tile columns in parallel, the loop
filter's superblock rows as a wavefront: the reference's threading), with
the one-core scalar figure inside, whose per-frame MD5s the GPU's output is
checked against (bit_exact).  "roofline.traffic" is the dominant kernel's
HBM bytes per launch from the committed counter profile.

End to end ("end_to_end"): the same stream from packets to packed frames in
host memory, through the decoder API (tile columns parsed in parallel) and
through the server with this one stream (parse running ahead of the GPU),
against the all-core CPU whole decode.

Config 1 ("config1"): the 352x288 sinusoid stream of 30 frames from seed 1,
decoded by the scalar CPU decoder on one thread (frames/s and per-frame
MD5s), and by the GPU decoder API (its MD5s must equal them).

Batched stages ("batched"): one inter frame of the stream reconstructed 64
times in one batch (every launch covers all 64 copies), per-stage GB/s of
algorithmic bytes against the 8 TB/s HBM peak: the kernels with enough work
to fill the GPU.

4K 10-bit ("uhd_10bit"): the 3840x2160 10-bit 4:2:0 stream of 48 frames from
seed 4 whose every 12th frame is coded at 1920x1080 (2:1 and 1:2 scaled
references), frames/s from resident records and its MD5s against the oracle.

Serving (in the same JSON line, "serving"): 32 streams from seed 5 arriving
Poisson at 4/s, mixed sizes (640x360 / 1280x720 / 1920x1080) and formats
(8-bit 4:2:0, 10-bit 4:2:0, 12-bit 4:4:4), 90-300 frames each, served by the
multi-stream server (vp9r_server_*): paced at 30 frames/s per stream from its
arrival (per-frame latency p50/p99 from packet submit to the frame in host
memory, aggregate frames/s), and saturated (every packet submitted at once:
the server's capacity).  Every served frame's MD5 is checked against the
oracle's decode of its stream, which, one process per stream on all host
cores, is also the serving CPU baseline.

Kernel microbenchmark ("microbench"): 2M inverse-transform blocks and 1M
motion-compensation blocks from seed 3 (tests/microbench.py), one batched
call each; kernel device time, algorithmic GB/s, bit-exactness against the
oracle over every block, and the oracle on one core as CPU baseline.  Its
"lf" part: 64 1080p key frames with random filter levels and sharpness,
filtered in one batch (vp9r_gpu_filter_batch), MD5s against the oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S] [--no-cpu] [--no-serve]
                    [--no-batched] [--no-uhd] [--no-micro] [--micro-only] [--no-extras]
"""
import argparse
import ctypes
import hashlib
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "rivet-vp9_amd")
TESTS = os.path.join(ROOT, "tests")

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
STAGES = ("inter_pred", "inter_residual", "intra", "loop_filter")
# inter_pred: the margin clear, the loop filter decisions and the inter
# prediction (one launch); intra: the fused launches holding an intra
# level (the loop filter levels placed beside them run there too);
# loop_filter: the filter levels after the last intra level (several per
# k_lf_cone launch, a few one per k_lf launch)
STAGE_KERNELS = ("k_prep_mc", "k_itx_inter", "k_wave", "k_lf_cone + k_lf")
STAGE_BYTES = ("mc_bytes", "itx_bytes", "intra_bytes", "lf_bytes")


def _package():
    if "rivet_vp9_amd" in sys.modules:
        return sys.modules["rivet_vp9_amd"]
    spec = importlib.util.spec_from_file_location("rivet_vp9_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rivet_vp9_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


class Schedule:
    """The reference slot chaining of the decoder (decoder/mod.rs:398-415)
    over a pool of HBM frame stores, for a list of parsed frames."""

    def __init__(self, vp, gpu, frames):
        self.vp, self.gpu, self.frames = vp, gpu, frames
        self.slots = [None] * 8
        self.pool = []

    def store(self, h):
        busy = {id(s) for s in self.slots if s is not None}
        for f in self.pool:
            if id(f) not in busy and (f.width, f.height, f.bit_depth, f.ss_x, f.ss_y) == (
                    h.width, h.height, h.bit_depth, h.ss_x, h.ss_y):
                return f
        f = self.gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        self.pool.append(f)
        return f

    def step(self, i, graphs=None):
        """Queues frame i; returns the frame store it shows (or None).  With
        `graphs` (a dict): through frame i's HIP graph, captured on first
        use and re-targeted when its frame stores differ."""
        fr = self.frames[i]
        h = fr["hdr"]
        if h.show_existing_frame:
            return self.slots[h.frame_to_show_map_idx]
        cur = self.store(h)
        refs = [None] * 3 if h.frame_is_intra else [self.slots[h.ref_frame_idx[k]] for k in range(3)]
        if graphs is None:
            self.gpu.run(fr["prep"], refs, cur)
        else:
            key = (tuple(id(r) for r in refs), id(cur))
            if i not in graphs:
                graphs[i] = [self.gpu.graph(fr["prep"], refs, cur), key]
            elif graphs[i][1] != key:
                graphs[i][0].update(refs, cur)
                graphs[i][1] = key
            graphs[i][0].run()
        for k in range(8):
            if (h.refresh_frame_flags >> k) & 1:
                self.slots[k] = cur
        return cur if h.show_frame else None


def cpu_baseline(frames_bytes, budget_s, threads=1):
    """The oracle's reconstruction + loop filter of the parsed records, best
    pass over the whole stream, repeated within the time budget; also the
    per-frame MD5s of the shown frames.  threads > 1: tile columns
    reconstructed in parallel and the loop filter's superblock rows as a
    wavefront (the reference's decode_tiles_parallel and filter_bands).
    Returns (recon+LF frames/s, MD5s, passes, seconds, whole-decode frames/s
    with the parse, tile columns parsed in parallel when threads > 1)."""
    vp = _package()
    if TESTS not in sys.path:
        sys.path.insert(0, TESTS)
    import oracle_lib as ol
    best, best_whole, md5s, passes, spent = None, None, None, 0, 0.0
    while passes < 1 or (spent < budget_s and passes < 5):
        parser = vp.Parser()
        if threads > 1:
            parser.set_threads(threads)
        slots = [None] * 8
        md5_pass, t_recon, t_parse = [], 0.0, 0.0
        for data in frames_bytes:
            tp = time.perf_counter()
            rec = parser.parse(data)
            t_parse += time.perf_counter() - tp
            h = rec.hdr
            if h.show_existing_frame:
                md5_pass.append(ol.md5(slots[h.frame_to_show_map_idx].packed()))
                continue
            cur = ol.Frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
            refs = (ctypes.POINTER(ol.OrcFrame) * 3)()
            if not h.frame_is_intra:
                for k in range(3):
                    refs[k] = slots[h.ref_frame_idx[k]].f
            t0 = time.perf_counter()
            if threads > 1:
                st = ol.lib().orc_reconstruct_mt(ctypes.byref(rec), refs, cur.f, threads)
            else:
                st = ol.lib().orc_reconstruct(ctypes.byref(rec), refs, cur.f)
            if st:
                raise RuntimeError("oracle reconstruction failed")
            if threads > 1:
                ol.lib().orc_loop_filter_mt(ctypes.byref(rec), cur.f, threads)
            else:
                ol.lib().orc_loop_filter(ctypes.byref(rec), cur.f)
            t_recon += time.perf_counter() - t0
            for k in range(8):
                if (h.refresh_frame_flags >> k) & 1:
                    slots[k] = cur
            if h.show_frame:
                md5_pass.append(ol.md5(cur.packed()))
        passes += 1
        spent += t_recon + t_parse
        fps = len(frames_bytes) / t_recon
        whole = len(frames_bytes) / (t_recon + t_parse)
        best = fps if best is None else max(best, fps)
        best_whole = whole if best_whole is None else max(best_whole, whole)
        md5s = md5_pass
    return best, md5s, passes, spent, best_whole


def cpu_baselines(frames_bytes, budget_s, label):
    """The one-core and all-core CPU baselines of a stream (recon + loop
    filter of parsed records, the GPU's scope, and the whole decode with the
    parse), as a cpu_baseline object: the all-core figure first (the
    comparison the target names), the one-core one inside."""
    cores = _host_cores()
    one, md5_one, p1, s1, whole1 = cpu_baseline(frames_bytes, budget_s, 1)
    alls, md5_all, pa, sa, whole_a = cpu_baseline(frames_bytes, budget_s / 2, cores)
    n = len(frames_bytes)
    out = {"value": round(alls, 3), "unit": "frames/s", "cores": cores, "kind": "port",
           "sample": f"oracle reconstruction + loop filter of all {n} frames of {label} (records parsed outside "
                     f"the timing), tile columns in parallel and the loop filter's superblock rows as a "
                     f"wavefront on {cores} threads, best of {pa} passes, {sa:.1f}s",
           "whole_decode_fps": round(whole_a, 3),
           "one_core": {"value": round(one, 3), "unit": "frames/s", "cores": 1, "kind": "port",
                        "sample": f"the same, scalar on one core, best of {p1} passes, {s1:.1f}s",
                        "whole_decode_fps": round(whole1, 3)}}
    return out, md5_one, md5_all


def profile_traffic(kernel, profile="recon_1080p_moving_rects.md"):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 summary
    under profiles/: the mean FETCH_SIZE and WRITE_SIZE per dispatch (KiB),
    each scaled by what the counter reads for a plain device copy of known
    size in the same file (its calibration tables).  None when absent."""
    import re
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    text = open(path).read()
    tables = {}
    for m in re.finditer(r"== \S*/(\w+)_(FETCH_SIZE|WRITE_SIZE)/\w+_results\.db\n(.*?)(?=\n==|\n```)", text, re.S):
        tables[(m.group(1), m.group(2))] = m.group(3)

    def mean(kind, counter, name):
        body = tables.get((kind, counter))
        if body is None:
            return None
        # the counter table only (a section may also hold a kernel trace
        # table, whose lines start with the same names)
        in_pmc = False
        for line in body.splitlines():
            if line.startswith("pmc"):
                in_pmc = True
                continue
            if not line.startswith(" "):
                in_pmc = False
            line = line.strip()
            if in_pmc and line.startswith(name) and line[len(name):][:1].isspace():
                return float(line.split()[-1])
        return None
    copy = re.search(r"a (\d+) MiB device-to-device copy", text)
    if not copy:
        return None
    copy_bytes = int(copy.group(1)) << 20
    out = 0.0
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = mean("cal", counter, "__amd_rocclr_copyBuffer")
        v = mean("pmc", counter, kernel)
        if cal is None or v is None:
            return None
        out += v * (copy_bytes / (cal * 1024.0)) * 1024.0
    return round(out)


def config1_bench(vp, check, passes=5):
    """The 352x288 8-bit 4:2:0 sinusoid stream of 30 frames from seed 1
    (base_q_idx 120, golden refresh every 8 frames, one tile column): the
    scalar CPU decoder (product parser + oracle, one thread, the
    VP9_FORCE_SCALAR=1 equivalent) over every frame, best of `passes`, with
    its per-frame MD5s (the oracle of this configuration); the same stream
    through the GPU decoder API, whose MD5s must equal them."""
    import streamgen
    import oracle_lib as ol
    frames = streamgen.generate(**streamgen.CONFIG1)
    best, md5s = None, None
    for _ in range(passes):
        d = ol.CpuDecoder(1)
        t0 = time.perf_counter()
        got = [o for f in frames for o in d.decode(f)]
        dt = time.perf_counter() - t0
        md5s = [ol.md5(o.packed()) for o in got]
        best = dt if best is None else min(best, dt)
    dec = vp.Decoder(0)
    best_gpu = None
    for p in range(passes + 1):  # one decoder; the first pass untimed
        t0 = time.perf_counter()
        gpu_frames = [fr for f in frames for fr in dec.decode_all(f)]
        dt = time.perf_counter() - t0
        if p:
            best_gpu = dt if best_gpu is None else min(best_gpu, dt)
    gpu_md5 = [ol.md5(fr.packed()) for fr in gpu_frames]
    out = {"workload": "352x288 8-bit 4:2:0 profile 0 sinusoid grid (periods 17 and 29 px) moving (+1.375, -0.625) "
                       "px/frame with noise, 30 frames, seed 1, base_q_idx 120, golden every 8, 1 tile column",
           "cpu_frames_per_s": round(len(frames) / best, 2), "cpu_threads": 1,
           "cpu_note": "whole decode (parse, scalar reconstruction, loop filter, packed output), best of "
                       f"{passes} passes",
           "frame_md5": md5s,
           "gpu_decoder_api_frames_per_s": round(len(frames) / best_gpu, 2),
           "gpu_md5_equal": gpu_md5 == md5s}
    _log(f"bench: config 1: CPU {out['cpu_frames_per_s']} frames/s (1 thread); GPU decoder API "
         f"{out['gpu_decoder_api_frames_per_s']} frames/s; MD5s equal {out['gpu_md5_equal']}")
    return out


def end_to_end(vp, frames_bytes, want_md5, cpu_whole_all):
    """Single-stream 1080p decode from packets to Frame bytes on the host:
    (a) the decoder API, one frame at a time (parse with tile columns in
    parallel on the host cores, upload, reconstruction, loop filter, packed
    download); (b) the server with this one stream (parse workers run ahead
    of the GPU, downloads overlap the next frames).  Best of 3 passes each;
    MD5s against the oracle's."""
    n = len(frames_bytes)
    # one decoder and one server for every pass (the stream restarts at its
    # key frame), the first pass untimed: context set-up and first
    # allocations are not per-frame work
    best_api = None
    dec = vp.Decoder(0)
    dec.set_threads(_host_cores())
    for p in range(4):
        t0 = time.perf_counter()
        got = [fr for f in frames_bytes for fr in dec.decode_all(f)]
        dt = time.perf_counter() - t0
        if p:
            best_api = dt if best_api is None else min(best_api, dt)
    api_ok = [hashlib.md5(fr.packed()).hexdigest() for fr in got] == want_md5
    del dec
    best_srv, srv_ok = None, True
    srv = vp.Server(0, parse_threads=4)
    for p in range(4):
        sid = srv.open(1920, 1080)
        got_srv = []
        t0 = time.perf_counter()
        for f in frames_bytes:
            srv.submit(sid, f)
        srv.close(sid)
        while True:
            fr = srv.poll(10000)
            if fr is None:
                raise RuntimeError("server stalled")
            if fr.eos:
                break
            # the frame copied out of the server's pinned buffer (as the
            # decoder API's Frame is); its MD5 after the timed region
            got_srv.append(bytes(fr.view()))
            fr.release()
        dt = time.perf_counter() - t0
        srv_ok = srv_ok and [hashlib.md5(b).hexdigest() for b in got_srv] == want_md5
        if p:
            best_srv = dt if best_srv is None else min(best_srv, dt)
    srv.close_server()
    out = {"decoder_api_fps": round(n / best_api, 2), "decoder_api_threads": _host_cores(),
           "server_one_stream_fps": round(n / best_srv, 2), "bit_exact": api_ok and srv_ok,
           "cpu_all_core_whole_decode_fps": cpu_whole_all,
           "scope": "packets in, parse, record upload, reconstruction, loop filter, packed Frame in host memory"}
    if cpu_whole_all:
        out["vs_cpu_all_core"] = round(max(out["decoder_api_fps"], out["server_one_stream_fps"]) / cpu_whole_all, 2)
    _log(f"bench: end to end: decoder API {out['decoder_api_fps']} frames/s, server (one stream) "
         f"{out['server_one_stream_fps']} frames/s, bit-exact {out['bit_exact']}")
    return out


# ---- the 4K 10-bit stream with scaled references

def uhd_bench(vp, gpu, check, passes=3, graphs_too=True):
    """The 3840x2160 10-bit 4:2:0 profile-2 stream of 48 frames from seed 4
    (moving rectangles at 10-bit range, 4 tile columns, q 80; every 12th
    frame coded at 1920x1080, so its neighbours predict through a 2:1 or 1:2
    scaled reference): frames/s of vp9r_gpu_run from resident records over
    `passes` passes of the stream, and its shown frames' MD5s against the
    oracle's decode (check)."""
    if TESTS not in sys.path:
        sys.path.insert(0, TESTS)
    import streamgen
    t0 = time.perf_counter()
    frames_bytes = streamgen.generate(**streamgen.CONFIG4)
    parser = vp.Parser()
    frames = []
    for data in frames_bytes:
        rec = parser.parse(data)
        hdr = vp.FrameHdr()
        ctypes.memmove(ctypes.byref(hdr), ctypes.byref(rec.hdr), ctypes.sizeof(vp.FrameHdr))
        frames.append({"hdr": hdr, "prep": None if hdr.show_existing_frame else gpu.prepare(rec)})
    n = len(frames)
    _log(f"bench: 4K 10-bit stream: {n} frames generated and prepared in {time.perf_counter() - t0:.1f}s")
    sched = Schedule(vp, gpu, frames)
    md5 = []
    for i in range(n):
        shown = sched.step(i)
        if shown is not None:
            md5.append(hashlib.md5(shown.packed()).hexdigest())
    for i in range(n):
        sched.step(i)
    gpu.synchronize()
    t0 = time.perf_counter()
    for k in range(passes * n):
        sched.step(k % n)
    gpu.synchronize()
    fps = passes * n / (time.perf_counter() - t0)
    # the same passes with each frame's launches replayed as a HIP graph
    # (captured in an untimed pass; graphs re-targeted when the stores of a
    # frame differ), and the MD5s of a graph pass
    graphs = {}
    gfps, gmd5, nodes = None, md5, 0
    if graphs_too:
        for i in range(n):
            sched.step(i, graphs)
        gpu.synchronize()
        t0 = time.perf_counter()
        for k in range(passes * n):
            sched.step(k % n, graphs)
        gpu.synchronize()
        gfps = passes * n / (time.perf_counter() - t0)
        gmd5 = []
        for i in range(n):
            shown = sched.step(i, graphs)
            if shown is not None:
                gmd5.append(hashlib.md5(shown.packed()).hexdigest())
        nodes = sum(g[0].nodes() for g in graphs.values()) / max(len(graphs), 1)
    del graphs
    out = {"workload": "3840x2160 10-bit 4:2:0 profile 2, 48 frames, seed 4, every 12th frame at 1920x1080 "
                       "(2:1 / 1:2 scaled references), 4 tile columns, base_q_idx 80 (stream generator)",
           "frames_per_s": round(fps, 2), "ms_per_frame": round(1000.0 / fps, 4), "timed_frames": passes * n,
           "launch": "direct (vp9r_gpu_run)",
           "graph": None if gfps is None else {
               "frames_per_s": round(gfps, 2), "ms_per_frame": round(1000.0 / gfps, 4),
               "kernel_nodes_per_frame": round(nodes, 1), "md5_equal_direct": gmd5 == md5,
               "launch": "one HIP graph per frame (vp9r_gpu_graph_run), captured once"},
           "bit_exact": None}
    if check:
        cpu, md5_one, md5_all = cpu_baselines(frames_bytes, 0.0, "the 4K 10-bit stream")
        out["bit_exact"] = md5_one == md5 and md5_all == md5 and gmd5 == md5
        out["cpu_baseline"] = cpu
        out["vs_cpu_all_core"] = round(fps / cpu["value"], 2)
    _log(f"bench: 4K 10-bit {fps:.1f} frames/s direct, {gfps} frames/s through graphs, "
         f"bit-exact {out['bit_exact']}")
    del sched, frames
    return out


# ---- kernel microbenchmark: transforms and motion compensation, no bitstream

def _micro_cpu(kind, oracle_fn, obj, sample, threads):
    """The oracle's batched form on one core over the first `sample` blocks
    (blocks/s), then its result over every block on `threads` cores."""
    t0 = time.perf_counter()
    oracle_fn(obj, 1, sample)
    spent = time.perf_counter() - t0
    cpu = {"value": round(sample / spent, 1), "unit": "blocks/s", "cores": 1, "kind": "port",
           "sample": f"oracle {kind} batch over the first {sample} blocks, one core, {spent:.1f}s"}
    return cpu, oracle_fn(obj, threads)


def micro_bench(vp, gpu, check, reps=5, tune=True):
    """Kernel microbenchmark (tests/microbench.py): 2M inverse-transform
    blocks and 1M motion-compensation blocks from seed 3, 8-bit, each batch
    one vp9r_itx_add_batch / vp9r_mc_batch call with its inputs resident in
    HBM; the kernels' device time from events around their launches
    (vp9r_gpu_seam_ms), median of `reps` calls after one untimed call, as
    algorithmic GB/s against the 8 TB/s HBM peak.  With check, the first
    call's output is compared with the oracle's over every block, and the
    oracle's batched forms on one core over a sample are the CPU baseline."""
    import numpy as np
    if TESTS not in sys.path:
        sys.path.insert(0, TESTS)
    import microbench as mb
    lib = vp.library()
    threads = _host_cores()
    out = {"workload": "2M inverse-transform blocks (sizes 4/8/16/32 at 40/30/20/10%, Laplace(0,40) coefficients, "
                       "85% zeros, eob U[1,n^2]) and 1M motion-compensation blocks (VP9 shapes 8..64, MVs +-64 px, "
                       "30% compound, four 1920x1080 references), seed 3, 8-bit", "bit_exact": None}

    def roof(nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(gbs / 8000.0, 4), "algorithmic_bytes": nbytes}

    # the launch-shape autotuner first: its choices apply to the calls below
    if tune:
        t0 = time.perf_counter()
        choice, times = gpu.autotune()
        names = {f"{k}x{k}": None for k in (4, 8, 16, 32)}
        itx_us = {}
        for c, cls in enumerate(("8bit", "hbd")):
            for k, size in enumerate(names):
                row = times[(c * 4 + k) * 12:(c * 4 + k + 1) * 12]
                itx_us[f"{cls} {size}"] = {vp.Gpu.shape_name(
                    (1, 2, 4)[j % 3] | (vp.SHAPE_REG if (j // 3) % 2 else 0) | (vp.SHAPE_PERSISTENT if j // 6 else 0)):
                    round(t, 1) for j, t in enumerate(row) if t >= 0}
        out["autotune"] = {"choices": choice, "seconds": round(time.perf_counter() - t0, 2),
                           "chosen_itx_shapes": {"8bit": [vp.Gpu.shape_name(c) for c in choice["itx_shape_8bit"]],
                                                 "hbd": [vp.Gpu.shape_name(c) for c in choice["itx_shape_hbd"]]},
                           "itx_us_by_class_size_shape": itx_us,
                           "mc_us_by_class_rows_32_16": [round(t, 1) for t in times[96:100]]}
        _log(f"bench: autotune {choice}")

    t0 = time.perf_counter()
    s = mb.ItxSet(2_000_000, 3)
    bufs = [gpu.upload(a) for a in (s.coefs, s.coef_off, s.tx_type, s.eob, s.dst_off)]
    d_plane = gpu.upload(s.plane.astype(np.uint8))
    _log(f"bench: microbench transforms generated and uploaded in {time.perf_counter() - t0:.1f}s")
    times, first = [], None
    for r in range(reps + 1):
        vp._check(lib.vp9r_itx_add_batch(gpu.handle, s.count, bufs[0].ptr, bufs[1].ptr, mb.ptr(s.tx), bufs[2].ptr,
                                         bufs[3].ptr, bufs[4].ptr, d_plane.ptr, mb.PLANE_STRIDE, 8, 0))
        if r == 0:
            first = d_plane.read(np.uint8, s.rows * mb.PLANE_STRIDE) if check else None
        else:
            times.append(gpu.seam_ms())
    ms = sorted(times)[len(times) // 2]
    itx = {"blocks": s.count, "kernel_ms": round(ms, 4), "blocks_per_s": round(s.count / (ms * 1e-3), 1),
           "roofline": roof(s.bytes(), ms), "bit_exact": None}
    del bufs, d_plane
    if check:
        itx["cpu_baseline"], want = _micro_cpu("inverse transform", mb.oracle_itx, s, 300_000, threads)
        itx["bit_exact"] = bool(np.array_equal(first.reshape(want.shape), want))
        del want
    _log(f"bench: microbench itx {ms:.3f} ms ({itx['roofline']['achieved']} GB/s), bit-exact {itx['bit_exact']}")
    del s, first

    t0 = time.perf_counter()
    m = mb.McSet(1_000_000, 3)
    d_refs = [gpu.upload(r.astype(np.uint8)) for r in m.refs]
    ref_ptrs = (ctypes.c_void_p * mb.N_REFS)(*[d.ptr.value for d in d_refs])
    d_out = gpu.upload(np.zeros(m.rows * mb.PLANE_STRIDE, dtype=np.uint8))
    _log(f"bench: microbench MC generated and uploaded in {time.perf_counter() - t0:.1f}s")
    times = []
    for r in range(reps + 1):
        vp._check(lib.vp9r_mc_batch(gpu.handle, m.count, mb.ptr(m.desc), ref_ptrs, mb.N_REFS, mb.REF_W, mb.REF_W - 1,
                                    mb.REF_H - 1, d_out.ptr, mb.PLANE_STRIDE, 8))
        if r:
            times.append(gpu.seam_ms())
    ms = sorted(times)[len(times) // 2]
    mc = {"blocks": m.count, "kernel_ms": round(ms, 4), "blocks_per_s": round(m.count / (ms * 1e-3), 1),
          "roofline": roof(m.bytes(), ms), "bit_exact": None,
          "note": "algorithmic bytes count every block's filter footprint; the four 2 MB reference planes "
                  "stay in cache, so HBM traffic is mostly the output"}
    if check:
        got = d_out.read(np.uint8, m.rows * mb.PLANE_STRIDE).reshape(m.rows, mb.PLANE_STRIDE)
        mc["cpu_baseline"], want = _micro_cpu("motion compensation", mb.oracle_mc, m, 60_000, threads)
        mc["bit_exact"] = bool(np.array_equal(got, want))
        del got, want
    _log(f"bench: microbench MC {ms:.3f} ms ({mc['roofline']['achieved']} GB/s), bit-exact {mc['bit_exact']}")
    del d_refs, d_out, m
    out["itx"], out["mc"] = itx, mc
    if check:
        out["bit_exact"] = itx["bit_exact"] and mc["bit_exact"]
    return out


# ---- loop filter microbenchmark: random levels and sharpness, 64 frames a launch

def lf_bench(vp, gpu, check, copies=64, reps=5):
    """64 one-frame 1920x1080 8-bit 4:2:0 streams (key frames of moving
    rectangles, 4 tile columns, q 100, seeds 300..363, so 64 partitions),
    each with its own loop filter level U[0, 63] and sharpness U[0, 7] (seed
    3).  Each of `reps` + 1 rounds reconstructs all 64 with the filter off
    (vp9r_gpu_run_batch of the records at level 0), then filters them in one
    batch (vp9r_gpu_filter_batch: every launch covers all 64) timed by events
    around its launches; median of the timed rounds, against 2 x the visible
    frame bytes per frame.  With check, every filtered frame's MD5 (first
    round) against the oracle's decode of its stream."""
    import concurrent.futures
    import multiprocessing
    import numpy as np
    import streamgen
    rng = np.random.default_rng(3)
    levels, sharp = rng.integers(0, 64, size=copies), rng.integers(0, 8, size=copies)
    cfgs = [dict(streamgen.CONFIG2, frames=1, seed=300 + i, lf_level=int(levels[i]), lf_sharpness=int(sharp[i]))
            for i in range(copies)]
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(_host_cores()) as ex:
        streams = list(ex.map(lambda c: streamgen.generate(**c), cfgs))
    p_lf, p_off, curs = [], [], []
    for fr in streams:
        parser = vp.Parser()
        rec = parser.parse(fr[0])
        p_lf.append(gpu.prepare(rec))
        rec.hdr.lf_level = 0
        p_off.append(gpu.prepare(rec))
        curs.append(gpu.frame(1920, 1080, 8, 1, 1))
    _log(f"bench: loop filter microbench: {copies} key frames generated and prepared in "
         f"{time.perf_counter() - t0:.1f}s")
    times, md5 = [], None
    for r in range(reps + 1):
        gpu.run_batch([(p, None, c) for p, c in zip(p_off, curs)])
        gpu.synchronize()
        ms = gpu.filter_batch(list(zip(p_lf, curs)))
        if r == 0:
            md5 = [hashlib.md5(c.packed()).hexdigest() for c in curs]
        else:
            times.append(ms)
    ms = sorted(times)[len(times) // 2]
    nbytes = sum(p.stats()["lf_bytes"] for p in p_lf)
    gbs = nbytes / (ms * 1e-3) / 1e9
    out = {"workload": f"{copies} 1920x1080 8-bit 4:2:0 key frames (seeds 300..{299 + copies}), filter level "
                       f"U[0,63] and sharpness U[0,7] per frame (seed 3), one vp9r_gpu_filter_batch",
           "frames": copies, "kernel_ms": round(ms, 4), "frames_per_s": round(copies / (ms * 1e-3), 1),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": nbytes},
           "sharpness_levels": sorted({int(v) for v in sharp}), "bit_exact": None,
           "parity_unpinned": ["lf sharpness 1-7: the GPU equals the oracle's restatement of loopfilter.rs:143-160; "
                               "no fixture pins it"]}
    if check:
        ctx = multiprocessing.get_context("spawn")
        with concurrent.futures.ProcessPoolExecutor(_host_cores(), mp_context=ctx) as ex:
            want = [m[0][0] for m in ex.map(_oracle_worker, streams)]
        out["bit_exact"] = want == md5
    _log(f"bench: loop filter microbench {ms:.3f} ms for {copies} frames ({gbs:.1f} GB/s), "
         f"bit-exact {out['bit_exact']}")
    del curs, p_lf, p_off
    return out


# ---- batched stages: the kernels at full occupancy

def batched_stages(vp, gpu, frames, copies=64, reps=5):
    """One inter frame of the stream reconstructed and loop filtered `copies`
    times in one batch (vp9r_gpu_run_batch: every launch covers all copies),
    `reps` batches timed per stage with HIP events: each stage's kernels with
    enough work to fill the GPU.  Every copy must equal the frame decoded
    alone."""
    n = len(frames)
    k = next(i for i in range(n // 3, n) if not frames[i]["hdr"].show_existing_frame
             and not frames[i]["hdr"].frame_is_intra)
    sched = Schedule(vp, gpu, frames)
    for i in range(k):
        sched.step(i)
    h, prep = frames[k]["hdr"], frames[k]["prep"]
    refs = [sched.slots[h.ref_frame_idx[j]] for j in range(3)]
    alone = gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
    gpu.run(prep, refs, alone)
    want = hashlib.md5(alone.packed()).hexdigest()
    curs = [gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y) for _ in range(copies)]
    jobs = [(prep, refs, c) for c in curs]
    gpu.run_batch(jobs)
    gpu.synchronize()
    exact = all(hashlib.md5(curs[i].packed()).hexdigest() == want for i in (0, copies // 2, copies - 1))
    gpu.set_timing(True)
    gpu.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        gpu.run_batch(jobs)
    gpu.synchronize()
    wall = time.perf_counter() - t0
    ms, nfr = gpu.stage_times()
    gpu.set_timing(False)
    st = prep.stats()
    out = {"frame_index": k, "copies_per_batch": copies, "batches": reps, "fps": round(nfr / wall, 1),
           "bit_exact": exact, "stages": {}}
    for s, name in enumerate(STAGES):
        b = st[STAGE_BYTES[s]] * nfr
        gbs = b / (ms[s] * 1e-3) / 1e9 if ms[s] > 0 else None
        out["stages"][name] = {"kernel": STAGE_KERNELS[s], "ms_per_frame": round(ms[s] / nfr, 4),
                               "GB_per_s": round(gbs, 1) if gbs else None,
                               "hbm_frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    _log(f"bench: batched x{copies}: {out}")
    return out


# ---- serving workload

def _gen_worker(cfg):
    if TESTS not in sys.path:
        sys.path.insert(0, TESTS)
    import streamgen
    return streamgen.generate(**cfg)


def _oracle_worker(frames):
    """The oracle's decode of one stream (product parser, scalar
    reconstruction and loop filter): its shown frames' MD5s and the time."""
    _package()
    if TESTS not in sys.path:
        sys.path.insert(0, TESTS)
    import oracle_lib
    t0 = time.perf_counter()
    dec = oracle_lib.CpuDecoder()
    md5s = [oracle_lib.md5(o.packed()) for f in frames for o in dec.decode(f)]
    return md5s, time.perf_counter() - t0


def _host_cores():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


def serve_run(vp, plan, streams, paced, parse_threads, fps=30.0):
    """Runs the streams through one server: opens each at its arrival (paced)
    or all at once, submits its packets at `fps` from its arrival (paced) or
    at once, closes it after its last packet.  Returns per-stream MD5 lists,
    per-stream status, per-frame latencies, wall time and server stats."""
    import concurrent.futures
    import threading
    srv = vp.Server(0, parse_threads=parse_threads)
    events = []
    for i, ((arrival, _), frames) in enumerate(zip(plan, streams)):
        t0 = arrival if paced else 0.0
        events.append((t0, 0, i, -1))
        for j in range(len(frames)):
            events.append((t0 + (j / fps if paced else 0.0), 1, i, j))
        events.append((t0 + ((len(frames) - 1) / fps if paced else 0.0), 2, i, -1))
    events.sort()
    sid_of = {}
    idx_of = {}
    start = time.perf_counter()
    err = []

    def submitter():
        try:
            for t, kind, i, j in events:
                d = start + t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                if kind == 0:
                    cfg = plan[i][1]
                    sid = srv.open(cfg["width"], cfg["height"], cfg.get("bit_depth", 8), cfg.get("ss_x", 1),
                                   cfg.get("ss_y", 1))
                    sid_of[i] = sid
                    idx_of[sid] = i
                elif kind == 1:
                    srv.submit(sid_of[i], streams[i][j])
                else:
                    srv.close(sid_of[i])
        except Exception as e:  # reported by the caller
            err.append(e)

    def hash_release(fr):
        d = hashlib.md5(fr.view()).hexdigest()
        fr.release()
        return fr.stream, fr.index, d

    th = threading.Thread(target=submitter)
    th.start()
    pool = concurrent.futures.ThreadPoolExecutor(8)
    futs, lat, status, lat_at = [], [], {}, []
    last = start
    while len(status) < len(streams):
        if err:
            break
        fr = srv.poll(500)
        if fr is None:
            if time.perf_counter() - start > 600:
                raise RuntimeError(f"serving did not finish: {srv.stats()}")
            continue
        if fr.eos:
            status[fr.stream] = fr.status
            continue
        last = time.perf_counter()
        lat.append(fr.latency_ms)
        lat_at.append((fr.stream, fr.index, fr.latency_ms))
        futs.append(pool.submit(hash_release, fr))
    th.join()
    if err:
        raise err[0]
    md5 = [dict() for _ in streams]
    for f in futs:
        sid, index, d = f.result()
        md5[idx_of[sid]][index] = d
    pool.shutdown()
    stats = srv.stats()
    srv.close_server()
    md5 = [[m[k] for k in sorted(m)] for m in md5]
    st = [status.get(sid_of[i]) for i in range(len(streams))]
    serve_run.latency_at = [(idx_of[sid], j, ms) for sid, j, ms in lat_at]
    return md5, st, lat, last - start, stats


def _pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(q * len(v)))] if v else None


def serving_bench(vp, n_streams, parse_threads):
    import concurrent.futures
    import multiprocessing
    import streamgen
    plan = streamgen.config5(seed=5, streams=n_streams)
    cores = _host_cores()
    ctx = multiprocessing.get_context("spawn")
    t0 = time.perf_counter()
    with concurrent.futures.ProcessPoolExecutor(cores, mp_context=ctx) as ex:
        streams = list(ex.map(_gen_worker, [cfg for _, cfg in plan]))
    total = sum(len(s) for s in streams)
    _log(f"bench: serving workload: {len(streams)} streams, {total} frames generated in "
         f"{time.perf_counter() - t0:.1f}s on {cores} processes")
    # CPU baseline and expected MD5s: the oracle, one decoder per stream, all cores
    t0 = time.perf_counter()
    with concurrent.futures.ProcessPoolExecutor(cores, mp_context=ctx) as ex:
        # longest streams first, so the pool is not left waiting on one
        order = sorted(range(len(streams)), key=lambda i: -len(streams[i]) * plan[i][1]["width"])
        res = dict(zip(order, ex.map(_oracle_worker, [streams[i] for i in order])))
    cpu_wall = time.perf_counter() - t0
    expected = [res[i][0] for i in range(len(streams))]
    shown_total = sum(len(e) for e in expected)
    cpu_fps = shown_total / cpu_wall
    _log(f"bench: serving CPU baseline (oracle, {cores} processes): {cpu_fps:.1f} frames/s")
    out = {"workload": f"{len(streams)} streams, seed 5, Poisson arrivals 4/s; 50% 640x360, 30% 1280x720, "
                       f"20% 1920x1080; 70% 8-bit 4:2:0, 20% 10-bit 4:2:0, 10% 12-bit 4:4:4; 90-300 frames",
           "streams": len(streams), "frames": shown_total,
           "cpu_baseline": {"value": round(cpu_fps, 2), "unit": "frames/s", "cores": cores, "kind": "port",
                            "sample": "oracle decode (product parser + scalar reconstruction and loop filter) of "
                                      "every stream, one process per stream"}}
    exact = True
    # paced: arrivals as planned; all_live: every stream arriving at once and
    # paced at 30 frames/s, so that all of them are live together until the
    # shortest ends; saturated: every packet at once (capacity)
    plan_all = [(0.0, cfg) for _, cfg in plan]
    for mode, paced, pl in (("paced", True, plan), ("paced_all_live", True, plan_all),
                            ("saturated", False, plan)):
        md5, st, lat, wall, stats = serve_run(vp, pl, streams, paced, parse_threads)
        ok = all(s == vp.OK for s in st) and md5 == expected
        exact &= ok
        conc = 0
        if paced:
            iv = [(a, a + (len(s) - 1) / 30.0) for (a, _), s in zip(pl, streams)]
            conc = max(sum(1 for a, b in iv if a <= t <= b) for t, _ in iv)
        out[mode] = {"aggregate_fps": round(len(lat) / wall, 2),
                     "latency_ms": {"p50": round(_pct(lat, 0.5), 3), "p99": round(_pct(lat, 0.99), 3),
                                    "max": round(max(lat), 3)},
                     "frames": len(lat), "wall_s": round(wall, 3), "batches": stats["batches"],
                     "mean_batch": round(stats["frames_decoded"] / max(stats["batches"], 1), 2),
                     "max_batch": stats["max_batch"], "bit_exact": ok}
        if paced:
            out[mode]["stream_fps"] = 30
            out[mode]["max_concurrent_streams"] = conc
        if mode == "paced_all_live":
            # the frames decoded while every stream was live (each stream's
            # first min-length frames)
            n_live = min(len(x) for x in streams)
            live = [ms for i, j, ms in serve_run.latency_at if j < n_live]
            out[mode]["all_live_window"] = {"streams": len(streams), "frames_per_stream": n_live,
                                            "seconds": round(n_live / 30.0, 2),
                                            "latency_ms": {"p50": round(_pct(live, 0.5), 3),
                                                           "p99": round(_pct(live, 0.99), 3),
                                                           "max": round(max(live), 3)}}
        _log(f"bench: serving {mode}: {out[mode]}")
    out["bit_exact"] = exact
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600, help="timed frames")
    ap.add_argument("--warmup", type=int, default=60, help="untimed frames before the timed region")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (and the MD5 check)")
    ap.add_argument("--no-extras", action="store_true", help="skip the PCIe and latency passes")
    ap.add_argument("--no-serve", action="store_true", help="skip the multi-stream serving workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-stage measurement")
    ap.add_argument("--no-uhd", action="store_true", help="skip the 4K 10-bit scaled-reference stream")
    ap.add_argument("--no-micro", action="store_true", help="skip the transform / MC kernel microbenchmark")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP graph passes")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner (default shapes)")
    ap.add_argument("--micro-only", action="store_true",
                    help="only run the kernel microbenchmark and print its JSON (a profiling aid)")
    ap.add_argument("--serve-streams", type=int, default=32, help="streams of the serving workload")
    ap.add_argument("--parse-threads", type=int, default=12, help="parse workers of the server")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the reference reading of the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty(n, dtype=torch.uint8, device="cuda")
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 5}), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: the single-stream reconstruction does not shard across GPUs "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    vp = _package()
    vp.library()
    if TESTS not in sys.path:
        sys.path.insert(0, TESTS)
    if args.micro_only:
        g = vp.Gpu(0)
        micro = micro_bench(vp, g, not args.no_cpu, tune=not args.no_tune)
        micro["lf"] = lf_bench(vp, g, not args.no_cpu)
        if micro["lf"]["bit_exact"] is False:
            micro["bit_exact"] = False
        print(json.dumps({"microbench": micro}), flush=True)
        sys.exit(1 if micro["bit_exact"] is False else 0)
    import streamgen

    cfg = dict(streamgen.CONFIG2)
    t0 = time.perf_counter()
    frames_bytes = streamgen.generate(**cfg)
    _log(f"bench: generated {len(frames_bytes)} frames ({sum(map(len, frames_bytes))} bytes) "
         f"in {time.perf_counter() - t0:.2f}s")

    gpu = vp.Gpu(0)
    parser = vp.Parser()
    frames = []
    t0 = time.perf_counter()
    t_parse = 0.0
    for data in frames_bytes:
        tp = time.perf_counter()
        rec = parser.parse(data)
        t_parse += time.perf_counter() - tp
        hdr = vp.FrameHdr()
        ctypes.memmove(ctypes.byref(hdr), ctypes.byref(rec.hdr), ctypes.sizeof(vp.FrameHdr))
        prep = None if hdr.show_existing_frame else gpu.prepare(rec)
        frames.append({"hdr": hdr, "prep": prep, "stats": prep.stats() if prep else None})
    n = len(frames)
    _log(f"bench: parsed and prepared {n} frames in {time.perf_counter() - t0:.2f}s "
         f"(parse {n / t_parse:.1f} frames/s on one core)")

    sched = Schedule(vp, gpu, frames)

    # bit-exactness pass: every shown frame's MD5
    gpu_md5 = []
    for i in range(n):
        shown = sched.step(i)
        if shown is not None:
            gpu_md5.append(hashlib.md5(shown.packed()).hexdigest())
    gpu.synchronize()
    _log(f"bench: verification pass done ({len(gpu_md5)} shown frames)")

    # warmup, then the timed region (HIP events per stage on the kernel stream)
    pos = 0
    for _ in range(args.warmup):
        sched.step(pos % n)
        pos += 1
    gpu.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sched.step(pos % n)
        pos += 1
    gpu.synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps / elapsed
    _log(f"bench: {args.steps} frames in {elapsed:.3f}s = {fps:.1f} frames/s")

    # the same frames through one HIP graph per frame (captured in an
    # untimed pass), reported beside the direct launches
    graph_fps = None
    if not args.no_graph:
        graphs = {}
        for i in range(n):
            sched.step(i, graphs)
        gpu.synchronize()
        t0 = time.perf_counter()
        for j in range(args.steps):
            sched.step(j % n, graphs)
        gpu.synchronize()
        graph_fps = args.steps / (time.perf_counter() - t0)
        pos = args.steps  # the next frame of the chain
        del graphs
        _log(f"bench: through graphs {graph_fps:.1f} frames/s")

    # the stage breakdown: the same number of frames again, with HIP events
    # between the stages on the kernel stream (each event record holds the
    # stream for a few microseconds, so the headline region above has none)
    gpu.set_timing(True)
    gpu.synchronize()
    t0 = time.perf_counter()
    timed_idx = []
    for _ in range(args.steps):
        timed_idx.append(pos % n)
        sched.step(pos % n)
        pos += 1
    gpu.synchronize()
    staged_fps = args.steps / (time.perf_counter() - t0)
    stage_ms, timed_frames = gpu.stage_times()
    gpu.set_timing(False)
    _log(f"bench: stage pass {staged_fps:.1f} frames/s")

    # per-stage totals over the timed frames
    # (the loop filter's bytes split between the two by their share of its
    # levels)
    work = {k: 0 for k in STAGE_BYTES}
    launches = [0, 0, 0, 0]
    for i in timed_idx:
        st = frames[i]["stats"]
        if st is None:
            continue
        tail = st["lf_tail_levels"] / st["lf_levels"] if st["lf_levels"] else 0.0
        work["mc_bytes"] += st["mc_bytes"]
        work["itx_bytes"] += st["itx_bytes"]
        work["intra_bytes"] += st["intra_bytes"] + (1.0 - tail) * st["lf_bytes"]
        work["lf_bytes"] += tail * st["lf_bytes"]
        launches[0] += 1 if st["mc_tiles"] else 0
        launches[1] += 1
        launches[2] += st["wave_launches"] + (1 if st["lf_levels"] else 0)
        launches[3] += st["lf_tail_launches"]
    stages = {}
    for s, name in enumerate(STAGES):
        ms = stage_ms[s]
        b = work[STAGE_BYTES[s]]
        stages[name] = {
            "kernel": STAGE_KERNELS[s],
            "ms_per_frame": ms / max(timed_frames, 1),
            "launches_per_frame": launches[s] / max(timed_frames, 1),
            "avg_launch_us": 1000.0 * ms / max(launches[s], 1),
            "bytes_per_frame": b / max(timed_frames, 1),
            "GB_per_s": (b / (ms * 1e-3) / 1e9) if ms > 0 else None,
        }
    dom = max(range(4), key=lambda s: stage_ms[s])
    dname = STAGES[dom]
    achieved = stages[dname]["GB_per_s"] or 0.0
    # the filter tail's deepest cone (five levels at 8 bits), else four
    kern = {"loop_filter": "vp9g::k_lf_cone<false, 5>", "inter_pred": "vp9g::k_prep_mc<false, 1>",
            "inter_residual": "vp9g::k_itx_inter<false>"}.get(dname)
    traffic = profile_traffic(kern) if kern else None
    if traffic is None and dname == "loop_filter":
        kern = "vp9g::k_lf_cone<false, 4>"
        traffic = profile_traffic(kern)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": STAGE_KERNELS[dom], "stage": dname,
                "traffic_source": "profiles/recon_1080p_moving_rects.md: FETCH_SIZE and WRITE_SIZE per dispatch, "
                                  "scaled by a device copy's counts; HBM bytes per launch"}
    if roofline["traffic"] and dname in stages and stages[dname]["launches_per_frame"]:
        roofline["algorithmic_bytes_per_launch"] = round(stages[dname]["bytes_per_frame"] /
                                                         stages[dname]["launches_per_frame"])

    extras = {}
    if not args.no_extras:
        # PCIe-inclusive: every shown frame also copied to pinned host memory
        size = vp.library().vp9r_packed_size(cfg["width"], cfg["height"], 8, 1, 1)
        hosts = [vp.HostBuffer(size) for _ in range(4)]
        pending = [None] * 4
        k = 0
        gpu.synchronize()
        t0 = time.perf_counter()
        for j in range(args.steps):
            shown = sched.step(pos % n)
            pos += 1
            if shown is not None:
                slot = k % 4
                if pending[slot] is not None:
                    pending[slot].wait()
                shown.packed_async(hosts[slot])
                pending[slot] = shown
                k += 1
        for f in pending:
            if f is not None:
                f.wait()
        gpu.synchronize()
        pcie_fps = args.steps / (time.perf_counter() - t0)
        # synchronous latency: one frame at a time, submit -> frame on the host
        lat = []
        for j in range(min(args.steps, 2 * n)):
            t1 = time.perf_counter()
            shown = sched.step(pos % n)
            pos += 1
            if shown is not None:
                shown.packed_async(hosts[0])
                shown.wait()
            else:
                gpu.synchronize()
            lat.append((time.perf_counter() - t1) * 1000.0)
        lat.sort()
        extras = {"pcie_inclusive_fps": round(pcie_fps, 2),
                  "latency_ms": {"p50": round(lat[len(lat) // 2], 3),
                                 "p99": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3),
                                 "frames": len(lat), "mode": "one frame in flight, incl. D2H"}}
        _log(f"bench: PCIe-inclusive {pcie_fps:.1f} frames/s; latency p50 {extras['latency_ms']['p50']} ms "
             f"p99 {extras['latency_ms']['p99']} ms")

    cpu = None
    bit_exact = None
    e2e = None
    config1 = None
    if not args.no_cpu:
        cpu, cpu_md5, cpu_md5_all = cpu_baselines(frames_bytes, args.cpu_seconds, "the same stream")
        bit_exact = cpu_md5 == gpu_md5 and cpu_md5_all == gpu_md5
        cpu["vs_gpu"] = round(fps / cpu["value"], 2)
        _log(f"bench: CPU oracle {cpu['one_core']['value']} frames/s on one core, {cpu['value']} on "
             f"{cpu['cores']}; GPU MD5s {'==' if bit_exact else '!='} oracle MD5s")
        if not bit_exact:
            bad = [i for i, (a, b) in enumerate(zip(gpu_md5, cpu_md5)) if a != b]
            _log(f"bench: MD5 MISMATCH at shown frames {bad[:10]}")
        if not args.no_extras:
            e2e = end_to_end(vp, frames_bytes, cpu_md5, cpu["whole_decode_fps"])
            config1 = config1_bench(vp, True)

    batched = None if args.no_batched else batched_stages(vp, gpu, frames)
    uhd = None if args.no_uhd else uhd_bench(vp, gpu, not args.no_cpu, graphs_too=not args.no_graph)
    micro = None if args.no_micro else micro_bench(vp, gpu, not args.no_cpu, tune=not args.no_tune)
    if micro is not None:
        micro["lf"] = lf_bench(vp, gpu, not args.no_cpu)
        if micro["lf"]["bit_exact"] is False:
            micro["bit_exact"] = False

    serving = None
    if not args.no_serve and not args.no_cpu:
        serving = serving_bench(vp, args.serve_streams, args.parse_threads)

    out = {
        "metric": "1080p decoded frames/s GPU vs CPU (bit-exact); p99 frame latency @32 streams",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "1920x1080 8-bit 4:2:0 profile 0, 60 frames, seed 2, 64 moving rectangles, "
                               "4 tile columns, base_q_idx 100 (stream generator)",
                   "frames_in_stream": n, "records_resident_in_hbm": True},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bit_exact": bit_exact,
        "stages": stages,
        "stage_pass_fps": round(staged_fps, 2),
        "graph_fps": None if graph_fps is None else round(graph_fps, 2),
        "parse_fps_one_core": round(n / t_parse, 2),
    }
    out.update(extras)
    if e2e is not None:
        out["end_to_end"] = e2e
    if config1 is not None:
        out["config1"] = config1
    # parity that no fixture or closed form pins: the GPU equals the oracle's
    # restatement of the reference there
    out["parity_unpinned"] = ["loop filter sharpness 1-7 (no committed vector has sharpness > 0)",
                              "more than one tile column (no committed vector has two)"]
    if batched is not None:
        out["batched"] = batched
    if uhd is not None:
        out["uhd_10bit"] = uhd
    if micro is not None:
        out["microbench"] = micro
    if serving is not None:
        out["serving"] = serving
    print(json.dumps(out), flush=True)
    if bit_exact is False or (serving is not None and not serving["bit_exact"]) or \
            (e2e is not None and not e2e["bit_exact"]) or (config1 is not None and not config1["gpu_md5_equal"]) or \
            (batched is not None and not batched["bit_exact"]) or (uhd is not None and uhd["bit_exact"] is False) or \
            (micro is not None and micro["bit_exact"] is False):
        sys.exit(1)


if __name__ == "__main__":
    main()
