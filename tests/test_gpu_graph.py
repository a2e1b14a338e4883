"""HIP graphs of a frame's launches (vp9r_gpu_graph_*): every frame of the
4K 10-bit stream with 2:1 and 1:2 scaled references and of the 1080p
workload, run through a graph captured for its frame stores, equals the
direct launches (vp9r_gpu_run) and the oracle.  The graphs are replayed
over several passes, with direct passes and synchronisations between them,
and every replayed frame is compared; a graph re-targeted to other frame
stores (vp9r_gpu_graph_update) gives the same frames."""
import ctypes

import pytest

import oracle_lib
import rivet_vp9_amd as vp
import streamgen

pytestmark = pytest.mark.gpu


def prepared_stream(gpu, frames):
    p = vp.Parser()
    out = []
    for f in frames:
        rec = p.parse(f)
        hdr = vp.FrameHdr()
        ctypes.memmove(ctypes.byref(hdr), ctypes.byref(rec.hdr), ctypes.sizeof(vp.FrameHdr))
        out.append((hdr, None if hdr.show_existing_frame else gpu.prepare(rec)))
    return out


def run_stream(gpu, prepped, pool, graphs=None, rotate=0):
    """Decodes the prepared frames with the decoder's slot chaining over a
    pool of stores (rotate: start the store search elsewhere, so the stores
    differ from an earlier pass); direct launches, or graphs (made on first
    use, re-targeted when the stores differ).  Returns the shown MD5s."""
    slots = [None] * 8
    shown = []
    for i, (h, prep) in enumerate(prepped):
        if h.show_existing_frame:
            shown.append(oracle_lib.md5(slots[h.frame_to_show_map_idx].packed()))
            continue
        busy = {id(s) for s in slots if s is not None}
        fmt = (h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        cands = [f for f in pool if id(f) not in busy and (f.width, f.height, f.bit_depth, f.ss_x, f.ss_y) == fmt]
        if len(cands) < 2:
            cands.append(gpu.frame(*fmt))
            pool.append(cands[-1])
        cur = cands[rotate % len(cands)]
        refs = [None] * 3 if h.frame_is_intra else [slots[h.ref_frame_idx[k]] for k in range(3)]
        if graphs is None:
            gpu.run(prep, refs, cur)
        else:
            key = (tuple(id(r) for r in refs), id(cur))
            if i not in graphs:
                graphs[i] = [gpu.graph(prep, refs, cur), key]
            elif graphs[i][1] != key:
                graphs[i][0].update(refs, cur)
                graphs[i][1] = key
            graphs[i][0].run()
        for k in range(8):
            if (h.refresh_frame_flags >> k) & 1:
                slots[k] = cur
        if h.show_frame:
            shown.append(oracle_lib.md5(cur.packed()))
    gpu.synchronize()
    return shown


@pytest.mark.parametrize("cfg", [dict(streamgen.CONFIG4, frames=6, resize_every=2), dict(streamgen.CONFIG2, frames=4)],
                         ids=["4k-10bit-scaled", "1080p"])
def test_graph_replays_equal_direct_and_oracle(cfg):
    frames = streamgen.generate(**cfg)
    ref = oracle_lib.CpuDecoder()
    want = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
    gpu = vp.Gpu(0)
    prepped = prepared_stream(gpu, frames)
    pool = []
    assert run_stream(gpu, prepped, pool) == want
    graphs = {}
    for rep in range(3):
        print(f"graph pass {rep}", flush=True)
        assert run_stream(gpu, prepped, pool, graphs, rotate=rep) == want, f"graph pass {rep}"
        assert run_stream(gpu, prepped, pool) == want, f"direct pass after graph pass {rep}"
    assert all(g[0].nodes() > 0 for g in graphs.values())
    del graphs
