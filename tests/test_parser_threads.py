"""Tile columns parsed in parallel (the reference's decode_tiles_parallel,
src/decoder/block.rs:1769-1866, chosen by set_threads,
src/decoder/mod.rs:171-181, 436-486): the record is byte for byte the
serial one, on generated streams of 2 and 4 tile columns (with tile rows
and backward adaptation off, as the generator writes them) and on the
committed streams."""
import ctypes
import os

import pytest

import rivet_vp9_amd as vp
import streamgen
import webm


def record_bytes(rec):
    h = rec.hdr
    parts = [bytes(h)]
    if h.show_existing_frame:
        return b"".join(parts)
    n_mi = h.mi_rows * h.mi_cols
    for ptr, n, size in ((rec.blocks, rec.n_blocks, ctypes.sizeof(vp.Block)),
                         (rec.txbs, rec.n_txbs, ctypes.sizeof(vp.Txb)),
                         (rec.coef_pos, rec.n_coefs, 2), (rec.coef_val, rec.n_coefs, 4),
                         (rec.mi, n_mi, ctypes.sizeof(vp.Mi))):
        parts.append(ctypes.string_at(ptr, n * size) if n else b"")
    return b"".join(parts)


def records(frames, threads):
    p = vp.Parser()
    p.set_threads(threads)
    out = []
    for pkt in frames:
        for f in vp.superframe_split(pkt):
            out.append(record_bytes(p.parse(f)))
    return out


@pytest.mark.parametrize("w,h,tiles,frames,seg", [(640, 360, 1, 5, 0), (1280, 720, 2, 4, 0), (1920, 1080, 2, 3, 0),
                                                  (1280, 720, 2, 4, 1)])
def test_parallel_tile_parse_equals_serial(w, h, tiles, frames, seg):
    s = streamgen.generate(w, h, frames, seed=21, tile_cols_log2=tiles, content=1, intra_percent=10,
                           segmentation=seg)
    one = records(s, 1)
    assert len(one) == frames
    for t in (2, 4, 8):
        assert records(s, t) == one


def test_parallel_parse_of_committed_streams():
    for path in webm.vector_paths():
        pk = list(webm.packets(path))
        assert records(pk, 4) == records(pk, 1), os.path.basename(path)
