"""The threaded CPU decoder (the all-core baseline): tile columns parsed and
reconstructed in parallel (decode_tiles_parallel, src/decoder/block.rs:
1769-1866) and the loop filter's superblock rows as a wavefront
(filter_bands, src/decoder/loopfilter.rs:195-265) give the serial decoder's
frames, on generated streams of 1, 2 and 4 tile columns (with several
superblock rows, filter levels and sharpness) and on the committed streams."""
import os

import pytest

import oracle_lib as ol
import streamgen
import webm


def md5s(frames, threads):
    d = ol.CpuDecoder(threads)
    return [ol.md5(o.packed()) for f in frames for o in d.decode(f)]


@pytest.mark.parametrize("cfg", [
    dict(width=640, height=360, frames=4, seed=51, tile_cols_log2=1, content=1, intra_percent=10),
    dict(width=1280, height=720, frames=3, seed=52, tile_cols_log2=2, content=1, intra_percent=10, lf_level=40,
         lf_sharpness=3),
    dict(width=352, height=288, frames=4, seed=53, content=0),
    dict(width=640, height=360, frames=3, seed=54, tile_cols_log2=1, bit_depth=10, content=1),
    dict(width=640, height=360, frames=4, seed=55, tile_cols_log2=1, content=1, intra_percent=10, segmentation=1),
], ids=["640x360-2tiles", "720p-4tiles-sharp3", "352x288-1tile", "640x360-10bit", "640x360-segmentation"])
def test_threaded_decoder_equals_serial(cfg):
    frames = streamgen.generate(**cfg)
    one = md5s(frames, 1)
    assert md5s(frames, 8) == one
    assert md5s(frames, 3) == one


def test_threaded_decoder_on_committed_streams():
    for path in webm.vector_paths():
        exp = webm.expected_md5s(path)
        got = md5s(list(webm.packets(path)), 8)
        assert got == exp, os.path.basename(path)
