"""The stream generator (the benchmark workloads' source): its streams parse
with the product parser and decode with the CPU oracle, it is deterministic
per seed, and it reaches the features the workloads need (tile columns,
scaled references, high bit depth, 4:4:4)."""
import pytest

import oracle_lib as ol
import rivet_vp9_amd as vp
import streamgen


def decode_all(frames):
    dec = ol.CpuDecoder()
    shown, sizes, recs = [], [], []

    def on_record(rec, refs):
        h = rec.hdr
        recs.append((h.width, h.height, h.frame_is_intra, h.tile_cols_log2, h.bit_depth, h.ss_x, h.ss_y,
                     rec.n_blocks, rec.n_txbs, rec.n_coefs,
                     tuple((h.ref_width[i], h.ref_height[i]) for i in range(3))))

    for f in frames:
        for out in dec.decode(f, on_record):
            shown.append(ol.md5(out.packed()))
            sizes.append((out.width, out.height))
    return shown, sizes, recs


def test_deterministic_per_seed():
    a = streamgen.generate(160, 96, 4, seed=11)
    b = streamgen.generate(160, 96, 4, seed=11)
    c = streamgen.generate(160, 96, 4, seed=12)
    assert a == b
    assert a != c


def test_sinusoid_stream_decodes():
    frames = streamgen.generate(352, 288, 6, seed=1, base_q_idx=120, content=0)
    shown, sizes, recs = decode_all(frames)
    assert len(shown) == 6 and set(sizes) == {(352, 288)}
    assert recs[0][2] == 1 and all(r[2] == 0 for r in recs[1:])  # key frame, then inter frames
    assert all(r[7] > 0 for r in recs)


def test_rectangles_with_tile_columns_decode():
    frames = streamgen.generate(640, 360, 4, seed=2, tile_cols_log2=1, content=1, intra_percent=10)
    shown, sizes, recs = decode_all(frames)
    assert len(shown) == 4
    assert all(r[3] == 1 for r in recs)
    assert len(set(shown)) == 4


def test_resized_frames_use_scaled_references():
    frames = streamgen.generate(200, 120, 7, seed=3, resize_every=3)
    shown, sizes, recs = decode_all(frames)
    assert (100, 60) in sizes and (200, 120) in sizes
    scaled = [r for r in recs if not r[2] and any(rs != (r[0], r[1]) for rs in r[10])]
    assert scaled, "no frame predicted through a scaled reference"


@pytest.mark.parametrize("bd,ssx,ssy", [(10, 1, 1), (12, 0, 0), (8, 0, 0), (8, 1, 0), (10, 0, 1)])
def test_formats_decode(bd, ssx, ssy):
    frames = streamgen.generate(160, 96, 3, seed=5, bit_depth=bd, ss_x=ssx, ss_y=ssy)
    shown, sizes, recs = decode_all(frames)
    assert len(shown) == 3
    assert all(r[4:7] == (bd, ssx, ssy) for r in recs)


def test_loop_filter_level_and_sharpness_as_set():
    for level, sharp in ((0, 0), (1, 3), (40, 5), (63, 7)):
        frames = streamgen.generate(160, 96, 2, seed=5, lf_level=level, lf_sharpness=sharp)
        p = vp.Parser()
        for f in frames:
            h = p.parse(f).hdr
            assert (h.lf_level, h.lf_sharpness) == (level, sharp)
    h = vp.Parser().parse(streamgen.generate(160, 96, 1, seed=5, base_q_idx=100)[0]).hdr
    assert (h.lf_level, h.lf_sharpness) == ((100 * 10 + 32) // 64, 0)


def test_benchmark_workload_headers():
    """The 1080p workload's first frames: size, tiles, quantiser and a mix
    of intra and inter blocks."""
    frames = streamgen.generate(**dict(streamgen.CONFIG2, frames=3))
    p = vp.Parser()
    for i, f in enumerate(frames):
        rec = p.parse(f)
        h = rec.hdr
        assert (h.width, h.height, h.bit_depth, h.ss_x, h.ss_y) == (1920, 1080, 8, 1, 1)
        assert h.tile_cols_log2 == 2 and h.base_q_idx == 100
        assert h.frame_is_intra == (i == 0)
        if i:
            blocks = rec.blocks[:rec.n_blocks]
            inter = sum(1 for b in blocks if b.is_inter)
            assert 0 < inter < rec.n_blocks


def test_segmentation_stream_reaches_every_segment_and_feature():
    """The generator's segmentation: every segment id in the map, ALT_Q and
    ALT_L in delta and absolute mode, the forced reference and skip, a map
    kept from the prediction, and the oracle decodes it."""
    frames = streamgen.generate(352, 288, 6, seed=61, content=1, intra_percent=10, segmentation=1)
    p = vp.Parser()
    seen, modes, kept = set(), set(), 0
    for f in frames:
        r = p.parse(f)
        h = r.hdr
        assert h.seg_enabled
        ids = {r.blocks[i].segment_id for i in range(r.n_blocks)}
        seen |= ids
        kept += ids == {0}
        modes.add(h.seg_abs_or_delta_update)
        assert h.seg_feature_enabled[1][0] and h.seg_feature_enabled[3][1] and h.seg_feature_enabled[5][3]
        if not h.frame_is_intra:
            assert h.seg_feature_enabled[4][2] and h.seg_feature_data[4][2] == 1
            for i in range(r.n_blocks):
                b = r.blocks[i]
                if b.segment_id == 4:
                    assert b.is_inter and b.ref_frame[0] == 1
                if b.segment_id == 5:
                    assert b.skip
    assert seen == set(range(8)) and modes == {0, 1} and kept >= 1
    shown, _, _ = decode_all(frames)
    assert len(shown) == 6
