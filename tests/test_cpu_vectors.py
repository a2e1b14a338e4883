"""The CPU oracle pinned end to end: the product parser plus the oracle's
reconstruction and loop filter decode the reference's 15 committed streams
(tests/data of safewords/rivet-vp9; tests/committed_vectors.rs:6-29) with every
shown frame's MD5 equal to the published one: 8, 10 and 12 bits, 4:2:0,
4:2:2, 4:4:0, 4:4:4, odd sizes, scaled references, lossless, bilinear,
show_existing_frame, a corrupt container tail, and the pre-final profile 1
stream whose chroma transforms are sized as 4:2:0
(committed_vectors.rs:30-45)."""
import os

import pytest

import oracle_lib
import webm

VECTORS = webm.vector_paths()


def test_all_fifteen_committed_streams_are_here():
    assert len(VECTORS) == 15


@pytest.mark.parametrize("path", VECTORS, ids=[os.path.basename(p) for p in VECTORS])
def test_oracle_decodes_committed_stream(path):
    exp = webm.expected_md5s(path)
    d = oracle_lib.CpuDecoder()
    got = []
    for pkt in webm.packets(path):
        for f in d.decode(pkt):
            got.append(oracle_lib.md5(f.packed()))
    assert got == exp


def test_legacy_profile1_detection():
    """vp91-2-04-yv444 needs the legacy 4:2:0 chroma sizing, yuv444 must not
    take it (committed_vectors.rs:30-45)."""
    seen = {}
    for name in ("vp91-2-04-yv444.webm", "vp91-2-04-yuv444.webm"):
        d = oracle_lib.CpuDecoder()
        legacy = []
        for pkt in webm.packets(os.path.join(webm.VECTORS, name)):
            d.decode(pkt, on_record=lambda rec, refs: legacy.append(rec.hdr.legacy_uv))
        seen[name] = any(legacy)
    assert seen == {"vp91-2-04-yv444.webm": True, "vp91-2-04-yuv444.webm": False}
