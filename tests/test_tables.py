"""The specification tables in rivet-vp9_amd/csrc/vp9_tables.h: every table
still has the weighted checksum recorded when it was generated
(csrc/gen_tables.py), and the tables obey what the specification fixes
independently of any transcription: scans are permutations, quantiser
lookups are increasing from their first value, every interpolation filter
phase sums to 128 with phase 0 the identity and bilinear phase 8 the
midpoint, probabilities are in 1..255, block lookups agree with each other."""
import os
import re

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rivet-vp9_amd", "csrc",
                   "vp9_tables.h")


def tables():
    text = open(HDR).read()
    out = {}
    pat = re.compile(r"/\* (\w+): [^*]*\(weighted checksum (\d+)\) \*/\s*static const \w+ VP9_\1((?:\[\d+\])+) = "
                     r"\{(.*?)\};", re.S)
    for m in pat.finditer(text):
        name, chk, shape, body = m.group(1), int(m.group(2)), m.group(3), m.group(4)
        dims = [int(d) for d in re.findall(r"\d+", shape)]
        vals = np.array([int(v) for v in re.findall(r"-?\d+", body)], dtype=np.int64)
        out[name] = (chk, vals.reshape(dims))
    return out


T = tables()


def test_every_table_is_parsed():
    assert len(T) >= 50
    for name in ("DC_QLOOKUP", "AC_QLOOKUP", "SUBPEL_FILTERS", "DEFAULT_SCAN_32X32", "DEFAULT_COEF_PROBS"):
        assert name in T


@pytest.mark.parametrize("name", sorted(T))
def test_weighted_checksum(name):
    chk, v = T[name]
    flat = v.reshape(-1)
    assert int(np.sum((np.arange(1, len(flat) + 1)) * flat) & 0xFFFFFFFF) == chk


def test_scans_are_permutations():
    for name, (_, v) in T.items():
        if "_SCAN_" in name:
            assert sorted(v.tolist()) == list(range(len(v))), name


def test_quantiser_lookups():
    dc, ac = T["DC_QLOOKUP"][1], T["AC_QLOOKUP"][1]
    # first and last entries of the 8/10/12-bit tables; increasing (the
    # 8-bit DC table repeats some steps: 4, 8, 8, 9, ...)
    assert dc[:, 0].tolist() == [4, 4, 4] and ac[:, 0].tolist() == [4, 4, 4]
    assert dc[0, :4].tolist() == [4, 8, 8, 9]
    assert ac[0, 255] == 1828 and ac[1, 255] == 7312 and ac[2, 255] == 29247
    assert dc[0, 255] == 1336 and dc[1, 255] == 5347 and dc[2, 255] == 21387
    assert np.all(np.diff(dc[0]) >= 0)
    for t in (dc[1:], ac):
        assert np.all(np.diff(t, axis=1) > 0)


def test_interpolation_filters():
    f = T["SUBPEL_FILTERS"][1]
    assert np.all(f.sum(axis=2) == 128)
    assert np.all(f[:, 0] == [0, 0, 0, 128, 0, 0, 0, 0])
    assert f[3, 8].tolist() == [0, 0, 0, 64, 64, 0, 0, 0]  # bilinear half sample
    assert np.all(f[3, :, [0, 1, 2, 5, 6, 7]] == 0)  # bilinear: two taps


def test_probabilities_in_range():
    # zeros only where a table is padded: the transform size probabilities
    # of 8x8 (one) and 16x16 (two) in rows of three, and band 0's three
    # unused contexts of the coefficient probabilities
    for name, (_, v) in T.items():
        if "PROB" not in name:
            continue
        if name == "DEFAULT_TX_PROBS":
            assert np.all(v[1, :, 1:] == 0) and np.all(v[2, :, 2] == 0) and np.all(v[0] == 0)
            v = np.concatenate([v[1, :, :1].ravel(), v[2, :, :2].ravel(), v[3].ravel()])
        elif name == "DEFAULT_COEF_PROBS":
            assert np.all(v[:, :, :, 0, 3:] == 0)
            v = np.concatenate([v[:, :, :, 0, :3].ravel(), v[:, :, :, 1:].ravel()])
        assert v.min() >= 1 and v.max() <= 255, name


def test_block_lookups_agree():
    bw, bh = T["B_WIDTH_LOG2"][1], T["B_HEIGHT_LOG2"][1]
    assert np.all(T["NUM_4X4_WIDE"][1] == 1 << bw) and np.all(T["NUM_4X4_HIGH"][1] == 1 << bh)
    assert np.all(T["NUM_8X8_WIDE"][1] == np.maximum(1, (1 << bw) >> 1))
    assert np.all(T["NUM_8X8_HIGH"][1] == np.maximum(1, (1 << bh) >> 1))
