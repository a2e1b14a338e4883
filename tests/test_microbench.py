"""The kernel microbenchmark's inputs (tests/microbench.py) and the oracle's
batched forms, on the CPU: the generator follows its stated distributions
(size ratio, DCT_DCT at 32x32, zeros from the scan's eob on, VP9 shapes of
aspect <= 2, 30% compound, positions (bx << 4) + 2 mv), the scans are
permutations starting at the DC position, and orc_itx_add_batch /
orc_mc_batch equal the per-block orc_itx_add / orc_inter_predict + orc_avg
that the decoder tests pin."""
import numpy as np

import microbench as mb
import oracle_lib


def test_scans_are_permutations():
    for k in range(4):
        for t in range(4 if k < 3 else 1):
            s = mb.scan(k, t)
            n = 4 << k
            assert s[0] == 0
            assert np.array_equal(np.sort(s), np.arange(n * n))
    # the default 4x4 scan of the VP9 specification (default_scan_4x4)
    assert list(mb.scan(0, 0)) == [0, 4, 1, 5, 8, 2, 12, 9, 3, 6, 13, 10, 7, 14, 11, 15]


def test_itx_set_distribution():
    s = mb.ItxSet(40_000, 3)
    frac = np.bincount(s.tx, minlength=4) / s.count
    assert np.allclose(frac, mb.ITX_SIZE_P, atol=0.01)
    assert (s.tx_type[s.tx == 3] == 0).all()
    assert set(np.unique(s.tx_type[s.tx < 3])) == {0, 1, 2, 3}
    nn = (4 << s.tx.astype(np.int64)) ** 2
    assert (s.eob >= 1).all() and (s.eob <= nn).all()
    nz = []
    for i in range(0, s.count, 37):
        n0 = 4 << int(s.tx[i])
        c = s.coefs[int(s.coef_off[i]):int(s.coef_off[i]) + n0 * n0]
        order = mb.scan(int(s.tx[i]), int(s.tx_type[i]))
        assert not c[order[int(s.eob[i]):]].any(), f"block {i}: coefficient after eob"
        nz.append(np.count_nonzero(c[order[:int(s.eob[i])]]) / int(s.eob[i]))
    # about 15% of the positions before eob are non-zero (rounding adds a few zeros)
    assert 0.12 < np.mean(nz) < 0.16
    # blocks are disjoint and inside the plane
    covered = np.zeros((s.rows, mb.PLANE_STRIDE), dtype=np.uint8)
    for i in range(s.count):
        n0 = 4 << int(s.tx[i])
        r, c = divmod(int(s.dst_off[i]), mb.PLANE_STRIDE)
        covered[r:r + n0, c:c + n0] += 1
    assert covered.max() == 1 and covered.sum() == int(nn.sum())
    assert s.bytes() == int((nn * 6).sum())


def test_mc_set_distribution():
    m = mb.McSet(20_000, 3)
    d = m.desc
    w, h = d["w"].astype(int), d["h"].astype(int)
    assert set(zip(w, h)) == set(mb.MC_SHAPES)
    assert (np.maximum(w, h) <= 2 * np.minimum(w, h)).all()
    assert abs((d["ref2_idx"] >= 0).mean() - 0.3) < 0.02
    assert (d["ref2_idx"] != d["ref_idx"]).all()
    assert set(np.unique(d["filter"])) == {0, 1, 2}
    assert (d["xs"] == 16).all() and (d["ys"] == 16).all()
    # x = (bx << 4) + 2 mv with 0 <= bx <= lim - side and |mv| <= 512; the
    # vectors of a compound block differ
    for x, side, lim in ((d["x"], w, mb.REF_W), (d["y"], h, mb.REF_H), (d["x2"], w, mb.REF_W),
                         (d["y2"], h, mb.REF_H)):
        x = x.astype(np.int64)
        assert (x % 2 == 0).all()
        assert (x >= -1024).all() and (x <= ((lim - side) << 4) + 1024).all()
        assert x.min() < 0 and ((x >> 4) + side > lim).any()
    assert (d["x"] != d["x2"]).mean() > 0.99
    covered = np.zeros((m.rows, mb.PLANE_STRIDE), dtype=np.uint8)
    for i in range(m.count):
        r, c = divmod(int(d["out_off"][i]), mb.PLANE_STRIDE)
        covered[r:r + h[i], c:c + w[i]] += 1
    assert covered.max() == 1


def test_oracle_batches_equal_single_calls():
    L, P = oracle_lib.lib(), oracle_lib.ptr
    s = mb.ItxSet(3000, 11)
    got = mb.oracle_itx(s, threads=3)
    want = s.plane.copy()
    for i in range(s.count):
        n0 = 4 << int(s.tx[i])
        r, c = divmod(int(s.dst_off[i]), mb.PLANE_STRIDE)
        sub = np.ascontiguousarray(want[r:r + n0, c:c + n0])
        blk = np.ascontiguousarray(s.coefs[int(s.coef_off[i]):int(s.coef_off[i]) + n0 * n0])
        L.orc_itx_add(P(blk), 2 + int(s.tx[i]), int(s.tx_type[i]), 0, int(s.eob[i]), 8, P(sub), n0)
        want[r:r + n0, c:c + n0] = sub
    assert np.array_equal(got, want)
    assert not np.array_equal(got, s.plane)

    m = mb.McSet(800, 11)
    got = mb.oracle_mc(m, threads=3)
    for i in range(m.count):
        d = m.desc[i]
        w, h = int(d["w"]), int(d["h"])
        a = np.zeros((h, w), dtype=np.uint16)
        L.orc_inter_predict(P(m.refs[int(d["ref_idx"])]), mb.REF_W, mb.REF_W - 1, mb.REF_H - 1, int(d["x"]),
                            int(d["y"]), 16, 16, w, h, int(d["filter"]), 8, P(a), w)
        if d["ref2_idx"] >= 0:
            b = np.zeros((h, w), dtype=np.uint16)
            L.orc_inter_predict(P(m.refs[int(d["ref2_idx"])]), mb.REF_W, mb.REF_W - 1, mb.REF_H - 1, int(d["x2"]),
                                int(d["y2"]), 16, 16, w, h, int(d["filter"]), 8, P(b), w)
            c = np.zeros((h, w), dtype=np.uint16)
            L.orc_avg(P(a), P(b), w, h, P(c), w)
            a = c
        r, c0 = divmod(int(d["out_off"]), mb.PLANE_STRIDE)
        assert np.array_equal(got[r:r + h, c0:c0 + w], a), f"descriptor {i}"


def test_subpel_taps_fit_signed_bytes():
    """The 8-bit unscaled MC path sums taps as signed bytes over samples
    biased by -128: every fractional phase's taps must fit [-128, 127] and
    every phase sum to 128."""
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rivet-vp9_amd", "csrc",
                        "vp9_tables.h")
    text = open(path).read()
    body = text[text.index("VP9_SUBPEL_FILTERS[4][16][8]"):]
    body = body[body.index("{") + 1:body.index("};")]
    taps = np.array([int(v) for v in re.findall(r"-?\d+", body)]).reshape(4, 16, 8)
    assert (taps.sum(axis=2) == 128).all()
    assert (taps[:, 1:] >= -128).all() and (taps[:, 1:] <= 127).all()
    assert (taps[:, 0] == [0, 0, 0, 128, 0, 0, 0, 0]).all()
