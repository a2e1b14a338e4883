"""Pins of the CPU oracle that do not go through the committed streams:
closed forms, invariants and exact identities that a plausible mistake in
the restatement (a wrong constant, index, sign or rounding) breaks.

The committed streams (test_cpu_vectors.py) pin the whole frame process at
conforming inputs; these tests pin the kernels one by one, including inputs
no committed stream reaches (scaled positions, every filter phase, every
intra mode at every size, lossless).
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

L = ol.lib()
P = ol.ptr


def impulse_matrix(fn, n_log2, scale=1 << 12):
    n = 1 << n_log2
    m = np.zeros((n, n))
    for k in range(n):
        t = np.zeros(n, dtype=np.int64)
        t[k] = scale
        fn(P(t), n_log2)
        m[:, k] = t / scale
    return m


# ---------------------------------------------------------------- transforms

@pytest.mark.parametrize("n_log2", [2, 3, 4, 5])
def test_spec_idct_is_the_inverse_dct(n_log2):
    """8.7.1.3: the butterfly network is the DCT-III with 14-bit cosines,
    x[i] = X[0]/sqrt(2) + sum_k X[k] cos(pi (2i+1) k / 2N)."""
    n = 1 << n_log2
    m = impulse_matrix(L.orc_spec_idct, n_log2)
    i, k = np.arange(n)[:, None], np.arange(n)[None, :]
    c = np.cos(np.pi * (2 * i + 1) * k / (2 * n))
    c[:, 0] /= np.sqrt(2)
    assert np.abs(m - c).max() < 1e-3
    rng = np.random.default_rng(n_log2)
    for _ in range(50):
        x = rng.integers(-4096, 4096, n).astype(np.int64)
        t = x.copy()
        L.orc_spec_idct(P(t), n_log2)
        # one rounding per butterfly stage
        assert np.abs(t - c @ x).max() <= 2 * n_log2


def test_spec_iadst4_is_the_sine_transform():
    """ADST4 (8.7.1.8): x[i] = (2 sqrt(2) / 3) sum_k X[k] sin(pi (i+1) (2k+1) / 9)."""
    m = impulse_matrix(L.orc_spec_iadst, 2)
    i, k = np.arange(4)[:, None], np.arange(4)[None, :]
    s = (2 * np.sqrt(2) / 3) * np.sin(np.pi * (i + 1) * (2 * k + 1) / 9)
    assert np.abs(m - s).max() < 1e-3


@pytest.mark.parametrize("n_log2", [3, 4])
def test_spec_iadst_is_the_dst4(n_log2):
    """ADST8/16 (8.7.1.6-7): x[i] = sum_k X[k] sin(pi (2i+1) (2k+1) / 4N)."""
    n = 1 << n_log2
    m = impulse_matrix(L.orc_spec_iadst, n_log2)
    i, k = np.arange(n)[:, None], np.arange(n)[None, :]
    s = np.sin(np.pi * (2 * i + 1) * (2 * k + 1) / (4 * n))
    assert np.abs(m - s).max() < 1e-3


def sparse_block(rng, n, eob_full=True, scale=40.0, zeros=0.85):
    c = np.round(rng.laplace(0, scale, n * n)).astype(np.int32)
    c[rng.random(n * n) < zeros] = 0
    return c


@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("n_log2", [2, 3, 4, 5])
def test_lane_transforms_equal_the_spec_process(bd, n_log2):
    """itx.rs's S32 (8-bit) and S64 (10/12-bit) lanes equal the 64-bit
    specification process (8.7.2) on conforming input, every transform type."""
    n = 1 << n_log2
    rng = np.random.default_rng(100 * bd + n_log2)
    mx = (1 << bd) - 1
    for trial in range(24):
        tx_type = 0 if n == 32 else trial % 4
        coefs = sparse_block(rng, n, scale=40.0 * (1 << (bd - 8)))
        pred = rng.integers(0, mx + 1, (n, n)).astype(np.uint16)
        a = pred.copy()
        L.orc_itx_add(P(coefs), n_log2, tx_type, 0, n * n, bd, P(a), n)
        b = pred.copy()
        L.orc_itx_spec_reconstruct(P(coefs), n_log2, tx_type, 0, P(b), mx)
        assert np.array_equal(a, b), (bd, n, tx_type, trial)


@pytest.mark.parametrize("n_log2", [2, 3, 4, 5])
def test_dc_only_shortcut_equals_full_transform(n_log2):
    """itx.rs:729-733: at 8 bits, eob == 1 with DCT_DCT adds one constant,
    which must be what the full 2D transform of a DC-only block gives."""
    n = 1 << n_log2
    rng = np.random.default_rng(7 + n_log2)
    for dc in list(rng.integers(-3000, 3000, 40)) + [0, 1, -1, 4095, -4096]:
        coefs = np.zeros(n * n, dtype=np.int32)
        coefs[0] = dc
        pred = rng.integers(0, 256, (n, n)).astype(np.uint16)
        a = pred.copy()
        L.orc_itx_add(P(coefs), n_log2, 0, 0, 1, 8, P(a), n)
        b = pred.copy()
        L.orc_itx_spec_reconstruct(P(coefs), n_log2, 0, 0, P(b), 255)
        assert np.array_equal(a, b), dc
        # and the added value is the same everywhere (before clamping)
        d = a.astype(int) - pred.astype(int)
        inside = (a > 0) & (a < 255)
        assert len(set(d[inside].ravel().tolist())) <= 1


def fwht4x4(x):
    """The lossless forward Walsh-Hadamard transform of the VP9 reference
    encoder (libvpx vp9_fwht4x4): columns then rows, lifting steps, output
    scaled by 4 (UNIT_QUANT_FACTOR)."""
    def lift(a1, b1, c1, d1):
        a1 += b1
        d1 -= c1
        e1 = (a1 - d1) >> 1
        b1 = e1 - b1
        c1 = e1 - c1
        a1 -= c1
        d1 += b1
        return a1, c1, d1, b1
    t = [[0] * 4 for _ in range(4)]
    for c in range(4):
        col = lift(x[0][c], x[1][c], x[2][c], x[3][c])
        for r in range(4):
            t[r][c] = col[r]
    out = [[0] * 4 for _ in range(4)]
    for r in range(4):
        row = lift(*t[r])
        for c in range(4):
            out[r][c] = row[c] * 4
    return out


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_lossless_wht_inverts_the_encoder_forward_wht(bd):
    """Lossless coding is exact: the inverse WHT (8.7.1.10) of the forward
    WHT of any residual restores the residual."""
    rng = np.random.default_rng(bd)
    mx = (1 << bd) - 1
    for _ in range(200):
        src = rng.integers(0, mx + 1, (4, 4))
        pred = rng.integers(0, mx + 1, (4, 4))
        res = (src - pred).tolist()
        coefs = np.array(fwht4x4(res), dtype=np.int32).ravel()
        out = pred.astype(np.uint16)
        L.orc_itx_add(P(coefs), 2, 0, 1, 16, bd, P(out), 4)
        assert np.array_equal(out.astype(int), src)


# ---------------------------------------------------------------- intra

def intra(mode, log2, above, left, have_left=1, have_above=1, bd=8):
    n = 1 << log2
    a = np.ascontiguousarray(above, dtype=np.int32)
    lft = np.ascontiguousarray(left, dtype=np.int32)
    assert len(a) >= 2 * n + 1 and len(lft) >= n
    out = np.zeros((n, n), dtype=np.uint16)
    L.orc_intra_predict(mode, log2, P(a), P(lft), have_left, have_above, bd, P(out), n)
    return out.astype(int)


@pytest.mark.parametrize("log2", [2, 3, 4, 5])
def test_intra_constant_edges_give_a_constant(log2):
    n = 1 << log2
    for v in (0, 77, 255):
        for mode in range(10):
            p = intra(mode, log2, [v] * (2 * n + 1), [v] * n)
            assert (p == v).all(), (mode, v)


@pytest.mark.parametrize("log2", [2, 3, 4, 5])
def test_intra_closed_forms(log2):
    n = 1 << log2
    rng = np.random.default_rng(log2)
    for _ in range(20):
        above = rng.integers(0, 256, 2 * n + 1)
        left = rng.integers(0, 256, n)
        a, tl = above[1:], above[0]
        assert (intra(1, log2, above, left) == a[None, :n]).all()  # V
        assert (intra(2, log2, above, left) == left[:n, None]).all()  # H
        assert (intra(9, log2, above, left) == np.clip(left[:n, None] + a[None, :n] - tl, 0, 255)).all()  # TM
        dc = (a[:n].sum() + left.sum() + n) // (2 * n)
        assert (intra(0, log2, above, left) == dc).all()
        assert (intra(0, log2, above, left, have_left=0) == (a[:n].sum() + n // 2) // n).all()
        assert (intra(0, log2, above, left, have_above=0) == (left.sum() + n // 2) // n).all()
        assert (intra(0, log2, above, left, 0, 0) == 128).all()
        assert (intra(0, log2, above, left, 0, 0, bd=10) == 512).all()


@pytest.mark.parametrize("log2", [2, 3, 4, 5])
def test_intra_directional_structure(log2):
    """Each directional predictor is constant along its direction, reads
    only the edges it should, and reproduces a linear ramp."""
    n = 1 << log2
    rng = np.random.default_rng(10 + log2)
    above = rng.integers(0, 256, 2 * n + 1)
    left = rng.integers(0, 256, n)
    other_above = rng.integers(0, 256, 2 * n + 1)
    other_left = rng.integers(0, 256, n)
    d45, d135 = intra(3, log2, above, left), intra(4, log2, above, left)
    d117, d153 = intra(5, log2, above, left), intra(6, log2, above, left)
    d207, d63 = intra(7, log2, above, left), intra(8, log2, above, left)
    for r in range(n - 1):
        for c in range(n - 1):
            assert d45[r + 1, c] == d45[r, c + 1]  # anti-diagonals
            assert d135[r + 1, c + 1] == d135[r, c]  # diagonals
    for r in range(n - 2):
        for c in range(1, n):
            assert d117[r + 2, c] == d117[r, c - 1]
        for c in range(n - 1):
            assert d63[r + 2, c] == d63[r, c + 1]
    for r in range(1, n):
        for c in range(n - 2):
            assert d153[r, c + 2] == d153[r - 1, c]
    for r in range(n - 1):
        for c in range(n - 2):
            assert d207[r, c + 2] == d207[r + 1, c]
    # edge dependence
    assert (intra(3, log2, above, other_left) == d45).all()
    assert (intra(8, log2, above, other_left) == d63).all()
    assert (intra(7, log2, other_above, left) == d207).all()
    assert not (intra(7, log2, above, other_left) == d207).all()
    # a linear ramp along the above row is reproduced by the 3-tap smoothing
    ramp = np.arange(2 * n + 1) * 3 + 10
    p = intra(3, log2, ramp, [0] * n)
    for r in range(n):
        for c in range(n):
            if r + c + 2 < 2 * n:
                assert p[r, c] == ramp[1 + r + c + 1]


# ---------------------------------------------------------------- inter

def predict(ref, x, y, xs, ys, w, h, filt, bd=8, last=None):
    rows, cols = ref.shape
    lx, ly = (cols - 1, rows - 1) if last is None else last
    out = np.zeros((h, w), dtype=np.uint16)
    r = np.ascontiguousarray(ref, dtype=np.uint16)
    L.orc_inter_predict(P(r), cols, lx, ly, x, y, xs, ys, w, h, filt, bd, P(out), w)
    return out.astype(int)


def test_inter_integer_positions_copy():
    rng = np.random.default_rng(1)
    ref = rng.integers(0, 256, (96, 96))
    for filt in range(4):
        for (w, h) in ((4, 4), (8, 16), (64, 64), (32, 8)):
            x, y = int(rng.integers(0, 96 - w)), int(rng.integers(0, 96 - h))
            p = predict(ref, x * 16, y * 16, 16, 16, w, h, filt)
            assert (p == ref[y:y + h, x:x + w]).all()


def test_inter_constant_reference_is_constant():
    """Every filter phase sums to 128 and rounds exactly on a constant."""
    for v in (0, 100, 255):
        ref = np.full((40, 40), v)
        for filt in range(4):
            for fx in range(16):
                for fy in (0, 5, 15):
                    p = predict(ref, 160 + fx, 160 + fy, 16, 16, 8, 8, filt)
                    assert (p == v).all()


def test_inter_bilinear_interpolates_a_ramp():
    """Filter 3 is bilinear: on a linear ramp it gives the ramp at the
    fractional position (exact when the product is a multiple of 16)."""
    xs = np.arange(64)
    ref = np.tile(xs * 2, (24, 1))
    for fx in range(0, 16, 8):
        p = predict(ref, 16 * 10 + fx, 16 * 4, 16, 16, 8, 8, 3)
        want = (xs[10:18] * 2 * 16 + 2 * fx) // 16
        assert (p == want[None, :]).all()


def test_inter_edges_repeat_the_border():
    """Positions beyond the last sample read the repeated edge sample."""
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 256, (32, 32))
    p = predict(ref, 16 * 100, 16 * 100, 16, 16, 8, 8, 0)
    assert (p == ref[31, 31]).all()
    p = predict(ref, -16 * 100, 16 * 5, 16, 16, 8, 8, 0)
    assert (p == ref[5:13, 0][:, None]).all()
    # the visible size, not the stored one, bounds the read
    p = predict(ref, 16 * 20, 16 * 2, 16, 16, 8, 8, 0, last=(23, 31))
    assert (p == ref[2:10, [min(20 + c, 23) for c in range(8)]]).all()


def test_inter_scaled_step_subsamples():
    """A 2:1 step at an integer position reads every other sample."""
    rng = np.random.default_rng(4)
    ref = rng.integers(0, 256, (64, 64))
    for filt in range(3):
        p = predict(ref, 16 * 6, 16 * 4, 32, 32, 8, 8, filt)
        assert (p == ref[4:20:2, 6:22:2]).all()


def test_inter_unscaled_formula_equals_scaled_path_at_step_16():
    """Step 16 through the general (scaled) formula equals the separable
    unscaled filter: a 2-tap-free check, the 1D filter applied twice."""
    rng = np.random.default_rng(5)
    ref = rng.integers(0, 256, (48, 48))
    taps = subpel_filters()
    for filt in range(3):
        for _ in range(10):
            fx, fy = int(rng.integers(0, 16)), int(rng.integers(0, 16))
            x0, y0 = 20, 18
            p = predict(ref, x0 * 16 + fx, y0 * 16 + fy, 16, 16, 8, 8, filt)
            # separable reference filter with per-pass rounding and clamping
            hrows = np.zeros((15, 8), dtype=int)
            for r in range(15):
                for c in range(8):
                    s = sum(taps[filt][fx][k] * ref[y0 - 3 + r, x0 - 3 + c + k] for k in range(8))
                    hrows[r, c] = min(max((s + 64) >> 7, 0), 255)
            want = np.zeros((8, 8), dtype=int)
            for r in range(8):
                for c in range(8):
                    s = sum(taps[filt][fy][k] * hrows[r + k, c] for k in range(8))
                    want[r, c] = min(max((s + 64) >> 7, 0), 255)
            assert (p == want).all()


def subpel_filters():
    """The filter bank read back through the oracle: the impulse response of
    a one-sample position at each phase (the rounding is exact for 128 * 1)."""
    taps = [[None] * 16 for _ in range(4)]
    for filt in range(4):
        for ph in range(16):
            t = []
            for k in range(8):
                ref = np.zeros((1, 32), dtype=int)
                ref[0, 12 - 3 + k] = 255  # the sample under tap k at position 12
                # one row, horizontal pass only: y integer
                p = predict(np.repeat(ref, 16, axis=0), 12 * 16 + ph, 16 * 4, 16, 16, 1, 1, filt)
                t.append(p[0, 0])
            taps[filt][ph] = t
    return taps_from_responses(taps)


def taps_from_responses(resp):
    """A clamped impulse response, Round2(255 * tap, 7), cannot give negative
    taps back, so the bank is read from the generated tables and the
    oracle's responses are checked against it."""
    bank = _read_bank()
    for filt in range(4):
        for ph in range(16):
            for k in range(8):
                t = bank[filt][ph][k]
                want = min(max((255 * t + 64) >> 7, 0), 255)
                assert resp[filt][ph][k] == want
    return bank


def _read_bank():
    import os
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(os.path.dirname(here), "rivet-vp9_amd", "csrc", "vp9_tables.h")
    text = open(path).read()
    m = re.search(r"VP9_SUBPEL_FILTERS\[4\]\[16\]\[8\]\s*=\s*\{(.*?)\};", text, re.S)
    nums = [int(v) for v in re.findall(r"-?\d+", m.group(1))]
    assert len(nums) == 4 * 16 * 8
    bank = [[nums[(f * 16 + p) * 8:(f * 16 + p) * 8 + 8] for p in range(16)] for f in range(4)]
    for f in range(4):
        for p in range(16):
            assert sum(bank[f][p]) == 128
        assert bank[f][0] == [0, 0, 0, 128, 0, 0, 0, 0]
    for p in range(16):  # bilinear: two taps, 128 - 8p and 8p
        assert bank[3][p][3] == 128 - 8 * p and bank[3][p][4] == 8 * p
    return bank


def test_avg_closed_form():
    rng = np.random.default_rng(6)
    a = rng.integers(0, 4096, (8, 16)).astype(np.uint16)
    b = rng.integers(0, 4096, (8, 16)).astype(np.uint16)
    out = np.zeros((8, 16), dtype=np.uint16)
    L.orc_avg(P(a), P(b), 16, 8, P(out), 16)
    assert (out == (a.astype(int) + b + 1) >> 1).all()


# ---------------------------------------------------------------- loop filter

def lf(buf, pos, step, fs, level, bd=8):
    limit = max(level >> 0, 1)
    blimit = 2 * (level + 2) + limit
    thresh = level >> 4
    b = np.ascontiguousarray(buf, dtype=np.uint16)
    L.orc_lf_filter(P(b), pos, step, fs, limit, blimit, thresh, bd)
    return b.astype(int)


@pytest.mark.parametrize("fs", [0, 1, 2])
def test_lf_constant_is_unchanged(fs):
    for v in (0, 60, 255):
        buf = np.full(32, v)
        assert (lf(buf, 16, 1, fs, 40) == v).all()


@pytest.mark.parametrize("fs", [0, 1, 2])
def test_lf_real_edge_is_kept(fs):
    """A step larger than blimit is an image edge: the mask leaves it."""
    buf = np.array([20] * 16 + [200] * 16)
    assert (lf(buf, 16, 1, fs, 10) == buf).all()


@pytest.mark.parametrize("fs", [0, 1, 2])
def test_lf_smooths_a_small_step_monotonically(fs):
    """A small step across a flat region: the output stays monotone, within
    the step's range, touches only the taps of its width, and narrows the
    step; the wide filters flatten it."""
    width = {0: 3, 1: 3, 2: 7}[fs]
    for lo, hi in ((100, 104), (50, 56), (200, 201)):
        buf = np.array([lo] * 16 + [hi] * 16)
        out = lf(buf, 16, 1, fs, 30)
        assert all(out[i] <= out[i + 1] for i in range(31))
        assert out.min() >= lo and out.max() <= hi
        assert (out[:16 - width] == lo).all() and (out[16 + width:] == hi).all()
        assert out[16] - out[15] <= hi - lo
        # mirror symmetry of the decision and the taps: reflecting the
        # profile (value v -> lo + hi - v, order reversed) reflects the
        # output exactly for the wide filters, whose taps are symmetric
        if fs:
            mirror = lf((lo + hi) - buf[::-1], 16, 1, fs, 30)
            assert (np.abs(mirror - ((lo + hi) - out[::-1])) <= 1).all()


def test_lf_strided_equals_contiguous():
    """Vertical edges (step 1) and horizontal edges (step = stride) are the
    same filter."""
    rng = np.random.default_rng(8)
    col = np.concatenate([rng.integers(90, 96, 16), rng.integers(100, 106, 16)])
    for fs in range(3):
        a = lf(col, 16, 1, fs, 40)
        grid = np.zeros((32, 5), dtype=int)
        grid[:, 2] = col
        g = lf(grid.ravel(), 16 * 5 + 2, 5, fs, 40).reshape(32, 5)
        assert (g[:, 2] == a).all() and (g[:, [0, 1, 3, 4]] == 0).all()
