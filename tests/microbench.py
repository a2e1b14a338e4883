"""Seeded inputs of the kernel microbenchmark (no bitstream): a batch of
inverse-transform blocks and a batch of motion-compensation blocks, in the
layouts of vp9r_itx_add_batch and vp9r_mc_batch.  Test infrastructure.

Inverse transforms (the workload of src/dsp/bench.rs:26-185 over many
blocks): sizes 4/8/16/32 in ratio 40/30/20/10; transform type uniform over
the four, DCT_DCT at 32x32 (itx.rs:463-469 has no 32-point ADST); an
end-of-block position eob uniform in [1, n*n]; coefficients rounded
Laplace(0, 40), each zero with probability 0.85, and zero at every scan
position from eob on (the decoder's scan of the size and type); a random
8-bit prediction under every block.  Blocks sit in one plane of
PLANE_STRIDE samples, each class in its own band of rows.

Motion compensation (inter::predict, inter.rs:289-395, unscaled): 1M blocks
of the VP9 shapes with sides in {8, 16, 32, 64} and aspect at most 2, at a
block position (bx, by) of a 1920x1080 frame; motion vectors uniform in
[-512, 512] eighth-samples, position x = (bx << 4) + 2 mv; 8-tap filter
uniform over regular, smooth and sharp; 30% compound (a second reference and
vector); four random 8-bit 1920x1080 reference planes.  Each block's
prediction lands in its own area of an output plane of PLANE_STRIDE
samples."""
import ctypes

import numpy as np

import streamgen

PLANE_STRIDE = 4096
ITX_SIZE_P = (0.4, 0.3, 0.2, 0.1)
MC_SHAPES = ((8, 8), (8, 16), (16, 8), (16, 16), (16, 32), (32, 16), (32, 32), (32, 64), (64, 32), (64, 64))
REF_W, REF_H, N_REFS = 1920, 1080, 4
_CHUNK = 1 << 22  # coefficients generated at a time
# vp9r_mc_desc (include/vp9r.h) with C alignment
MC_DESC = np.dtype([("x", "<i4"), ("y", "<i4"), ("x2", "<i4"), ("y2", "<i4"), ("xs", "<i2"), ("ys", "<i2"),
                    ("w", "u1"), ("h", "u1"), ("filter", "u1"), ("ref_idx", "i1"), ("ref2_idx", "i1"),
                    ("pad", "u1"), ("out_off", "<u8")], align=True)


def ptr(a):
    """A numpy array's data as a ctypes pointer."""
    return a.ctypes.data_as(ctypes.c_void_p)


def scan(tx_size, tx_type):
    """The decoder's scan of a size (0..3) and type: raster positions in
    scan order."""
    n = 4 << tx_size
    out = np.zeros(n * n, dtype=np.uint16)
    streamgen.lib().sg_scan(tx_size, tx_type, out.ctypes.data_as(ctypes.c_void_p))
    return out


def _bands(counts, sides):
    """Row bands of PLANE_STRIDE-wide planes for `counts[k]` blocks of side
    (w, h) = sides[k]: (first row of each band, total rows)."""
    first, rows = [], 0
    for c, (w, h) in zip(counts, sides):
        first.append(rows)
        per = PLANE_STRIDE // w
        rows += -(-c // per) * h
    return first, rows


class ItxSet:
    """n_blocks transform blocks (seeded): the vp9r_itx_add_batch inputs."""

    def __init__(self, n_blocks, seed, bit_depth=8):
        rng = np.random.default_rng(seed)
        self.count = n_blocks
        self.bit_depth = bit_depth
        self.tx = rng.choice(4, size=n_blocks, p=ITX_SIZE_P).astype(np.uint8)
        tt = rng.integers(0, 4, size=n_blocks).astype(np.uint8)
        tt[self.tx == 3] = 0
        self.tx_type = tt
        nn = (4 << self.tx.astype(np.int64)) ** 2
        self.eob = (1 + np.floor(rng.random(n_blocks) * nn)).astype(np.uint16)
        counts = [int((self.tx == k).sum()) for k in range(4)]
        # coefficients class by class; within a class in block order
        idx = [np.flatnonzero(self.tx == k) for k in range(4)]
        base = np.cumsum([0] + [c * (16 << (2 * k)) for k, c in enumerate(counts)])
        self.coefs = np.zeros(int(base[-1]), dtype=np.int32)
        self.coef_off = np.zeros(n_blocks, dtype=np.uint32)
        first, self.rows = _bands(counts, [(4 << k, 4 << k) for k in range(4)])
        self.dst_off = np.zeros(n_blocks, dtype=np.uint64)
        for k in range(4):
            n0 = 4 << k
            nsq = n0 * n0
            blocks = idx[k]
            self.coef_off[blocks] = base[k] + np.arange(len(blocks), dtype=np.int64) * nsq
            per = PLANE_STRIDE // n0
            j = np.arange(len(blocks), dtype=np.int64)
            self.dst_off[blocks] = ((first[k] + (j // per) * n0) * PLANE_STRIDE + (j % per) * n0).astype(np.uint64)
            # rank[t][p]: the scan index of raster position p under type t
            rank = np.zeros((4, nsq), dtype=np.int32)
            for t in range(4):
                rank[t][scan(k, t)] = np.arange(nsq)
            step = max(1, _CHUNK // nsq)
            for s in range(0, len(blocks), step):
                b = blocks[s:s + step]
                keep = rng.random((len(b), nsq), dtype=np.float32) >= 0.85
                keep &= rank[self.tx_type[b]] < self.eob[b, None].astype(np.int32)
                v = np.zeros((len(b), nsq), dtype=np.int32)
                v[keep] = np.rint(rng.laplace(0.0, 40.0, size=int(keep.sum())))
                o = int(base[k]) + s * nsq
                self.coefs[o:o + v.size] = v.ravel()
        self.plane = rng.integers(0, 1 << bit_depth, size=(self.rows, PLANE_STRIDE)).astype(np.uint16)

    def bytes(self):
        """Algorithmic HBM bytes: per block 4 n^2 coefficient bytes read and
        n^2 prediction samples read and written."""
        b = 1 if self.bit_depth == 8 else 2
        nn = (4 << self.tx.astype(np.int64)) ** 2
        return int((nn * (4 + 2 * b)).sum())


class McSet:
    """n_blocks motion-compensation descriptors (seeded) and their reference
    planes: the vp9r_mc_batch inputs."""

    def __init__(self, n_blocks, seed, bit_depth=8):
        rng = np.random.default_rng(seed)
        self.count = n_blocks
        self.bit_depth = bit_depth
        self.refs = [rng.integers(0, 1 << bit_depth, size=(REF_H, REF_W)).astype(np.uint16)
                     for _ in range(N_REFS)]
        shape = rng.integers(0, len(MC_SHAPES), size=n_blocks)
        sides = np.array(MC_SHAPES, dtype=np.int64)
        w, h = sides[shape, 0], sides[shape, 1]
        bx = w * np.floor(rng.random(n_blocks) * (REF_W // w)).astype(np.int64)
        by = h * np.floor(rng.random(n_blocks) * (REF_H // h)).astype(np.int64)
        mv = rng.integers(-512, 513, size=(n_blocks, 4))
        compound = rng.random(n_blocks) < 0.3
        ref = rng.integers(0, N_REFS, size=n_blocks)
        ref2 = (ref + rng.integers(1, N_REFS, size=n_blocks)) % N_REFS
        d = np.zeros(n_blocks, dtype=MC_DESC)
        d["x"] = (bx << 4) + 2 * mv[:, 0]
        d["y"] = (by << 4) + 2 * mv[:, 1]
        d["x2"] = (bx << 4) + 2 * mv[:, 2]
        d["y2"] = (by << 4) + 2 * mv[:, 3]
        d["xs"] = 16
        d["ys"] = 16
        d["w"], d["h"] = w, h
        d["filter"] = rng.integers(0, 3, size=n_blocks)
        d["ref_idx"] = ref
        d["ref2_idx"] = np.where(compound, ref2, -1)
        counts = [int((shape == s).sum()) for s in range(len(MC_SHAPES))]
        first, self.rows = _bands(counts, MC_SHAPES)
        for s, (sw, sh) in enumerate(MC_SHAPES):
            blocks = np.flatnonzero(shape == s)
            per = PLANE_STRIDE // sw
            j = np.arange(len(blocks), dtype=np.int64)
            d["out_off"][blocks] = (first[s] + (j // per) * sh) * PLANE_STRIDE + (j % per) * sw
        self.desc = d

    def bytes(self):
        """Algorithmic HBM bytes: per reference read, the filter footprint
        (w + 7 [fx != 0]) (h + 7 [fy != 0]) samples (inter.rs:248-281), and
        w h samples written."""
        b = 1 if self.bit_depth == 8 else 2
        d = self.desc
        w, h = d["w"].astype(np.int64), d["h"].astype(np.int64)

        def footprint(x, y):
            return (w + 7 * ((x & 15) != 0)) * (h + 7 * ((y & 15) != 0))

        rd = footprint(d["x"], d["y"]) + np.where(d["ref2_idx"] >= 0, footprint(d["x2"], d["y2"]), 0)
        return int(((rd + w * h) * b).sum())


def _chunks(count, threads):
    step = -(-count // max(1, threads))
    return [(a, min(count, a + step)) for a in range(0, count, step)]


def oracle_itx(s, threads=1, count=None):
    """The oracle's result of blocks [0, count) of an ItxSet over its
    prediction plane (u16), on `threads` host threads over disjoint blocks."""
    import concurrent.futures

    import oracle_lib
    L, P = oracle_lib.lib(), oracle_lib.ptr
    count = s.count if count is None else count
    plane = s.plane.copy()

    def run(ab):
        a, b = ab
        L.orc_itx_add_batch(b - a, P(s.coefs), P(s.coef_off[a:]), P(s.tx[a:]), P(s.tx_type[a:]), P(s.eob[a:]),
                            P(s.dst_off[a:]), P(plane), PLANE_STRIDE, s.bit_depth, 0)

    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        list(ex.map(run, _chunks(count, threads)))
    return plane


def oracle_mc(m, threads=1, count=None):
    """The oracle's predictions of descriptors [0, count) of an McSet in an
    output plane (u16, zero elsewhere), on `threads` host threads."""
    import concurrent.futures

    import oracle_lib
    L, P = oracle_lib.lib(), oracle_lib.ptr
    count = m.count if count is None else count
    out = np.zeros((m.rows, PLANE_STRIDE), dtype=np.uint16)
    refs = (ctypes.c_void_p * N_REFS)(*[r.ctypes.data for r in m.refs])

    def run(ab):
        a, b = ab
        L.orc_mc_batch(b - a, P(m.desc[a:]), refs, REF_W, REF_W - 1, REF_H - 1, P(out), PLANE_STRIDE,
                       m.bit_depth)

    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        list(ex.map(run, _chunks(count, threads)))
    return out
