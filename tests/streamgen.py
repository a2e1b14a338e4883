"""ctypes access to the stream generator (tests/streamgen): synthetic VP9
streams for the benchmark configurations.  Test infrastructure."""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
GEN_DIR = os.path.join(HERE, "streamgen")
GEN_LIB = os.path.join(GEN_DIR, "_build", "libstreamgen.so")


class Config(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "width", "height", "frames", "seed", "bit_depth", "ss_x", "ss_y", "base_q_idx", "tile_cols_log2",
        "golden_interval", "content", "resize_every", "intra_percent", "allow_hp", "lf_level", "lf_sharpness",
        "segmentation")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        # make rebuilds when the generator's or the parser's sources are newer
        subprocess.run(["make", "-s", "-C", GEN_DIR], check=True)
        L = ctypes.CDLL(GEN_LIB)
        L.sg_new.restype = ctypes.c_void_p
        L.sg_new.argtypes = [ctypes.POINTER(Config)]
        L.sg_free.argtypes = [ctypes.c_void_p]
        L.sg_error.restype = ctypes.c_char_p
        L.sg_next.restype = ctypes.c_long
        L.sg_next.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))]
        _lib = L
    return _lib


def generate(width, height, frames, seed, bit_depth=8, ss_x=1, ss_y=1, base_q_idx=100, tile_cols_log2=0,
             golden_interval=8, content=1, resize_every=0, intra_percent=3, allow_hp=1, lf_level=None,
             lf_sharpness=0, segmentation=0):
    """The frames (bytes) of a generated stream.  lf_level: the loop filter
    level (0..63; None derives it from the quantiser); lf_sharpness 0..7."""
    L = lib()
    if lf_level is not None and not 0 <= lf_level <= 63 or not 0 <= lf_sharpness <= 7:
        raise ValueError("loop filter level 0..63, sharpness 0..7")
    cfg = Config(width, height, frames, seed, bit_depth, ss_x, ss_y, base_q_idx, tile_cols_log2,
                 golden_interval, content, resize_every, intra_percent, allow_hp,
                 0 if lf_level is None else lf_level + 1, lf_sharpness, segmentation)
    g = L.sg_new(ctypes.byref(cfg))
    out = []
    try:
        while True:
            p = ctypes.POINTER(ctypes.c_uint8)()
            n = L.sg_next(g, ctypes.byref(p))
            if n < 0:
                raise RuntimeError("stream generator: " + L.sg_error().decode())
            if n == 0:
                break
            out.append(ctypes.string_at(p, n))
    finally:
        L.sg_free(g)
    return out


# The benchmark configurations (sizes, formats, seeds, quantisers, tiles);
# like the crate's encoder, no high-precision motion vectors
# (src/encoder/mod.rs:668, 909).
CONFIG1 = dict(width=352, height=288, frames=30, seed=1, base_q_idx=120, golden_interval=8, tile_cols_log2=0,
               content=0, intra_percent=1, allow_hp=0)
CONFIG2 = dict(width=1920, height=1080, frames=60, seed=2, base_q_idx=100, golden_interval=8, tile_cols_log2=2,
               content=1, intra_percent=3, allow_hp=0)
CONFIG4 = dict(width=3840, height=2160, frames=48, seed=4, bit_depth=10, base_q_idx=80, golden_interval=8,
               tile_cols_log2=2, content=1, resize_every=12, intra_percent=3, allow_hp=0)


def config5(seed=5, streams=32):
    """The serving workload: `streams` streams arriving Poisson at 4/s;
    resolutions 50% 640x360, 30% 1280x720, 20% 1920x1080; formats 70% 8-bit
    4:2:0 (profile 0), 20% 10-bit 4:2:0 (profile 2), 10% 12-bit 4:4:4
    (profile 3); 90 to 300 frames of moving rectangles each.  Returns a list
    of (arrival seconds, generator arguments), in arrival order."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = [(640, 360), (1280, 720), (1920, 1080)]
    formats = [(8, 1, 1), (10, 1, 1), (12, 0, 0)]
    t = 0.0
    out = []
    for i in range(streams):
        t += float(rng.exponential(0.25))
        w, h = sizes[int(rng.choice(3, p=[0.5, 0.3, 0.2]))]
        bd, sx, sy = formats[int(rng.choice(3, p=[0.7, 0.2, 0.1]))]
        n = int(rng.integers(90, 301))
        out.append((t, dict(width=w, height=h, frames=n, seed=seed * 1000 + i, bit_depth=bd, ss_x=sx, ss_y=sy,
                            base_q_idx=100, golden_interval=8, tile_cols_log2=2, content=1, intra_percent=3,
                            allow_hp=0)))
    return out
