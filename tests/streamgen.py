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
        "golden_interval", "content", "resize_every", "intra_percent", "allow_hp")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(GEN_LIB):
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
             golden_interval=8, content=1, resize_every=0, intra_percent=3, allow_hp=1):
    """The frames (bytes) of a generated stream."""
    L = lib()
    cfg = Config(width, height, frames, seed, bit_depth, ss_x, ss_y, base_q_idx, tile_cols_log2,
                 golden_interval, content, resize_every, intra_percent, allow_hp)
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
