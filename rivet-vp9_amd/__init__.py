"""rivet-vp9 reconstruction on MI355X: the Python side of the C ABI in
include/vp9r.h.

The decoder mirrors the reference crate's public API (src/decoder/mod.rs:150-216
and src/frame.rs): ``Decoder().decode(packet)`` / ``decode_all`` /
``set_threads`` / ``set_max_pixels`` / ``frames_decoded``, returning ``Frame``
objects whose ``packed()`` bytes are the crate's ``Frame::packed()`` layout.
Errors raise ``Vp9Error`` subclasses named after the crate's ``Error`` kinds
(src/lib.rs:67-81).

The compute path is libvp9r.so (HIP kernels for gfx950 plus the C++ parser);
there is no CPU fallback: without the library or a GPU, constructing a
``Decoder`` raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvp9r.so")

OK, ERR_BITSTREAM, ERR_UNSUPPORTED, ERR_INVALID_INPUT, ERR_DEVICE = 0, 1, 2, 3, 4


class Vp9Error(Exception):
    """Base of the decoder's errors (src/lib.rs:67-81)."""


class BitstreamError(Vp9Error):
    pass


class UnsupportedError(Vp9Error):
    pass


class InvalidInputError(Vp9Error):
    pass


class DeviceError(Vp9Error):
    """A HIP failure, or no usable GPU."""


_ERRORS = {
    ERR_BITSTREAM: BitstreamError,
    ERR_UNSUPPORTED: UnsupportedError,
    ERR_INVALID_INPUT: InvalidInputError,
    ERR_DEVICE: DeviceError,
}

u8, i8, u16, i16, u32, i32, u64 = (ctypes.c_uint8, ctypes.c_int8, ctypes.c_uint16, ctypes.c_int16,
                                   ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64)


class FrameHdr(ctypes.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("mi_cols", u32), ("mi_rows", u32),
        ("sb64_cols", u32), ("sb64_rows", u32), ("render_width", u32), ("render_height", u32),
        ("profile", u8), ("bit_depth", u8), ("ss_x", u8), ("ss_y", u8),
        ("frame_type", u8), ("show_frame", u8), ("show_existing_frame", u8),
        ("frame_to_show_map_idx", u8),
        ("intra_only", u8), ("frame_is_intra", u8), ("error_resilient_mode", u8), ("lossless", u8),
        ("refresh_frame_flags", u8), ("legacy_uv", u8), ("interp_filter", u8), ("tx_mode", u8),
        ("ref_frame_idx", u8 * 3), ("reference_mode", u8),
        ("color_space", u8), ("color_range", u8), ("tile_cols_log2", u8), ("tile_rows_log2", u8),
        ("base_q_idx", i32), ("delta_q_y_dc", i32), ("delta_q_uv_dc", i32), ("delta_q_uv_ac", i32),
        ("lf_level", u8), ("lf_sharpness", u8), ("lf_delta_enabled", u8), ("_pad0", u8),
        ("lf_ref_deltas", i8 * 4), ("lf_mode_deltas", i8 * 2),
        ("seg_enabled", u8), ("seg_abs_or_delta_update", u8),
        ("seg_feature_enabled", (u8 * 4) * 8), ("seg_feature_data", (i16 * 4) * 8),
        ("ref_width", u32 * 3), ("ref_height", u32 * 3),
    ]


class Block(ctypes.Structure):
    _fields_ = [
        ("mi_row", u16), ("mi_col", u16),
        ("mi_size", u8), ("skip", u8), ("tx_size", u8), ("is_inter", u8),
        ("y_mode", u8), ("uv_mode", u8), ("interp_filter", u8), ("segment_id", u8),
        ("sub_modes", u8 * 4), ("ref_frame", i8 * 2), ("avail_u", u8), ("avail_l", u8),
        ("mv", ((i32 * 2) * 4) * 2),
    ]


class Txb(ctypes.Structure):
    _fields_ = [
        ("x", u16), ("y", u16), ("plane", u8), ("tx_size", u8), ("tx_type", u8), ("_pad", u8),
        ("eob", u16), ("ncoef", u16), ("coef_start", u32), ("block", u32),
    ]


class Mi(ctypes.Structure):
    _fields_ = [("mi_size", u8), ("skip", u8), ("tx_size", u8), ("y_mode", u8),
                ("segment_id", u8), ("ref_frame0", i8), ("_pad", u8 * 2)]


class Record(ctypes.Structure):
    _fields_ = [
        ("hdr", FrameHdr),
        ("blocks", ctypes.POINTER(Block)), ("n_blocks", u32),
        ("txbs", ctypes.POINTER(Txb)), ("n_txbs", u32),
        ("coef_pos", ctypes.POINTER(u16)), ("coef_val", ctypes.POINTER(i32)), ("n_coefs", u32),
        ("mi", ctypes.POINTER(Mi)),
    ]


class FrameDesc(ctypes.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("bit_depth", u32), ("ss_x", u32), ("ss_y", u32),
        ("color_space", u32), ("full_range", u32), ("render_width", u32), ("render_height", u32),
        ("data", ctypes.POINTER(u8)), ("size", ctypes.c_size_t),
        ("plane_offset", ctypes.c_size_t * 3), ("plane_width", u32 * 3), ("plane_height", u32 * 3),
    ]


class McDesc(ctypes.Structure):
    _fields_ = [("x", i32), ("y", i32), ("x2", i32), ("y2", i32), ("xs", i16), ("ys", i16),
                ("w", u8), ("h", u8), ("filter", u8), ("ref_idx", i8), ("ref2_idx", i8),
                ("_pad", u8), ("out_off", u64)]


_lib = None


def library():
    """libvp9r.so, loaded once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P, V = ctypes.c_void_p, ctypes.c_void_p
    sz = ctypes.c_size_t
    lib.vp9r_last_error.restype = ctypes.c_char_p
    lib.vp9r_superframe_split.restype = sz
    lib.vp9r_superframe_split.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(sz), ctypes.POINTER(sz), sz]
    lib.vp9r_parser_new.restype = P
    lib.vp9r_parser_free.argtypes = [P]
    lib.vp9r_parser_set_max_pixels.argtypes = [P, u64]
    lib.vp9r_parser_parse_frame.argtypes = [P, ctypes.c_char_p, sz, ctypes.POINTER(Record)]
    lib.vp9r_packed_size.restype = sz
    lib.vp9r_packed_size.argtypes = [u32, u32, u32, u32, u32]
    _lib = lib
    return lib


def _check(status):
    if status != OK:
        msg = library().vp9r_last_error().decode(errors="replace")
        raise _ERRORS.get(status, Vp9Error)(msg or f"status {status}")


def superframe_split(packet):
    """Annex B: the frames of a packet (src/superframe.rs:6-20)."""
    lib = library()
    offs = (ctypes.c_size_t * 8)()
    sizes = (ctypes.c_size_t * 8)()
    n = lib.vp9r_superframe_split(packet, len(packet), offs, sizes, 8)
    return [packet[offs[i]:offs[i] + sizes[i]] for i in range(min(n, 8))]


class Parser:
    """The CPU parser: one frame in, the per-frame reconstruction record out."""

    def __init__(self):
        lib = library()
        self._p = lib.vp9r_parser_new()
        self._rec = Record()

    def __del__(self):
        if getattr(self, "_p", None):
            library().vp9r_parser_free(self._p)
            self._p = None

    def set_max_pixels(self, pixels):
        library().vp9r_parser_set_max_pixels(self._p, pixels)

    def parse(self, frame):
        """Parses one frame; the returned Record is valid until the next call."""
        _check(library().vp9r_parser_parse_frame(self._p, frame, len(frame), ctypes.byref(self._rec)))
        return self._rec


def ivf_frames(data):
    """The frames of an IVF file (src/ivf.rs:34-95): 32-byte header, then per
    frame a 12-byte header (little-endian size, timestamp) and the data."""
    if len(data) < 32 or data[:4] != b"DKIF":
        raise InvalidInputError("not an IVF file")
    pos = max(int.from_bytes(data[6:8], "little"), 32)
    out = []
    while pos + 12 <= len(data):
        size = int.from_bytes(data[pos:pos + 4], "little")
        pos += 12
        if pos + size > len(data):
            raise InvalidInputError("IVF frame runs past the end of the file")
        out.append(data[pos:pos + size])
        pos += size
    return out
