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

OK, ERR_BITSTREAM, ERR_UNSUPPORTED, ERR_INVALID_INPUT, ERR_DEVICE, AGAIN = 0, 1, 2, 3, 4, 5


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


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", u64), ("parse_threads", u32), ("max_batch", u32), ("max_inflight", u32),
                ("parse_ahead", u32)]


class ServerFrame(ctypes.Structure):
    _fields_ = [("stream", u32), ("index", u32), ("eos", u32), ("status", i32), ("latency_ms", ctypes.c_double),
                ("handle", u64), ("frame", FrameDesc)]


class McDesc(ctypes.Structure):
    _fields_ = [("x", i32), ("y", i32), ("x2", i32), ("y2", i32), ("xs", i16), ("ys", i16),
                ("w", u8), ("h", u8), ("filter", u8), ("ref_idx", i8), ("ref2_idx", i8),
                ("_pad", u8), ("out_off", u64)]


class RefPlane(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("stride", u32), ("last_x", i32), ("last_y", i32), ("_pad", u32)]


class AvgDesc(ctypes.Structure):
    _fields_ = [("a_off", u64), ("b_off", u64), ("dst_off", u64), ("w", u8), ("h", u8), ("_pad", u8 * 6)]


class AddConstDesc(ctypes.Structure):
    _fields_ = [("dst_off", u64), ("v", i32), ("n", u8), ("_pad", u8 * 3)]


class IntraDesc(ctypes.Structure):
    _fields_ = [("dst_off", u64), ("edge_off", u32), ("mode", u8), ("log2", u8), ("have_left", u8),
                ("have_above", u8)]


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
    lib.vp9r_parser_set_threads.argtypes = [P, u32]
    lib.vp9r_parser_parse_frame.argtypes = [P, ctypes.c_char_p, sz, ctypes.POINTER(Record)]
    lib.vp9r_packed_size.restype = sz
    lib.vp9r_packed_size.argtypes = [u32, u32, u32, u32, u32]
    PP = ctypes.POINTER(ctypes.c_void_p)
    lib.vp9r_gpu_new.argtypes = [ctypes.c_int, PP]
    lib.vp9r_gpu_free.argtypes = [P]
    lib.vp9r_gpu_frame_new.argtypes = [P, u32, u32, u32, u32, u32, PP]
    lib.vp9r_gpu_frame_free.argtypes = [P]
    lib.vp9r_gpu_frame_upload_u16.argtypes = [P, ctypes.c_int, V]
    lib.vp9r_gpu_frame_download_u16.argtypes = [P, ctypes.c_int, V]
    lib.vp9r_gpu_frame_geometry.argtypes = [P, ctypes.c_int, ctypes.POINTER(u32), ctypes.POINTER(u32)]
    lib.vp9r_gpu_reconstruct.argtypes = [P, ctypes.POINTER(Record), ctypes.c_void_p * 3, P]
    lib.vp9r_gpu_loop_filter.argtypes = [P, ctypes.POINTER(Record), P]
    lib.vp9r_gpu_download_packed.argtypes = [P, P, V]
    lib.vp9r_gpu_download_packed_async.argtypes = [P, P, V]
    lib.vp9r_gpu_frame_wait.argtypes = [P]
    lib.vp9r_host_alloc.argtypes = [sz, PP]
    lib.vp9r_host_free.argtypes = [V]
    lib.vp9r_gpu_synchronize.argtypes = [P]
    lib.vp9r_gpu_set_timing.argtypes = [P, ctypes.c_int]
    lib.vp9r_gpu_stage_times.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(u32)]
    lib.vp9r_itx_add_batch.argtypes = [P, u32, V, V, V, V, V, V, V, u32, u32, ctypes.c_int]
    lib.vp9r_mc_batch.argtypes = [P, u32, V, V, u32, u32, i32, i32, V, u32, u32]
    lib.vp9r_lf_edges_batch.argtypes = [P, u32, V, V, V, V, V, V, V, V, u32, u32, u32]
    lib.vp9r_mc_batch_planes.argtypes = [P, u32, V, V, u32, V, u32, u32]
    lib.vp9r_mc_batch_planes_noalloc.argtypes = [P, u32, V, V, u32, V, u32, u32]
    lib.vp9r_avg_batch.argtypes = [P, u32, V, V, V, V, u32, u32]
    lib.vp9r_add_const_batch.argtypes = [P, u32, V, V, u32]
    lib.vp9r_intra_predict_batch.argtypes = [P, u32, V, V, V, u32, u32]
    lib.vp9r_itx_add_host.argtypes = [P, V, u32, u32, ctypes.c_int, u32, u32, V, u32]
    lib.vp9r_intra_predict_host.argtypes = [P, u32, u32, V, V, ctypes.c_int, ctypes.c_int, u32, V, u32]
    lib.vp9r_gpu_prepare.argtypes = [P, ctypes.POINTER(Record), PP]
    lib.vp9r_prepared_free.argtypes = [P]
    lib.vp9r_gpu_run.argtypes = [P, P, ctypes.c_void_p * 3, P]
    lib.vp9r_gpu_autotune.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
    lib.vp9r_gpu_tuning.argtypes = [P, ctypes.POINTER(u32)]
    lib.vp9r_gpu_set_tuning.argtypes = [P, ctypes.POINTER(u32)]
    lib.vp9r_gpu_graph_new.argtypes = [P, P, ctypes.c_void_p * 3, P, PP]
    lib.vp9r_gpu_graph_update.argtypes = [P, ctypes.c_void_p * 3, P]
    lib.vp9r_gpu_graph_run.argtypes = [P]
    lib.vp9r_gpu_graph_nodes.restype = u32
    lib.vp9r_gpu_graph_nodes.argtypes = [P]
    lib.vp9r_gpu_graph_free.argtypes = [P]
    lib.vp9r_gpu_run_batch.argtypes = [P, u32, ctypes.POINTER(V), ctypes.POINTER(V), ctypes.POINTER(V)]
    lib.vp9r_gpu_filter_batch.argtypes = [P, u32, ctypes.POINTER(V), ctypes.POINTER(V)]
    lib.vp9r_gpu_prepared_stats.argtypes = [P, ctypes.POINTER(u64)]
    lib.vp9r_gpu_prepared_schedule.argtypes = [P, ctypes.POINTER(ctypes.c_uint32)]
    lib.vp9r_gpu_malloc.argtypes = [P, sz, PP]
    lib.vp9r_gpu_mfree.argtypes = [P, V]
    lib.vp9r_gpu_to_device.argtypes = [P, V, V, sz]
    lib.vp9r_gpu_to_host.argtypes = [P, V, V, sz]
    lib.vp9r_gpu_seam_ms.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
    lib.vp9r_gpu_frame_plane.restype = V
    lib.vp9r_gpu_frame_plane.argtypes = [P, ctypes.c_int]
    lib.vp9r_decoder_new.argtypes = [ctypes.c_int, PP]
    lib.vp9r_decoder_free.argtypes = [P]
    lib.vp9r_decoder_set_threads.argtypes = [P, u32]
    lib.vp9r_decoder_threads.restype = u32
    lib.vp9r_decoder_threads.argtypes = [P]
    lib.vp9r_decoder_set_max_pixels.argtypes = [P, u64]
    lib.vp9r_decoder_decode.argtypes = [P, ctypes.c_char_p, sz, ctypes.POINTER(u32)]
    lib.vp9r_decoder_frame.argtypes = [P, u32, ctypes.POINTER(FrameDesc)]
    lib.vp9r_decoder_frames_decoded.restype = u64
    lib.vp9r_decoder_frames_decoded.argtypes = [P]
    lib.vp9r_decoder_test_fail_stores.argtypes = [P, u32]
    lib.vp9r_server_new.argtypes = [ctypes.c_int, ctypes.POINTER(ServerConfig), PP]
    lib.vp9r_server_free.argtypes = [P]
    lib.vp9r_frame_store_bytes.restype = u64
    lib.vp9r_frame_store_bytes.argtypes = [u32] * 5
    lib.vp9r_server_open.argtypes = [P, u32, u32, u32, u32, u32, ctypes.POINTER(u32)]
    lib.vp9r_server_submit.argtypes = [P, u32, ctypes.c_char_p, sz]
    lib.vp9r_server_close.argtypes = [P, u32]
    lib.vp9r_server_poll.argtypes = [P, ctypes.POINTER(ServerFrame), ctypes.c_int]
    lib.vp9r_server_release.argtypes = [P, u64]
    lib.vp9r_server_stats.argtypes = [P, ctypes.POINTER(u64)]
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

    def set_threads(self, threads):
        """Threads that parse tile columns at once (the record is the same)."""
        library().vp9r_parser_set_threads(self._p, threads)

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


# transform launch shape flags (vp9r_gpu_set_tuning)
SHAPE_REG = 16
SHAPE_PERSISTENT = 32


class Gpu:
    """A GPU reconstruction context on one HIP device (vp9r_gpu_*)."""

    def __init__(self, device=0):
        self._g = ctypes.c_void_p()
        _check(library().vp9r_gpu_new(device, ctypes.byref(self._g)))

    def __del__(self):
        if getattr(self, "_g", None):
            library().vp9r_gpu_free(self._g)
            self._g = None

    @property
    def handle(self):
        return self._g

    def frame(self, width, height, bit_depth, ss_x, ss_y):
        return GpuFrame(self, width, height, bit_depth, ss_x, ss_y)

    @staticmethod
    def _refs(refs):
        return (ctypes.c_void_p * 3)(*[(r.handle if r is not None else None) for r in refs])

    TUNE_TIMES = 100

    def autotune(self):
        """Runs the launch-shape autotuner; returns (choices, candidate times
        in us: transforms [class][size][persistent][registers][waves 1, 2, 4],
        then MC [class][rows 32, 16]; -1 where a size lacks the shape)."""
        t = (ctypes.c_float * self.TUNE_TIMES)()
        _check(library().vp9r_gpu_autotune(self._g, t))
        return self.tuning(), list(t)

    @staticmethod
    def shape_name(code):
        """A transform launch shape code as text: waves, transposes, grid."""
        return (f"{code & 15}w" + ("+reg" if code & SHAPE_REG else "+lds") +
                ("+persistent" if code & SHAPE_PERSISTENT else ""))

    def tuning(self):
        out = (u32 * 11)()
        _check(library().vp9r_gpu_tuning(self._g, out))
        v = list(out)
        return {"itx_shape_8bit": v[0:4], "itx_shape_hbd": v[4:8], "mc_tile_rows_8bit": v[8],
                "mc_tile_rows_hbd": v[9], "tuned": bool(v[10])}

    def set_tuning(self, itx_shape_8bit, itx_shape_hbd, mc_rows_8bit, mc_rows_hbd):
        """Transform shapes per size 4x4 .. 32x32: waves per workgroup (1, 2
        or 4) | SHAPE_REG (register transposes, up to 16x16) |
        SHAPE_PERSISTENT; MC tile rows 16 or 32."""
        arr = (u32 * 10)(*(list(itx_shape_8bit) + list(itx_shape_hbd) + [mc_rows_8bit, mc_rows_hbd]))
        _check(library().vp9r_gpu_set_tuning(self._g, arr))

    def graph(self, prepared, refs, cur):
        """The launches of run(prepared, refs, cur) captured as a Graph."""
        return Graph(self, prepared, refs, cur)

    def reconstruct(self, rec, refs, cur):
        arr = (ctypes.c_void_p * 3)(*[(r.handle if r is not None else None) for r in refs])
        _check(library().vp9r_gpu_reconstruct(self._g, ctypes.byref(rec), arr, cur.handle))

    def loop_filter(self, rec, cur):
        _check(library().vp9r_gpu_loop_filter(self._g, ctypes.byref(rec), cur.handle))

    def synchronize(self):
        _check(library().vp9r_gpu_synchronize(self._g))

    def set_timing(self, on):
        library().vp9r_gpu_set_timing(self._g, 1 if on else 0)

    def stage_times(self):
        """(ms per stage summed over the frames timed since the last read:
        inter prediction, inter residual, intra, loop filter, whole; frames)."""
        out = (ctypes.c_float * 5)()
        n = u32()
        _check(library().vp9r_gpu_stage_times(self._g, out, ctypes.byref(n)))
        return list(out), n.value

    def prepare(self, rec):
        """The record's inputs copied to HBM (a Prepared), for run()."""
        return Prepared(self, rec)

    def run(self, prepared, refs, cur):
        arr = (ctypes.c_void_p * 3)(*[(r.handle if r is not None else None) for r in refs])
        _check(library().vp9r_gpu_run(self._g, prepared.handle, arr, cur.handle))

    def run_batch(self, jobs):
        """Runs independent frames together (vp9r_gpu_run_batch): `jobs` is a
        list of (prepared, refs, cur) as for run()."""
        n = len(jobs)
        ps = (ctypes.c_void_p * max(n, 1))(*[j[0].handle for j in jobs])
        rs = (ctypes.c_void_p * max(3 * n, 1))()
        for i, (_, refs, _) in enumerate(jobs):
            for k in range(3):
                r = refs[k] if refs else None
                rs[3 * i + k] = r.handle if r is not None else None
        cs = (ctypes.c_void_p * max(n, 1))(*[j[2].handle for j in jobs])
        _check(library().vp9r_gpu_run_batch(self._g, n, ps, rs, cs))

    def filter_batch(self, jobs):
        """The loop filter alone of reconstructed frames (vp9r_gpu_filter_batch):
        `jobs` is a list of (prepared, cur); returns the launches' device ms."""
        n = len(jobs)
        ps = (ctypes.c_void_p * max(n, 1))(*[j[0].handle for j in jobs])
        cs = (ctypes.c_void_p * max(n, 1))(*[j[1].handle for j in jobs])
        _check(library().vp9r_gpu_filter_batch(self._g, n, ps, cs))
        return self.seam_ms()

    def buffer(self, nbytes):
        return DevBuffer(self, nbytes)

    def upload(self, arr):
        """A device copy of a numpy array (a DevBuffer)."""
        import numpy as np
        a = np.ascontiguousarray(arr)
        b = DevBuffer(self, a.nbytes)
        _check(library().vp9r_gpu_to_device(self._g, b.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
        return b

    def seam_ms(self):
        """Device time (ms) of the last batched seam call's kernels."""
        ms = ctypes.c_float()
        _check(library().vp9r_gpu_seam_ms(self._g, ctypes.byref(ms)))
        return ms.value


class Prepared:
    """A record whose inputs (and work lists) are resident in HBM."""

    def __init__(self, gpu, rec):
        self.gpu = gpu
        self.hdr = FrameHdr()
        ctypes.memmove(ctypes.byref(self.hdr), ctypes.byref(rec.hdr), ctypes.sizeof(FrameHdr))
        self._p = ctypes.c_void_p()
        _check(library().vp9r_gpu_prepare(gpu.handle, ctypes.byref(rec), ctypes.byref(self._p)))

    def __del__(self):
        if getattr(self, "_p", None):
            library().vp9r_prepared_free(self._p)
            self._p = None

    @property
    def handle(self):
        return self._p

    def stats(self):
        out = (u64 * 12)()
        library().vp9r_gpu_prepared_stats(self._p, out)
        keys = ("mc_tiles", "itx4", "itx8", "itx16", "itx32", "intra_groups", "intra_levels", "lf_levels",
                "mc_bytes", "itx_bytes", "intra_bytes", "lf_bytes")
        d = dict(zip(keys, list(out)))
        sched = (ctypes.c_uint32 * 4)()
        library().vp9r_gpu_prepared_schedule(self._p, sched)
        d["wave_launches"], d["lf_tail_launches"], d["lf_tail_levels"] = sched[0], sched[1], sched[3]
        return d


class Graph:
    """A prepared frame's launches captured as a HIP graph
    (vp9r_gpu_graph_*): run() replays them for the frame stores it was made
    for; update() re-targets it to other stores."""

    def __init__(self, gpu, prepared, refs, cur):
        self.gpu, self.prepared = gpu, prepared
        self._keep = (list(refs), cur)
        self._h = ctypes.c_void_p()
        _check(library().vp9r_gpu_graph_new(gpu.handle, prepared.handle, Gpu._refs(refs), cur.handle,
                                            ctypes.byref(self._h)))

    def update(self, refs, cur):
        _check(library().vp9r_gpu_graph_update(self._h, Gpu._refs(refs), cur.handle))
        self._keep = (list(refs), cur)

    def run(self):
        _check(library().vp9r_gpu_graph_run(self._h))

    def nodes(self):
        return library().vp9r_gpu_graph_nodes(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            library().vp9r_gpu_graph_free(self._h)
            self._h = None


class DevBuffer:
    """Device memory from vp9r_gpu_malloc."""

    def __init__(self, gpu, nbytes):
        self.gpu = gpu
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        _check(library().vp9r_gpu_malloc(gpu.handle, nbytes, ctypes.byref(p)))
        self.ptr = p
        # the context stays valid while this buffer holds its reference
        self._g = ctypes.c_void_p(gpu.handle.value)

    def __del__(self):
        if getattr(self, "ptr", None):
            library().vp9r_gpu_mfree(self._g, self.ptr)
            self.ptr = None

    def read(self, dtype, count):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        _check(library().vp9r_gpu_to_host(self._g, out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                          out.nbytes))
        return out

    def write(self, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        _check(library().vp9r_gpu_to_device(self._g, self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                            a.nbytes))


class HostBuffer:
    """Page-locked host memory (vp9r_host_alloc)."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        _check(library().vp9r_host_alloc(nbytes, ctypes.byref(p)))
        self.ptr = p

    def __del__(self):
        if getattr(self, "ptr", None):
            library().vp9r_host_free(self.ptr)
            self.ptr = None

    def bytes(self, n=None):
        return ctypes.string_at(self.ptr, self.nbytes if n is None else n)


class GpuFrame:
    """An HBM-resident frame store."""

    def __init__(self, gpu, width, height, bit_depth, ss_x, ss_y):
        self.gpu = gpu
        self.width, self.height, self.bit_depth, self.ss_x, self.ss_y = width, height, bit_depth, ss_x, ss_y
        self._f = ctypes.c_void_p()
        _check(library().vp9r_gpu_frame_new(gpu.handle, width, height, bit_depth, ss_x, ss_y,
                                            ctypes.byref(self._f)))
        # the context stays valid while this frame store holds its reference
        self._g = ctypes.c_void_p(gpu.handle.value)

    def __del__(self):
        if getattr(self, "_f", None):
            library().vp9r_gpu_frame_free(self._f)
            self._f = None

    @property
    def handle(self):
        return self._f

    def plane_ptr(self, plane):
        return library().vp9r_gpu_frame_plane(self._f, plane)

    def geometry(self, plane):
        s, r = u32(), u32()
        library().vp9r_gpu_frame_geometry(self._f, plane, ctypes.byref(s), ctypes.byref(r))
        return s.value, r.value

    def packed(self):
        n = library().vp9r_packed_size(self.width, self.height, self.bit_depth, self.ss_x, self.ss_y)
        buf = (ctypes.c_uint8 * n)()
        _check(library().vp9r_gpu_download_packed(self._g, self._f, buf))
        return bytes(buf)

    def packed_async(self, host):
        """Queues the packed download into `host` (a HostBuffer) on the copy
        stream; wait() ends it."""
        n = library().vp9r_packed_size(self.width, self.height, self.bit_depth, self.ss_x, self.ss_y)
        assert host.nbytes >= n
        _check(library().vp9r_gpu_download_packed_async(self._g, self._f, host.ptr))

    def wait(self):
        _check(library().vp9r_gpu_frame_wait(self._f))

    def download_plane(self, plane):
        """The whole plane (stride x rows) as a list-like array of u16."""
        import numpy as np
        s, r = self.geometry(plane)
        out = np.empty((r, s), dtype=np.uint16)
        _check(library().vp9r_gpu_frame_download_u16(self._f, plane, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def upload_plane(self, plane, arr):
        import numpy as np
        s, r = self.geometry(plane)
        a = np.ascontiguousarray(arr, dtype=np.uint16)
        assert a.shape == (r, s)
        _check(library().vp9r_gpu_frame_upload_u16(self._f, plane, a.ctypes.data_as(ctypes.c_void_p)))


class Frame:
    """A decoded frame (src/frame.rs:112-137): packed Y, U, V."""

    def __init__(self, desc):
        self.width, self.height, self.bit_depth = desc.width, desc.height, desc.bit_depth
        self.ss_x, self.ss_y = desc.ss_x, desc.ss_y
        self.color_space, self.full_range = desc.color_space, bool(desc.full_range)
        self.render_width, self.render_height = desc.render_width, desc.render_height
        self.data = ctypes.string_at(desc.data, desc.size)
        self.planes = [(desc.plane_offset[i], desc.plane_width[i], desc.plane_height[i]) for i in range(3)]

    def packed(self):
        return self.data

    def bytes_per_sample(self):
        return 2 if self.bit_depth > 8 else 1

    def plane(self, i):
        off, w, h = self.planes[i]
        return self.data[off:off + w * h * self.bytes_per_sample()]


class Decoder:
    """The crate's Decoder (src/decoder/mod.rs:150-216) on the GPU backend."""

    def __init__(self, device=0):
        self._d = ctypes.c_void_p()
        _check(library().vp9r_decoder_new(device, ctypes.byref(self._d)))

    def __del__(self):
        if getattr(self, "_d", None):
            library().vp9r_decoder_free(self._d)
            self._d = None

    def set_threads(self, threads):
        library().vp9r_decoder_set_threads(self._d, threads)

    def threads(self):
        return library().vp9r_decoder_threads(self._d)

    def set_max_pixels(self, pixels):
        library().vp9r_decoder_set_max_pixels(self._d, pixels)

    def decode_all(self, packet):
        n = u32()
        _check(library().vp9r_decoder_decode(self._d, packet, len(packet), ctypes.byref(n)))
        out = []
        for i in range(n.value):
            d = FrameDesc()
            _check(library().vp9r_decoder_frame(self._d, i, ctypes.byref(d)))
            out.append(Frame(d))
        return out

    def decode(self, packet):
        frames = self.decode_all(packet)
        return frames[-1] if frames else None

    def frames_decoded(self):
        return library().vp9r_decoder_frames_decoded(self._d)

    def test_fail_stores(self, n):
        """Test hook: refuse the next n frame stores (VP9R_ERR_DEVICE)."""
        library().vp9r_decoder_test_fail_stores(self._d, n)


class ServedFrame:
    """A frame (or end of stream) handed out by Server.poll().  The packed
    data stays in the server's pinned memory until release()."""

    def __init__(self, server, rec):
        self._server = server
        self.stream, self.index, self.eos = rec.stream, rec.index, bool(rec.eos)
        self.status, self.latency_ms, self.handle = rec.status, rec.latency_ms, rec.handle
        self.desc = rec.frame
        self.size = rec.frame.size if not self.eos else 0

    def view(self):
        """The packed frame as a memoryview over pinned memory (no copy)."""
        arr = (ctypes.c_uint8 * self.size).from_address(ctypes.addressof(self.desc.data.contents))
        return memoryview(arr).cast("B")

    def frame(self):
        """A copy as a Frame."""
        return Frame(self.desc)

    def release(self):
        if self.handle:
            library().vp9r_server_release(self._server._s, self.handle)
            self.handle = 0


class Server:
    """Many streams on one GPU (vp9r_server_*): admission by HBM footprint,
    parse workers, cross-stream batches, per-frame latency."""

    def __init__(self, device=0, hbm_budget=0, parse_threads=0, max_batch=0, max_inflight=0, parse_ahead=0):
        cfg = ServerConfig(hbm_budget, parse_threads, max_batch, max_inflight, parse_ahead)
        self._s = ctypes.c_void_p()
        _check(library().vp9r_server_new(device, ctypes.byref(cfg), ctypes.byref(self._s)))

    def close_server(self):
        if getattr(self, "_s", None):
            library().vp9r_server_free(self._s)
            self._s = None

    def __del__(self):
        self.close_server()

    @staticmethod
    def footprint(width, height, bit_depth=8, ss_x=1, ss_y=1):
        return 10 * library().vp9r_frame_store_bytes(width, height, bit_depth, ss_x, ss_y)

    def open(self, max_width, max_height, bit_depth=8, ss_x=1, ss_y=1):
        sid = u32()
        _check(library().vp9r_server_open(self._s, max_width, max_height, bit_depth, ss_x, ss_y, ctypes.byref(sid)))
        return sid.value

    def submit(self, stream, packet):
        _check(library().vp9r_server_submit(self._s, stream, packet, len(packet)))

    def close(self, stream):
        _check(library().vp9r_server_close(self._s, stream))

    def poll(self, timeout_ms=-1):
        """The next ServedFrame, or None when none came within timeout_ms."""
        rec = ServerFrame()
        st = library().vp9r_server_poll(self._s, ctypes.byref(rec), timeout_ms)
        if st == AGAIN:
            return None
        _check(st)
        return ServedFrame(self, rec)

    def stats(self):
        out = (u64 * 8)()
        _check(library().vp9r_server_stats(self._s, out))
        keys = ["frames_shown", "frames_decoded", "batches", "max_batch", "streams_admitted", "streams_waiting",
                "store_bytes", "hbm_budget"]
        return dict(zip(keys, list(out)))
