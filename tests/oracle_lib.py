"""ctypes access to the CPU oracle (oracle/_build/liboracle.so) and a CPU
decoder built from the product parser plus the oracle's reconstruction: the
checker every GPU result is compared with.  Test infrastructure only."""
import ctypes
import hashlib
import os
import subprocess

import numpy as np
import rivet_vp9_amd as vp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")


class Edges(ctypes.Structure):
    _fields_ = [("fs", (ctypes.c_int8 * 16) * 16), ("limit", (ctypes.c_uint8 * 16) * 16),
                ("blimit", (ctypes.c_uint8 * 16) * 16), ("thresh", (ctypes.c_uint8 * 16) * 16)]


class OrcFrame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("bit_depth", ctypes.c_uint32),
                ("ss_x", ctypes.c_uint32), ("ss_y", ctypes.c_uint32),
                ("stride", ctypes.c_uint32 * 3), ("rows", ctypes.c_uint32 * 3),
                ("plane", ctypes.POINTER(ctypes.c_uint16) * 3)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    L = ctypes.CDLL(ORACLE_LIB)
    P = ctypes.c_void_p
    FP = ctypes.POINTER(OrcFrame)
    L.orc_frame_new.restype = FP
    L.orc_frame_new.argtypes = [ctypes.c_uint32] * 5
    L.orc_frame_free.argtypes = [FP]
    L.orc_reconstruct.argtypes = [ctypes.POINTER(vp.Record), ctypes.POINTER(FP), FP]
    L.orc_loop_filter.argtypes = [ctypes.POINTER(vp.Record), FP]
    L.orc_reconstruct_mt.argtypes = [ctypes.POINTER(vp.Record), ctypes.POINTER(FP), FP, ctypes.c_int]
    L.orc_loop_filter_mt.argtypes = [ctypes.POINTER(vp.Record), FP, ctypes.c_int]
    L.orc_packed_size.restype = ctypes.c_size_t
    L.orc_packed_size.argtypes = [FP]
    L.orc_pack.argtypes = [FP, P]
    L.orc_itx_add.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              P, ctypes.c_size_t]
    L.orc_itx_spec_reconstruct.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int]
    L.orc_spec_idct.argtypes = [P, ctypes.c_int]
    L.orc_spec_iadst.argtypes = [P, ctypes.c_int]
    L.orc_inter_predict.argtypes = [P, ctypes.c_size_t] + [ctypes.c_int] * 10 + [P, ctypes.c_size_t]
    L.orc_avg.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t]
    L.orc_itx_add_batch.argtypes = [ctypes.c_uint32] + [P] * 7 + [ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    L.orc_mc_batch.argtypes = [ctypes.c_uint32, P, P, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, P,
                               ctypes.c_size_t, ctypes.c_int]
    L.orc_add_const.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    L.orc_intra_predict.argtypes = [ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, P, ctypes.c_size_t]
    L.orc_lf_filter.argtypes = [P, ctypes.c_ssize_t, ctypes.c_ssize_t] + [ctypes.c_int] * 5
    L.orc_lf_filter_edges.argtypes = [P, ctypes.c_size_t] + [ctypes.c_int] * 6 + [ctypes.POINTER(Edges),
                                                                                  ctypes.c_int]
    L.orc_lf_superblock.restype = ctypes.c_int
    L.orc_lf_superblock.argtypes = [ctypes.POINTER(vp.Record), ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.POINTER(Edges), ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int)]
    _lib = L
    return L


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Frame:
    """An oracle frame store (u16 planes of whole superblocks)."""

    def __init__(self, w, h, bd, ssx, ssy):
        self.f = lib().orc_frame_new(w, h, bd, ssx, ssy)
        c = self.f.contents
        self.width, self.height, self.bit_depth, self.ss_x, self.ss_y = w, h, bd, ssx, ssy

    def __del__(self):
        if getattr(self, "f", None):
            lib().orc_frame_free(self.f)
            self.f = None

    def plane(self, p):
        c = self.f.contents
        return np.ctypeslib.as_array(c.plane[p], shape=(c.rows[p], c.stride[p]))

    def packed(self):
        n = lib().orc_packed_size(self.f)
        out = np.empty(n, dtype=np.uint8)
        lib().orc_pack(self.f, ptr(out))
        return out.tobytes()


class CpuDecoder:
    """The scalar CPU decoder: product parser + oracle reconstruction and loop
    filter, with the reference's 8 slots (decoder/mod.rs:218-422).  With
    threads > 1: tile columns parsed and reconstructed in parallel and the
    loop filter's superblock rows as a wavefront (the reference's threading;
    the same frames)."""

    def __init__(self, threads=1):
        self.parser = vp.Parser()
        self.threads = threads
        if threads > 1:
            self.parser.set_threads(threads)
        self.slots = [None] * 8
        self.frames_decoded = 0

    def decode_frame(self, frame, on_record=None):
        rec = self.parser.parse(frame)
        h = rec.hdr
        if h.show_existing_frame:
            self.frames_decoded += 1
            return self.slots[h.frame_to_show_map_idx]
        cur = Frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        refs = (ctypes.POINTER(OrcFrame) * 3)()
        if not h.frame_is_intra:
            for i in range(3):
                refs[i] = self.slots[h.ref_frame_idx[i]].f
        if on_record is not None:
            on_record(rec, [None if h.frame_is_intra else self.slots[h.ref_frame_idx[i]] for i in range(3)])
        if self.threads > 1:
            st = lib().orc_reconstruct_mt(ctypes.byref(rec), refs, cur.f, self.threads)
        else:
            st = lib().orc_reconstruct(ctypes.byref(rec), refs, cur.f)
        if st:
            raise vp.BitstreamError("oracle reconstruction failed")
        if self.threads > 1:
            lib().orc_loop_filter_mt(ctypes.byref(rec), cur.f, self.threads)
        else:
            lib().orc_loop_filter(ctypes.byref(rec), cur.f)
        for i in range(8):
            if (h.refresh_frame_flags >> i) & 1:
                self.slots[i] = cur
        self.frames_decoded += 1
        return cur if h.show_frame else None

    def decode(self, packet, on_record=None):
        out = []
        for f in vp.superframe_split(packet):
            r = self.decode_frame(f, on_record)
            if r is not None:
                out.append(r)
        return out


def md5(b):
    return hashlib.md5(b).hexdigest()
