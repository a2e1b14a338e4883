"""Test support: decode a stream on the GPU and with the CPU oracle in
lockstep, from the same parsed records, comparing every frame after
reconstruction and again after the loop filter."""
import numpy as np

import oracle_lib
import rivet_vp9_amd as vp
import webm


def plane_region(h, plane):
    sx = h.ss_x if plane else 0
    sy = h.ss_y if plane else 0
    return (h.mi_rows * 8) >> sy, (h.mi_cols * 8) >> sx


def first_diff(a, b):
    idx = np.argwhere(a != b)
    if len(idx) == 0:
        return None
    y, x = idx[0]
    return int(y), int(x), int(a[y, x]), int(b[y, x]), len(idx)


class Lockstep:
    """GPU frame stores in 8 slots beside the oracle decoder's."""

    def __init__(self, gpu):
        self.gpu = gpu
        self.cpu = oracle_lib.CpuDecoder()
        self.slots = [None] * 8
        self.report = []

    def compare(self, rec, gf, of, stage):
        h = rec.hdr
        for p in range(3):
            rows, cols = plane_region(h, p)
            g = gf.download_plane(p)[:rows, :cols]
            o = of.plane(p)[:rows, :cols]
            d = first_diff(g, o)
            if d is not None:
                return f"{stage} plane {p}: first diff at (y={d[0]}, x={d[1]}) gpu={d[2]} oracle={d[3]}, {d[4]} samples differ"
        return None

    def decode_frame(self, frame):
        rec = self.cpu.parser.parse(frame)
        h = rec.hdr
        if h.show_existing_frame:
            self.cpu.frames_decoded += 1
            return self.slots[h.frame_to_show_map_idx], self.cpu.slots[h.frame_to_show_map_idx], None
        import ctypes
        cur_o = oracle_lib.Frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        refs_o = (ctypes.POINTER(oracle_lib.OrcFrame) * 3)()
        refs_g = [None, None, None]
        if not h.frame_is_intra:
            for i in range(3):
                refs_o[i] = self.cpu.slots[h.ref_frame_idx[i]].f
                refs_g[i] = self.slots[h.ref_frame_idx[i]]
        lib = oracle_lib.lib()
        assert lib.orc_reconstruct(ctypes.byref(rec), refs_o, cur_o.f) == 0
        cur_g = self.gpu.frame(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)
        self.gpu.reconstruct(rec, refs_g, cur_g)
        self.gpu.synchronize()
        err = self.compare(rec, cur_g, cur_o, "reconstruction")
        lib.orc_loop_filter(ctypes.byref(rec), cur_o.f)
        self.gpu.loop_filter(rec, cur_g)
        self.gpu.synchronize()
        if err is None:
            err = self.compare(rec, cur_g, cur_o, "loop filter")
        for i in range(8):
            if (h.refresh_frame_flags >> i) & 1:
                self.slots[i] = cur_g
                self.cpu.slots[i] = cur_o
        self.cpu.frames_decoded += 1
        if h.show_frame:
            return cur_g, cur_o, err
        return None, None, err


def run_vector(path, gpu):
    """Returns (matched MD5s, expected count, first error message or None)."""
    exp = webm.expected_md5s(path)
    ls = Lockstep(gpu)
    shown = 0
    matched = 0
    for pi, pkt in enumerate(webm.packets(path)):
        for f in vp.superframe_split(pkt):
            g, o, err = ls.decode_frame(f)
            if err is not None:
                return matched, len(exp), f"packet {pi}: {err}"
            if g is None:
                continue
            gm = oracle_lib.md5(g.packed())
            om = oracle_lib.md5(o.packed())
            if gm != om:
                return matched, len(exp), f"packet {pi}: packed frame differs from the oracle's"
            if shown < len(exp) and gm == exp[shown]:
                matched += 1
            else:
                return matched, len(exp), f"frame {shown}: md5 {gm} != expected"
            shown += 1
    return matched, len(exp), None


def run_frames(frames, gpu):
    """Lockstep over a list of frames (no published MD5s): returns
    (frames compared, first error message or None)."""
    ls = Lockstep(gpu)
    n = 0
    for i, f in enumerate(frames):
        g, o, err = ls.decode_frame(f)
        if err is not None:
            return n, f"frame {i}: {err}"
        if g is not None and oracle_lib.md5(g.packed()) != oracle_lib.md5(o.packed()):
            return n, f"frame {i}: packed frame differs from the oracle's"
        n += 1
    return n, None
