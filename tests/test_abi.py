"""The C ABI: libvp9r.so loads without a GPU, exports every function
include/vp9r.h declares, and fails loudly (VP9R_ERR_DEVICE) instead of
falling back to the CPU when no GPU is usable; the record structures have the
layout the header gives them."""
import ctypes
import os
import re
import subprocess

import pytest

import rivet_vp9_amd as vp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vp9r.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(vp9r_\w+)\s*\(", text))
    return sorted(names)


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", vp.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("vp9r_decoder_decode", "vp9r_gpu_reconstruct", "vp9r_gpu_loop_filter",
                 "vp9r_parser_parse_frame", "vp9r_itx_add_batch", "vp9r_mc_batch", "vp9r_lf_edges_batch"):
        assert must in names


def test_library_exports_every_declared_function():
    missing = [n for n in declared_functions() if n not in exported_symbols()]
    assert not missing, f"declared but not exported: {missing}"


def test_every_declared_function_resolves():
    lib = vp.library()
    for n in declared_functions():
        assert getattr(lib, n) is not None


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_no_gpu_fails_loudly():
    with pytest.raises(vp.DeviceError):
        vp.Decoder(0)
    with pytest.raises(vp.DeviceError):
        vp.Gpu(0)


def test_record_layout_matches_the_header():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "vp9r.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(vp9r_frame_hdr), sizeof(vp9r_block), sizeof(vp9r_txb),
         sizeof(vp9r_mi), sizeof(vp9r_record), sizeof(vp9r_frame), sizeof(vp9r_mc_desc));
  printf("%zu %zu %zu\n", offsetof(vp9r_frame_hdr, seg_feature_data), offsetof(vp9r_block, mv),
         offsetof(vp9r_record, mi));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(vp9r_ref_plane), sizeof(vp9r_avg_desc), sizeof(vp9r_add_const_desc),
         sizeof(vp9r_intra_desc), offsetof(vp9r_intra_desc, mode), offsetof(vp9r_ref_plane, last_y));
  return 0;
}
"""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out]
    want = [ctypes.sizeof(vp.FrameHdr), ctypes.sizeof(vp.Block), ctypes.sizeof(vp.Txb), ctypes.sizeof(vp.Mi),
            ctypes.sizeof(vp.Record), ctypes.sizeof(vp.FrameDesc), ctypes.sizeof(vp.McDesc),
            vp.FrameHdr.seg_feature_data.offset, vp.Block.mv.offset, vp.Record.mi.offset,
            ctypes.sizeof(vp.RefPlane), ctypes.sizeof(vp.AvgDesc), ctypes.sizeof(vp.AddConstDesc),
            ctypes.sizeof(vp.IntraDesc), vp.IntraDesc.mode.offset, vp.RefPlane.last_y.offset]
    assert sizes == want


def test_superframe_split():
    # two frames of 3 and 2 bytes with a 1-byte-size index (Annex B)
    marker = 0xc0 | (0 << 3) | 1
    pkt = bytes([1, 2, 3, 4, 5, marker, 3, 2, marker])
    assert vp.superframe_split(pkt) == [bytes([1, 2, 3]), bytes([4, 5])]
    assert vp.superframe_split(b"\x01\x02") == [b"\x01\x02"]
