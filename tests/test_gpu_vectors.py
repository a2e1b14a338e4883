"""GPU parity on the reference's 15 committed streams (tests/data of
safewords/rivet-vp9, tests/committed_vectors.rs:6-29): every frame of the GPU
reconstruction equals the oracle's after reconstruction and after the loop
filter, and every shown frame's MD5 equals the published one.  Through the
C ABI (vp9r_gpu_reconstruct / vp9r_gpu_loop_filter), and through the
decoder API (vp9r_decoder_*)."""
import os
import sys

import pytest

import gpu_lockstep
import oracle_lib
import rivet_vp9_amd as vp
import webm

pytestmark = pytest.mark.gpu

VECTORS = webm.vector_paths()


@pytest.fixture(scope="module")
def gpu():
    return vp.Gpu(0)


@pytest.mark.parametrize("path", VECTORS, ids=[os.path.basename(p) for p in VECTORS])
def test_lockstep_vector(gpu, path):
    print(f"lockstep {os.path.basename(path)}", flush=True)
    matched, expected, err = gpu_lockstep.run_vector(path, gpu)
    assert err is None, f"{os.path.basename(path)}: {matched}/{expected} then {err}"
    assert matched == expected


@pytest.mark.parametrize("path", VECTORS, ids=[os.path.basename(p) for p in VECTORS])
def test_decoder_api_vector(path):
    print(f"decoder {os.path.basename(path)}", flush=True)
    exp = webm.expected_md5s(path)
    d = vp.Decoder(0)
    got = []
    for pkt in webm.packets(path):
        for f in d.decode_all(pkt):
            got.append(oracle_lib.md5(f.packed()))
    assert got == exp
    assert d.frames_decoded() >= len(exp)
