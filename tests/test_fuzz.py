"""Malformed input (the reference's tests/fuzz.rs:28-64): arbitrary bytes,
and the committed streams with bits flipped, a packet cut and garbage
spliced in, through the product parser (one thread and four, which must
agree) and the oracle's reconstruction, built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/fuzz).  Every frame must end in a status;
the driver runs in its own process, so a crash fails the test instead of
ending pytest."""
import os
import struct
import subprocess

import rivet_vp9_amd as vp
import webm

FUZZ_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fuzz")
DRIVER = os.path.join(FUZZ_DIR, "_build", "parser_fuzz")


def _packets_file(path):
    streams = [list(webm.packets(p)) for p in webm.vector_paths()]
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(streams)))
        for s in streams:
            f.write(struct.pack("<I", len(s)))
            for p in s:
                f.write(struct.pack("<I", len(p)))
                f.write(p)


def test_malformed_input_ends_in_a_status(tmp_path):
    subprocess.run(["make", "-s", "-C", FUZZ_DIR], check=True)
    pk = os.path.join(tmp_path, "packets.bin")
    _packets_file(pk)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([DRIVER, pk, "20261204", "150"], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, f"driver exit {r.returncode}\n{r.stdout}\n{r.stderr[-4000:]}"
    assert "cases" in r.stdout
    # the mutations reach both outcomes
    n_ok, n_err = [int(w) for w in r.stdout.split() if w.isdigit()][1:3]
    assert n_ok > 0 and n_err > 0


def test_max_pixels_refuses_larger_frames():
    """Decoder::set_max_pixels (decoder/mod.rs:183-188): a frame of more luma
    samples than the limit is Error::Unsupported, and the parser keeps no
    state from it (the next key frame under the limit parses)."""
    path = os.path.join(webm.VECTORS, "vp90-2-02-size-66x66.webm")
    pk = list(webm.packets(path))
    p = vp.Parser()
    p.set_max_pixels(66 * 66 - 1)
    try:
        p.parse(vp.superframe_split(pk[0])[0])
        raise AssertionError("a 66x66 frame passed a limit of 66*66-1 samples")
    except vp.UnsupportedError:
        pass
    p.set_max_pixels(66 * 66)
    rec = p.parse(vp.superframe_split(pk[0])[0])
    assert (rec.hdr.width, rec.hdr.height) == (66, 66)
