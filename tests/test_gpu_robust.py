"""Failure handling on the GPU path.

- A reconstruction that fails after its frame was parsed (the frame store
  refused by the decoder's test hook) leaves the decoder and the parser
  agreeing that the frame's slots are empty: a later show_existing_frame of
  such a slot, or an inter frame predicting from it, is an error, not a
  crash, and the stream decodes again from the next key frame
  (the reference: show_existing_frame of an empty slot is
  Error::Bitstream, src/decoder/mod.rs:224-231).
- The decoder gives the same frames with one thread and with several.
- vp9r_gpu_prepare checks every record it is given: records with corrupted
  fields are refused with a status (never a crash).  Only prepared, not
  run."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib
import rivet_vp9_amd as vp
import streamgen
import webm

pytestmark = pytest.mark.gpu


def frame_headers(frames):
    p = vp.Parser()
    out = []
    for f in frames:
        for sub in vp.superframe_split(f):
            h = p.parse(sub).hdr
            out.append((h.show_existing_frame, h.frame_to_show_map_idx, h.refresh_frame_flags, h.frame_is_intra,
                        tuple(h.ref_frame_idx)))
    return out


def test_failed_store_then_show_existing_is_an_error():
    path = os.path.join(webm.VECTORS, "vp90-2-10-show-existing-frame2.webm")
    pk = list(webm.packets(path))
    flat = [sub for p in pk for sub in vp.superframe_split(p)]
    hdrs = frame_headers(pk)
    # a frame whose refreshed slot is shown later by show_existing_frame,
    # before any refresh of that slot
    victim = None
    for k, (se, _, rf, _, _) in enumerate(hdrs):
        if se or not rf:
            continue
        for j in range(k + 1, len(hdrs)):
            se2, idx2, rf2, _, _ = hdrs[j]
            if se2 and (rf >> idx2) & 1:
                victim = (k, j)
                break
            if rf2 & rf:
                break
        if victim:
            break
    assert victim, "no frame refreshes a slot that is shown later"
    k, j = victim
    print(f"store refused at frame {k}; show_existing at frame {j}", flush=True)
    dec = vp.Decoder(0)
    errors = {}
    for i, f in enumerate(flat):
        if i == k:
            dec.test_fail_stores(1)
        try:
            dec.decode_all(f)
        except vp.Vp9Error as e:
            errors[i] = type(e)
    assert errors.get(k) is vp.DeviceError
    assert errors.get(j) is vp.BitstreamError
    # the next key frame after j decodes, and so does everything after it
    later_key = [i for i in range(j + 1, len(hdrs)) if hdrs[i][3] and not hdrs[i][0]]
    if later_key:
        assert all(i not in errors for i in range(later_key[0], len(flat)))


def test_decoder_threads_give_the_same_frames():
    frames = streamgen.generate(1280, 720, 4, seed=31, tile_cols_log2=2, content=1, intra_percent=10)
    ref = oracle_lib.CpuDecoder()
    want = [oracle_lib.md5(o.packed()) for f in frames for o in ref.decode(f)]
    for t in (1, 4):
        dec = vp.Decoder(0)
        dec.set_threads(t)
        assert dec.threads() == t
        got = [oracle_lib.md5(fr.packed()) for f in frames for fr in dec.decode_all(f)]
        assert got == want, f"threads {t}"


class OwnedRecord:
    """A parsed record copied into arrays this object owns (so it can be
    corrupted), with a Record pointing at them."""

    def __init__(self, rec):
        h = rec.hdr
        n_mi = h.mi_rows * h.mi_cols
        self.blocks = (vp.Block * max(1, rec.n_blocks))()
        ctypes.memmove(self.blocks, rec.blocks, ctypes.sizeof(vp.Block) * rec.n_blocks)
        self.txbs = (vp.Txb * max(1, rec.n_txbs))()
        ctypes.memmove(self.txbs, rec.txbs, ctypes.sizeof(vp.Txb) * rec.n_txbs)
        self.cpos = (ctypes.c_uint16 * max(1, rec.n_coefs))()
        ctypes.memmove(self.cpos, rec.coef_pos, 2 * rec.n_coefs)
        self.cval = (ctypes.c_int32 * max(1, rec.n_coefs))()
        ctypes.memmove(self.cval, rec.coef_val, 4 * rec.n_coefs)
        self.mi = (vp.Mi * max(1, n_mi))()
        ctypes.memmove(self.mi, rec.mi, ctypes.sizeof(vp.Mi) * n_mi)
        self.rec = vp.Record()
        self.rec.hdr = h
        self.rec.blocks = ctypes.cast(self.blocks, type(self.rec.blocks))
        self.rec.n_blocks = rec.n_blocks
        self.rec.txbs = ctypes.cast(self.txbs, type(self.rec.txbs))
        self.rec.n_txbs = rec.n_txbs
        self.rec.coef_pos = ctypes.cast(self.cpos, type(self.rec.coef_pos))
        self.rec.coef_val = ctypes.cast(self.cval, type(self.rec.coef_val))
        self.rec.n_coefs = rec.n_coefs
        self.rec.mi = ctypes.cast(self.mi, type(self.rec.mi))


def corrupt(o, rng):
    """One random field of the record set to a random value."""
    r = o.rec
    kind = int(rng.integers(0, 8))
    if kind == 0 and r.n_blocks:
        b = o.blocks[int(rng.integers(0, r.n_blocks))]
        field = ["mi_row", "mi_col", "mi_size", "tx_size", "y_mode", "uv_mode", "interp_filter",
                 "segment_id"][int(rng.integers(0, 8))]
        setattr(b, field, int(rng.integers(0, 1 << 16 if field in ("mi_row", "mi_col") else 256)))
    elif kind == 1 and r.n_blocks:
        b = o.blocks[int(rng.integers(0, r.n_blocks))]
        b.ref_frame[int(rng.integers(0, 2))] = int(rng.integers(-128, 128))
        b.is_inter = int(rng.integers(0, 2))
    elif kind in (2, 3) and r.n_txbs:
        t = o.txbs[int(rng.integers(0, r.n_txbs))]
        field = ["x", "y", "plane", "tx_size", "tx_type", "eob", "ncoef", "coef_start", "block"][
            int(rng.integers(0, 9))]
        big = {"x": 1 << 16, "y": 1 << 16, "eob": 1 << 16, "ncoef": 1 << 16, "coef_start": 1 << 32,
               "block": 1 << 32}.get(field, 256)
        setattr(t, field, int(rng.integers(0, big)))
    elif kind == 4:
        r.n_coefs = int(rng.integers(0, r.n_coefs + 1))
    elif kind == 5:
        r.n_txbs = int(rng.integers(0, r.n_txbs + 1))
    elif kind == 6 and r.n_blocks:
        r.hdr.width = int(rng.integers(1, 4096))
    else:
        r.hdr.ref_width[int(rng.integers(0, 3))] = int(rng.integers(0, 8192))


def corrupted_prepare_outcomes(cases, seed, prepare=True):
    """Prepares `cases` corrupted copies of a stream's records; returns the
    count of each outcome ("ok" or the error class)."""
    frames = streamgen.generate(352, 288, 3, seed=seed, content=0, intra_percent=20)
    p = vp.Parser()
    recs = [OwnedRecord(p.parse(f)) for f in frames]
    g = vp.Gpu(0) if prepare else None
    rng = np.random.default_rng(seed)
    outcomes = {}
    for case in range(cases):
        o = OwnedRecord(recs[case % len(recs)].rec)
        for _ in range(int(rng.integers(1, 4))):
            corrupt(o, rng)
        if not prepare:
            continue
        try:
            vp.Prepared(g, o.rec)
            outcomes["ok"] = outcomes.get("ok", 0) + 1
        except vp.Vp9Error as e:
            outcomes[type(e).__name__] = outcomes.get(type(e).__name__, 0) + 1
    return outcomes


def test_prepare_refuses_corrupted_records():
    # in a child process: a crash fails this test instead of ending the run
    import json
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    code = ("import sys, json; sys.path.insert(0, %r); import conftest, test_gpu_robust as t; "
            "print(json.dumps(t.corrupted_prepare_outcomes(300, 41)))" % here)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, f"exit {r.returncode}\n{r.stdout}\n{r.stderr[-3000:]}"
    outcomes = json.loads(r.stdout.strip().splitlines()[-1])
    print(outcomes, flush=True)
    assert outcomes.get("InvalidInputError", 0) > 0
