"""Test support: the first video track's frames of a WebM file, and the
packets / expected MD5s of a committed test vector (the reference's
tests/common/mod.rs:12-115, restated)."""
import os

VECTORS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors")


def _vint(data, pos, keep_marker):
    if pos >= len(data):
        return None
    first = data[pos]
    length = 8 - first.bit_length() + 1 if first else 9
    if length > 8 or pos + length > len(data):
        return None
    v = first if keep_marker else first & ((1 << (8 - length)) - 1)
    for i in range(1, length):
        v = (v << 8) | data[pos + i]
    return v, length


def _walk(data, out, track):
    pos = 0
    while pos < len(data):
        r = _vint(data, pos, True)
        if r is None:
            return
        eid, il = r
        r = _vint(data, pos + il, False)
        if r is None:
            return
        size, sl = r
        start = pos + il + sl
        unknown = size == (1 << (7 * sl)) - 1
        if not unknown and start + size > len(data):
            return  # an element larger than its parent: corrupt from here
        end = len(data) if unknown else start + size
        if eid in (0x18538067, 0x1F43B675, 0xA0):  # Segment, Cluster, BlockGroup
            _walk(data[start:end], out, track)
        elif eid in (0xA3, 0xA1):  # SimpleBlock, Block
            b = data[start:end]
            r = _vint(b, 0, False)
            if r is not None:
                tn, tl = r
                if track[0] is None:
                    track[0] = tn
                if tn == track[0] and len(b) >= tl + 3:
                    out.append(bytes(b[tl + 3:]))
        pos = end


def webm_frames(data):
    out = []
    _walk(data, out, [None])
    return out


def packets(path):
    import rivet_vp9_amd as vp
    data = open(path, "rb").read()
    if data[:4] == b"DKIF":
        return vp.ivf_frames(data)
    return webm_frames(data)


def expected_md5s(path):
    with open(path + ".md5") as f:
        return [l.split()[0] for l in f if l.strip()]


def vector_paths():
    return sorted(os.path.join(VECTORS, n) for n in os.listdir(VECTORS)
                  if n.endswith(".webm") or n.endswith(".ivf"))
