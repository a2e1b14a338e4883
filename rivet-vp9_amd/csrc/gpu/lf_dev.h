// Loop filter device code shared by the per-level kernel (k_lf, lf.hip) and
// the fused wavefront kernel (k_wave, intra.hip): the filters (src/dsp/lf.rs
// 24-143), the unit region staged in LDS, and one plane of one work item of
// a level (lf_unit).  The order of the work items is set out in lf.hip.
#pragma once
#include <type_traits>

#include "dev_common.h"
#include "launch.h"
#include "vp9_consts.h"

namespace vp9g {

// LDS pitch of a unit's region in samples (8 + 128 columns + slack, an odd
// number of dwords): samples are held in LDS as bytes at 8 bits, as u16
// above.
template <class T>
struct LfPitch {
  static constexpr int BP = sizeof(T) == 1 ? 140 : 138;
};

// |a - b| of two samples (< 2^16) in one instruction (v_sad_u16).
DEV uint32_t adiff(int a, int b) { return __builtin_amdgcn_sad_u16((uint32_t)a, (uint32_t)b, 0u); }
DEV int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }  // lo <= hi

// The filters of one edge position, on samples reached through `s`: s(k)
// is the sample k positions across the edge from the first q sample
// (k = -8 .. 7), a reference into LDS or into a register array.

// The wide filters (8.8.5.3): output i in [-n, n) is
// Round2(s[i] + T[i], LOG2) with the window sum
// T[i] = sum_{j=-n..n} s[clamp(i + j, -n - 1, n)] and n = 2^(LOG2-1) - 1.
// T slides: T[i + 1] = T[i] - s[clamp(i - n)] + s[clamp(i + n + 1)].  Every
// T is formed from the inputs before any output is written, so each output
// replaces its input in place.
template <int LOG2, class S>
DEV void lf_wide(S s) {
  constexpr int n = (1 << (LOG2 - 1)) - 1;
  auto at = [&](int k) -> int { return s(k < -(n + 1) ? -(n + 1) : (k > n ? n : k)); };
  int t[2 * n];
  // T[-n]: s[-n - 1] n times (the clamped positions) and s[-n] .. s[0]
  // once; a 24-bit multiply (full rate) for the n copies
  int tt = __mul24(at(-(n + 1)), n);
#pragma unroll
  for (int k = -n; k <= 0; k++) tt += at(k);
#pragma unroll
  for (int i = -n; i < n; i++) {
    t[i + n] = tt;
    tt += at(i + n + 1) - at(i - n);
  }
#pragma unroll
  for (int i = -n; i < n; i++) s(i) = (s(i) + t[i + n] + (1 << (LOG2 - 1))) >> LOG2;
}

// lf::filter (src/dsp/lf.rs:24-86): the mask, then the narrow filter, or
// the 8- or 16-wide filter where the samples are flat.  The mask and
// flatness tests are maxima of absolute differences against one limit
// each (the reference's chains of comparisons, without branches).  The
// three filters are exclusive per lane and each rewrites its samples in
// place: three independent conditional blocks, so that no sample is copied
// to merge the paths.
// fs >= 0: callers filter only where the decision has a filter.
template <class S>
DEV void lf_apply(S s, int fs, int limit, int blimit, int thresh, int bd) {
  const int q0 = s(0), q1 = s(1), q2 = s(2), q3 = s(3);
  const int p0 = s(-1), p1 = s(-2), p2 = s(-3), p3 = s(-4);
  const int sh = bd - 8;
  const uint32_t d10 = adiff(p1, p0), e10 = adiff(q1, q0);
  const uint32_t m = max(max(max(adiff(p3, p2), adiff(p2, p1)), d10), max(max(e10, adiff(q2, q1)), adiff(q3, q2)));
  const uint32_t edge = adiff(p0, q0) * 2 + (adiff(p1, q1) >> 1);
  const bool on = (m <= (uint32_t)(limit << sh)) & (edge <= (uint32_t)(blimit << sh));
  const bool hev = max(d10, e10) > (uint32_t)(thresh << sh);
  const uint32_t one = 1u << sh;
  // the flatness maximum for every lane (no branch around it: most waves
  // have a lane that needs it)
  const uint32_t fl =
      max(max(max(d10, e10), max(adiff(p2, p0), adiff(q2, q0))), max(adiff(p3, p0), adiff(q3, q0)));
  const bool flat = on & (fs >= 1) & (fl <= one);
  bool flat2 = false;
  if (flat && fs >= 2) {
    uint32_t f2 = 0;
#pragma unroll
    for (int k = 4; k < 8; k++) f2 = max(f2, max(adiff(s(-(k + 1)), p0), adiff(s(k), q0)));
    flat2 = f2 <= one;
  }
  if (on && !flat) {
    // narrow filter (8.8.5.2).  The reference works on samples offset by
    // 0x80 << (bd - 8) into [lo, hi]; differences of two samples lose the
    // offset, and clamp(x - off, lo, hi) + off = clamp(x, 0, max), so the
    // outputs are formed on the samples themselves.
    const int lo = -(1 << (bd - 1)), hi = (1 << (bd - 1)) - 1, mx = (1 << bd) - 1;
    int f = hev ? clampi(p1 - q1, lo, hi) : 0;
    f = clampi(f + __mul24(q0 - p0, 3), lo, hi);
    // f >= lo, so f + 4 and f + 3 only need the upper clamp
    const int f1 = min(f + 4, hi) >> 3, f2 = min(f + 3, hi) >> 3;
    s(0) = clampi(q0 - f1, 0, mx);
    s(-1) = clampi(p0 + f2, 0, mx);
    if (!hev) {
      const int g = (f1 + 1) >> 1;
      s(1) = clampi(q1 - g, 0, mx);
      s(-2) = clampi(p1 + g, 0, mx);
    }
  }
  // wide filter (8.8.5.3): log2 3 (8 taps) or 4 (16 taps)
  if (flat && !flat2) lf_wide<3>(s);
  if (flat2) lf_wide<4>(s);
}

// lf::filter on an LDS buffer: across the edge whose first q sample is
// b[pos], `step` apart.
DEV void lf_filter(uint16_t *b, int pos, int step, int fs, int limit, int blimit, int thresh, int bd) {
  lf_apply([&](int k) -> uint16_t & { return b[pos + k * step]; }, fs, limit, blimit, thresh, bd);
}


// ---- one work item of a wavefront level: the region in LDS
//
// A unit (r, c) of one plane is the horizontal pass of superblock (r, c)
// followed by the vertical pass of superblock (r, c + 1): both passes touch
// the rows of superblock row r (the horizontal pass also the 8 rows above),
// the first the columns of superblock c, the second those of c + 1 and the 8
// left of them.  The region staged in LDS is that L shape: the 8 rows above
// over superblock c's columns, and superblock row r over both superblocks.
// LDS row 8 + y holds plane row y0 + y; LDS column 8 + x holds plane column
// x0 + x, so that a frame-edge vertical pass (superblock column 0) has an
// 8-sample window margin on its left.  The rows past the region are slack
// for lf_stream's reads past a line's end.

template <class T>
struct __align__(16) LfShared {
  static constexpr int BP = LfPitch<T>::BP;
  T buf[80 * BP];
  uint32_t dq[3][272];  // packed decisions of the item's passes (16 slack entries each)
};

// The region in parts of 8-sample vectors (nt = 1 << lnt vectors per
// superblock row): part 0 is what the first vertical and the horizontal
// pass read, the 8 rows above (8 nt vectors, none in superblock row 0 and
// for a lone vertical pass) and then superblock c's sh rows; part 1 is
// superblock c + 1's rows, which only the second vertical pass reads (it is
// written to LDS after the horizontal pass, its loads in flight meanwhile);
// part 2 is superblock c's last vector of each row, the 8 samples left of
// superblock c + 1 that the second vertical pass rewrites (the rest of part
// 0 goes back to the frame before that pass runs).
struct LfItem {
  int x0, y0, top, lnt, sh;
};

template <int PART>
DEV bool lf_vec(const LfItem &it, int k, int &r, int &q) {
  const int nt = 1 << it.lnt;
  if (PART == 0) {
    if (k < it.top) {
      r = k >> it.lnt;
      q = k & (nt - 1);
      return true;
    }
    k -= it.top;
    r = 8 + (k >> it.lnt);
    q = k & (nt - 1);
    return k < (it.sh << it.lnt);
  }
  if (PART == 1) {
    r = 8 + (k >> it.lnt);
    q = nt + (k & (nt - 1));
    return k < (it.sh << it.lnt);
  }
  r = 8 + k;
  q = nt - 1;
  return k < it.sh;
}

// Part 0: at most 8 x 8 + 64 x 8 vectors; part 1: 64 x 8; part 2: 64.
template <class P, int N>
struct LfRegion {
  typedef typename std::conditional<sizeof(P) == 1, uint2, uint4>::type V;  // 8 samples
  V v[N];
};

// Issues every load of a part (into registers) at once.
template <class P, int PART, int N>
DEV void lf_region_load(const P *pl, uint32_t stride, const LfItem &it, LfRegion<P, N> &rg, int lane) {
  typedef typename LfRegion<P, N>::V V;
#pragma unroll
  for (int i = 0; i < N; i++) {
    int r, q;
    if (lf_vec<PART>(it, lane + 64 * i, r, q))
      rg.v[i] = gld((const V *)(pl + (size_t)(it.y0 - 8 + r) * stride + it.x0 + 8 * q));
  }
}

// The region's 8-sample vectors into LDS (as bytes at 8 bits: two dword
// stores; as u16 above: four), and back.  BP and the column offsets keep
// every store dword-aligned.
template <class P, int PART, int N>
DEV void lf_region_to_lds(const LfItem &it, const LfRegion<P, N> &rg, P *buf, int lane) {
  constexpr int BP = LfPitch<P>::BP;
#pragma unroll
  for (int i = 0; i < N; i++) {
    int r, q;
    if (!lf_vec<PART>(it, lane + 64 * i, r, q)) continue;
    uint32_t *l = (uint32_t *)(buf + r * BP + 8 + 8 * q);
    if (sizeof(P) == 1) {
      const uint2 w = *(const uint2 *)&rg.v[i];
      l[0] = w.x;
      l[1] = w.y;
    } else {
      const uint4 w = *(const uint4 *)&rg.v[i];
      l[0] = w.x;
      l[1] = w.y;
      l[2] = w.z;
      l[3] = w.w;
    }
  }
}

// Stores the vectors of a part; skip_last leaves out the last vector of
// each superblock row (part 0 before the second vertical pass).  Every LDS
// read is issued before the first store (an LDS address stays inside the
// buffer for a vector the part does not have), so the reads' latency is paid
// once.
template <class P, int PART, int N>
DEV void lf_region_store(P *pl, uint32_t stride, const LfItem &it, const P *buf, bool skip_last, int lane) {
  constexpr int BP = LfPitch<P>::BP;
  typedef typename LfRegion<P, N>::V V;
  V v[N];
  int rr[N], qq[N];
  bool ok[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    ok[i] = lf_vec<PART>(it, lane + 64 * i, rr[i], qq[i]) && !(skip_last && rr[i] >= 8 && qq[i] == (1 << it.lnt) - 1);
    const uint32_t *l = (const uint32_t *)(buf + (ok[i] ? rr[i] * BP + 8 + 8 * qq[i] : 8));
    if (sizeof(P) == 1)
      *(uint2 *)&v[i] = make_uint2(l[0], l[1]);
    else
      *(uint4 *)&v[i] = make_uint4(l[0], l[1], l[2], l[3]);
  }
#pragma unroll
  for (int i = 0; i < N; i++)
    if (ok[i]) gst((V *)(pl + (size_t)(it.y0 - 8 + rr[i]) * stride + it.x0 + 8 * qq[i]), v[i]);
}

// ---- the edge filters of one pass, streamed through a register window
//
// The vertical pass gives each lane one row and slides a 16-sample register
// window (p7 .. q7 of the current edge) along it: after edge e the 4 samples
// that leave the window are final (no later edge of the row reaches them)
// and go back to LDS, 4 new ones come in.  The horizontal pass does the same
// down each column.

// One pass: lane-owned lines of `len` samples, line i at buf + i * line,
// samples `step` apart; edges at 8 + 4 e for e < n_edges (a multiple of 4).
// dq[rn * 16 + e] packs the decision of edge e, run rn: (filter size + 1) |
// limit << 8 | blimit << 16 | thresh << 24, 0 for no filter; the four of
// edges 4 m .. 4 m + 3 are one 16-byte LDS read.  The window holds the 16
// samples p7 .. q7 of the current edge, position x of the line in w[x & 15]:
// with the edge loop unrolled by 4 every index is a constant, and sliding
// the window is renaming.  The 4 samples entering the window are fetched
// before the filter (the filter of edge e writes at most up to position
// 4 e + 14).
// ROW: the line's samples are contiguous (step 1) and its start 4-byte
// aligned, so the 4 samples entering and the 4 leaving the window at each
// edge are one dword (8 bits) or two (above): one LDS access each way, with
// the unpacking and packing in registers (every window value is a sample,
// in [0, max]).
template <class T, bool ROW>
DEV void lf_load4(const T *p, int step, int (&v)[4]) {
  if (ROW && sizeof(T) == 1) {
    const uint32_t q = *(const uint32_t *)p;
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = (int)__builtin_amdgcn_ubfe(q, 8 * i, 8);
  } else if (ROW) {
    const uint32_t q0 = ((const uint32_t *)p)[0], q1 = ((const uint32_t *)p)[1];
    v[0] = (int)(q0 & 0xffff);
    v[1] = (int)(q0 >> 16);
    v[2] = (int)(q1 & 0xffff);
    v[3] = (int)(q1 >> 16);
  } else {
    // zero-extended here, beside the load (a later block would mask it again)
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = (int)__builtin_amdgcn_ubfe((uint32_t)p[i * step], 0, 8 * sizeof(T));
  }
}

template <class T, bool ROW>
DEV void lf_store4(T *p, int step, int a, int b, int c, int d) {
  if (ROW && sizeof(T) == 1) {
    *(uint32_t *)p = (uint32_t)a | (uint32_t)b << 8 | (uint32_t)c << 16 | (uint32_t)d << 24;
  } else if (ROW) {
    ((uint32_t *)p)[0] = (uint32_t)a | (uint32_t)b << 16;
    ((uint32_t *)p)[1] = (uint32_t)c | (uint32_t)d << 16;
  } else {
    p[0] = (T)a;
    p[step] = (T)b;
    p[2 * step] = (T)c;
    p[3 * step] = (T)d;
  }
}

// Edge e = 4 m + K of lf_stream (K a constant: every window index is one).
template <int K, bool ROW, class T>
DEV void lf_step(int (&w)[16], T *ln, int e, int step, uint32_t d, int bd) {
  // Reads past the line's end (positions >= len, at most 4 n_edges + 11)
  // stay inside the LDS buffer and never reach a result: they enter the
  // window after the last edge that could read them (the last edge is 4
  // apart from the line's end, a narrow filter reading q0 .. q3 only) and
  // are not written back.
  int nw[4];
  lf_load4<T, ROW>(ln + (4 * e + 16) * step, step, nw);
  if (d)
    lf_apply([&](int kk) -> int & { return w[(4 * K + 8 + kk) & 15]; }, (int)(d & 0xff) - 1, (d >> 8) & 0xff,
             (d >> 16) & 0xff, d >> 24, bd);
  lf_store4<T, ROW>(ln + 4 * e * step, step, w[4 * K], w[4 * K + 1], w[4 * K + 2], w[4 * K + 3]);
#pragma unroll
  for (int i = 0; i < 4; i++) w[4 * K + i] = nw[i];
}

template <class T, bool ROW = false>
DEV void lf_stream(T *buf, int line_off, int step, int len, int n_edges, const uint32_t *dq, int rn, int bd) {
  int w[16];
  T *ln = buf + line_off;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int v[4];
    lf_load4<T, ROW>(ln + 4 * i * step, step, v);
#pragma unroll
    for (int k = 0; k < 4; k++) w[4 * i + k] = v[k];
  }
  // the next group's decisions are read before this group's filters; past
  // the last group the read lands in the next run's entries or the 16-entry
  // slack and is not used
  const uint4 *dr = (const uint4 *)(dq + rn * 16);
  uint4 dn = dr[0];
#pragma unroll 1
  for (int e = 0; e < n_edges; e += 4) {
    const uint4 d = dn;
    dn = dr[(e >> 2) + 1];
    lf_step<0, ROW>(w, ln, e, step, d.x, bd);
    lf_step<1, ROW>(w, ln, e + 1, step, d.y, bd);
    lf_step<2, ROW>(w, ln, e + 2, step, d.z, bd);
    lf_step<3, ROW>(w, ln, e + 3, step, d.w, bd);
  }
  // the last window: positions 4 n_edges + i at w[i]
  const int base = 4 * n_edges;
#pragma unroll
  for (int i = 0; i < 16; i += 4)
    if (base + i < len) lf_store4<T, ROW>(ln + (base + i) * step, step, w[i], w[i + 1], w[i + 2], w[i + 3]);
}

// ---- one plane of one work item of level f.t

// item: the item's index in its level; plane: the plane this wave filters
// (lanes 0..63), at plane_ptr with row stride `stride` (samples); dq, geo
// and t as in LfOne.  An inactive wave (the fused kernel's fourth) loads and
// stores nothing and only takes part in the workgroup's barriers, which
// depend on the item alone.
template <bool HBD>
DEV void lf_unit(uint8_t *plane_ptr, uint32_t stride, const uint32_t *dq, uint32_t geo, uint32_t t_level, int item,
                 int plane, bool active, int lane, LfShared<typename std::conditional<HBD, uint16_t, uint8_t>::type> &sh) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  constexpr int BP = LfPitch<P>::BP;
  // Level t: items below n_units are the units (r, t - r) from the first
  // row the level reaches; item n_units, when t + 1 < R, is the vertical
  // pass of superblock (t + 1, 0), which no later item of row t + 1
  // includes (unit (0, 0) does its own).
  const int t = (int)t_level;
  const int C = (int)(geo & 0xfff), R = (int)((geo >> 12) & 0xfff);
  const int r_lo = t >= C ? t - C + 1 : 0, n_units = min(R - 1, t) - r_lo + 1;
  const bool lone_v = item >= n_units;
  const int sb_row = lone_v ? t + 1 : r_lo + item, sb_col = lone_v ? 0 : t - sb_row;
  const bool first_v = lone_v || (sb_row == 0 && sb_col == 0);
  const bool right_v = !lone_v && sb_col + 1 < C;
  const int sx = plane ? (int)((geo >> 24) & 1) : 0, sy = plane ? (int)((geo >> 25) & 1) : 0;
  const int bit_depth = (int)(geo >> 26);
  const int SW = 64 >> sx, SH = 64 >> sy, lnt = 3 - sx;
  LfItem it;
  it.x0 = sb_col * SW;
  it.y0 = sb_row * SH;
  it.top = lone_v || sb_row == 0 ? 0 : 8 << lnt;
  it.lnt = lnt;
  it.sh = SH;
  P *pl = (P *)plane_ptr;
  // every global load at once: the passes' decisions and the region
  const size_t n_sb = (size_t)R * C, sb = (size_t)sb_row * C + sb_col;
  const uint4 *dv = (const uint4 *)(dq + (plane * 2 + 0) * n_sb * 256);
  const uint4 *dh = (const uint4 *)(dq + (plane * 2 + 1) * n_sb * 256);
  uint4 d0 = make_uint4(0, 0, 0, 0), d1 = d0, d2 = d0;
  LfRegion<P, 9> part0;
  LfRegion<P, 8> part1;
  if (active) {
    if (first_v) d0 = gld(dv + sb * 64 + lane);
    if (!lone_v) d1 = gld(dh + sb * 64 + lane);
    if (right_v) d2 = gld(dv + (sb + 1) * 64 + lane);
    lf_region_load<P, 0>(pl, stride, it, part0, lane);
    if (right_v) lf_region_load<P, 1>(pl, stride, it, part1, lane);
    ((uint4 *)sh.dq[0])[lane] = d0;
    ((uint4 *)sh.dq[1])[lane] = d1;
    ((uint4 *)sh.dq[2])[lane] = d2;
    lf_region_to_lds<P, 0>(it, part0, sh.buf, lane);
    // superblock c + 1 too: the horizontal pass works on superblock c's
    // columns only (the first vertical pass reads past its line's end into
    // these columns but never writes them), and its registers are then free
    // for the passes
    if (right_v) lf_region_to_lds<P, 1>(it, part1, sh.buf, lane);
  }
  __syncthreads();
  // pass 0: the vertical edges of superblock (r, c) (c = 0 only); pass 1 its
  // horizontal edges; pass 2 the vertical edges of (r, c + 1).  One copy of
  // the edge loop for every pass and geometry.
#pragma unroll 1
  for (int pass = 0; pass < 3; pass++) {
    if (pass == 0 ? !first_v : pass == 1 ? lone_v : !right_v) continue;
    if (pass == 2) {
      // superblock c's finished samples out
      if (active) lf_region_store<P, 0, 9>(pl, stride, it, sh.buf, true, lane);
      __syncthreads();
    }
    // vertical passes: rows, 4-byte aligned (BP and SW are multiples of 4
    // bytes), packed LDS accesses; the horizontal pass: columns
    if (active && pass != 1 && lane < SH)
      lf_stream<P, true>(sh.buf, (8 + lane) * BP + (pass == 2 ? SW : 0), 1, SW + 8, SW / 4, sh.dq[pass], lane >> 2,
                         HBD ? bit_depth : 8);
    if (active && pass == 1 && lane < SW)
      lf_stream<P, false>(sh.buf, 8 + lane, BP, SH + 8, SH / 4, sh.dq[pass], lane >> 2, HBD ? bit_depth : 8);
    __syncthreads();
  }
  if (!active) return;
  if (right_v) {
    lf_region_store<P, 1, 8>(pl, stride, it, sh.buf, false, lane);
    lf_region_store<P, 2, 1>(pl, stride, it, sh.buf, false, lane);
  } else {
    lf_region_store<P, 0, 9>(pl, stride, it, sh.buf, false, lane);
  }
}

}  // namespace vp9g
