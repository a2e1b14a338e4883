// The loop filter (8.8): edge decisions per superblock (one parallel
// launch), then the sample filters as a superblock wavefront.
//
// Decisions follow superblock() (src/decoder/loopfilter.rs:466-620), one
// lane per (edge, 4-sample run) entry, with LvlLookup and the limits of
// Ctx::new (loopfilter.rs:15-47, 135-170) precomputed by the host.  The
// filters follow src/dsp/lf.rs:24-143.
//
// Order: the reference filters each plane superblock by superblock in raster
// order, all vertical edges of a superblock (left to right) then all its
// horizontal edges (top to bottom) (loopfilter.rs:66-80).  The vertical pass
// V(r, c) touches superblock (r, c) and the 8 samples left of it, the
// horizontal pass H(r, c) superblock (r, c) and the 8 samples above it.  So
// V(r, c) follows H(r, c - 1) only, and H(r, c) follows V(r, c), H(r - 1, c)
// and V(r - 1, c + 1).  Grouping each row's sequence into units
// U(r, c) = H(r, c) V(r, c + 1), U(r, c) follows U(r, c - 1) and U(r - 1, c)
// only, and units with equal r + c touch disjoint samples: one launch per
// r + c (R + C - 1 launches, against 2R + C - 2 for whole superblocks).  The
// first vertical pass of each row, V(r, 0), depends on nothing and runs
// alone one level early (unit (0, 0) runs V(0, 0) itself).  Inside a pass
// the edges are sequential along each row (vertical) or column (horizontal)
// and the lines independent: one lane per line, on the region staged in LDS.
#include <type_traits>

#include "dev_common.h"
#include "launch.h"
#include "vp9_consts.h"

namespace vp9g {

constexpr int BP = 138;  // LDS pitch of a unit's region (8 + 128 columns + slack, odd in dwords)

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
  int tt = 0;
#pragma unroll
  for (int j = -n; j <= n; j++) tt += at(-n + j);
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
template <class S>
DEV void lf_apply(S s, int fs, int limit, int blimit, int thresh, int bd) {
  const int q0 = s(0), q1 = s(1), q2 = s(2), q3 = s(3);
  const int p0 = s(-1), p1 = s(-2), p2 = s(-3), p3 = s(-4);
  const int sh = bd - 8;
  const uint32_t d10 = adiff(p1, p0), e10 = adiff(q1, q0);
  const uint32_t m = max(max(max(adiff(p3, p2), adiff(p2, p1)), d10), max(max(e10, adiff(q2, q1)), adiff(q3, q2)));
  const uint32_t edge = adiff(p0, q0) * 2 + (adiff(p1, q1) >> 1);
  const bool on = fs >= 0 && m <= (uint32_t)(limit << sh) && edge <= (uint32_t)(blimit << sh);
  const bool hev = max(d10, e10) > (uint32_t)(thresh << sh);
  const uint32_t one = 1u << sh;
  const bool flat = on && fs >= 1 &&
                    max(max(max(d10, e10), max(adiff(p2, p0), adiff(q2, q0))), max(adiff(p3, p0), adiff(q3, q0))) <= one;
  bool flat2 = false;
  if (flat && fs >= 2) {
    uint32_t f2 = 0;
#pragma unroll
    for (int k = 4; k < 8; k++) f2 = max(f2, max(adiff(s(-(k + 1)), p0), adiff(s(k), q0)));
    flat2 = f2 <= one;
  }
  if (on && !flat) {
    // narrow filter (8.8.5.2)
    const int lo = -(1 << (bd - 1)), hi = (1 << (bd - 1)) - 1, off = 0x80 << sh;
    const int ps1 = p1 - off, ps0 = p0 - off, qs0 = q0 - off, qs1 = q1 - off;
    int f = hev ? clampi(ps1 - qs1, lo, hi) : 0;
    f = clampi(f + 3 * (qs0 - ps0), lo, hi);
    const int f1 = clampi(f + 4, lo, hi) >> 3, f2 = clampi(f + 3, lo, hi) >> 3;
    s(0) = clampi(qs0 - f1, lo, hi) + off;
    s(-1) = clampi(ps0 + f2, lo, hi) + off;
    if (!hev) {
      const int g = (f1 + 1) >> 1;
      s(1) = clampi(qs1 - g, lo, hi) + off;
      s(-2) = clampi(ps1 + g, lo, hi) + off;
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

// ---- decisions

// The filter size (-1: none) and level of entry (ed, rn) of one superblock
// pass (superblock(), loopfilter.rs:466-620): edge ed (4 samples apart
// across), run rn (4 samples along).
DEV void lf_decide(const LfArgs &a, int plane, int pass, uint32_t sb_row, uint32_t sb_col, int ed, int rn,
                   int &fs_out, int &lvl_out) {
  const FrameGeom &g = a.g;
  const uint32_t row = sb_row * 8, col = sb_col * 8;
  const int sub_x = plane ? g.ss_x : 0, sub_y = plane ? g.ss_y : 0;
  const int sub = pass == 0 ? sub_x : sub_y;
  const int edge_len = pass == 0 ? 64 >> sub_y : 64 >> sub_x;
  const int n_edges = 16 >> sub, n_runs = edge_len / 4;
  const uint32_t mi_cols = g.mi_cols, mi_rows = g.mi_rows;
  fs_out = -1;
  lvl_out = 0;
  auto at_x = [&](int e, int r) -> uint32_t {
    return pass == 0 ? col * 8 + e * (4u << sub_x) : col * 8 + ((4u * r) << sub_x);
  };
  auto at_y = [&](int e, int r) -> uint32_t {
    return pass == 0 ? row * 8 + ((4u * r) << sub_y) : row * 8 + e * (4u << sub_y);
  };
  if (ed >= n_edges || rn >= n_runs) return;
  const int ca = ed >> 1, cb = rn >> 1, k = rn & 1;
  uint32_t x = at_x(2 * ca, 2 * cb), y = at_y(2 * ca, 2 * cb);
  if (x >= 8 * mi_cols || y >= 8 * mi_rows) return;
  const uint32_t loop_col = ((x >> 3) >> sub_x) << sub_x, loop_row = ((y >> 3) >> sub_y) << sub_y;
  const vp9r_mi m = a.mi[loop_row * mi_cols + loop_col];
  const int mode_type = m.y_mode == NEARESTMV || m.y_mode == NEARMV || m.y_mode == NEWMV;
  const int rf = m.ref_frame0 > 0 ? m.ref_frame0 : 0;
  const int l = a.p.lvl[m.segment_id & 7][rf & 3][mode_type];
  if (l == 0) return;
  int tx_sz = m.tx_size;
  if (plane > 0) {
    if (m.mi_size < BLOCK_8X8) {
      tx_sz = TX_4X4;
    } else {
      // get_uv_tx_size: min(tx_size, max_txsize_lookup[ss_size_lookup[...]]),
      // i.e. the largest transform fitting the block's chroma
      const int ux = g.legacy_uv ? 1 : g.ss_x, uy = g.legacy_uv ? 1 : g.ss_y;
      const int wl = ((kW4P >> (3 * m.mi_size)) & 7) - ux, hl = ((kH4P >> (3 * m.mi_size)) & 7) - uy;
      const int uv_max = min(min(wl, hl), 3);
      tx_sz = min(tx_sz, uv_max);
    }
  }
  const int sb_size = sub == 0 ? m.mi_size : max((int)m.mi_size, (int)BLOCK_16X16);
  const uint32_t block_mask = pass == 0 ? 8u * num8x8_wide(sb_size) - 1 : 8u * num8x8_high(sb_size) - 1;
  const bool coded = m.ref_frame0 <= INTRA_FRAME || !m.skip;
  x = at_x(ed, 2 * cb);
  y = at_y(ed, 2 * cb);
  if ((pass == 0 && x == 0) || (pass == 1 && y == 0)) return;
  const uint32_t across = pass == 0 ? x : y;
  const bool is_block_edge = (across & block_mask) == 0;
  const bool tx_edge = (ed & ((1 << tx_sz) - 1)) == 0;
  if (!is_block_edge && !(tx_edge && coded)) return;
  const bool is_32_edge = ed % 8 == 0;
  const int base_size = (tx_sz == TX_4X4 && is_32_edge) ? TX_8X8 : min(tx_sz, (int)TX_16X16);
  const int filter_size =
      ((pass == 0 && sub_x == 1 && base_size == TX_16X16 && (x >> 3) == mi_cols - 1) ||
       (pass == 1 && sub_y == 1 && base_size == TX_16X16 && (y >> 3) == mi_rows - 1))
          ? TX_8X8
          : base_size;
  const uint32_t xx = at_x(ed, rn), yy = at_y(ed, rn);
  if (k == 1 && (xx >= 8 * mi_cols || yy >= 8 * mi_rows)) return;
  const bool is_tx_edge =
      (pass == 1 && sub_x == 1 && (mi_cols & 1) && (ed & 1) && (xx + 8) >= mi_cols * 8) ? false : tx_edge;
  if (!(is_block_edge || (is_tx_edge && coded))) return;
  fs_out = filter_size;
  lvl_out = l;
}

// Per frame, workgroup j of 6 * n_sb: superblock j % n_sb, pass and plane
// from j / n_sb.
__global__ __launch_bounds__(256) void k_lf_mask(const LfArgs *fa, const Range *rg, uint32_t nf) {
  const uint32_t fi = range_frame(rg, nf, blockIdx.x);
  const LfArgs &a = fa[fi];
  const FrameGeom &g = a.g;
  const int n_sb = g.sb64_rows * g.sb64_cols;
  const int j = blockIdx.x - rg[fi].wg, sb = j % n_sb, pp = j / n_sb, pass = pp & 1, plane = pp >> 1;
  int fs, lvl;
  lf_decide(a, plane, pass, sb / g.sb64_cols, sb % g.sb64_cols, threadIdx.x >> 4, threadIdx.x & 15, fs, lvl);
  const size_t o = ((size_t)(plane * 2 + pass) * n_sb + sb) * 256 + threadIdx.x;
  a.dq[o] = fs < 0 ? 0u
                   : (uint32_t)(fs + 1) | (uint32_t)a.p.limit[lvl] << 8 | (uint32_t)a.p.blimit[lvl] << 16 |
                         (uint32_t)a.p.thresh[lvl] << 24;
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

struct __align__(16) LfShared {
  uint16_t buf[80 * BP];
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
DEV void lf_region_load(const P *pl, uint32_t stride, const LfItem &it, LfRegion<P, N> &rg) {
  typedef typename LfRegion<P, N>::V V;
#pragma unroll
  for (int i = 0; i < N; i++) {
    int r, q;
    if (lf_vec<PART>(it, threadIdx.x + 64 * i, r, q))
      rg.v[i] = gld((const V *)(pl + (size_t)(it.y0 - 8 + r) * stride + it.x0 + 8 * q));
  }
}

// Samples as 16-bit values, two to a dword store (BP and the column offsets
// are even).
template <class P, int PART, int N>
DEV void lf_region_to_lds(const LfItem &it, const LfRegion<P, N> &rg, uint16_t *buf) {
#pragma unroll
  for (int i = 0; i < N; i++) {
    int r, q;
    if (!lf_vec<PART>(it, threadIdx.x + 64 * i, r, q)) continue;
    uint32_t *l = (uint32_t *)(buf + r * BP + 8 + 8 * q);
    if (sizeof(P) == 1) {
      const uint2 w = *(const uint2 *)&rg.v[i];
      l[0] = (w.x & 0xff) | (w.x & 0xff00) << 8;
      l[1] = (w.x >> 16 & 0xff) | (w.x >> 8 & 0xff0000);
      l[2] = (w.y & 0xff) | (w.y & 0xff00) << 8;
      l[3] = (w.y >> 16 & 0xff) | (w.y >> 8 & 0xff0000);
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
// each superblock row (part 0 before the second vertical pass).
template <class P, int PART, int N>
DEV void lf_region_store(P *pl, uint32_t stride, const LfItem &it, const uint16_t *buf, bool skip_last) {
#pragma unroll
  for (int i = 0; i < N; i++) {
    int r, q;
    if (!lf_vec<PART>(it, threadIdx.x + 64 * i, r, q)) continue;
    if (skip_last && r >= 8 && q == (1 << it.lnt) - 1) continue;
    const uint32_t *l = (const uint32_t *)(buf + r * BP + 8 + 8 * q);
    P *g = pl + (size_t)(it.y0 - 8 + r) * stride + it.x0 + 8 * q;
    if (sizeof(P) == 1) {
      uint2 w;
      w.x = (l[0] & 0xff) | (l[0] >> 8 & 0xff00) | (l[1] & 0xff) << 16 | (l[1] >> 16) << 24;
      w.y = (l[2] & 0xff) | (l[2] >> 8 & 0xff00) | (l[3] & 0xff) << 16 | (l[3] >> 16) << 24;
      gst((uint2 *)g, w);
    } else {
      gst((uint4 *)g, make_uint4(l[0], l[1], l[2], l[3]));
    }
  }
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
// dq[e * 16 + rn] packs the edge's decision: (filter size + 1) | limit << 8
// | blimit << 16 | thresh << 24, 0 for no filter.  The window holds the 16
// samples p7 .. q7 of the current edge, position x of the line in w[x & 15]:
// with the edge loop unrolled by 4 every index is a constant, and sliding
// the window is renaming.  The next edge's decision and the 4 samples
// entering the window are fetched before the filter (the filter of edge e
// writes at most up to position 4 e + 14).
// Edge e = 4 m + K of lf_stream (K a constant: every window index is one).
template <int K>
DEV void lf_step(int (&w)[16], uint16_t *ln, int e, int step, int n_edges, const uint32_t *dq, int rn,
                 uint32_t &d_next, int bd) {
  // Reads past the line's end (positions >= len, at most 4 n_edges + 11)
  // stay inside LfShared and never reach a result: they enter the window
  // after the last edge that could read them (the last edge is 4 apart from
  // the line's end, a narrow filter reading q0 .. q3 only) and are not
  // written back.  Likewise the decision after the last edge.
  const uint32_t d = d_next;
  d_next = dq[(e + 1) * 16 + rn];
  int nw[4];
#pragma unroll
  for (int i = 0; i < 4; i++) nw[i] = ln[(4 * e + 16 + i) * step];
  if (d)
    lf_apply([&](int kk) -> int & { return w[(4 * K + 8 + kk) & 15]; }, (int)(d & 0xff) - 1, (d >> 8) & 0xff,
             (d >> 16) & 0xff, d >> 24, bd);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    ln[(4 * e + i) * step] = (uint16_t)w[4 * K + i];
    w[4 * K + i] = nw[i];
  }
}

DEV void lf_stream(uint16_t *buf, int line_off, int step, int len, int n_edges, const uint32_t *dq, int rn,
                   int bd) {
  int w[16];
  uint16_t *ln = buf + line_off;
#pragma unroll
  for (int i = 0; i < 16; i++) w[i] = ln[i * step];
  uint32_t d_next = dq[rn];
#pragma unroll 1
  for (int e = 0; e < n_edges; e += 4) {
    lf_step<0>(w, ln, e, step, n_edges, dq, rn, d_next, bd);
    lf_step<1>(w, ln, e + 1, step, n_edges, dq, rn, d_next, bd);
    lf_step<2>(w, ln, e + 2, step, n_edges, dq, rn, d_next, bd);
    lf_step<3>(w, ln, e + 3, step, n_edges, dq, rn, d_next, bd);
  }
  // the last window: positions 4 n_edges + i at w[i]
  const int base = 4 * n_edges;
#pragma unroll
  for (int i = 0; i < 16; i++)
    if (base + i < len) ln[(base + i) * step] = (uint16_t)w[i];
}

// ---- one wavefront level per launch (decisions from k_lf_mask)

template <bool HBD, bool ONE>
__global__ __launch_bounds__(64) void k_lf(uint8_t *p0, uint8_t *p1, uint8_t *p2, uint32_t s0, uint32_t s1,
                                           uint32_t s2, uint32_t *dq, uint32_t geo, uint32_t t1, const LfArgs *fa,
                                           const Range *rg, uint32_t nf) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ LfShared sh;
  const int lane = threadIdx.x;
  // a launch of one frame (ONE) passes what the workgroups need as
  // arguments (preloaded, so the region loads need no load before them, and
  // known to be global memory: LfOne); a batch finds each frame's through
  // the range table
  int j = (int)blockIdx.x;
  if (!ONE) {
    const uint32_t fi = range_frame(rg, nf, blockIdx.x);
    const LfOne o = lf_one(fa[fi], rg[fi].first);
    p0 = o.p[0];
    p1 = o.p[1];
    p2 = o.p[2];
    s0 = o.stride[0];
    s1 = o.stride[1];
    s2 = o.stride[2];
    dq = o.dq;
    geo = o.geo;
    t1 = o.t;
    j = (int)(blockIdx.x - rg[fi].wg);
  }
  const LfOne f = {{p0, p1, p2}, {s0, s1, s2}, dq, geo, t1};
  // Level t: items j / 3 (plane j % 3) below n_units are the units (r, t - r)
  // from the first row the level reaches; item n_units, when t + 1 < R, is
  // the vertical pass of superblock (t + 1, 0), which no later item of row
  // t + 1 includes (unit (0, 0) does its own).
  const int t = (int)f.t;
  const int C = (int)(f.geo & 0xfff), R = (int)((f.geo >> 12) & 0xfff);
  const int r_lo = t >= C ? t - C + 1 : 0, n_units = min(R - 1, t) - r_lo + 1;
  const int plane = j % 3, item = j / 3;
  const bool lone_v = item >= n_units;
  const int sb_row = lone_v ? t + 1 : r_lo + item, sb_col = lone_v ? 0 : t - sb_row;
  const bool first_v = lone_v || (sb_row == 0 && sb_col == 0);
  const bool right_v = !lone_v && sb_col + 1 < C;
  const int sx = plane ? (int)((f.geo >> 24) & 1) : 0, sy = plane ? (int)((f.geo >> 25) & 1) : 0;
  const int bit_depth = (int)(f.geo >> 26);
  const int SW = 64 >> sx, SH = 64 >> sy, lnt = 3 - sx;
  LfItem it;
  it.x0 = sb_col * SW;
  it.y0 = sb_row * SH;
  it.top = lone_v || sb_row == 0 ? 0 : 8 << lnt;
  it.lnt = lnt;
  it.sh = SH;
  const uint32_t stride = plane == 0 ? s0 : plane == 1 ? s1 : s2;
  P *pl = (P *)(plane == 0 ? p0 : plane == 1 ? p1 : p2);
  // every global load at once: the passes' decisions and the region
  const size_t n_sb = (size_t)R * C, sb = (size_t)sb_row * C + sb_col;
  const uint4 *dv = (const uint4 *)(f.dq + (plane * 2 + 0) * n_sb * 256);
  const uint4 *dh = (const uint4 *)(f.dq + (plane * 2 + 1) * n_sb * 256);
  uint4 d0 = make_uint4(0, 0, 0, 0), d1 = d0, d2 = d0;
  if (first_v) d0 = gld(dv + sb * 64 + lane);
  if (!lone_v) d1 = gld(dh + sb * 64 + lane);
  if (right_v) d2 = gld(dv + (sb + 1) * 64 + lane);
  LfRegion<P, 9> part0;
  LfRegion<P, 8> part1;
  lf_region_load<P, 0>(pl, stride, it, part0);
  if (right_v) lf_region_load<P, 1>(pl, stride, it, part1);
  ((uint4 *)sh.dq[0])[lane] = d0;
  ((uint4 *)sh.dq[1])[lane] = d1;
  ((uint4 *)sh.dq[2])[lane] = d2;
  lf_region_to_lds<P, 0>(it, part0, sh.buf);
  __syncthreads();
  // pass 0: the vertical edges of superblock (r, c) (c = 0 only); pass 1 its
  // horizontal edges; pass 2 the vertical edges of (r, c + 1).  One copy of
  // the edge loop for every pass and geometry.
#pragma unroll 1
  for (int pass = 0; pass < 3; pass++) {
    if (pass == 0 ? !first_v : pass == 1 ? lone_v : !right_v) continue;
    if (pass == 2) {
      // superblock c + 1 into LDS, superblock c's finished samples out
      lf_region_to_lds<P, 1>(it, part1, sh.buf);
      lf_region_store<P, 0, 9>(pl, stride, it, sh.buf, true);
      __syncthreads();
    }
    const bool vert = pass != 1;
    if (lane < (vert ? SH : SW))
      lf_stream(sh.buf, vert ? (8 + lane) * BP + (pass == 2 ? SW : 0) : 8 + lane, vert ? 1 : BP,
                (vert ? SW : SH) + 8, (vert ? SW : SH) / 4, sh.dq[pass], lane >> 2, HBD ? bit_depth : 8);
    __syncthreads();
  }
  if (right_v) {
    lf_region_store<P, 1, 8>(pl, stride, it, sh.buf, false);
    lf_region_store<P, 2, 1>(pl, stride, it, sh.buf, false);
  } else {
    lf_region_store<P, 0, 9>(pl, stride, it, sh.buf, false);
  }
}

void launch_lf_mask(hipStream_t s, const LfArgs *fa, const Range *r, uint32_t n, uint32_t wgs) {
  if (!wgs) return;
  hipLaunchKernelGGL(k_lf_mask, dim3(wgs), dim3(256), 0, s, fa, r, n);
}

void launch_lf_level(hipStream_t s, const LfArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd,
                     const LfOne &one) {
  if (!wgs) return;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(wgs), dim3(64), 0, s, one.p[0], one.p[1], one.p[2], one.stride[0],
                       one.stride[1], one.stride[2], one.dq, one.geo, one.t, fa, r, n);
  };
  if (one.p[0]) {
    if (hbd)
      go(k_lf<true, true>);
    else
      go(k_lf<false, true>);
  } else {
    if (hbd)
      go(k_lf<true, false>);
    else
      go(k_lf<false, false>);
  }
}

// Per-kernel seam: lf::filter_edges on global samples, one pass per
// workgroup launch (passes are ordered by the stream).
template <bool HBD>
__global__ __launch_bounds__(64) void k_lf_edges(const uint8_t *e, int x0, int y0, int vertical,
                                                 int n_edges, int n_runs, void *buf_, uint32_t stride,
                                                 int bd) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  // stage the pass's region (edges reach 8 back and 8 forward) in LDS
  __shared__ uint16_t buf[80 * 80];
  P *g = (P *)buf_;
  const int8_t *fs = (const int8_t *)e;
  const uint8_t *lim = e + 256, *blim = e + 512, *thr = e + 768;
  const int ext_a = 4 * n_edges + 8, ext_r = 4 * n_runs;  // across, along
  const int W = vertical ? ext_a + 8 : ext_r, H = vertical ? ext_r : ext_a + 8;
  const int gx0 = vertical ? x0 - 8 : x0, gy0 = vertical ? y0 : y0 - 8;
  for (int k = threadIdx.x; k < W * H; k += 64) {
    int r = k / W, c = k - r * W;
    if (gy0 + r >= 0 && gx0 + c >= 0) buf[r * 80 + c] = g[(size_t)(gy0 + r) * stride + gx0 + c];
  }
  __syncthreads();
  const int lane = threadIdx.x;
  if (lane < ext_r) {
    const int rn = lane >> 2;
    for (int ed = 0; ed < n_edges; ed++) {
      const int f = fs[ed * 16 + rn];
      if (f < 0) continue;
      const int i = ed * 16 + rn;
      if (vertical)
        lf_filter(buf, lane * 80 + 8 + 4 * ed, 1, f, lim[i], blim[i], thr[i], bd);
      else
        lf_filter(buf, (8 + 4 * ed) * 80 + lane, 80, f, lim[i], blim[i], thr[i], bd);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < W * H; k += 64) {
    int r = k / W, c = k - r * W;
    if (gy0 + r >= 0 && gx0 + c >= 0) g[(size_t)(gy0 + r) * stride + gx0 + c] = (P)buf[r * 80 + c];
  }
}

// lf_filter's limits are unshifted values; limit / blimit / thresh tables
// per level are not used here: the Edges tables carry the values.
void launch_lf_edges(hipStream_t s, uint32_t count, const uint8_t *edges, const uint16_t *x0,
                     const uint16_t *y0, const uint8_t *vertical, const uint8_t *n_edges,
                     const uint8_t *n_runs, void *buf, uint32_t stride, uint32_t bit_depth) {
  // host arrays for the geometry, device pointer for the tables
  for (uint32_t i = 0; i < count; i++) {
    if (bit_depth > 8)
      hipLaunchKernelGGL(k_lf_edges<true>, dim3(1), dim3(64), 0, s, edges + 1024 * (size_t)i, x0[i], y0[i],
                         vertical[i], n_edges[i], n_runs[i], buf, stride, (int)bit_depth);
    else
      hipLaunchKernelGGL(k_lf_edges<false>, dim3(1), dim3(64), 0, s, edges + 1024 * (size_t)i, x0[i], y0[i],
                         vertical[i], n_edges[i], n_runs[i], buf, stride, (int)bit_depth);
  }
}

}  // namespace vp9g
