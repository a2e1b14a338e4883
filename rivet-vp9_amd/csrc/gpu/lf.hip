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
#include "frame_dev.h"
#include "lf_dev.h"

namespace vp9g {

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

// Per frame, K = 1: workgroup j of 6 n_sb takes entry (plane, pass) = j /
// n_sb of superblock j % n_sb; K = 6: workgroup j of n_sb takes the six
// entries of superblock j in turn (for large batches, where a workgroup per
// entry would spend more on dispatch than on the decisions).
template <int K>
DEV void lf_mask_wg(const LfArgs *fa, const Range *rg, uint32_t nf, uint32_t wg) {
  const uint32_t fi = range_frame(rg, nf, wg);
  const LfArgs &a = fa[fi];
  const FrameGeom &g = a.g;
  const int n_sb = g.sb64_rows * g.sb64_cols;
  const int j = wg - rg[fi].wg, sb = K == 1 ? j % n_sb : j;
  const uint32_t sb_row = sb / g.sb64_cols, sb_col = sb % g.sb64_cols;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int pp = K == 1 ? j / n_sb : k, pass = pp & 1, plane = pp >> 1;
    int fs, lvl;
    // entry rn * 16 + ed: a run's decisions along its line are contiguous
    lf_decide(a, plane, pass, sb_row, sb_col, threadIdx.x & 15, threadIdx.x >> 4, fs, lvl);
    const size_t o = ((size_t)pp * n_sb + sb) * 256 + threadIdx.x;
    a.dq[o] = fs < 0 ? 0u
                     : (uint32_t)(fs + 1) | (uint32_t)a.p.limit[lvl] << 8 | (uint32_t)a.p.blimit[lvl] << 16 |
                           (uint32_t)a.p.thresh[lvl] << 24;
  }
}

template <int K>
__global__ __launch_bounds__(256) void k_lf_mask(const LfArgs *fa, const Range *rg, uint32_t nf) {
  lf_mask_wg<K>(fa, rg, nf, blockIdx.x);
}

// The margin clear and the loop filter decisions of a batch in one launch,
// its first: workgroups below n_clear clear margins (frame_dev.h), the rest
// decide.  The decisions read mode info only and write the decision
// buffer, the clear writes frame samples no other part of the launch
// touches.
template <class P, int K>
__global__ __launch_bounds__(256) void k_prep(const ClearArgs *fc, const Range *rc, uint32_t n_clear,
                                              const LfArgs *fl, const Range *rl, uint32_t nf) {
  if (blockIdx.x < n_clear)
    clear_margins_wg<P>(fc, rc, nf, blockIdx.x);
  else
    lf_mask_wg<K>(fl, rl, nf, blockIdx.x - n_clear);
}

void launch_prep(hipStream_t s, const ClearArgs *fc, const Range *rc, uint32_t n_clear, const LfArgs *fl,
                 const Range *rl, uint32_t n_mask, uint32_t n, int per_sb, bool hbd) {
  if (!n_clear && !n_mask) return;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(n_clear + n_mask), dim3(256), 0, s, fc, rc, n_clear, fl, rl, n);
  };
  if (hbd)
    per_sb == 6 ? go(k_prep<uint16_t, 6>) : go(k_prep<uint16_t, 1>);
  else
    per_sb == 6 ? go(k_prep<uint8_t, 6>) : go(k_prep<uint8_t, 1>);
}

// ---- one wavefront level per launch (decisions from k_lf_mask)

template <bool HBD, bool ONE>
__global__ __launch_bounds__(64) void k_lf(uint8_t *p0, uint8_t *p1, uint8_t *p2, uint32_t s0, uint32_t s1,
                                           uint32_t s2, uint32_t *dq, uint32_t geo, uint32_t t1, const LfArgs *fa,
                                           const Range *rg, uint32_t nf) {
  __shared__ LfShared<typename std::conditional<HBD, uint16_t, uint8_t>::type> sh;
  // a launch of one frame (ONE) passes what the workgroups need as
  // arguments (preloaded, so the region loads need no load before them, and
  // known to be global memory: LfOne); a batch finds each frame's through
  // the range table.  A frame's 3 n workgroups are the luma planes of its n
  // items, then the two chroma planes: the workgroups that take longest are
  // dispatched first.
  if (!ONE) {
    const uint32_t fi = range_frame(rg, nf, blockIdx.x);
    const LfArgs &a = fa[fi];
    const int j = (int)(blockIdx.x - rg[fi].wg), items = (int)(rg[fi].count / 3);
    const int plane = j / items;
    const LfOne o = lf_one(a, rg[fi].first);
    lf_unit<HBD>((uint8_t *)a.cur.p[plane], a.cur.stride[plane], o.dq, o.geo, o.t, j - plane * items, plane, true,
                 (int)threadIdx.x, sh);
    return;
  }
  const int items = (int)(gridDim.x / 3), plane = (int)blockIdx.x / items, item = (int)blockIdx.x - plane * items;
  lf_unit<HBD>(plane == 0 ? p0 : plane == 1 ? p1 : p2, plane == 0 ? s0 : plane == 1 ? s1 : s2, dq, geo, t1, item,
               plane, true, (int)threadIdx.x, sh);
}

void launch_lf_mask(hipStream_t s, const LfArgs *fa, const Range *r, uint32_t n, uint32_t wgs, int per_sb) {
  if (!wgs) return;
  if (per_sb == 6)
    hipLaunchKernelGGL(k_lf_mask<6>, dim3(wgs), dim3(256), 0, s, fa, r, n);
  else
    hipLaunchKernelGGL(k_lf_mask<1>, dim3(wgs), dim3(256), 0, s, fa, r, n);
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

// Per-kernel seam: lf::filter_edges (lf.rs:178-218) over a sequence of
// passes, in order, by one workgroup: each pass stages its region (edges
// reach 8 samples back and 8 forward) in LDS, filters edges first ..
// n_edges - 1 of it (one lane per line along the edges), and writes the
// region back before the next pass reads the plane.
template <bool HBD>
__global__ __launch_bounds__(64) void k_lf_edges(const uint8_t *tables, const LfPassDesc *passes, uint32_t count,
                                                 void *buf_, uint32_t stride, int bd) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ uint16_t buf[80 * 80];
  P *g = (P *)buf_;
  for (uint32_t i = 0; i < count; i++) {
    const LfPassDesc ps = passes[i];
    const int8_t *fs = (const int8_t *)(tables + 1024 * (size_t)i);
    const uint8_t *lim = tables + 1024 * (size_t)i + 256, *blim = lim + 256, *thr = lim + 512;
    const int x0 = ps.x0, y0 = ps.y0, vertical = ps.vertical, n_edges = ps.n_edges, n_runs = ps.n_runs;
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
      for (int ed = ps.first; ed < n_edges; ed++) {
        const int f = fs[ed * 16 + rn];
        if (f < 0) continue;
        const int j = ed * 16 + rn;
        if (vertical)
          lf_filter(buf, lane * 80 + 8 + 4 * ed, 1, f, lim[j], blim[j], thr[j], bd);
        else
          lf_filter(buf, (8 + 4 * ed) * 80 + lane, 80, f, lim[j], blim[j], thr[j], bd);
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < W * H; k += 64) {
      int r = k / W, c = k - r * W;
      if (gy0 + r >= 0 && gx0 + c >= 0) g[(size_t)(gy0 + r) * stride + gx0 + c] = (P)buf[r * 80 + c];
    }
    __syncthreads();  // the next pass reads what this one wrote
  }
}

// lf_filter's limits are unshifted values (the Edges tables carry them).
void launch_lf_edges(hipStream_t s, const uint8_t *tables, const LfPassDesc *passes, uint32_t count, void *buf,
                     uint32_t stride, uint32_t bit_depth) {
  if (!count) return;
  if (bit_depth > 8)
    hipLaunchKernelGGL(k_lf_edges<true>, dim3(1), dim3(64), 0, s, tables, passes, count, buf, stride,
                       (int)bit_depth);
  else
    hipLaunchKernelGGL(k_lf_edges<false>, dim3(1), dim3(64), 0, s, tables, passes, count, buf, stride,
                       (int)bit_depth);
}

}  // namespace vp9g
