// The loop filter's edge decisions (device part shared by the decision
// launch, the fused prep launch in lf.hip and the prep + inter prediction
// launch in mc.hip).
//
// Decisions follow superblock() (src/decoder/loopfilter.rs:466-620), one
// lane per (edge, 4-sample run) entry, with LvlLookup and the limits of
// Ctx::new (loopfilter.rs:15-47, 135-170) precomputed by the host.
#pragma once

#include "dev_common.h"
#include "launch.h"
#include "vp9_consts.h"

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
// entry would spend more on dispatch than on the decisions).  NT threads,
// the 256 (edge, run) pairs of an entry NT apart.
template <int K, int NT>
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
#pragma unroll
    for (int e = threadIdx.x; e < 256; e += NT) {
      int fs, lvl;
      // entry rn * 16 + ed: a run's decisions along its line are contiguous
      lf_decide(a, plane, pass, sb_row, sb_col, e & 15, e >> 4, fs, lvl);
      const size_t o = ((size_t)pp * n_sb + sb) * 256 + e;
      a.dq[o] = fs < 0 ? 0u
                       : (uint32_t)(fs + 1) | (uint32_t)a.p.limit[lvl] << 8 | (uint32_t)a.p.blimit[lvl] << 16 |
                             (uint32_t)a.p.thresh[lvl] << 24;
    }
  }
}

}  // namespace vp9g
