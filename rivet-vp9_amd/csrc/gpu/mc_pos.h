// Inter prediction positions (8.5.2.1-8.5.2.3): motion vector selection,
// clamping and scaling (src/decoder/recon.rs:106-170, 284-292;
// src/decoder/mod.rs:495-541).  Host and device: the kernels use it, and the
// host uses it to count the reference footprints a frame reads.
#pragma once

#include "dev_common.h"
#include "launch.h"
#include "vp9_consts.h"

namespace vp9g {

// A reference position: start (1/16 sample), steps, clamp bounds, plane.
struct RefPos {
  int32_t x, y, xs, ys, last_x, last_y;
};

HDEV int32_t q2(int32_t v) { return (v < 0 ? v - 1 : v + 1) / 2; }
HDEV int32_t q4(int32_t v) { return (v < 0 ? v - 2 : v + 2) / 4; }

// Motion vector selection (8.5.2.1): reference list rl's vector for plane
// region `sub` of block b.
HDEV void mc_select_mv(const FrameGeom &g, const vp9r_block &b, int plane, int sub, int rl, int32_t &mv0,
                       int32_t &mv1) {
  if (plane == 0 || b.mi_size >= BLOCK_8X8) {
    mv0 = b.mv[rl][sub][0];
    mv1 = b.mv[rl][sub][1];
  } else {
    int usx = g.legacy_uv ? 1 : g.ss_x, usy = g.legacy_uv ? 1 : g.ss_y;
    const int32_t(*bm)[2] = b.mv[rl];
    if (!usx && !usy) {
      mv0 = bm[sub][0];
      mv1 = bm[sub][1];
    } else if (!usx) {
      mv0 = q2(bm[sub][0] + bm[sub + 2][0]);
      mv1 = q2(bm[sub][1] + bm[sub + 2][1]);
    } else if (!usy) {
      mv0 = q2(bm[sub][0] + bm[sub + 1][0]);
      mv1 = q2(bm[sub][1] + bm[sub + 1][1]);
    } else {
      mv0 = q4(bm[0][0] + bm[1][0] + bm[2][0] + bm[3][0]);
      mv1 = q4(bm[0][1] + bm[1][1] + bm[2][1] + bm[3][1]);
    }
  }
}

// Clamping (8.5.2.2) and scaling (8.5.2.3) of vector (mv0, mv1) of a block
// at (mi_row, mi_col) of size mi_size predicting from ref_frame, for the
// plane region at (x, y).
HDEV RefPos mc_position_mv(const FrameGeom &g, const uint32_t ref_w[3], const uint32_t ref_h[3], int32_t mv0,
                           int32_t mv1, int mi_row, int mi_col, int mi_size, int ref_frame, int plane, int x,
                           int y) {
  const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
  const int bh = num8x8_high(mi_size), bw = num8x8_wide(mi_size);
  const int to_top = -((mi_row * 8) * 16) >> sy;
  const int to_bottom = ((((int)g.mi_rows - bh - mi_row) * 8) * 16) >> sy;
  const int to_left = -((mi_col * 8) * 16) >> sx;
  const int to_right = ((((int)g.mi_cols - bw - mi_col) * 8) * 16) >> sx;
  const int spel_left = (INTERP_EXTEND + ((bw * 8) >> sx)) << 4, spel_right = spel_left - 16;
  const int spel_top = (INTERP_EXTEND + ((bh * 8) >> sy)) << 4, spel_bottom = spel_top - 16;
  const int32_t c0 = iclamp((2 * mv0) >> sy, to_top - spel_top, to_bottom + spel_bottom);
  const int32_t c1 = iclamp((2 * mv1) >> sx, to_left - spel_left, to_right + spel_right);
  const int ri = ref_frame - LAST_FRAME;
  const uint32_t rw = ref_w[ri], rh = ref_h[ri];
  const int64_t xsc = ((int64_t)rw << 14) / g.width, ysc = ((int64_t)rh << 14) / g.height;
  const int64_t base_x = ((int64_t)x * xsc) >> 14, base_y = ((int64_t)y * ysc) >> 14;
  const int64_t lx = plane ? (int64_t)x << g.ss_x : x, ly = plane ? (int64_t)y << g.ss_y : y;
  const int64_t frac_x = ((16 * lx * xsc) >> 14) & 15, frac_y = ((16 * ly * ysc) >> 14) & 15;
  const int64_t dx = (((int64_t)c1 * xsc) >> 14) + frac_x, dy = (((int64_t)c0 * ysc) >> 14) + frac_y;
  RefPos r;
  r.xs = (int32_t)((16 * xsc) >> 14);
  r.ys = (int32_t)((16 * ysc) >> 14);
  r.x = (int32_t)((base_x << 4) + dx);
  r.y = (int32_t)((base_y << 4) + dy);
  r.last_x = (int)((rw + sx) >> sx) - 1;
  r.last_y = (int)((rh + sy) >> sy) - 1;
  return r;
}

// Selection, clamping and scaling for reference list rl of block b, plane
// region `sub` at (x, y).
HDEV RefPos mc_position(const FrameGeom &g, const uint32_t ref_w[3], const uint32_t ref_h[3], const vp9r_block &b,
                        int plane, int x, int y, int sub, int rl) {
  int32_t mv0, mv1;
  mc_select_mv(g, b, plane, sub, rl, mv0, mv1);
  return mc_position_mv(g, ref_w, ref_h, mv0, mv1, b.mi_row, b.mi_col, b.mi_size, b.ref_frame[rl], plane, x, y);
}

}  // namespace vp9g
