// Clearing of a frame store's margins (device part shared by the clear
// launch and the fused prep launch, lf.hip).
//
// The reference reconstructs into a zero-filled frame (src/decoder/pool.rs:
// buffers are handed out zero-filled) and some of its reads reach past the
// 8x8-aligned area that reconstruction writes: the 4:2:0 loop filter's
// internal 4x4 chroma edge in the last, odd, 8x8 row or column reads the 4
// samples beyond it (loopfilter.rs:558-600).  Frame stores here are reused,
// so every reconstruction first clears what lies outside that area.
#pragma once

#include "dev_common.h"
#include "launch.h"

namespace vp9g {

// Per frame, workgroup j of count = 3 * groups (NT lanes): plane j /
// groups, and j % groups takes kClearRows rows of the plane's row list:
// first the rows below the 8x8-aligned area (whole rows), then the rows of
// the area (their part right of it).
template <class P, int NT>
DEV void clear_margins_wg(const ClearArgs *fa, const Range *rg, uint32_t nf, uint32_t wg) {
  const uint32_t fi = range_frame(rg, nf, wg);
  const ClearArgs &a = fa[fi];
  const uint32_t j = wg - rg[fi].wg, per = rg[fi].count / 3;
  const int plane = j / per;
  const uint32_t r0 = (j - plane * per) * kClearRows;
  const FrameGeom &g = a.g;
  const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
  const uint32_t mi_w = (g.mi_cols * 8) >> sx, mi_h = (g.mi_rows * 8) >> sy;
  const uint32_t rows = a.f.rows[plane], stride = a.f.stride[plane];
  const uint32_t n_bottom = rows - mi_h, n_right = stride > mi_w ? mi_h : 0;
  P *pl = (P *)a.f.p[plane];
  // whole rows below the area
  const uint32_t b_end = min(r0 + kClearRows, n_bottom);
  for (uint32_t r = r0; r < b_end; r++) {
    P *row = pl + (size_t)(mi_h + r) * stride;
    for (uint32_t x = threadIdx.x; x < stride; x += NT) row[x] = 0;
  }
  // the part right of the area of the group's list entries lo .. hi - 1
  // (area rows lo - n_bottom ..), flattened over the group
  const uint32_t lo = max(r0, n_bottom), hi = min(r0 + kClearRows, n_bottom + n_right);
  if (hi > lo) {
    const uint32_t mw = stride - mi_w, n = (hi - lo) * mw;
    P *first = pl + (size_t)(lo - n_bottom) * stride + mi_w;
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const uint32_t r = i / mw;
      first[(size_t)r * stride + (i - r * mw)] = 0;
    }
  }
}

}  // namespace vp9g
