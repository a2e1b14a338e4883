// Frame store set-up for a reconstruction.
//
// The reference reconstructs into a zero-filled frame (src/decoder/pool.rs:
// buffers are handed out zero-filled) and some of its reads reach past the
// 8x8-aligned area that reconstruction writes: the 4:2:0 loop filter's
// internal 4x4 chroma edge in the last, odd, 8x8 row or column reads the 4
// samples beyond it (loopfilter.rs:558-600).  Frame stores here are reused,
// so every reconstruction first clears what lies outside that area.
#include <type_traits>

#include "dev_common.h"
#include "launch.h"

namespace vp9g {

// Per frame, workgroup j of count = 3 * rows: plane j / rows, and j % rows
// counts first the rows below the 8x8-aligned area (whole rows), then the
// rows of the area (their part right of it).
template <class P>
__global__ __launch_bounds__(256) void k_clear_margins(const ClearArgs *fa, const Range *rg, uint32_t nf) {
  const uint32_t fi = range_frame(rg, nf, blockIdx.x);
  const ClearArgs &a = fa[fi];
  const uint32_t j = blockIdx.x - rg[fi].wg, per = rg[fi].count / 3;
  const int plane = j / per;
  const uint32_t r = j - plane * per;
  const FrameGeom &g = a.g;
  const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
  const uint32_t mi_w = (g.mi_cols * 8) >> sx, mi_h = (g.mi_rows * 8) >> sy;
  const uint32_t rows = a.f.rows[plane], stride = a.f.stride[plane];
  const uint32_t n_bottom = rows - mi_h;
  P *pl = (P *)a.f.p[plane];
  if (r < n_bottom) {
    P *row = pl + (size_t)(mi_h + r) * stride;
    for (uint32_t x = threadIdx.x; x < stride; x += 256) row[x] = 0;
  } else if (r - n_bottom < mi_h && stride > mi_w) {
    P *row = pl + (size_t)(r - n_bottom) * stride;
    for (uint32_t x = mi_w + threadIdx.x; x < stride; x += 256) row[x] = 0;
  }
}

// Workgroups per plane the margins of a frame need (count of its range).
uint32_t clear_margin_rows(const Planes &f, const FrameGeom &g) {
  uint32_t n = 0;
  for (int p = 0; p < 3; p++) {
    const int sx = p ? g.ss_x : 0, sy = p ? g.ss_y : 0;
    const uint32_t mi_w = (g.mi_cols * 8) >> sx, mi_h = (g.mi_rows * 8) >> sy;
    uint32_t np = f.rows[p] - mi_h;
    if (f.stride[p] > mi_w) np += mi_h;
    n = np > n ? np : n;
  }
  return n;
}

void launch_clear_margins(hipStream_t s, const ClearArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd) {
  if (!wgs) return;
  if (hbd)
    hipLaunchKernelGGL(k_clear_margins<uint16_t>, dim3(wgs), dim3(256), 0, s, fa, r, n);
  else
    hipLaunchKernelGGL(k_clear_margins<uint8_t>, dim3(wgs), dim3(256), 0, s, fa, r, n);
}

}  // namespace vp9g
