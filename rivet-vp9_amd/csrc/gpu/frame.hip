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

// blockIdx.y: plane; blockIdx.x: first the rows below the 8x8-aligned area
// (whole rows), then the rows of the area (their part right of it).
template <class P>
__global__ __launch_bounds__(256) void k_clear_margins(Planes f, FrameGeom g) {
  const int plane = blockIdx.y;
  const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
  const uint32_t mi_w = (g.mi_cols * 8) >> sx, mi_h = (g.mi_rows * 8) >> sy;
  const uint32_t rows = f.rows[plane], stride = f.stride[plane];
  const uint32_t n_bottom = rows - mi_h;
  P *pl = (P *)f.p[plane];
  const uint32_t r = blockIdx.x;
  if (r < n_bottom) {
    P *row = pl + (size_t)(mi_h + r) * stride;
    for (uint32_t x = threadIdx.x; x < stride; x += 256) row[x] = 0;
  } else if (r - n_bottom < mi_h && stride > mi_w) {
    P *row = pl + (size_t)(r - n_bottom) * stride;
    for (uint32_t x = mi_w + threadIdx.x; x < stride; x += 256) row[x] = 0;
  }
}

void launch_clear_margins(hipStream_t s, const Planes &f, const FrameGeom &g) {
  uint32_t n = 0;
  for (int p = 0; p < 3; p++) {
    const int sx = p ? g.ss_x : 0, sy = p ? g.ss_y : 0;
    const uint32_t mi_w = (g.mi_cols * 8) >> sx, mi_h = (g.mi_rows * 8) >> sy;
    uint32_t np = f.rows[p] - mi_h;
    if (f.stride[p] > mi_w) np += mi_h;
    n = np > n ? np : n;
  }
  if (!n) return;
  if (g.bit_depth > 8)
    hipLaunchKernelGGL(k_clear_margins<uint16_t>, dim3(n, 3), dim3(256), 0, s, f, g);
  else
    hipLaunchKernelGGL(k_clear_margins<uint8_t>, dim3(n, 3), dim3(256), 0, s, f, g);
}

}  // namespace vp9g
