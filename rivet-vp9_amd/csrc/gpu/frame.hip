// Frame store set-up for a reconstruction: the margin clear (frame_dev.h).
#include <type_traits>

#include "dev_common.h"
#include "frame_dev.h"
#include "launch.h"

namespace vp9g {

template <class P>
__global__ __launch_bounds__(256) void k_clear_margins(const ClearArgs *fa, const Range *rg, uint32_t nf) {
  clear_margins_wg<P, 256>(fa, rg, nf, blockIdx.x);
}

// Workgroups per plane the margins of a frame need (count of its range
// over 3): its longest row list in groups of kClearRows.
uint32_t clear_margin_rows(const Planes &f, const FrameGeom &g) {
  uint32_t n = 0;
  for (int p = 0; p < 3; p++) {
    const int sx = p ? g.ss_x : 0, sy = p ? g.ss_y : 0;
    const uint32_t mi_w = (g.mi_cols * 8) >> sx, mi_h = (g.mi_rows * 8) >> sy;
    uint32_t np = f.rows[p] - mi_h;
    if (f.stride[p] > mi_w) np += mi_h;
    np = (np + kClearRows - 1) / kClearRows;
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
