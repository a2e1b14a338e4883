// src/dsp/pixel.rs as batched launches for the per-kernel seam: the
// compound average (pixel::avg, pixel.rs:22-50) and the DC-only add
// (pixel::add_const, pixel.rs:9-16), one workgroup per block.  Samples are
// u8 at 8 bits and u16 above, as in the frame stores.
#include <type_traits>

#include "dev_common.h"
#include "launch.h"

namespace vp9g {

// dst = (a + b + 1) >> 1 over w x h samples; a and b have row stride w.
template <class P>
__global__ __launch_bounds__(256) void k_avg_batch(const vp9r_avg_desc *desc, const P *a, const P *b, P *dst,
                                                   uint32_t stride) {
  const vp9r_avg_desc d = desc[blockIdx.x];
  const int w = d.w, n = d.w * d.h;
  for (int k = threadIdx.x; k < n; k += 256) {
    const int r = k / w, c = k - r * w;
    dst[d.dst_off + (size_t)r * stride + c] = (P)(((uint32_t)a[d.a_off + k] + b[d.b_off + k] + 1) >> 1);
  }
}

// The n x n samples at dst + off plus v (v clamped to +-65536 first), the
// result clamped to 0 .. 255 (8 bits only, as the reference).
__global__ __launch_bounds__(256) void k_add_const_batch(const vp9r_add_const_desc *desc, uint8_t *dst,
                                                         uint32_t stride) {
  const vp9r_add_const_desc d = desc[blockIdx.x];
  const int n = d.n, v = min(max(d.v, -65536), 65536);
  for (int k = threadIdx.x; k < n * n; k += 256) {
    const int r = k / n, c = k - r * n;
    uint8_t *p = dst + d.dst_off + (size_t)r * stride + c;
    *p = (uint8_t)min(max((int)*p + v, 0), 255);
  }
}

void launch_avg_batch(hipStream_t s, uint32_t count, const vp9r_avg_desc *desc, const void *a, const void *b,
                      void *dst, uint32_t stride, uint32_t bit_depth) {
  if (!count) return;
  if (bit_depth > 8)
    hipLaunchKernelGGL(k_avg_batch<uint16_t>, dim3(count), dim3(256), 0, s, desc, (const uint16_t *)a,
                       (const uint16_t *)b, (uint16_t *)dst, stride);
  else
    hipLaunchKernelGGL(k_avg_batch<uint8_t>, dim3(count), dim3(256), 0, s, desc, (const uint8_t *)a,
                       (const uint8_t *)b, (uint8_t *)dst, stride);
}

void launch_add_const_batch(hipStream_t s, uint32_t count, const vp9r_add_const_desc *desc, void *dst,
                            uint32_t stride) {
  if (!count) return;
  hipLaunchKernelGGL(k_add_const_batch, dim3(count), dim3(256), 0, s, desc, (uint8_t *)dst, stride);
}

}  // namespace vp9g
