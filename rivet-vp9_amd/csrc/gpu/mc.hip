// Inter prediction (8.5.2): motion vector selection, clamping and scaling
// (src/decoder/recon.rs:106-217, src/decoder/mod.rs:495-541), the separable
// 8-tap filter with per-pass rounding and clamping (src/dsp/inter.rs:289-395)
// and compound averaging (src/dsp/pixel.rs:22-50).
//
// One 64-lane workgroup predicts one tile of at most 8 x 8 samples of a
// prediction unit: its reference footprint (up to 22 x 22 for a 2:1 scaled
// reference) is staged in LDS with the reference's edge samples repeated
// (the specification's clamped coordinates), the horizontal pass writes an
// LDS intermediate, the vertical pass gives each lane its sample.  One code
// path serves every step: the specification's formula, of which the
// reference's unscaled fast path is a special case (its own test,
// inter.rs:486-564, checks they agree).
#include <type_traits>

#include "dev_common.h"
#include "launch.h"
#include "mc_pos.h"
#include "vp9_consts.h"
#include "vp9_tables.h"

namespace vp9g {

__constant__ int16_t c_filters[4][16][8];

static bool filters_ready = false;
static void upload_filters() {
  if (filters_ready) return;
  int16_t f[4][16][8];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 16; b++)
      for (int c = 0; c < 8; c++) f[a][b][c] = (int16_t)VP9_SUBPEL_FILTERS[a][b][c];
  hipMemcpyToSymbol(HIP_SYMBOL(c_filters), f, sizeof f);
  filters_ready = true;
}

constexpr int FP_W = 24;  // footprint row pitch in LDS (>= 22)
constexpr int FP_H = 24;

// Predicts the 8 x 8 (or smaller) tile whose first sample is at (x0, y0)
// (1/16 sample) with steps (xs, ys) from `ref`; returns lane (r, c)'s sample.
// All 64 lanes call it (it holds barriers).
template <class P>
DEV int mc_tile(const P *ref, uint32_t stride, int last_x, int last_y, int x0, int y0, int xs, int ys,
                int th, int filter, int max, uint16_t *fp, int16_t *hp) {
  const int lane = threadIdx.x;
  const int fx0 = (x0 >> 4) - 3, fy0 = (y0 >> 4) - 3;
  const int ih = ((((th - 1) * ys + 15) >> 4) + 8);
  for (int k = lane; k < ih * FP_W; k += 64) {
    int rr = k / FP_W, cc = k - rr * FP_W;
    int sy = iclamp(fy0 + rr, 0, last_y), sx = iclamp(fx0 + cc, 0, last_x);
    fp[k] = ref[(size_t)sy * stride + sx];
  }
  __syncthreads();
  for (int k = lane; k < ih * 8; k += 64) {
    int rr = k >> 3, cc = k & 7;
    int p = x0 + xs * cc;
    const int16_t *f = c_filters[filter][p & 15];
    const uint16_t *s = fp + rr * FP_W + (p >> 4) - 3 - fx0;
    int sum = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) sum += f[i] * (int)s[i];
    hp[k] = (int16_t)iclamp((sum + 64) >> 7, 0, max);
  }
  __syncthreads();
  const int r = lane >> 3, c = lane & 7;
  int p = (y0 & 15) + ys * r;
  const int16_t *f = c_filters[filter][p & 15];
  const int16_t *s = hp + (p >> 4) * 8 + c;
  int sum = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) sum += f[i] * (int)s[i * 8];
  __syncthreads();
  return iclamp((sum + 64) >> 7, 0, max);
}

template <bool HBD>
__global__ __launch_bounds__(64) void k_mc(McArgs a) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ uint16_t fp[FP_W * FP_H];
  __shared__ int16_t hp[FP_H * 8 + 8 * 8];
  const McTile t = a.tiles[blockIdx.x];
  const vp9r_block &b = a.blocks[t.block];
  const int max = (1 << a.g.bit_depth) - 1;
  const bool compound = b.ref_frame[1] > INTRA_FRAME;
  int out[2] = {0, 0};
  for (int rl = 0; rl < 1 + (int)compound; rl++) {
    RefPos rp = mc_position(a.g, a.ref_w, a.ref_h, b, t.plane, t.x, t.y, t.sub, rl);
    const int ri = b.ref_frame[rl] - LAST_FRAME;
    const P *ref = (const P *)a.ref[ri].p[t.plane];
    out[rl] = mc_tile<P>(ref, a.ref[ri].stride[t.plane], rp.last_x, rp.last_y, rp.x + rp.xs * t.tx,
                         rp.y + rp.ys * t.ty, rp.xs, rp.ys, t.th, b.interp_filter, max, fp, hp);
  }
  const int r = threadIdx.x >> 3, c = threadIdx.x & 7;
  if (r < t.th && c < t.tw) {
    int v = compound ? (out[0] + out[1] + 1) >> 1 : out[0];
    P *dst = (P *)a.cur.p[t.plane];
    dst[(size_t)(t.y + t.ty + r) * a.cur.stride[t.plane] + t.x + t.tx + c] = (P)v;
  }
}

// Per-kernel seam: explicit positions (vp9r_mc_desc), tiles expanded on
// the host as (desc index, tile x, tile y).
struct DescTile {
  uint32_t desc;
  uint16_t tx, ty;
};

template <bool HBD>
__global__ __launch_bounds__(64) void k_mc_desc(const vp9r_mc_desc *descs, const DescTile *tiles,
                                                const void *const *refs, uint32_t ref_stride,
                                                int32_t last_x, int32_t last_y, void *out,
                                                uint32_t out_stride, int bd) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ uint16_t fp[FP_W * FP_H];
  __shared__ int16_t hp[FP_H * 8 + 8 * 8];
  const DescTile t = tiles[blockIdx.x];
  const vp9r_mc_desc d = descs[t.desc];
  const int max = (1 << bd) - 1;
  const int th = d.h - t.ty < 8 ? d.h - t.ty : 8, tw = d.w - t.tx < 8 ? d.w - t.tx : 8;
  int o0 = mc_tile<P>((const P *)refs[d.ref_idx], ref_stride, last_x, last_y, d.x + d.xs * t.tx,
                      d.y + d.ys * t.ty, d.xs, d.ys, th, d.filter, max, fp, hp);
  int v = o0;
  if (d.ref2_idx >= 0) {
    int o1 = mc_tile<P>((const P *)refs[d.ref2_idx], ref_stride, last_x, last_y, d.x2 + d.xs * t.tx,
                        d.y2 + d.ys * t.ty, d.xs, d.ys, th, d.filter, max, fp, hp);
    v = (o0 + o1 + 1) >> 1;
  }
  const int r = threadIdx.x >> 3, c = threadIdx.x & 7;
  if (r < th && c < tw) ((P *)out)[d.out_off + (size_t)(t.ty + r) * out_stride + t.tx + c] = (P)v;
}

void launch_mc(hipStream_t s, const McArgs &a, uint32_t n_tiles) {
  upload_filters();
  if (!n_tiles) return;
  if (a.g.bit_depth > 8)
    hipLaunchKernelGGL(k_mc<true>, dim3(n_tiles), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(k_mc<false>, dim3(n_tiles), dim3(64), 0, s, a);
}

void launch_mc_desc_tiles(hipStream_t s, const vp9r_mc_desc *desc, const void *tiles, uint32_t n_tiles,
                          const void *const *refs, uint32_t ref_stride, int32_t last_x,
                          int32_t last_y, void *out, uint32_t out_stride, uint32_t bd) {
  upload_filters();
  if (!n_tiles) return;
  if (bd > 8)
    hipLaunchKernelGGL(k_mc_desc<true>, dim3(n_tiles), dim3(64), 0, s, desc, (const DescTile *)tiles,
                       refs, ref_stride, last_x, last_y, out, out_stride, (int)bd);
  else
    hipLaunchKernelGGL(k_mc_desc<false>, dim3(n_tiles), dim3(64), 0, s, desc, (const DescTile *)tiles,
                       refs, ref_stride, last_x, last_y, out, out_stride, (int)bd);
}

}  // namespace vp9g
