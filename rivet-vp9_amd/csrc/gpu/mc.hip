// Inter prediction (8.5.2): motion vector selection, clamping and scaling
// (src/decoder/recon.rs:106-217, src/decoder/mod.rs:495-541), the separable
// 8-tap filter with per-pass rounding and clamping (src/dsp/inter.rs:289-395)
// and compound averaging (src/dsp/pixel.rs:22-50).
//
// One 64-lane workgroup predicts one tile of at most 16 x 16 samples of a
// prediction unit: its reference footprint (up to 38 x 38 for a 2:1 scaled
// reference) is staged in LDS with the reference's edge samples repeated
// (the specification's clamped coordinates), the horizontal pass writes an
// LDS intermediate, the vertical pass gives each lane up to 4 samples.  One code
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

// A tile is at most 16 x 16 samples; its footprint at most 38 x 38 (a 2:1
// step: ((16 - 1) * 32 + 15) / 16 + 8).
constexpr int TILE = 16;
constexpr int FPW = 40;  // footprint row pitch in LDS
constexpr int FPH = 38;

struct McShared {
  uint16_t fp[FPH * FPW];  // reference footprint, edge samples repeated
  int16_t hp[FPH * TILE];  // after the horizontal pass
};

// Predicts the tw x th (<= 16 x 16) tile whose first sample is at (x0, y0)
// (1/16 sample) with steps (xs, ys) from `ref`, whose bottom-right sample is
// (last_x, last_y).  Output o = lane + 64 i (row o / tw, column o % tw) goes
// to out[i].  The specification's formula (8.5.2.3): each pass rounds by 7
// and clamps; a pass whose positions are all whole samples (step 16, phase
// 0) is the identity and is skipped, with its 7 extra footprint samples.
// All 64 lanes call it (it holds barriers).
template <class P>
DEV void mc_tile(const P *ref, uint32_t stride, int last_x, int last_y, int x0, int y0, int xs, int ys,
                 int tw, int th, int filter, int max, McShared &sh, int out[4]) {
  const int lane = threadIdx.x;
  const bool fx = xs != 16 || (x0 & 15), fy = ys != 16 || (y0 & 15);
  const int fx0 = (x0 >> 4) - (fx ? 3 : 0), fy0 = (y0 >> 4) - (fy ? 3 : 0);
  const int iw = fx ? ((((tw - 1) * xs + (x0 & 15)) >> 4) + 8) : tw;
  const int ih = fy ? ((((th - 1) * ys + (y0 & 15)) >> 4) + 8) : th;
  for (int k = lane; k < ih * iw; k += 64) {
    const int rr = k / iw, cc = k - rr * iw;
    const int sy = iclamp(fy0 + rr, 0, last_y), sx = iclamp(fx0 + cc, 0, last_x);
    sh.fp[rr * FPW + cc] = ref[(size_t)sy * stride + sx];
  }
  __syncthreads();
  for (int k = lane; k < ih * tw; k += 64) {
    const int rr = k / tw, cc = k - rr * tw;
    int v;
    if (fx) {
      const int p = x0 + xs * cc;
      const int16_t *f = c_filters[filter][p & 15];
      const uint16_t *src = sh.fp + rr * FPW + (p >> 4) - 3 - fx0;
      int sum = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) sum += f[i] * (int)src[i];
      v = iclamp((sum + 64) >> 7, 0, max);
    } else {
      v = sh.fp[rr * FPW + cc];
    }
    sh.hp[rr * TILE + cc] = (int16_t)v;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int o = lane + 64 * i;
    int v = 0;
    if (o < th * tw) {
      const int r = o / tw, c = o - r * tw;
      if (fy) {
        const int p = (y0 & 15) + ys * r;
        const int16_t *f = c_filters[filter][p & 15];
        const int16_t *src = sh.hp + (p >> 4) * TILE + c;
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) sum += f[k] * (int)src[k * TILE];
        v = iclamp((sum + 64) >> 7, 0, max);
      } else {
        v = sh.hp[r * TILE + c];
      }
    }
    out[i] = v;
  }
  __syncthreads();
}

template <bool HBD>
__global__ __launch_bounds__(64) void k_mc(const McArgs *fa, const Range *rg, uint32_t nf) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ McShared sh;
  const uint32_t fi = range_frame(rg, nf, blockIdx.x);
  const McArgs &a = fa[fi];
  const McTile t = a.tiles[rg[fi].first + blockIdx.x - rg[fi].wg];
  const vp9r_block &b = a.blocks[t.block];
  const int max = (1 << a.g.bit_depth) - 1;
  const bool compound = b.ref_frame[1] > INTRA_FRAME;
  int out[2][4];
  for (int rl = 0; rl < 1 + (int)compound; rl++) {
    RefPos rp = mc_position(a.g, a.ref_w, a.ref_h, b, t.plane, t.x, t.y, t.sub, rl);
    const int ri = b.ref_frame[rl] - LAST_FRAME;
    const P *ref = (const P *)a.ref[ri].p[t.plane];
    mc_tile<P>(ref, a.ref[ri].stride[t.plane], rp.last_x, rp.last_y, rp.x + rp.xs * t.tx, rp.y + rp.ys * t.ty,
               rp.xs, rp.ys, t.tw, t.th, b.interp_filter, max, sh, out[rl]);
  }
  P *dst = (P *)a.cur.p[t.plane];
  const uint32_t stride = a.cur.stride[t.plane];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int o = threadIdx.x + 64 * i;
    if (o < t.th * t.tw) {
      const int r = o / t.tw, c = o - r * t.tw;
      const int v = compound ? (out[0][i] + out[1][i] + 1) >> 1 : out[0][i];
      dst[(size_t)(t.y + t.ty + r) * stride + t.x + t.tx + c] = (P)v;
    }
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
  __shared__ McShared sh;
  const DescTile t = tiles[blockIdx.x];
  const vp9r_mc_desc d = descs[t.desc];
  const int max = (1 << bd) - 1;
  const int th = d.h - t.ty < TILE ? d.h - t.ty : TILE, tw = d.w - t.tx < TILE ? d.w - t.tx : TILE;
  int o0[4], o1[4];
  mc_tile<P>((const P *)refs[d.ref_idx], ref_stride, last_x, last_y, d.x + d.xs * t.tx, d.y + d.ys * t.ty,
             d.xs, d.ys, tw, th, d.filter, max, sh, o0);
  const bool compound = d.ref2_idx >= 0;
  if (compound)
    mc_tile<P>((const P *)refs[d.ref2_idx], ref_stride, last_x, last_y, d.x2 + d.xs * t.tx,
               d.y2 + d.ys * t.ty, d.xs, d.ys, tw, th, d.filter, max, sh, o1);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int o = threadIdx.x + 64 * i;
    if (o < th * tw) {
      const int r = o / tw, c = o - r * tw;
      const int v = compound ? (o0[i] + o1[i] + 1) >> 1 : o0[i];
      ((P *)out)[d.out_off + (size_t)(t.ty + r) * out_stride + t.tx + c] = (P)v;
    }
  }
}

void launch_mc(hipStream_t s, const McArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd) {
  upload_filters();
  if (!wgs) return;
  if (hbd)
    hipLaunchKernelGGL(k_mc<true>, dim3(wgs), dim3(64), 0, s, fa, r, n);
  else
    hipLaunchKernelGGL(k_mc<false>, dim3(wgs), dim3(64), 0, s, fa, r, n);
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
