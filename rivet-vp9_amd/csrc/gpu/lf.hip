// The loop filter (8.8): edge decisions per superblock (one parallel
// launch), then the sample filters as a superblock wavefront.
//
// Decisions (lf_mask.h) follow superblock() (src/decoder/loopfilter.rs:
// 466-620), one lane per (edge, 4-sample run) entry, with LvlLookup and the
// limits of Ctx::new (loopfilter.rs:15-47, 135-170) precomputed by the host.
// The filters follow src/dsp/lf.rs:24-143.
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
#include "lf_mask.h"

namespace vp9g {

template <int K>
__global__ __launch_bounds__(256) void k_lf_mask(const LfArgs *fa, const Range *rg, uint32_t nf) {
  lf_mask_wg<K, 256>(fa, rg, nf, blockIdx.x);
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
    clear_margins_wg<P, 256>(fc, rc, nf, blockIdx.x);
  else
    lf_mask_wg<K, 256>(fl, rl, nf, blockIdx.x - n_clear);
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
