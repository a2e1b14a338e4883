// Inverse transform + reconstruction add of the transform blocks of inter
// blocks, batched by size class: a 64-lane workgroup takes 64 / n0 blocks of
// n0 x n0, one lane per row then per column (src/dsp/itx.rs:701-749 with the
// dequantisation of src/decoder/recon.rs:233-281).  Inter blocks' residuals
// are independent of each other, so a whole frame's are one launch per size.
#include <type_traits>

#include "itx_dev.h"
#include "launch.h"

namespace vp9g {

// Adds residuals r[0..n0) to the n0 samples at `row`, clamped to [0, max]:
// as dwords when the row is 4-byte aligned, else sample by sample.
template <class P, int n0>
DEV void add_row(P *row, const int32_t *r, int max) {
  if (((uintptr_t)row & 3) == 0) {
    constexpr int PER = 4 / (int)sizeof(P), ND = n0 / PER;
    uint32_t w[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) w[d] = gld((const uint32_t *)row + d);
#pragma unroll
    for (int d = 0; d < ND; d++) {
      uint32_t o = 0;
#pragma unroll
      for (int e = 0; e < PER; e++) {
        const int sh = 32 / PER * e;
        const int v = iclamp((int)((w[d] >> sh) & (0xffffffffu >> (32 - 32 / PER))) + r[d * PER + e], 0, max);
        o |= (uint32_t)v << sh;
      }
      gst((uint32_t *)row + d, o);
    }
  } else {
#pragma unroll
    for (int k = 0; k < n0; k++) row[k] = (P)iclamp((int)row[k] + r[k], 0, max);
  }
}

// Stores residuals r[0..n0) as int16 at `row` (clamped): as dwords when the
// row is 4-byte aligned.
template <int n0>
DEV void store_res_row(int16_t *row, const int32_t *r) {
  if (((uintptr_t)row & 3) == 0) {
#pragma unroll
    for (int d = 0; d < n0 / 2; d++) {
      const uint32_t lo = (uint16_t)iclamp(r[2 * d], -32768, 32767), hi = (uint16_t)iclamp(r[2 * d + 1], -32768, 32767);
      gst((uint32_t *)row + d, lo | hi << 16);
    }
  } else {
#pragma unroll
    for (int k = 0; k < n0; k++) row[k] = (int16_t)iclamp(r[k], -32768, 32767);
  }
}

// The 2D transform of the block in `tile` (row stride n0 + 1) by its team's
// lanes, ending row-wise: rows, then columns (residuals back into the tile,
// each lane its own column), then lane k writes row k of the residuals:
// added into the frame at base (clamped to [0, max]) or, with res set,
// stored as int16.  dc_only: every residual is dc_v.  Every lane of the
// workgroup calls this (it holds barriers).
template <class L, int n, class P>
DEV void itx_2d_rows(typename L::T *tile, int lane, bool active, int tx_type, bool lossless, bool dc_only,
                     int32_t dc_v, P *base, int16_t *res, uint32_t stride, int max) {
  typedef typename L::T T;
  constexpr int n0 = 1 << n, S = n0 + 1;
  T t[n0];
  if (active && !dc_only) {
    T *row = tile + lane * S;
    bool zero = true;
#pragma unroll
    for (int k = 0; k < n0; k++) {
      t[k] = row[k];
      zero = zero && t[k] == 0;
    }
    if (!zero || lossless) {
      tx1d<L, n>(t, (tx_type & 2) == 0, lossless, 2);
#pragma unroll
      for (int k = 0; k < n0; k++) row[k] = t[k];
    }
  }
  __syncthreads();
  if (active && !dc_only) {
#pragma unroll
    for (int k = 0; k < n0; k++) t[k] = tile[k * S + lane];
    tx1d<L, n>(t, (tx_type & 1) == 0, lossless, 0);
    constexpr int shift = n + 2 < 6 ? n + 2 : 6;
#pragma unroll
    for (int k = 0; k < n0; k++) {
      const T v = lossless ? t[k] : L::rnd(t[k], shift);
      tile[k * S + lane] = (T)lclamp((int64_t)v, -65536, 65536);
    }
  }
  __syncthreads();
  if (active) {
    int32_t r[n0];
#pragma unroll
    for (int k = 0; k < n0; k++) r[k] = dc_only ? iclamp(dc_v, -65536, 65536) : (int32_t)tile[lane * S + k];
    if (res)
      store_res_row<n0>(res + (size_t)lane * stride, r);
    else
      add_row<P, n0>(base + (size_t)lane * stride, r, max);
  }
}

// Workgroup wg of size class n: per frame of the batch, its range's
// workgroups take BPW consecutive jobs each (count = the frame's jobs of
// this size); tiles: BPW * n0 * (n0 + 1) elements of LDS.
template <int n, bool HBD>
DEV void itx_inter_wg(const ItxArgs *fa, const Range *rg, uint32_t nf, uint32_t wg,
                      typename std::conditional<HBD, Lane64, Lane32>::type::T *tiles) {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type L;
  typedef typename L::T T;
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  constexpr int n0 = 1 << n, BPW = 64 / n0, S = n0 + 1;
  const uint32_t fi = range_frame(rg, nf, wg);
  const ItxArgs &a = fa[fi];
  const int team = threadIdx.x / n0, lane = threadIdx.x % n0;
  const uint32_t local = (wg - rg[fi].wg) * BPW + team;
  const bool active = local < rg[fi].count;
  const uint32_t job = rg[fi].first + local;
  T *tile = tiles + team * (n0 * S);
#pragma unroll
  for (int k = 0; k < S; k++) tile[lane * S + k] = 0;
  __syncthreads();
  vp9r_txb t = {};
  int32_t dcq = 0, acq = 0;
  bool intra = false;
  if (active) {
    const uint32_t jv = a.jobs[n - 2][job];
    intra = (jv & kItxIntra) != 0;
    t = a.txbs[jv & ~kItxIntra];
    const int seg = a.blocks[t.block].segment_id;
    const int pt = t.plane > 0;
    dcq = a.q[seg][pt][0];
    acq = a.q[seg][pt][1];
    const uint16_t *pos = a.coef_pos + t.coef_start;
    const int32_t *val = a.coef_val + t.coef_start;
    for (int k = lane; k < t.ncoef; k += n0) {
      int p = pos[k] & (n0 * n0 - 1);
      tile[(p >> n) * S + (p & (n0 - 1))] = (T)dequant(val[k], p, dcq, acq, n == 5);
    }
  }
  __syncthreads();
  const bool lossless = a.g.lossless;
  // an all-zero block's residual is zero (only intra blocks' are listed)
  const bool dc_only = active && (t.eob == 0 || (!HBD && t.eob == 1 && t.tx_type == DCT_DCT && !lossless));
  int32_t dc_v = 0;
  if (dc_only && t.eob) dc_v = dc_only_value((int32_t)tile[0], n + 0);
  const uint32_t stride = a.cur.stride[t.plane];
  P *base = (P *)a.cur.p[t.plane] + (size_t)t.y * stride + t.x;
  int16_t *res = intra ? a.res[t.plane] + (size_t)t.y * stride + t.x : nullptr;
  const int max = (1 << a.g.bit_depth) - 1;
  itx_2d_rows<L, n, P>(tile, lane, active, t.tx_type, lossless, dc_only, dc_v, base, res, stride,
                       dc_only && !HBD ? 255 : max);
}

// All four size classes in one launch (they are independent): workgroups
// [0, b4) take 32x32 blocks, [b4, b8) 16x16, [b8, b16) 8x8, the rest 4x4
// (the largest, slowest, first).  LDS: the largest class's tiles.
template <bool HBD>
__global__ __launch_bounds__(64) void k_itx_inter(const ItxArgs *fa, const Range *r32, const Range *r16,
                                                  const Range *r8, const Range *r4, uint32_t nf, uint32_t b4,
                                                  uint32_t b8, uint32_t b16) {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type L;
  __shared__ typename L::T tiles[2 * 32 * 33];
  const uint32_t w = blockIdx.x;
  if (w < b4)
    itx_inter_wg<5, HBD>(fa, r32, nf, w, tiles);
  else if (w < b8)
    itx_inter_wg<4, HBD>(fa, r16, nf, w - b4, tiles);
  else if (w < b16)
    itx_inter_wg<3, HBD>(fa, r8, nf, w - b8, tiles);
  else
    itx_inter_wg<2, HBD>(fa, r4, nf, w - b16, tiles);
}

// Per-kernel seam: dense dequantised coefficients per block.  A team of n0
// lanes takes a block: lane k transforms row k straight from memory, the
// columns go through the LDS tile, the column residuals go back into the
// tile, and lane k adds row k of them to the destination (one dword access
// per 4 bytes of the row).
// W waves per workgroup (the autotuner's choice per size and bit depth class).
template <int n, bool HBD, int W>
__global__ __launch_bounds__(64 * W) void k_itx_dense(const int32_t *coefs, const uint32_t *coef_off,
                                                  const uint8_t *tx_type, const uint16_t *eob,
                                                  const uint64_t *dst_off, void *dst_base,
                                                  uint32_t stride, int bd, int lossless,
                                                  const uint32_t *jobs, uint32_t n_jobs) {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type L;
  typedef typename L::T T;
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  constexpr int n0 = 1 << n, BPW = 64 * W / n0, S = n0 + 1;
  __shared__ T tiles[BPW][n0 * S];
  const int team = threadIdx.x / n0, lane = threadIdx.x % n0;
  const uint32_t job = xcd_order(blockIdx.x, gridDim.x) * BPW + team;
  const bool active = job < n_jobs;
  T *tile = tiles[team];
  const uint32_t i = active ? jobs[job] : 0;
  const int tt = active ? tx_type[i] : 0;
  const int32_t *c = coefs + (active ? coef_off[i] : 0);
  const bool dc_only = !HBD && active && eob[i] == 1 && tt == DCT_DCT && !lossless;
  T t[n0];
  if (active && !dc_only) {
    bool zero = true;
#pragma unroll
    for (int k = 0; k < n0; k++) {
      t[k] = (T)c[lane * n0 + k];
      zero = zero && t[k] == 0;
    }
    if (!zero || lossless) tx1d<L, n>(t, (tt & 2) == 0, lossless != 0, 2);
#pragma unroll
    for (int k = 0; k < n0; k++) tile[lane * S + k] = t[k];
  }
  __syncthreads();
  const int max = (1 << bd) - 1;
  if (active && !dc_only) {
    // column `lane`; only this lane reads or writes it until the barrier
#pragma unroll
    for (int k = 0; k < n0; k++) t[k] = tile[k * S + lane];
    tx1d<L, n>(t, (tt & 1) == 0, lossless != 0, 0);
    constexpr int shift = n + 2 < 6 ? n + 2 : 6;
#pragma unroll
    for (int k = 0; k < n0; k++) {
      const T v = lossless ? t[k] : L::rnd(t[k], shift);
      tile[k * S + lane] = (T)lclamp((int64_t)v, -65536, 65536);
    }
  }
  __syncthreads();
  if (active) {
    int32_t r[n0];
    if (dc_only) {
      const int32_t v = iclamp(dc_only_value(c[0], n), -65536, 65536);
#pragma unroll
      for (int k = 0; k < n0; k++) r[k] = v;
    } else {
#pragma unroll
      for (int k = 0; k < n0; k++) r[k] = (int32_t)tile[lane * S + k];
    }
    P *row = (P *)dst_base + dst_off[i] + (size_t)lane * stride;
    add_row<P, n0>(row, r, HBD ? max : 255);
  }
}

void launch_itx_inter(hipStream_t s, const ItxArgs *fa, const Range *const r[4], uint32_t nf,
                      const uint32_t wgs[4], bool hbd) {
  const uint32_t b4 = wgs[3], b8 = b4 + wgs[2], b16 = b8 + wgs[1], total = b16 + wgs[0];
  if (!total) return;
  if (hbd)
    hipLaunchKernelGGL(k_itx_inter<true>, dim3(total), dim3(64), 0, s, fa, r[3], r[2], r[1], r[0], nf, b4, b8, b16);
  else
    hipLaunchKernelGGL(k_itx_inter<false>, dim3(total), dim3(64), 0, s, fa, r[3], r[2], r[1], r[0], nf, b4, b8, b16);
}

template <int n, int W>
static void itx_dense_nw(hipStream_t s, const int32_t *coefs, const uint32_t *coef_off, const uint8_t *tx_type,
                         const uint16_t *eob, const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bd,
                         int lossless, const uint32_t *jobs, uint32_t n_jobs) {
  constexpr int BPW = (64 * W) >> n;
  dim3 grid((n_jobs + BPW - 1) / BPW);
  if (bd > 8)
    hipLaunchKernelGGL((k_itx_dense<n, true, W>), grid, dim3(64 * W), 0, s, coefs, coef_off, tx_type, eob,
                       dst_off, dst, stride, (int)bd, lossless, jobs, n_jobs);
  else
    hipLaunchKernelGGL((k_itx_dense<n, false, W>), grid, dim3(64 * W), 0, s, coefs, coef_off, tx_type, eob,
                       dst_off, dst, stride, (int)bd, lossless, jobs, n_jobs);
}

// waves: 1, 2 or 4 per workgroup (4 only where the workgroup's tiles fit
// 64 KB of LDS: 32 x 32 at 10/12 bits takes 2).
template <int n>
static void itx_dense_n(hipStream_t s, const int32_t *coefs, const uint32_t *coef_off, const uint8_t *tx_type,
                        const uint16_t *eob, const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bd,
                        int lossless, const uint32_t *jobs, uint32_t n_jobs, int waves) {
  if (waves >= 4 && !(n == 5 && bd > 8))
    itx_dense_nw<n, 4>(s, coefs, coef_off, tx_type, eob, dst_off, dst, stride, bd, lossless, jobs, n_jobs);
  else if (waves >= 2)
    itx_dense_nw<n, 2>(s, coefs, coef_off, tx_type, eob, dst_off, dst, stride, bd, lossless, jobs, n_jobs);
  else
    itx_dense_nw<n, 1>(s, coefs, coef_off, tx_type, eob, dst_off, dst, stride, bd, lossless, jobs, n_jobs);
}

void launch_itx_dense(hipStream_t s, uint32_t, const int32_t *coefs, const uint32_t *coef_off,
                      const uint8_t *, const uint8_t *tx_type, const uint16_t *eob,
                      const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bit_depth,
                      int lossless, int tx_size, const uint32_t *jobs, uint32_t n_jobs, int waves) {
  if (!n_jobs) return;
  switch (tx_size) {
    case 0: itx_dense_n<2>(s, coefs, coef_off, tx_type, eob, dst_off, dst, stride, bit_depth, lossless, jobs, n_jobs, waves); break;
    case 1: itx_dense_n<3>(s, coefs, coef_off, tx_type, eob, dst_off, dst, stride, bit_depth, lossless, jobs, n_jobs, waves); break;
    case 2: itx_dense_n<4>(s, coefs, coef_off, tx_type, eob, dst_off, dst, stride, bit_depth, lossless, jobs, n_jobs, waves); break;
    default: itx_dense_n<5>(s, coefs, coef_off, tx_type, eob, dst_off, dst, stride, bit_depth, lossless, jobs, n_jobs, waves); break;
  }
}

}  // namespace vp9g
