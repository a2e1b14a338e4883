// Inverse transform + reconstruction add of the transform blocks of inter
// blocks, batched by size class: a 64-lane workgroup takes 64 / n0 blocks of
// n0 x n0, one lane per row then per column (src/dsp/itx.rs:701-749 with the
// dequantisation of src/decoder/recon.rs:233-281).  Inter blocks' residuals
// are independent of each other, so a whole frame's are one launch per size.
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

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

// ---- per-kernel seam: dense dequantised coefficients per block
//
// A team of n0 lanes takes a block: lane k transforms row k straight from
// memory, the rows become columns, lane k transforms column k, the columns
// become rows again, and lane k adds row k of the residuals to the
// destination (one dword access per 4 bytes of the row).  Two ways to turn
// rows into columns (the autotuner picks per size class):
//   - through an LDS tile of the team's own (row stride n0 + 1), ordered by
//     the wave alone (each team lies inside one wave);
//   - in registers (team_transpose): log2(n0) rounds of lane exchanges,
//     DPP quad permutations and row rotations or ds_swizzle, no LDS.

// v from lane (this lane ^ K), K < 32: DPP within quads (1, 2) and rows (8:
// a rotation by 8 of a 16-lane row), ds_swizzle's bit mode elsewhere.
template <int K>
DEV int32_t lane_xor32(int32_t v) {
  if constexpr (K == 1)
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1, 0, 3, 2]
  else if constexpr (K == 2)
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2, 3, 0, 1]
  else if constexpr (K == 8)
    return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, true);  // row_ror 8
  else
    return __builtin_amdgcn_ds_swizzle(v, 0x1F | (K << 10));  // and 0x1f, or 0, xor K
}
template <int K, class T>
DEV T lane_xor(T v) {
  if constexpr (sizeof(T) == 4) {
    return (T)lane_xor32<K>((int32_t)v);
  } else {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)lane_xor32<K>((int32_t)(uint32_t)u);
    const uint32_t hi = (uint32_t)lane_xor32<K>((int32_t)(uint32_t)(u >> 32));
    return (T)((uint64_t)hi << 32 | lo);
  }
}

// One round of team_transpose: element j and lane l trade bit K where they
// differ, so lane l gives its elements whose bit K differs from its own to
// lane l ^ K and takes that lane's in their place.  m is all ones on lanes
// with bit K set, else zero: each choice is one bitfield select (v_bfi_b32)
// per dword.
template <class T>
DEV T pick(uint32_t m, T a, T b) {  // m ? a : b
  if constexpr (sizeof(T) == 4) {
    return (T)((m & (uint32_t)a) | (~m & (uint32_t)b));
  } else {
    const uint64_t mm = (uint64_t)m << 32 | m;
    return (T)((mm & (uint64_t)a) | (~mm & (uint64_t)b));
  }
}
template <int K, int n0, class T>
DEV void transpose_round(T (&t)[n0], uint32_t m) {
  T send[n0 / 2], recv[n0 / 2];
#pragma unroll
  for (int j = 0, p = 0; j < n0; j++)
    if (!(j & K)) send[p++] = pick(m, t[j], t[j | K]);
#pragma unroll
  for (int p = 0; p < n0 / 2; p++) recv[p] = lane_xor<K>(send[p]);
#pragma unroll
  for (int j = 0, p = 0; j < n0; j++)
    if (!(j & K)) {
      t[j] = pick(m, recv[p], t[j]);
      t[j | K] = pick(m, t[j | K], recv[p]);
      p++;
    }
}

// Lane l of an aligned team of n0 lanes holds row l of an n0 x n0 matrix in
// t; afterwards it holds column l.  Every lane of the wave must take part.
template <int n0, class T>
DEV void team_transpose(T (&t)[n0], int lane) {
  auto mask = [&](int k) { return (lane & k) ? ~0u : 0u; };
  transpose_round<1>(t, mask(1));
  if constexpr (n0 > 2) transpose_round<2>(t, mask(2));
  if constexpr (n0 > 4) transpose_round<4>(t, mask(4));
  if constexpr (n0 > 8) transpose_round<8>(t, mask(8));
  if constexpr (n0 > 16) transpose_round<16>(t, mask(16));
}

struct ItxDenseArgs {
  const int32_t *coefs;
  const uint32_t *coef_off;
  const uint8_t *tx_type;
  const uint16_t *eob;
  const uint64_t *dst_off;
  void *dst;
  const uint32_t *jobs;
  uint32_t stride, n_jobs;
  int bd, lossless;
};

// Block jobs[job] by the team that holds `lane` (its index in the team).
// REG: the transposes in registers (every lane of the wave calls this);
// else through `tile`, the team's n0 x (n0 + 1) LDS tile.
template <int n, bool HBD, bool REG>
DEV void itx_dense_block(const ItxDenseArgs &a, uint32_t job, int lane,
                         typename std::conditional<HBD, Lane64, Lane32>::type::T *tile) {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type L;
  typedef typename L::T T;
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  constexpr int n0 = 1 << n, S = n0 + 1;
  const bool active = job < a.n_jobs;
  const uint32_t i = active ? a.jobs[job] : 0;
  const int tt = active ? a.tx_type[i] : 0;
  const int32_t *c = a.coefs + (active ? a.coef_off[i] : 0);
  const bool dc_only = !HBD && active && a.eob[i] == 1 && tt == DCT_DCT && !a.lossless;
  const bool full = active && !dc_only;
  T t[n0];
  if (full) {
    bool zero = true;
#pragma unroll
    for (int k = 0; k < n0; k++) {
      t[k] = (T)c[lane * n0 + k];
      zero = zero && t[k] == 0;
    }
    if (!zero || a.lossless) tx1d<L, n>(t, (tt & 2) == 0, a.lossless != 0, 2);
  } else {
#pragma unroll
    for (int k = 0; k < n0; k++) t[k] = 0;
  }
  if constexpr (REG) {
    team_transpose<n0>(t, lane);
  } else {
    if (full)
#pragma unroll
      for (int k = 0; k < n0; k++) tile[lane * S + k] = t[k];
    wave_lds_sync();
    if (full)
#pragma unroll
      for (int k = 0; k < n0; k++) t[k] = tile[k * S + lane];
  }
  if (full) {
    tx1d<L, n>(t, (tt & 1) == 0, a.lossless != 0, 0);
    constexpr int shift = n + 2 < 6 ? n + 2 : 6;
#pragma unroll
    for (int k = 0; k < n0; k++) {
      const T v = a.lossless ? t[k] : L::rnd(t[k], shift);
      t[k] = (T)lclamp((int64_t)v, -65536, 65536);
    }
  }
  if constexpr (REG) {
    team_transpose<n0>(t, lane);
  } else {
    // column `lane` back into the tile (only this lane touched it since the
    // last sync), then row `lane` out
    if (full)
#pragma unroll
      for (int k = 0; k < n0; k++) tile[k * S + lane] = t[k];
    wave_lds_sync();
    if (full)
#pragma unroll
      for (int k = 0; k < n0; k++) t[k] = tile[lane * S + k];
    wave_lds_sync();  // the tile's next use writes it
  }
  if (active) {
    int32_t r[n0];
    if (dc_only) {
      const int32_t v = iclamp(dc_only_value(c[0], n), -65536, 65536);
#pragma unroll
      for (int k = 0; k < n0; k++) r[k] = v;
    } else {
#pragma unroll
      for (int k = 0; k < n0; k++) r[k] = (int32_t)t[k];
    }
    P *row = (P *)a.dst + a.dst_off[i] + (size_t)lane * a.stride;
    add_row<P, n0>(row, r, HBD ? (1 << a.bd) - 1 : 255);
  }
}

template <int n, bool HBD, bool REG, int W>
struct ItxDenseShared {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type::T T;
  static constexpr int n0 = 1 << n, TEAMS = REG ? 1 : 64 * W / n0, TILE = REG ? 1 : n0 * (n0 + 1);
  T tiles[TEAMS][TILE];
};

// One workgroup of W waves per 64 W / n0 blocks, in XCD-contiguous runs
// (xcd_order).
template <int n, bool HBD, int W, bool REG>
__global__ __launch_bounds__(64 * W) void k_itx_dense(ItxDenseArgs a) {
  constexpr int n0 = 1 << n, BPW = 64 * W / n0;
  __shared__ ItxDenseShared<n, HBD, REG, W> sh;
  const int team = threadIdx.x / n0, lane = threadIdx.x % n0;
  const uint32_t job = xcd_order(blockIdx.x, gridDim.x) * BPW + team;
  itx_dense_block<n, HBD, REG>(a, job, lane, sh.tiles[REG ? 0 : team]);
}

// Persistent form: as many workgroups as stay resident at once (a multiple
// of 8), each wave looping over units of 64 / n0 blocks.  The units are cut
// into 8 contiguous segments, segment x for the workgroups the hardware
// places on XCD x (blockIdx % 8 == x), whose waves take its units in turn,
// so that the blocks in flight on one XCD are neighbours (one L2).
template <int n, bool HBD, int W, bool REG>
__global__ __launch_bounds__(64 * W) void k_itx_dense_p(ItxDenseArgs a) {
  constexpr int n0 = 1 << n, BPV = 64 / n0;
  __shared__ ItxDenseShared<n, HBD, REG, W> sh;
  const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63, team = l64 / n0, lane = l64 % n0;
  const uint32_t units = (a.n_jobs + BPV - 1) / BPV;
  const uint32_t x = blockIdx.x & 7, wx = (blockIdx.x >> 3) * W + wave, nw = (gridDim.x >> 3) * W;
  const uint32_t lo = (uint32_t)((uint64_t)units * x / 8), hi = (uint32_t)((uint64_t)units * (x + 1) / 8);
  for (uint32_t u = lo + wx; u < hi; u += nw)
    itx_dense_block<n, HBD, REG>(a, u * BPV + team, lane, sh.tiles[REG ? 0 : wave * BPV + team]);
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

// The resident workgroups of a kernel on this device (occupancy x CUs,
// rounded up to a multiple of 8), kept per kernel and device.
template <class K>
static uint32_t resident_grid(K kern, int threads) {
  static std::mutex mu;
  static std::map<std::pair<const void *, int>, uint32_t> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair((const void *)kern, dev);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  const uint32_t g = ((uint32_t)(per_cu * cus) + 7) & ~7u;
  cache[key] = g;
  return g;
}

template <int n, bool HBD, int W, bool REG>
static void itx_dense_go(hipStream_t s, const ItxDenseArgs &a, bool persistent) {
  constexpr int BPW = (64 * W) >> n;
  const uint32_t wgs = (a.n_jobs + BPW - 1) / BPW;
  if (persistent) {
    const uint32_t g = std::min(resident_grid(k_itx_dense_p<n, HBD, W, REG>, 64 * W), (wgs + 7) & ~7u);
    hipLaunchKernelGGL((k_itx_dense_p<n, HBD, W, REG>), dim3(g), dim3(64 * W), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_itx_dense<n, HBD, W, REG>), dim3(wgs), dim3(64 * W), 0, s, a);
  }
}

// shape: waves per workgroup (1, 2 or 4; 4 only where the workgroup's
// tiles fit 64 KB of LDS: 32x32 at 10/12 bits takes 2) | kItxShapeReg
// (transposes in registers: sizes 4x4 .. 16x16) | kItxShapePersistent.
template <int n, bool HBD>
static void itx_dense_n(hipStream_t s, const ItxDenseArgs &a, int shape) {
  const int waves = shape & 7;
  const bool reg = (shape & kItxShapeReg) != 0 && n <= 4, pers = (shape & kItxShapePersistent) != 0;
  if (reg) {
    if constexpr (n <= 4) {
      if (waves >= 4)
        itx_dense_go<n, HBD, 4, true>(s, a, pers);
      else if (waves >= 2)
        itx_dense_go<n, HBD, 2, true>(s, a, pers);
      else
        itx_dense_go<n, HBD, 1, true>(s, a, pers);
    }
    return;
  }
  if (waves >= 4 && !(n == 5 && HBD))
    itx_dense_go<n, HBD, (n == 5 && HBD) ? 2 : 4, false>(s, a, pers);
  else if (waves >= 2)
    itx_dense_go<n, HBD, 2, false>(s, a, pers);
  else
    itx_dense_go<n, HBD, 1, false>(s, a, pers);
}

void launch_itx_dense(hipStream_t s, uint32_t, const int32_t *coefs, const uint32_t *coef_off,
                      const uint8_t *, const uint8_t *tx_type, const uint16_t *eob,
                      const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bit_depth,
                      int lossless, int tx_size, const uint32_t *jobs, uint32_t n_jobs, int shape) {
  if (!n_jobs) return;
  const ItxDenseArgs a{coefs, coef_off, tx_type, eob, dst_off, dst, jobs, stride, n_jobs, (int)bit_depth, lossless};
  const bool hbd = bit_depth > 8;
  switch (tx_size) {
    case 0: hbd ? itx_dense_n<2, true>(s, a, shape) : itx_dense_n<2, false>(s, a, shape); break;
    case 1: hbd ? itx_dense_n<3, true>(s, a, shape) : itx_dense_n<3, false>(s, a, shape); break;
    case 2: hbd ? itx_dense_n<4, true>(s, a, shape) : itx_dense_n<4, false>(s, a, shape); break;
    default: hbd ? itx_dense_n<5, true>(s, a, shape) : itx_dense_n<5, false>(s, a, shape); break;
  }
}

}  // namespace vp9g
