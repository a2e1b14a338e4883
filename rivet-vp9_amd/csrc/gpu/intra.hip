// Intra prediction and reconstruction as a dependency wavefront.
//
// A workgroup owns one intra block (k_intra_blocks) or one (superblock,
// plane) group (k_intra): it stages the region and the row above / column
// left of it in LDS, then, for every transform block in decode order,
// builds the edges (src/decoder/recon.rs:17-103), predicts
// (src/dsp/intra.rs:19-175: each directional mode in its per-sample closed
// form), dequantises and adds the inverse transform (itx_dev.h), and finally
// writes the region back.  A launch runs every item of one dependency
// level: intra blocks read the reconstruction of the blocks left, above and
// above-left of them only (above-right reads stay inside the block,
// recon.rs:61-69), so the host levels the items by the longest chain of
// such dependencies between intra items (inter blocks are complete when
// intra starts).  A launch may cover the items of several frames.
#include <type_traits>

#include "itx_dev.h"
#include "launch.h"

namespace vp9g {

constexpr int TS = 65;  // LDS tile pitch: column x0 - 1 .. x0 + 63

DEV int r2(int x, int n) { return (x + (1 << (n - 1))) >> n; }

// Sample (i, j) of the prediction for `mode` (8.5.1.2); A[-1..2*size-1] is
// the above row, L[0..size-1] the left column.
DEV int intra_sample(int mode, int i, int j, int size, const int32_t *A, const int32_t *L, int dc,
                     int max) {
  switch (mode) {
    case V_PRED:
      return A[j];
    case H_PRED:
      return L[i];
    case D207_PRED: {
      int k = j >> 1, jj = j & 1, ii = i + k;
      if (ii >= size - 1) return L[size - 1];
      if (jj == 0) return r2(L[ii] + L[ii + 1], 1);
      if (ii == size - 2) return r2(L[size - 2] + 3 * L[size - 1], 2);
      return r2(L[ii] + 2 * L[ii + 1] + L[ii + 2], 2);
    }
    case D45_PRED: {
      int k = i + j;
      return k + 2 < 2 * size ? r2(A[k] + 2 * A[k + 1] + A[k + 2], 2) : A[2 * size - 1];
    }
    case D63_PRED: {
      int k = (i >> 1) + j;
      return (i & 1) ? r2(A[k] + 2 * A[k + 1] + A[k + 2], 2) : r2(A[k] + A[k + 1], 1);
    }
    case D117_PRED: {
      int k = min(i >> 1, j);
      i -= 2 * k;
      j -= k;
      if (i == 0) return r2(A[j - 1] + A[j], 1);
      if (i == 1) return j == 0 ? r2(L[0] + 2 * A[-1] + A[0], 2) : r2(A[j - 2] + 2 * A[j - 1] + A[j], 2);
      if (i == 2) return r2(A[-1] + 2 * L[0] + L[1], 2);
      return r2(L[i - 3] + 2 * L[i - 2] + L[i - 1], 2);
    }
    case D135_PRED: {
      int k = min(i, j);
      i -= k;
      j -= k;
      if (i == 0) return j == 0 ? r2(L[0] + 2 * A[-1] + A[0], 2) : r2(A[j - 2] + 2 * A[j - 1] + A[j], 2);
      if (i == 1) return r2(A[-1] + 2 * L[0] + L[1], 2);
      return r2(L[i - 2] + 2 * L[i - 1] + L[i], 2);
    }
    case D153_PRED: {
      int k = min(i, j >> 1);
      i -= k;
      j -= 2 * k;
      if (j == 0) return i == 0 ? r2(L[0] + A[-1], 1) : r2(L[i - 1] + L[i], 1);
      if (j == 1) {
        if (i == 0) return r2(L[0] + 2 * A[-1] + A[0], 2);
        if (i == 1) return r2(A[-1] + 2 * L[0] + L[1], 2);
        return r2(L[i - 2] + 2 * L[i - 1] + L[i], 2);
      }
      return r2(A[j - 3] + 2 * A[j - 2] + A[j - 1], 2);  // i == 0
    }
    case TM_PRED:
      return iclamp(A[j] + L[i] - A[-1], 0, max);
    default:
      return dc;
  }
}

// DC_PRED's value (intra.rs:145-168).
DEV int dc_value(int log2, const int32_t *A, const int32_t *L, bool have_left, bool have_above, int bd) {
  int size = 1 << log2, s = 0;
  if (have_left && have_above) {
    for (int k = 0; k < size; k++) s += L[k] + A[k];
    return (s + size) >> (log2 + 1);
  }
  if (have_left || have_above) {
    const int32_t *e = have_left ? L : A;
    for (int k = 0; k < size; k++) s += e[k];
    return (s + (1 << (log2 - 1))) >> log2;
  }
  return 1 << (bd - 1);
}

struct LdsDst {
  uint16_t *base;
  DEV void add(int i, int j, int r, int max) {
    uint16_t *p = base + i * TS + j;
    *p = (uint16_t)iclamp((int)*p + r, 0, max);
  }
};

template <class L, int n>
DEV void intra_residual(typename L::T *coef, uint16_t *tile, int lx, int ly, const vp9r_txb &t,
                        const IntraArgs &a, int seg, bool hbd, const uint16_t *pos, const int32_t *val) {
  typedef typename L::T T;
  constexpr int n0 = 1 << n, S = n0 + 1;
  const int lane = threadIdx.x;
  const int pt = t.plane > 0;
  const int32_t dcq = a.q[seg][pt][0], acq = a.q[seg][pt][1];
  // the first 64 coefficients are fetched before the tile is cleared, so
  // their latency overlaps the clearing
  int p0 = -1;
  int32_t v0 = 0;
  if (lane < t.ncoef) {
    p0 = pos[lane] & (n0 * n0 - 1);
    v0 = val[lane];
  }
  for (int k = lane; k < n0 * S; k += 64) coef[k] = 0;
  __syncthreads();
  if (p0 >= 0) coef[(p0 >> n) * S + (p0 & (n0 - 1))] = (T)dequant(v0, p0, dcq, acq, n == 5);
  for (int k = lane + 64; k < t.ncoef; k += 64) {
    int p = pos[k] & (n0 * n0 - 1);
    coef[(p >> n) * S + (p & (n0 - 1))] = (T)dequant(val[k], p, dcq, acq, n == 5);
  }
  __syncthreads();
  const bool lossless = a.g.lossless;
  const bool dc_only = !hbd && t.eob == 1 && t.tx_type == DCT_DCT && !lossless;
  const int32_t dc_v = dc_only ? dc_only_value((int32_t)coef[0], n) : 0;
  LdsDst dst{tile + ly * TS + lx};
  itx_2d<L, n>(coef, lane, t.tx_type, lossless, dc_only, dc_v, dst, (1 << a.g.bit_depth) - 1);
}

// Staging of an intra block's transform block records and coefficients
// (they are read in a dependent chain per transform block otherwise).
constexpr int kStageTxbs = 192, kStageCoefs = 2048;

struct IntraShared {
  uint16_t tile[TS * TS];  // rows y0 - 1 .., columns x0 - 1 .. of the region
  int32_t edge[65 + 32];
  vp9r_txb txb[kStageTxbs];
  uint16_t pos[kStageCoefs];
  int32_t val[kStageCoefs];
};

// The fields of a block that intra reconstruction reads, in registers (the
// sub-block modes packed, so that no array is indexed at run time).
struct IntraBlk {
  int mi_row, mi_col, mi_size, y_mode, uv_mode, segment_id;
  bool avail_u, avail_l;
  uint32_t sub_modes;  // 4 x 8 bits
};

DEV IntraBlk intra_blk(const vp9r_block *p) {
  IntraBlk b;
  b.mi_row = p->mi_row;
  b.mi_col = p->mi_col;
  b.mi_size = p->mi_size;
  b.y_mode = p->y_mode;
  b.uv_mode = p->uv_mode;
  b.segment_id = p->segment_id & 7;
  b.avail_u = p->avail_u;
  b.avail_l = p->avail_l;
  b.sub_modes = *(const uint32_t *)p->sub_modes;
  return b;
}

// Predicts and reconstructs transform block t in the LDS tile whose sample
// (0, 0) is plane position (x0 - 1, y0 - 1) (edges: recon.rs:17-103).
template <class L, bool HBD>
DEV void intra_txb(const IntraArgs &a, IntraShared &sh, typename L::T *coef, const vp9r_txb &t,
                   const IntraBlk &b, int x0, int y0, const uint16_t *pos, const int32_t *val) {
  const FrameGeom &g = a.g;
  const int lane = threadIdx.x;
  const int plane = t.plane;
  const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
  const int max_x = (int)((g.mi_cols * 8) >> sx) - 1, max_y = (int)((g.mi_rows * 8) >> sy) - 1;
  const int base = 1 << (g.bit_depth - 1);
  const int maxv = (1 << g.bit_depth) - 1;
  uint16_t *tile = sh.tile;
  int32_t *above = sh.edge, *left = sh.edge + 65;
  const int32_t *A = above + 1;
#define TILE(py, px) ((int)tile[((py) - y0 + 1) * TS + ((px) - x0 + 1)])
  const int tx = t.tx_size, log2 = tx + 2, size = 1 << log2;
  const int bsize = b.mi_size < BLOCK_8X8 ? BLOCK_8X8 : b.mi_size;
  const int n4w = num4x4_wide(bsize) >> sx;
  const int base_x = (b.mi_col * 8) >> sx, base_y = (b.mi_row * 8) >> sy;
  const int xi = (t.x - base_x) >> 2, yi = (t.y - base_y) >> 2;
  const int step = 1 << tx;
  const bool have_left = b.avail_l || xi > 0, have_above = b.avail_u || yi > 0;
  const bool not_on_right = xi + step < n4w;
  const int block_idx = (yi / step) * ((n4w + step - 1) / step) + xi / step;
  const int mode =
      plane > 0 ? b.uv_mode : (b.mi_size >= BLOCK_8X8 ? b.y_mode : (int)((b.sub_modes >> (8 * (block_idx & 3))) & 0xff));
  const int x = t.x, y = t.y;
  // aboveRow[-1 .. 2 * size - 1]: 2 * size + 1 entries, 65 at 32x32, so
  // lane 0 also fills the last one
  for (int k = lane; k <= 2 * size; k += 64) {
    int v;
    if (!have_above) {
      v = base - 1;
    } else if (k == 0) {
      v = have_left ? TILE(y - 1, min(max_x, x - 1)) : base + 1;
    } else {
      int i = k - 1;
      if (i >= size && !(not_on_right && tx == TX_4X4)) i = size - 1;
      v = TILE(y - 1, min(max_x, x + i));
    }
    above[k] = v;
  }
  if (lane < size) left[lane] = have_left ? TILE(min(max_y, y + lane), x - 1) : base + 1;
#undef TILE
  __syncthreads();
  const int dc = mode == DC_PRED ? dc_value(log2, A, left, have_left, have_above, g.bit_depth) : 0;
  const int lx = x - x0 + 1, ly = y - y0 + 1;
  for (int k = lane; k < size * size; k += 64) {
    int i = k >> log2, j = k & (size - 1);
    tile[(ly + i) * TS + lx + j] = (uint16_t)intra_sample(mode, i, j, size, A, left, dc, maxv);
  }
  __syncthreads();
  if (t.eob > 0) {
    const int seg = b.segment_id;
    switch (tx) {
      case 0: intra_residual<L, 2>(coef, tile, lx, ly, t, a, seg, HBD, pos, val); break;
      case 1: intra_residual<L, 3>(coef, tile, lx, ly, t, a, seg, HBD, pos, val); break;
      case 2: intra_residual<L, 4>(coef, tile, lx, ly, t, a, seg, HBD, pos, val); break;
      default: intra_residual<L, 5>(coef, tile, lx, ly, t, a, seg, HBD, pos, val); break;
    }
    __syncthreads();
  }
}

// ---- one wavefront level of (superblock, plane) groups per launch

template <bool HBD>
__global__ __launch_bounds__(64) void k_intra(const IntraArgs *fa, const Range *rg, uint32_t nf) {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type L;
  typedef typename L::T T;
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ IntraShared sh;
  __shared__ T coef[32 * 33];
  const int lane = threadIdx.x;
  const uint32_t fi = range_frame(rg, nf, blockIdx.x);
  const IntraArgs &a = fa[fi];
  const IntraGroup gr = a.groups[rg[fi].first + blockIdx.x - rg[fi].wg];
  const FrameGeom &g = a.g;
  const int plane = gr.plane;
  const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
  const int SW = 64 >> sx, SH = 64 >> sy;
  const int x0 = gr.sb_col * SW, y0 = gr.sb_row * SH;
  const uint32_t stride = a.cur.stride[plane];
  P *pl = (P *)a.cur.p[plane];
  const uint32_t dsw = SW + 1, msw = ((1u << 20) + dsw - 1) / dsw;  // k / (SW + 1)
  for (int k = lane; k < (SH + 1) * (SW + 1); k += 64) {
    const int r = (int)(((uint32_t)k * msw) >> 20), c = k - r * (SW + 1);
    const int py = y0 - 1 + r, px = x0 - 1 + c;
    if (py >= 0 && px >= 0) sh.tile[r * TS + c] = pl[(size_t)py * stride + px];
  }
  __syncthreads();
  for (uint32_t ti = 0; ti < gr.count; ti++) {
    const vp9r_txb t = a.txbs[a.txb_index[gr.first + ti]];
    intra_txb<L, HBD>(a, sh, coef, t, intra_blk(a.blocks + t.block), x0, y0, a.coef_pos + t.coef_start,
                      a.coef_val + t.coef_start);
  }
  const int lsw = 31 - __clz(SW);  // SW is 64 or 32
  for (int k = lane; k < SH * SW; k += 64) {
    const int r = k >> lsw, c = k & (SW - 1);
    pl[(size_t)(y0 + r) * stride + x0 + c] = (P)sh.tile[(r + 1) * TS + c + 1];
  }
}

// Reconstructs intra block `it` (all its planes, transform blocks in
// order) in LDS: the row above it (from above-left) and the column left of
// it come from the frame, the block goes back to the frame.
template <class L, bool HBD>
DEV void intra_block(const IntraArgs &a, IntraShared &sh, typename L::T *coef, const IntraItem &it) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  const FrameGeom &g = a.g;
  const int lane = threadIdx.x;
  const IntraBlk b = intra_blk(a.blocks + it.block);
  const int mi_row = b.mi_row, mi_col = b.mi_col;
  const int bw8 = b.mi_size < BLOCK_8X8 ? 1 : num8x8_wide(b.mi_size);
  const int bh8 = b.mi_size < BLOCK_8X8 ? 1 : num8x8_high(b.mi_size);
  // stage the transform block records and the coefficients they cover
  const uint32_t n_st = it.count < (uint32_t)kStageTxbs ? it.count : (uint32_t)kStageTxbs;
  for (uint32_t k = lane; k < n_st; k += 64) sh.txb[k] = a.txbs[it.first + k];
  const uint32_t c0 = a.txbs[it.first].coef_start;
  const vp9r_txb last = a.txbs[it.first + it.count - 1];
  uint32_t c1 = last.coef_start + last.ncoef;
  if (c1 < c0) c1 = c0;
  const uint32_t n_c = c1 - c0 < (uint32_t)kStageCoefs ? c1 - c0 : (uint32_t)kStageCoefs;
  for (uint32_t k = lane; k < n_c; k += 64) {
    sh.pos[k] = a.coef_pos[c0 + k];
    sh.val[k] = a.coef_val[c0 + k];
  }
  __syncthreads();
  uint32_t k0 = 0;
  for (int plane = 0; plane < 3; plane++) {
    const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
    const int x0 = (mi_col * 8) >> sx, y0 = (mi_row * 8) >> sy;
    const int BW = (bw8 * 8) >> sx, BH = (bh8 * 8) >> sy;
    const uint32_t stride = a.cur.stride[plane];
    P *pl = (P *)a.cur.p[plane];
    // the row above (x0 - 1 .. x0 + BW - 1) and the column left; the block
    // itself from the frame too (its samples outside the 8x8-aligned area
    // are not reconstructed and keep the frame's)
    const uint32_t d = BW + 1, m = ((1u << 20) + d - 1) / d;  // k / d as (k * m) >> 20
    for (int k = lane; k < (BH + 1) * (BW + 1); k += 64) {
      const int r = (int)(((uint32_t)k * m) >> 20), c = k - r * (BW + 1);
      const int py = y0 - 1 + r, px = x0 - 1 + c;
      if (py >= 0 && px >= 0) sh.tile[r * TS + c] = pl[(size_t)py * stride + px];
    }
    __syncthreads();
    const uint32_t k_plane = k0;
    while (k0 < it.count) {
      const vp9r_txb t = k0 < n_st ? sh.txb[k0] : a.txbs[it.first + k0];
      if (t.plane != plane) break;
      const bool staged = t.coef_start >= c0 && t.coef_start - c0 + t.ncoef <= n_c;
      const uint16_t *pos = staged ? sh.pos + (t.coef_start - c0) : a.coef_pos + t.coef_start;
      const int32_t *val = staged ? sh.val + (t.coef_start - c0) : a.coef_val + t.coef_start;
      intra_txb<L, HBD>(a, sh, coef, t, b, x0, y0, pos, val);
      k0++;
    }
    if (k0 > k_plane) {
      const int lw = 31 - __clz(BW);  // BW is a power of two
      for (int k = lane; k < BW * BH; k += 64) {
        const int r = k >> lw, c = k & (BW - 1);
        pl[(size_t)(y0 + r) * stride + x0 + c] = (P)sh.tile[(r + 1) * TS + c + 1];
      }
    }
    __syncthreads();
  }
}

// ---- one launch per level of intra blocks (levels from the host: a block
// is one level above the deepest intra block it reads)

template <bool HBD>
__global__ __launch_bounds__(64) void k_intra_blocks(const IntraArgs *fa, const Range *rg, uint32_t nf) {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type L;
  typedef typename L::T T;
  __shared__ IntraShared sh;
  __shared__ T coef[32 * 33];
  const uint32_t fi = range_frame(rg, nf, blockIdx.x);
  const IntraArgs &a = fa[fi];
  intra_block<L, HBD>(a, sh, coef, a.items[rg[fi].first + blockIdx.x - rg[fi].wg]);
}

void launch_intra(hipStream_t s, const IntraArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd) {
  if (!wgs) return;
  if (hbd)
    hipLaunchKernelGGL(k_intra<true>, dim3(wgs), dim3(64), 0, s, fa, r, n);
  else
    hipLaunchKernelGGL(k_intra<false>, dim3(wgs), dim3(64), 0, s, fa, r, n);
}

void launch_intra_blocks(hipStream_t s, const IntraArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd) {
  if (!wgs) return;
  if (hbd)
    hipLaunchKernelGGL(k_intra_blocks<true>, dim3(wgs), dim3(64), 0, s, fa, r, n);
  else
    hipLaunchKernelGGL(k_intra_blocks<false>, dim3(wgs), dim3(64), 0, s, fa, r, n);
}

// Per-kernel seam: one intra::predict call (intra.rs:19-29) on host-given
// edges, for the parity tests.
__global__ void k_intra_pred(int mode, int log2, const int32_t *above, const int32_t *left, int have_left,
                             int have_above, int bd, uint16_t *dst, uint32_t stride) {
  __shared__ int32_t e[65 + 32];
  const int size = 1 << log2;
  for (int k = threadIdx.x; k <= 2 * size; k += blockDim.x) e[k] = above[k];
  for (int k = threadIdx.x; k < size; k += blockDim.x) e[65 + k] = left[k];
  __syncthreads();
  const int dc = dc_value(log2, e + 1, e + 65, have_left, have_above, bd);
  for (int k = threadIdx.x; k < size * size; k += blockDim.x) {
    int i = k >> log2, j = k & (size - 1);
    dst[i * stride + j] = (uint16_t)intra_sample(mode, i, j, size, e + 1, e + 65, dc, (1 << bd) - 1);
  }
}

void launch_intra_pred(hipStream_t s, uint32_t mode, uint32_t log2, const int32_t *above,
                       const int32_t *left, int have_left, int have_above, uint32_t bit_depth,
                       uint16_t *dst, uint32_t stride) {
  hipLaunchKernelGGL(k_intra_pred, dim3(1), dim3(64), 0, s, (int)mode, (int)log2, above, left, have_left,
                     have_above, (int)bit_depth, dst, stride);
}

}  // namespace vp9g
