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

// Sample (i, j) of the prediction for `mode` (8.5.1.2); A(k), k = -1 ..
// 2 * size - 1, is the above row, L(k), k = 0 .. size - 1, the left column.
template <class EA, class EL>
DEV int intra_sample_at(int mode, int i, int j, int size, EA A, EL L, int dc, int max) {
  switch (mode) {
    case V_PRED:
      return A(j);
    case H_PRED:
      return L(i);
    case D207_PRED: {
      int k = j >> 1, jj = j & 1, ii = i + k;
      if (ii >= size - 1) return L(size - 1);
      if (jj == 0) return r2(L(ii) + L(ii + 1), 1);
      if (ii == size - 2) return r2(L(size - 2) + 3 * L(size - 1), 2);
      return r2(L(ii) + 2 * L(ii + 1) + L(ii + 2), 2);
    }
    case D45_PRED: {
      int k = i + j;
      return k + 2 < 2 * size ? r2(A(k) + 2 * A(k + 1) + A(k + 2), 2) : A(2 * size - 1);
    }
    case D63_PRED: {
      int k = (i >> 1) + j;
      return (i & 1) ? r2(A(k) + 2 * A(k + 1) + A(k + 2), 2) : r2(A(k) + A(k + 1), 1);
    }
    case D117_PRED: {
      int k = min(i >> 1, j);
      i -= 2 * k;
      j -= k;
      if (i == 0) return r2(A(j - 1) + A(j), 1);
      if (i == 1) return j == 0 ? r2(L(0) + 2 * A(-1) + A(0), 2) : r2(A(j - 2) + 2 * A(j - 1) + A(j), 2);
      if (i == 2) return r2(A(-1) + 2 * L(0) + L(1), 2);
      return r2(L(i - 3) + 2 * L(i - 2) + L(i - 1), 2);
    }
    case D135_PRED: {
      int k = min(i, j);
      i -= k;
      j -= k;
      if (i == 0) return j == 0 ? r2(L(0) + 2 * A(-1) + A(0), 2) : r2(A(j - 2) + 2 * A(j - 1) + A(j), 2);
      if (i == 1) return r2(A(-1) + 2 * L(0) + L(1), 2);
      return r2(L(i - 2) + 2 * L(i - 1) + L(i), 2);
    }
    case D153_PRED: {
      int k = min(i, j >> 1);
      i -= k;
      j -= 2 * k;
      if (j == 0) return i == 0 ? r2(L(0) + A(-1), 1) : r2(L(i - 1) + L(i), 1);
      if (j == 1) {
        if (i == 0) return r2(L(0) + 2 * A(-1) + A(0), 2);
        if (i == 1) return r2(A(-1) + 2 * L(0) + L(1), 2);
        return r2(L(i - 2) + 2 * L(i - 1) + L(i), 2);
      }
      return r2(A(j - 3) + 2 * A(j - 2) + A(j - 1), 2);  // i == 0
    }
    case TM_PRED:
      return iclamp(A(j) + L(i) - A(-1), 0, max);
    default:
      return dc;
  }
}

// DC_PRED's value (intra.rs:145-168).
template <class EA, class EL>
DEV int dc_value_at(int log2, EA A, EL L, bool have_left, bool have_above, int bd) {
  int size = 1 << log2, s = 0;
  if (have_left && have_above) {
    for (int k = 0; k < size; k++) s += L(k) + A(k);
    return (s + size) >> (log2 + 1);
  }
  if (have_left || have_above) {
    for (int k = 0; k < size; k++) s += have_left ? L(k) : A(k);
    return (s + (1 << (log2 - 1))) >> log2;
  }
  return 1 << (bd - 1);
}

// The array forms (edges staged in LDS).
DEV int intra_sample(int mode, int i, int j, int size, const int32_t *A, const int32_t *L, int dc, int max) {
  return intra_sample_at(
      mode, i, j, size, [&](int k) { return (int)A[k]; }, [&](int k) { return (int)L[k]; }, dc, max);
}
DEV int dc_value(int log2, const int32_t *A, const int32_t *L, bool have_left, bool have_above, int bd) {
  return dc_value_at(
      log2, [&](int k) { return (int)A[k]; }, [&](int k) { return (int)L[k]; }, have_left, have_above, bd);
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

// ---- intra blocks (k_intra_blocks): one workgroup of 256 lanes per block
//
// Per plane, the block's region (with the row above and the column left of
// it) is staged in LDS.  The residuals of all its transform blocks are
// computed first, in parallel (teams of n0 lanes, as in the inter residual
// kernel), into an LDS residual plane: the reference adds the clamped
// residual to the prediction with a clip to the sample range
// (itx.rs:606-614), so the residual does not depend on the prediction.
// Residuals are kept as int16: beyond +-32767 the clip gives 0 or the
// maximum whatever the prediction, as it does for the full value.  Then
// prediction and the add run as a wavefront over the plane's grid of
// transform blocks: block (r, c) reads the reconstruction left, above and
// above-left of it, and at 4x4 also above-right (recon.rs:61-69), so it is
// at level c + r (c + 2 r at 4x4) and the blocks of a level are
// independent.  Edges are read from the LDS tile as recon.rs:17-103 builds
// them.

constexpr int kIbThreads = 256;
constexpr int kIbStageTxbs = 256, kIbStageCoefs = 2048, kIbCoefTile = 2 * 32 * 33;

struct IntraBlockShared {
  uint16_t tile[TS * TS];  // plane region: rows y0 - 1 .., columns x0 - 1 ..
  int16_t res[64 * 64];    // residual of sample (x, y) at (y - y0) * 64 + x - x0
  vp9r_txb txb[kIbStageTxbs];
  uint16_t pos[kIbStageCoefs];
  int32_t val[kIbStageCoefs];
  uint16_t tmap[16 * 16];  // grid cell r * 16 + c -> transform block (offset in the plane), 0xffff: none
  int32_t dcv[16];         // DC_PRED values of the current level's blocks
};

struct ResDst {
  int16_t *base;
  DEV void add(int i, int j, int r, int) { base[i * 64 + j] = (int16_t)min(max(r, -32768), 32767); }
};

// The residuals of a plane's np transform blocks (size 4 << (n - 2)),
// `teams` blocks a pass.
template <class L, int n, class TX>
DEV void ib_residuals(const IntraArgs &a, IntraBlockShared &sh, typename L::T *ctile, TX txb_at, uint32_t k0,
                      uint32_t np, int plane, int x0, int y0, int seg, uint32_t c0, uint32_t n_c, bool hbd) {
  typedef typename L::T T;
  constexpr int n0 = 1 << n, S = n0 + 1;
  constexpr int teams = n == 2 ? 64 : n == 3 ? 16 : n == 4 ? 4 : 2;
  const int tid = threadIdx.x, team = tid / n0, ln = tid % n0;
  const bool in_team = team < teams;
  T *tile = ctile + (in_team ? team : 0) * (n0 * S);
  const int pt = plane > 0;
  const int32_t dcq = a.q[seg][pt][0], acq = a.q[seg][pt][1];
  const bool lossless = a.g.lossless;
  for (uint32_t base = 0; base < np; base += teams) {
    const uint32_t j = base + team;
    vp9r_txb t = {};
    bool active = false;
    if (in_team && j < np) {
      t = txb_at(k0 + j);
      active = t.eob > 0;
    }
    if (in_team)
      for (int k = ln; k < n0 * S; k += n0) tile[k] = 0;
    __syncthreads();
    if (active) {
      const bool staged = t.coef_start >= c0 && t.coef_start - c0 + t.ncoef <= n_c;
      const uint16_t *pos = staged ? sh.pos + (t.coef_start - c0) : a.coef_pos + t.coef_start;
      const int32_t *val = staged ? sh.val + (t.coef_start - c0) : a.coef_val + t.coef_start;
      for (int k = ln; k < t.ncoef; k += n0) {
        const int p = pos[k] & (n0 * n0 - 1);
        tile[(p >> n) * S + (p & (n0 - 1))] = (T)dequant(val[k], p, dcq, acq, n == 5);
      }
    }
    __syncthreads();
    const bool dc_only = !hbd && active && t.eob == 1 && t.tx_type == DCT_DCT && !lossless;
    const int32_t dc_v = dc_only ? dc_only_value((int32_t)tile[0], n) : 0;
    ResDst dst{sh.res + (active ? (t.y - y0) * 64 + (t.x - x0) : 0)};
    itx_2d<L, n>(tile, active ? ln : n0, t.tx_type, lossless, dc_only, dc_v, dst, (1 << a.g.bit_depth) - 1);
    __syncthreads();
  }
}

template <bool HBD>
__global__ __launch_bounds__(kIbThreads) void k_intra_blocks(const IntraArgs *fa, const Range *rg, uint32_t nf) {
  typedef typename std::conditional<HBD, Lane64, Lane32>::type L;
  typedef typename L::T T;
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ IntraBlockShared sh;
  __shared__ T ctile[kIbCoefTile];
  const int tid = threadIdx.x;
  const uint32_t fi = range_frame(rg, nf, blockIdx.x);
  const IntraArgs &a = fa[fi];
  const FrameGeom &g = a.g;
  const IntraBlockItem it = a.items[rg[fi].first + blockIdx.x - rg[fi].wg];
  // stage the transform block records and the coefficients they cover
  const uint32_t n_st = min(it.count, (uint32_t)kIbStageTxbs), n_c = min(it.coef_count, (uint32_t)kIbStageCoefs);
  const uint32_t c0 = it.coef_first;
  for (uint32_t k = tid; k < n_st; k += kIbThreads) sh.txb[k] = a.txbs[it.first + k];
  for (uint32_t k = tid; k < n_c; k += kIbThreads) {
    sh.pos[k] = a.coef_pos[c0 + k];
    sh.val[k] = a.coef_val[c0 + k];
  }
  auto txb_at = [&](uint32_t k) -> vp9r_txb { return k < n_st ? sh.txb[k] : a.txbs[it.first + k]; };
  const int mi_size = it.mi_size, seg = it.flags & 7;
  const bool avail_u = (it.flags & 8) != 0, avail_l = (it.flags & 16) != 0;
  const int bsize = mi_size < BLOCK_8X8 ? BLOCK_8X8 : mi_size;
  const int bw8 = num8x8_wide(bsize), bh8 = num8x8_high(bsize);
  const int bd = g.bit_depth, maxv = (1 << bd) - 1, base = 1 << (bd - 1);
  uint32_t k0 = 0;
  for (int plane = 0; plane < 3; plane++) {
    const uint32_t np = plane == 0 ? it.n_y : plane == 1 ? it.n_u : it.count - it.n_y - it.n_u;
    if (np == 0) continue;
    const int sx = plane ? g.ss_x : 0, sy = plane ? g.ss_y : 0;
    const int x0 = (it.mi_col * 8) >> sx, y0 = (it.mi_row * 8) >> sy;
    const int BW = (bw8 * 8) >> sx, BH = (bh8 * 8) >> sy;
    const int lw = 31 - __clz(BW);  // BW is a power of two
    const uint32_t stride = a.cur.stride[plane];
    P *pl = (P *)a.cur.p[plane];
    // the region: the block, the row above it (from above-left) and the
    // column left of it; samples of the block outside the 8x8-aligned area
    // are not reconstructed and keep the frame's
    const uint32_t dw = BW + 1, mw = ((1u << 20) + dw - 1) / dw;  // k / (BW + 1) as (k * mw) >> 20
    for (int k = tid; k < (BH + 1) * (BW + 1); k += kIbThreads) {
      const int r = (int)(((uint32_t)k * mw) >> 20), c = k - r * (BW + 1);
      const int py = y0 - 1 + r, px = x0 - 1 + c;
      if (py >= 0 && px >= 0) sh.tile[r * TS + c] = pl[(size_t)py * stride + px];
    }
    for (int k = tid; k < BW * BH; k += kIbThreads) sh.res[(k >> lw) * 64 + (k & (BW - 1))] = 0;
    for (int k = tid; k < 256; k += kIbThreads) sh.tmap[k] = 0xffff;
    __syncthreads();
    const int tx = txb_at(k0).tx_size, log2 = tx + 2, size = 1 << log2;
    for (uint32_t j = tid; j < np; j += kIbThreads) {
      const vp9r_txb t = txb_at(k0 + j);
      sh.tmap[((t.y - y0) >> log2) * 16 + ((t.x - x0) >> log2)] = (uint16_t)j;
    }
    switch (tx) {  // ends with a barrier
      case 0: ib_residuals<L, 2>(a, sh, ctile, txb_at, k0, np, plane, x0, y0, seg, c0, n_c, HBD); break;
      case 1: ib_residuals<L, 3>(a, sh, ctile, txb_at, k0, np, plane, x0, y0, seg, c0, n_c, HBD); break;
      case 2: ib_residuals<L, 4>(a, sh, ctile, txb_at, k0, np, plane, x0, y0, seg, c0, n_c, HBD); break;
      default: ib_residuals<L, 5>(a, sh, ctile, txb_at, k0, np, plane, x0, y0, seg, c0, n_c, HBD); break;
    }
    // the wavefront over the grid of transform blocks
    const int W = max(BW >> log2, 1), H = max(BH >> log2, 1);
    const int kk = tx == TX_4X4 ? 2 : 1, n_lv = (W - 1) + kk * (H - 1) + 1;
    const int max_x = (int)((g.mi_cols * 8) >> sx) - 1, max_y = (int)((g.mi_rows * 8) >> sy) - 1;
    const int n4w = num4x4_wide(bsize) >> sx, step4 = 1 << tx;
    const int mode_u = plane > 0 ? it.uv_mode : (mi_size >= BLOCK_8X8 ? it.y_mode : -1);
    bool any_dc = mode_u == DC_PRED;
    if (mode_u < 0)
      for (int q = 0; q < 4; q++) any_dc = any_dc || ((it.sub_modes >> (8 * q)) & 0xff) == DC_PRED;
    auto mode_of = [&](const vp9r_txb &t) -> int {
      if (mode_u >= 0) return mode_u;
      const int xi = (t.x - x0) >> 2, yi = (t.y - y0) >> 2;
      const int block_idx = (yi / step4) * ((n4w + step4 - 1) / step4) + xi / step4;
      return (int)((it.sub_modes >> (8 * (block_idx & 3))) & 0xff);
    };
#define TILE(py, px) ((int)sh.tile[((py) - y0 + 1) * TS + ((px) - x0 + 1)])
    // the edges of transform block t (recon.rs:17-103)
    auto edges = [&](const vp9r_txb &t, bool &have_left, bool &have_above, bool &not_on_right) {
      const int xi = (t.x - x0) >> 2, yi = (t.y - y0) >> 2;
      have_left = avail_l || xi > 0;
      have_above = avail_u || yi > 0;
      not_on_right = xi + step4 < n4w;
    };
    auto above_of = [&](int x, int y, bool have_left, bool have_above, bool not_on_right) {
      return [=](int k) -> int {
        if (!have_above) return base - 1;
        if (k < 0) return have_left ? TILE(y - 1, min(max_x, x - 1)) : base + 1;
        const int ii = (k >= size && !(not_on_right && tx == TX_4X4)) ? size - 1 : k;
        return TILE(y - 1, min(max_x, x + ii));
      };
    };
    auto left_of = [&](int x, int y, bool have_left) {
      return [=](int k) -> int { return have_left ? TILE(min(max_y, y + k), x - 1) : base + 1; };
    };
    for (int lv = 0; lv < n_lv; lv++) {
      const int r_lo = lv > W - 1 ? (lv - (W - 1) + kk - 1) / kk : 0;
      const int r_hi = min(H - 1, lv / kk);
      const int n_l = r_hi - r_lo + 1;
      if (n_l <= 0) continue;
      if (any_dc) {
        if (tid < n_l) {
          const int r = r_lo + tid, c = lv - kk * r;
          const uint16_t ix = sh.tmap[r * 16 + c];
          if (ix != 0xffff) {
            const vp9r_txb t = txb_at(k0 + ix);
            if (mode_of(t) == DC_PRED) {
              bool hl, ha, nr;
              edges(t, hl, ha, nr);
              sh.dcv[tid] = dc_value_at(log2, above_of(t.x, t.y, hl, ha, nr), left_of(t.x, t.y, hl), hl, ha, bd);
            }
          }
        }
        __syncthreads();
      }
      for (int item = tid; item < (n_l << (2 * log2)); item += kIbThreads) {
        const int jt = item >> (2 * log2), s = item & ((1 << (2 * log2)) - 1);
        const int r = r_lo + jt, c = lv - kk * r;
        const uint16_t ix = sh.tmap[r * 16 + c];
        if (ix == 0xffff) continue;
        const vp9r_txb t = txb_at(k0 + ix);
        bool hl, ha, nr;
        edges(t, hl, ha, nr);
        const int mode = mode_of(t);
        const int i = s >> log2, j = s & (size - 1);
        const int pv = intra_sample_at(mode, i, j, size, above_of(t.x, t.y, hl, ha, nr), left_of(t.x, t.y, hl),
                                    mode == DC_PRED ? sh.dcv[jt] : 0, maxv);
        const int ry = t.y - y0 + i, rx = t.x - x0 + j;
        sh.tile[(ry + 1) * TS + rx + 1] = (uint16_t)iclamp(pv + sh.res[ry * 64 + rx], 0, maxv);
      }
      __syncthreads();
    }
#undef TILE
    for (int k = tid; k < BW * BH; k += kIbThreads) {
      const int r = k >> lw, c = k & (BW - 1);
      pl[(size_t)(y0 + r) * stride + x0 + c] = (P)sh.tile[(r + 1) * TS + c + 1];
    }
    __syncthreads();  // the next plane's region reuses the tile
    k0 += np;
  }
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
    hipLaunchKernelGGL(k_intra_blocks<true>, dim3(wgs), dim3(kIbThreads), 0, s, fa, r, n);
  else
    hipLaunchKernelGGL(k_intra_blocks<false>, dim3(wgs), dim3(kIbThreads), 0, s, fa, r, n);
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
