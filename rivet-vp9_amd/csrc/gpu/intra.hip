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
#include "lf_dev.h"

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

// ---- intra blocks (k_intra_blocks): one workgroup of ib_threads lanes
// per block
//
// The block's three planes are independent (each reads only its own
// plane), so they are processed together.  Each plane's region (the block
// with the row above and the column left of it) is staged in LDS, with the
// residuals of its transform blocks: the reference adds the clamped
// residual to the prediction with a clip to the sample range
// (itx.rs:606-614), so the residual does not depend on the prediction and
// the inter residual launches compute it beforehand (k_itx_inter, jobs
// flagged kItxIntra, int16 planes: beyond +-32767 the clip gives 0 or the
// maximum whatever the prediction, as it does for the full value).  Then
// prediction and the add run as a wavefront over each plane's grid of
// transform blocks: block (r, c) reads the reconstruction left, above and
// above-left of it, and at 4x4 also above-right (recon.rs:61-69), so it is
// at level c + r (c + 2 r at 4x4) and the blocks of a level, in all three
// planes, are independent.  Edges are read from the LDS tiles as
// recon.rs:17-103 builds them.  The LDS layout depends on the frame's
// subsampling (ib_layout); the launch passes the largest of its frames'.

// Geometry of one plane of the block (in LDS).
struct IbGeo {
  int x0, y0, BW, BH;         // plane block origin and size
  int log2, kk, W, H, nlv;    // transform size, level step, grid, levels
  int tpitch, rpitch;         // LDS tile and residual pitches
  int max_x, max_y, n4w;      // last sample of the 8x8-aligned area; width in 4x4s
  int mode_u;                 // the plane's mode (-1: luma of a sub-8x8 block, per 4x4)
  uint32_t tile_off, res_off; // LDS byte offsets
};
static_assert(sizeof(IbGeo) <= kIbGeoBytes, "IbGeo");

// The region of one plane of an intra block staged in LDS: the block (BW x
// BH at (x0, y0)) and the row above it, in vectors of V = min(8, BW)
// samples (a block is aligned to its size), then the column left of it
// (x0 - 1), sample by sample.  Item k < nvec is vector (k >> lq) of row
// (k >> lq) -- rows from y0 - 1 --, the rest the left column.
struct IbRegion {
  int x0, y0, BW, BH, V, lq;  // lq: log2 of the vectors per row
  int nvec, n;                // vector items; all items (0 when the plane is absent)
};

DEV IbRegion ib_region(const IntraBlockItem &it, const FrameGeom &g, int p, int bw8, int bh8, bool present) {
  IbRegion r;
  const int sx = p ? g.ss_x : 0, sy = p ? g.ss_y : 0;
  r.x0 = (it.mi_col * 8) >> sx;
  r.y0 = (it.mi_row * 8) >> sy;
  r.BW = (bw8 * 8) >> sx;
  r.BH = (bh8 * 8) >> sy;
  r.V = r.BW < 8 ? r.BW : 8;
  r.lq = 31 - __clz(r.BW / r.V);
  r.nvec = (r.BH + 1) << r.lq;
  r.n = present ? r.nvec + r.BH + 1 : 0;
  return r;
}

// The intra block workgroup's LDS (laid out by ib_layout).
extern __shared__ __align__(16) uint8_t ib_lds[];

// One transform block of the intra block, set up before the wavefront: its
// place in the LDS tile and residual plane (16-bit element indices into the
// workgroup's LDS), its mode and edge availability, and the clamps of its
// edge reads (recon.rs:17-103): the above row is read up to max_x - x
// samples right of x (and, past the block's width, repeats its last sample
// unless it is a 4x4 not on the block's right: a4), the left column up to
// max_y - y rows down.  Grid cell (plane, r, c) of the block holds it.
struct IbTx {
  int tile, tp, res, rp;  // tile[tile + i tp + j] = sample (i, j); res likewise
  int bits;               // mode | log2 << 4 | above << 8 | left << 9 | a4 << 10 | valid << 11
  int lim_a, lim_l, _pad;
  DEV int mode() const { return bits & 15; }
  DEV bool valid() const { return (bits >> 11) & 1; }
  DEV int T(int idx) const { return ((const uint16_t *)ib_lds)[idx]; }
  DEV int A(int k, int size, int base) const {
    if (!((bits >> 8) & 1)) return base - 1;
    if (k < 0) return ((bits >> 9) & 1) ? T(tile - tp - 1) : base + 1;
    return T(tile - tp + min((k >= size && !((bits >> 10) & 1)) ? size - 1 : k, lim_a));
  }
  DEV int L(int k, int base) const { return ((bits >> 9) & 1) ? T(tile - 1 + min(k, lim_l) * tp) : base + 1; }
  // DC_PRED's value (intra.rs:145-168) for size 4 << (LOG2 - 2)
  template <int LOG2>
  DEV int dc(int bd) const {
    constexpr int size = 1 << LOG2;
    const bool have_above = (bits >> 8) & 1, have_left = (bits >> 9) & 1;
    int sa = 0, sl = 0;
    if (have_above) {
#pragma unroll
      for (int k = 0; k < size; k++) sa += T(tile - tp + min(k, lim_a));
    }
    if (have_left) {
#pragma unroll
      for (int k = 0; k < size; k++) sl += T(tile - 1 + min(k, lim_l) * tp);
    }
    if (have_left && have_above) return (sa + sl + size) >> (LOG2 + 1);
    if (have_left || have_above) return (sa + sl + (size >> 1)) >> LOG2;
    return 1 << (bd - 1);
  }
};

// Stages the regions of the three planes in LDS (layout ly): each region
// rg (the block or superblock, the row above it from above-left and the
// column left of it; samples outside the 8x8-aligned area are not
// reconstructed and keep the frame's), and the residuals of its area.
template <bool HBD, int NT>
DEV void ib_stage(const FrameGeom &g, const IbLayout &ly, const IbRegion &rg0, const IbRegion &rg1,
                  const IbRegion &rg2, uint8_t *const (&pl_)[3], const int16_t *const (&rs_)[3],
                  const uint32_t (&st_)[3], int tid) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  auto cur_plane = [&](int p) -> uint8_t * { return p == 0 ? pl_[0] : p == 1 ? pl_[1] : pl_[2]; };
  auto res_plane = [&](int p) -> const int16_t * { return p == 0 ? rs_[0] : p == 1 ? rs_[1] : rs_[2]; };
  auto stride_of = [&](int p) -> size_t { return p == 0 ? st_[0] : p == 1 ? st_[1] : st_[2]; };
  // All loads of the
  // three planes are in flight before the first LDS write: at most
  // kIbRegionSlots items per lane (4:4:4 64x64: 3 x (65 x 8 + 65) items).
  constexpr int kIbRegionSlots = (3 * (65 * 8 + 65) + NT - 1) / NT;
  constexpr int kIbResSlots = (3 * 512 + NT - 1) / NT;
  const int nr0 = rg0.n ? rg0.BH << rg0.lq : 0, nr1 = rg1.n ? rg1.BH << rg1.lq : 0, nr2 = rg2.n ? rg2.BH << rg2.lq : 0;
  {
    uint4 v[kIbRegionSlots], rv[kIbResSlots];
#pragma unroll
    for (int i = 0; i < kIbRegionSlots; i++) {
      const int k = tid + NT * i;
      const int p = k < rg0.n ? 0 : k < rg0.n + rg1.n ? 1 : 2;
      const IbRegion R = p == 0 ? rg0 : p == 1 ? rg1 : rg2;
      const int local = k - (p == 0 ? 0 : p == 1 ? rg0.n : rg0.n + rg1.n);
      if (local >= R.n) continue;
      const P *pl = (const P *)cur_plane(p);
      const size_t stride = stride_of(p);
      if (local < R.nvec) {
        const int r = local >> R.lq, q = local & ((1 << R.lq) - 1), py = R.y0 - 1 + r;
        if (py < 0) continue;
        // one dword load per 4 bytes, each straight into its own component
        // (loads of different widths into one vector would be merged with
        // a wait for each)
        const uint32_t *src = (const uint32_t *)(pl + (size_t)py * stride + R.x0 + q * R.V);
        const int nb = R.V * (int)sizeof(P);
        v[i].x = src[0];
        if (nb >= 8) v[i].y = src[1];
        if (nb == 16) {
          v[i].z = src[2];
          v[i].w = src[3];
        }
      } else {
        const int r = local - R.nvec, py = R.y0 - 1 + r;
        if (py < 0 || R.x0 == 0) continue;
        v[i] = make_uint4(pl[(size_t)py * stride + R.x0 - 1], 0, 0, 0);
      }
    }
    // the residuals of the block (k_itx_inter wrote every transform block's;
    // samples outside the transform blocks are never read), in vectors of
    // V samples: at most kIbResSlots per lane (4:4:4 64x64: 3 x 512)
#pragma unroll
    for (int i = 0; i < kIbResSlots; i++) {
      const int k = tid + NT * i;
      const int p = k < nr0 ? 0 : k < nr0 + nr1 ? 1 : 2;
      const int local = k - (p == 0 ? 0 : p == 1 ? nr0 : nr0 + nr1);
      if (local >= (p == 0 ? nr0 : p == 1 ? nr1 : nr2)) continue;
      const IbRegion R = p == 0 ? rg0 : p == 1 ? rg1 : rg2;
      const int r = local >> R.lq, q = local & ((1 << R.lq) - 1);
      const uint32_t *src = (const uint32_t *)(res_plane(p) + (size_t)(R.y0 + r) * stride_of(p) + R.x0 + q * R.V);
      rv[i].x = src[0];
      rv[i].y = src[1];
      if (R.V == 8) {
        rv[i].z = src[2];
        rv[i].w = src[3];
      }
    }
#pragma unroll
    for (int i = 0; i < kIbResSlots; i++) {
      const int k = tid + NT * i;
      const int p = k < nr0 ? 0 : k < nr0 + nr1 ? 1 : 2;
      const int local = k - (p == 0 ? 0 : p == 1 ? nr0 : nr0 + nr1);
      if (local >= (p == 0 ? nr0 : p == 1 ? nr1 : nr2)) continue;
      const IbRegion R = p == 0 ? rg0 : p == 1 ? rg1 : rg2;
      const int r = local >> R.lq, q = local & ((1 << R.lq) - 1), rp = 64 >> (p ? g.ss_x : 0);
      uint8_t *dst = ib_lds + (p == 0 ? ly.res[0] : p == 1 ? ly.res[1] : ly.res[2]) + (r * rp + q * R.V) * 2;
      if (R.V == 8)
        *(uint4 *)dst = rv[i];
      else
        *(uint2 *)dst = make_uint2(rv[i].x, rv[i].y);
    }
#pragma unroll
    for (int i = 0; i < kIbRegionSlots; i++) {
      const int k = tid + NT * i;
      const int p = k < rg0.n ? 0 : k < rg0.n + rg1.n ? 1 : 2;
      const IbRegion R = p == 0 ? rg0 : p == 1 ? rg1 : rg2;
      const int local = k - (p == 0 ? 0 : p == 1 ? rg0.n : rg0.n + rg1.n);
      if (local >= R.n) continue;
      const int tp = (64 >> (p ? g.ss_x : 0)) + 1;
      uint16_t *tile = (uint16_t *)(ib_lds + (p == 0 ? ly.tile[0] : p == 1 ? ly.tile[1] : ly.tile[2]));
      if (local < R.nvec) {
        const int r = local >> R.lq, q = local & ((1 << R.lq) - 1);
        if (R.y0 - 1 + r < 0) continue;
        uint16_t *dst = tile + r * tp + 1 + q * R.V;
        const uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
          if (j >= R.V) continue;
          dst[j] = sizeof(P) == 1 ? (uint16_t)((w[j >> 2] >> (8 * (j & 3))) & 0xff)
                                  : (uint16_t)((w[j >> 1] >> (16 * (j & 1))) & 0xffff);
        }
      } else {
        const int r = local - R.nvec;
        if (R.y0 - 1 + r < 0 || R.x0 == 0) continue;
        tile[r * tp] = (uint16_t)v[i].x;
      }
    }
  }
}

// Writes the regions' areas back from LDS to the frame.
template <bool HBD, int NT>
DEV void ib_store(const FrameGeom &g, const IbLayout &ly, const IbRegion &rg0, const IbRegion &rg1,
                  const IbRegion &rg2, uint8_t *const (&pl_)[3], const uint32_t (&st_)[3], int tid) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  // the blocks back to the frame, in vectors of V samples
#pragma unroll
  for (int p = 0; p < 3; p++) {
    const IbRegion R = p == 0 ? rg0 : p == 1 ? rg1 : rg2;
    if (!R.n) continue;
    const int tp = (64 >> (p ? g.ss_x : 0)) + 1;
    const uint16_t *tile = (const uint16_t *)(ib_lds + ly.tile[p]);
    const size_t stride = st_[p];
    P *pl = (P *)pl_[p];
    for (int k = tid; k < (R.BH << R.lq); k += NT) {
      const int r = k >> R.lq, q = k & ((1 << R.lq) - 1);
      const uint16_t *src = tile + (r + 1) * tp + 1 + q * R.V;
      P *dst = pl + (size_t)(R.y0 + r) * stride + R.x0 + q * R.V;
      uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (j >= R.V) continue;
        if (sizeof(P) == 1)
          w[j >> 2] |= (uint32_t)src[j] << (8 * (j & 3));
        else
          w[j >> 1] |= (uint32_t)src[j] << (16 * (j & 1));
      }
      const int nb = R.V * (int)sizeof(P);
      if (nb == 16)
        *(uint4 *)dst = make_uint4(w[0], w[1], w[2], w[3]);
      else if (nb == 8)
        *(uint2 *)dst = make_uint2(w[0], w[1]);
      else
        *(uint32_t *)dst = w[0];
    }
  }
}

// A superblock item: every intra block of one superblock in one workgroup.
// The superblock's region in all three planes (with the row above and the
// column left of it) and its residuals are staged in LDS once; then the
// item's levels run in order, each predicting and reconstructing its
// transform blocks (whose edges come only from samples of earlier levels or
// of earlier items) with a barrier between levels; the region goes back at
// the end.  The host levels the transform blocks (intra_sb_items) and the
// items by the superblocks left, above and above-left of them.
template <bool HBD, int NT>
DEV void intra_sb(const IntraArgs &a, const IntraBlockItem &it, uint8_t *const (&pl_)[3],
                  const int16_t *const (&rs_)[3], const uint32_t (&st_)[3]) {
  const int tid = threadIdx.x;
  const FrameGeom &g = a.g;
  const IbLayout ly = ib_layout(g.ss_x, g.ss_y);
  const uint32_t M = ib_max_txbs(g.ss_x, g.ss_y);
  IbTx *txs = (IbTx *)(ib_lds + ly.txs);
  uint32_t *units = (uint32_t *)(ib_lds + ly.txb);
  int32_t *dcv = (int32_t *)(ib_lds + ly.txb + 4 * M);
  uint2 *levels = (uint2 *)(ib_lds + ly.txb + 8 * M);
  const int bd = g.bit_depth, maxv = (1 << bd) - 1, base = 1 << (bd - 1);
  const IbRegion rg0 = ib_region(it, g, 0, 8, 8, true), rg1 = ib_region(it, g, 1, 8, 8, true),
                 rg2 = ib_region(it, g, 2, 8, 8, true);
  // descriptors, units and levels first: their loads overlap the region's
  for (uint32_t k = tid; k < it.count; k += NT) {
    const IntraSbTx d = a.sb_tx[it.first + k];
    const int p = d.plane, sx = p ? g.ss_x : 0;
    const uint32_t tile_off = p == 0 ? ly.tile[0] : p == 1 ? ly.tile[1] : ly.tile[2];
    const uint32_t res_off = p == 0 ? ly.res[0] : p == 1 ? ly.res[1] : ly.res[2];
    IbTx t;
    t.tp = (64 >> sx) + 1;
    t.rp = 64 >> sx;
    t.tile = (int)(tile_off >> 1) + (d.y + 1) * t.tp + d.x + 1;
    t.res = (int)(res_off >> 1) + d.y * t.rp + d.x;
    t.bits = d.mode | d.log2 << 4 | (d.flags & 1) << 8 | ((d.flags >> 1) & 1) << 9 | ((d.flags >> 2) & 1) << 10 | 1 << 11;
    t.lim_a = d.lim_a;
    t.lim_l = d.lim_l;
    t._pad = 0;
    txs[k] = t;
  }
  for (uint32_t k = tid; k < it.coef_count; k += NT) units[k] = a.sb_units[it.coef_first + k];
  for (uint32_t k = tid; k < it.n_u; k += NT) levels[k] = a.sb_levels[it.sub_modes + k];
  ib_stage<HBD, NT>(g, ly, rg0, rg1, rg2, pl_, rs_, st_, tid);
  __syncthreads();
  const int16_t *res = (const int16_t *)ib_lds;
  uint16_t *tile = (uint16_t *)ib_lds;
  // Each wave takes chunks of the level in turn: up to four 4x4 units of
  // one transform block, a 4x4 per 16 lanes.  A wave's lanes then predict
  // one mode, so the mode's branch and its edge reads run once per wave (a
  // wave spanning several blocks ran every block's branch in turn, their
  // reads one after another).
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  for (uint32_t lv = 0; lv < it.n_u; lv++) {
    const uint2 L = levels[lv];
    const int d0 = (int)(L.x & 0xffff), nd = (int)((L.x >> 16) & 0x7fff);
    const int c0 = (int)(L.y & 0xffff), nc = (int)(L.y >> 16);
    const bool dc_pass = (L.x >> 31) != 0;
    if (dc_pass) {
      // DC values of the level's DC_PRED blocks (it has one of 16x16 or
      // larger), every edge read in flight at once
      if (tid < nd) {
        const IbTx d = txs[d0 + tid];
        if (d.mode() == DC_PRED) {
          int v;
          switch ((d.bits >> 4) & 15) {
            case 2: v = d.dc<2>(bd); break;
            case 3: v = d.dc<3>(bd); break;
            case 4: v = d.dc<4>(bd); break;
            default: v = d.dc<5>(bd); break;
          }
          dcv[d0 + tid] = v;
        }
      }
      __syncthreads();
    }
    for (int c = wave; c < nc; c += NT / 64) {
      const uint32_t ch = units[c0 + c];
      if (lane >= 16 * (int)(ch >> 18)) continue;
      const int di = (int)(ch & 0x3ff), sub = (int)((ch >> 10) & 0xff) + (lane >> 4), s16 = lane & 15;
      const IbTx d = txs[di];
      const int l2 = (d.bits >> 4) & 15, size = 1 << l2, mode = d.mode();
      const int i = ((sub >> (l2 - 2)) << 2) + (s16 >> 2), j = ((sub & ((size >> 2) - 1)) << 2) + (s16 & 3);
      // DC from the DC pass, else (4x4 and 8x8 only) computed here
      int dc = 0;
      if (mode == DC_PRED) dc = dc_pass ? dcv[di] : (l2 == 2 ? d.dc<2>(bd) : d.dc<3>(bd));
      const int pv = intra_sample_at(
          mode, i, j, size, [&](int k) { return d.A(k, size, base); }, [&](int k) { return d.L(k, base); }, dc,
          maxv);
      tile[d.tile + i * d.tp + j] = (uint16_t)iclamp(pv + res[d.res + i * d.rp + j], 0, maxv);
    }
    __syncthreads();
  }
  ib_store<HBD, NT>(g, ly, rg0, rg1, rg2, pl_, st_, tid);
}

// Intra block wg of the launch (NT lanes: ib_threads(ONE)).
template <bool HBD, bool ONE, int NT = ib_threads(ONE)>
DEV void intra_block(const IntraArgs *fa, const Range *rg, uint32_t nf, const IntraBlockItem *items0, uint8_t *cp0,
                     uint8_t *cp1, uint8_t *cp2, const int16_t *rp0, const int16_t *rp1, const int16_t *rp2,
                     const vp9r_txb *txbs1, uint32_t wg) {
  const int tid = threadIdx.x;
  const uint32_t fi = range_frame(rg, nf, wg);
  const IntraArgs &a = fa[fi];
  const FrameGeom &g = a.g;
  // a launch of one frame (ONE) passes its items (the item loads with the
  // frame's arguments instead of after them) and its planes, residual
  // planes and transform block records (kernel arguments: known to be
  // global memory)
  const IntraBlockItem it = ONE ? items0[wg] : a.items[rg[fi].first + wg - rg[fi].wg];
  // the planes, residual planes and strides as scalars, read once: the
  // staging loops pick a plane per lane, and a per-lane read of the frame's
  // arguments would hold up every load behind it
  uint8_t *const pl_[3] = {ONE ? cp0 : (uint8_t *)a.cur.p[0], ONE ? cp1 : (uint8_t *)a.cur.p[1],
                           ONE ? cp2 : (uint8_t *)a.cur.p[2]};
  const int16_t *const rs_[3] = {ONE ? rp0 : a.res[0], ONE ? rp1 : a.res[1], ONE ? rp2 : a.res[2]};
  const uint32_t st_[3] = {a.cur.stride[0], a.cur.stride[1], a.cur.stride[2]};
  if (it.flags & kIbSuperblock) {
    intra_sb<HBD, NT>(a, it, pl_, rs_, st_);
    return;
  }
  const vp9r_txb *txbs = ONE ? txbs1 : a.txbs;
  const IbLayout ly = ib_layout(g.ss_x, g.ss_y);
  vp9r_txb *stx = (vp9r_txb *)(ib_lds + ly.txb);
  IbTx *txs = (IbTx *)(ib_lds + ly.txs);
  int32_t *dcv = (int32_t *)(ib_lds + ly.dcv);
  // the transform block records, all in LDS (at most ib_max_txbs; the host
  // checks each plane's grid, which bounds the count)
  for (uint32_t k = tid; k < it.count; k += NT) stx[k] = txbs[it.first + k];
  auto txb_at = [&](uint32_t k) -> vp9r_txb { return stx[k]; };
  const int mi_size = it.mi_size;
  const bool avail_u = (it.flags & 8) != 0, avail_l = (it.flags & 16) != 0;
  const int bsize = mi_size < BLOCK_8X8 ? BLOCK_8X8 : mi_size;
  const int bw8 = num8x8_wide(bsize), bh8 = num8x8_high(bsize);
  const int bd = g.bit_depth, maxv = (1 << bd) - 1, base = 1 << (bd - 1);
  const uint32_t np0 = it.n_y, np1 = it.n_u, np2 = it.count - it.n_y - it.n_u;
  const IbRegion rg0 = ib_region(it, g, 0, bw8, bh8, np0 > 0), rg1 = ib_region(it, g, 1, bw8, bh8, np1 > 0),
                 rg2 = ib_region(it, g, 2, bw8, bh8, np2 > 0);
  ib_stage<HBD, NT>(g, ly, rg0, rg1, rg2, pl_, rs_, st_, tid);
  for (int k = tid; k < 3 * 256; k += NT) txs[k].bits = 0;
  __syncthreads();
  // the planes' grid geometry, by every lane (a plane's transform size from
  // its first record, a broadcast LDS read), in registers: no barrier for a
  // table of it (selected per plane, not indexed: an indexed private array
  // is memory)
  auto sel3 = [](int p, int a, int b, int c) { return p == 0 ? a : p == 1 ? b : c; };
  const int tx0 = np0 ? txb_at(0).tx_size : 0, tx1 = np1 ? txb_at(np0).tx_size : 0,
            tx2 = np2 ? txb_at(np0 + np1).tx_size : 0;
  const int lg0 = tx0 + 2, lg1 = tx1 + 2, lg2 = tx2 + 2;
  const int kk0 = tx0 == TX_4X4 ? 2 : 1, kk1 = tx1 == TX_4X4 ? 2 : 1, kk2 = tx2 == TX_4X4 ? 2 : 1;
  const int bw0 = bw8 * 8, bh0 = bh8 * 8, bwc = (bw8 * 8) >> g.ss_x, bhc = (bh8 * 8) >> g.ss_y;
  const int gw0 = max(bw0 >> lg0, 1), gw1 = max(bwc >> lg1, 1), gw2 = max(bwc >> lg2, 1);
  const int gh0 = max(bh0 >> lg0, 1), gh1 = max(bhc >> lg1, 1), gh2 = max(bhc >> lg2, 1);
  const int gn0 = np0 ? (gw0 - 1) + kk0 * (gh0 - 1) + 1 : 0, gn1 = np1 ? (gw1 - 1) + kk1 * (gh1 - 1) + 1 : 0,
            gn2 = np2 ? (gw2 - 1) + kk2 * (gh2 - 1) + 1 : 0;
  const int mode_y = mi_size >= BLOCK_8X8 ? it.y_mode : -1;
  // every transform block's descriptor in its grid cell
  for (uint32_t j = tid; j < it.count; j += NT) {
    const vp9r_txb t = txb_at(j);
    const int p = t.plane, sx = p ? g.ss_x : 0, sy = p ? g.ss_y : 0;
    const int x0 = (it.mi_col * 8) >> sx, y0 = (it.mi_row * 8) >> sy, log2 = sel3(p, lg0, lg1, lg2);
    const int tpitch = (64 >> sx) + 1, rpitch = 64 >> sx, n4w = num4x4_wide(bsize) >> sx;
    const uint32_t tile_off = p == 0 ? ly.tile[0] : p == 1 ? ly.tile[1] : ly.tile[2];
    const uint32_t res_off = p == 0 ? ly.res[0] : p == 1 ? ly.res[1] : ly.res[2];
    const int tx = t.x - x0, ty = t.y - y0, xi = tx >> 2, yi = ty >> 2, step4 = 1 << (log2 - 2);
    int mode = p > 0 ? it.uv_mode : mode_y;
    if (mode < 0) {
      const int block_idx = (yi / step4) * ((n4w + step4 - 1) / step4) + xi / step4;
      mode = (int)((it.sub_modes >> (8 * (block_idx & 3))) & 0xff);
    }
    IbTx d;
    d.tp = tpitch;
    d.tile = (int)(tile_off >> 1) + (ty + 1) * tpitch + tx + 1;
    d.rp = rpitch;
    d.res = (int)(res_off >> 1) + ty * rpitch + tx;
    const bool have_left = avail_l || xi > 0, have_above = avail_u || yi > 0;
    const bool a4 = log2 == 2 && xi + step4 < n4w;
    d.bits = mode | log2 << 4 | (int)have_above << 8 | (int)have_left << 9 | (int)a4 << 10 | 1 << 11;
    d.lim_a = (int)((g.mi_cols * 8) >> sx) - 1 - (int)t.x;
    d.lim_l = (int)((g.mi_rows * 8) >> sy) - 1 - (int)t.y;
    d._pad = 0;
    txs[p * 256 + (ty >> log2) * 16 + (tx >> log2)] = d;
  }
  __syncthreads();
  const int nlv = max(max(gn0, gn1), gn2);
  // DC values: a 4x4 or 8x8 block's (8 or 16 edge reads) by each of its
  // lanes in the sample pass; a larger block's once, in a pass of its own
  // before the samples, run only by an item with such a block (the luma
  // modes of a block below 8x8 are 4x4 transforms)
  const bool any_dc = (np0 && lg0 >= 4 && mode_y == DC_PRED) ||
                      (((np1 && lg1 >= 4) || (np2 && lg2 >= 4)) && it.uv_mode == DC_PRED);
  // level lv of plane p: rows r_lo .. of the grid, n_l blocks
  auto level = [&](int p, int lv, int &r_lo) -> int {
    const int W = sel3(p, gw0, gw1, gw2), H = sel3(p, gh0, gh1, gh2), k = sel3(p, kk0, kk1, kk2);
    if (lv >= sel3(p, gn0, gn1, gn2)) return 0;
    r_lo = lv > W - 1 ? (lv - (W - 1) + k - 1) / k : 0;
    const int r_hi = min(H - 1, lv / k);
    return max(r_hi - r_lo + 1, 0);
  };
  for (int lv = 0; lv < nlv; lv++) {
    int r0 = 0, r1 = 0, r2 = 0;
    const int n0 = level(0, lv, r0), n1 = level(1, lv, r1), n2 = level(2, lv, r2);
    if (any_dc) {
      if (tid < n0 + n1 + n2) {
        const int p = tid < n0 ? 0 : tid < n0 + n1 ? 1 : 2;
        const int jt = tid - (p == 0 ? 0 : p == 1 ? n0 : n0 + n1);
        const int r = (p == 0 ? r0 : p == 1 ? r1 : r2) + jt, c = lv - sel3(p, kk0, kk1, kk2) * r;
        const IbTx d = txs[p * 256 + r * 16 + c];
        const int l2 = sel3(p, lg0, lg1, lg2);
        if (d.valid() && d.mode() == DC_PRED && l2 >= 4)  // every edge read in flight at once
          dcv[p * 16 + jt] = l2 == 4 ? d.dc<4>(bd) : d.dc<5>(bd);
      }
      __syncthreads();
    }
    const int i0 = n0 << (2 * lg0), i1 = n1 << (2 * lg1), i2 = n2 << (2 * lg2);
    for (int item = tid; item < i0 + i1 + i2; item += NT) {
      const int p = item < i0 ? 0 : item < i0 + i1 ? 1 : 2;
      const int local = item - (p == 0 ? 0 : p == 1 ? i0 : i0 + i1);
      const int l2 = sel3(p, lg0, lg1, lg2), jt = local >> (2 * l2), s = local & ((1 << (2 * l2)) - 1);
      const int r = (p == 0 ? r0 : p == 1 ? r1 : r2) + jt, c = lv - sel3(p, kk0, kk1, kk2) * r;
      const IbTx d = txs[p * 256 + r * 16 + c];
      if (!d.valid()) continue;
      const int mode = d.mode();
      const int size = 1 << l2, i = s >> l2, j = s & (size - 1);
      const int pv = intra_sample_at(
          mode, i, j, size, [&](int k) { return d.A(k, size, base); }, [&](int k) { return d.L(k, base); },
          mode != DC_PRED ? 0 : l2 >= 4 ? dcv[p * 16 + jt] : l2 == 3 ? d.dc<3>(bd) : d.dc<2>(bd), maxv);
      const int16_t *res = (const int16_t *)ib_lds;
      uint16_t *tile = (uint16_t *)ib_lds;
      tile[d.tile + i * d.tp + j] = (uint16_t)iclamp(pv + res[d.res + i * d.rp + j], 0, maxv);
    }
    __syncthreads();
  }
  ib_store<HBD, NT>(g, ly, rg0, rg1, rg2, pl_, st_, tid);
}

void launch_intra(hipStream_t s, const IntraArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd) {
  if (!wgs) return;
  if (hbd)
    hipLaunchKernelGGL(k_intra<true>, dim3(wgs), dim3(64), 0, s, fa, r, n);
  else
    hipLaunchKernelGGL(k_intra<false>, dim3(wgs), dim3(64), 0, s, fa, r, n);
}

// ---- the fused wavefront launch (k_wave)
//
// One launch holds one intra dependency level and one loop filter level,
// both on the kernel stream: workgroups below n_iw are intra blocks of the
// level (intra_block), the rest are loop filter work items of the level
// (lf_unit: waves 0..2 filter the three planes, wave 3 only joins the
// barriers).  The host places loop filter level j in a launch after the
// intra levels it needs (lf_need) and after level j - 1, so the two chains
// overlap without cross-stream events.  A launch of one frame (ONE) passes
// the frame's planes, decisions, level, items, residual planes and records
// as arguments (the first 16 dwords preloaded, all known to be global
// memory); a batch goes through the range tables of both parts.
template <bool HBD, bool ONE>
__global__ __launch_bounds__(ib_threads(ONE)) void k_wave(uint32_t n_iw, uint32_t geo, uint32_t t1, uint8_t *p0,
                                                     uint8_t *p1, uint8_t *p2, uint32_t *dq, uint32_t s0, uint32_t s1,
                                                     uint32_t s2, const IntraBlockItem *items0, const int16_t *rp0,
                                                     const int16_t *rp1, const int16_t *rp2, const vp9r_txb *txbs1,
                                                     const IntraArgs *fa_i, const Range *rg_i, const LfArgs *fa_l,
                                                     const Range *rg_l, uint32_t nf) {
  if (blockIdx.x < n_iw) {
    intra_block<HBD, ONE>(fa_i, rg_i, nf, items0, p0, p1, p2, rp0, rp1, rp2, txbs1, blockIdx.x);
    return;
  }
  const uint32_t w = blockIdx.x - n_iw;
  // the wave's plane, uniform across the wave
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), plane = wave < 3 ? wave : 0;
  typedef LfShared<typename std::conditional<HBD, uint16_t, uint8_t>::type> LfS;
  LfS &sh = ((LfS *)ib_lds)[plane];
  const int lane = (int)(threadIdx.x & 63);
  if (!ONE) {
    const uint32_t fi = range_frame(rg_l, nf, w);
    const LfArgs &a = fa_l[fi];
    const LfOne o = lf_one(a, rg_l[fi].first);
    lf_unit<HBD>((uint8_t *)a.cur.p[plane], a.cur.stride[plane], o.dq, o.geo, o.t, (int)(w - rg_l[fi].wg), plane,
                 wave < 3, lane, sh);
    return;
  }
  lf_unit<HBD>(plane == 0 ? p0 : plane == 1 ? p1 : p2, plane == 0 ? s0 : plane == 1 ? s1 : s2, dq, geo, t1, (int)w,
               plane, wave < 3, lane, sh);
}

uint32_t wave_lds_bytes(int ss_x, int ss_y) {
  // the u16 form, the larger: one size for both sample types
  return max(ib_layout(ss_x, ss_y).bytes, (uint32_t)(3 * sizeof(LfShared<uint16_t>)));
}

void launch_wave(hipStream_t s, const WaveLaunch &w, bool hbd) {
  if (!w.n_iw && !w.n_lw) return;
  // dynamic LDS above 64 KB needs the attribute
  static bool attr_set[2][2] = {{false, false}, {false, false}};
  const bool one = w.one_i || w.one_l;
  auto go = [&](auto kernel, bool h) {
    if (!attr_set[h][one]) {
      (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      attr_set[h][one] = true;
    }
    LfOne lo = w.one_l ? lf_one(*w.one_l, w.lf_level) : LfOne{};
    uint8_t *p[3] = {lo.p[0], lo.p[1], lo.p[2]};
    const int16_t *rp[3] = {nullptr, nullptr, nullptr};
    const vp9r_txb *txbs = nullptr;
    if (w.one_i) {
      for (int k = 0; k < 3; k++) {
        p[k] = (uint8_t *)w.one_i->cur.p[k];
        rp[k] = w.one_i->res[k];
      }
      txbs = w.one_i->txbs;
    }
    hipLaunchKernelGGL(kernel, dim3(w.n_iw + w.n_lw), dim3(ib_threads(one)), w.lds_bytes, s, w.n_iw, lo.geo, lo.t, p[0],
                       p[1], p[2], lo.dq, lo.stride[0], lo.stride[1], lo.stride[2], w.items0, rp[0], rp[1], rp[2],
                       txbs, w.fa_i, w.rg_i, w.fa_l, w.rg_l, w.nf);
  };
  if (one) {
    if (hbd)
      go(k_wave<true, true>, true);
    else
      go(k_wave<false, true>, false);
  } else {
    if (hbd)
      go(k_wave<true, false>, true);
    else
      go(k_wave<false, false>, false);
  }
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

// Per-kernel seam, batched: one intra::predict (intra.rs:19-29) per
// workgroup, block i's edges at edges + desc[i].edge_off (aboveRow[-1 ..
// 2 size - 1], then leftCol[0 .. size - 1]), its prediction written at
// dst + desc[i].dst_off samples with row stride `stride`.
template <bool HBD>
__global__ __launch_bounds__(64) void k_intra_pred_batch(const vp9r_intra_desc *desc, const int32_t *edges,
                                                         void *dst_, uint32_t stride, int bd) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ int32_t e[65 + 32];
  const vp9r_intra_desc d = desc[blockIdx.x];
  const int log2 = d.log2, size = 1 << log2;
  const int32_t *src = edges + d.edge_off;
  for (int k = threadIdx.x; k <= 2 * size; k += 64) e[k] = src[k];
  for (int k = threadIdx.x; k < size; k += 64) e[65 + k] = src[2 * size + 1 + k];
  __syncthreads();
  const int dc = dc_value(log2, e + 1, e + 65, d.have_left, d.have_above, bd);
  P *dst = (P *)dst_ + d.dst_off;
  for (int k = threadIdx.x; k < size * size; k += 64) {
    const int i = k >> log2, j = k & (size - 1);
    dst[(size_t)i * stride + j] = (P)intra_sample(d.mode, i, j, size, e + 1, e + 65, dc, (1 << bd) - 1);
  }
}

void launch_intra_pred_batch(hipStream_t s, uint32_t count, const vp9r_intra_desc *desc, const int32_t *edges,
                             void *dst, uint32_t stride, uint32_t bit_depth) {
  if (!count) return;
  if (bit_depth > 8)
    hipLaunchKernelGGL(k_intra_pred_batch<true>, dim3(count), dim3(64), 0, s, desc, edges, dst, stride,
                       (int)bit_depth);
  else
    hipLaunchKernelGGL(k_intra_pred_batch<false>, dim3(count), dim3(64), 0, s, desc, edges, dst, stride,
                       (int)bit_depth);
}

}  // namespace vp9g
