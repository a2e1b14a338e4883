// Kernel argument blocks and launchers: the interface between the host
// orchestration (host/recon.cpp) and the gfx950 kernels (gpu/*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vp9r.h"

namespace vp9g {

// One HBM-resident plane set: u8 samples at 8 bits, u16 above.
struct Planes {
  void *p[3];
  uint32_t stride[3];  // in samples
  uint32_t rows[3];
};

// The frame geometry every kernel reads.
struct FrameGeom {
  uint32_t width, height, mi_cols, mi_rows, sb64_cols, sb64_rows;
  uint8_t bit_depth, ss_x, ss_y, legacy_uv;
  uint8_t lossless, _pad[3];
};

// An inter prediction tile: a tw x th piece (at most 16 x 16) at offset
// (tx, ty) of one prediction unit (a block's plane, or one 4x4 sub-block)
// of size w x h at plane position (x, y), with what its reference position
// needs of its block, so that a workgroup reads its tile and then the
// references: the unit's motion vector per reference list (selected, and
// averaged for chroma sub-blocks, 8.5.2.1), the block's position, size,
// reference frames and filter.
struct McTile {
  int32_t mv[2][2];
  uint32_t block;
  uint16_t x, y;
  uint16_t mi_row, mi_col;
  uint8_t plane, sub, w, h;
  uint8_t tx, ty, tw, th;
  uint8_t mi_size, interp_filter;
  int8_t ref_frame[2];
};

// Cross-frame batching.  One launch covers the work items of several
// frames (of different streams): frame i of a launch takes workgroups
// [r[i].wg, r[i + 1].wg) and, in its own item list, the items
// [r[i].first, r[i].first + r[i].count); r[n].wg is the launch's workgroup
// count.  Per-frame arguments are arrays in device memory indexed by frame.
struct Range {
  uint32_t wg, first, count, _pad;
};

// The frame of workgroup w: the last i < n with r[i].wg <= w (r sorted by
// wg).  Each lane of the wave reads one entry and a ballot counts those at
// or below w, 64 entries per memory round trip (a binary search would wait
// for one round trip per halving).  Called by whole waves in uniform control
// flow.
__device__ __forceinline__ uint32_t range_frame(const Range *r, uint32_t n, uint32_t w) {
  if (n <= 1) return 0;
  const uint32_t lane = __lane_id();
  for (uint32_t base = 0;; base += 64) {
    const uint32_t i = base + lane;
    const uint64_t le = __ballot(i < n && r[i].wg <= w);
    const uint32_t c = (uint32_t)__popcll(le);
    if (c < 64 || base + 64 >= n) return base + c > 0 ? base + c - 1 : 0;
  }
}

// Workgroup b of a grid of g, renumbered so that each XCD takes a
// contiguous run of the work: the hardware hands workgroups to the 8 XCDs in
// turn (b % 8), each XCD has its own L2, and neighbouring tiles share
// reference lines and output lines.  XCD x's workgroups take the x-th
// eighth, in order (a bijection on [0, g)).
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t g) {
  const uint32_t q = g >> 3, r = g & 7, x = b & 7, k = b >> 3;
  return x * q + min(x, r) + k;
}

struct ClearArgs {
  Planes f;
  FrameGeom g;
};

struct McArgs {
  FrameGeom g;
  const vp9r_block *blocks;
  const McTile *tiles;
  Planes cur;
  Planes ref[3];
  uint32_t ref_w[3], ref_h[3];
};

// Dequantisation and inverse transform of the transform blocks of inter
// blocks (one size class per launch).
struct ItxArgs {
  FrameGeom g;
  const vp9r_block *blocks;
  const vp9r_txb *txbs;
  const uint32_t *jobs[4];  // txb indices per size class
  const uint16_t *coef_pos;
  const int32_t *coef_val;
  Planes cur;
  int32_t q[8][2][2];  // [segment][plane > 0][dc, ac]
  // Jobs flagged kItxIntra are transform blocks of intra blocks: their
  // residuals go to res (int16 planes with cur's strides), not into cur,
  // since the prediction they add to comes later (k_intra_blocks).
  int16_t *res[3];
};
constexpr uint32_t kItxIntra = 0x80000000u;

// The intra wavefront: one group per (superblock, plane) with intra work;
// a launch runs the groups of one dependency level.
struct IntraGroup {
  uint16_t sb_row, sb_col;
  uint8_t plane, _pad[3];
  uint32_t first, count;  // into the intra txb index list
};

// An intra block: its transform blocks are txbs[first .. first + count)
// (planes in order).
struct IntraItem {
  uint32_t block, first, count;
};

// An intra block for k_intra_blocks: its transform blocks are
// txbs[first .. first + count) (n_y luma, n_u Cb, then Cr; one transform
// size per plane), their coefficients coef_*[coef_first .. + coef_count);
// their residuals are in IntraArgs::res (written by k_itx_inter).
struct IntraBlockItem {
  uint32_t first, count, coef_first, coef_count;
  uint16_t mi_row, mi_col, n_y, n_u;
  uint8_t mi_size, y_mode, uv_mode, flags;  // flags: segment_id | avail_u << 3 | avail_l << 4
  uint32_t sub_modes;                       // 4 x 8 bits
};

// A superblock item (an IntraBlockItem with kIbSuperblock in flags) holds
// every intra block of one superblock: its transform blocks (all planes)
// as descriptors sb_tx[first .. first + count), sorted by their level
// inside the superblock; chunks sb_units[coef_first .. + coef_count), by
// level, each up to four 4x4 units of one transform block (descriptor
// index | first unit << 10 | units << 18, units in raster order inside the
// block), so that a wave's lanes predict one mode; levels
// sb_levels[sub_modes .. sub_modes + n_u), each x = first descriptor |
// count << 16 | (a DC_PRED block of 16x16 or larger) << 31, y = first chunk
// | count << 16 (relative to the item's first); (mi_row, mi_col) the
// superblock's origin.
constexpr uint8_t kIbSuperblock = 0x80;
// The most levels a superblock item may have (the host keeps frames with
// more in the per-block form).
constexpr uint32_t kIbSbMaxLevels = 128;
// A transform block of a superblock item: plane, log2 of its size, its
// prediction mode, flags (have_above 1, have_left 2, above-right read
// inside its block 4: recon.rs:61-69), its plane position inside the
// superblock and the clamps of its edge reads (the frame's last column
// and row minus its position, recon.rs:42-43).
struct IntraSbTx {
  uint8_t plane, log2, mode, flags;
  uint8_t x, y;
  int16_t lim_a, lim_l;
  uint16_t _pad[3];
};

// LDS of one k_intra_blocks workgroup for a frame of
// subsampling (ss_x, ss_y): byte offsets of the transform block records
// (all of the block's), the three transform grids' descriptors (32 B per
// cell), DC values, plane geometry, and per plane the region tile ((64 >> sx) + 1 columns, u16) and
// residuals ((64 >> sx) columns, int16).  The launch passes `bytes` (the
// largest of its frames') as dynamic LDS.
// Lanes of an intra item's workgroup: a single frame's launches take 1024
// (16 waves, the most a workgroup takes: the large transform blocks' levels,
// up to 3 x 1024 samples, are SIMD-bound at 8 waves and their LDS reads wait
// less with 16; 1080p intra stage 0.147 -> 0.141 ms per frame against 512),
// a batch's 512 (with items of many frames on the chip, 1024-lane
// workgroups made a batch of 64 frames 12% slower).
constexpr int kIbThreadsOne = 1024, kIbThreadsBatch = 512, kIbGeoBytes = 80;
constexpr int ib_threads(bool one) { return one ? kIbThreadsOne : kIbThreadsBatch; }
// The most transform blocks an intra block has: 64x64 at 4x4 in every plane.
__host__ __device__ inline uint32_t ib_max_txbs(int ss_x, int ss_y) { return 256 + 2 * (256u >> (ss_x + ss_y)); }
struct IbLayout {
  uint32_t txb, txs, dcv, geo, tile[3], res[3], bytes;
};
__host__ __device__ inline IbLayout ib_layout(int ss_x, int ss_y) {
  IbLayout l{};
  uint32_t o = 0;
  auto take = [&o](uint32_t n) {
    const uint32_t r = o;
    o = (o + n + 15u) & ~15u;
    return r;
  };
  l.txb = take(ib_max_txbs(ss_x, ss_y) * (uint32_t)sizeof(vp9r_txb));
  l.txs = take(3 * 256 * 32);
  l.dcv = take(3 * 16 * 4);
  l.geo = take(3 * kIbGeoBytes);
  for (int p = 0; p < 3; p++) {
    const int sx = p ? ss_x : 0, sy = p ? ss_y : 0;
    l.tile[p] = take((uint32_t)((64 >> sx) + 1) * (uint32_t)((64 >> sy) + 1) * 2);
  }
  for (int p = 0; p < 3; p++) {
    const int sx = p ? ss_x : 0, sy = p ? ss_y : 0;
    l.res[p] = take((uint32_t)(64 >> sx) * (uint32_t)(64 >> sy) * 2);
  }
  l.bytes = o;
  return l;
}

struct IntraArgs {
  FrameGeom g;
  const vp9r_block *blocks;
  const vp9r_txb *txbs;
  const uint32_t *txb_index;  // groups: the intra txbs of each group
  const IntraGroup *groups;   // by level
  const IntraBlockItem *items;  // intra blocks by level
  const uint16_t *coef_pos;
  const int32_t *coef_val;
  Planes cur;
  int32_t q[8][2][2];
  const int16_t *res[3];  // the intra transform blocks' residuals (k_itx_inter)
  // superblock items (kIbSuperblock)
  const IntraSbTx *sb_tx;
  const uint32_t *sb_units;
  const uint2 *sb_levels;
};

// Loop filter: per superblock, plane and pass, 16 x 16 entries of filter
// size (int8, -1 = none) and filter level (uint8).
struct LfParams {
  uint8_t lvl[8][4][2];     // LvlLookup[segment][ref][mode type]
  uint8_t limit[64], blimit[64], thresh[64];
};

// k_lf_mask's ranges cover 6 workgroups per superblock, or one (large
// batches); k_lf's
// ranges have first = the level t and count = 3 x its n items (workgroup j:
// plane j / n of the level's (j % n)-th item: the units (r, t - r), rows
// ascending, then the lone vertical pass of superblock (t + 1, 0)).
struct LfArgs {
  FrameGeom g;
  const vp9r_mi *mi;
  // [plane][pass][sb][16 edges][16 runs]: (filter size + 1) | limit << 8 |
  // blimit << 16 | thresh << 24, 0 for no filter
  uint32_t *dq;
  Planes cur;
  LfParams p;
};

// What a k_lf workgroup needs of its frame: the planes, the decisions, the
// superblock grid, subsampling and bit depth (geo: cols | rows << 12 |
// ss_x << 24 | ss_y << 25 | bit_depth << 26), and the level t.  13 dwords,
// first among the kernel's arguments, so that they arrive preloaded.
struct LfOne {
  uint8_t *p[3];
  uint32_t stride[3];
  uint32_t *dq;
  uint32_t geo, t;
};
__host__ __device__ inline LfOne lf_one(const LfArgs &a, uint32_t t) {
  LfOne o;
  for (int k = 0; k < 3; k++) {
    o.p[k] = (uint8_t *)a.cur.p[k];
    o.stride[k] = a.cur.stride[k];
  }
  o.dq = a.dq;
  o.geo = a.g.sb64_cols | a.g.sb64_rows << 12 | a.g.ss_x << 24 | a.g.ss_y << 25 | a.g.bit_depth << 26;
  o.t = t;
  return o;
}

// Launchers: fa = per-frame arguments, r = ranges (n + 1 entries), all in
// device memory; wgs = r[n].wg; hbd: samples are u16.
constexpr uint32_t kClearRows = 16;  // rows per margin clear workgroup
uint32_t clear_margin_rows(const Planes &f, const FrameGeom &g);
void launch_clear_margins(hipStream_t s, const ClearArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd);
// one_tiles: for a launch of one frame, its one_count tiles (device
// memory), read straight from a kernel argument.  A launch of one frame
// takes workgroups of kMcWavesOne waves, a tile each (half the workgroups
// to dispatch for the few thousand tiles of a frame: 12.3 to 11.4 us at
// 1080p; 4 waves 11.5-11.8, 8 waves 12.1); a batch one wave per workgroup
// (its ranges count mc_tiles_per_wg(n) tiles per workgroup), as the
// per-kernel seam (4 waves: 1.66 to 1.79 ms for the microbenchmark).
constexpr int kMcWavesOne = 2;
inline uint32_t mc_tiles_per_wg(uint32_t n_frames) { return n_frames == 1 ? kMcWavesOne : 1; }
void launch_mc(hipStream_t s, const McArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd,
               const McTile *one_tiles, uint32_t one_count);
// A single frame's first launch: its margin clear (n_clear workgroups, as
// launch_clear_margins), loop filter decisions (n_mask, as launch_lf_mask
// with per_sb entries per workgroup; 0 for none) and inter prediction
// (n_tiles tiles, as launch_mc of one frame).
void launch_prep_mc(hipStream_t s, const ClearArgs *fc, const Range *rc, uint32_t n_clear, const LfArgs *fl,
                    const Range *rl, uint32_t n_mask, int per_sb, const McArgs *fa, const Range *rg,
                    const McTile *tiles, uint32_t n_tiles, bool hbd);
// r[k], wgs[k]: the ranges and workgroups of transform size k (4x4 .. 32x32)
void launch_itx_inter(hipStream_t s, const ItxArgs *fa, const Range *const r[4], uint32_t n, const uint32_t wgs[4],
                      bool hbd);
void launch_intra(hipStream_t s, const IntraArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd);
// One fused wavefront launch (k_wave): n_iw intra block workgroups (ranges
// rg_i over fa_i) and n_lw loop filter work items (ranges rg_l over fa_l,
// first = the frame's level).  For a launch of one frame, one_i / one_l are
// host copies of its arguments (whose pointers become kernel arguments),
// items0 its items of the intra level and lf_level its loop filter level;
// else they are null and every workgroup goes through the ranges.
struct WaveLaunch {
  uint32_t n_iw = 0, n_lw = 0, nf = 0, lds_bytes = 0, lf_level = 0;
  const IntraArgs *fa_i = nullptr;
  const Range *rg_i = nullptr;
  const LfArgs *fa_l = nullptr;
  const Range *rg_l = nullptr;
  const IntraArgs *one_i = nullptr;
  const LfArgs *one_l = nullptr;
  const IntraBlockItem *items0 = nullptr;
};
void launch_wave(hipStream_t s, const WaveLaunch &w, bool hbd);
// dynamic LDS of a k_wave workgroup for a frame of subsampling (ss_x, ss_y)
uint32_t wave_lds_bytes(int ss_x, int ss_y);
// per_sb: 1 (a workgroup per superblock entry, 6 per superblock) or 6 (a
// workgroup per superblock); the ranges count workgroups accordingly
void launch_lf_mask(hipStream_t s, const LfArgs *fa, const Range *r, uint32_t n, uint32_t wgs, int per_sb);
// The margin clear (n_clear workgroups, ranges rc over fc) and the loop
// filter decisions (n_mask workgroups, ranges rl over fl, per_sb as
// launch_lf_mask) of a batch of n frames as one launch.
void launch_prep(hipStream_t s, const ClearArgs *fc, const Range *rc, uint32_t n_clear, const LfArgs *fl,
                 const Range *rl, uint32_t n_mask, uint32_t n, int per_sb, bool hbd);
// one: for a launch of one frame, lf_one of its arguments and the level
// (else p[0] null: each workgroup finds its frame through the range table)
void launch_lf_level(hipStream_t s, const LfArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd,
                     const LfOne &one);
// Levels one.t .. one.t + depth - 1 of one frame in one launch
// (lf_cone.hip; depth 2 or 3, 4 at 8 bits), from the planes one.p (the
// frame after level one.t - 1) into the planes `out` (same geometry), so
// that `out` holds the frame after level one.t + depth - 1: prev = 0 when
// `out` holds nothing of the frame yet, else the depth of the previous
// launch, which read `out` (the frame after level one.t - prev - 1).
void launch_lf_cone(hipStream_t s, const LfOne &one, const Planes &out, int depth, int prev, bool hbd);

// Per-kernel seam (vp9r.h batched forms).
// A transform launch shape: waves per workgroup (1, 2 or 4) | the flags.
constexpr int kItxShapeReg = 16;         // rows to columns in registers (4x4 .. 16x16), not LDS
constexpr int kItxShapePersistent = 32;  // resident workgroups looping over the blocks
void launch_itx_dense(hipStream_t s, uint32_t count, const int32_t *coefs, const uint32_t *coef_off,
                      const uint8_t *tx, const uint8_t *tx_type, const uint16_t *eob,
                      const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bit_depth,
                      int lossless, int tx_size, const uint32_t *jobs, uint32_t n_jobs, int shape = 1);
// Copies the interpolation taps to the current device if not yet there (the
// MC launches do it themselves; a graph capture calls it first).
void mc_prepare();
// tiles of 16 columns by tile_rows (16 or 32) rows
void launch_mc_desc_tiles(hipStream_t s, const vp9r_mc_desc *desc, const void *tiles, uint32_t n_tiles,
                          const vp9r_ref_plane *planes, void *out, uint32_t out_stride, uint32_t bd, int tile_rows);
void launch_mc_desc_expand(hipStream_t s, const vp9r_mc_desc *desc, uint32_t count, const vp9r_ref_plane *planes,
                           void *out, uint32_t out_stride, uint32_t bd);
// one pass of lf::filter_edges: edges [first, n_edges) of the region at (x0, y0)
struct LfPassDesc {
  uint16_t x0, y0;
  uint8_t vertical, first, n_edges, n_runs;
};
void launch_lf_edges(hipStream_t s, const uint8_t *tables, const LfPassDesc *passes, uint32_t count, void *buf,
                     uint32_t stride, uint32_t bit_depth);
void launch_intra_pred(hipStream_t s, uint32_t mode, uint32_t log2, const int32_t *above,
                       const int32_t *left, int have_left, int have_above, uint32_t bit_depth,
                       uint16_t *dst, uint32_t stride);
void launch_intra_pred_batch(hipStream_t s, uint32_t count, const vp9r_intra_desc *desc, const int32_t *edges,
                             void *dst, uint32_t stride, uint32_t bit_depth);
void launch_avg_batch(hipStream_t s, uint32_t count, const vp9r_avg_desc *desc, const void *a, const void *b,
                      void *dst, uint32_t stride, uint32_t bit_depth);
void launch_add_const_batch(hipStream_t s, uint32_t count, const vp9r_add_const_desc *desc, void *dst,
                            uint32_t stride);

}  // namespace vp9g
