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
// of size w x h at plane position (x, y).
struct McTile {
  uint32_t block;
  uint16_t x, y;
  uint8_t plane, sub, w, h;
  uint8_t tx, ty, tw, th;
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
  const uint32_t *jobs;  // txb indices of this size class
  const uint16_t *coef_pos;
  const int32_t *coef_val;
  Planes cur;
  int32_t q[8][2][2];  // [segment][plane > 0][dc, ac]
};

// The intra wavefront: one group per (superblock, plane) with intra work;
// a launch runs the groups of one dependency level.
struct IntraGroup {
  uint16_t sb_row, sb_col;
  uint8_t plane, _pad[3];
  uint32_t first, count;  // into the intra txb index list
};

// Device-side scheduling of a wavefront in one launch: tickets hand out
// work items in a topological order, and each item's completion flag holds
// the epoch (frame number) that completed it.
struct WaveSync {
  uint32_t *ticket;  // zero at launch
  uint32_t *done;    // per item: epoch of the last frame that completed it
  uint32_t *status;  // set to nonzero when a wait exceeds its spin limit
  uint32_t epoch;    // nonzero
};

// Polls of a completion flag before a wait gives up (with s_sleep between
// polls: on the order of a second).
constexpr uint32_t kWaveSpinLimit = 1u << 22;

// An intra block of the one-launch intra wavefront: its transform blocks
// are txbs[first .. first + count) (planes in order).
struct IntraItem {
  uint32_t block, first, count;
};

struct IntraArgs {
  FrameGeom g;
  const vp9r_block *blocks;
  const vp9r_txb *txbs;
  const uint32_t *txb_index;
  const IntraGroup *groups;
  const uint16_t *coef_pos;
  const int32_t *coef_val;
  Planes cur;
  int32_t q[8][2][2];
  const vp9r_mi *mi;         // one-launch wavefront: which 8x8 positions are intra
  const IntraItem *items;    // one-launch wavefront: intra blocks in decode order
  uint32_t n_items;
  WaveSync sync;             // flags per 8x8 position
};

// Loop filter: per superblock, plane and pass, 16 x 16 entries of filter
// size (int8, -1 = none) and filter level (uint8).
struct LfParams {
  uint8_t lvl[8][4][2];     // LvlLookup[segment][ref][mode type]
  uint8_t limit[64], blimit[64], thresh[64];
};

struct LfArgs {
  FrameGeom g;
  const vp9r_mi *mi;
  int8_t *mask_fs;   // [plane][pass][sb][16][16] (level launches)
  uint8_t *mask_lvl;
  Planes cur;
  LfParams p;
  const uint32_t *sbs;  // (plane << 28 | sb_row << 14 | sb_col): per level, or all in order
  uint32_t n_items;     // one-launch wavefront: entries of sbs
  WaveSync sync;        // one-launch wavefront: flags per (plane, superblock)
};

void launch_clear_margins(hipStream_t s, const Planes &f, const FrameGeom &g);
void launch_mc(hipStream_t s, const McArgs &a, uint32_t n_tiles);
void launch_itx_inter(hipStream_t s, const ItxArgs &a, int tx_size, uint32_t n_jobs);
void launch_intra(hipStream_t s, const IntraArgs &a, uint32_t first_group, uint32_t n_groups);
void launch_intra_wave(hipStream_t s, const IntraArgs &a, uint32_t n_groups);
void launch_intra_blocks(hipStream_t s, const IntraArgs &a, uint32_t first, uint32_t n);
void launch_lf_mask(hipStream_t s, const LfArgs &a);
void launch_lf_level(hipStream_t s, const LfArgs &a, uint32_t first, uint32_t n);
void launch_lf_wave(hipStream_t s, const LfArgs &a, uint32_t n_groups);

// Per-kernel seam (vp9r.h batched forms).
void launch_itx_dense(hipStream_t s, uint32_t count, const int32_t *coefs, const uint32_t *coef_off,
                      const uint8_t *tx, const uint8_t *tx_type, const uint16_t *eob,
                      const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bit_depth,
                      int lossless, int tx_size, const uint32_t *jobs, uint32_t n_jobs);
void launch_mc_desc(hipStream_t s, uint32_t count, const vp9r_mc_desc *desc, const void *const *refs,
                    uint32_t ref_stride, int32_t last_x, int32_t last_y, void *out,
                    uint32_t out_stride, uint32_t bit_depth);
void launch_lf_edges(hipStream_t s, uint32_t count, const uint8_t *edges, const uint16_t *x0,
                     const uint16_t *y0, const uint8_t *vertical, const uint8_t *n_edges,
                     const uint8_t *n_runs, void *buf, uint32_t stride, uint32_t bit_depth);
void launch_intra_pred(hipStream_t s, uint32_t mode, uint32_t log2, const int32_t *above,
                       const int32_t *left, int have_left, int have_above, uint32_t bit_depth,
                       uint16_t *dst, uint32_t stride);

}  // namespace vp9g
