// GPU reconstruction, host side: frame stores, per-frame work lists and the
// launch sequence.  A batch of records (one frame, or independent frames of
// several streams) goes in on one HIP stream:
//   1. one staged host-to-device copy per record of its inputs and work
//      lists, and one copy of the batch's kernel arguments and ranges;
//   2. margin clearing and inter prediction of every inter block (one launch
//      each: inter prediction reads only the references);
//   3. inter residuals, one launch per transform size;
//   4. intra prediction + residual, one launch per dependency level;
//   5. the loop filter decisions (one launch) and the filter wavefront, one
//      launch per level r + c of (horizontal pass, next vertical pass) units.
// Every launch covers all frames of the batch: a workgroup finds its frame
// in the launch's range table.
// This is the order of the reference's replay (src/decoder/block.rs:1405-1498)
// up to reorderings that change no sample: inter pixels depend only on
// references, so they may all precede the intra chain, which reads
// reconstructed neighbours only.
#include "recon.h"

#include "../gpu/mc_pos.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "vp9_consts.h"
#include "vp9_tables.h"

namespace vp9r {

hipError_t DevBuf::ensure(size_t n) {
  if (n <= cap) return hipSuccess;
  if (p) hipFree(p);
  p = nullptr;
  cap = 0;
  size_t want = std::max<size_t>(n + n / 2, 1 << 20);
  hipError_t e = hipMalloc(&p, want);
  if (e == hipSuccess) cap = want;
  return e;
}
DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

hipError_t HostBuf::ensure(size_t n) {
  if (n <= cap) return hipSuccess;
  if (p) hipHostFree(p);
  p = nullptr;
  cap = 0;
  size_t want = std::max<size_t>(n + n / 2, 1 << 20);
  hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e == hipSuccess) cap = want;
  return e;
}
HostBuf::~HostBuf() {
  if (p) (void)hipHostFree(p);
}

static int hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return VP9R_ERR_DEVICE;
}

#define HIP_TRY(expr, what)                 \
  do {                                      \
    hipError_t e_ = (expr);                 \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

// qindex (src/decoder/recon.rs:219-230) and the dc/ac factors of reconstruct
// (recon.rs:240-250) for every segment and plane type.
static void q_table(const vp9r_frame_hdr &h, int32_t q[8][2][2]) {
  int bd_idx = (h.bit_depth - 8) >> 1;
  for (int s = 0; s < 8; s++) {
    int qi = h.base_q_idx;
    if (h.seg_enabled && h.seg_feature_enabled[s][SEG_LVL_ALT_Q]) {
      int d = h.seg_feature_data[s][SEG_LVL_ALT_Q];
      if (!h.seg_abs_or_delta_update) d += h.base_q_idx;
      qi = iclamp(d, 0, 255);
    }
    q[s][0][0] = VP9_DC_QLOOKUP[bd_idx][iclamp(qi + h.delta_q_y_dc, 0, 255)];
    q[s][0][1] = VP9_AC_QLOOKUP[bd_idx][iclamp(qi, 0, 255)];
    q[s][1][0] = VP9_DC_QLOOKUP[bd_idx][iclamp(qi + h.delta_q_uv_dc, 0, 255)];
    q[s][1][1] = VP9_AC_QLOOKUP[bd_idx][iclamp(qi + h.delta_q_uv_ac, 0, 255)];
  }
}

// lvl_lookup (src/decoder/loopfilter.rs:15-47) and the limits of Ctx::new
// (loopfilter.rs:135-170).
static void lf_params(const vp9r_frame_hdr &h, vp9g::LfParams &p) {
  int n_shift = h.lf_level >> 5;
  for (int s = 0; s < 8; s++) {
    int lvl_seg = h.lf_level;
    if (h.seg_enabled && h.seg_feature_enabled[s][SEG_LVL_ALT_L]) {
      int d = h.seg_feature_data[s][SEG_LVL_ALT_L];
      lvl_seg = iclamp(h.seg_abs_or_delta_update ? d : d + h.lf_level, 0, MAX_LOOP_FILTER);
    }
    if (!h.lf_delta_enabled) {
      for (int r = 0; r < 4; r++) p.lvl[s][r][0] = p.lvl[s][r][1] = (uint8_t)lvl_seg;
    } else {
      int intra = lvl_seg + h.lf_ref_deltas[INTRA_FRAME] * (1 << n_shift);
      p.lvl[s][0][0] = p.lvl[s][0][1] = (uint8_t)iclamp(intra, 0, MAX_LOOP_FILTER);
      for (int rf = 1; rf < 4; rf++)
        for (int m = 0; m < 2; m++)
          p.lvl[s][rf][m] = (uint8_t)iclamp(
              lvl_seg + h.lf_ref_deltas[rf] * (1 << n_shift) + h.lf_mode_deltas[m] * (1 << n_shift), 0,
              MAX_LOOP_FILTER);
    }
  }
  int sharp = h.lf_sharpness, shift = sharp > 4 ? 2 : sharp > 0 ? 1 : 0;
  for (int l = 0; l < 64; l++) {
    int limit = sharp > 0 ? iclamp(l >> shift, 1, 9 - sharp) : std::max(l >> shift, 1);
    p.limit[l] = (uint8_t)limit;
    p.blimit[l] = (uint8_t)(2 * (l + 2) + limit);
    p.thresh[l] = (uint8_t)(l >> 4);
  }
}

static vp9g::FrameGeom geom_of(const vp9r_frame_hdr &h) {
  vp9g::FrameGeom g{};
  g.width = h.width;
  g.height = h.height;
  g.mi_cols = h.mi_cols;
  g.mi_rows = h.mi_rows;
  g.sb64_cols = h.sb64_cols;
  g.sb64_rows = h.sb64_rows;
  g.bit_depth = h.bit_depth;
  g.ss_x = h.ss_x;
  g.ss_y = h.ss_y;
  g.legacy_uv = h.legacy_uv;
  g.lossless = h.lossless;
  return g;
}

// The inter prediction tiles of the record (block.rs:1425-1448): each
// block's planes as one prediction unit, or one 4x4 unit per sub-block
// below 8x8; units split into tiles of at most 16 x 16.
// A prediction unit's tiles are 16 columns by 16 rows, or by 32 rows when
// every reference it reads is unscaled (a step of one sample: the kernel's
// fast paths, which take 32-row tiles; the general path takes 16).
static void build_mc_tiles(Builder *g, const vp9r_record *rec) {
  const vp9r_frame_hdr &h = rec->hdr;
  const vp9g::FrameGeom geo = geom_of(h);
  bool unscaled[3];
  for (int ri = 0; ri < 3; ri++) {
    const int64_t xsc = ((int64_t)h.ref_width[ri] << 14) / h.width, ysc = ((int64_t)h.ref_height[ri] << 14) / h.height;
    unscaled[ri] = ((16 * xsc) >> 14) == 16 && ((16 * ysc) >> 14) == 16;
  }
  auto &tiles = g->tiles;
  tiles.clear();
  for (uint32_t bi = 0; bi < rec->n_blocks; bi++) {
    const vp9r_block &b = rec->blocks[bi];
    if (!b.is_inter) continue;
    const bool compound = b.ref_frame[1] > INTRA_FRAME;
    const int r0 = b.ref_frame[0] - LAST_FRAME, r1 = compound ? b.ref_frame[1] - LAST_FRAME : r0;
    const int tile_rows = r0 >= 0 && r0 < 3 && r1 >= 0 && r1 < 3 && unscaled[r0] && unscaled[r1] ? 32 : 16;
    int bsize = b.mi_size < BLOCK_8X8 ? BLOCK_8X8 : b.mi_size;
    for (int plane = 0; plane < 3; plane++) {
      int sx = plane ? h.ss_x : 0, sy = plane ? h.ss_y : 0;
      int psz = VP9_SS_SIZE_LOOKUP[bsize][sx][sy];
      int n4w = VP9_NUM_4X4_WIDE[psz], n4h = VP9_NUM_4X4_HIGH[psz];
      int base_x = (b.mi_col * 8) >> sx, base_y = (b.mi_row * 8) >> sy;
      auto unit = [&](int x, int y, int w, int hh, int sub) {
        for (int ty = 0; ty < hh; ty += tile_rows)
          for (int tx = 0; tx < w; tx += 16) {
            vp9g::McTile t = {};
            for (int rl = 0; rl < (compound ? 2 : 1); rl++)
              vp9g::mc_select_mv(geo, b, plane, sub, rl, t.mv[rl][0], t.mv[rl][1]);
            t.block = bi;
            t.x = (uint16_t)x;
            t.y = (uint16_t)y;
            t.mi_row = (uint16_t)b.mi_row;
            t.mi_col = (uint16_t)b.mi_col;
            t.plane = (uint8_t)plane;
            t.sub = (uint8_t)sub;
            t.w = (uint8_t)w;
            t.h = (uint8_t)hh;
            t.tx = (uint8_t)tx;
            t.ty = (uint8_t)ty;
            t.tw = (uint8_t)std::min(16, w - tx);
            t.th = (uint8_t)std::min(tile_rows, hh - ty);
            t.mi_size = (uint8_t)b.mi_size;
            t.interp_filter = (uint8_t)b.interp_filter;
            t.ref_frame[0] = (int8_t)b.ref_frame[0];
            t.ref_frame[1] = (int8_t)b.ref_frame[1];
            tiles.push_back(t);
          }
      };
      if (b.mi_size < BLOCK_8X8) {
        for (int y = 0; y < n4h; y++)
          for (int x = 0; x < n4w; x++) unit(base_x + 4 * x, base_y + 4 * y, 4, 4, y * n4w + x);
      } else {
        unit(base_x, base_y, n4w * 4, n4h * 4, 0);
      }
    }
  }
}

// Intra groups per (plane, superblock) and their dependency levels.
static void build_intra(Builder *g, const vp9r_record *rec) {
  const vp9r_frame_hdr &h = rec->hdr;
  const uint32_t n_sb = h.sb64_rows * h.sb64_cols;
  g->intra_cnt.assign(3 * n_sb, 0);
  g->intra_fill.assign(3 * n_sb, 0);
  auto key_of = [&](const vp9r_txb &t) {
    int sx = t.plane ? h.ss_x : 0, sy = t.plane ? h.ss_y : 0;
    uint32_t c = t.x / (64u >> sx), r = t.y / (64u >> sy);
    return t.plane * n_sb + r * h.sb64_cols + c;
  };
  uint32_t n_intra = 0;
  for (uint32_t i = 0; i < rec->n_txbs; i++) {
    const vp9r_txb &t = rec->txbs[i];
    if (rec->blocks[t.block].is_inter) continue;
    g->intra_cnt[key_of(t)]++;
    n_intra++;
  }
  g->intra_first.resize(3 * n_sb);
  uint32_t acc = 0;
  for (uint32_t k = 0; k < 3 * n_sb; k++) {
    g->intra_first[k] = acc;
    acc += g->intra_cnt[k];
  }
  g->intra_index.resize(n_intra);
  for (uint32_t i = 0; i < rec->n_txbs; i++) {
    const vp9r_txb &t = rec->txbs[i];
    if (rec->blocks[t.block].is_inter) continue;
    uint32_t k = key_of(t);
    g->intra_index[g->intra_first[k] + g->intra_fill[k]++] = i;
  }
  // the intra blocks in decode order (the record keeps a block's transform
  // blocks together, planes in order)
  g->intra_items.clear();
  for (uint32_t i = 0; i < rec->n_txbs; i++) {
    const vp9r_txb &t = rec->txbs[i];
    if (rec->blocks[t.block].is_inter) continue;
    if (g->intra_items.empty() || g->intra_items.back().block != t.block)
      g->intra_items.push_back(vp9g::IntraItem{t.block, i, 0});
    g->intra_items.back().count++;
  }
  // Levels of the intra blocks: one above the deepest intra block covering
  // an 8x8 position the block reads (the row above from above-left to its
  // top-right, the column left); inter positions are complete before the
  // intra launches.  Items are ordered by level (a topological order).
  {
    const uint32_t cols = h.mi_cols, rows = h.mi_rows;
    g->cell_level.assign((size_t)rows * cols, -1);
    g->item_level.resize(g->intra_items.size());
    int max_lv = -1;
    for (size_t k = 0; k < g->intra_items.size(); k++) {
      const vp9r_block &b = rec->blocks[g->intra_items[k].block];
      const int bw8 = b.mi_size < BLOCK_8X8 ? 1 : VP9_NUM_8X8_WIDE[b.mi_size];
      const int bh8 = b.mi_size < BLOCK_8X8 ? 1 : VP9_NUM_8X8_HIGH[b.mi_size];
      int lv = 0;
      auto dep = [&](int r, int c) {
        if (r < 0 || c < 0 || r >= (int)rows || c >= (int)cols) return;
        const int l = g->cell_level[(size_t)r * cols + c];
        if (l >= 0) lv = std::max(lv, l + 1);
      };
      if (b.avail_u)
        for (int i = -1; i < bw8; i++) dep((int)b.mi_row - 1, (int)b.mi_col + i);
      if (b.avail_l)
        for (int i = 0; i < bh8; i++) dep((int)b.mi_row + i, (int)b.mi_col - 1);
      g->item_level[k] = lv;
      max_lv = std::max(max_lv, lv);
      for (int r = b.mi_row; r < (int)b.mi_row + bh8 && r < (int)rows; r++)
        for (int c = b.mi_col; c < (int)b.mi_col + bw8 && c < (int)cols; c++)
          g->cell_level[(size_t)r * cols + c] = lv;
    }
    g->item_level_start.assign(max_lv + 2, 0);
    for (int lv : g->item_level) g->item_level_start[lv + 1]++;
    for (int l = 0; l <= max_lv; l++) g->item_level_start[l + 1] += g->item_level_start[l];
    std::vector<uint32_t> fill(g->item_level_start.begin(), g->item_level_start.end());
    g->items_sorted.resize(g->intra_items.size());
    for (size_t k = 0; k < g->intra_items.size(); k++) g->items_sorted[fill[g->item_level[k]]++] = g->intra_items[k];
  }
  // level = 1 + the deepest of the left, above and above-left groups
  g->intra_level.assign(3 * n_sb, -1);
  int max_level = -1;
  for (int p = 0; p < 3; p++)
    for (uint32_t r = 0; r < h.sb64_rows; r++)
      for (uint32_t c = 0; c < h.sb64_cols; c++) {
        uint32_t k = p * n_sb + r * h.sb64_cols + c;
        if (!g->intra_cnt[k]) continue;
        int lv = 0;
        auto dep = [&](int rr, int cc) {
          if (rr < 0 || cc < 0) return;
          uint32_t d = p * n_sb + rr * h.sb64_cols + cc;
          if (g->intra_cnt[d]) lv = std::max(lv, g->intra_level[d] + 1);
        };
        dep((int)r, (int)c - 1);
        dep((int)r - 1, (int)c);
        dep((int)r - 1, (int)c - 1);
        g->intra_level[k] = lv;
        max_level = std::max(max_level, lv);
      }
  g->groups.clear();
  g->group_level_start.assign(max_level + 2, 0);
  for (uint32_t k = 0; k < 3 * n_sb; k++)
    if (g->intra_cnt[k]) g->group_level_start[g->intra_level[k] + 1]++;
  for (int l = 0; l <= max_level; l++) g->group_level_start[l + 1] += g->group_level_start[l];
  g->groups.resize(g->group_level_start[max_level + 1]);
  std::vector<uint32_t> fill(g->group_level_start.begin(), g->group_level_start.end());
  for (uint32_t k = 0; k < 3 * n_sb; k++) {
    if (!g->intra_cnt[k]) continue;
    vp9g::IntraGroup gr{};
    uint32_t sb = k % n_sb;
    gr.plane = (uint8_t)(k / n_sb);
    gr.sb_row = (uint16_t)(sb / h.sb64_cols);
    gr.sb_col = (uint16_t)(sb % h.sb64_cols);
    gr.first = g->intra_first[k];
    gr.count = g->intra_cnt[k];
    g->groups[fill[g->intra_level[k]]++] = gr;
  }
}

// Superblock items (kIbSuperblock): every intra block of a superblock in
// one item.  An item is levelled one above the deepest of the items left,
// above and above-left of it (an intra block reads samples of its own
// superblock and of those three only: the row above from above-left to its
// top-right, the column left of it, recon.rs:17-103).  Inside an item each
// transform block is levelled one above the deepest transform block of the
// item whose samples it reads (4x4 cells of each plane); inter samples are
// complete before intra starts.  The frame takes this form when it has at
// most half the levels of the per-block form (a key frame: 338 block
// levels against 46 superblock diagonals at 1080p); false keeps the
// per-block form, also for an item over kIbSbMaxLevels levels.  On success
// block_items, item_level_start and item_level (per intra block, for
// lf_need) describe the items.  The record's transform blocks were checked
// by the per-block form's pass (one size per plane and block, no repeated
// positions, inside the block).
static bool build_intra_sb(Builder *g, const vp9r_record *rec) {
  const vp9r_frame_hdr &h = rec->hdr;
  const int C = (int)h.sb64_cols, R = (int)h.sb64_rows;
  const size_t n_items = g->intra_items.size();
  if (!n_items) return false;
  std::vector<int32_t> &sb_level = g->sb_level;
  sb_level.assign((size_t)R * C, -1);
  auto sb_of = [&](const vp9r_block &b) { return (int)(b.mi_row / 8) * C + (int)(b.mi_col / 8); };
  for (size_t k = 0; k < n_items; k++) sb_level[sb_of(rec->blocks[g->intra_items[k].block])] = 0;
  int max_lv = -1;
  for (int r = 0; r < R; r++)
    for (int c = 0; c < C; c++) {
      int32_t &lv = sb_level[(size_t)r * C + c];
      if (lv < 0) continue;
      auto dep = [&](int rr, int cc) {
        if (rr >= 0 && cc >= 0 && sb_level[(size_t)rr * C + cc] >= 0) lv = std::max(lv, sb_level[(size_t)rr * C + cc] + 1);
      };
      dep(r, c - 1);
      dep(r - 1, c);
      dep(r - 1, c - 1);
      max_lv = std::max(max_lv, lv);
    }
  const size_t block_levels = g->item_level_start.empty() ? 0 : g->item_level_start.size() - 1;
  if ((size_t)(max_lv + 1) * 2 > block_levels) return false;
  // items in superblock level order (raster order within a level)
  std::vector<uint32_t> &order = g->sb_order;
  order.clear();
  std::vector<uint32_t> lstart(max_lv + 2, 0);
  for (int32_t lv : sb_level)
    if (lv >= 0) lstart[lv + 1]++;
  for (int l = 0; l <= max_lv; l++) lstart[l + 1] += lstart[l];
  order.resize(lstart[max_lv + 1]);
  {
    std::vector<uint32_t> fill(lstart.begin(), lstart.end());
    for (int k = 0; k < R * C; k++)
      if (sb_level[k] >= 0) order[fill[sb_level[k]]++] = (uint32_t)k;
  }
  // the intra blocks of each superblock (decode order keeps a superblock's
  // blocks together)
  std::vector<uint32_t> first_item((size_t)R * C, 0), n_item((size_t)R * C, 0);
  for (size_t k = 0; k < n_items; k++) {
    const int sb = sb_of(rec->blocks[g->intra_items[k].block]);
    if (!n_item[sb]) first_item[sb] = (uint32_t)k;
    if (first_item[sb] + n_item[sb] != k) return false;  // not contiguous: keep the per-block form
    n_item[sb]++;
  }
  const uint32_t M = vp9g::ib_max_txbs(h.ss_x, h.ss_y);
  g->sb_tx.clear();
  g->sb_units.clear();
  g->sb_levels.clear();
  std::vector<vp9g::IntraBlockItem> items(order.size());
  std::vector<int32_t> &cells = g->sb_cells;
  std::vector<vp9g::IntraSbTx> descs;
  std::vector<int32_t> &dlv = g->sb_tx_level;
  for (size_t oi = 0; oi < order.size(); oi++) {
    const int sb = (int)order[oi], sr = sb / C, sc = sb % C;
    // 4x4 cells of the three planes: luma 16 x 16, chroma (16 >> ss_x) x (16 >> ss_y)
    const int cw[3] = {16, 16 >> h.ss_x, 16 >> h.ss_x}, ch[3] = {16, 16 >> h.ss_y, 16 >> h.ss_y};
    const int coff[3] = {0, 256, 256 + cw[1] * ch[1]};
    cells.assign((size_t)coff[2] + cw[2] * ch[2], -1);
    descs.clear();
    dlv.clear();
    int n_lv = 0;
    for (uint32_t k = first_item[sb]; k < first_item[sb] + n_item[sb]; k++) {
      const vp9g::IntraItem &it = g->intra_items[k];
      const vp9r_block &b = rec->blocks[it.block];
      const int bsize = b.mi_size < BLOCK_8X8 ? BLOCK_8X8 : b.mi_size;
      for (uint32_t i = it.first; i < it.first + it.count; i++) {
        const vp9r_txb &t = rec->txbs[i];
        const int p = t.plane, sx = p ? h.ss_x : 0, sy = p ? h.ss_y : 0;
        const int log2 = t.tx_size + 2, size = 1 << log2;
        const int x0 = (sc * 64) >> sx, y0 = (sr * 64) >> sy;
        const int x = (int)t.x - x0, y = (int)t.y - y0;
        if (x < 0 || y < 0 || (x & 3) || (y & 3) || x + size > (64 >> sx) || y + size > (64 >> sy)) return false;
        const int bx = (b.mi_col * 8) >> sx, by = (b.mi_row * 8) >> sy;
        const int xi = ((int)t.x - bx) >> 2, yi = ((int)t.y - by) >> 2, step4 = 1 << t.tx_size;
        const int n4w = vp9g::num4x4_wide(bsize) >> sx;
        int mode;
        if (p > 0)
          mode = b.uv_mode;
        else if (b.mi_size >= BLOCK_8X8)
          mode = b.y_mode;
        else
          mode = b.sub_modes[((yi / step4) * ((n4w + step4 - 1) / step4) + xi / step4) & 3];
        const bool have_left = b.avail_l || xi > 0, have_above = b.avail_u || yi > 0;
        const bool a4 = log2 == 2 && xi + step4 < n4w;
        const int max_x = (int)((h.mi_cols * 8) >> sx) - 1, max_y = (int)((h.mi_rows * 8) >> sy) - 1;
        // the level: one above the deepest cell of this superblock read
        int lv = 0;
        auto dep = [&](int px, int py) {  // plane position relative to the superblock
          if (px < 0 || py < 0) return;
          const int l = cells[coff[p] + (py >> 2) * cw[p] + (px >> 2)];
          if (l >= 0) lv = std::max(lv, l + 1);
        };
        if (have_above) {
          const int n = a4 ? 2 * size : size;
          for (int k2 = 0; k2 < n; k2++) dep(std::min(max_x - x0, x + k2), y - 1);
          if (have_left) dep(x - 1, y - 1);
        }
        if (have_left)
          for (int k2 = 0; k2 < size; k2++) dep(x - 1, std::min(max_y - y0, y + k2));
        for (int cy = y >> 2; cy < (y + size) >> 2; cy++)
          for (int cx = x >> 2; cx < (x + size) >> 2; cx++) {
            int32_t &cell = cells[coff[p] + cy * cw[p] + cx];
            if (cell >= 0) return false;  // repeated position
            cell = lv;
          }
        vp9g::IntraSbTx d{};
        d.plane = (uint8_t)p;
        d.log2 = (uint8_t)log2;
        d.mode = (uint8_t)mode;
        d.flags = (uint8_t)((have_above ? 1 : 0) | (have_left ? 2 : 0) | (a4 ? 4 : 0));
        d.x = (uint8_t)x;
        d.y = (uint8_t)y;
        d.lim_a = (int16_t)std::max(-32768, std::min(32767, max_x - (int)t.x));
        d.lim_l = (int16_t)std::max(-32768, std::min(32767, max_y - (int)t.y));
        descs.push_back(d);
        dlv.push_back(lv);
        n_lv = std::max(n_lv, lv + 1);
      }
    }
    if (descs.size() > M || (uint32_t)n_lv > vp9g::kIbSbMaxLevels) return false;
    // descriptors by level (stable), their units, the level table
    std::vector<uint32_t> ls(n_lv + 1, 0);
    for (int32_t l : dlv) ls[l + 1]++;
    for (int l = 0; l < n_lv; l++) ls[l + 1] += ls[l];
    std::vector<uint32_t> pos(descs.size());
    {
      std::vector<uint32_t> fill(ls.begin(), ls.end());
      for (size_t j = 0; j < descs.size(); j++) pos[fill[dlv[j]]++] = (uint32_t)j;
    }
    vp9g::IntraBlockItem &o = items[oi];
    o = vp9g::IntraBlockItem{};
    o.first = (uint32_t)g->sb_tx.size();
    o.count = (uint32_t)descs.size();
    o.coef_first = (uint32_t)g->sb_units.size();
    o.sub_modes = (uint32_t)g->sb_levels.size();
    o.n_u = (uint16_t)n_lv;
    o.mi_row = (uint16_t)(sr * 8);
    o.mi_col = (uint16_t)(sc * 8);
    o.mi_size = BLOCK_64X64;
    o.flags = vp9g::kIbSuperblock;
    uint32_t chunks = 0;
    for (int l = 0; l < n_lv; l++) {
      bool dc = false;  // a DC_PRED block of 16x16 or larger: a DC pass first
      const uint32_t c0 = chunks;
      for (uint32_t q = ls[l]; q < ls[l + 1]; q++) {
        const vp9g::IntraSbTx &d = descs[pos[q]];
        g->sb_tx.push_back(d);
        dc = dc || (d.mode == DC_PRED && d.log2 >= 4);
        // chunks of at most four 4x4 units of one transform block
        const uint32_t n4 = 1u << (2 * (d.log2 - 2));
        for (uint32_t sub = 0; sub < n4; sub += 4, chunks++)
          g->sb_units.push_back(q | sub << 10 | std::min(4u, n4 - sub) << 18);
      }
      g->sb_levels.push_back(make_uint2(ls[l] | (ls[l + 1] - ls[l]) << 16 | (dc ? 1u << 31 : 0u),
                                        c0 | (chunks - c0) << 16));
    }
    if (chunks > M) return false;
    o.coef_count = chunks;
  }
  g->block_items = items;
  g->item_level_start = lstart;
  for (size_t k = 0; k < n_items; k++) g->item_level[k] = sb_level[sb_of(rec->blocks[g->intra_items[k].block])];
  return true;
}

// The loop filter's wavefront levels (lf.hip): level t holds the units
// (r, t - r) and, while t + 1 < rows, the lone vertical pass of superblock
// (t + 1, 0); lf_level_start[t] = 3 x the items of the levels before t (one
// workgroup per item and plane).
static void build_lf_levels(Builder *g, uint32_t rows, uint32_t cols) {
  if (g->lf_rows == rows && g->lf_cols == cols) return;
  g->lf_rows = rows;
  g->lf_cols = cols;
  g->lf_level_start.clear();
  const int R = (int)rows, C = (int)cols, T = R + C - 1;
  uint32_t n = 0;
  for (int t = 0; t < T; t++) {
    g->lf_level_start.push_back(n);
    const int units = std::min(R - 1, t) - std::max(0, t - C + 1) + 1;
    n += 3 * (uint32_t)(units + (t + 1 < R ? 1 : 0));
  }
  g->lf_level_start.push_back(n);
}

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// The algorithmic HBM bytes of each stage of a frame (what the roofline
// divides by time), with b = bytes per sample:
//   inter prediction, per prediction unit (block plane, or 4x4 sub-block
//     below 8x8) and reference: the filter footprint read,
//     (w + 7[x fractional]) x (h + 7[y fractional]) x b unscaled
//     (src/dsp/inter.rs:248-281) or the scaled footprint (inter.rs:354-359),
//     plus w x h x b written once;
//   inter residual, per transform block with eob > 0: n^2 x 4 coefficient
//     bytes plus the prediction read and written, 2 x n^2 x b;
//   intra, per transform block: the edges read, (3n + 1) x b, the block
//     written, n^2 x b, and n^2 x 4 coefficient bytes when eob > 0;
//   loop filter: every visible sample read and written once, 2 x frame bytes.
static void work_bytes(const vp9r_record *rec, Prepared &p) {
  const vp9r_frame_hdr &h = rec->hdr;
  const uint64_t b = h.bit_depth > 8 ? 2 : 1;
  for (auto &v : p.work_bytes) v = 0;
  for (uint32_t bi = 0; bi < rec->n_blocks; bi++) {
    const vp9r_block &blk = rec->blocks[bi];
    if (!blk.is_inter) continue;
    const int bsize = blk.mi_size < BLOCK_8X8 ? BLOCK_8X8 : blk.mi_size;
    const int nref = 1 + (blk.ref_frame[1] > INTRA_FRAME);
    for (int plane = 0; plane < 3; plane++) {
      int sx = plane ? h.ss_x : 0, sy = plane ? h.ss_y : 0;
      int psz = VP9_SS_SIZE_LOOKUP[bsize][sx][sy];
      int n4w = VP9_NUM_4X4_WIDE[psz], n4h = VP9_NUM_4X4_HIGH[psz];
      int base_x = (blk.mi_col * 8) >> sx, base_y = (blk.mi_row * 8) >> sy;
      auto unit = [&](int x, int y, int w, int hh, int sub) {
        for (int rl = 0; rl < nref; rl++) {
          vp9g::RefPos r = vp9g::mc_position(p.geo, h.ref_width, h.ref_height, blk, plane, x, y, sub, rl);
          uint64_t rw = (r.xs == 16 && (r.x & 15) == 0) ? w : ((((w - 1) * r.xs + (r.x & 15)) >> 4) + 8);
          uint64_t rh = (r.ys == 16 && (r.y & 15) == 0) ? hh : ((((hh - 1) * r.ys + (r.y & 15)) >> 4) + 8);
          p.work_bytes[0] += rw * rh * b;
        }
        p.work_bytes[0] += (uint64_t)w * hh * b;
      };
      if (blk.mi_size < BLOCK_8X8) {
        for (int y = 0; y < n4h; y++)
          for (int x = 0; x < n4w; x++) unit(base_x + 4 * x, base_y + 4 * y, 4, 4, y * n4w + x);
      } else {
        unit(base_x, base_y, n4w * 4, n4h * 4, 0);
      }
    }
  }
  for (uint32_t i = 0; i < rec->n_txbs; i++) {
    const vp9r_txb &t = rec->txbs[i];
    const uint64_t n = 4u << t.tx_size;
    if (rec->blocks[t.block].is_inter) {
      if (t.eob) p.work_bytes[1] += n * n * (4 + 2 * b);
    } else {
      p.work_bytes[2] += (3 * n + 1) * b + n * n * b + (t.eob ? n * n * 4 : 0);
    }
  }
  if (h.lf_level) {
    uint64_t cw = (h.width + h.ss_x) >> h.ss_x, ch = (h.height + h.ss_y) >> h.ss_y;
    p.work_bytes[3] = 2 * b * ((uint64_t)h.width * h.height + 2 * cw * ch);
  }
}

// Adds the stage times of every frame timed since the last fold to the
// running sums (waits for the last of them).
hipError_t fold_stage_times(vp9r_gpu *g) {
  if (g->ev_used == 0) return hipSuccess;
  hipError_t e = hipEventSynchronize(g->ev_sets[g->ev_used - 1][4]);
  for (size_t k = 0; k < g->ev_used && e == hipSuccess; k++) {
    const auto &ev = g->ev_sets[k];
    float ms = 0;
    for (int i = 0; i < 4 && e == hipSuccess; i++) {
      e = hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
      g->stage_acc[i] += ms;
    }
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[4]);
    g->stage_acc[4] += ms;
    g->frames_acc += g->ev_frames[k];
  }
  g->ev_used = 0;
  g->ev = nullptr;
  return e;
}

// Takes the event set of the next timed call, which covers `frames` frames.
hipError_t begin_timed_frame(vp9r_gpu *g, uint32_t frames) {
  if (g->ev_used == 1024) {
    hipError_t e = fold_stage_times(g);
    if (e != hipSuccess) return e;
  }
  if (g->ev_used == g->ev_sets.size()) {
    std::array<hipEvent_t, 5> set{};
    for (auto &ev : set) {
      hipError_t e = hipEventCreate(&ev);
      if (e != hipSuccess) return e;
    }
    g->ev_sets.push_back(set);
    g->ev_frames.push_back(0);
  }
  g->ev_frames[g->ev_used] = frames;
  g->ev = g->ev_sets[g->ev_used++].data();
  return hipSuccess;
}

// Orders work on `s` that writes frame f after f's download in flight, if any.
hipError_t order_after_copy(hipStream_t s, vp9r_gpu_frame *f) {
  if (!f->copy_pending) return hipSuccess;
  return hipStreamWaitEvent(s, f->copy_done, 0);
}

}  // namespace vp9r

using namespace vp9r;

extern "C" {

int vp9r_gpu_new(int device, vp9r_gpu **out) {
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= device || device < 0) {
    set_error("no usable GPU for the reconstruction backend");
    return VP9R_ERR_DEVICE;
  }
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  vp9r_gpu *g = new vp9r_gpu();
  g->device = device;
  // The one-launch schedule (device-side waits between wavefront levels)
  // hung on its first key frame and was removed; asking for it is an error.
  const char *wf = getenv("VP9R_WAVEFRONT");
  if (wf && strcmp(wf, "one") == 0) {
    delete g;
    set_error("VP9R_WAVEFRONT=one: the one-launch wavefront schedule was removed");
    return VP9R_ERR_UNSUPPORTED;
  }
  // VP9R_INTRA=groups: intra frames as (superblock, plane) groups
  const char *im = getenv("VP9R_INTRA");
  g->intra_blocks = !(im && strcmp(im, "groups") == 0);
  // VP9R_LF=levels: one loop filter level per launch throughout; =2 .. 5:
  // at most that many per launch
  const char *lm = getenv("VP9R_LF");
  if (lm && strcmp(lm, "levels") == 0)
    g->lf_depth = 1;
  else if (lm && lm[0] >= '2' && lm[0] <= '5' && !lm[1])
    g->lf_depth = lm[0] - '0';
  e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->copy_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->lf_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->up_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    if (g->stream) (void)hipStreamDestroy(g->stream);
    if (g->copy_stream) (void)hipStreamDestroy(g->copy_stream);
    if (g->lf_stream) (void)hipStreamDestroy(g->lf_stream);
    delete g;
    return hip_fail(e, "hipStreamCreate");
  }
  e = hipEventCreateWithFlags(&g->ev_upload, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_up, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_pre, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_lf_done, hipEventDisableTiming);
  for (int i = 0; i < kBatchSlots && e == hipSuccess; i++)
    e = hipEventCreateWithFlags(&g->slots[i].ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    vp9r_gpu_free(g);
    return hip_fail(e, "hipEventCreate");
  }
  *out = g;
  return VP9R_OK;
}

void vp9r_gpu_free(vp9r_gpu *g) {
  if (g) vp9r::gpu_release(g);
}

}  // extern "C"

void vp9r::gpu_retain(vp9r_gpu *g) { g->refs.fetch_add(1); }

hipError_t vp9r::seam_begin(vp9r_gpu *g) {
  g->seam_ms = -1.0f;
  for (hipEvent_t &e : g->seam_ev)
    if (!e) {
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
    }
  return hipEventRecord(g->seam_ev[0], g->stream);
}

hipError_t vp9r::seam_end(vp9r_gpu *g) { return hipEventElapsedTime(&g->seam_ms, g->seam_ev[0], g->seam_ev[1]); }

void vp9r::gpu_release(vp9r_gpu *g) {
  if (g->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(g->device);
  (void)hipStreamSynchronize(g->stream);
  (void)hipStreamSynchronize(g->copy_stream);
  if (g->lf_stream) (void)hipStreamSynchronize(g->lf_stream);
  if (g->up_stream) (void)hipStreamSynchronize(g->up_stream);
  if (g->ev_upload) (void)hipEventDestroy(g->ev_upload);
  for (hipEvent_t e : {g->ev_up, g->ev_pre, g->ev_lf_done})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : g->intra_events) (void)hipEventDestroy(e);
  for (auto &s : g->slots)
    if (s.ev) (void)hipEventDestroy(s.ev);
  for (auto &set : g->ev_sets)
    for (auto ev : set) (void)hipEventDestroy(ev);
  for (hipEvent_t e : g->seam_ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(g->stream);
  (void)hipStreamDestroy(g->copy_stream);
  if (g->lf_stream) (void)hipStreamDestroy(g->lf_stream);
  if (g->up_stream) (void)hipStreamDestroy(g->up_stream);
  delete g;
}

extern "C" {

int vp9r_gpu_frame_new(vp9r_gpu *g, uint32_t width, uint32_t height, uint32_t bd, uint32_t ss_x,
                       uint32_t ss_y, vp9r_gpu_frame **out) {
  *out = nullptr;
  if (width == 0 || height == 0 || width > 65536 || height > 65536 || (bd != 8 && bd != 10 && bd != 12) ||
      ss_x > 1 || ss_y > 1) {
    set_error("invalid frame geometry");
    return VP9R_ERR_INVALID_INPUT;
  }
  vp9r_gpu_frame *f = new vp9r_gpu_frame();
  f->g = g;
  f->width = width;
  f->height = height;
  f->bit_depth = bd;
  f->ss_x = ss_x;
  f->ss_y = ss_y;
  f->mi_cols = (width + 7) >> 3;
  f->mi_rows = (height + 7) >> 3;
  f->sb64_cols = (f->mi_cols + 7) >> 3;
  f->sb64_rows = (f->mi_rows + 7) >> 3;
  size_t bps = bd > 8 ? 2 : 1, off[3], total = 0;
  for (int p = 0; p < 3; p++) {
    f->planes.stride[p] = (f->sb64_cols * 64) >> (p ? ss_x : 0);
    f->planes.rows[p] = (f->sb64_rows * 64) >> (p ? ss_y : 0);
    off[p] = total;
    total = align256(total + (size_t)f->planes.stride[p] * f->planes.rows[p] * bps);
  }
  hipError_t e = hipMalloc(&f->base, total);
  if (e != hipSuccess) {
    delete f;
    return hip_fail(e, "hipMalloc(frame)");
  }
  e = hipMemsetAsync(f->base, 0, total, g->stream);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&f->ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&f->copy_done, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(g->stream);
    if (f->ready) (void)hipEventDestroy(f->ready);
    (void)hipFree(f->base);
    delete f;
    return hip_fail(e, "frame set-up");
  }
  f->bytes = total;
  for (int p = 0; p < 3; p++) f->planes.p[p] = (uint8_t *)f->base + off[p];
  gpu_retain(g);
  *out = f;
  return VP9R_OK;
}

void vp9r_gpu_frame_free(vp9r_gpu_frame *f) {
  if (!f) return;
  if (f->g) {
    (void)hipStreamSynchronize(f->g->stream);
    (void)hipStreamSynchronize(f->g->copy_stream);
  }
  (void)hipEventDestroy(f->ready);
  (void)hipEventDestroy(f->copy_done);
  (void)hipFree(f->base);
  if (f->g) gpu_release(f->g);
  delete f;
}

void vp9r_gpu_frame_geometry(const vp9r_gpu_frame *f, int plane, uint32_t *stride, uint32_t *rows) {
  *stride = f->planes.stride[plane];
  *rows = f->planes.rows[plane];
}

int vp9r_gpu_frame_upload_u16(vp9r_gpu_frame *f, int plane, const uint16_t *src) {
  size_t n = (size_t)f->planes.stride[plane] * f->planes.rows[plane];
  hipStream_t s = f->g->stream;
  HIP_TRY(vp9r::order_after_copy(s, f), "hipStreamWaitEvent");
  if (f->bit_depth > 8) {
    HIP_TRY(hipMemcpyAsync(f->planes.p[plane], src, n * 2, hipMemcpyHostToDevice, s), "upload");
  } else {
    std::vector<uint8_t> tmp(n);
    for (size_t i = 0; i < n; i++) tmp[i] = (uint8_t)src[i];
    HIP_TRY(hipMemcpyAsync(f->planes.p[plane], tmp.data(), n, hipMemcpyHostToDevice, s), "upload");
    HIP_TRY(hipStreamSynchronize(s), "upload sync");
  }
  HIP_TRY(hipStreamSynchronize(s), "upload sync");
  return VP9R_OK;
}

int vp9r_gpu_frame_download_u16(vp9r_gpu_frame *f, int plane, uint16_t *dst) {
  size_t n = (size_t)f->planes.stride[plane] * f->planes.rows[plane];
  hipStream_t s = f->g->stream;
  if (f->bit_depth > 8) {
    HIP_TRY(hipMemcpyAsync(dst, f->planes.p[plane], n * 2, hipMemcpyDeviceToHost, s), "download");
    HIP_TRY(hipStreamSynchronize(s), "download sync");
  } else {
    std::vector<uint8_t> tmp(n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), f->planes.p[plane], n, hipMemcpyDeviceToHost, s), "download");
    HIP_TRY(hipStreamSynchronize(s), "download sync");
    for (size_t i = 0; i < n; i++) dst[i] = tmp[i];
  }
  return VP9R_OK;
}

int vp9r_gpu_synchronize(vp9r_gpu *g) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
  return VP9R_OK;
}

void vp9r_gpu_set_timing(vp9r_gpu *g, int enable) {
  g->timing = enable != 0;
  g->ev_used = 0;
  g->ev = nullptr;
  g->frames_acc = 0;
  for (double &v : g->stage_acc) v = 0;
}

int vp9r_gpu_stage_times(vp9r_gpu *g, float out[5], uint32_t *frames) {
  if (!g->timing) {
    set_error("timing is not enabled");
    return VP9R_ERR_INVALID_INPUT;
  }
  HIP_TRY(vp9r::fold_stage_times(g), "stage timing");
  for (int i = 0; i < 5; i++) out[i] = (float)g->stage_acc[i];
  if (frames) *frames = g->frames_acc;
  g->frames_acc = 0;
  for (double &v : g->stage_acc) v = 0;
  return VP9R_OK;
}

}  // extern "C"

namespace vp9r {

// Checks the record against the frame geometry and the reference sizes it
// names, builds the work lists, and writes everything to `stage` (host) at
// the offsets it records in p.  Every index the kernels follow is checked
// here, so that no record can send a kernel outside its frame stores.
int build_record(Builder &w, const vp9r_record *rec, Prepared &p, HostBuf &stage) {
  Builder *g = &w;
  const vp9r_frame_hdr &h = rec->hdr;
  if (h.show_existing_frame) {
    set_error("a show_existing_frame record has nothing to reconstruct");
    return VP9R_ERR_INVALID_INPUT;
  }
  uint32_t mi_cols = (h.width + 7) >> 3, mi_rows = (h.height + 7) >> 3;
  if (h.width == 0 || h.height == 0 || h.mi_cols != mi_cols || h.mi_rows != mi_rows ||
      h.sb64_cols != ((mi_cols + 7) >> 3) || h.sb64_rows != ((mi_rows + 7) >> 3) ||
      (h.bit_depth != 8 && h.bit_depth != 10 && h.bit_depth != 12) || h.ss_x > 1 || h.ss_y > 1) {
    set_error("malformed record header");
    return VP9R_ERR_INVALID_INPUT;
  }
  p.hdr = h;
  p.geo = geom_of(h);
  q_table(h, p.q);
  lf_params(h, p.lfp);
  const uint32_t plane_w[3] = {h.sb64_cols * 64, (h.sb64_cols * 64) >> h.ss_x, (h.sb64_cols * 64) >> h.ss_x};
  const uint32_t plane_h[3] = {h.sb64_rows * 64, (h.sb64_rows * 64) >> h.ss_y, (h.sb64_rows * 64) >> h.ss_y};
  bool ref_ok[3] = {false, false, false};
  if (!h.frame_is_intra)
    for (int i = 0; i < 3; i++) {
      uint32_t rw = h.ref_width[i], rh = h.ref_height[i];
      ref_ok[i] = rw && rh && !(2 * h.width < rw || 2 * h.height < rh || h.width > 16 * rw || h.height > 16 * rh);
    }
  for (uint32_t bi = 0; bi < rec->n_blocks; bi++) {
    const vp9r_block &b = rec->blocks[bi];
    if (b.mi_size > BLOCK_64X64 || b.mi_row >= h.mi_rows || b.mi_col >= h.mi_cols ||
        (b.mi_row & (VP9_NUM_8X8_HIGH[b.mi_size] - 1)) || (b.mi_col & (VP9_NUM_8X8_WIDE[b.mi_size] - 1)) ||
        b.interp_filter > 3 || b.uv_mode > TM_PRED || b.y_mode > NEWMV || b.segment_id > 7) {
      set_error("malformed record block");
      return VP9R_ERR_INVALID_INPUT;
    }
    if (b.is_inter) {
      for (int l = 0; l < 1 + (b.ref_frame[1] > INTRA_FRAME); l++) {
        int rf = b.ref_frame[l];
        if (rf < LAST_FRAME || rf > ALTREF_FRAME || h.frame_is_intra) {
          set_error("malformed record block reference");
          return VP9R_ERR_INVALID_INPUT;
        }
        if (!ref_ok[rf - 1]) {
          set_error("reference size outside the 2x-down / 16x-up scaling range");
          return VP9R_ERR_BITSTREAM;
        }
      }
    } else {
      for (int k = 0; k < 4; k++)
        if (b.sub_modes[k] > TM_PRED) {
          set_error("malformed record block");
          return VP9R_ERR_INVALID_INPUT;
        }
    }
  }
  build_mc_tiles(g, rec);
  for (auto &j : g->jobs) j.clear();
  for (auto &j : g->ijobs) j.clear();
  for (uint32_t i = 0; i < rec->n_txbs; i++) {
    const vp9r_txb &t = rec->txbs[i];
    const uint32_t n0 = 4u << t.tx_size;
    bool ok = t.tx_size <= 3 && t.plane < 3 && t.tx_type <= 3 && t.block < rec->n_blocks &&
              (uint64_t)t.coef_start + t.ncoef <= rec->n_coefs && t.eob <= n0 * n0 && (t.x & (n0 - 1)) == 0 &&
              (t.y & (n0 - 1)) == 0 && t.x + n0 <= plane_w[t.plane] && t.y + n0 <= plane_h[t.plane];
    if (ok) {
      const vp9r_block &b = rec->blocks[t.block];
      int sx = t.plane ? h.ss_x : 0, sy = t.plane ? h.ss_y : 0;
      uint32_t bx = (b.mi_col * 8u) >> sx, by = (b.mi_row * 8u) >> sy;
      const int psz = VP9_SS_SIZE_LOOKUP[b.mi_size < BLOCK_8X8 ? BLOCK_8X8 : b.mi_size][sx][sy];
      ok = psz != BLOCK_INVALID && t.x >= bx && t.y >= by &&
           t.x + n0 <= bx + 4u * VP9_NUM_4X4_WIDE[psz] && t.y + n0 <= by + 4u * VP9_NUM_4X4_HIGH[psz] &&
           (!h.lossless || t.tx_size == 0);
    }
    if (!ok) {
      set_error("malformed record transform block");
      return VP9R_ERR_INVALID_INPUT;
    }
    if (!rec->blocks[t.block].is_inter)
      g->ijobs[t.tx_size].push_back(i | vp9g::kItxIntra);
    else if (t.eob > 0)
      g->jobs[t.tx_size].push_back(i);
  }
  build_intra(g, rec);
  {
    // each intra block once, its planes in order, blocks in decode order
    uint32_t prev_block = 0;
    for (size_t k = 0; k < g->intra_items.size(); k++) {
      const vp9g::IntraItem &it = g->intra_items[k];
      if (k && it.block <= prev_block) {
        set_error("malformed record: intra transform blocks out of decode order");
        return VP9R_ERR_INVALID_INPUT;
      }
      prev_block = it.block;
      for (uint32_t i = it.first + 1; i < it.first + it.count; i++)
        if (rec->txbs[i].plane < rec->txbs[i - 1].plane) {
          set_error("malformed record: transform block planes out of order");
          return VP9R_ERR_INVALID_INPUT;
        }
    }
    // the kernel's view of each block (by level): one transform size per
    // plane, each grid position of a plane's transform blocks at most once
    g->block_items.resize(g->items_sorted.size());
    uint8_t seen[3][256];
    for (size_t k = 0; k < g->items_sorted.size(); k++) {
      const vp9g::IntraItem &it = g->items_sorted[k];
      const vp9r_block &b = rec->blocks[it.block];
      vp9g::IntraBlockItem &o = g->block_items[k];
      o = vp9g::IntraBlockItem{};
      o.first = it.first;
      o.count = it.count;
      const vp9r_txb &t0 = rec->txbs[it.first], &t1 = rec->txbs[it.first + it.count - 1];
      o.coef_first = t0.coef_start;
      o.coef_count = t1.coef_start + t1.ncoef > t0.coef_start ? t1.coef_start + t1.ncoef - t0.coef_start : 0;
      o.mi_row = b.mi_row;
      o.mi_col = b.mi_col;
      o.mi_size = b.mi_size;
      o.y_mode = b.y_mode;
      o.uv_mode = b.uv_mode;
      o.flags = (uint8_t)((b.segment_id & 7) | (b.avail_u ? 8 : 0) | (b.avail_l ? 16 : 0));
      memcpy(&o.sub_modes, b.sub_modes, 4);
      int tx[3] = {-1, -1, -1};
      uint32_t n[3] = {0, 0, 0};
      memset(seen, 0, sizeof seen);
      for (uint32_t i = it.first; i < it.first + it.count; i++) {
        const vp9r_txb &t = rec->txbs[i];
        const int sx = t.plane ? h.ss_x : 0, sy = t.plane ? h.ss_y : 0;
        const uint32_t bx = (b.mi_col * 8u) >> sx, by = (b.mi_row * 8u) >> sy;
        const uint32_t cell = ((t.y - by) >> (2 + t.tx_size)) * 16 + ((t.x - bx) >> (2 + t.tx_size));
        if ((tx[t.plane] >= 0 && tx[t.plane] != t.tx_size) || cell >= 256 || seen[t.plane][cell]) {
          set_error("malformed record: intra transform blocks of mixed sizes or repeated positions");
          return VP9R_ERR_INVALID_INPUT;
        }
        tx[t.plane] = t.tx_size;
        seen[t.plane][cell] = 1;
        n[t.plane]++;
      }
      if (it.count > vp9g::ib_max_txbs(h.ss_x, h.ss_y)) {
        set_error("malformed record: too many transform blocks in an intra block");
        return VP9R_ERR_INVALID_INPUT;
      }
      if (n[1] && n[2] && tx[1] != tx[2]) {
        set_error("malformed record: intra chroma transform blocks of two sizes");
        return VP9R_ERR_INVALID_INPUT;
      }
      o.n_y = (uint16_t)n[0];
      o.n_u = (uint16_t)n[1];
    }
  }
  // a frame with many more intra levels than superblock diagonals (a key
  // frame) takes superblock items
  g->sb_mode = build_intra_sb(g, rec);
  p.intra_sb = g->sb_mode;
  work_bytes(rec, p);
  const bool do_lf = h.lf_level != 0;
  if (do_lf) build_lf_levels(g, h.sb64_rows, h.sb64_cols);
  const size_t n_mi = (size_t)h.mi_rows * h.mi_cols;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  p.o_blocks = take(sizeof(vp9r_block) * rec->n_blocks);
  p.o_txbs = take(sizeof(vp9r_txb) * rec->n_txbs);
  p.o_cpos = take(sizeof(uint16_t) * rec->n_coefs);
  p.o_cval = take(sizeof(int32_t) * rec->n_coefs);
  p.o_mi = take(sizeof(vp9r_mi) * n_mi);
  p.o_tiles = take(sizeof(vp9g::McTile) * g->tiles.size());
  for (int k = 0; k < 4; k++) p.o_jobs[k] = take(sizeof(uint32_t) * (g->jobs[k].size() + g->ijobs[k].size()));
  p.o_iidx = take(sizeof(uint32_t) * g->intra_index.size());
  p.o_groups = take(sizeof(vp9g::IntraGroup) * g->groups.size());
  p.o_items = take(sizeof(vp9g::IntraBlockItem) * g->block_items.size());
  const size_t n_sbtx = g->sb_mode ? g->sb_tx.size() : 0, n_sbunits = g->sb_mode ? g->sb_units.size() : 0,
               n_sblevels = g->sb_mode ? g->sb_levels.size() : 0;
  p.o_sbtx = take(sizeof(vp9g::IntraSbTx) * n_sbtx);
  p.o_sbunits = take(sizeof(uint32_t) * n_sbunits);
  p.o_sblevels = take(sizeof(uint2) * n_sblevels);
  p.bytes = off;
  p.n_tiles = (uint32_t)g->tiles.size();
  for (int k = 0; k < 4; k++) {
    p.n_jobs[k] = (uint32_t)g->jobs[k].size();
    p.n_ijobs[k] = (uint32_t)g->ijobs[k].size();
  }
  p.n_groups = (uint32_t)g->groups.size();
  p.n_items = (uint32_t)g->intra_items.size();
  p.item_level_start = g->item_level_start;
  p.group_level_start = g->group_level_start;
  if (do_lf)
    p.lf_level_start = g->lf_level_start;
  else
    p.lf_level_start.clear();
  // The loop filter's passes on superblock (r, c) read and write samples in
  // mi rows 8r - 2 .. 8r + 7 and columns 8c - 2 .. 8c + 7 (the chroma halo
  // is 8 samples at half resolution).  An intra block writes its own area
  // and reads the row above it (above-right reads stay inside the block's
  // columns, recon.rs:61-69) and the column left of it.  Since the loop filter follows
  // the whole frame's reconstruction (decoder/mod.rs:355-358), a level may
  // run once every intra block meeting one of its superblocks is done:
  // lf_need[t] is the highest intra level among them, over the levels <= t.
  p.lf_need.clear();
  if (do_lf) {
    const int R = (int)h.sb64_rows, C = (int)h.sb64_cols;
    g->sb_need.assign((size_t)R * C, -1);
    auto meet = [&](int row0, int row1, int col0, int col1, int32_t level) {
      const int r_lo = std::max(0, row0) / 8, r_hi = std::min(R - 1, (row1 + 2) / 8);
      const int c_lo = std::max(0, col0) / 8, c_hi = std::min(C - 1, (col1 + 2) / 8);
      for (int r = r_lo; r <= r_hi; r++)
        for (int c = c_lo; c <= c_hi; c++) {
          int32_t &v = g->sb_need[(size_t)r * C + c];
          v = std::max(v, level);
        }
    };
    for (size_t k = 0; k < g->intra_items.size(); k++) {
      const vp9r_block &b = rec->blocks[g->intra_items[k].block];
      const int bw8 = b.mi_size < BLOCK_8X8 ? 1 : VP9_NUM_8X8_WIDE[b.mi_size];
      const int bh8 = b.mi_size < BLOCK_8X8 ? 1 : VP9_NUM_8X8_HIGH[b.mi_size];
      const int row = (int)b.mi_row, col = (int)b.mi_col;
      meet(row - 1, row + bh8 - 1, col - 1, col + bw8 - 1, g->item_level[k]);
    }
    auto need = [&](int r, int c) { return g->sb_need[(size_t)r * C + c]; };
    const int T = R + C - 1;
    p.lf_need.assign(T, -1);
    int32_t run = -1;
    for (int t = 0; t < T; t++) {
      for (int r = std::max(0, t - C + 1); r <= std::min(R - 1, t); r++) {
        const int c = t - r;
        run = std::max(run, need(r, c));
        if (c + 1 < C) run = std::max(run, need(r, c + 1));
      }
      if (t + 1 < R) run = std::max(run, need(t + 1, 0));
      p.lf_need[t] = run;
    }
  }
  if (stage.ensure(p.bytes) != hipSuccess) {
    set_error("hipHostMalloc(staging) failed");
    return VP9R_ERR_DEVICE;
  }
  uint8_t *st = (uint8_t *)stage.p;
  auto put = [&](size_t o, const void *src, size_t bytes) {
    if (bytes) memcpy(st + o, src, bytes);
  };
  put(p.o_blocks, rec->blocks, sizeof(vp9r_block) * rec->n_blocks);
  put(p.o_txbs, rec->txbs, sizeof(vp9r_txb) * rec->n_txbs);
  put(p.o_cpos, rec->coef_pos, sizeof(uint16_t) * rec->n_coefs);
  put(p.o_cval, rec->coef_val, sizeof(int32_t) * rec->n_coefs);
  put(p.o_mi, rec->mi, sizeof(vp9r_mi) * n_mi);
  put(p.o_tiles, g->tiles.data(), sizeof(vp9g::McTile) * g->tiles.size());
  for (int k = 0; k < 4; k++) {
    put(p.o_jobs[k], g->jobs[k].data(), sizeof(uint32_t) * g->jobs[k].size());
    put(p.o_jobs[k] + sizeof(uint32_t) * g->jobs[k].size(), g->ijobs[k].data(),
        sizeof(uint32_t) * g->ijobs[k].size());
  }
  put(p.o_iidx, g->intra_index.data(), sizeof(uint32_t) * g->intra_index.size());
  put(p.o_groups, g->groups.data(), sizeof(vp9g::IntraGroup) * g->groups.size());
  put(p.o_items, g->block_items.data(), sizeof(vp9g::IntraBlockItem) * g->block_items.size());
  put(p.o_sbtx, g->sb_tx.data(), sizeof(vp9g::IntraSbTx) * n_sbtx);
  put(p.o_sbunits, g->sb_units.data(), sizeof(uint32_t) * n_sbunits);
  put(p.o_sblevels, g->sb_levels.data(), sizeof(uint2) * n_sblevels);
  return VP9R_OK;
}

static int check_cur(const Prepared &p, const vp9r_gpu_frame *cur) {
  const vp9r_frame_hdr &h = p.hdr;
  if (!cur || cur->width != h.width || cur->height != h.height || cur->bit_depth != h.bit_depth ||
      cur->ss_x != h.ss_x || cur->ss_y != h.ss_y) {
    set_error("the frame store does not match the record's geometry");
    return VP9R_ERR_INVALID_INPUT;
  }
  return VP9R_OK;
}

static int check_frames(const Prepared &p, vp9r_gpu_frame *const refs[3], vp9r_gpu_frame *cur) {
  const vp9r_frame_hdr &h = p.hdr;
  if (int st = check_cur(p, cur)) return st;
  if (!h.frame_is_intra)
    for (int i = 0; i < 3; i++) {
      const vp9r_gpu_frame *r = refs ? refs[i] : nullptr;
      if (!r || r->bit_depth != h.bit_depth || r->ss_x != h.ss_x || r->ss_y != h.ss_y) {
        set_error("missing reference, or a reference of another format");
        return VP9R_ERR_BITSTREAM;
      }
      if (r->width != h.ref_width[i] || r->height != h.ref_height[i]) {
        set_error("a reference's size is not the one the record names");
        return VP9R_ERR_INVALID_INPUT;
      }
      if (r == cur) {
        set_error("the frame being reconstructed is one of its references");
        return VP9R_ERR_INVALID_INPUT;
      }
    }
  return VP9R_OK;
}

// ---- batched launches
//
// Per bit-depth class (8-bit and above share no kernel instance), the
// frames' arguments and the ranges of every launch are staged in one pinned
// buffer, copied to a device slot with one transfer, and the launches read
// them there.  Slots form a ring; a slot is reused once its last launch
// has finished (its event).

namespace {

// Superblocks in a batch from which the loop filter decisions take one
// workgroup per superblock (four 1080p frames, or one 4K frame: 15.9
// against 18.7 us; 64 1080p frames: 201 against 295 us): below it, a
// workgroup per superblock entry finishes sooner (one 1080p frame: 7.5
// against 9.8 us).
constexpr size_t kMaskPerSbMin = 2040;

template <class T>
size_t stage_take(std::vector<uint8_t> &buf, size_t n) {
  size_t o = (buf.size() + 255) & ~(size_t)255;
  buf.resize(o + sizeof(T) * n);
  return o;
}

// The ranges of one launch: frame i takes ceil(count_i / per_wg) workgroups.
struct LaunchTable {
  size_t off = 0;        // in the staging buffer
  uint32_t wgs = 0;
};

LaunchTable add_ranges(std::vector<uint8_t> &buf, int n, const uint32_t *first, const uint32_t *count,
                       uint32_t per_wg) {
  LaunchTable t;
  t.off = stage_take<vp9g::Range>(buf, n + 1);
  vp9g::Range *r = (vp9g::Range *)(buf.data() + t.off);
  uint32_t wg = 0;
  for (int i = 0; i < n; i++) {
    r[i].wg = wg;
    r[i].first = first[i];
    r[i].count = count[i];
    r[i]._pad = 0;
    wg += (count[i] + per_wg - 1) / per_wg;
  }
  r[n].wg = wg;
  r[n].first = r[n].count = r[n]._pad = 0;
  t.wgs = wg;
  return t;
}

}  // namespace

// The fused wavefront schedule of one frame: its intra levels (launches 0
// .. n_int - 1), and per loop filter level j the launch lf_at[j]: after
// level j - 1 and after the intra levels it needs (lf_need).  Returns the
// number of launches.
size_t wave_schedule(const Prepared &p, bool recon, bool lf, size_t &n_int, std::vector<int> &lf_at) {
  n_int = recon && p.item_level_start.size() > 1 ? p.item_level_start.size() - 1 : 0;
  const size_t n_lf = lf && p.lf_level_start.size() > 1 ? p.lf_level_start.size() - 1 : 0;
  lf_at.clear();
  int prev = -1;
  for (size_t j = 0; j < n_lf; j++) {
    const int need = recon && j < p.lf_need.size() ? p.lf_need[j] : -1;
    prev = std::max(prev + 1, need + 1);
    lf_at.push_back(prev);
  }
  return std::max(n_int, n_lf ? (size_t)lf_at.back() + 1 : 0);
}

// The m loop filter levels of a single frame after its intra levels: the
// first `singles` one per launch in place (k_lf), then launches of `depths`
// levels each (k_lf_cone), an even number of them: they alternate between
// the frame store and a second copy of its planes, and the last must write
// the frame store.  The split minimises the launches' summed time, taken
// from 1080p measurements (microseconds from one launch's start to the
// next's, by depth 1 .. 5); max_depth 1 keeps one level per launch.
LfPlan lf_tail_plan(size_t m, int max_depth) {
  static const double cost[6] = {0.0, 10.8, 15.6, 19.0, 23.3, 28.1};
  LfPlan best;
  best.singles = m;
  if (max_depth < 2 || m < 4) return best;
  // f[x][p]: least time of x levels in cone launches whose count has parity
  // p, and the depth of the first of them
  const double inf = 1e30;
  std::vector<std::array<double, 2>> f(m + 1, {inf, inf});
  std::vector<std::array<int, 2>> first(m + 1, {0, 0});
  f[0][0] = 0.0;
  for (size_t x = 2; x <= m; x++)
    for (int p = 0; p < 2; p++)
      for (int d = 2; d <= max_depth && (size_t)d <= x; d++)
        if (f[x - d][p ^ 1] + cost[d] < f[x][p]) {
          f[x][p] = f[x - d][p ^ 1] + cost[d];
          first[x][p] = d;
        }
  double t_best = cost[1] * (double)m;
  for (size_t sgl = 0; sgl < m; sgl++) {
    const double t = cost[1] * (double)sgl + f[m - sgl][0];
    if (t < t_best) {
      t_best = t;
      best.singles = sgl;
    }
  }
  best.depths.clear();
  for (size_t x = m - best.singles, p = 0; x > 0;) {
    const int d = first[x][p];
    best.depths.push_back(d);
    x -= (size_t)d;
    p ^= 1;
  }
  return best;
}

int launch_batch(vp9r_gpu *g, const BatchFrame *fr, int n_all, bool recon, bool lf, GraphTarget *gt) {
  hipStream_t s = g->stream;
  if (gt && (n_all != 1 || g->timing || !g->intra_blocks)) {
    set_error("graph capture: one frame, stage timing off, leveled intra blocks");
    return VP9R_ERR_INVALID_INPUT;
  }
  for (int cls = 0; cls < 2; cls++) {
    const bool hbd = cls == 1;
    std::vector<const BatchFrame *> f;
    for (int i = 0; i < n_all; i++)
      if ((fr[i].p->hdr.bit_depth > 8) == hbd) f.push_back(&fr[i]);
    const int n = (int)f.size();
    if (!n) continue;
    std::vector<uint8_t> &st = g->batch_stage;
    st.clear();
    const size_t o_clear = stage_take<vp9g::ClearArgs>(st, n), o_mc = stage_take<vp9g::McArgs>(st, n),
                 o_itx = stage_take<vp9g::ItxArgs>(st, n), o_intra = stage_take<vp9g::IntraArgs>(st, n),
                 o_lf = stage_take<vp9g::LfArgs>(st, n);
    // LF decisions: per frame n_sb * 6 * 256 packed words
    size_t mask_bytes = 0;
    std::vector<size_t> mask_off(n);
    for (int i = 0; i < n; i++) {
      mask_off[i] = mask_bytes;
      mask_bytes += (size_t)f[i]->p->hdr.sb64_rows * f[i]->p->hdr.sb64_cols * 6 * 256 * 4;
    }
    if (lf && mask_bytes > g->lf_mask.cap) {
      HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
      HIP_TRY(hipStreamSynchronize(g->lf_stream), "hipStreamSynchronize");
      HIP_TRY(g->lf_mask.ensure(mask_bytes), "hipMalloc(lf masks)");
    }
    // intra residual planes (int16, the frame store's strides) of the
    // frames with intra blocks; read and written on the kernel stream only
    size_t res_bytes = 0;
    std::vector<size_t> res_off(3 * n, 0);
    if (recon && g->intra_blocks)
      for (int i = 0; i < n; i++) {
        if (!f[i]->p->n_items) continue;
        for (int pl = 0; pl < 3; pl++) {
          res_off[3 * i + pl] = res_bytes;
          res_bytes = align256(res_bytes + (size_t)f[i]->cur->planes.stride[pl] * f[i]->cur->planes.rows[pl] * 2);
        }
      }
    if (res_bytes > g->intra_res.cap) {
      HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
      HIP_TRY(g->intra_res.ensure(res_bytes), "hipMalloc(intra residuals)");
    }
    std::vector<uint32_t> first(n), count(n);
    auto table = [&](auto get_first, auto get_count, uint32_t per_wg) {
      for (int i = 0; i < n; i++) {
        first[i] = get_first(i);
        count[i] = get_count(i);
      }
      return add_ranges(st, n, first.data(), count.data(), per_wg);
    };
    LaunchTable t_clear, t_mc, t_itx[4], t_mask;
    int mask_per_sb = 1;
    std::vector<LaunchTable> t_intra, t_lf;
    bool groups = false;
    const bool fused = g->intra_blocks;
    if (recon) {
      t_clear = table([](int) { return 0u; },
                      [&](int i) { return 3 * vp9g::clear_margin_rows(f[i]->cur->planes, f[i]->p->geo); }, 1);
      t_mc = table([](int) { return 0u; }, [&](int i) { return f[i]->p->n_tiles; }, vp9g::mc_tiles_per_wg(n));
      // intra: leveled blocks, whose residuals the inter residual launches
      // compute, or (superblock, plane) groups, which compute their own
      groups = !g->intra_blocks;
      for (int k = 0; k < 4; k++)
        t_itx[k] = table([](int) { return 0u; },
                         [&](int i) { return f[i]->p->n_jobs[k] + (groups ? 0u : f[i]->p->n_ijobs[k]); },
                         64u >> (2 + k));
      size_t levels = 0;
      for (int i = 0; i < n; i++) {
        const auto &ls = groups ? f[i]->p->group_level_start : f[i]->p->item_level_start;
        if (ls.size() > 1) levels = std::max(levels, ls.size() - 1);
      }
      if (fused) levels = 0;
      for (size_t l = 0; l < levels; l++) {
        t_intra.push_back(table(
            [&](int i) {
              const auto &ls = groups ? f[i]->p->group_level_start : f[i]->p->item_level_start;
              return l + 1 < ls.size() ? ls[l] : 0u;
            },
            [&](int i) {
              const auto &ls = groups ? f[i]->p->group_level_start : f[i]->p->item_level_start;
              return l + 1 < ls.size() ? ls[l + 1] - ls[l] : 0u;
            },
            1));
      }
    }
    // Fused wavefront launches (leveled intra blocks): launch k holds intra
    // level k of every frame and, per frame, the loop filter level placed at
    // k: level j goes after level j - 1 and after the intra levels it needs
    // (lf_need), all on the kernel stream.  Each filter work item is one
    // workgroup (three planes).
    std::vector<LaunchTable> tw_i, tw_l;
    std::vector<std::vector<int>> lf_lvl;  // [frame][launch]: its loop filter level there, or -1
    size_t n_wave = 0, n_wave_intra = 0;
    if (fused) {
      std::vector<size_t> n_int(n);
      std::vector<std::vector<int>> lf_at(n);
      for (int i = 0; i < n; i++) {
        n_wave = std::max(n_wave, wave_schedule(*f[i]->p, recon, lf, n_int[i], lf_at[i]));
        n_wave_intra = std::max(n_wave_intra, n_int[i]);
      }
      for (int i = 0; i < n; i++) {
        lf_lvl.emplace_back(n_wave, -1);
        for (size_t j = 0; j < lf_at[i].size(); j++) lf_lvl[i][lf_at[i][j]] = (int)j;
      }
      for (size_t k = 0; k < n_wave; k++) {
        tw_i.push_back(table(
            [&](int i) { return k < n_int[i] ? f[i]->p->item_level_start[k] : 0u; },
            [&](int i) {
              const auto &ls = f[i]->p->item_level_start;
              return k < n_int[i] ? ls[k + 1] - ls[k] : 0u;
            },
            1));
        // with intra work: one k_wave workgroup per item; after the last
        // intra level: k_lf, one workgroup per item and plane
        const bool with_intra = k < n_wave_intra;
        tw_l.push_back(table([&](int i) { return lf_lvl[i][k] >= 0 ? (uint32_t)lf_lvl[i][k] : 0u; },
                             [&](int i) {
                               const int j = lf_lvl[i][k];
                               const auto &ls = f[i]->p->lf_level_start;
                               return j >= 0 ? (ls[j + 1] - ls[j]) / (with_intra ? 3 : 1) : 0u;
                             },
                             1));
      }
    }
    if (lf) {
      // decisions: a workgroup per superblock entry, or per superblock once
      // the batch is large enough to keep the chip busy that way
      size_t batch_sb = 0;
      for (int i = 0; i < n; i++) batch_sb += (size_t)f[i]->p->hdr.sb64_rows * f[i]->p->hdr.sb64_cols;
      mask_per_sb = batch_sb >= kMaskPerSbMin ? 6 : 1;
      t_mask = table([](int) { return 0u; },
                     [&](int i) {
                       const auto &h = f[i]->p->hdr;
                       return f[i]->p->lf_level_start.empty() ? 0u : (6 / mask_per_sb) * h.sb64_rows * h.sb64_cols;
                     },
                     1);
      size_t levels = 0;
      for (int i = 0; i < n && !fused; i++)
        if (f[i]->p->lf_level_start.size() > 1) levels = std::max(levels, f[i]->p->lf_level_start.size() - 1);
      for (size_t l = 0; l < levels; l++)
        t_lf.push_back(table([&](int) { return (uint32_t)l; },
            [&](int i) {
              const auto &ls = f[i]->p->lf_level_start;
              return l + 1 < ls.size() ? ls[l + 1] - ls[l] : 0u;
            },
            1));
    }
    // per-frame arguments
    for (int i = 0; i < n; i++) {
      const Prepared &p = *f[i]->p;
      const uint8_t *dv = f[i]->dv;
      vp9r_gpu_frame *cur = f[i]->cur;
      vp9g::ClearArgs &ca = ((vp9g::ClearArgs *)(st.data() + o_clear))[i];
      ca.f = cur->planes;
      ca.g = p.geo;
      vp9g::McArgs &ma = ((vp9g::McArgs *)(st.data() + o_mc))[i];
      ma = vp9g::McArgs{};
      ma.g = p.geo;
      ma.blocks = (const vp9r_block *)(dv + p.o_blocks);
      ma.tiles = (const vp9g::McTile *)(dv + p.o_tiles);
      ma.cur = cur->planes;
      for (int k = 0; k < 3; k++)
        if (f[i]->refs[k]) {
          ma.ref[k] = f[i]->refs[k]->planes;
          ma.ref_w[k] = f[i]->refs[k]->width;
          ma.ref_h[k] = f[i]->refs[k]->height;
        }
      vp9g::ItxArgs &ia = ((vp9g::ItxArgs *)(st.data() + o_itx))[i];
      ia = vp9g::ItxArgs{};
      ia.g = p.geo;
      ia.blocks = ma.blocks;
      ia.txbs = (const vp9r_txb *)(dv + p.o_txbs);
      for (int k = 0; k < 4; k++) ia.jobs[k] = (const uint32_t *)(dv + p.o_jobs[k]);
      ia.coef_pos = (const uint16_t *)(dv + p.o_cpos);
      ia.coef_val = (const int32_t *)(dv + p.o_cval);
      ia.cur = cur->planes;
      memcpy(ia.q, p.q, sizeof p.q);
      if (res_bytes && p.n_items)
        for (int pl = 0; pl < 3; pl++) ia.res[pl] = (int16_t *)((uint8_t *)g->intra_res.p + res_off[3 * i + pl]);
      vp9g::IntraArgs &na = ((vp9g::IntraArgs *)(st.data() + o_intra))[i];
      na = vp9g::IntraArgs{};
      na.g = p.geo;
      na.blocks = ma.blocks;
      na.txbs = ia.txbs;
      na.txb_index = (const uint32_t *)(dv + p.o_iidx);
      na.groups = (const vp9g::IntraGroup *)(dv + p.o_groups);
      na.items = (const vp9g::IntraBlockItem *)(dv + p.o_items);
      na.sb_tx = (const vp9g::IntraSbTx *)(dv + p.o_sbtx);
      na.sb_units = (const uint32_t *)(dv + p.o_sbunits);
      na.sb_levels = (const uint2 *)(dv + p.o_sblevels);
      na.coef_pos = ia.coef_pos;
      na.coef_val = ia.coef_val;
      na.cur = cur->planes;
      for (int pl = 0; pl < 3; pl++) na.res[pl] = ia.res[pl];
      memcpy(na.q, p.q, sizeof p.q);
      vp9g::LfArgs &la = ((vp9g::LfArgs *)(st.data() + o_lf))[i];
      la = vp9g::LfArgs{};
      la.g = p.geo;
      la.mi = (const vp9r_mi *)(dv + p.o_mi);
      la.dq = (uint32_t *)((uint8_t *)g->lf_mask.p + mask_off[i]);
      la.cur = cur->planes;
      la.p = p.lfp;
    }
    const uint8_t *d = nullptr;
    BatchSlot *slotp = nullptr;
    if (gt) {
      // the graph's own table copy, then the capture of the launches alone
      vp9g::mc_prepare();  // no synchronous copy may happen under the capture
      HIP_TRY(hipGetLastError(), "interpolation taps");
      HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
      HIP_TRY(gt->tables.ensure(st.size()), "hipMalloc(graph tables)");
      HIP_TRY(hipMemcpy(gt->tables.p, st.data(), st.size(), hipMemcpyHostToDevice), "upload graph tables");
      d = (const uint8_t *)gt->tables.p;
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    } else {
      // one transfer into the next device slot
      BatchSlot &slot = g->slots[g->slot_next];
      slotp = &slot;
      g->slot_next = (g->slot_next + 1) % kBatchSlots;
      if (slot.pending) {
        HIP_TRY(hipEventSynchronize(slot.ev), "batch slot reuse");
        slot.pending = false;
      }
      HIP_TRY(slot.host.ensure(st.size()), "hipHostMalloc(batch)");
      HIP_TRY(slot.dev.ensure(st.size()), "hipMalloc(batch)");
      memcpy(slot.host.p, st.data(), st.size());
      // on its own stream: the copy runs while earlier batches still
      // filter, and the host waits for it before queueing the launches.  A
      // stream wait on the copy's event instead left the kernel stream idle
      // about 10 us at every batch even when the copy had long finished
      // (1080p stream: 2011 -> 2030 frames/s without it); the host is
      // frames ahead of the device here.
      HIP_TRY(hipMemcpyAsync(slot.dev.p, slot.host.p, st.size(), hipMemcpyHostToDevice, g->up_stream),
              "upload batch");
      HIP_TRY(hipEventRecord(g->ev_up, g->up_stream), "hipEventRecord");
      HIP_TRY(hipEventSynchronize(g->ev_up), "upload batch");
      d = (const uint8_t *)slot.dev.p;
    }
    auto R = [&](const LaunchTable &t) { return (const vp9g::Range *)(d + t.off); };
    if (recon) {
      if (!gt)  // a graph's launch waits for the downloads before it
        for (int i = 0; i < n; i++) HIP_TRY(order_after_copy(s, f[i]->cur), "hipStreamWaitEvent");
      // the margin clear, with the loop filter decisions in the same launch
      // when the filter follows in this call on this stream; a single
      // frame's inter prediction in that launch too
      const bool masks = fused && lf;
      if (n == 1) {
        vp9g::launch_prep_mc(s, (const vp9g::ClearArgs *)(d + o_clear), R(t_clear), t_clear.wgs,
                             (const vp9g::LfArgs *)(d + o_lf), R(t_mask), masks ? t_mask.wgs : 0u, mask_per_sb,
                             (const vp9g::McArgs *)(d + o_mc), R(t_mc),
                             (const vp9g::McTile *)(f[0]->dv + f[0]->p->o_tiles), f[0]->p->n_tiles, hbd);
      } else {
        if (masks)
          vp9g::launch_prep(s, (const vp9g::ClearArgs *)(d + o_clear), R(t_clear), t_clear.wgs,
                            (const vp9g::LfArgs *)(d + o_lf), R(t_mask), t_mask.wgs, n, mask_per_sb, hbd);
        else
          vp9g::launch_clear_margins(s, (const vp9g::ClearArgs *)(d + o_clear), R(t_clear), n, t_clear.wgs, hbd);
        vp9g::launch_mc(s, (const vp9g::McArgs *)(d + o_mc), R(t_mc), n, t_mc.wgs, hbd, nullptr, 0u);
      }
      if (g->timing) HIP_TRY(hipEventRecord(g->ev[1], s), "hipEventRecord");
      {
        const vp9g::Range *rk[4] = {R(t_itx[0]), R(t_itx[1]), R(t_itx[2]), R(t_itx[3])};
        const uint32_t wk[4] = {t_itx[0].wgs, t_itx[1].wgs, t_itx[2].wgs, t_itx[3].wgs};
        vp9g::launch_itx_inter(s, (const vp9g::ItxArgs *)(d + o_itx), rk, n, wk, hbd);
      }
      if (g->timing) HIP_TRY(hipEventRecord(g->ev[2], s), "hipEventRecord");
    }
    if (recon && !fused) {
      HIP_TRY(hipEventRecord(g->ev_pre, s), "hipEventRecord");
      // the intra levels the loop filter waits for (an event after each)
      std::vector<char> mark(t_intra.size(), 0);
      if (lf && !t_intra.empty()) {
        if (groups) {
          mark.back() = 1;
        } else {
          for (size_t l = 0; l < t_lf.size(); l++)
            for (int i = 0; i < n; i++) {
              const auto &nd = f[i]->p->lf_need;
              if (l < nd.size() && nd[l] >= 0) mark[nd[l]] = 1;
            }
        }
      }
      while (g->intra_events.size() < t_intra.size()) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        g->intra_events.push_back(e);
      }
      for (size_t l = 0; l < t_intra.size(); l++) {
        const LaunchTable &t = t_intra[l];
        vp9g::launch_intra(s, (const vp9g::IntraArgs *)(d + o_intra), R(t), n, t.wgs, hbd);
        if (mark[l]) HIP_TRY(hipEventRecord(g->intra_events[l], s), "hipEventRecord");
      }
      if (g->timing) HIP_TRY(hipEventRecord(g->ev[3], s), "hipEventRecord");
    }
    if (fused) {
      // everything on the kernel stream: the decisions, then the fused levels
      if (!recon && !gt)
        for (int i = 0; i < n; i++) HIP_TRY(order_after_copy(s, f[i]->cur), "hipStreamWaitEvent");
      if (lf && !recon)  // else in the prep launch
        vp9g::launch_lf_mask(s, (const vp9g::LfArgs *)(d + o_lf), R(t_mask), n, t_mask.wgs, mask_per_sb);
      if (recon && g->timing && !n_wave_intra) HIP_TRY(hipEventRecord(g->ev[3], s), "hipEventRecord");
      uint32_t lds = 0;
      for (int i = 0; i < n; i++) lds = std::max(lds, vp9g::wave_lds_bytes(f[i]->p->hdr.ss_x, f[i]->p->hdr.ss_y));
      // A single frame's levels after its intra levels: the first ones one
      // per launch in place, the rest several per launch (k_lf_cone), from
      // the frame store to lf_tmp and back (lf_tail_plan).
      size_t cone_from = n_wave;
      vp9g::Planes tmp{};
      LfPlan plan;
      if (n == 1 && !gt && lf && g->lf_depth > 1 && n_wave > n_wave_intra) {
        plan = lf_tail_plan(n_wave - n_wave_intra, std::min(g->lf_depth, hbd ? 3 : 5));
        if (!plan.depths.empty()) {
          const vp9r_gpu_frame *cf = f[0]->cur;
          if (cf->bytes > g->lf_tmp.cap) {
            HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
            HIP_TRY(g->lf_tmp.ensure(cf->bytes), "hipMalloc(loop filter planes)");
          }
          tmp = cf->planes;
          for (int p = 0; p < 3; p++)
            tmp.p[p] = (uint8_t *)g->lf_tmp.p + ((const uint8_t *)cf->planes.p[p] - (const uint8_t *)cf->base);
          cone_from = n_wave_intra + plan.singles;
        }
      }
      for (size_t k = 0; k < n_wave; k++) {
        if (k == cone_from) {
          const vp9g::LfArgs &la = ((const vp9g::LfArgs *)(st.data() + o_lf))[0];
          size_t kk = k;
          int prev = 0;
          for (size_t i = 0; i < plan.depths.size(); i++) {
            vp9g::LfOne one = vp9g::lf_one(la, (uint32_t)lf_lvl[0][kk]);
            const bool back = i % 2 == 1;  // lf_tmp to the frame store
            if (back)
              for (int p = 0; p < 3; p++) one.p[p] = (uint8_t *)tmp.p[p];
            vp9g::launch_lf_cone(s, one, back ? la.cur : tmp, plan.depths[i], prev, hbd);
            prev = plan.depths[i];
            kk += (size_t)prev;
          }
          break;
        }
        if (k >= n_wave_intra) {
          vp9g::launch_lf_level(s, (const vp9g::LfArgs *)(d + o_lf), R(tw_l[k]), n, tw_l[k].wgs, hbd,
                                n == 1 && lf_lvl[0][k] >= 0
                                    ? vp9g::lf_one(((const vp9g::LfArgs *)(st.data() + o_lf))[0], (uint32_t)lf_lvl[0][k])
                                    : vp9g::LfOne{});
          continue;
        }
        vp9g::WaveLaunch w;
        w.n_iw = tw_i[k].wgs;
        w.n_lw = tw_l[k].wgs;
        w.nf = (uint32_t)n;
        w.lds_bytes = lds;
        w.fa_i = (const vp9g::IntraArgs *)(d + o_intra);
        w.rg_i = R(tw_i[k]);
        w.fa_l = (const vp9g::LfArgs *)(d + o_lf);
        w.rg_l = R(tw_l[k]);
        if (n == 1) {
          // host copies of the frame's arguments: pointers as kernel arguments
          w.one_i = (const vp9g::IntraArgs *)(st.data() + o_intra);
          w.one_l = (const vp9g::LfArgs *)(st.data() + o_lf);
          if (w.n_iw)
            w.items0 = (const vp9g::IntraBlockItem *)(f[0]->dv + f[0]->p->o_items) + f[0]->p->item_level_start[k];
          w.lf_level = lf_lvl[0][k] >= 0 ? (uint32_t)lf_lvl[0][k] : 0u;
        }
        vp9g::launch_wave(s, w, hbd);
        if (recon && g->timing && k + 1 == n_wave_intra) HIP_TRY(hipEventRecord(g->ev[3], s), "hipEventRecord");
      }
    } else if (lf) {
      // The loop filter runs on its own stream: the decisions and levels
      // once the inter work is done, each level once its intra levels
      // (lf_need) are, overlapping the others.
      // The main stream then waits for it (later frames read this one).
      hipStream_t sl = g->lf_stream;
      if (!recon) HIP_TRY(hipEventRecord(g->ev_pre, s), "hipEventRecord");
      for (int i = 0; i < n; i++) HIP_TRY(order_after_copy(sl, f[i]->cur), "hipStreamWaitEvent");
      // ev_pre follows, on the kernel stream, the batch tables, the frames'
      // records (uploaded on that stream) and the inter stages
      HIP_TRY(hipStreamWaitEvent(sl, g->ev_pre, 0), "hipStreamWaitEvent");
      vp9g::launch_lf_mask(sl, (const vp9g::LfArgs *)(d + o_lf), R(t_mask), n, t_mask.wgs, mask_per_sb);
      int waited = -1;
      for (size_t l = 0; l < t_lf.size(); l++) {
        int need = -1;
        if (recon && !t_intra.empty()) {
          if (groups) {
            need = (int)t_intra.size() - 1;
          } else {
            for (int i = 0; i < n; i++) {
              const auto &nd = f[i]->p->lf_need;
              if (l < nd.size()) need = std::max(need, (int)nd[l]);
            }
          }
        }
        if (need > waited) {
          HIP_TRY(hipStreamWaitEvent(sl, g->intra_events[need], 0), "hipStreamWaitEvent");
          waited = need;
        }
        vp9g::launch_lf_level(sl, (const vp9g::LfArgs *)(d + o_lf), R(t_lf[l]), n, t_lf[l].wgs, hbd,
                              n == 1 ? vp9g::lf_one(((const vp9g::LfArgs *)(st.data() + o_lf))[0], (uint32_t)l)
                                     : vp9g::LfOne{});
      }
      HIP_TRY(hipEventRecord(g->ev_lf_done, sl), "hipEventRecord");
      HIP_TRY(hipStreamWaitEvent(s, g->ev_lf_done, 0), "hipStreamWaitEvent");
    }
    if (gt) {
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(s, &gt->graph);
      HIP_TRY(le, "kernel launch (capture)");
      HIP_TRY(ce, "hipStreamEndCapture");
    } else {
      HIP_TRY(hipEventRecord(slotp->ev, s), "hipEventRecord");
      slotp->pending = true;
    }
  }
  if (lf && g->timing) HIP_TRY(hipEventRecord(g->ev[4], s), "hipEventRecord");
  HIP_TRY(hipGetLastError(), "kernel launch");
  return VP9R_OK;
}

}  // namespace vp9r

struct vp9r_prepared {
  vp9r::Prepared p;
  vp9r::DevBuf dev;
};

extern "C" {

int vp9r_gpu_reconstruct(vp9r_gpu *g, const vp9r_record *rec, vp9r_gpu_frame *const refs[3],
                         vp9r_gpu_frame *cur) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (g->upload_pending) {
    HIP_TRY(hipEventSynchronize(g->ev_upload), "staging reuse");
    g->upload_pending = false;
  }
  Prepared &p = g->last;
  int st = build_record(g->work, rec, p, g->stage);
  if (st) return st;
  st = check_frames(p, refs, cur);
  if (st) return st;
  HIP_TRY(g->arena.ensure(p.bytes), "hipMalloc(arena)");
  hipStream_t s = g->stream;
  if (g->timing) {
    HIP_TRY(begin_timed_frame(g, 1), "stage timing");
    HIP_TRY(hipEventRecord(g->ev[0], s), "hipEventRecord");
  }
  HIP_TRY(hipMemcpyAsync(g->arena.p, g->stage.p, p.bytes, hipMemcpyHostToDevice, s), "upload record");
  HIP_TRY(hipEventRecord(g->ev_upload, s), "hipEventRecord");
  g->upload_pending = true;
  g->last_valid = true;
  BatchFrame bf{&p, (const uint8_t *)g->arena.p, {nullptr, nullptr, nullptr}, cur};
  if (!p.hdr.frame_is_intra)
    for (int k = 0; k < 3; k++) bf.refs[k] = refs[k];
  return launch_batch(g, &bf, 1, true, false);
}

int vp9r_gpu_loop_filter(vp9r_gpu *g, const vp9r_record *rec, vp9r_gpu_frame *cur) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  // the record just reconstructed: its inputs are on the device already
  if (!g->last_valid || memcmp(&g->last.hdr, &rec->hdr, sizeof rec->hdr) != 0) {
    set_error("vp9r_gpu_loop_filter takes the record last given to vp9r_gpu_reconstruct");
    return VP9R_ERR_INVALID_INPUT;
  }
  if (!cur || cur->width != rec->hdr.width || cur->height != rec->hdr.height ||
      cur->bit_depth != rec->hdr.bit_depth || cur->ss_x != rec->hdr.ss_x || cur->ss_y != rec->hdr.ss_y) {
    set_error("the frame store does not match the record's geometry");
    return VP9R_ERR_INVALID_INPUT;
  }
  if (g->timing && !g->ev) {
    set_error("timing was enabled between vp9r_gpu_reconstruct and vp9r_gpu_loop_filter");
    return VP9R_ERR_INVALID_INPUT;
  }
  BatchFrame bf{&g->last, (const uint8_t *)g->arena.p, {nullptr, nullptr, nullptr}, cur};
  return launch_batch(g, &bf, 1, false, true);
}

int vp9r_gpu_prepare(vp9r_gpu *g, const vp9r_record *rec, vp9r_prepared **out) {
  *out = nullptr;
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (g->upload_pending) {
    HIP_TRY(hipEventSynchronize(g->ev_upload), "staging reuse");
    g->upload_pending = false;
  }
  vp9r_prepared *pr = new vp9r_prepared();
  int st = build_record(g->work, rec, pr->p, g->stage);
  if (st) {
    delete pr;
    return st;
  }
  hipError_t e = pr->dev.ensure(pr->p.bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(pr->dev.p, g->stage.p, pr->p.bytes, hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) {
    delete pr;
    return hip_fail(e, "upload prepared record");
  }
  *out = pr;
  return VP9R_OK;
}

void vp9r_prepared_free(vp9r_prepared *p) { delete p; }

int vp9r_gpu_run(vp9r_gpu *g, const vp9r_prepared *pr, vp9r_gpu_frame *const refs[3], vp9r_gpu_frame *cur) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  const vp9r_prepared *one[1] = {pr};
  vp9r_gpu_frame *const cur1[1] = {cur};
  vp9r_gpu_frame *none[3] = {nullptr, nullptr, nullptr};
  return vp9r_gpu_run_batch(g, 1, one, refs ? refs : none, cur1);
}

int vp9r_gpu_run_batch(vp9r_gpu *g, uint32_t n, const vp9r_prepared *const *pr, vp9r_gpu_frame *const *refs,
                       vp9r_gpu_frame *const *cur) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (n == 0) return VP9R_OK;
  if (!pr || !cur || !refs) {
    set_error("vp9r_gpu_run_batch: null argument");
    return VP9R_ERR_INVALID_INPUT;
  }
  std::vector<BatchFrame> bf(n);
  for (uint32_t i = 0; i < n; i++) {
    if (!pr[i]) {
      set_error("vp9r_gpu_run_batch: null prepared record");
      return VP9R_ERR_INVALID_INPUT;
    }
    const Prepared &p = pr[i]->p;
    int st = check_frames(p, refs + 3 * i, cur[i]);
    if (st) return st;
    bf[i] = BatchFrame{&p, (const uint8_t *)pr[i]->dev.p, {nullptr, nullptr, nullptr}, cur[i]};
    if (!p.hdr.frame_is_intra)
      for (int k = 0; k < 3; k++) bf[i].refs[k] = refs[3 * i + k];
  }
  // frames of one batch run side by side: none may read or write another's output
  for (uint32_t i = 0; i < n; i++)
    for (uint32_t j = 0; j < n; j++) {
      if (i != j && bf[i].cur == bf[j].cur) {
        set_error("vp9r_gpu_run_batch: two frames of the batch write the same frame store");
        return VP9R_ERR_INVALID_INPUT;
      }
      for (int k = 0; k < 3; k++)
        if (i != j && bf[i].refs[k] == bf[j].cur) {
          set_error("vp9r_gpu_run_batch: a frame of the batch references another frame's output");
          return VP9R_ERR_INVALID_INPUT;
        }
    }
  if (g->timing) {
    HIP_TRY(begin_timed_frame(g, n), "stage timing");
    HIP_TRY(hipEventRecord(g->ev[0], g->stream), "hipEventRecord");
  }
  return launch_batch(g, bf.data(), (int)n, true, true);
}

}  // extern "C"

// A frame's launches captured once and replayed: the graph holds its own
// batch tables and the frame stores it was made for.
struct vp9r_graph {
  vp9r_gpu *g = nullptr;
  const vp9r_prepared *p = nullptr;
  vp9r_gpu_frame *refs[3] = {nullptr, nullptr, nullptr};
  vp9r_gpu_frame *cur = nullptr;
  vp9r::GraphTarget target;
  hipGraphExec_t exec = nullptr;
  uint32_t nodes = 0;
  ~vp9r_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (target.graph) (void)hipGraphDestroy(target.graph);
  }
};

namespace {

// Captures the launches of (p, refs, cur) into gr->target (tables and graph).
int capture_frame(vp9r_gpu *g, const vp9r_prepared *pr, vp9r_gpu_frame *const refs[3], vp9r_gpu_frame *cur,
                  vp9r_graph *gr) {
  const Prepared &p = pr->p;
  vp9r_gpu_frame *none[3] = {nullptr, nullptr, nullptr};
  int st = check_frames(p, refs ? refs : none, cur);
  if (st) return st;
  BatchFrame bf{&p, (const uint8_t *)pr->dev.p, {nullptr, nullptr, nullptr}, cur};
  if (!p.hdr.frame_is_intra)
    for (int k = 0; k < 3; k++) bf.refs[k] = refs[k];
  if (gr->target.graph) {
    (void)hipGraphDestroy(gr->target.graph);
    gr->target.graph = nullptr;
  }
  st = launch_batch(g, &bf, 1, true, true, &gr->target);
  if (st) return st;
  size_t n = 0;
  HIP_TRY(hipGraphGetNodes(gr->target.graph, nullptr, &n), "hipGraphGetNodes");
  gr->nodes = (uint32_t)n;
  gr->p = pr;
  for (int k = 0; k < 3; k++) gr->refs[k] = refs ? refs[k] : nullptr;
  gr->cur = cur;
  return VP9R_OK;
}

}  // namespace

extern "C" {

int vp9r_gpu_graph_new(vp9r_gpu *g, const vp9r_prepared *pr, vp9r_gpu_frame *const refs[3], vp9r_gpu_frame *cur,
                       vp9r_graph **out) {
  *out = nullptr;
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (!pr) {
    set_error("vp9r_gpu_graph_new: null prepared record");
    return VP9R_ERR_INVALID_INPUT;
  }
  vp9r_graph *gr = new (std::nothrow) vp9r_graph();
  if (!gr) return VP9R_ERR_INVALID_INPUT;
  gr->g = g;
  int st = capture_frame(g, pr, refs, cur, gr);
  if (!st) {
    hipError_t e = hipGraphInstantiate(&gr->exec, gr->target.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      st = VP9R_ERR_DEVICE;
    }
  }
  if (st) {
    delete gr;
    return st;
  }
  gpu_retain(g);
  *out = gr;
  return VP9R_OK;
}

int vp9r_gpu_graph_update(vp9r_graph *gr, vp9r_gpu_frame *const refs[3], vp9r_gpu_frame *cur) {
  vp9r_gpu *g = gr->g;
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  // a fresh capture with the new stores (tables re-uploaded), its kernel
  // parameters moved into the instantiated graph; a fresh instantiation if
  // the topology differs
  int st = capture_frame(g, gr->p, refs, cur, gr);
  if (st) return st;
  hipGraphNode_t err_node = nullptr;
  hipGraphExecUpdateResult res;
  hipError_t e = hipGraphExecUpdate(gr->exec, gr->target.graph, &err_node, &res);
  if (e != hipSuccess || res != hipGraphExecUpdateSuccess) {
    (void)hipGetLastError();
    (void)hipGraphExecDestroy(gr->exec);
    gr->exec = nullptr;
    e = hipGraphInstantiate(&gr->exec, gr->target.graph, nullptr, nullptr, 0);
    HIP_TRY(e, "hipGraphInstantiate");
  }
  return VP9R_OK;
}

int vp9r_gpu_graph_run(vp9r_graph *gr) {
  vp9r_gpu *g = gr->g;
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (g->timing) {
    set_error("vp9r_gpu_graph_run: stage timing is on");
    return VP9R_ERR_INVALID_INPUT;
  }
  // the downloads in flight of the store it writes, then the graph
  HIP_TRY(order_after_copy(g->stream, gr->cur), "hipStreamWaitEvent");
  HIP_TRY(hipGraphLaunch(gr->exec, g->stream), "hipGraphLaunch");
  return VP9R_OK;
}

uint32_t vp9r_gpu_graph_nodes(const vp9r_graph *gr) { return gr->nodes; }

void vp9r_gpu_graph_free(vp9r_graph *gr) {
  if (!gr) return;
  vp9r_gpu *g = gr->g;
  (void)hipStreamSynchronize(g->stream);
  delete gr;
  gpu_release(g);
}

int vp9r_gpu_filter_batch(vp9r_gpu *g, uint32_t n, const vp9r_prepared *const *pr, vp9r_gpu_frame *const *cur) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (n == 0) return VP9R_OK;
  if (!pr || !cur) {
    set_error("vp9r_gpu_filter_batch: null argument");
    return VP9R_ERR_INVALID_INPUT;
  }
  if (g->timing) {
    set_error("vp9r_gpu_filter_batch: stage timing is on");
    return VP9R_ERR_INVALID_INPUT;
  }
  std::vector<BatchFrame> bf(n);
  for (uint32_t i = 0; i < n; i++) {
    if (!pr[i]) {
      set_error("vp9r_gpu_filter_batch: null prepared record");
      return VP9R_ERR_INVALID_INPUT;
    }
    const Prepared &p = pr[i]->p;
    if (int st = check_cur(p, cur[i])) return st;
    bf[i] = BatchFrame{&p, (const uint8_t *)pr[i]->dev.p, {nullptr, nullptr, nullptr}, cur[i]};
    for (uint32_t j = 0; j < i; j++)
      if (bf[j].cur == cur[i]) {
        set_error("vp9r_gpu_filter_batch: two frames of the batch filter the same frame store");
        return VP9R_ERR_INVALID_INPUT;
      }
  }
  HIP_TRY(seam_begin(g), "hipEventRecord");
  int st = launch_batch(g, bf.data(), (int)n, false, true);
  if (st) return st;
  HIP_TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
  HIP_TRY(hipStreamSynchronize(g->stream), "filter sync");
  HIP_TRY(seam_end(g), "hipEventElapsedTime");
  return VP9R_OK;
}

int vp9r_gpu_prepared_schedule(const vp9r_prepared *pr, uint32_t out[4]) {
  size_t n_int = 0;
  std::vector<int> lf_at;
  const size_t n_wave = vp9r::wave_schedule(pr->p, true, pr->p.hdr.lf_level != 0, n_int, lf_at);
  const vp9r::LfPlan plan =
      vp9r::lf_tail_plan(n_wave - n_int, std::min(vp9r::kLfDepthDefault, pr->p.hdr.bit_depth > 8 ? 3 : 5));
  out[0] = (uint32_t)n_int;
  out[1] = (uint32_t)(plan.singles + plan.depths.size());
  out[2] = (uint32_t)lf_at.size();
  out[3] = (uint32_t)(n_wave - n_int);
  return VP9R_OK;
}

int vp9r_gpu_prepared_stats(const vp9r_prepared *pr, uint64_t out[12]) {
  const vp9r::Prepared &p = pr->p;
  for (int k = 0; k < 4; k++) out[8 + k] = p.work_bytes[k];
  out[0] = p.n_tiles;
  for (int k = 0; k < 4; k++) out[1 + k] = p.n_jobs[k];
  const bool blocks = true;
  out[5] = blocks ? p.n_items : p.n_groups;
  out[6] = blocks ? (p.item_level_start.empty() ? 0 : p.item_level_start.size() - 1)
                  : (p.group_level_start.empty() ? 0 : p.group_level_start.size() - 1);
  out[7] = p.lf_level_start.empty() ? 0 : p.lf_level_start.size() - 1;
  return VP9R_OK;
}

int vp9r_gpu_download_packed(vp9r_gpu *g, vp9r_gpu_frame *f, uint8_t *dst) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (f->g != g) {
    set_error("the frame store belongs to another context");
    return VP9R_ERR_INVALID_INPUT;
  }
  size_t bps = f->bit_depth > 8 ? 2 : 1, off = 0;
  for (int p = 0; p < 3; p++) {
    size_t w = p ? (f->width + f->ss_x) >> f->ss_x : f->width;
    size_t hh = p ? (f->height + f->ss_y) >> f->ss_y : f->height;
    HIP_TRY(hipMemcpy2DAsync(dst + off, w * bps, f->planes.p[p], (size_t)f->planes.stride[p] * bps, w * bps,
                             hh, hipMemcpyDeviceToHost, g->stream),
            "download frame");
    off += w * hh * bps;
  }
  HIP_TRY(hipStreamSynchronize(g->stream), "download sync");
  return VP9R_OK;
}

int vp9r_gpu_download_packed_async(vp9r_gpu *g, vp9r_gpu_frame *f, uint8_t *dst) {
  HIP_TRY(hipSetDevice(g->device), "hipSetDevice");
  if (f->g != g) {
    set_error("the frame store belongs to another context");
    return VP9R_ERR_INVALID_INPUT;
  }
  // the copy starts once the work queued so far on the compute stream is done
  HIP_TRY(hipEventRecord(f->ready, g->stream), "hipEventRecord");
  HIP_TRY(hipStreamWaitEvent(g->copy_stream, f->ready, 0), "hipStreamWaitEvent");
  size_t bps = f->bit_depth > 8 ? 2 : 1, off = 0;
  for (int p = 0; p < 3; p++) {
    size_t w = p ? (f->width + f->ss_x) >> f->ss_x : f->width;
    size_t hh = p ? (f->height + f->ss_y) >> f->ss_y : f->height;
    HIP_TRY(hipMemcpy2DAsync(dst + off, w * bps, f->planes.p[p], (size_t)f->planes.stride[p] * bps, w * bps,
                             hh, hipMemcpyDeviceToHost, g->copy_stream),
            "download frame");
    off += w * hh * bps;
  }
  HIP_TRY(hipEventRecord(f->copy_done, g->copy_stream), "hipEventRecord");
  f->copy_pending = true;
  return VP9R_OK;
}

int vp9r_gpu_frame_wait(vp9r_gpu_frame *f) {
  if (!f->copy_pending) return VP9R_OK;
  HIP_TRY(hipEventSynchronize(f->copy_done), "hipEventSynchronize");
  f->copy_pending = false;
  return VP9R_OK;
}

int vp9r_host_alloc(size_t bytes, void **out) {
  *out = nullptr;
  HIP_TRY(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
  return VP9R_OK;
}

void vp9r_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

}  // extern "C"
