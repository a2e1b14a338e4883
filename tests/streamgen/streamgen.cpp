// Stream generator for the benchmark and parity workloads: writes VP9
// bitstreams whose syntax statistics follow the configurations the
// reference's own encoder would be run on (its encoder cannot run here).
// TEST INFRASTRUCTURE.
//
// The bits are produced by the product parser's own syntax walk in write
// mode (vp9r::Parser::write_frame), so every context, motion vector
// candidate list and token context is the decoder's; this file supplies the
// headers and the decisions (a SyntaxPlan):
//   - content 0: a translating sinusoid grid (one global motion vector);
//   - content 1: textured rectangles moving with their own 1/8-sample
//     velocities over a slowly panning background: partitions follow the
//     rectangles' edges, each block takes the motion of what covers it,
//     newly uncovered areas are intra coded;
//   - residuals are open-loop: quantised coefficients drawn per transform
//     block (sparse, Laplacian, more for intra), not a transform of a
//     difference, since no reconstruction runs here.
// Frame headers mirror the crate's encoder (src/encoder/mod.rs:558-697):
// key frames, then error-resilient, frame-parallel inter frames without
// context refresh, EIGHTTAP, single reference, TX_MODE_SELECT, loop filter
// level (q * 10 + 32) / 64 unless set, sharpness as set (0 by default), LAST refreshed
// every frame and GOLDEN every golden_interval frames.  Optional resizing
// codes every Nth frame at half size (scaled references both ways).
// Optional segmentation (segmentation = 1): a segment map written every
// frame but every third inter frame (which then keeps the predicted map,
// all 0 after the error-resilient reset), temporally predicted on odd
// frames, with the features of seg_features() in delta mode on even frames
// and absolute mode on odd ones: quantiser (ALT_Q) and filter level
// (ALT_L) changes, a forced LAST reference (REF_FRAME, inter frames) and
// forced skip (SKIP).
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "host/parser.h"
#include "vp9_consts.h"
#include "vp9_tables.h"

using namespace vp9r;

// The coefficient scan of a transform size and type (the decoder's tables).
static const uint16_t *scan_of(int tx_size, int tx_type) {
  return tx_size == TX_4X4     ? (tx_type == ADST_DCT ? VP9_ROW_SCAN_4X4 : tx_type == DCT_ADST ? VP9_COL_SCAN_4X4 : VP9_DEFAULT_SCAN_4X4)
         : tx_size == TX_8X8   ? (tx_type == ADST_DCT ? VP9_ROW_SCAN_8X8 : tx_type == DCT_ADST ? VP9_COL_SCAN_8X8 : VP9_DEFAULT_SCAN_8X8)
         : tx_size == TX_16X16 ? (tx_type == ADST_DCT ? VP9_ROW_SCAN_16X16 : tx_type == DCT_ADST ? VP9_COL_SCAN_16X16 : VP9_DEFAULT_SCAN_16X16)
                               : VP9_DEFAULT_SCAN_32X32;
}

extern "C" {
typedef struct sg_config {
  uint32_t width, height, frames, seed;
  uint32_t bit_depth, ss_x, ss_y;
  uint32_t base_q_idx, tile_cols_log2, golden_interval;
  uint32_t content;       // 0 sinusoid grid, 1 moving rectangles
  uint32_t resize_every;  // 0 never, else every Nth frame at half size
  uint32_t intra_percent; // share of inter-frame blocks coded intra
  uint32_t allow_hp;
  uint32_t lf_level;      // 0 derived from the quantiser, else the filter level + 1
  uint32_t lf_sharpness;  // 0..7
  uint32_t segmentation;  // 0 off; 1 segment map and features (below)
} sg_config;
}

namespace {

class BitWriter {
 public:
  std::vector<uint8_t> out;
  void f(int n, uint32_t v) {
    for (int i = n - 1; i >= 0; i--) {
      if ((pos_ & 7) == 0) out.push_back(0);
      if ((v >> i) & 1) out.back() |= (uint8_t)(0x80 >> (pos_ & 7));
      pos_++;
    }
  }
  void align() { pos_ = (pos_ + 7) & ~(size_t)7; }

 private:
  size_t pos_ = 0;
};

struct Rect {
  double x, y, w, h, vx, vy;  // samples, samples per frame
};

class Plan : public SyntaxPlan {
 public:
  Plan(const sg_config &c, uint32_t frame_w, uint32_t frame_h, int t, bool key, const std::vector<Rect> &rects,
       double bg_vx, double bg_vy, std::mt19937_64 &rng, bool seg_map)
      : c_(c), w_(frame_w), h_(frame_h), t_(t), key_(key), rects_(rects), bgx_(bg_vx), bgy_(bg_vy), rng_(rng),
        seg_map_(seg_map) {}

  // The segment of a block: from its position and the frame when the map is
  // written, else the predicted one (0: the error-resilient reset map).
  uint8_t segment_of(uint32_t mi_row, uint32_t mi_col) const {
    if (!c_.segmentation || !seg_map_) return 0;
    const uint32_t v = (mi_row / 2) * 5 + (mi_col / 2) * 3 + (uint32_t)t_ * 7;
    const uint8_t sid = (uint8_t)((v * 2654435761u) >> 29);  // 0..7
    return key_ && sid == 4 ? 0 : sid;  // no forced reference on an intra frame
  }

  // the object covering luma sample (x, y): -1 background, else a rectangle
  int object_at(double x, double y) const {
    for (int i = (int)rects_.size() - 1; i >= 0; i--) {
      const Rect &r = rects_[i];
      double rx = r.x + r.vx * t_, ry = r.y + r.vy * t_;
      if (x >= rx && x < rx + r.w && y >= ry && y < ry + r.h) return i;
    }
    return -1;
  }
  // an object near (x, y) that was not there one frame earlier: uncovered
  bool uncovered(double x, double y) const {
    if (c_.content != 1 || t_ == 0) return false;
    int now = object_at(x, y);
    Plan prev = *this;
    prev.t_ = t_ - 1;
    return prev.object_at(x, y) != now;
  }

  int partition(uint32_t mi_row, uint32_t mi_col, int bsize) override {
    int p = choose_partition(mi_row, mi_col, bsize);
    // at the frame's right / bottom edge the syntax only offers SPLIT or
    // the one partition that fits (decoder: has_rows / has_cols)
    const uint32_t mi_cols = (w_ + 7) >> 3, mi_rows = (h_ + 7) >> 3;
    const uint32_t half = VP9_NUM_8X8_WIDE[bsize] >> 1;
    const bool has_rows = mi_row + half < mi_rows, has_cols = mi_col + half < mi_cols;
    if (!has_rows && has_cols && p != PARTITION_SPLIT) p = PARTITION_HORZ;
    if (has_rows && !has_cols && p != PARTITION_SPLIT) p = PARTITION_VERT;
    // 4:2:2 has no chroma for blocks taller than wide, 4:4:0 none for wider
    if (p == PARTITION_VERT && c_.ss_x && !c_.ss_y) p = PARTITION_SPLIT;
    if (p == PARTITION_HORZ && c_.ss_y && !c_.ss_x) p = PARTITION_SPLIT;
    return p;
  }

  int choose_partition(uint32_t mi_row, uint32_t mi_col, int bsize) {
    int n = VP9_NUM_8X8_WIDE[bsize] * 8;  // luma samples
    double x0 = mi_col * 8.0, y0 = mi_row * 8.0;
    if (bsize == BLOCK_8X8) return PARTITION_NONE;  // no blocks below 8x8
    if (key_) {
      // intra: 64 rarely, mostly 32/16/8 by a pseudo-texture
      std::uniform_int_distribution<int> u(0, 99);
      int r = u(rng_);
      if (bsize == BLOCK_64X64) return r < 10 ? PARTITION_NONE : PARTITION_SPLIT;
      if (bsize == BLOCK_32X32) return r < 35 ? PARTITION_NONE : r < 45 ? PARTITION_HORZ : PARTITION_SPLIT;
      return r < 50 ? PARTITION_NONE : r < 60 ? PARTITION_VERT : PARTITION_SPLIT;
    }
    if (c_.content == 0) {
      std::uniform_int_distribution<int> u(0, 99);
      int r = u(rng_);
      if (bsize == BLOCK_64X64) return r < 30 ? PARTITION_NONE : PARTITION_SPLIT;
      return r < 60 ? PARTITION_NONE : PARTITION_SPLIT;
    }
    // homogeneous quadrants: the object at each corner and the centre
    auto obj = [&](double fx, double fy) { return object_at(x0 + fx * n, y0 + fy * n); };
    int tl = obj(0.1, 0.1), tr = obj(0.9, 0.1), bl = obj(0.1, 0.9), br = obj(0.9, 0.9), cc = obj(0.5, 0.5);
    if (tl == tr && tl == bl && tl == br && tl == cc) return PARTITION_NONE;
    if (tl == tr && bl == br) return PARTITION_HORZ;
    if (tl == bl && tr == br) return PARTITION_VERT;
    return PARTITION_SPLIT;
  }

  void block(uint32_t mi_row, uint32_t mi_col, int bsize, bool frame_is_intra, PlanBlock *o) override {
    std::uniform_int_distribution<int> u(0, 99);
    int bw = VP9_NUM_8X8_WIDE[bsize] * 8, bh = VP9_NUM_8X8_HIGH[bsize] * 8;
    double cx = mi_col * 8.0 + bw / 2.0, cy = mi_row * 8.0 + bh / 2.0;
    int max_tx = VP9_MAX_TXSIZE_LOOKUP[bsize];
    o->tx_size = (uint8_t)(u(rng_) < 70 ? max_tx : (max_tx > 0 ? max_tx - 1 : 0));
    o->interp_filter = EIGHTTAP;
    o->segment_id = segment_of(mi_row, mi_col);
    // segment 4 forces the LAST reference (an inter block), segment 5 skip
    const bool seg_ref = c_.segmentation && o->segment_id == 4;
    const bool seg_skip = c_.segmentation && o->segment_id == 5;
    bool intra = !seg_ref && (frame_is_intra || (int)u(rng_) < (int)c_.intra_percent || uncovered(cx, cy));
    cur_intra_ = intra;
    if (intra) {
      static const uint8_t modes[10] = {DC_PRED, DC_PRED, DC_PRED, TM_PRED, TM_PRED, V_PRED, H_PRED,
                                        D45_PRED, D135_PRED, D207_PRED};
      static const uint8_t extra[4] = {D117_PRED, D153_PRED, D63_PRED, V_PRED};
      int r = u(rng_);
      o->y_mode = r < 90 ? modes[r % 10] : extra[r % 4];
      for (int k = 0; k < 4; k++) o->sub_modes[k] = o->y_mode;
      o->uv_mode = u(rng_) < 70 ? DC_PRED : o->y_mode;
      o->is_inter = 0;
      o->ref_frame[0] = INTRA_FRAME;
      o->ref_frame[1] = REF_NONE;
      o->skip = u(rng_) < 5 || seg_skip;
      cur_skip_ = o->skip;
      return;
    }
    o->is_inter = 1;
    o->ref_frame[0] = (!seg_ref && c_.golden_interval && u(rng_) < 8) ? GOLDEN_FRAME : LAST_FRAME;
    o->ref_frame[1] = REF_NONE;
    // motion: content moving by (vx, vy) per frame predicts from (-vx, -vy)
    double vx = bgx_, vy = bgy_;
    if (c_.content == 1) {
      int obj = object_at(cx, cy);
      if (obj >= 0) {
        vx = rects_[obj].vx;
        vy = rects_[obj].vy;
      }
    }
    double frames_back = o->ref_frame[0] == GOLDEN_FRAME ? 1.0 + (t_ % std::max(1u, c_.golden_interval)) : 1.0;
    o->mv[0][0] = (int32_t)std::lround(-vy * 8.0 * frames_back);
    o->mv[0][1] = (int32_t)std::lround(-vx * 8.0 * frames_back);
    bool boundary = partition_boundary(mi_row, mi_col, bw, bh);
    o->skip = (!boundary && u(rng_) < (c_.content == 0 ? 20 : 60)) || seg_skip;
    cur_skip_ = o->skip;
  }

  bool partition_boundary(uint32_t mi_row, uint32_t mi_col, int bw, int bh) const {
    if (c_.content != 1) return false;
    double x0 = mi_col * 8.0, y0 = mi_row * 8.0;
    int a = object_at(x0 + 0.5, y0 + 0.5), b = object_at(x0 + bw - 0.5, y0 + bh - 0.5);
    return a != b;
  }

  void coefs(int plane, uint32_t, uint32_t, int tx_size, int tx_type, int32_t *raster) override {
    std::uniform_int_distribution<int> u(0, 99);
    const int n0 = 4 << tx_size, nn = n0 * n0;
    // how many nonzero coefficients, and their scale
    int p_any = cur_intra_ ? (plane ? 70 : 95) : (plane ? 25 : 55);
    if (c_.content == 0) p_any = cur_intra_ ? 95 : (plane ? 60 : 85);
    if (u(rng_) >= p_any) return;
    double mean_k = cur_intra_ ? 6.0 + 2.0 * tx_size : 1.5 + 1.0 * tx_size;
    std::geometric_distribution<int> gk(1.0 / mean_k);
    int k = 1 + gk(rng_);
    if (k > nn) k = nn;
    double scale = cur_intra_ ? 6.0 : 1.6;
    scale *= 100.0 / std::max(40u, c_.base_q_idx);
    std::exponential_distribution<double> lap(1.0 / scale);
    const uint16_t *scan = scan_of(tx_size, tx_type);
    // positions: the first scan indices mostly, spread over about 3k
    int span = std::min(nn, 3 * k + 2);
    std::uniform_int_distribution<int> pos(0, span - 1);
    for (int i = 0; i < k; i++) {
      int s = i == 0 ? 0 : pos(rng_);
      int v = 1 + (int)lap(rng_);
      if (s == 0 && cur_intra_) v += (int)lap(rng_) * 4;
      int maxv = 1 << (c_.bit_depth + 6);
      if (v > maxv) v = maxv;
      raster[scan[s]] = u(rng_) < 50 ? v : -v;
    }
  }

 private:
  const sg_config &c_;
  uint32_t w_, h_;
  int t_;
  bool key_;
  const std::vector<Rect> &rects_;
  double bgx_, bgy_;
  std::mt19937_64 &rng_;
  bool seg_map_ = false;
  bool cur_intra_ = false, cur_skip_ = false;
};

struct Gen {
  sg_config c;
  Parser dec;
  std::mt19937_64 rng;
  std::vector<Rect> rects;
  double bgx = 0, bgy = 0;
  uint32_t t = 0;
  std::vector<uint8_t> frame;
  uint32_t slot_w[8] = {}, slot_h[8] = {};

  explicit Gen(const sg_config &cfg) : c(cfg), rng(cfg.seed * 0x9e3779b97f4a7c15ull + 1) {
    if (c.content == 1) {
      std::uniform_real_distribution<double> side(16, 256), px(0, c.width), py(0, c.height);
      std::uniform_int_distribution<int> vel(-64, 64);  // 1/8 sample per frame
      for (int i = 0; i < 64; i++) {
        Rect r;
        r.w = side(rng);
        r.h = side(rng);
        r.x = px(rng) - r.w / 2;
        r.y = py(rng) - r.h / 2;
        r.vx = vel(rng) / 8.0;
        r.vy = vel(rng) / 8.0;
        rects.push_back(r);
      }
      bgx = 0.25;  // slow pan
      bgy = -0.125;
    } else {
      bgx = 1.375;  // the sinusoid grid's translation per frame
      bgy = -0.625;
    }
  }

  uint32_t profile() const {
    bool non420 = !(c.ss_x == 1 && c.ss_y == 1);
    return (c.bit_depth > 8 ? 2 : 0) + (non420 ? 1 : 0);
  }

  // Half size on every resize_every-th frame (not the first).
  void frame_size(uint32_t i, uint32_t *w, uint32_t *h) const {
    *w = c.width;
    *h = c.height;
    if (c.resize_every && i > 0 && i % c.resize_every == 0) {
      *w = (c.width + 1) / 2;
      *h = (c.height + 1) / 2;
    }
  }

  // Segmentation (6.2.11): the map is written on key frames and on two of
  // every three inter frames.
  bool seg_map(bool key) const { return key || t % 3 != 2; }
  // Features per segment, delta or absolute: 1 and 7 quantiser only, 2 and
  // 6 quantiser and filter level, 3 filter level only, 4 the LAST reference
  // (inter frames), 5 skip, 0 none.
  static void seg_features(bool key, bool abs, bool on[8][4], int16_t data[8][4]) {
    memset(on, 0, sizeof(bool) * 32);
    memset(data, 0, sizeof(int16_t) * 32);
    auto set = [&](int s, int f, int v) {
      on[s][f] = true;
      data[s][f] = (int16_t)v;
    };
    if (abs) {
      set(1, 0, 60);
      set(2, 0, 140);
      set(2, 1, 10);
      set(3, 1, 36);
      set(6, 0, 90);
      set(6, 1, 0);
      set(7, 0, 200);
    } else {
      set(1, 0, -24);
      set(2, 0, 40);
      set(2, 1, -12);
      set(3, 1, 9);
      set(6, 0, -8);
      set(6, 1, 20);
      set(7, 0, 60);
    }
    if (!key) set(4, 2, 1);  // LAST_FRAME
    set(5, 3, 0);
  }

  std::vector<uint8_t> header(bool key, uint32_t w, uint32_t h, uint32_t refresh, uint32_t lf_level) {
    BitWriter b;
    uint32_t prof = profile();
    b.f(2, 2);
    b.f(1, prof & 1);
    b.f(1, prof >> 1);
    if (prof == 3) b.f(1, 0);
    b.f(1, 0);            // show_existing_frame
    b.f(1, key ? 0 : 1);  // frame_type
    b.f(1, 1);            // show_frame
    b.f(1, key ? 0 : 1);  // error_resilient_mode
    auto frame_size_bits = [&]() {
      b.f(16, w - 1);
      b.f(16, h - 1);
    };
    if (key) {
      b.f(8, 0x49);
      b.f(8, 0x83);
      b.f(8, 0x42);
      if (prof >= 2) b.f(1, c.bit_depth == 12);
      b.f(3, 1);  // BT.601
      b.f(1, 0);  // color_range
      if (prof == 1 || prof == 3) {
        b.f(1, c.ss_x);
        b.f(1, c.ss_y);
        b.f(1, 0);
      }
      frame_size_bits();
      b.f(1, 0);  // render_and_frame_size_different
    } else {
      b.f(8, refresh);
      for (int i = 0; i < 3; i++) {
        b.f(3, i);  // ref_frame_idx
        b.f(1, 0);  // sign bias
      }
      int found = -1;
      for (int i = 0; i < 3 && found < 0; i++) {
        bool same = slot_w[i] == w && slot_h[i] == h;
        b.f(1, same);
        if (same) found = i;
      }
      if (found < 0) frame_size_bits();
      b.f(1, 0);  // render_and_frame_size_different
      b.f(1, c.allow_hp);
      b.f(1, 0);  // is_filter_switchable
      b.f(2, 1);  // literal_to_type[1] = EIGHTTAP
    }
    if (key) {
      b.f(1, 0);  // refresh_frame_context
      b.f(1, 1);  // frame_parallel_decoding_mode
    }
    b.f(2, 0);  // frame_context_idx
    b.f(6, lf_level);
    b.f(3, c.lf_sharpness);
    b.f(1, 1);  // mode_ref_delta_enabled
    b.f(1, 0);  // mode_ref_delta_update
    b.f(8, c.base_q_idx);
    b.f(1, 0);
    b.f(1, 0);
    b.f(1, 0);  // no delta_q
    if (!c.segmentation) {
      b.f(1, 0);  // segmentation_enabled
    } else {
      b.f(1, 1);
      const bool map = seg_map(key);
      b.f(1, map);  // update_map
      if (map) {
        for (int i = 0; i < 7; i++) {
          const bool coded = (i + t) % 3 != 0;
          b.f(1, coded);
          if (coded) b.f(8, 64 + ((i * 37 + t * 11) % 128));
        }
        const bool temporal = !key && (t & 1);
        b.f(1, temporal);  // temporal_update
        if (temporal)
          for (int i = 0; i < 3; i++) {
            b.f(1, 1);
            b.f(8, 160 + 20 * i);
          }
      }
      b.f(1, 1);  // update_data (every frame: error resilience clears the features)
      const bool abs = t & 1;
      b.f(1, abs);
      int16_t data[8][4];
      bool on[8][4];
      seg_features(key, abs, on, data);
      static const int bits[4] = {8, 6, 2, 0}, sgn[4] = {1, 1, 0, 0};
      for (int i = 0; i < 8; i++)
        for (int j = 0; j < 4; j++) {
          b.f(1, on[i][j]);
          if (!on[i][j]) continue;
          const int v = data[i][j];
          if (bits[j]) b.f(bits[j], (uint32_t)(v < 0 ? -v : v));
          if (sgn[j]) b.f(1, v < 0);
        }
    }
    uint32_t sb64_cols = (((w + 7) >> 3) + 7) >> 3;
    uint32_t min_log2 = 0, max_log2 = 1;
    while ((64u << min_log2) < sb64_cols) min_log2++;
    while ((sb64_cols >> max_log2) >= 4) max_log2++;
    max_log2--;
    uint32_t tl = std::max(min_log2, std::min(c.tile_cols_log2, max_log2));
    for (uint32_t k = min_log2; k < tl; k++) b.f(1, 1);
    if (tl < max_log2) b.f(1, 0);
    b.f(1, 0);   // tile_rows_log2
    b.f(16, 1);  // header_size_in_bytes, patched by the writer
    b.align();
    return b.out;
  }

  // The next frame; returns its size (0 when done).
  size_t next() {
    if (t >= c.frames) return 0;
    bool key = t == 0;
    uint32_t w, h;
    frame_size(t, &w, &h);
    uint32_t refresh = key ? 0xff : (1u | ((c.golden_interval && t % c.golden_interval == 0) ? 2u : 0u));
    uint32_t lf_level = c.lf_level ? std::min(63u, c.lf_level - 1) : std::min(63u, (c.base_q_idx * 10 + 32) / 64);
    std::vector<uint8_t> hdr = header(key, w, h, refresh, lf_level);
    Plan plan(c, w, h, (int)t, key, rects, bgx, bgy, rng, seg_map(key));
    Parser sim = dec;
    frame = sim.write_frame(hdr.data(), hdr.size(), TX_MODE_SELECT, SINGLE_REFERENCE, &plan);
    vp9r_record rec;
    dec.parse_frame(frame.data(), frame.size(), &rec);  // advances the decoder; throws if unreadable
    for (int i = 0; i < 8; i++)
      if ((refresh >> i) & 1) {
        slot_w[i] = w;
        slot_h[i] = h;
      }
    t++;
    return frame.size();
  }
};

thread_local std::string g_err;

}  // namespace

extern "C" {

void *sg_new(const sg_config *cfg) { return new Gen(*cfg); }
void sg_free(void *g) { delete (Gen *)g; }
const char *sg_error(void) { return g_err.c_str(); }

// The scan order of transform size tx_size (0..3) and type tx_type: out
// receives (4 << tx_size)^2 raster positions in scan order.
void sg_scan(int tx_size, int tx_type, uint16_t *out) {
  const int n = 4 << tx_size;
  memcpy(out, scan_of(tx_size, tx_type), sizeof(uint16_t) * n * n);
}

// Writes the next frame; returns its length (0 at the end, -1 on error) and
// points *data at it (valid until the next call).
long sg_next(void *gp, const uint8_t **data) {
  Gen *g = (Gen *)gp;
  try {
    size_t n = g->next();
    *data = g->frame.data();
    return (long)n;
  } catch (const ParseError &e) {
    g_err = e.msg;
    return -1;
  }
}

}  // extern "C"
