// The CPU parser: see parser.h.  Section numbers are the VP9 specification's;
// file:line citations are the reference crate's (safewords/rivet-vp9).
#include "parser.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "vp9_consts.h"
#include "vp9_tables.h"

namespace vp9r {

[[noreturn]] static void fail(int code, const char *msg) { throw ParseError{code, msg}; }
[[noreturn]] static void bitstream(const char *msg) { fail(VP9R_ERR_BITSTREAM, msg); }

// --------------------------------------------------------------------------
// f(n) reader of the uncompressed header (src/bits.rs:7-58).

class BitReader {
 public:
  BitReader(const uint8_t *d, size_t n) : d_(d), n_(n) {}
  uint32_t bit() {
    size_t byte = pos_ >> 3;
    if (byte >= n_) bitstream("frame data ends inside the uncompressed header");
    uint32_t b = (d_[byte] >> (7 - (pos_ & 7))) & 1;
    pos_++;
    return b;
  }
  uint32_t f(int n) {
    uint32_t x = 0;
    for (int i = 0; i < n; i++) x = (x << 1) | bit();
    return x;
  }
  bool flag() { return bit() != 0; }
  int32_t s(int n) {
    int32_t v = (int32_t)f(n);
    return flag() ? -v : v;
  }
  void byte_align() { pos_ = (pos_ + 7) & ~(size_t)7; }
  size_t position() const { return pos_; }

 private:
  const uint8_t *d_;
  size_t n_;
  size_t pos_ = 0;
};

// --------------------------------------------------------------------------
// Boolean decoder (9.2; src/bool_coder.rs:12-127).  `win` holds BoolValue in
// its top byte and the following stream bits below it; `avail` counts the
// valid bits of `win` from the top.  Bits past the end of the data are zero.

class BoolDecoder {
 public:
  static constexpr bool kWrite = false;
  BoolDecoder(const uint8_t *d, size_t n) : p_(d), end_(d + n) {
    if (n == 0) bitstream("bool-coded data has no bytes");
    refill();
    if (read(128)) bitstream("bool decoder: the first decoded bit must be 0");
  }
  inline bool read(uint32_t prob) {
    uint32_t split = 1 + (((range_ - 1) * prob) >> 8);
    uint64_t big = (uint64_t)split << 56;
    bool bit;
    if (win_ < big) {
      range_ = split;
      bit = false;
    } else {
      range_ -= split;
      win_ -= big;
      bit = true;
    }
    int shift = __builtin_clz(range_) - 24;
    range_ <<= shift;
    win_ <<= shift;
    avail_ -= shift;
    if (avail_ < 16) refill();
    return bit;
  }
  uint32_t literal(int n, uint32_t = 0) {
    uint32_t x = 0;
    for (int i = 0; i < n; i++) x = (x << 1) | (uint32_t)read(128);
    return x;
  }
  // The syntax walk calls bit/tree with the value a writer would code; a
  // reader ignores it.
  inline bool bit(uint32_t prob, bool = false) { return read(prob); }
  // tree decode (9.3.3): probability of node n>>1 from prob(node)
  template <typename F>
  inline int tree(const int8_t *t, F prob, int = 0) {
    int n = 0;
    do {
      n = t[n + (int)read(prob(n >> 1))];
    } while (n > 0);
    return -n;
  }
  // exit_bool (9.2.3): the padding after the last symbol must be zero
  bool padding_is_zero() const {
    if ((win_ & ((1ull << 56) - 1)) != 0) return false;
    for (const uint8_t *q = p_; q < end_; q++)
      if (*q) return false;
    return true;
  }

 private:
  void refill() {
    while (avail_ <= 56) {
      if (p_ < end_) {
        win_ |= (uint64_t)(*p_++) << (56 - avail_);
        avail_ += 8;
      } else {
        avail_ = 64;  // zeros from here on
        break;
      }
    }
  }
  const uint8_t *p_, *end_;
  uint64_t win_ = 0;
  int avail_ = 0;
  uint32_t range_ = 255;
};

// --------------------------------------------------------------------------
// Probability contexts and adaptation (src/probs.rs).

void Probs::set_defaults() {
  memcpy(tx, VP9_DEFAULT_TX_PROBS, sizeof tx);
  memcpy(coef, VP9_DEFAULT_COEF_PROBS, sizeof coef);
  memcpy(skip, VP9_DEFAULT_SKIP_PROB, sizeof skip);
  memcpy(inter_mode, VP9_DEFAULT_INTER_MODE_PROBS, sizeof inter_mode);
  memcpy(interp_filter, VP9_DEFAULT_INTERP_FILTER_PROBS, sizeof interp_filter);
  memcpy(is_inter, VP9_DEFAULT_IS_INTER_PROB, sizeof is_inter);
  memcpy(comp_mode, VP9_DEFAULT_COMP_MODE_PROB, sizeof comp_mode);
  memcpy(single_ref, VP9_DEFAULT_SINGLE_REF_PROB, sizeof single_ref);
  memcpy(comp_ref, VP9_DEFAULT_COMP_REF_PROB, sizeof comp_ref);
  memcpy(y_mode, VP9_DEFAULT_Y_MODE_PROBS, sizeof y_mode);
  memcpy(uv_mode, VP9_DEFAULT_UV_MODE_PROBS, sizeof uv_mode);
  memcpy(partition, VP9_DEFAULT_PARTITION_PROBS, sizeof partition);
  memcpy(mv_joint, VP9_DEFAULT_MV_JOINT_PROBS, sizeof mv_joint);
  memcpy(mv_sign, VP9_DEFAULT_MV_SIGN_PROB, sizeof mv_sign);
  memcpy(mv_class, VP9_DEFAULT_MV_CLASS_PROBS, sizeof mv_class);
  memcpy(mv_class0_bit, VP9_DEFAULT_MV_CLASS0_BIT_PROB, sizeof mv_class0_bit);
  memcpy(mv_bits, VP9_DEFAULT_MV_BITS_PROB, sizeof mv_bits);
  memcpy(mv_class0_fr, VP9_DEFAULT_MV_CLASS0_FR_PROBS, sizeof mv_class0_fr);
  memcpy(mv_fr, VP9_DEFAULT_MV_FR_PROBS, sizeof mv_fr);
  memcpy(mv_class0_hp, VP9_DEFAULT_MV_CLASS0_HP_PROB, sizeof mv_class0_hp);
  memcpy(mv_hp, VP9_DEFAULT_MV_HP_PROB, sizeof mv_hp);
}

// merge_prob (8.4.1)
static uint8_t merge_prob(uint8_t pre, uint32_t ct0, uint32_t ct1, uint32_t sat, uint32_t upd) {
  uint32_t den = ct0 + ct1;
  uint32_t prob = 128;
  if (den) {
    uint64_t p = ((uint64_t)ct0 * 256 + (den >> 1)) / den;
    prob = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(p, 1), 255);
  }
  uint32_t count = std::min(den, sat);
  uint32_t factor = upd * count / sat;
  return (uint8_t)((pre * (256 - factor) + prob * factor + 128) >> 8);
}

// merge_probs (8.4.2): adapts probs along `tree` from node index i
static uint32_t merge_probs(const int8_t *tree, int i, uint8_t *probs, const uint32_t *counts,
                            uint32_t sat, uint32_t upd) {
  int s = tree[i];
  uint32_t left = s <= 0 ? counts[-s] : merge_probs(tree, s, probs, counts, sat, upd);
  int r = tree[i + 1];
  uint32_t right = r <= 0 ? counts[-r] : merge_probs(tree, r, probs, counts, sat, upd);
  probs[i >> 1] = merge_prob(probs[i >> 1], left, right, sat, upd);
  return left + right;
}

static void adapt_tree(const int8_t *tree, uint8_t *probs, const uint32_t *counts) {
  merge_probs(tree, 0, probs, counts, 20, 128);
}
static void adapt_bool(uint8_t *prob, const uint32_t *counts) {
  *prob = merge_prob(*prob, counts[0], counts[1], 20, 128);
}

// adapt_coef_probs (8.4.3)
static void adapt_coef_probs(Probs &p, const Counts &c, bool frame_is_intra, bool last_was_key) {
  uint32_t upd = frame_is_intra ? 112 : (last_was_key ? 128 : 112);
  for (int t = 0; t < 4; t++)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int k = 0; k < 6; k++)
          for (int l = 0; l < (k == 0 ? 3 : 6); l++) {
            uint8_t *pr = p.coef[t][i][j][k][l];
            merge_probs(VP9_SMALL_TOKEN_TREE, 2, pr, c.token[t][i][j][k][l], 24, upd);
            merge_probs(VP9_BINARY_TREE, 0, pr, c.more_coefs[t][i][j][k][l], 24, upd);
          }
}

// adapt_noncoef_probs (8.4.4)
static void adapt_noncoef_probs(Probs &p, const Counts &c, bool switchable, bool tx_select,
                                bool allow_hp) {
  for (int i = 0; i < 4; i++) adapt_bool(&p.is_inter[i], c.is_inter[i]);
  for (int i = 0; i < 5; i++) adapt_bool(&p.comp_mode[i], c.comp_mode[i]);
  for (int i = 0; i < 5; i++) adapt_bool(&p.comp_ref[i], c.comp_ref[i]);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 2; j++) adapt_bool(&p.single_ref[i][j], c.single_ref[i][j]);
  for (int i = 0; i < 7; i++) adapt_tree(VP9_INTER_MODE_TREE, p.inter_mode[i], c.inter_mode[i]);
  for (int i = 0; i < 4; i++) adapt_tree(VP9_INTRA_MODE_TREE, p.y_mode[i], c.intra_mode[i]);
  for (int i = 0; i < 10; i++) adapt_tree(VP9_INTRA_MODE_TREE, p.uv_mode[i], c.uv_mode[i]);
  for (int i = 0; i < 16; i++) adapt_tree(VP9_PARTITION_TREE, p.partition[i], c.partition[i]);
  for (int i = 0; i < 3; i++) adapt_bool(&p.skip[i], c.skip[i]);
  if (switchable)
    for (int i = 0; i < 4; i++)
      adapt_tree(VP9_INTERP_FILTER_TREE, p.interp_filter[i], c.interp_filter[i]);
  if (tx_select)
    for (int i = 0; i < 2; i++) {
      adapt_tree(VP9_TX_SIZE_8_TREE, p.tx[1][i], c.tx[1][i]);
      adapt_tree(VP9_TX_SIZE_16_TREE, p.tx[2][i], c.tx[2][i]);
      adapt_tree(VP9_TX_SIZE_32_TREE, p.tx[3][i], c.tx[3][i]);
    }
  adapt_tree(VP9_MV_JOINT_TREE, p.mv_joint, c.mv_joint);
  for (int i = 0; i < 2; i++) {
    adapt_bool(&p.mv_sign[i], c.mv_sign[i]);
    adapt_tree(VP9_MV_CLASS_TREE, p.mv_class[i], c.mv_class[i]);
    adapt_bool(&p.mv_class0_bit[i], c.mv_class0_bit[i]);
    for (int j = 0; j < 10; j++) adapt_bool(&p.mv_bits[i][j], c.mv_bits[i][j]);
    for (int j = 0; j < 2; j++) adapt_tree(VP9_MV_FR_TREE, p.mv_class0_fr[i][j], c.mv_class0_fr[i][j]);
    adapt_tree(VP9_MV_FR_TREE, p.mv_fr[i], c.mv_fr[i]);
    if (allow_hp) {
      adapt_bool(&p.mv_class0_hp[i], c.mv_class0_hp[i]);
      adapt_bool(&p.mv_hp[i], c.mv_hp[i]);
    }
  }
}

// --------------------------------------------------------------------------
// Uncompressed header (6.2; src/header.rs:245-469).

enum { RESET_NONE = 0, RESET_ALL = 1, RESET_ONE = 2 };

static void frame_sync_code(BitReader &r) {
  uint32_t a = r.f(8), b = r.f(8), c = r.f(8);
  if (a != 0x49 || b != 0x83 || c != 0x42) bitstream("frame sync bytes are not 0x49 0x83 0x42");
}

static void color_config(BitReader &r, FrameHeader &h) {
  h.bit_depth = h.profile >= 2 ? (r.flag() ? 12 : 10) : 8;
  h.color_space = r.f(3);
  if (h.color_space != 7) {  // not sRGB
    h.color_range = r.flag();
    if (h.profile == 1 || h.profile == 3) {
      h.ss_x = r.f(1);
      h.ss_y = r.f(1);
      if (r.flag()) bitstream("a reserved bit is set");
      if (h.ss_x == 1 && h.ss_y == 1) bitstream("profiles 1 and 3 carry no 4:2:0 video");
    } else {
      h.ss_x = h.ss_y = 1;
    }
  } else {
    h.color_range = true;
    if (h.profile == 1 || h.profile == 3) {
      h.ss_x = h.ss_y = 0;
      if (r.flag()) bitstream("a reserved bit is set");
    } else {
      bitstream("profiles 0 and 2 carry no RGB video");
    }
  }
}

static void frame_size(BitReader &r, FrameHeader &h) {
  h.width = r.f(16) + 1;
  h.height = r.f(16) + 1;
  h.compute_image_size();
}

static void render_size(BitReader &r, FrameHeader &h) {
  if (r.flag()) {
    h.render_width = r.f(16) + 1;
    h.render_height = r.f(16) + 1;
  } else {
    h.render_width = h.width;
    h.render_height = h.height;
  }
}

static int32_t read_delta_q(BitReader &r) { return r.flag() ? r.s(4) : 0; }
static uint8_t read_prob(BitReader &r) { return r.flag() ? (uint8_t)r.f(8) : 255; }

void Parser::parse_uncompressed(const uint8_t *data, size_t len, int *reset_kind) {
  BitReader r(data, len);
  FrameHeader &h = h_;
  h = FrameHeader();
  *reset_kind = RESET_NONE;
  if (r.f(2) != 2) bitstream("the frame marker bits are not binary 10");
  uint32_t lo = r.f(1), hi = r.f(1);
  h.profile = (hi << 1) + lo;
  if (h.profile == 3 && r.flag()) bitstream("a reserved bit is set");
  h.show_existing_frame = r.flag();
  if (h.show_existing_frame) {
    h.frame_to_show_map_idx = r.f(3);
    r.byte_align();
    h.uncompressed_size = r.position() / 8;
    return;
  }
  h.frame_type = r.f(1);
  h.show_frame = r.flag();
  h.error_resilient_mode = r.flag();
  h.last_frame_type = st_last_frame_type_;
  st_last_frame_type_ = h.frame_type;
  h.bit_depth = st_bit_depth_;
  h.ss_x = st_ss_x_;
  h.ss_y = st_ss_y_;
  h.color_space = st_color_space_;
  h.color_range = st_color_range_;
  if (h.frame_type == KEY_FRAME) {
    frame_sync_code(r);
    color_config(r, h);
    frame_size(r, h);
    render_size(r, h);
    h.refresh_frame_flags = 0xff;
    h.frame_is_intra = true;
  } else {
    h.intra_only = h.show_frame ? false : r.flag();
    h.frame_is_intra = h.intra_only;
    h.reset_frame_context = h.error_resilient_mode ? 0 : r.f(2);
    if (h.intra_only) {
      frame_sync_code(r);
      if (h.profile > 0) {
        color_config(r, h);
      } else {
        h.color_space = 1;  // BT.601
        h.ss_x = h.ss_y = 1;
        h.bit_depth = 8;
      }
      h.refresh_frame_flags = r.f(8);
      frame_size(r, h);
      render_size(r, h);
    } else {
      h.refresh_frame_flags = r.f(8);
      for (int i = 0; i < 3; i++) {
        h.ref_frame_idx[i] = r.f(3);
        h.ref_frame_sign_bias[LAST_FRAME + i] = r.flag();
      }
      bool found = false;
      for (int i = 0; i < 3; i++) {
        if (r.flag()) {
          uint32_t w = ref_sizes_[h.ref_frame_idx[i]][0], ht = ref_sizes_[h.ref_frame_idx[i]][1];
          if (w == 0) bitstream("size copied from a reference slot that holds no frame");
          h.width = w;
          h.height = ht;
          found = true;
          break;
        }
      }
      if (!found)
        frame_size(r, h);
      else
        h.compute_image_size();
      render_size(r, h);
      h.allow_high_precision_mv = r.flag();
      h.interp_filter = r.flag() ? SWITCHABLE : VP9_LITERAL_TO_TYPE[r.f(2)];
    }
  }
  st_bit_depth_ = h.bit_depth;
  st_ss_x_ = h.ss_x;
  st_ss_y_ = h.ss_y;
  st_color_space_ = h.color_space;
  st_color_range_ = h.color_range;
  if (!h.error_resilient_mode) {
    h.refresh_frame_context = r.flag();
    h.frame_parallel_decoding_mode = r.flag();
  } else {
    h.refresh_frame_context = false;
    h.frame_parallel_decoding_mode = true;
  }
  h.frame_context_idx = r.f(2);
  if (h.frame_is_intra || h.error_resilient_mode) {
    // setup_past_independence()
    h.past_independence = true;
    memset(seg_.feature_data, 0, sizeof seg_.feature_data);
    memset(seg_.feature_enabled, 0, sizeof seg_.feature_enabled);
    seg_.abs_or_delta_update = false;
    lf_.delta_enabled = true;
    lf_.ref_deltas[0] = 1;
    lf_.ref_deltas[1] = 0;
    lf_.ref_deltas[2] = -1;
    lf_.ref_deltas[3] = -1;
    lf_.mode_deltas[0] = lf_.mode_deltas[1] = 0;
    if (h.frame_type == KEY_FRAME || h.error_resilient_mode || h.reset_frame_context == 3)
      *reset_kind = RESET_ALL;
    else if (h.reset_frame_context == 2)
      *reset_kind = RESET_ONE;  // of the frame_context_idx just read
    reset_idx_ = h.frame_context_idx;
    h.frame_context_idx = 0;
  }
  // loop_filter_params()
  lf_.level = (uint8_t)r.f(6);
  lf_.sharpness = (uint8_t)r.f(3);
  lf_.delta_enabled = r.flag();
  if (lf_.delta_enabled && r.flag()) {
    for (int i = 0; i < 4; i++)
      if (r.flag()) lf_.ref_deltas[i] = (int8_t)r.s(6);
    for (int i = 0; i < 2; i++)
      if (r.flag()) lf_.mode_deltas[i] = (int8_t)r.s(6);
  }
  // quantization_params()
  h.base_q_idx = (int32_t)r.f(8);
  h.delta_q_y_dc = read_delta_q(r);
  h.delta_q_uv_dc = read_delta_q(r);
  h.delta_q_uv_ac = read_delta_q(r);
  h.lossless = h.base_q_idx == 0 && h.delta_q_y_dc == 0 && h.delta_q_uv_dc == 0 &&
               h.delta_q_uv_ac == 0;
  // segmentation_params()
  seg_.enabled = r.flag();
  seg_.update_map = false;
  seg_.temporal_update = false;
  if (seg_.enabled) {
    seg_.update_map = r.flag();
    if (seg_.update_map) {
      for (int i = 0; i < 7; i++) seg_.tree_probs[i] = read_prob(r);
      seg_.temporal_update = r.flag();
      for (int i = 0; i < 3; i++) seg_.pred_probs[i] = seg_.temporal_update ? read_prob(r) : 255;
    }
    if (r.flag()) {
      seg_.abs_or_delta_update = r.flag();
      for (int i = 0; i < MAX_SEGMENTS; i++)
        for (int j = 0; j < 4; j++) {
          int16_t v = 0;
          bool en = r.flag();
          seg_.feature_enabled[i][j] = en;
          if (en) {
            v = (int16_t)r.f(VP9_SEGMENTATION_FEATURE_BITS[j]);
            if (VP9_SEGMENTATION_FEATURE_SIGNED[j] == 1 && r.flag()) v = (int16_t)-v;
          }
          seg_.feature_data[i][j] = v;
        }
    }
  }
  // tile_info()
  uint32_t min_log2 = 0;
  while (((uint32_t)MAX_TILE_WIDTH_B64 << min_log2) < h.sb64_cols) min_log2++;
  uint32_t max_log2 = 1;
  while ((h.sb64_cols >> max_log2) >= (uint32_t)MIN_TILE_WIDTH_B64) max_log2++;
  max_log2--;
  h.tile_cols_log2 = min_log2;
  while (h.tile_cols_log2 < max_log2) {
    if (r.flag())
      h.tile_cols_log2++;
    else
      break;
  }
  h.tile_rows_log2 = r.f(1);
  if (h.tile_rows_log2 == 1) h.tile_rows_log2 += r.f(1);
  h.header_size_pos = r.position();
  h.header_size_in_bytes = r.f(16);
  r.byte_align();
  h.uncompressed_size = r.position() / 8;
}

// --------------------------------------------------------------------------
// Compressed header (6.3; src/header.rs:470-680).

template <class Coder>
static uint32_t decode_term_subexp(Coder &d) {
  if (d.literal(1) == 0) return d.literal(4);
  if (d.literal(1) == 0) return d.literal(4) + 16;
  if (d.literal(1) == 0) return d.literal(5) + 32;
  uint32_t v = d.literal(7);
  if (v < 65) return v + 64;
  uint32_t bit = d.literal(1);
  return (v << 1) - 1 + bit;
}

static int32_t inv_recenter_nonneg(int32_t v, int32_t m) {
  if (v > 2 * m) return v;
  if (v & 1) return m - ((v + 1) >> 1);
  return m + (v >> 1);
}

static uint8_t inv_remap_prob(uint32_t delta, uint8_t prob) {
  int32_t v = VP9_INV_MAP_TABLE[std::min<uint32_t>(delta, 254)];
  int32_t m = (int32_t)prob - 1;
  int32_t r;
  if ((m << 1) <= 255)
    r = 1 + inv_recenter_nonneg(v, m);
  else
    r = 255 - inv_recenter_nonneg(v, 255 - 1 - m);
  return (uint8_t)std::min(std::max(r, 1), 255);
}

// (a writer codes "no update" for every probability)
template <class Coder>
static void diff_update_prob(Coder &d, uint8_t *p) {
  if (d.bit(252, false)) *p = inv_remap_prob(decode_term_subexp(d), *p);
}

template <class Coder>
static void update_mv_prob(Coder &d, uint8_t *p) {
  if (d.bit(252, false)) *p = (uint8_t)((d.literal(7) << 1) | 1);
}

void Parser::parse_compressed(const uint8_t *data, size_t len) {
  BoolDecoder d(data, len);
  compressed_syntax(d);
}

template <class Coder>
void Parser::compressed_syntax(Coder &d) {
  FrameHeader &h = h_;
  const uint32_t want_reference_mode = h.reference_mode;
  Probs &p = probs_;
  if (h.lossless) {
    h.tx_mode = ONLY_4X4;
  } else {
    uint32_t want = h.tx_mode;
    h.tx_mode = d.literal(2, std::min<uint32_t>(want, ALLOW_32X32));
    if (h.tx_mode == ALLOW_32X32) h.tx_mode += d.literal(1, want == TX_MODE_SELECT);
  }
  if (h.tx_mode == TX_MODE_SELECT) {
    for (int i = 0; i < 2; i++) diff_update_prob(d, &p.tx[1][i][0]);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) diff_update_prob(d, &p.tx[2][i][j]);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 3; j++) diff_update_prob(d, &p.tx[3][i][j]);
  }
  int max_tx = VP9_TX_MODE_TO_BIGGEST_TX_SIZE[h.tx_mode];
  for (int tx = 0; tx <= max_tx; tx++) {
    if (d.literal(1, 0) == 1)
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
          for (int k = 0; k < 6; k++)
            for (int l = 0; l < (k == 0 ? 3 : 6); l++)
              for (int m = 0; m < 3; m++) diff_update_prob(d, &p.coef[tx][i][j][k][l][m]);
  }
  for (int i = 0; i < 3; i++) diff_update_prob(d, &p.skip[i]);
  h.reference_mode = SINGLE_REFERENCE;
  if (h.frame_is_intra) return;
  for (int i = 0; i < 7; i++)
    for (int j = 0; j < 3; j++) diff_update_prob(d, &p.inter_mode[i][j]);
  if (h.interp_filter == SWITCHABLE)
    for (int j = 0; j < 4; j++)
      for (int i = 0; i < 2; i++) diff_update_prob(d, &p.interp_filter[j][i]);
  for (int i = 0; i < 4; i++) diff_update_prob(d, &p.is_inter[i]);
  // frame_reference_mode()
  const bool *sb = h.ref_frame_sign_bias;
  bool compound_allowed = sb[2] != sb[1] || sb[3] != sb[1];
  if (compound_allowed) {
    uint32_t want = want_reference_mode;
    if (d.literal(1, want != SINGLE_REFERENCE) == 0) {
      h.reference_mode = SINGLE_REFERENCE;
    } else {
      h.reference_mode = d.literal(1, want == REFERENCE_MODE_SELECT) == 0 ? COMPOUND_REFERENCE : REFERENCE_MODE_SELECT;
      // setup_compound_reference_mode()
      if (sb[LAST_FRAME] == sb[GOLDEN_FRAME]) {
        h.comp_fixed_ref = ALTREF_FRAME;
        h.comp_var_ref[0] = LAST_FRAME;
        h.comp_var_ref[1] = GOLDEN_FRAME;
      } else if (sb[LAST_FRAME] == sb[ALTREF_FRAME]) {
        h.comp_fixed_ref = GOLDEN_FRAME;
        h.comp_var_ref[0] = LAST_FRAME;
        h.comp_var_ref[1] = ALTREF_FRAME;
      } else {
        h.comp_fixed_ref = LAST_FRAME;
        h.comp_var_ref[0] = GOLDEN_FRAME;
        h.comp_var_ref[1] = ALTREF_FRAME;
      }
    }
  }
  if (h.reference_mode == REFERENCE_MODE_SELECT)
    for (int i = 0; i < 5; i++) diff_update_prob(d, &p.comp_mode[i]);
  if (h.reference_mode != COMPOUND_REFERENCE)
    for (int i = 0; i < 5; i++) {
      diff_update_prob(d, &p.single_ref[i][0]);
      diff_update_prob(d, &p.single_ref[i][1]);
    }
  if (h.reference_mode != SINGLE_REFERENCE)
    for (int i = 0; i < 5; i++) diff_update_prob(d, &p.comp_ref[i]);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 9; j++) diff_update_prob(d, &p.y_mode[i][j]);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 3; j++) diff_update_prob(d, &p.partition[i][j]);
  for (int j = 0; j < 3; j++) update_mv_prob(d, &p.mv_joint[j]);
  for (int i = 0; i < 2; i++) {
    update_mv_prob(d, &p.mv_sign[i]);
    for (int j = 0; j < 10; j++) update_mv_prob(d, &p.mv_class[i][j]);
    update_mv_prob(d, &p.mv_class0_bit[i]);
    for (int j = 0; j < 10; j++) update_mv_prob(d, &p.mv_bits[i][j]);
  }
  for (int i = 0; i < 2; i++) {
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 3; k++) update_mv_prob(d, &p.mv_class0_fr[i][j][k]);
    for (int k = 0; k < 3; k++) update_mv_prob(d, &p.mv_fr[i][k]);
  }
  if (h.allow_high_precision_mv)
    for (int i = 0; i < 2; i++) {
      update_mv_prob(d, &p.mv_class0_hp[i]);
      update_mv_prob(d, &p.mv_hp[i]);
    }
}

// --------------------------------------------------------------------------
// Scans and coefficient contexts (6.4.24, 6.4.25; block.rs:1636-1715).

static const uint16_t *get_scan(int tx_sz, int tx_type) {
  switch (tx_sz) {
    case TX_4X4:
      return tx_type == ADST_DCT ? VP9_ROW_SCAN_4X4
             : tx_type == DCT_ADST ? VP9_COL_SCAN_4X4
                                   : VP9_DEFAULT_SCAN_4X4;
    case TX_8X8:
      return tx_type == ADST_DCT ? VP9_ROW_SCAN_8X8
             : tx_type == DCT_ADST ? VP9_COL_SCAN_8X8
                                   : VP9_DEFAULT_SCAN_8X8;
    case TX_16X16:
      return tx_type == ADST_DCT ? VP9_ROW_SCAN_16X16
             : tx_type == DCT_ADST ? VP9_COL_SCAN_16X16
                                   : VP9_DEFAULT_SCAN_16X16;
    default:
      return VP9_DEFAULT_SCAN_32X32;
  }
}

// For scan index c > 0, the two raster positions whose token energies give
// the context of coefficient c: above and left, or twice the one a 1D-ADST
// direction favours.
struct NeighborTables {
  std::vector<uint16_t> nb[4][3];  // [tx_sz][kind]: 2 entries per scan index
  NeighborTables() {
    const int types[3] = {DCT_DCT, DCT_ADST, ADST_DCT};
    for (int t = 0; t < 4; t++)
      for (int k = 0; k < 3; k++) {
        int n = 4 << t, tx_type = types[k];
        const uint16_t *sc = get_scan(t, tx_type);
        std::vector<uint16_t> &v = nb[t][k];
        v.resize(2 * n * n);
        for (int c = 0; c < n * n; c++) {
          int pos = sc[c], i = pos / n, j = pos % n, a, b;
          if (i > 0 && j > 0) {
            int up = (i - 1) * n + j, lf = i * n + j - 1;
            if (tx_type == DCT_ADST)
              a = b = up;
            else if (tx_type == ADST_DCT)
              a = b = lf;
            else
              a = up, b = lf;
          } else if (i > 0) {
            a = b = (i - 1) * n + j;
          } else {
            a = b = j > 0 ? j - 1 : 0;
          }
          v[2 * c] = (uint16_t)a;
          v[2 * c + 1] = (uint16_t)b;
        }
      }
  }
  const uint16_t *get(int tx_sz, int tx_type) const {
    int k = tx_type == DCT_ADST ? 1 : tx_type == ADST_DCT ? 2 : 0;
    return nb[tx_sz][k].data();
  }
};
static const NeighborTables kNeighbors;

// pareto (9.3.2): probability of token-tree node `node` from its band's
// model probability
static inline uint32_t pareto(int node, uint32_t prob) {
  if (node < 2) return prob;
  int x = ((int)std::max<uint32_t>(prob, 1) - 1) / 2;
  if (prob & 1) return VP9_PARETO_TABLE[x][node - 2];
  return ((uint32_t)VP9_PARETO_TABLE[x][node - 2] + VP9_PARETO_TABLE[std::min(x + 1, 127)][node - 2]) >> 1;
}

static inline bool use_mv_hp(const int32_t *mv) {
  return (std::abs(mv[0]) >> 3) < COMPANDED_MVREF_THRESH &&
         (std::abs(mv[1]) >> 3) < COMPANDED_MVREF_THRESH;
}

// get_tile_offset (6.4.1)
static uint32_t tile_offset(uint32_t tile_num, uint32_t mis, uint32_t log2) {
  uint32_t sbs = (mis + 7) >> 3;
  uint32_t off = ((tile_num * sbs) >> log2) << 3;
  return std::min(off, mis);
}

// inter_scale validity (decoder/mod.rs:503-506)
static bool scale_ok(uint32_t rw, uint32_t rh, uint32_t w, uint32_t h) {
  return !(2 * w < rw || 2 * h < rh || w > 16 * rw || h > 16 * rh);
}

// --------------------------------------------------------------------------
// Tile syntax (6.4; block.rs:380-1266, 1540-1715; mvpred.rs).

struct Block {
  uint32_t mi_row, mi_col;
  uint8_t mi_size;
  bool avail_u, avail_l;
  uint8_t segment_id;
  bool skip;
  uint8_t tx_size;
  bool is_inter;
  uint8_t y_mode, sub_modes[4], uv_mode;
  int8_t ref_frame[2];
  uint8_t interp_filter;
  int32_t block_mvs[2][4][2];
  const MiInfo *above, *left;
  uint8_t mode_context[4];
  int32_t nearest_mv[2][2], near_mv[2][2], best_mv[2][2], mv[2][2];
};

// The tile syntax walk (6.4), over a reader (BoolDecoder) or, for the stream
// generator, a writer whose values come from a SyntaxPlan.  Every context,
// motion vector candidate list and token context is the same code either
// way, so what is written is what is read.
template <class Coder>
class TileWalker {
 public:
  static constexpr bool W = Coder::kWrite;
  SyntaxPlan *plan = nullptr;
  // Records and symbol counts go to `out` (the parser's own record and
  // counts, or one tile's part of them when tiles are parsed in parallel).
  TileWalker(Parser &p, TileOut &out)
      : P(p), h(p.h_), seg(p.seg_), probs(p.probs_), O(out), counts(out.counts) {
    uint32_t n4w = h.sb64_cols * 16 + 16, n4h = h.sb64_rows * 16 + 16;
    for (int i = 0; i < 3; i++) {
      above_nz[i].assign(n4w, 0);
      left_nz[i].assign(n4h, 0);
    }
    above_part.assign(h.sb64_cols * 8 + 8, 0);
    left_part.assign(h.sb64_rows * 8 + 8, 0);
    above_segp.assign(h.sb64_cols * 8 + 8, 0);
    left_segp.assign(h.sb64_rows * 8 + 8, 0);
    memset(coefs, 0, sizeof coefs);
    memset(token_cache, 0, sizeof token_cache);
    count_tokens = !h.error_resilient_mode && !h.frame_parallel_decoding_mode;
  }

  void walk_tile(Coder &d, uint32_t row, uint32_t col) {
    row_start = tile_offset(row, h.mi_rows, h.tile_rows_log2);
    row_end = tile_offset(row + 1, h.mi_rows, h.tile_rows_log2);
    col_start = tile_offset(col, h.mi_cols, h.tile_cols_log2);
    col_end = tile_offset(col + 1, h.mi_cols, h.tile_cols_log2);
    for (uint32_t r = row_start; r < row_end; r += 8) {
      for (int p = 0; p < 3; p++) std::fill(left_nz[p].begin(), left_nz[p].end(), 0);
      std::fill(left_part.begin(), left_part.end(), 0);
      std::fill(left_segp.begin(), left_segp.end(), 0);
      for (uint32_t c = col_start; c < col_end; c += 8) decode_partition(d, r, c, BLOCK_64X64);
    }
  }

  void decode_tile(const uint8_t *data, size_t len, uint32_t row, uint32_t col) {
    BoolDecoder d(data, len);
    walk_tile(d, row, col);
    padding_ok = padding_ok && d.padding_is_zero();
  }

  bool padding_ok = true;

 private:
  Parser &P;
  const FrameHeader &h;
  const SegParams &seg;
  const Probs &probs;
  TileOut &O;
  Counts &counts;
  std::vector<uint8_t> above_nz[3], left_nz[3], above_part, left_part, above_segp, left_segp;
  uint32_t row_start = 0, row_end = 0, col_start = 0, col_end = 0;
  Block b;
  PlanBlock pb{};  // the generator's decisions for the block (writing only)
  uint32_t eob_total = 0;
  bool count_tokens;
  int32_t coefs[1024];
  uint8_t token_cache[1024];

  MiInfo &mi_at(uint32_t r, uint32_t c) { return P.mi_[(size_t)r * h.mi_cols + c]; }

  void decode_partition(Coder &d, uint32_t r, uint32_t c, int bsize) {
    if (r >= h.mi_rows || c >= h.mi_cols) return;
    uint32_t num8x8 = VP9_NUM_8X8_WIDE[bsize];
    uint32_t half = num8x8 >> 1;
    bool has_rows = (r + half) < h.mi_rows;
    bool has_cols = (c + half) < h.mi_cols;
    uint32_t bsl = VP9_MI_WIDTH_LOG2[bsize];
    uint32_t boffset = VP9_MI_WIDTH_LOG2[BLOCK_64X64] - bsl;
    uint32_t above = 0, left = 0;
    for (uint32_t i = 0; i < num8x8; i++) {
      above |= above_part[c + i];
      left |= left_part[r + i];
    }
    bool a = (above & (1u << boffset)) != 0, l = (left & (1u << boffset)) != 0;
    uint32_t ctx = bsl * 4 + (l ? 2 : 0) + (a ? 1 : 0);
    const uint8_t *pp = h.frame_is_intra ? VP9_KF_PARTITION_PROBS[ctx] : probs.partition[ctx];
    int want = 0;
    if constexpr (W) want = plan->partition(r, c, bsize);
    int partition;
    if (has_rows && has_cols)
      partition = d.tree(VP9_PARTITION_TREE, [&](int n) { return pp[n]; }, want);
    else if (has_cols)
      partition = d.bit(pp[1], want == PARTITION_SPLIT) ? PARTITION_SPLIT : PARTITION_HORZ;
    else if (has_rows)
      partition = d.bit(pp[2], want == PARTITION_SPLIT) ? PARTITION_SPLIT : PARTITION_VERT;
    else
      partition = PARTITION_SPLIT;
    if (!h.frame_is_intra) counts.partition[ctx][partition]++;
    int subsize = VP9_SUBSIZE_LOOKUP[partition][bsize];
    if (subsize == BLOCK_INVALID) bitstream("partition symbol out of range");
    if (subsize < BLOCK_8X8 || partition == PARTITION_NONE) {
      decode_block(d, r, c, subsize);
    } else if (partition == PARTITION_HORZ) {
      decode_block(d, r, c, subsize);
      if (has_rows) decode_block(d, r + half, c, subsize);
    } else if (partition == PARTITION_VERT) {
      decode_block(d, r, c, subsize);
      if (has_cols) decode_block(d, r, c + half, subsize);
    } else {
      decode_partition(d, r, c, subsize);
      decode_partition(d, r, c + half, subsize);
      decode_partition(d, r + half, c, subsize);
      decode_partition(d, r + half, c + half, subsize);
    }
    if (bsize == BLOCK_8X8 || partition != PARTITION_SPLIT) {
      uint8_t av = (uint8_t)(15 >> VP9_B_WIDTH_LOG2[subsize]);
      uint8_t lv = (uint8_t)(15 >> VP9_B_HEIGHT_LOG2[subsize]);
      for (uint32_t i = 0; i < num8x8; i++) {
        above_part[c + i] = av;
        left_part[r + i] = lv;
      }
    }
  }

  void decode_block(Coder &d, uint32_t r, uint32_t c, int subsize) {
    if (h.ss_x + h.ss_y > 0 && subsize >= BLOCK_8X8 &&
        VP9_SS_SIZE_LOOKUP[subsize][h.ss_x][h.ss_y] == BLOCK_INVALID)
      bitstream("this block size has no chroma block at this subsampling");
    memset(&b, 0, sizeof b);
    b.mi_row = r;
    b.mi_col = c;
    b.mi_size = (uint8_t)subsize;
    b.avail_u = r > 0;
    b.avail_l = c > col_start;
    b.above = b.avail_u ? &mi_at(r - 1, c) : nullptr;
    b.left = b.avail_l ? &mi_at(r, c - 1) : nullptr;
    b.ref_frame[0] = INTRA_FRAME;
    b.ref_frame[1] = REF_NONE;
    if constexpr (W) {
      pb = PlanBlock{};
      plan->block(r, c, subsize, h.frame_is_intra, &pb);
    }
    if (h.frame_is_intra)
      intra_frame_mode_info(d);
    else
      inter_frame_mode_info(d);
    eob_total = 0;
    // the record's block
    uint32_t bi = (uint32_t)O.blocks.size();
    vp9r_block rb;
    memset(&rb, 0, sizeof rb);
    rb.mi_row = (uint16_t)r;
    rb.mi_col = (uint16_t)c;
    rb.mi_size = b.mi_size;
    rb.skip = b.skip;
    rb.tx_size = b.tx_size;
    rb.is_inter = b.is_inter;
    rb.y_mode = b.y_mode;
    rb.uv_mode = b.uv_mode;
    rb.interp_filter = b.interp_filter;
    rb.segment_id = b.segment_id;
    memcpy(rb.sub_modes, b.sub_modes, 4);
    rb.ref_frame[0] = b.ref_frame[0];
    rb.ref_frame[1] = b.ref_frame[1];
    rb.avail_u = b.avail_u;
    rb.avail_l = b.avail_l;
    memcpy(rb.mv, b.block_mvs, sizeof rb.mv);
    O.blocks.push_back(rb);
    residual(d, bi);
    bool skip = b.skip;
    if (b.is_inter && subsize >= BLOCK_8X8 && eob_total == 0) skip = true;
    MiInfo info;
    info.mi_size = (uint8_t)subsize;
    info.skip = skip;
    info.tx_size = b.tx_size;
    info.y_mode = b.y_mode;
    memcpy(info.sub_modes, b.sub_modes, 4);
    info.segment_id = b.segment_id;
    info.interp_filter = b.interp_filter;
    info.ref_frame[0] = b.ref_frame[0];
    info.ref_frame[1] = b.ref_frame[1];
    memcpy(info.mv, b.block_mvs, sizeof info.mv);
    uint32_t bh = VP9_NUM_8X8_HIGH[subsize], bw = VP9_NUM_8X8_WIDE[subsize];
    uint32_t ye = std::min(r + bh, h.mi_rows), xe = std::min(c + bw, h.mi_cols);
    for (uint32_t y = r; y < ye; y++)
      for (uint32_t x = c; x < xe; x++) mi_at(y, x) = info;
  }

  // ---- mode info (6.4.5 - 6.4.17)

  void intra_frame_mode_info(Coder &d) {
    b.segment_id = (seg.enabled && seg.update_map)
                       ? (uint8_t)d.tree(VP9_SEGMENT_TREE, [&](int n) { return seg.tree_probs[n]; }, pb.segment_id)
                       : 0;
    read_skip(d);
    read_tx_size(d, true);
    b.ref_frame[0] = INTRA_FRAME;
    b.ref_frame[1] = REF_NONE;
    b.is_inter = false;
    if (b.mi_size >= BLOCK_8X8) {
      int am = b.above ? b.above->sub_modes[2] : DC_PRED;
      int lm = b.left ? b.left->sub_modes[1] : DC_PRED;
      const uint8_t *p = VP9_KF_Y_MODE_PROBS[am][lm];
      int mode = d.tree(VP9_INTRA_MODE_TREE, [&](int n) { return p[n]; }, pb.y_mode);
      b.y_mode = (uint8_t)mode;
      for (int i = 0; i < 4; i++) b.sub_modes[i] = (uint8_t)mode;
    } else {
      int n4w = VP9_NUM_4X4_WIDE[b.mi_size], n4h = VP9_NUM_4X4_HIGH[b.mi_size];
      int mode = DC_PRED;
      for (int idy = 0; idy < 2; idy += n4h)
        for (int idx = 0; idx < 2; idx += n4w) {
          int am = idy > 0 ? b.sub_modes[idx] : (b.above ? b.above->sub_modes[2 + idx] : DC_PRED);
          int lm = idx > 0 ? b.sub_modes[idy * 2] : (b.left ? b.left->sub_modes[1 + idy * 2] : DC_PRED);
          const uint8_t *p = VP9_KF_Y_MODE_PROBS[am][lm];
          mode = d.tree(VP9_INTRA_MODE_TREE, [&](int n) { return p[n]; }, pb.sub_modes[idy * 2 + idx]);
          for (int y2 = 0; y2 < n4h; y2++)
            for (int x2 = 0; x2 < n4w; x2++) b.sub_modes[(idy + y2) * 2 + idx + x2] = (uint8_t)mode;
        }
      b.y_mode = (uint8_t)mode;
    }
    const uint8_t *p = VP9_KF_UV_MODE_PROBS[b.y_mode];
    b.uv_mode = (uint8_t)d.tree(VP9_INTRA_MODE_TREE, [&](int n) { return p[n]; }, pb.uv_mode);
  }

  void read_skip(Coder &d) {
    if (seg.active(b.segment_id, SEG_LVL_SKIP)) {
      b.skip = true;
    } else {
      int ctx = (b.above ? b.above->skip : 0) + (b.left ? b.left->skip : 0);
      b.skip = d.bit(probs.skip[ctx], pb.skip);
      counts.skip[ctx][b.skip]++;
    }
  }

  void read_tx_size(Coder &d, bool allow_select) {
    int max_tx = VP9_MAX_TXSIZE_LOOKUP[b.mi_size];
    if (allow_select && h.tx_mode == TX_MODE_SELECT && b.mi_size >= BLOCK_8X8) {
      int above = max_tx, left = max_tx;
      if (b.above && !b.above->skip) above = b.above->tx_size;
      if (b.left && !b.left->skip) left = b.left->tx_size;
      if (!b.avail_l) left = above;
      if (!b.avail_u) above = left;
      int ctx = (above + left) > max_tx ? 1 : 0;
      const uint8_t *p = probs.tx[max_tx][ctx];
      const int8_t *tree = max_tx == TX_32X32   ? VP9_TX_SIZE_32_TREE
                           : max_tx == TX_16X16 ? VP9_TX_SIZE_16_TREE
                                                : VP9_TX_SIZE_8_TREE;
      int v = d.tree(tree, [&](int n) { return p[n]; }, std::min<int>(pb.tx_size, max_tx));
      counts.tx[max_tx][ctx][v]++;
      b.tx_size = (uint8_t)v;
    } else {
      b.tx_size = (uint8_t)std::min<int>(max_tx, VP9_TX_MODE_TO_BIGGEST_TX_SIZE[h.tx_mode]);
    }
  }

  void inter_frame_mode_info(Coder &d) {
    inter_segment_id(d);
    read_skip(d);
    read_is_inter(d);
    read_tx_size(d, !b.skip || !b.is_inter);
    if (b.is_inter)
      inter_block_mode_info(d);
    else
      intra_block_mode_info(d);
  }

  uint8_t predicted_segment_id() {
    uint32_t bw = VP9_NUM_8X8_WIDE[b.mi_size], bh = VP9_NUM_8X8_HIGH[b.mi_size];
    uint32_t xmis = std::min(h.mi_cols - b.mi_col, bw), ymis = std::min(h.mi_rows - b.mi_row, bh);
    uint8_t s = 7;
    for (uint32_t y = 0; y < ymis; y++)
      for (uint32_t x = 0; x < xmis; x++) {
        size_t i = (size_t)(b.mi_row + y) * h.mi_cols + b.mi_col + x;
        uint8_t v = i < P.prev_segment_ids_.size() ? P.prev_segment_ids_[i] : 0;
        s = std::min(s, v);
      }
    return s;
  }

  void inter_segment_id(Coder &d) {
    if (!seg.enabled) {
      b.segment_id = 0;
      return;
    }
    uint8_t predicted = predicted_segment_id();
    if (!seg.update_map) {
      b.segment_id = predicted;
      return;
    }
    if (seg.temporal_update) {
      int ctx = left_segp[b.mi_row] + above_segp[b.mi_col];
      bool pred = d.bit(seg.pred_probs[ctx], pb.segment_id == predicted);
      b.segment_id =
          pred ? predicted
               : (uint8_t)d.tree(VP9_SEGMENT_TREE, [&](int n) { return seg.tree_probs[n]; }, pb.segment_id);
      uint32_t bw = VP9_NUM_8X8_WIDE[b.mi_size], bh = VP9_NUM_8X8_HIGH[b.mi_size];
      for (uint32_t i = 0; i < bw; i++) above_segp[b.mi_col + i] = pred;
      for (uint32_t i = 0; i < bh; i++) left_segp[b.mi_row + i] = pred;
    } else {
      b.segment_id = (uint8_t)d.tree(VP9_SEGMENT_TREE, [&](int n) { return seg.tree_probs[n]; }, pb.segment_id);
    }
  }

  // references of the left / above neighbours (INTRA/NONE when absent)
  void left_ref(int8_t *o) const {
    o[0] = b.left ? b.left->ref_frame[0] : INTRA_FRAME;
    o[1] = b.left ? b.left->ref_frame[1] : REF_NONE;
  }
  void above_ref(int8_t *o) const {
    o[0] = b.above ? b.above->ref_frame[0] : INTRA_FRAME;
    o[1] = b.above ? b.above->ref_frame[1] : REF_NONE;
  }

  void read_is_inter(Coder &d) {
    if (seg.active(b.segment_id, SEG_LVL_REF_FRAME)) {
      b.is_inter = seg.feature_data[b.segment_id][SEG_LVL_REF_FRAME] != INTRA_FRAME;
      return;
    }
    int8_t l[2], a[2];
    left_ref(l);
    above_ref(a);
    bool li = l[0] <= INTRA_FRAME, ai = a[0] <= INTRA_FRAME;
    bool au = b.avail_u, al = b.avail_l;
    int ctx;
    if (au && al)
      ctx = (li && ai) ? 3 : (li || ai) ? 1 : 0;
    else if (au || al)
      ctx = 2 * (au ? ai : li);
    else
      ctx = 0;
    b.is_inter = d.bit(probs.is_inter[ctx], pb.is_inter);
    counts.is_inter[ctx][b.is_inter]++;
  }

  void intra_block_mode_info(Coder &d) {
    b.ref_frame[0] = INTRA_FRAME;
    b.ref_frame[1] = REF_NONE;
    if (b.mi_size >= BLOCK_8X8) {
      int ctx = VP9_SIZE_GROUP[b.mi_size];
      const uint8_t *p = probs.y_mode[ctx];
      int mode = d.tree(VP9_INTRA_MODE_TREE, [&](int n) { return p[n]; }, pb.y_mode);
      counts.intra_mode[ctx][mode]++;
      b.y_mode = (uint8_t)mode;
      for (int i = 0; i < 4; i++) b.sub_modes[i] = (uint8_t)mode;
    } else {
      int n4w = VP9_NUM_4X4_WIDE[b.mi_size], n4h = VP9_NUM_4X4_HIGH[b.mi_size];
      const uint8_t *p = probs.y_mode[0];
      int mode = DC_PRED;
      for (int idy = 0; idy < 2; idy += n4h)
        for (int idx = 0; idx < 2; idx += n4w) {
          mode = d.tree(VP9_INTRA_MODE_TREE, [&](int n) { return p[n]; }, pb.sub_modes[idy * 2 + idx]);
          counts.intra_mode[0][mode]++;
          for (int y2 = 0; y2 < n4h; y2++)
            for (int x2 = 0; x2 < n4w; x2++) b.sub_modes[(idy + y2) * 2 + idx + x2] = (uint8_t)mode;
        }
      b.y_mode = (uint8_t)mode;
    }
    const uint8_t *p = probs.uv_mode[b.y_mode];
    int uv = d.tree(VP9_INTRA_MODE_TREE, [&](int n) { return p[n]; }, pb.uv_mode);
    counts.uv_mode[b.y_mode][uv]++;
    b.uv_mode = (uint8_t)uv;
  }

  void inter_block_mode_info(Coder &d) {
    read_ref_frames(d);
    for (int j = 0; j < 2; j++)
      if (b.ref_frame[j] > INTRA_FRAME) find_best_ref_mvs(j);
    bool is_compound = b.ref_frame[1] > INTRA_FRAME;
    if (seg.active(b.segment_id, SEG_LVL_SKIP))
      b.y_mode = ZEROMV;
    else if (b.mi_size >= BLOCK_8X8)
      b.y_mode = (uint8_t)(NEARESTMV + read_inter_mode(d));
    if (h.interp_filter == SWITCHABLE) {
      int8_t l[2], a[2];
      left_ref(l);
      above_ref(a);
      int left_i = (b.avail_l && l[0] > INTRA_FRAME) ? b.left->interp_filter : 3;
      int above_i = (b.avail_u && a[0] > INTRA_FRAME) ? b.above->interp_filter : 3;
      int ctx;
      if (left_i == above_i)
        ctx = left_i;
      else if (left_i == 3 && above_i != 3)
        ctx = above_i;
      else if (left_i != 3 && above_i == 3)
        ctx = left_i;
      else
        ctx = 3;
      const uint8_t *p = probs.interp_filter[ctx];
      int f = d.tree(VP9_INTERP_FILTER_TREE, [&](int n) { return p[n]; }, pb.interp_filter);
      counts.interp_filter[ctx][f]++;
      b.interp_filter = (uint8_t)f;
    } else {
      b.interp_filter = (uint8_t)h.interp_filter;
    }
    int nl = is_compound ? 2 : 1;
    if (b.mi_size < BLOCK_8X8) {
      int n4w = VP9_NUM_4X4_WIDE[b.mi_size], n4h = VP9_NUM_4X4_HIGH[b.mi_size];
      for (int idy = 0; idy < 2; idy += n4h)
        for (int idx = 0; idx < 2; idx += n4w) {
          b.y_mode = (uint8_t)(NEARESTMV + read_inter_mode(d));
          if (b.y_mode == NEARESTMV || b.y_mode == NEARMV)
            for (int j = 0; j < nl; j++) append_sub8x8_mvs(idy * 2 + idx, j);
          assign_mv(d, is_compound);
          for (int y2 = 0; y2 < n4h; y2++)
            for (int x2 = 0; x2 < n4w; x2++) {
              int blk = (idy + y2) * 2 + idx + x2;
              for (int rl = 0; rl < nl; rl++) {
                b.block_mvs[rl][blk][0] = b.mv[rl][0];
                b.block_mvs[rl][blk][1] = b.mv[rl][1];
              }
            }
        }
    } else {
      assign_mv(d, is_compound);
      for (int rl = 0; rl < nl; rl++)
        for (int k = 0; k < 4; k++) {
          b.block_mvs[rl][k][0] = b.mv[rl][0];
          b.block_mvs[rl][k][1] = b.mv[rl][1];
        }
    }
  }

  // The inter mode a writer codes: the candidate equal to the planned motion
  // vector, else NEWMV.
  int planned_inter_mode() {
    int32_t t[2][2];
    bool compound = b.ref_frame[1] > INTRA_FRAME;
    for (int l = 0; l < 2; l++) planned_mv(l, t[l]);
    auto eq = [&](int32_t m[2][2]) {
      return m[0][0] == t[0][0] && m[0][1] == t[0][1] && (!compound || (m[1][0] == t[1][0] && m[1][1] == t[1][1]));
    };
    int32_t zero[2][2] = {{0, 0}, {0, 0}};
    if (eq(b.nearest_mv)) return NEARESTMV - NEARESTMV;
    if (eq(b.near_mv)) return NEARMV - NEARESTMV;
    if (eq(zero)) return ZEROMV - NEARESTMV;
    return NEWMV - NEARESTMV;
  }
  // The planned motion vector of list l, made codable: even components when
  // high precision is unavailable (the reader then assumes the last bit).
  void planned_mv(int l, int32_t *out) {
    out[0] = pb.mv[l][0];
    out[1] = pb.mv[l][1];
    if (!(h.allow_high_precision_mv && use_mv_hp(b.best_mv[l])))
      for (int k = 0; k < 2; k++) out[k] = (out[k] / 2) * 2;
  }

  int read_inter_mode(Coder &d) {
    int ctx = b.mode_context[b.ref_frame[0]];
    const uint8_t *p = probs.inter_mode[ctx];
    int want = 0;
    if constexpr (W) want = planned_inter_mode();
    int v = d.tree(VP9_INTER_MODE_TREE, [&](int n) { return p[n]; }, want);
    counts.inter_mode[ctx][v]++;
    return v;
  }

  void read_ref_frames(Coder &d) {
    if (seg.active(b.segment_id, SEG_LVL_REF_FRAME)) {
      int rf = seg.feature_data[b.segment_id][SEG_LVL_REF_FRAME];
      if (rf < 1 || rf > 3) bitstream("segment feature forces a reference on an intra-only frame");
      b.ref_frame[0] = (int8_t)rf;
      b.ref_frame[1] = REF_NONE;
      return;
    }
    int comp_mode;
    if (h.reference_mode == REFERENCE_MODE_SELECT) {
      int ctx = comp_mode_ctx();
      bool v = d.bit(probs.comp_mode[ctx], pb.ref_frame[1] > INTRA_FRAME);
      counts.comp_mode[ctx][v]++;
      comp_mode = v ? COMPOUND_REFERENCE : SINGLE_REFERENCE;
    } else {
      comp_mode = (int)h.reference_mode;
    }
    if (comp_mode == COMPOUND_REFERENCE) {
      int idx = h.ref_frame_sign_bias[h.comp_fixed_ref] ? 1 : 0;
      int ctx = comp_ref_ctx();
      bool v = d.bit(probs.comp_ref[ctx], pb.ref_frame[1 - idx] == h.comp_var_ref[1]);
      counts.comp_ref[ctx][v]++;
      b.ref_frame[idx] = (int8_t)h.comp_fixed_ref;
      b.ref_frame[1 - idx] = (int8_t)h.comp_var_ref[v ? 1 : 0];
    } else {
      int ctx = single_ref_p1_ctx();
      bool p1 = d.bit(probs.single_ref[ctx][0], pb.ref_frame[0] != LAST_FRAME);
      counts.single_ref[ctx][0][p1]++;
      if (p1) {
        int ctx2 = single_ref_p2_ctx();
        bool p2 = d.bit(probs.single_ref[ctx2][1], pb.ref_frame[0] == ALTREF_FRAME);
        counts.single_ref[ctx2][1][p2]++;
        b.ref_frame[0] = p2 ? ALTREF_FRAME : GOLDEN_FRAME;
      } else {
        b.ref_frame[0] = LAST_FRAME;
      }
      b.ref_frame[1] = REF_NONE;
    }
  }

  // contexts of the reference syntax (6.4.14 - 6.4.16; block.rs:998-1234)
  int comp_mode_ctx() const {
    int8_t a[2], l[2];
    above_ref(a);
    left_ref(l);
    bool au = b.avail_u, al = b.avail_l;
    bool as = a[1] <= REF_NONE, ls = l[1] <= REF_NONE;
    bool ai = a[0] <= INTRA_FRAME, li = l[0] <= INTRA_FRAME;
    int fixed = h.comp_fixed_ref;
    if (au && al) {
      if (as && ls) return (a[0] == fixed) ^ (l[0] == fixed);
      if (as) return 2 + (a[0] == fixed || ai);
      if (ls) return 2 + (l[0] == fixed || li);
      return 4;
    }
    if (au) return as ? (a[0] == fixed) : 3;
    if (al) return ls ? (l[0] == fixed) : 3;
    return 1;
  }

  int comp_ref_ctx() const {
    int8_t a[2], l[2];
    above_ref(a);
    left_ref(l);
    bool au = b.avail_u, al = b.avail_l;
    bool as = a[1] <= REF_NONE, ls = l[1] <= REF_NONE;
    bool ai = a[0] <= INTRA_FRAME, li = l[0] <= INTRA_FRAME;
    int fix_idx = h.ref_frame_sign_bias[h.comp_fixed_ref] ? 1 : 0;
    int var_idx = 1 - fix_idx;
    int var1 = h.comp_var_ref[1], var0 = h.comp_var_ref[0], fixed = h.comp_fixed_ref;
    if (au && al) {
      if (ai && li) return 2;
      if (li) return as ? 1 + 2 * (a[0] != var1) : 1 + 2 * (a[var_idx] != var1);
      if (ai) return ls ? 1 + 2 * (l[0] != var1) : 1 + 2 * (l[var_idx] != var1);
      int vrfa = as ? a[0] : a[var_idx];
      int vrfl = ls ? l[0] : l[var_idx];
      if (vrfa == vrfl && var1 == vrfa) return 0;
      if (ls && as) {
        if ((vrfa == fixed && vrfl == var0) || (vrfl == fixed && vrfa == var0)) return 4;
        if (vrfa == vrfl) return 3;
        return 1;
      }
      if (ls || as) {
        int vrfc = ls ? vrfa : vrfl;
        int rfs = as ? vrfa : vrfl;
        if (vrfc == var1 && rfs != var1) return 1;
        if (rfs == var1 && vrfc != var1) return 2;
        return 4;
      }
      return vrfa == vrfl ? 4 : 2;
    }
    if (au) {
      if (ai) return 2;
      return as ? 3 * (a[0] != var1) : 4 * (a[var_idx] != var1);
    }
    if (al) {
      if (li) return 2;
      return ls ? 3 * (l[0] != var1) : 4 * (l[var_idx] != var1);
    }
    return 2;
  }

  int single_ref_p1_ctx() const {
    int8_t a[2], l[2];
    above_ref(a);
    left_ref(l);
    bool au = b.avail_u, al = b.avail_l;
    bool as = a[1] <= REF_NONE, ls = l[1] <= REF_NONE;
    bool ai = a[0] <= INTRA_FRAME, li = l[0] <= INTRA_FRAME;
    const int LF = LAST_FRAME;
    if (au && al) {
      if (ai && li) return 2;
      if (li) return as ? 4 * (a[0] == LF) : 1 + (a[0] == LF || a[1] == LF);
      if (ai) return ls ? 4 * (l[0] == LF) : 1 + (l[0] == LF || l[1] == LF);
      if (as && ls) return 2 * (a[0] == LF) + 2 * (l[0] == LF);
      if (!as && !ls) return 1 + (a[0] == LF || a[1] == LF || l[0] == LF || l[1] == LF);
      int rfs = as ? a[0] : l[0];
      int crf1 = as ? l[0] : a[0];
      int crf2 = as ? l[1] : a[1];
      if (rfs == LF) return 3 + (crf1 == LF || crf2 == LF);
      return (crf1 == LF || crf2 == LF);
    }
    if (au) {
      if (ai) return 2;
      return as ? 4 * (a[0] == LF) : 1 + (a[0] == LF || a[1] == LF);
    }
    if (al) {
      if (li) return 2;
      return ls ? 4 * (l[0] == LF) : 1 + (l[0] == LF || l[1] == LF);
    }
    return 2;
  }

  int single_ref_p2_ctx() const {
    int8_t a[2], l[2];
    above_ref(a);
    left_ref(l);
    bool au = b.avail_u, al = b.avail_l;
    bool as = a[1] <= REF_NONE, ls = l[1] <= REF_NONE;
    bool ai = a[0] <= INTRA_FRAME, li = l[0] <= INTRA_FRAME;
    const int LF = LAST_FRAME, GF = GOLDEN_FRAME, AF = ALTREF_FRAME;
    if (au && al) {
      if (ai && li) return 2;
      if (li) {
        if (as) return a[0] == LF ? 3 : 4 * (a[0] == GF);
        return 1 + 2 * (a[0] == GF || a[1] == GF);
      }
      if (ai) {
        if (ls) return l[0] == LF ? 3 : 4 * (l[0] == GF);
        return 1 + 2 * (l[0] == GF || l[1] == GF);
      }
      if (as && ls) {
        if (a[0] == LF && l[0] == LF) return 3;
        if (a[0] == LF) return 4 * (l[0] == GF);
        if (l[0] == LF) return 4 * (a[0] == GF);
        return 2 * (a[0] == GF) + 2 * (l[0] == GF);
      }
      if (!as && !ls) {
        if (a[0] == l[0] && a[1] == l[1]) return 3 * (a[0] == GF || a[1] == GF);
        return 2;
      }
      int rfs = as ? a[0] : l[0];
      int crf1 = as ? l[0] : a[0];
      int crf2 = as ? l[1] : a[1];
      if (rfs == GF) return 3 + (crf1 == GF || crf2 == GF);
      if (rfs == AF) return (crf1 == GF || crf2 == GF);
      return 1 + 2 * (crf1 == GF || crf2 == GF);
    }
    if (au) {
      if (ai || (a[0] == LF && as)) return 2;
      if (as) return 4 * (a[0] == GF);
      return 3 * (a[0] == GF || a[1] == GF);
    }
    if (al) {
      if (li || (l[0] == LF && ls)) return 2;
      if (ls) return 4 * (l[0] == GF);
      return 3 * (l[0] == GF || l[1] == GF);
    }
    return 2;
  }

  void assign_mv(Coder &d, bool is_compound) {
    b.mv[1][0] = b.mv[1][1] = 0;
    for (int i = 0; i < (is_compound ? 2 : 1); i++) {
      switch (b.y_mode) {
        case NEWMV:
          read_mv(d, i);
          break;
        case NEARESTMV:
          b.mv[i][0] = b.nearest_mv[i][0];
          b.mv[i][1] = b.nearest_mv[i][1];
          break;
        case NEARMV:
          b.mv[i][0] = b.near_mv[i][0];
          b.mv[i][1] = b.near_mv[i][1];
          break;
        default:
          b.mv[i][0] = b.mv[i][1] = 0;
      }
    }
  }

  void read_mv(Coder &d, int r) {
    const int32_t *best = b.best_mv[r];
    bool use_hp = h.allow_high_precision_mv && use_mv_hp(best);
    int32_t diff[2] = {0, 0}, want[2] = {0, 0};
    if constexpr (W) {
      int32_t t[2];
      planned_mv(r, t);
      want[0] = t[0] - best[0];
      want[1] = t[1] - best[1];
    }
    int joint = d.tree(VP9_MV_JOINT_TREE, [&](int n) { return probs.mv_joint[n]; },
                       (want[0] != 0 ? 2 : 0) | (want[1] != 0 ? 1 : 0));
    counts.mv_joint[joint]++;
    if (joint == MV_JOINT_HZVNZ || joint == MV_JOINT_HNZVNZ) diff[0] = read_mv_component(d, 0, use_hp, want[0]);
    if (joint == MV_JOINT_HNZVZ || joint == MV_JOINT_HNZVNZ) diff[1] = read_mv_component(d, 1, use_hp, want[1]);
    b.mv[r][0] = best[0] + diff[0];
    b.mv[r][1] = best[1] + diff[1];
  }

  // want: the component a writer codes (nonzero; even when !use_hp)
  int32_t read_mv_component(Coder &d, int comp, bool use_hp, int32_t want) {
    int w_cls = 0;
    uint32_t w_off = 0;
    if constexpr (W) {
      w_off = (uint32_t)std::abs(want) - 1;
      while (w_cls < 10 && w_off >= (16u << w_cls)) w_cls++;
      if (w_cls) w_off -= 8u << w_cls;
    }
    bool sign = d.bit(probs.mv_sign[comp], want < 0);
    counts.mv_sign[comp][sign]++;
    const uint8_t *pc = probs.mv_class[comp];
    int cls = d.tree(VP9_MV_CLASS_TREE, [&](int n) { return pc[n]; }, w_cls);
    counts.mv_class[comp][cls]++;
    int32_t mag;
    if (cls == 0) {
      int c0 = d.bit(probs.mv_class0_bit[comp], (w_off >> 3) & 1);
      counts.mv_class0_bit[comp][c0]++;
      const uint8_t *pf = probs.mv_class0_fr[comp][c0];
      int fr = d.tree(VP9_MV_FR_TREE, [&](int n) { return pf[n]; }, (w_off >> 1) & 3);
      counts.mv_class0_fr[comp][c0][fr]++;
      int hp = use_hp ? (int)d.bit(probs.mv_class0_hp[comp], w_off & 1) : 1;
      counts.mv_class0_hp[comp][hp]++;
      mag = ((c0 << 3) | (fr << 1) | hp) + 1;
    } else {
      int32_t dd = 0;
      for (int i = 0; i < cls; i++) {
        int bit = d.bit(probs.mv_bits[comp][i], (w_off >> (3 + i)) & 1);
        counts.mv_bits[comp][i][bit]++;
        dd |= bit << i;
      }
      int32_t base = 2 << (cls + 2);
      const uint8_t *pf = probs.mv_fr[comp];
      int fr = d.tree(VP9_MV_FR_TREE, [&](int n) { return pf[n]; }, (w_off >> 1) & 3);
      counts.mv_fr[comp][fr]++;
      int hp = use_hp ? (int)d.bit(probs.mv_hp[comp], w_off & 1) : 1;
      counts.mv_hp[comp][hp]++;
      mag = base + ((dd << 3) | (fr << 1) | hp) + 1;
    }
    return sign ? -mag : mag;
  }

  // ---- motion vector prediction (6.5; mvpred.rs)

  struct Search {
    int count = 0;
    int32_t list[2][2] = {{0, 0}, {0, 0}};
    int32_t cand_mv[2][2] = {{0, 0}, {0, 0}};
    int8_t cand_frame[2] = {0, 0};
    void add(int rl) {
      if (count >= 2) return;
      if (count > 0 && cand_mv[rl][0] == list[0][0] && cand_mv[rl][1] == list[0][1]) return;
      list[count][0] = cand_mv[rl][0];
      list[count][1] = cand_mv[rl][1];
      count++;
    }
  };

  bool is_inside(int r, int c) const {
    return r >= 0 && r < (int)h.mi_rows && c >= (int)col_start && c < (int)col_end;
  }

  void get_block_mv(Search &s, int r, int c, int rl, bool use_prev) {
    if (use_prev) {
      size_t idx = (size_t)r * h.mi_cols + c;
      PrevMv pm;
      if (idx < P.prev_mvs_.size()) pm = P.prev_mvs_[idx];
      s.cand_mv[rl][0] = pm.mv[rl][0];
      s.cand_mv[rl][1] = pm.mv[rl][1];
      s.cand_frame[rl] = pm.ref_frame[rl];
    } else {
      const MiInfo &m = mi_at(r, c);
      s.cand_mv[rl][0] = m.mv[rl][3][0];
      s.cand_mv[rl][1] = m.mv[rl][3][1];
      s.cand_frame[rl] = m.ref_frame[rl];
    }
  }

  void if_same_ref_frame_add(Search &s, int r, int c, int ref_frame, bool use_prev) {
    for (int j = 0; j < 2; j++) {
      get_block_mv(s, r, c, j, use_prev);
      if (s.cand_frame[j] == ref_frame) {
        s.add(j);
        return;
      }
    }
  }

  void scale_mv(Search &s, int rl, int ref_frame) {
    int cand = s.cand_frame[rl];
    if (h.ref_frame_sign_bias[cand] != h.ref_frame_sign_bias[ref_frame]) {
      s.cand_mv[rl][0] *= -1;
      s.cand_mv[rl][1] *= -1;
    }
  }

  void if_diff_ref_frame_add(Search &s, int r, int c, int ref_frame, bool use_prev) {
    for (int j = 0; j < 2; j++) get_block_mv(s, r, c, j, use_prev);
    bool mvs_same = s.cand_mv[0][0] == s.cand_mv[1][0] && s.cand_mv[0][1] == s.cand_mv[1][1];
    if (s.cand_frame[0] > INTRA_FRAME && s.cand_frame[0] != ref_frame) {
      scale_mv(s, 0, ref_frame);
      s.add(0);
    }
    if (s.cand_frame[1] > INTRA_FRAME && s.cand_frame[1] != ref_frame && !mvs_same) {
      scale_mv(s, 1, ref_frame);
      s.add(1);
    }
  }

  int32_t clamp_mv_row(int32_t v, int32_t border) const {
    int32_t bh = VP9_NUM_8X8_HIGH[b.mi_size], mi_row = (int32_t)b.mi_row;
    int32_t to_top = -((mi_row * 8) * 8);
    int32_t to_bottom = (((int32_t)h.mi_rows - bh - mi_row) * 8) * 8;
    return std::min(std::max(v, to_top - border), to_bottom + border);
  }
  int32_t clamp_mv_col(int32_t v, int32_t border) const {
    int32_t bw = VP9_NUM_8X8_WIDE[b.mi_size], mi_col = (int32_t)b.mi_col;
    int32_t to_left = -((mi_col * 8) * 8);
    int32_t to_right = (((int32_t)h.mi_cols - bw - mi_col) * 8) * 8;
    return std::min(std::max(v, to_left - border), to_right + border);
  }

  // find_mv_refs (6.5.1): fills list
  void find_mv_refs(int ref_frame, int block, int32_t list[2][2]) {
    Search s;
    int context_counter = 0;
    const int8_t(*search)[2] = VP9_MV_REF_BLOCKS[b.mi_size];
    int mi_row = (int)b.mi_row, mi_col = (int)b.mi_col;
    for (int k = 0; k < 2; k++) {
      int r = mi_row + search[k][0], c = mi_col + search[k][1];
      if (is_inside(r, c)) {
        const MiInfo &m = mi_at(r, c);
        context_counter += VP9_MODE_2_COUNTER[m.y_mode];
        for (int j = 0; j < 2; j++) {
          if (m.ref_frame[j] == ref_frame) {
            int idx = block >= 0 ? VP9_IDX_N_COLUMN_TO_SUBBLOCK[block][search[k][1] == 0 ? 1 : 0] : 3;
            s.cand_mv[j][0] = m.mv[j][idx][0];
            s.cand_mv[j][1] = m.mv[j][idx][1];
            s.add(j);
            break;
          }
        }
      }
    }
    for (int k = 2; k < 8; k++) {
      int r = mi_row + search[k][0], c = mi_col + search[k][1];
      if (is_inside(r, c)) if_same_ref_frame_add(s, r, c, ref_frame, false);
    }
    bool use_prev = P.use_prev_mvs_;
    if (use_prev) if_same_ref_frame_add(s, mi_row, mi_col, ref_frame, true);
    for (int k = 0; k < 8; k++) {
      int r = mi_row + search[k][0], c = mi_col + search[k][1];
      if (is_inside(r, c)) if_diff_ref_frame_add(s, r, c, ref_frame, false);
    }
    if (use_prev) if_diff_ref_frame_add(s, mi_row, mi_col, ref_frame, true);
    b.mode_context[ref_frame] = VP9_COUNTER_TO_CONTEXT[context_counter];
    for (int i = 0; i < 2; i++) {
      list[i][0] = clamp_mv_row(s.list[i][0], MV_BORDER);
      list[i][1] = clamp_mv_col(s.list[i][1], MV_BORDER);
    }
  }

  // find_best_ref_mvs (6.5.12)
  void find_best_ref_mvs(int rl) {
    int32_t list[2][2];
    find_mv_refs(b.ref_frame[rl], -1, list);
    int32_t border = (BORDERINPIXELS - INTERP_EXTEND) << 3;
    for (int i = 0; i < 2; i++) {
      int32_t dr = list[i][0], dc = list[i][1];
      if (!h.allow_high_precision_mv || !use_mv_hp(list[i])) {
        if (dr & 1) dr += dr > 0 ? -1 : 1;
        if (dc & 1) dc += dc > 0 ? -1 : 1;
      }
      list[i][0] = clamp_mv_row(dr, border);
      list[i][1] = clamp_mv_col(dc, border);
    }
    b.nearest_mv[rl][0] = list[0][0];
    b.nearest_mv[rl][1] = list[0][1];
    b.near_mv[rl][0] = list[1][0];
    b.near_mv[rl][1] = list[1][1];
    b.best_mv[rl][0] = list[0][0];
    b.best_mv[rl][1] = list[0][1];
  }

  // append_sub8x8_mvs (6.5.14)
  void append_sub8x8_mvs(int block, int rl) {
    int32_t list[2][2];
    find_mv_refs(b.ref_frame[rl], block, list);
    int32_t sub[2][2] = {{0, 0}, {0, 0}};
    int dst = 0;
    int32_t(*bm)[2] = b.block_mvs[rl];
    auto same = [](const int32_t *x, const int32_t *y) { return x[0] == y[0] && x[1] == y[1]; };
    if (block == 0) {
      memcpy(sub, list, sizeof sub);
      dst = 2;
    } else if (block <= 2) {
      sub[dst][0] = bm[0][0];
      sub[dst][1] = bm[0][1];
      dst++;
    } else {
      sub[dst][0] = bm[2][0];
      sub[dst][1] = bm[2][1];
      dst++;
      for (int idx = 1; idx >= 0 && dst < 2; idx--) {
        if (!same(bm[idx], sub[0])) {
          sub[dst][0] = bm[idx][0];
          sub[dst][1] = bm[idx][1];
          dst++;
        }
      }
    }
    for (int n = 0; n < 2 && dst < 2; n++) {
      if (!same(list[n], sub[0])) {
        sub[dst][0] = list[n][0];
        sub[dst][1] = list[n][1];
        dst++;
      }
    }
    if (dst < 2) sub[dst][0] = sub[dst][1] = 0;
    b.nearest_mv[rl][0] = sub[0][0];
    b.nearest_mv[rl][1] = sub[0][1];
    b.near_mv[rl][0] = sub[1][0];
    b.near_mv[rl][1] = sub[1][1];
  }

  // ---- residual (6.4.21 - 6.4.26)

  void uv_subsampling(int *sx, int *sy) const {
    if (h.legacy_uv) {
      *sx = *sy = 1;
    } else {
      *sx = (int)h.ss_x;
      *sy = (int)h.ss_y;
    }
  }

  int uv_tx_size() const {
    if (b.mi_size < BLOCK_8X8) return TX_4X4;
    int sx, sy;
    uv_subsampling(&sx, &sy);
    int uv = VP9_SS_SIZE_LOOKUP[b.mi_size][sx][sy];
    return std::min<int>(b.tx_size, VP9_MAX_TXSIZE_LOOKUP[uv]);
  }

  int tx_type(int plane, int tx_sz, int block_idx) const {
    if (plane > 0 || tx_sz == TX_32X32) return DCT_DCT;
    if (tx_sz == TX_4X4) {
      if (h.lossless || b.is_inter) return DCT_DCT;
      return VP9_MODE2TXFM_MAP[b.mi_size < BLOCK_8X8 ? b.sub_modes[block_idx] : b.y_mode];
    }
    return VP9_MODE2TXFM_MAP[b.y_mode];
  }

  void check_refs() {
    int n = b.ref_frame[1] > INTRA_FRAME ? 2 : 1;
    for (int i = 0; i < n; i++) {
      int rf = b.ref_frame[i];
      uint32_t ok = P.ref_ok_[rf - LAST_FRAME];
      if (ok == 0) bitstream("inter block names a reference that is not loaded");
      if (ok == 2) bitstream("reference size outside the 2x-down / 16x-up scaling range");
    }
  }

  void residual(Coder &d, uint32_t bi) {
    int bsize = b.mi_size < BLOCK_8X8 ? BLOCK_8X8 : b.mi_size;
    for (int plane = 0; plane < 3; plane++) {
      int tx_sz = plane > 0 ? uv_tx_size() : b.tx_size;
      int step = 1 << tx_sz;
      int sx = plane > 0 ? (int)h.ss_x : 0, sy = plane > 0 ? (int)h.ss_y : 0;
      int plane_sz = VP9_SS_SIZE_LOOKUP[bsize][sx][sy];
      if (plane_sz == BLOCK_INVALID) bitstream("chroma plane has no block of this size");
      int n4w = VP9_NUM_4X4_WIDE[plane_sz], n4h = VP9_NUM_4X4_HIGH[plane_sz];
      uint32_t base_x = (b.mi_col * 8) >> sx, base_y = (b.mi_row * 8) >> sy;
      if (b.is_inter && plane == 0) check_refs();
      uint32_t max_x = (h.mi_cols * 8) >> sx, max_y = (h.mi_rows * 8) >> sy;
      int block_idx = 0;
      for (int y = 0; y < n4h; y += step) {
        for (int x = 0; x < n4w; x += step) {
          uint32_t start_x = base_x + 4 * x, start_y = base_y + 4 * y;
          bool nonzero = false;
          if (start_x < max_x && start_y < max_y) {
            int tt = tx_type(plane, tx_sz, block_idx);
            uint32_t eob = 0;
            uint32_t cstart = (uint32_t)O.coef_pos.size();
            if (!b.skip) {
              eob = tokens(d, plane, start_x, start_y, tx_sz, tt);
              nonzero = eob > 0;
              // keep the nonzero coefficients, and clear them
              const uint16_t *sc = get_scan(tx_sz, tt);
              for (uint32_t k = 0; k < eob; k++) {
                int pos = sc[k];
                if (coefs[pos]) {
                  O.coef_pos.push_back((uint16_t)pos);
                  O.coef_val.push_back(coefs[pos]);
                  coefs[pos] = 0;
                }
              }
            }
            if (!b.is_inter || eob > 0) {
              vp9r_txb t;
              t.x = (uint16_t)start_x;
              t.y = (uint16_t)start_y;
              t.plane = (uint8_t)plane;
              t.tx_size = (uint8_t)tx_sz;
              t.tx_type = (uint8_t)tt;
              t._pad = 0;
              t.eob = (uint16_t)eob;
              t.ncoef = (uint16_t)(O.coef_pos.size() - cstart);
              t.coef_start = cstart;
              t.block = bi;
              O.txbs.push_back(t);
            }
          }
          uint8_t *an = above_nz[plane].data(), *ln = left_nz[plane].data();
          for (int i = 0; i < step; i++) {
            an[(start_x >> 2) + i] = nonzero;
            ln[(start_y >> 2) + i] = nonzero;
          }
          block_idx++;
        }
      }
    }
  }

  uint32_t tokens(Coder &d, int plane, uint32_t start_x, uint32_t start_y, int tx_sz,
                  int tt) {
    uint32_t seg_eob = 16u << (tx_sz << 1);
    const uint16_t *scan = get_scan(tx_sz, tt);
    int ref_type = b.is_inter ? 1 : 0;
    int ptype = plane > 0 ? 1 : 0;
    int sx = plane > 0 ? (int)h.ss_x : 0, sy = plane > 0 ? (int)h.ss_y : 0;
    uint32_t max_x4 = (2 * h.mi_cols) >> sx, max_y4 = (2 * h.mi_rows) >> sy;
    uint32_t x4 = start_x >> 2, y4 = start_y >> 2;
    uint32_t above = 0, left = 0;
    for (int i = 0; i < (1 << tx_sz); i++) {
      if (x4 + i < max_x4) above |= above_nz[plane][x4 + i];
      if (y4 + i < max_y4) left |= left_nz[plane][y4 + i];
    }
    int ctx = (int)(above + left);
    const uint16_t *nb = kNeighbors.get(tx_sz, tt);
    const uint8_t *bands = tx_sz == TX_4X4 ? VP9_COEFBAND_4X4 : VP9_COEFBAND_8X8PLUS;
    const uint8_t(*cp)[6][3] = probs.coef[tx_sz][ptype][ref_type];
    bool check_eob = true;
    uint32_t c = 0;
    int bd = (int)h.bit_depth;
    uint32_t w_eob = 0;
    if constexpr (W) {
      plan->coefs(plane, start_x, start_y, tx_sz, tt, coefs);
      for (uint32_t k = 0; k < seg_eob; k++)
        if (coefs[scan[k]]) w_eob = k + 1;
    }
    while (c < seg_eob) {
      int pos = scan[c];
      int band = bands[c];
      if (c > 0) ctx = (1 + token_cache[nb[2 * c]] + token_cache[nb[2 * c + 1]]) >> 1;
      const uint8_t *p = cp[band][ctx];
      if (check_eob) {
        bool more = d.bit(p[0], c < w_eob);
        if (count_tokens) counts.more_coefs[tx_sz][ptype][ref_type][band][ctx][more]++;
        if (!more) break;
      }
      int w_tok = 0;
      if constexpr (W) w_tok = token_of(coefs[pos]);
      int token = d.tree(VP9_TOKEN_TREE, [&](int node) { return pareto(node, p[std::min(1 + node, 2)]); }, w_tok);
      if (count_tokens) counts.token[tx_sz][ptype][ref_type][band][ctx][std::min(token, 2)]++;
      token_cache[pos] = VP9_ENERGY_CLASS[token];
      if (token == ZERO_TOKEN) {
        check_eob = false;
      } else {
        int32_t coef = read_coef(d, token, bd, std::abs(coefs[pos]));
        bool sign = d.bit(128, coefs[pos] < 0);
        coefs[pos] = sign ? -coef : coef;
        check_eob = true;
      }
      c++;
    }
    if (c > 0) eob_total++;
    return c;
  }

  // want: the magnitude a writer codes
  static int32_t read_coef(Coder &d, int token, int bd, int32_t want) {
    int cat = VP9_EXTRA_BITS[token][0], num_extra = VP9_EXTRA_BITS[token][1];
    int32_t coef = VP9_EXTRA_BITS[token][2];
    const int32_t extra = want - coef;
    if (token == DCT_VAL_CAT6)
      for (int e = 0; e < bd - 8; e++) {
        int32_t high = d.bit(255, (extra >> (5 + bd - e)) & 1);
        coef += high << (5 + bd - e);
      }
    const uint8_t *p = VP9_CAT_PROBS[cat];
    for (int e = 0; e < num_extra; e++) {
      int32_t bit = d.bit(p[e], (extra >> (num_extra - 1 - e)) & 1);
      coef += bit << (num_extra - 1 - e);
    }
    return coef;
  }

  // the token of a coefficient value (6.4.24's token tree leaves)
  static int token_of(int32_t v) {
    int32_t a = std::abs(v);
    if (a <= 4) return a;
    if (a <= 6) return 5;
    if (a <= 10) return 6;
    if (a <= 18) return 7;
    if (a <= 34) return 8;
    if (a <= 66) return 9;
    return DCT_VAL_CAT6;
  }
};

// --------------------------------------------------------------------------
// Frame process (decoder/mod.rs:218-422, minus samples).

Parser::Parser() {
  for (auto &c : contexts_) c.set_defaults();
}

bool Parser::decode_tiles(const uint8_t *data, size_t len) {
  mi_.assign((size_t)h_.mi_rows * h_.mi_cols, MiInfo());
  const uint32_t tile_cols = 1u << h_.tile_cols_log2, tile_rows = 1u << h_.tile_rows_log2;
  // every tile's place in the data (each but the last is preceded by its
  // 4-byte size), found before any tile is parsed
  struct Span {
    size_t pos, size;
  };
  std::vector<Span> span(tile_rows * tile_cols);
  size_t pos = 0;
  for (uint32_t i = 0; i < tile_rows * tile_cols; i++) {
    size_t size;
    if (i + 1 == tile_rows * tile_cols) {
      size = len - pos;
    } else {
      if (pos + 4 > len) bitstream("tile length field reaches beyond the frame data");
      size = ((size_t)data[pos] << 24) | ((size_t)data[pos + 1] << 16) | ((size_t)data[pos + 2] << 8) | data[pos + 3];
      pos += 4;
    }
    if (pos + size > len) bitstream("tile data reaches beyond the frame data");
    span[i] = {pos, size};
    pos += size;
  }
  const uint32_t workers = std::min(tile_threads ? tile_threads : 1u, tile_cols);
  if (workers <= 1) {
    // serial: straight into the record
    TileOut out;
    out.blocks.swap(blocks_);
    out.txbs.swap(txbs_);
    out.coef_pos.swap(coef_pos_);
    out.coef_val.swap(coef_val_);
    out.clear();
    TileWalker<BoolDecoder> td(*this, out);
    ParseError err{0, ""};
    try {
      for (uint32_t i = 0; i < tile_rows * tile_cols; i++)
        td.decode_tile(data + span[i].pos, span[i].size, i / tile_cols, i % tile_cols);
    } catch (const ParseError &e) {
      err = e;
    }
    out.blocks.swap(blocks_);
    out.txbs.swap(txbs_);
    out.coef_pos.swap(coef_pos_);
    out.coef_val.swap(coef_val_);
    counts_ = out.counts;
    if (err.code) throw err;
    return td.padding_ok;
  }
  // Tile columns in parallel (the reference's decode_tiles_parallel,
  // src/decoder/block.rs:1769-1866): a column's tiles read only mode info
  // of their own columns (above context, left context reset per tile,
  // motion vector candidates inside the tile's columns), so columns are
  // independent; tile rows of one column stay in order.  Each tile writes
  // its own part of the record; the parts are joined in the serial order
  // (tile rows, then columns), and the symbol counts summed, so the record
  // is the serial one.  The error reported is the first in that order.
  tile_out_.resize(tile_rows * tile_cols);
  if (!pool_.p || pool_.p->threads() != workers) pool_.p.reset(new TaskPool(workers));
  std::vector<ParseError> errs(tile_rows * tile_cols, ParseError{0, ""});
  std::vector<uint8_t> pad(tile_cols, 1);
  pool_.p->run(tile_cols, [&](uint32_t col) {
    for (uint32_t row = 0; row < tile_rows; row++) {
      const uint32_t i = row * tile_cols + col;
      TileOut &o = tile_out_[i];
      o.clear();
      try {
        TileWalker<BoolDecoder> td(*this, o);
        td.decode_tile(data + span[i].pos, span[i].size, row, col);
        pad[col] &= td.padding_ok;
      } catch (const ParseError &e) {
        errs[i] = e;
        return;  // the later tiles of this column read this one's mode info
      }
    }
  });
  for (uint32_t i = 0; i < tile_rows * tile_cols; i++)
    if (errs[i].code) throw errs[i];
  // the parts' places in the record, then every part copied at once
  const uint32_t n_tiles = tile_rows * tile_cols;
  std::vector<size_t> b0(n_tiles + 1, 0), t0(n_tiles + 1, 0), c0(n_tiles + 1, 0);
  for (uint32_t i = 0; i < n_tiles; i++) {
    b0[i + 1] = b0[i] + tile_out_[i].blocks.size();
    t0[i + 1] = t0[i] + tile_out_[i].txbs.size();
    c0[i + 1] = c0[i] + tile_out_[i].coef_pos.size();
  }
  blocks_.resize(b0[n_tiles]);
  txbs_.resize(t0[n_tiles]);
  coef_pos_.resize(c0[n_tiles]);
  coef_val_.resize(c0[n_tiles]);
  pool_.p->run(n_tiles, [&](uint32_t i) {
    const TileOut &o = tile_out_[i];
    if (!o.blocks.empty()) memcpy(&blocks_[b0[i]], o.blocks.data(), sizeof(vp9r_block) * o.blocks.size());
    for (size_t k = 0; k < o.txbs.size(); k++) {
      vp9r_txb t = o.txbs[k];
      t.block += (uint32_t)b0[i];
      t.coef_start += (uint32_t)c0[i];
      txbs_[t0[i] + k] = t;
    }
    if (!o.coef_pos.empty()) {
      memcpy(&coef_pos_[c0[i]], o.coef_pos.data(), sizeof(uint16_t) * o.coef_pos.size());
      memcpy(&coef_val_[c0[i]], o.coef_val.data(), sizeof(int32_t) * o.coef_val.size());
    }
  });
  memset(&counts_, 0, sizeof counts_);
  for (uint32_t i = 0; i < n_tiles; i++) {
    uint32_t *dst = (uint32_t *)&counts_;
    const uint32_t *src = (const uint32_t *)&tile_out_[i].counts;
    for (size_t k = 0; k < sizeof(Counts) / sizeof(uint32_t); k++) dst[k] += src[k];
  }
  bool ok = true;
  for (uint32_t c = 0; c < tile_cols; c++) ok = ok && pad[c];
  return ok;
}

TaskPool::TaskPool(uint32_t threads) {
  for (uint32_t t = 1; t < threads; t++) workers_.emplace_back([this] { work(); });
}

TaskPool::~TaskPool() {
  {
    std::lock_guard<std::mutex> l(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto &t : workers_) t.join();
}

void TaskPool::run(uint32_t n, const std::function<void(uint32_t)> &fn) {
  std::unique_lock<std::mutex> l(mu_);
  fn_ = &fn;
  n_ = n;
  next_ = 0;
  gen_++;
  cv_.notify_all();
  // this thread takes indices too
  while (next_ < n_) {
    uint32_t i = next_++;
    busy_++;
    l.unlock();
    fn(i);
    l.lock();
    busy_--;
  }
  done_cv_.wait(l, [&] { return busy_ == 0; });
  fn_ = nullptr;
}

void TaskPool::work() {
  std::unique_lock<std::mutex> l(mu_);
  uint64_t seen = 0;
  for (;;) {
    cv_.wait(l, [&] { return stop_ || (gen_ != seen && fn_ && next_ < n_); });
    if (stop_) return;
    seen = gen_;
    while (fn_ && next_ < n_) {
      uint32_t i = next_++;
      busy_++;
      const std::function<void(uint32_t)> *fn = fn_;
      l.unlock();
      (*fn)(i);
      l.lock();
      busy_--;
    }
    if (busy_ == 0) done_cv_.notify_all();
  }
}

void Parser::forget_slots(uint32_t mask) {
  for (int i = 0; i < 8; i++)
    if ((mask >> i) & 1) {
      slots_[i] = SlotInfo();
      ref_sizes_[i][0] = ref_sizes_[i][1] = 0;
    }
}

void Parser::parse_frame(const uint8_t *data, size_t len, vp9r_record *out) {
  if (len == 0) bitstream("zero-length frame");
  uint32_t saved_last_type = st_last_frame_type_;
  int reset_kind;
  parse_uncompressed(data, len, &reset_kind);
  FrameHeader &h = h_;
  memset(out, 0, sizeof *out);
  if (h.show_existing_frame) {
    st_last_frame_type_ = saved_last_type;
    const SlotInfo &s = slots_[h.frame_to_show_map_idx];
    if (!s.valid) bitstream("show_existing_frame names a slot that holds no frame");
    frames++;
    vp9r_frame_hdr &o = out->hdr;
    o.show_existing_frame = 1;
    o.frame_to_show_map_idx = (uint8_t)h.frame_to_show_map_idx;
    o.show_frame = 1;
    o.width = s.width;
    o.height = s.height;
    o.bit_depth = (uint8_t)s.bit_depth;
    o.ss_x = (uint8_t)s.ss_x;
    o.ss_y = (uint8_t)s.ss_y;
    return;
  }
  setup_frame(reset_kind);
  size_t mi_len = (size_t)h.mi_rows * h.mi_cols;
  size_t comp_start = h.uncompressed_size;
  size_t comp_end = comp_start + h.header_size_in_bytes;
  if (comp_end > len) bitstream("frame data ends inside the compressed header");
  parse_compressed(data + comp_start, h.header_size_in_bytes);
  const uint8_t *tiles = data + comp_end;
  size_t tiles_len = len - comp_end;
  parse_tiles_and_finish(tiles, tiles_len, mi_len, out);
}

// The frame process between the uncompressed header and the compressed one:
// format and size checks, the active references, compute_image_size()
// semantics (7.2.6), context resets and load_probs (decoder/mod.rs:232-291).
void Parser::setup_frame(int reset_kind) {
  FrameHeader &h = h_;
  if ((uint64_t)h.width * h.height > max_pixels)
    fail(VP9R_ERR_UNSUPPORTED, "frame size is above the decoder's limit");
  if (h.header_size_in_bytes == 0) bitstream("header_size_in_bytes is 0");
  // the active references
  for (int i = 0; i < 3; i++) ref_ok_[i] = 0;
  if (!h.frame_is_intra) {
    for (int i = 0; i < 3; i++) {
      const SlotInfo &s = slots_[h.ref_frame_idx[i]];
      if (!s.valid) bitstream("active reference slot holds no frame");
      if (s.ss_x != h.ss_x || s.ss_y != h.ss_y || s.bit_depth != h.bit_depth)
        bitstream("active reference has another bit depth or subsampling");
      ref_ok_[i] = scale_ok(s.width, s.height, h.width, h.height) ? 1 : 2;
    }
  }
  // compute_image_size() semantics (7.2.6)
  bool first = !have_last_size_;
  bool same_size = have_last_size_ && last_w_ == h.width && last_h_ == h.height;
  size_t mi_len = (size_t)h.mi_rows * h.mi_cols;
  use_prev_mvs_ =
      !first && same_size && last_show_frame_ && !h.error_resilient_mode && !h.frame_is_intra;
  if (!same_size) prev_segment_ids_.assign(mi_len, 0);
  have_last_size_ = true;
  last_w_ = h.width;
  last_h_ = h.height;
  last_show_frame_ = h.show_frame;
  if (h.past_independence) std::fill(prev_segment_ids_.begin(), prev_segment_ids_.end(), 0);
  if (reset_kind == RESET_ALL)
    for (auto &c : contexts_) c.set_defaults();
  else if (reset_kind == RESET_ONE)
    contexts_[reset_idx_].set_defaults();
  probs_ = contexts_[h.frame_context_idx];
}

void Parser::parse_tiles_and_finish(const uint8_t *tiles, size_t tiles_len, size_t mi_len, vp9r_record *out) {
  FrameHeader &h = h_;
  // A non-4:2:0 stream may be one of the pre-final profile 1 streams that
  // size chroma transforms as 4:2:0 (decided at intra frames: the legacy
  // rules are taken if the specification's leave a tile unparsable or with
  // nonzero padding and the legacy ones parse cleanly; decoder/mod.rs:297-349).
  bool non420 = !(h.ss_x == 1 && h.ss_y == 1);
  h.legacy_uv = non420 && !h.frame_is_intra && legacy_uv_;
  bool ok = false;
  ParseError err{0, ""};
  try {
    ok = decode_tiles(tiles, tiles_len);
    err.code = 0;
  } catch (const ParseError &e) {
    err = e;
  }
  if (h.frame_is_intra) {
    if (non420 && !(err.code == 0 && ok)) {
      h.legacy_uv = true;
      bool ok2 = false;
      bool good = false;
      try {
        ok2 = decode_tiles(tiles, tiles_len);
        good = ok2;
      } catch (const ParseError &) {
        good = false;
      }
      if (good) {
        err.code = 0;
      } else {
        h.legacy_uv = false;
        if (err.code == 0) {
          // the specification's parse succeeded (nonzero padding only): redo it
          decode_tiles(tiles, tiles_len);
        } else {
          throw err;
        }
      }
    }
    legacy_uv_ = h.legacy_uv;
  }
  if (err.code) throw err;
  // backward adaptation (refresh_probs, 6.1.2)
  if (!h.error_resilient_mode && !h.frame_parallel_decoding_mode) {
    const Probs &pre = contexts_[h.frame_context_idx];
    Probs adapted = pre;
    memcpy(adapted.tx, probs_.tx, sizeof adapted.tx);
    memcpy(adapted.skip, probs_.skip, sizeof adapted.skip);
    adapt_coef_probs(adapted, counts_, h.frame_is_intra, h.last_frame_type == KEY_FRAME);
    if (!h.frame_is_intra) {
      memcpy(adapted.tx, pre.tx, sizeof adapted.tx);
      memcpy(adapted.skip, pre.skip, sizeof adapted.skip);
      adapt_noncoef_probs(adapted, counts_, h.interp_filter == SWITCHABLE,
                          h.tx_mode == TX_MODE_SELECT, h.allow_high_precision_mv);
    }
    probs_ = adapted;
  }
  if (h.refresh_frame_context) contexts_[h.frame_context_idx] = probs_;
  // the segmentation map and motion vectors the next frame reads
  if (seg_.enabled && seg_.update_map) {
    prev_segment_ids_.resize(mi_len);
    for (size_t i = 0; i < mi_len; i++) prev_segment_ids_[i] = mi_[i].segment_id;
  }
  prev_mvs_.resize(mi_len);
  for (size_t i = 0; i < mi_len; i++) {
    PrevMv &pm = prev_mvs_[i];
    pm.ref_frame[0] = mi_[i].ref_frame[0];
    pm.ref_frame[1] = mi_[i].ref_frame[1];
    for (int l = 0; l < 2; l++) {
      pm.mv[l][0] = mi_[i].mv[l][3][0];
      pm.mv[l][1] = mi_[i].mv[l][3][1];
    }
  }
  // the record
  vp9r_frame_hdr &o = out->hdr;
  o.width = h.width;
  o.height = h.height;
  o.mi_cols = h.mi_cols;
  o.mi_rows = h.mi_rows;
  o.sb64_cols = h.sb64_cols;
  o.sb64_rows = h.sb64_rows;
  o.render_width = h.render_width;
  o.render_height = h.render_height;
  o.profile = (uint8_t)h.profile;
  o.bit_depth = (uint8_t)h.bit_depth;
  o.ss_x = (uint8_t)h.ss_x;
  o.ss_y = (uint8_t)h.ss_y;
  o.frame_type = (uint8_t)h.frame_type;
  o.show_frame = h.show_frame;
  o.intra_only = h.intra_only;
  o.frame_is_intra = h.frame_is_intra;
  o.error_resilient_mode = h.error_resilient_mode;
  o.lossless = h.lossless;
  o.refresh_frame_flags = (uint8_t)h.refresh_frame_flags;
  o.legacy_uv = h.legacy_uv;
  o.interp_filter = (uint8_t)h.interp_filter;
  o.tx_mode = (uint8_t)h.tx_mode;
  for (int i = 0; i < 3; i++) o.ref_frame_idx[i] = (uint8_t)h.ref_frame_idx[i];
  o.reference_mode = (uint8_t)h.reference_mode;
  o.color_space = (uint8_t)h.color_space;
  o.color_range = h.color_range;
  o.tile_cols_log2 = (uint8_t)h.tile_cols_log2;
  o.tile_rows_log2 = (uint8_t)h.tile_rows_log2;
  o.base_q_idx = h.base_q_idx;
  o.delta_q_y_dc = h.delta_q_y_dc;
  o.delta_q_uv_dc = h.delta_q_uv_dc;
  o.delta_q_uv_ac = h.delta_q_uv_ac;
  o.lf_level = lf_.level;
  o.lf_sharpness = lf_.sharpness;
  o.lf_delta_enabled = lf_.delta_enabled;
  memcpy(o.lf_ref_deltas, lf_.ref_deltas, 4);
  memcpy(o.lf_mode_deltas, lf_.mode_deltas, 2);
  o.seg_enabled = seg_.enabled;
  o.seg_abs_or_delta_update = seg_.abs_or_delta_update;
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 4; j++) {
      o.seg_feature_enabled[i][j] = seg_.feature_enabled[i][j];
      o.seg_feature_data[i][j] = seg_.feature_data[i][j];
    }
  if (!h.frame_is_intra)
    for (int i = 0; i < 3; i++) {
      o.ref_width[i] = slots_[h.ref_frame_idx[i]].width;
      o.ref_height[i] = slots_[h.ref_frame_idx[i]].height;
    }
  lf_mi_.resize(mi_len);
  for (size_t i = 0; i < mi_len; i++) {
    const MiInfo &m = mi_[i];
    vp9r_mi &l = lf_mi_[i];
    l.mi_size = m.mi_size;
    l.skip = m.skip;
    l.tx_size = m.tx_size;
    l.y_mode = m.y_mode;
    l.segment_id = m.segment_id;
    l.ref_frame0 = m.ref_frame[0];
    l._pad[0] = l._pad[1] = 0;
  }
  out->blocks = blocks_.data();
  out->n_blocks = (uint32_t)blocks_.size();
  out->txbs = txbs_.data();
  out->n_txbs = (uint32_t)txbs_.size();
  out->coef_pos = coef_pos_.data();
  out->coef_val = coef_val_.data();
  out->n_coefs = (uint32_t)coef_pos_.size();
  out->mi = lf_mi_.data();
  // reference slot bookkeeping (8.10)
  for (int i = 0; i < 8; i++)
    if ((h.refresh_frame_flags >> i) & 1) {
      slots_[i].valid = true;
      slots_[i].width = h.width;
      slots_[i].height = h.height;
      slots_[i].ss_x = h.ss_x;
      slots_[i].ss_y = h.ss_y;
      slots_[i].bit_depth = h.bit_depth;
      ref_sizes_[i][0] = h.width;
      ref_sizes_[i][1] = h.height;
    }
  frames++;
}

// --------------------------------------------------------------------------
// Stream generation: the boolean encoder (the inverse of 9.2) and the frame
// writer over the same syntax walk.

class BoolWriter {
 public:
  static constexpr bool kWrite = true;
  std::vector<uint8_t> out;
  BoolWriter() { put(false, 128); }  // the marker bit
  inline bool bit(uint32_t prob, bool v) {
    put(v, prob);
    return v;
  }
  uint32_t literal(int n, uint32_t v = 0) {
    for (int i = n - 1; i >= 0; i--) put((v >> i) & 1, 128);
    return v;
  }
  // codes leaf `value` of tree t
  template <typename F>
  int tree(const int8_t *t, F prob, int value) {
    int path[32], depth = 0;
    if (!find(t, 0, value, path, &depth)) bitstream("stream generator: value outside its tree");
    int n = 0;
    for (int k = 0; k < depth; k++) {
      put(path[k], prob(n >> 1));
      n = t[n + path[k]];
    }
    return value;
  }
  // exit: 32 zero bits flush the value; never end on a byte that looks like
  // a superframe marker
  void finish() {
    for (int i = 0; i < 32; i++) put(false, 128);
    if (!out.empty() && (out.back() & 0xe0) == 0xc0) out.push_back(0);
  }

 private:
  static bool find(const int8_t *t, int n, int value, int *path, int *depth) {
    for (int b = 0; b < 2; b++) {
      int next = t[n + b];
      path[*depth] = b;
      (*depth)++;
      if (next <= 0) {
        if (-next == value) return true;
      } else if (find(t, next, value, path, depth)) {
        return true;
      }
      (*depth)--;
    }
    return false;
  }
  // low: the coded value's pending low bits (24 + carry), `bits`: how many
  // more normalisation shifts before the next byte is complete
  void put(bool v, uint32_t prob) {
    uint32_t split = 1 + (((range_ - 1) * prob) >> 8);
    if (v) {
      low_ += split;
      range_ -= split;
    } else {
      range_ = split;
    }
    int shift = __builtin_clz(range_) - 24;
    range_ <<= shift;
    bits_ += shift;
    if (bits_ >= 0) {
      int offset = shift - bits_;
      if ((low_ << (offset - 1)) & 0x80000000u) {
        // carry into the bytes already written
        size_t k = out.size();
        while (k > 0 && out[k - 1] == 0xff) out[--k] = 0;
        if (k > 0) out[k - 1]++;
      }
      out.push_back((uint8_t)(low_ >> (24 - offset)));
      low_ <<= offset;
      shift = bits_;
      low_ &= 0xffffff;
      bits_ -= 8;
    }
    low_ <<= shift;
  }
  uint32_t low_ = 0, range_ = 255;
  int bits_ = -24;
};

std::vector<uint8_t> Parser::write_frame(const uint8_t *uncompressed, size_t len, uint32_t tx_mode,
                                         uint32_t reference_mode, SyntaxPlan *plan) {
  int reset_kind;
  parse_uncompressed(uncompressed, len, &reset_kind);
  if (h_.show_existing_frame) bitstream("stream generator: nothing to write for show_existing_frame");
  if (h_.uncompressed_size != len) bitstream("stream generator: uncompressed header length mismatch");
  setup_frame(reset_kind);
  h_.tx_mode = tx_mode;
  h_.reference_mode = reference_mode;
  BoolWriter cw;
  compressed_syntax(cw);
  cw.finish();
  h_.legacy_uv = false;
  blocks_.clear();
  txbs_.clear();
  coef_pos_.clear();
  coef_val_.clear();
  memset(&counts_, 0, sizeof counts_);
  mi_.assign((size_t)h_.mi_rows * h_.mi_cols, MiInfo());
  TileOut wout;
  wout.blocks.swap(blocks_);
  wout.txbs.swap(txbs_);
  wout.coef_pos.swap(coef_pos_);
  wout.coef_val.swap(coef_val_);
  wout.clear();
  TileWalker<BoolWriter> tw(*this, wout);
  tw.plan = plan;
  std::vector<uint8_t> frame(uncompressed, uncompressed + len);
  // header_size_in_bytes (16 bits at header_size_pos)
  if (cw.out.size() > 0xffff) bitstream("stream generator: compressed header too large");
  for (int i = 0; i < 16; i++) {
    size_t bitpos = h_.header_size_pos + i;
    uint8_t mask = (uint8_t)(0x80 >> (bitpos & 7));
    if ((cw.out.size() >> (15 - i)) & 1)
      frame[bitpos >> 3] |= mask;
    else
      frame[bitpos >> 3] &= (uint8_t)~mask;
  }
  frame.insert(frame.end(), cw.out.begin(), cw.out.end());
  uint32_t tile_cols = 1u << h_.tile_cols_log2, tile_rows = 1u << h_.tile_rows_log2;
  for (uint32_t row = 0; row < tile_rows; row++)
    for (uint32_t col = 0; col < tile_cols; col++) {
      BoolWriter tw_d;
      tw.walk_tile(tw_d, row, col);
      tw_d.finish();
      bool last = row == tile_rows - 1 && col == tile_cols - 1;
      if (!last) {
        uint32_t n = (uint32_t)tw_d.out.size();
        uint8_t be[4] = {(uint8_t)(n >> 24), (uint8_t)(n >> 16), (uint8_t)(n >> 8), (uint8_t)n};
        frame.insert(frame.end(), be, be + 4);
      }
      frame.insert(frame.end(), tw_d.out.begin(), tw_d.out.end());
    }
  wout.blocks.swap(blocks_);
  wout.txbs.swap(txbs_);
  wout.coef_pos.swap(coef_pos_);
  wout.coef_val.swap(coef_val_);
  counts_ = wout.counts;
  return frame;
}

}  // namespace vp9r
