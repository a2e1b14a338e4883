// The CPU side of the decoder: everything that turns a VP9 frame into the
// per-frame reconstruction record without touching samples (uncompressed
// and compressed headers, the boolean decoder, probability contexts and
// backward adaptation, partition/mode/motion-vector/token syntax, the
// segmentation map and the previous frame's motion vectors).
//
// It follows the crate's decoder: src/header.rs, src/bool_coder.rs,
// src/probs.rs, src/decoder/block.rs (parse half), src/decoder/mvpred.rs and
// the frame process of src/decoder/mod.rs:218-422, minus the sample work,
// which the record hands to the GPU.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "vp9r.h"

namespace vp9r {

struct ParseError {
  int code;
  std::string msg;
};

typedef int32_t Mv[2];

// One frame context (src/probs.rs:13-86).
struct Probs {
  uint8_t tx[4][2][3];
  uint8_t coef[4][2][2][6][6][3];
  uint8_t skip[3];
  uint8_t inter_mode[7][3];
  uint8_t interp_filter[4][2];
  uint8_t is_inter[4];
  uint8_t comp_mode[5];
  uint8_t single_ref[5][2];
  uint8_t comp_ref[5];
  uint8_t y_mode[4][9];
  uint8_t uv_mode[10][9];
  uint8_t partition[16][3];
  uint8_t mv_joint[3];
  uint8_t mv_sign[2];
  uint8_t mv_class[2][10];
  uint8_t mv_class0_bit[2];
  uint8_t mv_bits[2][10];
  uint8_t mv_class0_fr[2][2][3];
  uint8_t mv_fr[2][3];
  uint8_t mv_class0_hp[2];
  uint8_t mv_hp[2];
  void set_defaults();
};

// Symbol counts of section 8.3 (src/probs.rs:88-118).
struct Counts {
  uint32_t intra_mode[4][10];
  uint32_t uv_mode[10][10];
  uint32_t partition[16][4];
  uint32_t interp_filter[4][3];
  uint32_t inter_mode[7][4];
  uint32_t tx[4][2][4];
  uint32_t is_inter[4][2];
  uint32_t comp_mode[5][2];
  uint32_t single_ref[5][2][2];
  uint32_t comp_ref[5][2];
  uint32_t skip[3][2];
  uint32_t mv_joint[4];
  uint32_t mv_sign[2][2];
  uint32_t mv_class[2][11];
  uint32_t mv_class0_bit[2][2];
  uint32_t mv_class0_fr[2][2][4];
  uint32_t mv_class0_hp[2][2];
  uint32_t mv_bits[2][10][2];
  uint32_t mv_fr[2][4];
  uint32_t mv_hp[2][2];
  uint32_t token[4][2][2][6][6][3];
  uint32_t more_coefs[4][2][2][6][6][2];
};

// A fixed set of worker threads that run the indices of a task together
// with the calling thread (for the tile columns of a frame).
class TaskPool {
 public:
  explicit TaskPool(uint32_t threads);  // threads - 1 workers
  ~TaskPool();
  uint32_t threads() const { return (uint32_t)workers_.size() + 1; }
  // fn(i) for every i < n, on the workers and this thread; returns when all
  // are done.  fn must not throw.
  void run(uint32_t n, const std::function<void(uint32_t)> &fn);

 private:
  void work();
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(uint32_t)> *fn_ = nullptr;
  uint32_t n_ = 0, next_ = 0, busy_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// A parser's pool, which a copy of the parser does not share (a copy starts
// without one and makes its own on first use).
struct PoolHolder {
  std::unique_ptr<TaskPool> p;
  PoolHolder() = default;
  PoolHolder(const PoolHolder &) {}
  PoolHolder &operator=(const PoolHolder &) {
    p.reset();
    return *this;
  }
};

// One tile's part of the record and its symbol counts.
struct TileOut {
  std::vector<vp9r_block> blocks;
  std::vector<vp9r_txb> txbs;
  std::vector<uint16_t> coef_pos;
  std::vector<int32_t> coef_val;
  Counts counts;
  void clear() {
    blocks.clear();
    txbs.clear();
    coef_pos.clear();
    coef_val.clear();
    memset(&counts, 0, sizeof counts);
  }
};

struct LoopFilterParams {
  uint8_t level = 0, sharpness = 0;
  bool delta_enabled = true;
  int8_t ref_deltas[4] = {1, 0, -1, -1};
  int8_t mode_deltas[2] = {0, 0};
};

struct SegParams {
  bool enabled = false, update_map = false, temporal_update = false, abs_or_delta_update = false;
  uint8_t tree_probs[7] = {255, 255, 255, 255, 255, 255, 255};
  uint8_t pred_probs[3] = {255, 255, 255};
  bool feature_enabled[8][4] = {};
  int16_t feature_data[8][4] = {};
  bool active(int seg, int feature) const { return enabled && feature_enabled[seg][feature]; }
};

// Everything the uncompressed and compressed headers say about a frame.
struct FrameHeader {
  uint32_t profile = 0;
  bool show_existing_frame = false;
  uint32_t frame_to_show_map_idx = 0;
  uint32_t frame_type = 0, last_frame_type = 0;
  bool show_frame = false, error_resilient_mode = false, intra_only = false;
  uint32_t reset_frame_context = 0;
  uint32_t bit_depth = 8, color_space = 1;
  bool color_range = false;
  uint32_t ss_x = 1, ss_y = 1;
  uint32_t refresh_frame_flags = 0;
  uint32_t ref_frame_idx[3] = {0, 0, 0};
  bool ref_frame_sign_bias[4] = {false, false, false, false};
  uint32_t width = 0, height = 0, render_width = 0, render_height = 0;
  bool allow_high_precision_mv = false;
  uint32_t interp_filter = 0;
  bool refresh_frame_context = false, frame_parallel_decoding_mode = false;
  uint32_t frame_context_idx = 0;
  int32_t base_q_idx = 0, delta_q_y_dc = 0, delta_q_uv_dc = 0, delta_q_uv_ac = 0;
  bool lossless = false;
  uint32_t tile_cols_log2 = 0, tile_rows_log2 = 0;
  uint32_t header_size_in_bytes = 0;
  size_t header_size_pos = 0;  // bit position of header_size_in_bytes
  size_t uncompressed_size = 0;
  bool past_independence = false;
  bool frame_is_intra = false;
  uint32_t mi_cols = 0, mi_rows = 0, sb64_cols = 0, sb64_rows = 0;
  uint32_t tx_mode = 0, reference_mode = 0;
  int32_t comp_fixed_ref = 0;
  int32_t comp_var_ref[2] = {0, 0};
  bool legacy_uv = false;
  void compute_image_size() {
    mi_cols = (width + 7) >> 3;
    mi_rows = (height + 7) >> 3;
    sb64_cols = (mi_cols + 7) >> 3;
    sb64_rows = (mi_rows + 7) >> 3;
  }
};

// What the frame remembers of each 8x8 position (block.rs:19-31).
struct MiInfo {
  uint8_t mi_size = 0, skip = 0, tx_size = 0, y_mode = 0;
  uint8_t sub_modes[4] = {0, 0, 0, 0};
  uint8_t segment_id = 0, interp_filter = 0;
  int8_t ref_frame[2] = {0, -1};
  int32_t mv[2][4][2] = {};
};

struct PrevMv {
  int8_t ref_frame[2] = {0, 0};
  int32_t mv[2][2] = {};
};

// Format of a reference slot, as the parser tracks it (no samples).
struct SlotInfo {
  bool valid = false;
  uint32_t width = 0, height = 0, ss_x = 0, ss_y = 0, bit_depth = 0;
};

// The decisions a stream generator supplies to the syntax walk when it
// writes instead of reads (see TileWalker in parser.cpp): partitions, each
// block's modes and motion vectors, each transform block's quantised
// coefficients.  Not used when decoding.
struct PlanBlock {
  uint8_t segment_id, skip, tx_size, is_inter;
  uint8_t y_mode, sub_modes[4], uv_mode, interp_filter;
  int8_t ref_frame[2];
  int32_t mv[2][2];  // the motion vector of each list (1/8 sample)
};

class SyntaxPlan {
 public:
  virtual ~SyntaxPlan() {}
  virtual int partition(uint32_t mi_row, uint32_t mi_col, int bsize) = 0;
  virtual void block(uint32_t mi_row, uint32_t mi_col, int bsize, bool frame_is_intra, PlanBlock *out) = 0;
  // fills the raster coefficients (zeroed on entry) of a transform block
  virtual void coefs(int plane, uint32_t x, uint32_t y, int tx_size, int tx_type, int32_t *raster) = 0;
};

class Parser {
 public:
  Parser();
  // Parses one frame into the record (see vp9r.h); throws ParseError.
  void parse_frame(const uint8_t *data, size_t len, vp9r_record *out);
  uint64_t max_pixels = 8192ull * 8192ull;
  uint64_t frames = 0;
  // Threads that parse tile columns at once (1: the calling thread only).
  // The record is the same with any number.
  uint32_t tile_threads = 1;
  // Marks the slots of `mask` empty.  A caller whose reconstruction of a
  // parsed frame failed calls it with the frame's refresh_frame_flags, so
  // that the parser's view of the slots stays the decoder's: a later frame
  // that names those slots is then a bitstream error, as with slots never
  // filled.
  void forget_slots(uint32_t mask);
  // Stream generation: given a frame's uncompressed header (its
  // header_size_in_bytes any nonzero placeholder), writes the compressed
  // header (tx_mode and reference_mode as given, no probability updates) and
  // the tile data with the plan's decisions, and returns the whole frame.
  // Changes this parser's state as parsing the header would: call it on a
  // copy of the decoder that will read the frame.
  std::vector<uint8_t> write_frame(const uint8_t *uncompressed, size_t len, uint32_t tx_mode,
                                   uint32_t reference_mode, SyntaxPlan *plan);

 private:
  // persistent state (header.rs:58-84, decoder/mod.rs:122-139)
  LoopFilterParams lf_;
  SegParams seg_;
  uint32_t ref_sizes_[8][2] = {};
  uint32_t st_last_frame_type_ = KEY_FRAME_;
  uint32_t st_bit_depth_ = 8, st_ss_x_ = 1, st_ss_y_ = 1, st_color_space_ = 1;
  bool st_color_range_ = false;
  Probs contexts_[4];
  SlotInfo slots_[8];
  std::vector<uint8_t> prev_segment_ids_;
  std::vector<PrevMv> prev_mvs_;
  bool have_last_size_ = false;
  uint32_t last_w_ = 0, last_h_ = 0;
  bool last_show_frame_ = false;
  bool legacy_uv_ = false;
  static constexpr uint32_t KEY_FRAME_ = 0;

  // the frame being parsed
  FrameHeader h_;
  Probs probs_;
  Counts counts_;
  std::vector<MiInfo> mi_;
  bool use_prev_mvs_ = false;
  uint32_t reset_idx_ = 0;
  uint32_t ref_ok_[3] = {0, 0, 0};  // 1 if the active reference may be used

  // record
  std::vector<vp9r_block> blocks_;
  std::vector<vp9r_txb> txbs_;
  std::vector<uint16_t> coef_pos_;
  std::vector<int32_t> coef_val_;
  std::vector<vp9r_mi> lf_mi_;
  std::vector<TileOut> tile_out_;  // per tile, when tile columns are parsed in parallel
  PoolHolder pool_;  // tile_threads threads, made on first use

  void parse_uncompressed(const uint8_t *data, size_t len, int *reset_kind);
  void setup_frame(int reset_kind);
  void parse_tiles_and_finish(const uint8_t *tiles, size_t tiles_len, size_t mi_len, vp9r_record *out);
  void parse_compressed(const uint8_t *data, size_t len);
  // returns true if every tile's padding was zero
  bool decode_tiles(const uint8_t *data, size_t len);
  void fill_record(vp9r_record *out);
  template <class Coder>
  friend class TileWalker;
  template <class Coder>
  void compressed_syntax(Coder &d);
};

}  // namespace vp9r

// The C ABI's parser handle.
struct vp9r_parser {
  vp9r::Parser p;
};
