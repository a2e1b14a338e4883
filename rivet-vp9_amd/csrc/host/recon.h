// Host side of the GPU reconstruction: the device context, frame stores and
// the per-frame launch sequence (include/vp9r.h, vp9r_gpu_*).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <string>
#include <vector>

#include "../gpu/launch.h"
#include "vp9r.h"

namespace vp9r {

void set_error(const std::string &m);

// A grow-only device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n);
  ~DevBuf();
};

// A grow-only pinned host buffer.
struct HostBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n);
  ~HostBuf();
};

// The device inputs of one frame: the record and its work lists in one
// device buffer (at the offsets below), with the launch geometry.
struct Prepared {
  vp9r_frame_hdr hdr;
  vp9g::FrameGeom geo;
  int32_t q[8][2][2];
  vp9g::LfParams lfp;
  size_t o_blocks = 0, o_txbs = 0, o_cpos = 0, o_cval = 0, o_mi = 0, o_tiles = 0, o_jobs[4] = {},
         o_iidx = 0, o_groups = 0, o_items = 0, o_sbtx = 0, o_sbunits = 0, o_sblevels = 0;
  // the intra items are superblock items (kIbSuperblock), not blocks
  bool intra_sb = false;
  // jobs[k] lists the inter transform blocks with coefficients (n_jobs[k]),
  // then every transform block of the intra blocks (n_ijobs[k], flagged
  // kItxIntra), whose residuals the intra block kernel reads
  uint32_t n_tiles = 0, n_jobs[4] = {}, n_ijobs[4] = {}, n_groups = 0, n_items = 0;
  std::vector<uint32_t> group_level_start, lf_level_start, item_level_start;
  // per loop filter level: the last intra level (of leveled intra blocks)
  // that must be complete before it, -1 for none
  std::vector<int32_t> lf_need;
  size_t bytes = 0;
  // algorithmic HBM bytes of the frame per stage: inter prediction, inter
  // residual, intra, loop filter
  uint64_t work_bytes[4] = {};
};

// Host work lists of one record, reused from record to record (one per
// thread that prepares records).
struct Builder {
  std::vector<vp9g::McTile> tiles;
  std::vector<uint32_t> jobs[4], ijobs[4];
  std::vector<uint32_t> intra_cnt, intra_first, intra_fill, intra_index;
  std::vector<vp9g::IntraItem> intra_items, items_sorted;
  std::vector<vp9g::IntraBlockItem> block_items;  // items_sorted as the kernel reads them
  std::vector<int32_t> cell_level, item_level;
  std::vector<uint32_t> item_level_start;
  std::vector<int32_t> intra_level;
  std::vector<vp9g::IntraGroup> groups;
  std::vector<uint32_t> group_level_start;
  std::vector<uint32_t> lf_level_start;
  uint32_t lf_rows = 0, lf_cols = 0;
  std::vector<int32_t> sb_need;
  // superblock intra items (kIbSuperblock), when the frame uses them
  bool sb_mode = false;
  std::vector<vp9g::IntraSbTx> sb_tx;
  std::vector<uint32_t> sb_units;
  std::vector<uint2> sb_levels;
  std::vector<int32_t> sb_level, sb_cells, sb_tx_level;
  std::vector<uint32_t> sb_order;
};

// Checks a record and writes its inputs and work lists to `stage` (pinned
// host memory) at the offsets it records in p.
int build_record(Builder &w, const vp9r_record *rec, Prepared &p, HostBuf &stage);

// One frame of a batched launch: its prepared inputs (device copy at dv)
// and its frame stores.
struct BatchFrame {
  const Prepared *p;
  const uint8_t *dv;
  vp9r_gpu_frame *refs[3];
  vp9r_gpu_frame *cur;
};

// A device copy of one batch's arguments and launch ranges.
struct BatchSlot {
  DevBuf dev;
  HostBuf host;
  hipEvent_t ev = nullptr;
  bool pending = false;
};
constexpr int kBatchSlots = 4;

// A capture of one frame's launches as a HIP graph: the batch tables go to
// `tables` (owned by the graph, uploaded before the capture begins), and
// only kernel launches are captured (no copies, events or waits).
struct GraphTarget {
  DevBuf tables;
  hipGraph_t graph = nullptr;  // the capture's result
};

// Queues reconstruction (recon) and/or the loop filter (lf) of n frames on
// the context's stream, each launch covering every frame.  With gt (one
// frame, timing off): captures the launches into gt->graph instead.
int launch_batch(vp9r_gpu *g, const BatchFrame *f, int n, bool recon, bool lf, GraphTarget *gt = nullptr);

// Takes / releases a reference to a context (the last release destroys it).
void gpu_retain(vp9r_gpu *g);
void gpu_release(vp9r_gpu *g);

// Events around a timed call's launches (created on first use) on the
// context's stream; after the stream has synchronised, seam_end reads their
// elapsed time into g->seam_ms.
hipError_t seam_begin(vp9r_gpu *g);
hipError_t seam_end(vp9r_gpu *g);

// The launches of a single frame's m loop filter levels after its intra
// levels: `singles` one-level launches, then launches of depths[i] levels
// (an even number of them), at most max_depth each.
constexpr int kLfDepthDefault = 5;  // 3 at 10/12 bits (the cone's LDS)
struct LfPlan {
  size_t singles = 0;
  std::vector<int> depths;
};
LfPlan lf_tail_plan(size_t m, int max_depth);

}  // namespace vp9r

struct vp9r_gpu_frame {
  vp9r_gpu *g = nullptr;
  uint32_t width = 0, height = 0, bit_depth = 8, ss_x = 1, ss_y = 1;
  uint32_t mi_cols = 0, mi_rows = 0, sb64_cols = 0, sb64_rows = 0;
  vp9g::Planes planes{};
  void *base = nullptr;
  size_t bytes = 0;
  // an asynchronous download in flight on the copy stream (the compute
  // stream waits for it before writing the frame again)
  hipEvent_t ready = nullptr, copy_done = nullptr;
  bool copy_pending = false;
};

struct vp9r_gpu {
  // the context and each frame store or device buffer made from it hold a
  // reference: vp9r_gpu_free releases the context's own, and the last
  // release destroys it, so the objects may be freed in any order
  std::atomic<int> refs{1};
  int device = 0;
  hipStream_t stream = nullptr;       // kernels and record uploads
  hipStream_t copy_stream = nullptr;  // frame downloads
  hipStream_t up_stream = nullptr;    // batch table uploads (ahead of the kernels)
  // the loop filter's stream (it overlaps the intra levels it does not
  // need) and its events
  hipStream_t lf_stream = nullptr;
  hipEvent_t ev_up = nullptr, ev_pre = nullptr, ev_lf_done = nullptr;
  std::vector<hipEvent_t> intra_events;
  vp9r::DevBuf arena, lf_mask, intra_res;
  vp9r::HostBuf stage;
  hipEvent_t ev_upload = nullptr;
  bool upload_pending = false;
  // the record of the last vp9r_gpu_reconstruct (its inputs are in `arena`)
  vp9r::Prepared last;
  bool last_valid = false;
  // stage timing: one set of 5 events per timed call (a frame, or a batch
  // of ev_frames[k] frames) since the last read
  bool timing = false;
  std::vector<std::array<hipEvent_t, 5>> ev_sets;
  std::vector<uint32_t> ev_frames;
  size_t ev_used = 0;
  hipEvent_t *ev = nullptr;  // the current call's set
  double stage_acc[5] = {};   // folded sums (ms) of sets already read back
  uint32_t frames_acc = 0;
  // host work lists, reused
  vp9r::Builder work;
  bool intra_blocks = true;  // intra as leveled blocks (else (superblock, plane) groups)
  // a single frame's loop filter levels after its intra levels up to
  // lf_depth per launch, through a second copy of the frame's planes
  int lf_depth = vp9r::kLfDepthDefault;
  vp9r::DevBuf lf_tmp;
  // batched launches
  std::vector<uint8_t> batch_stage;
  vp9r::BatchSlot slots[vp9r::kBatchSlots];
  int slot_next = 0;
  // per-kernel seam scratch, and events around the last batched seam call's
  // kernels (seam_ms: their elapsed time)
  vp9r::DevBuf seam_a, seam_b, seam_c;
  hipEvent_t seam_ev[2] = {};
  float seam_ms = -1.0f;
  // per-kernel seam launch shapes (vp9r_gpu_autotune / vp9r_gpu_set_tuning):
  // waves per workgroup of the inverse transforms per [8-bit, 10/12-bit]
  // [4x4 .. 32x32], and rows per motion compensation tile per [8-bit,
  // 10/12-bit]
  // transform launch shape per bit depth class and size: waves per
  // workgroup | register transposes (16) | persistent (32)
  uint8_t tune_itx_shape[2][4] = {{1, 1, 1, 1}, {1, 1, 1, 1}};
  uint8_t tune_mc_rows[2] = {32, 32};
  bool tuned = false;
};
