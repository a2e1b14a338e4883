// Host side of the GPU reconstruction: the device context, frame stores and
// the per-frame launch sequence (include/vp9r.h, vp9r_gpu_*).
#pragma once

#include <hip/hip_runtime.h>

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

}  // namespace vp9r

struct vp9r_gpu_frame {
  vp9r_gpu *g = nullptr;
  uint32_t width = 0, height = 0, bit_depth = 8, ss_x = 1, ss_y = 1;
  uint32_t mi_cols = 0, mi_rows = 0, sb64_cols = 0, sb64_rows = 0;
  vp9g::Planes planes{};
  void *base = nullptr;
  size_t bytes = 0;
};

struct vp9r_gpu {
  int device = 0;
  hipStream_t stream = nullptr;
  vp9r::DevBuf arena, lf_arena, lf_mask;
  vp9r::HostBuf stage, lf_stage;
  hipEvent_t ev_upload = nullptr, ev_lf_upload = nullptr;
  bool upload_pending = false, lf_upload_pending = false;
  bool timing = false;
  hipEvent_t ev[5] = {};
  float stage_ms[5] = {};
  // host work lists, reused
  std::vector<vp9g::McTile> tiles;
  std::vector<uint32_t> jobs[4];
  std::vector<uint32_t> intra_cnt, intra_first, intra_fill, intra_index;
  std::vector<int32_t> intra_level;
  std::vector<vp9g::IntraGroup> groups;
  std::vector<uint32_t> group_level_start;
  std::vector<uint32_t> lf_codes, lf_level_start;
  uint32_t lf_rows = 0, lf_cols = 0;
  // per-kernel seam scratch
  vp9r::DevBuf seam_a, seam_b, seam_c;
};
