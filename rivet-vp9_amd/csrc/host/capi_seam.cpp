// The per-kernel seam of include/vp9r.h: src/dsp's kernels as batched GPU
// launches (for the kernel microbenchmarks and parity tests).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "recon.h"

using namespace vp9r;

static int dev_err(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return VP9R_ERR_DEVICE;
}
#define TRY(expr, what)                         \
  do {                                          \
    hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return dev_err(e_, what); \
  } while (0)

static bool bd_ok(uint32_t bd) { return bd == 8 || bd == 10 || bd == 12; }

// ---- the allocation-free motion compensation entry point
//
// Its descriptors and reference planes are staged in `scratch`, a
// file-scope device array of fixed size: each GPU holds its own copy until
// the process ends, found once per device by hipGetSymbolAddress (the name
// is a host-side stand-in, never a device address).  Calls that overlap on
// one GPU share it and race.
constexpr uint32_t kNoallocMaxDescs = 65536;
constexpr uint32_t kNoallocMaxRefs = 64;
constexpr size_t kNoallocDescBytes = sizeof(vp9r_mc_desc) * kNoallocMaxDescs;
constexpr size_t kScratchBytes = kNoallocDescBytes + sizeof(vp9r_ref_plane) * kNoallocMaxRefs;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static uint8_t *scratch_ptr[kMaxDevices];  // per device, once found

// The device address of this device's scratch (the device current), or
// null with the error set.
static uint8_t *scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    scratch_ptr[device] = (uint8_t *)p;
  }
  return scratch_ptr[device];
}


extern "C" {

int vp9r_gpu_malloc(vp9r_gpu *g, size_t bytes, void **out) {
  *out = nullptr;
  TRY(hipSetDevice(g->device), "hipSetDevice");
  TRY(hipMalloc(out, bytes ? bytes : 1), "hipMalloc");
  gpu_retain(g);
  return VP9R_OK;
}

void vp9r_gpu_mfree(vp9r_gpu *g, void *p) {
  if (!p) return;
  (void)hipStreamSynchronize(g->stream);
  (void)hipFree(p);
  gpu_release(g);
}

int vp9r_gpu_to_device(vp9r_gpu *g, void *dst, const void *src, size_t bytes) {
  TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g->stream), "copy to device");
  TRY(hipStreamSynchronize(g->stream), "copy sync");
  return VP9R_OK;
}

int vp9r_gpu_to_host(vp9r_gpu *g, void *dst, const void *src, size_t bytes) {
  TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g->stream), "copy to host");
  TRY(hipStreamSynchronize(g->stream), "copy sync");
  return VP9R_OK;
}

int vp9r_gpu_seam_ms(vp9r_gpu *g, float *ms) {
  *ms = g->seam_ms;
  if (g->seam_ms < 0) {
    set_error("no timed batched call");
    return VP9R_ERR_INVALID_INPUT;
  }
  return VP9R_OK;
}

void *vp9r_gpu_frame_plane(vp9r_gpu_frame *f, int plane) {
  return (plane >= 0 && plane < 3) ? f->planes.p[plane] : nullptr;
}

int vp9r_itx_add_batch(vp9r_gpu *g, uint32_t count, const int32_t *coefs, const uint32_t *coef_off,
                       const uint8_t *tx, const uint8_t *tx_type, const uint16_t *eob,
                       const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bit_depth,
                       int lossless) {
  if (!bd_ok(bit_depth)) {
    set_error("bit depth must be 8, 10 or 12");
    return VP9R_ERR_INVALID_INPUT;
  }
  std::vector<uint32_t> jobs[4];
  for (uint32_t i = 0; i < count; i++) {
    if (tx[i] > 3 || (lossless && tx[i] != 0)) {
      set_error("transform size out of range (lossless is 4x4 only)");
      return VP9R_ERR_INVALID_INPUT;
    }
    jobs[tx[i]].push_back(i);
  }
  TRY(hipSetDevice(g->device), "hipSetDevice");
  // the job lists of the four sizes, largest first, in one upload; then the
  // launches back to back
  std::vector<uint32_t> all;
  all.reserve(count);
  uint32_t start[4];
  for (int k = 3; k >= 0; k--) {
    start[k] = (uint32_t)all.size();
    all.insert(all.end(), jobs[k].begin(), jobs[k].end());
  }
  TRY(g->seam_a.ensure(sizeof(uint32_t) * std::max<uint32_t>(count, 1)), "hipMalloc");
  uint32_t *dj = (uint32_t *)g->seam_a.p;
  if (count) TRY(hipMemcpyAsync(dj, all.data(), 4 * (size_t)count, hipMemcpyHostToDevice, g->stream), "upload jobs");
  TRY(seam_begin(g), "hipEventRecord");
  for (int k = 3; k >= 0; k--)
    vp9g::launch_itx_dense(g->stream, count, coefs, coef_off, nullptr, tx_type, eob, dst_off, dst, stride,
                           bit_depth, lossless, k, dj + start[k], (uint32_t)jobs[k].size(),
                           g->tune_itx_shape[bit_depth > 8][k]);
  TRY(hipGetLastError(), "itx launch");
  TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
  TRY(hipStreamSynchronize(g->stream), "itx sync");
  TRY(seam_end(g), "hipEventElapsedTime");
  return VP9R_OK;
}

int vp9r_mc_batch_planes(vp9r_gpu *g, uint32_t count, const vp9r_mc_desc *desc, const vp9r_ref_plane *planes,
                         uint32_t n_refs, void *out, uint32_t out_stride, uint32_t bit_depth) {
  if (!bd_ok(bit_depth) || n_refs == 0 || n_refs > 64) {
    set_error("invalid motion compensation batch");
    return VP9R_ERR_INVALID_INPUT;
  }
  for (uint32_t k = 0; k < n_refs; k++)
    if (!planes[k].data || planes[k].last_x < 0 || planes[k].last_y < 0 ||
        (uint32_t)planes[k].last_x >= planes[k].stride) {
      set_error("invalid reference plane");
      return VP9R_ERR_INVALID_INPUT;
    }
  struct DescTile {
    uint32_t desc;
    uint16_t tx, ty;
  };
  const int rows = g->tune_mc_rows[bit_depth > 8] == 16 ? 16 : 32;
  std::vector<DescTile> tiles;
  for (uint32_t i = 0; i < count; i++) {
    const vp9r_mc_desc &d = desc[i];
    bool ok = d.w >= 4 && d.w <= 64 && d.h >= 4 && d.h <= 64 && d.xs >= 1 && d.xs <= 32 && d.ys >= 1 &&
              d.ys <= 32 && d.filter < 4 && d.ref_idx >= 0 && (uint32_t)d.ref_idx < n_refs &&
              d.ref2_idx < (int)n_refs;
    if (!ok) {
      set_error("motion compensation descriptor out of range");
      return VP9R_ERR_INVALID_INPUT;
    }
    // tiles of 16 columns by up to `rows` rows (16 or 32, the tuned
    // choice; the unscaled paths filter all rows in one pass, the general
    // path in halves of 16)
    for (int ty = 0; ty < d.h; ty += rows)
      for (int tx = 0; tx < d.w; tx += 16) tiles.push_back({i, (uint16_t)tx, (uint16_t)ty});
  }
  TRY(hipSetDevice(g->device), "hipSetDevice");
  size_t dbytes = sizeof(vp9r_mc_desc) * count, tbytes = sizeof(DescTile) * tiles.size();
  TRY(g->seam_a.ensure(dbytes + 256), "hipMalloc");
  TRY(g->seam_b.ensure(tbytes + 256), "hipMalloc");
  TRY(g->seam_c.ensure(sizeof(vp9r_ref_plane) * n_refs), "hipMalloc");
  TRY(hipMemcpyAsync(g->seam_a.p, desc, dbytes, hipMemcpyHostToDevice, g->stream), "upload");
  TRY(hipMemcpyAsync(g->seam_b.p, tiles.data(), tbytes, hipMemcpyHostToDevice, g->stream), "upload");
  TRY(hipMemcpyAsync(g->seam_c.p, planes, sizeof(vp9r_ref_plane) * n_refs, hipMemcpyHostToDevice, g->stream),
      "upload");
  TRY(seam_begin(g), "hipEventRecord");
  vp9g::launch_mc_desc_tiles(g->stream, (const vp9r_mc_desc *)g->seam_a.p, g->seam_b.p, (uint32_t)tiles.size(),
                             (const vp9r_ref_plane *)g->seam_c.p, out, out_stride, bit_depth, rows);
  TRY(hipGetLastError(), "mc launch");
  TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
  TRY(hipStreamSynchronize(g->stream), "mc sync");
  TRY(seam_end(g), "hipEventElapsedTime");
  return VP9R_OK;
}

int vp9r_mc_batch_planes_noalloc(vp9r_gpu *g, uint32_t count, const vp9r_mc_desc *desc,
                                 const vp9r_ref_plane *planes, uint32_t n_refs, void *out, uint32_t out_stride,
                                 uint32_t bit_depth) {
  if (count > kNoallocMaxDescs || n_refs > kNoallocMaxRefs) {
    set_error("vp9r_mc_batch_planes_noalloc: at most 65536 descriptors and 64 reference planes");
    return VP9R_ERR_INVALID_INPUT;
  }
  if (!bd_ok(bit_depth) || n_refs == 0) {
    set_error("invalid motion compensation batch");
    return VP9R_ERR_INVALID_INPUT;
  }
  for (uint32_t k = 0; k < n_refs; k++)
    if (!planes[k].data || planes[k].last_x < 0 || planes[k].last_y < 0 ||
        (uint32_t)planes[k].last_x >= planes[k].stride) {
      set_error("invalid reference plane");
      return VP9R_ERR_INVALID_INPUT;
    }
  for (uint32_t i = 0; i < count; i++) {
    const vp9r_mc_desc &d = desc[i];
    bool ok = d.w >= 4 && d.w <= 64 && d.h >= 4 && d.h <= 64 && d.xs >= 1 && d.xs <= 32 && d.ys >= 1 &&
              d.ys <= 32 && d.filter < 4 && d.ref_idx >= 0 && (uint32_t)d.ref_idx < n_refs &&
              d.ref2_idx < (int)n_refs;
    if (!ok) {
      set_error("motion compensation descriptor out of range");
      return VP9R_ERR_INVALID_INPUT;
    }
  }
  TRY(hipSetDevice(g->device), "hipSetDevice");
  uint8_t *base = scratch_on(g->device);
  if (!base) return VP9R_ERR_DEVICE;
  // every byte the kernels read is written here first (nothing needs to
  // start at zero)
  if (count)
    TRY(hipMemcpyAsync(base, desc, sizeof(vp9r_mc_desc) * count, hipMemcpyHostToDevice, g->stream), "upload");
  TRY(hipMemcpyAsync(base + kNoallocDescBytes, planes, sizeof(vp9r_ref_plane) * n_refs, hipMemcpyHostToDevice,
                     g->stream),
      "upload");
  vp9g::launch_mc_desc_expand(g->stream, (const vp9r_mc_desc *)base, count,
                              (const vp9r_ref_plane *)(base + kNoallocDescBytes), out, out_stride, bit_depth);
  TRY(hipGetLastError(), "mc launch");
  TRY(hipStreamSynchronize(g->stream), "mc sync");
  return VP9R_OK;
}

int vp9r_mc_batch(vp9r_gpu *g, uint32_t count, const vp9r_mc_desc *desc, const void *const *refs,
                  uint32_t n_refs, uint32_t ref_stride, int32_t last_x, int32_t last_y, void *out,
                  uint32_t out_stride, uint32_t bit_depth) {
  if (n_refs == 0 || n_refs > 64) {
    set_error("invalid motion compensation batch");
    return VP9R_ERR_INVALID_INPUT;
  }
  std::vector<vp9r_ref_plane> planes(n_refs);
  for (uint32_t k = 0; k < n_refs; k++) planes[k] = {refs[k], ref_stride, last_x, last_y, 0};
  return vp9r_mc_batch_planes(g, count, desc, planes.data(), n_refs, out, out_stride, bit_depth);
}

int vp9r_avg_batch(vp9r_gpu *g, uint32_t count, const vp9r_avg_desc *desc, const void *a, const void *b,
                   void *dst, uint32_t stride, uint32_t bit_depth) {
  if (!bd_ok(bit_depth)) {
    set_error("bit depth must be 8, 10 or 12");
    return VP9R_ERR_INVALID_INPUT;
  }
  for (uint32_t i = 0; i < count; i++)
    if (desc[i].w < 1 || desc[i].w > 64 || desc[i].h < 1 || desc[i].h > 64 || desc[i].w > stride) {
      set_error("average block out of range");
      return VP9R_ERR_INVALID_INPUT;
    }
  TRY(hipSetDevice(g->device), "hipSetDevice");
  TRY(g->seam_a.ensure(sizeof(vp9r_avg_desc) * std::max<uint32_t>(count, 1)), "hipMalloc");
  if (count)
    TRY(hipMemcpyAsync(g->seam_a.p, desc, sizeof(vp9r_avg_desc) * count, hipMemcpyHostToDevice, g->stream),
        "upload");
  TRY(seam_begin(g), "hipEventRecord");
  vp9g::launch_avg_batch(g->stream, count, (const vp9r_avg_desc *)g->seam_a.p, a, b, dst, stride, bit_depth);
  TRY(hipGetLastError(), "avg launch");
  TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
  TRY(hipStreamSynchronize(g->stream), "avg sync");
  TRY(seam_end(g), "hipEventElapsedTime");
  return VP9R_OK;
}

int vp9r_add_const_batch(vp9r_gpu *g, uint32_t count, const vp9r_add_const_desc *desc, void *dst,
                         uint32_t stride) {
  for (uint32_t i = 0; i < count; i++)
    if (desc[i].n < 1 || desc[i].n > 64 || desc[i].n > stride) {
      set_error("block out of range");
      return VP9R_ERR_INVALID_INPUT;
    }
  TRY(hipSetDevice(g->device), "hipSetDevice");
  TRY(g->seam_a.ensure(sizeof(vp9r_add_const_desc) * std::max<uint32_t>(count, 1)), "hipMalloc");
  if (count)
    TRY(hipMemcpyAsync(g->seam_a.p, desc, sizeof(vp9r_add_const_desc) * count, hipMemcpyHostToDevice, g->stream),
        "upload");
  TRY(seam_begin(g), "hipEventRecord");
  vp9g::launch_add_const_batch(g->stream, count, (const vp9r_add_const_desc *)g->seam_a.p, dst, stride);
  TRY(hipGetLastError(), "add_const launch");
  TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
  TRY(hipStreamSynchronize(g->stream), "add_const sync");
  TRY(seam_end(g), "hipEventElapsedTime");
  return VP9R_OK;
}

int vp9r_intra_predict_batch(vp9r_gpu *g, uint32_t count, const vp9r_intra_desc *desc, const int32_t *edges,
                             void *dst, uint32_t stride, uint32_t bit_depth) {
  if (!bd_ok(bit_depth)) {
    set_error("bit depth must be 8, 10 or 12");
    return VP9R_ERR_INVALID_INPUT;
  }
  for (uint32_t i = 0; i < count; i++)
    if (desc[i].mode > 9 || desc[i].log2 < 2 || desc[i].log2 > 5 || (1u << desc[i].log2) > stride) {
      set_error("invalid intra prediction");
      return VP9R_ERR_INVALID_INPUT;
    }
  TRY(hipSetDevice(g->device), "hipSetDevice");
  TRY(g->seam_a.ensure(sizeof(vp9r_intra_desc) * std::max<uint32_t>(count, 1)), "hipMalloc");
  if (count)
    TRY(hipMemcpyAsync(g->seam_a.p, desc, sizeof(vp9r_intra_desc) * count, hipMemcpyHostToDevice, g->stream),
        "upload");
  TRY(seam_begin(g), "hipEventRecord");
  vp9g::launch_intra_pred_batch(g->stream, count, (const vp9r_intra_desc *)g->seam_a.p, edges, dst, stride,
                                bit_depth);
  TRY(hipGetLastError(), "intra launch");
  TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
  TRY(hipStreamSynchronize(g->stream), "intra sync");
  TRY(seam_end(g), "hipEventElapsedTime");
  return VP9R_OK;
}

int vp9r_lf_edges_batch(vp9r_gpu *g, uint32_t count, const uint8_t *edges, const uint16_t *x0,
                        const uint16_t *y0, const uint8_t *vertical, const uint8_t *first,
                        const uint8_t *n_edges, const uint8_t *n_runs, void *buf, uint32_t stride,
                        uint32_t rows, uint32_t bit_depth) {
  if (!bd_ok(bit_depth)) {
    set_error("bit depth must be 8, 10 or 12");
    return VP9R_ERR_INVALID_INPUT;
  }
  std::vector<vp9g::LfPassDesc> passes(count);
  for (uint32_t i = 0; i < count; i++) {
    uint32_t across = vertical[i] ? x0[i] : y0[i], along = vertical[i] ? y0[i] : x0[i];
    uint32_t lim_a = vertical[i] ? stride : rows, lim_r = vertical[i] ? rows : stride;
    const uint32_t f = first ? first[i] : 0;
    bool ok = n_edges[i] >= 1 && n_edges[i] <= 16 && n_runs[i] >= 1 && n_runs[i] <= 16 && f <= n_edges[i] &&
              across + 4u * n_edges[i] + 8 <= lim_a && along + 4u * n_runs[i] <= lim_r;
    if (!ok) {
      set_error("loop filter pass out of range of the plane");
      return VP9R_ERR_INVALID_INPUT;
    }
    passes[i] = {x0[i], y0[i], (uint8_t)(vertical[i] ? 1 : 0), (uint8_t)f, n_edges[i], n_runs[i]};
  }
  TRY(hipSetDevice(g->device), "hipSetDevice");
  TRY(g->seam_a.ensure(1024 * (size_t)std::max<uint32_t>(count, 1)), "hipMalloc");
  TRY(g->seam_b.ensure(sizeof(vp9g::LfPassDesc) * std::max<uint32_t>(count, 1)), "hipMalloc");
  if (count) {
    TRY(hipMemcpyAsync(g->seam_a.p, edges, 1024 * (size_t)count, hipMemcpyHostToDevice, g->stream), "upload");
    TRY(hipMemcpyAsync(g->seam_b.p, passes.data(), sizeof(vp9g::LfPassDesc) * count, hipMemcpyHostToDevice,
                       g->stream),
        "upload");
  }
  TRY(seam_begin(g), "hipEventRecord");
  vp9g::launch_lf_edges(g->stream, (const uint8_t *)g->seam_a.p, (const vp9g::LfPassDesc *)g->seam_b.p, count, buf,
                        stride, bit_depth);
  TRY(hipGetLastError(), "lf launch");
  TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
  TRY(hipStreamSynchronize(g->stream), "lf sync");
  TRY(seam_end(g), "hipEventElapsedTime");
  return VP9R_OK;
}

int vp9r_itx_add_host(vp9r_gpu *g, const int32_t *coefs, uint32_t n, uint32_t tx_type, int lossless,
                      uint32_t eob, uint32_t bit_depth, uint16_t *dst, uint32_t stride) {
  if (n < 2 || n > 5 || tx_type > 3 || !bd_ok(bit_depth) || (lossless && n != 2) || stride < (1u << n)) {
    set_error("invalid transform");
    return VP9R_ERR_INVALID_INPUT;
  }
  const uint32_t n0 = 1u << n;
  TRY(hipSetDevice(g->device), "hipSetDevice");
  // device: coefs, then the block as samples of the bit depth, then job data
  size_t bps = bit_depth > 8 ? 2 : 1;
  TRY(g->seam_a.ensure(4 * n0 * n0 + bps * n0 * n0 + 64), "hipMalloc");
  uint8_t *d = (uint8_t *)g->seam_a.p;
  int32_t *dc = (int32_t *)d;
  uint8_t *dpix = d + 4 * n0 * n0;
  struct Meta {
    uint32_t coef_off, job;
    uint8_t tx_type, pad;
    uint16_t eob;
    uint64_t dst_off;
  } meta = {0, 0, (uint8_t)tx_type, 0, (uint16_t)eob, 0};
  uint8_t *dmeta = dpix + bps * n0 * n0;
  dmeta = (uint8_t *)(((uintptr_t)dmeta + 15) & ~(uintptr_t)15);
  std::vector<uint8_t> pix(bps * n0 * n0);
  for (uint32_t i = 0; i < n0; i++)
    for (uint32_t j = 0; j < n0; j++) {
      uint16_t v = dst[i * stride + j];
      if (bps == 2)
        memcpy(&pix[2 * (i * n0 + j)], &v, 2);
      else
        pix[i * n0 + j] = (uint8_t)v;
    }
  TRY(hipMemcpyAsync(dc, coefs, 4 * n0 * n0, hipMemcpyHostToDevice, g->stream), "upload");
  TRY(hipMemcpyAsync(dpix, pix.data(), pix.size(), hipMemcpyHostToDevice, g->stream), "upload");
  TRY(hipMemcpyAsync(dmeta, &meta, sizeof meta, hipMemcpyHostToDevice, g->stream), "upload");
  Meta *m = (Meta *)dmeta;
  vp9g::launch_itx_dense(g->stream, 1, dc, &m->coef_off, nullptr, &m->tx_type, &m->eob, &m->dst_off, dpix, n0,
                         bit_depth, lossless, (int)n - 2, &m->job, 1);
  TRY(hipGetLastError(), "itx launch");
  TRY(hipMemcpyAsync(pix.data(), dpix, pix.size(), hipMemcpyDeviceToHost, g->stream), "download");
  TRY(hipStreamSynchronize(g->stream), "itx sync");
  for (uint32_t i = 0; i < n0; i++)
    for (uint32_t j = 0; j < n0; j++) {
      uint16_t v;
      if (bps == 2)
        memcpy(&v, &pix[2 * (i * n0 + j)], 2);
      else
        v = pix[i * n0 + j];
      dst[i * stride + j] = v;
    }
  return VP9R_OK;
}

int vp9r_intra_predict_host(vp9r_gpu *g, uint32_t mode, uint32_t log2, const int32_t *above,
                            const int32_t *left, int have_left, int have_above, uint32_t bit_depth,
                            uint16_t *dst, uint32_t stride) {
  if (mode > 9 || log2 < 2 || log2 > 5 || !bd_ok(bit_depth) || stride < (1u << log2)) {
    set_error("invalid intra prediction");
    return VP9R_ERR_INVALID_INPUT;
  }
  const uint32_t size = 1u << log2;
  TRY(hipSetDevice(g->device), "hipSetDevice");
  TRY(g->seam_a.ensure(4 * (2 * size + 1) + 4 * size + 2 * size * size + 64), "hipMalloc");
  int32_t *da = (int32_t *)g->seam_a.p, *dl = da + 2 * size + 1;
  uint16_t *dd = (uint16_t *)(dl + size);
  TRY(hipMemcpyAsync(da, above, 4 * (2 * size + 1), hipMemcpyHostToDevice, g->stream), "upload");
  TRY(hipMemcpyAsync(dl, left, 4 * size, hipMemcpyHostToDevice, g->stream), "upload");
  vp9g::launch_intra_pred(g->stream, mode, log2, da, dl, have_left, have_above, bit_depth, dd, size);
  TRY(hipGetLastError(), "intra launch");
  std::vector<uint16_t> out(size * size);
  TRY(hipMemcpyAsync(out.data(), dd, 2 * size * size, hipMemcpyDeviceToHost, g->stream), "download");
  TRY(hipStreamSynchronize(g->stream), "intra sync");
  for (uint32_t i = 0; i < size; i++)
    for (uint32_t j = 0; j < size; j++) dst[i * stride + j] = out[i * size + j];
  return VP9R_OK;
}

// ---- the autotuner of the per-kernel seam's launch shapes
//
// For each class it times every candidate shape on a synthetic batch of that
// class (best of three, device time between events around the launch) and
// keeps the fastest; the arithmetic is the same for every shape.
//   transforms, per size (4x4 .. 32x32) and bit depth class, on 262144
//   (4x4) .. 32768 (32x32) blocks of dense random coefficients: 1, 2
//   or 4 waves per workgroup, the transposes through LDS or in registers
//   (4x4 .. 16x16), one workgroup per 64 W / n blocks or resident
//   (persistent) workgroups;
//   motion compensation: tiles of 16 or 32 rows per bit depth class, on 8192
//   blocks of 16x16 .. 64x64 from a random 1920x1080 reference.
// There is no counterpart in the reference (src/dsp/mod.rs:32-96 picks its
// kernels by CPU feature, and src/dsp/bench.rs times them by hand).

static bool itx_shape_ok(uint32_t v, int size) {
  const uint32_t w = v & 15, flags = v & ~15u;
  if (w != 1 && w != 2 && w != 4) return false;
  if (flags & ~(uint32_t)(vp9g::kItxShapeReg | vp9g::kItxShapePersistent)) return false;
  return !((flags & vp9g::kItxShapeReg) && size > 2);
}

// candidate c of the 12 per size: (persistent, registers, waves 1/2/4)
static int itx_candidate(int c) {
  static const int w[3] = {1, 2, 4};
  return w[c % 3] | ((c / 3) % 2 ? vp9g::kItxShapeReg : 0) | (c / 6 ? vp9g::kItxShapePersistent : 0);
}

int vp9r_gpu_set_tuning(vp9r_gpu *g, const uint32_t in[10]) {
  for (int i = 0; i < 8; i++)
    if (!itx_shape_ok(in[i], i % 4)) {
      set_error("transform shape: 1, 2 or 4 waves per workgroup, plus 16 (register transposes, 4x4 .. 16x16) "
                "and 32 (persistent)");
      return VP9R_ERR_INVALID_INPUT;
    }
  for (int i = 8; i < 10; i++)
    if (in[i] != 16 && in[i] != 32) {
      set_error("motion compensation tile rows: 16 or 32");
      return VP9R_ERR_INVALID_INPUT;
    }
  for (int c = 0; c < 2; c++) {
    for (int k = 0; k < 4; k++) g->tune_itx_shape[c][k] = (uint8_t)in[4 * c + k];
    g->tune_mc_rows[c] = (uint8_t)in[8 + c];
  }
  return VP9R_OK;
}

int vp9r_gpu_tuning(const vp9r_gpu *g, uint32_t out[11]) {
  for (int c = 0; c < 2; c++) {
    for (int k = 0; k < 4; k++) out[4 * c + k] = g->tune_itx_shape[c][k];
    out[8 + c] = g->tune_mc_rows[c];
  }
  out[10] = g->tuned ? 1 : 0;
  return VP9R_OK;
}

int vp9r_gpu_autotune(vp9r_gpu *g, float times_us[VP9R_TUNE_TIMES]) {
  TRY(hipSetDevice(g->device), "hipSetDevice");
  uint64_t rng = 0x9e3779b97f4a7c15ull;
  auto next = [&]() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };
  float tmp[VP9R_TUNE_TIMES];
  float *tu = times_us ? times_us : tmp;
  for (int i = 0; i < VP9R_TUNE_TIMES; i++) tu[i] = -1.0f;
  // transforms
  for (int c = 0; c < 2; c++) {
    const uint32_t bd = c ? 10 : 8, bps = c ? 2 : 1;
    for (int k = 0; k < 4; k++) {
      const uint32_t nb = 262144u >> k;  // 16 MB (4x4) .. 128 MB (32x32) of coefficients
      const uint32_t n0 = 4u << k, stride = 1024;
      const uint32_t per_row = stride / n0, rows_blk = (nb + per_row - 1) / per_row;
      std::vector<int32_t> coefs((size_t)nb * n0 * n0);
      for (auto &v : coefs) v = (int32_t)(next() % 97) - 48;
      std::vector<uint32_t> coff(nb), jobs(nb);
      std::vector<uint8_t> ttype(nb);
      std::vector<uint16_t> eob(nb);
      std::vector<uint64_t> doff(nb);
      for (uint32_t i = 0; i < nb; i++) {
        coff[i] = i * n0 * n0;
        ttype[i] = k == 3 ? 0 : (uint8_t)(next() & 3);
        eob[i] = (uint16_t)(n0 * n0);
        doff[i] = (uint64_t)(i / per_row) * n0 * stride + (i % per_row) * n0;
        jobs[i] = i;
      }
      void *d_coef = nullptr, *d_meta = nullptr, *d_dst = nullptr;
      const size_t meta = 4 * nb + nb + 2 * nb + 8 * nb + 4 * nb + 64;
      TRY(hipMalloc(&d_coef, 4 * coefs.size()), "hipMalloc(tune)");
      TRY(hipMalloc(&d_meta, meta), "hipMalloc(tune)");
      TRY(hipMalloc(&d_dst, (size_t)rows_blk * n0 * stride * bps), "hipMalloc(tune)");
      uint8_t *m = (uint8_t *)d_meta;
      uint32_t *d_coff = (uint32_t *)m;
      uint64_t *d_doff = (uint64_t *)(((uintptr_t)(m + 4 * nb) + 7) & ~(uintptr_t)7);
      uint32_t *d_jobs = (uint32_t *)(d_doff + nb);
      uint16_t *d_eob = (uint16_t *)(d_jobs + nb);
      uint8_t *d_tt = (uint8_t *)(d_eob + nb);
      TRY(hipMemcpy(d_coef, coefs.data(), 4 * coefs.size(), hipMemcpyHostToDevice), "tune upload");
      TRY(hipMemcpy(d_coff, coff.data(), 4 * nb, hipMemcpyHostToDevice), "tune upload");
      TRY(hipMemcpy(d_doff, doff.data(), 8 * nb, hipMemcpyHostToDevice), "tune upload");
      TRY(hipMemcpy(d_jobs, jobs.data(), 4 * nb, hipMemcpyHostToDevice), "tune upload");
      TRY(hipMemcpy(d_eob, eob.data(), 2 * nb, hipMemcpyHostToDevice), "tune upload");
      TRY(hipMemcpy(d_tt, ttype.data(), nb, hipMemcpyHostToDevice), "tune upload");
      TRY(hipMemset(d_dst, 0x40, (size_t)rows_blk * n0 * stride * bps), "tune memset");
      float best = 1e30f;
      int best_shape = 1;
      for (int cand = 0; cand < 12; cand++) {
        const int shape = itx_candidate(cand);
        if (!itx_shape_ok((uint32_t)shape, k)) continue;
        float t_best = 1e30f;
        for (int rep = 0; rep < 4; rep++) {
          TRY(seam_begin(g), "hipEventRecord");
          vp9g::launch_itx_dense(g->stream, nb, (const int32_t *)d_coef, d_coff, nullptr, d_tt, d_eob, d_doff, d_dst,
                                 stride, bd, 0, k, d_jobs, nb, shape);
          TRY(hipGetLastError(), "tune itx launch");
          TRY(hipEventRecord(g->seam_ev[1], g->stream), "hipEventRecord");
          TRY(hipStreamSynchronize(g->stream), "tune sync");
          TRY(seam_end(g), "hipEventElapsedTime");
          if (rep) t_best = std::min(t_best, g->seam_ms);  // the first is a warm-up
        }
        tu[(c * 4 + k) * 12 + cand] = 1000.0f * t_best;
        if (t_best < best * 0.97f) {  // a candidate must win clearly to move off the default
          best = t_best;
          best_shape = shape;
        }
      }
      g->tune_itx_shape[c][k] = (uint8_t)best_shape;
      (void)hipFree(d_coef);
      (void)hipFree(d_meta);
      (void)hipFree(d_dst);
    }
  }
  // motion compensation
  const uint32_t rw = 1920, rh = 1080, nd = 8192, ostride = 4096;
  for (int c = 0; c < 2; c++) {
    const uint32_t bd = c ? 10 : 8, bps = c ? 2 : 1;
    std::vector<uint8_t> ref((size_t)rw * rh * bps);
    for (auto &v : ref) v = (uint8_t)next();
    if (c)
      for (size_t i = 1; i < ref.size(); i += 2) ref[i] &= 3;  // 10-bit samples
    std::vector<vp9r_mc_desc> desc(nd);
    const int sizes[4] = {16, 32, 64, 32};
    for (uint32_t i = 0; i < nd; i++) {
      vp9r_mc_desc &d = desc[i];
      d = vp9r_mc_desc{};
      d.w = (uint8_t)sizes[i & 3];
      d.h = (uint8_t)sizes[(i >> 2) & 3];
      d.x = (int32_t)(next() % ((rw - 80) * 16)) + 128;
      d.y = (int32_t)(next() % ((rh - 80) * 16)) + 128;
      d.x2 = d.x + 40;
      d.y2 = d.y - 24;
      d.xs = d.ys = 16;
      d.filter = (uint8_t)(next() % 3);
      d.ref_idx = 0;
      d.ref2_idx = (i % 10) < 3 ? 0 : -1;
      d.out_off = (uint64_t)(i / 64) * 64 * ostride + (i % 64) * 64;
    }
    void *d_ref = nullptr, *d_out = nullptr;
    const size_t out_bytes = (size_t)(nd / 64 + 1) * 64 * ostride * bps;
    TRY(hipMalloc(&d_ref, ref.size()), "hipMalloc(tune)");
    TRY(hipMalloc(&d_out, out_bytes), "hipMalloc(tune)");
    TRY(hipMemcpy(d_ref, ref.data(), ref.size(), hipMemcpyHostToDevice), "tune upload");
    vp9r_ref_plane plane{d_ref, rw, (int32_t)rw - 1, (int32_t)rh - 1, 0};
    float best = 1e30f;
    int best_rows = 32;
    const int cand[2] = {32, 16};
    for (int r : cand) {
      g->tune_mc_rows[c] = (uint8_t)r;
      float t_best = 1e30f;
      for (int rep = 0; rep < 4; rep++) {
        int st = vp9r_mc_batch_planes(g, nd, desc.data(), &plane, 1, d_out, ostride, bd);
        if (st) {
          (void)hipFree(d_ref);
          (void)hipFree(d_out);
          return st;
        }
        if (rep) t_best = std::min(t_best, g->seam_ms);
      }
      tu[96 + 2 * c + (r == 32 ? 0 : 1)] = 1000.0f * t_best;
      if (t_best < best * 0.97f) {
        best = t_best;
        best_rows = r;
      }
    }
    g->tune_mc_rows[c] = (uint8_t)best_rows;
    (void)hipFree(d_ref);
    (void)hipFree(d_out);
  }
  g->tuned = true;
  g->seam_ms = -1.0f;
  return VP9R_OK;
}

}  // extern "C"
