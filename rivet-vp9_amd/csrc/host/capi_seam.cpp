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
                           bit_depth, lossless, k, dj + start[k], (uint32_t)jobs[k].size());
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
    // tiles of 16 columns by up to 32 rows (the kernel's 8-bit unscaled
    // path filters the 32 rows in one pass; other paths take two halves)
    for (int ty = 0; ty < d.h; ty += 32)
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
                             (const vp9r_ref_plane *)g->seam_c.p, out, out_stride, bit_depth);
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

}  // extern "C"
