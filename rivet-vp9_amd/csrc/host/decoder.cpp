// The decoder API of include/vp9r.h: the crate's Decoder
// (src/decoder/mod.rs:122-216) with the CPU parser and the GPU
// reconstruction.  Reference slots hold HBM-resident frame stores shared
// between slots (the crate's Arc<RefFrame>, decoder/mod.rs:125, 398-415);
// frame stores are recycled through a pool once no slot holds them (the
// crate's decoder/pool.rs, here device memory).  Shown frames are copied to
// the host in the Frame::packed layout (decoder/mod.rs:79-106).
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "parser.h"
#include "recon.h"

using namespace vp9r;

namespace {

struct FrameStore {
  vp9r_gpu_frame *f = nullptr;
  uint32_t color_space = 0, color_range = 0, render_w = 0, render_h = 0;
  ~FrameStore() { vp9r_gpu_frame_free(f); }
};

// A shown frame of the current packet: its description and a pinned host
// buffer (reused from packet to packet) that its download lands in.
struct OutFrame {
  vp9r_frame desc{};
  vp9r::HostBuf *buf = nullptr;
  vp9r_gpu_frame *f = nullptr;  // the store being downloaded
};

}  // namespace

struct vp9r_decoder {
  vp9r_gpu *g = nullptr;
  vp9r_parser parser;
  std::shared_ptr<FrameStore> slots[8];
  std::vector<std::shared_ptr<FrameStore>> pool;
  std::vector<OutFrame> out;
  std::vector<std::unique_ptr<vp9r::HostBuf>> out_bufs;  // pinned, one per shown frame of a packet
  uint32_t threads = 0;
  uint64_t frames = 0;

  std::shared_ptr<FrameStore> take(uint32_t w, uint32_t h, uint32_t bd, uint32_t sx, uint32_t sy, int *st) {
    for (size_t i = 0; i < pool.size(); i++) {
      vp9r_gpu_frame *f = pool[i]->f;
      if (pool[i].use_count() == 1 && f->width == w && f->height == h && f->bit_depth == bd && f->ss_x == sx &&
          f->ss_y == sy) {
        bool in_slot = false;
        for (auto &s : slots) in_slot |= s == pool[i];
        if (!in_slot) return pool[i];
      }
    }
    auto fs = std::make_shared<FrameStore>();
    *st = vp9r_gpu_frame_new(g, w, h, bd, sx, sy, &fs->f);
    if (*st) return nullptr;
    pool.push_back(fs);
    // forget stores of other geometries nobody holds
    std::vector<std::shared_ptr<FrameStore>> keep;
    for (auto &p : pool) {
      bool held = p.use_count() > 1;
      for (auto &s : slots) held |= s == p;
      if (held || (p->f->width == w && p->f->height == h)) keep.push_back(p);
    }
    pool.swap(keep);
    return fs;
  }

  // Queues the packed download of a shown frame into a pinned buffer on
  // the copy stream (behind the frame's reconstruction); vp9r_decoder_decode
  // waits for all of a packet's downloads before it returns.
  int emit(const std::shared_ptr<FrameStore> &fs) {
    vp9r_gpu_frame *f = fs->f;
    OutFrame o;
    size_t n = vp9r_packed_size(f->width, f->height, f->bit_depth, f->ss_x, f->ss_y);
    if (out_bufs.size() <= out.size()) out_bufs.emplace_back(new vp9r::HostBuf());
    o.buf = out_bufs[out.size()].get();
    if (o.buf->ensure(n) != hipSuccess) {
      set_error("hipHostMalloc(frame output) failed");
      return VP9R_ERR_DEVICE;
    }
    int st = vp9r_gpu_download_packed_async(g, f, (uint8_t *)o.buf->p);
    if (st) return st;
    o.f = f;
    vp9r_frame &d = o.desc;
    d.width = f->width;
    d.height = f->height;
    d.bit_depth = f->bit_depth;
    d.ss_x = f->ss_x;
    d.ss_y = f->ss_y;
    d.color_space = fs->color_space;
    d.full_range = fs->color_range;
    d.render_width = fs->render_w;
    d.render_height = fs->render_h;
    d.size = n;
    size_t bps = f->bit_depth > 8 ? 2 : 1, off = 0;
    for (int p = 0; p < 3; p++) {
      d.plane_width[p] = p ? (f->width + f->ss_x) >> f->ss_x : f->width;
      d.plane_height[p] = p ? (f->height + f->ss_y) >> f->ss_y : f->height;
      d.plane_offset[p] = off;
      off += (size_t)d.plane_width[p] * d.plane_height[p] * bps;
    }
    out.push_back(o);
    return VP9R_OK;
  }

  // Waits for every download queued by emit() (each in its own buffer).
  int finish_out() {
    int st = VP9R_OK;
    for (OutFrame &o : out) {
      int s2 = vp9r_gpu_frame_wait(o.f);
      if (s2 && !st) st = s2;
    }
    return st;
  }

  int decode_frame(const uint8_t *data, size_t len) {
    vp9r_record rec;
    int st = vp9r_parser_parse_frame(&parser, data, len, &rec);
    if (st) return st;
    const vp9r_frame_hdr &h = rec.hdr;
    frames++;
    if (h.show_existing_frame) {
      // the parser refuses slots it never filled; this one may also have
      // been emptied by a failed reconstruction (forget_slots below)
      if (!slots[h.frame_to_show_map_idx]) {
        set_error("show_existing_frame names a slot that holds no frame");
        return VP9R_ERR_BITSTREAM;
      }
      return emit(slots[h.frame_to_show_map_idx]);
    }
    st = reconstruct(rec);
    if (st) {
      // The parser has already refreshed its slots with this frame.  Empty
      // them on both sides, so that no later frame predicts from or shows a
      // store that does not hold this frame.
      for (int i = 0; i < 8; i++)
        if ((h.refresh_frame_flags >> i) & 1) slots[i].reset();
      parser.p.forget_slots(h.refresh_frame_flags);
      return st;
    }
    return h.show_frame ? emit(last) : VP9R_OK;
  }

  // Reconstruction and loop filter of a parsed frame into a store from the
  // pool, then the slot refresh (decoder/mod.rs:398-415); `last` gets the
  // store.
  std::shared_ptr<FrameStore> last;
  int reconstruct(const vp9r_record &rec) {
    const vp9r_frame_hdr &h = rec.hdr;
    int st = VP9R_OK;
    vp9r_gpu_frame *refs[3] = {nullptr, nullptr, nullptr};
    if (!h.frame_is_intra)
      for (int i = 0; i < 3; i++) {
        if (!slots[h.ref_frame_idx[i]]) {
          set_error("active reference slot holds no frame");
          return VP9R_ERR_BITSTREAM;
        }
        refs[i] = slots[h.ref_frame_idx[i]]->f;
      }
    auto cur = test_fail_stores > 0 ? nullptr : take(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y, &st);
    if (!cur) {
      if (test_fail_stores > 0) {
        test_fail_stores--;
        set_error("frame store allocation refused (test hook)");
        return VP9R_ERR_DEVICE;
      }
      return st;
    }
    cur->color_space = h.color_space;
    cur->color_range = h.color_range;
    cur->render_w = h.render_width;
    cur->render_h = h.render_height;
    st = vp9r_gpu_reconstruct(g, &rec, refs, cur->f);
    if (st) return st;
    st = vp9r_gpu_loop_filter(g, &rec, cur->f);
    if (st) return st;
    for (int i = 0; i < 8; i++)
      if ((h.refresh_frame_flags >> i) & 1) slots[i] = cur;
    last = cur;
    return VP9R_OK;
  }
  // vp9r_decoder_test_fail_stores: frame store requests still to refuse
  uint32_t test_fail_stores = 0;
};

extern "C" {

int vp9r_decoder_new(int device, vp9r_decoder **out) {
  *out = nullptr;
  vp9r_decoder *d = new (std::nothrow) vp9r_decoder();
  if (!d) return VP9R_ERR_INVALID_INPUT;
  int st = vp9r_gpu_new(device, &d->g);
  if (st) {
    delete d;
    return st;
  }
  vp9r_decoder_set_threads(d, 0);
  *out = d;
  return VP9R_OK;
}

void vp9r_decoder_free(vp9r_decoder *d) {
  if (!d) return;
  if (d->g) {
    (void)hipStreamSynchronize(d->g->stream);
    (void)hipStreamSynchronize(d->g->copy_stream);
  }
  for (auto &s : d->slots) s.reset();
  d->pool.clear();
  vp9r_gpu_free(d->g);
  delete d;
}

uint32_t vp9r_decoder_threads(const vp9r_decoder *d) {
  if (d->threads) return d->threads;
  unsigned n = std::thread::hardware_concurrency();
  return n ? n : 1;
}

void vp9r_decoder_set_threads(vp9r_decoder *d, uint32_t threads) {
  d->threads = threads;
  d->parser.p.tile_threads = vp9r_decoder_threads(d);
}

void vp9r_decoder_set_max_pixels(vp9r_decoder *d, uint64_t pixels) { d->parser.p.max_pixels = pixels; }

int vp9r_decoder_decode(vp9r_decoder *d, const uint8_t *packet, size_t len, uint32_t *n_frames) {
  d->out.clear();
  *n_frames = 0;
  size_t offs[8], sizes[8];
  size_t n = vp9r_superframe_split(packet, len, offs, sizes, 8);
  for (size_t i = 0; i < n && i < 8; i++) {
    int st = d->decode_frame(packet + offs[i], sizes[i]);
    if (st) {
      (void)d->finish_out();
      d->out.clear();
      return st;
    }
  }
  int st = d->finish_out();
  if (st) {
    d->out.clear();
    return st;
  }
  *n_frames = (uint32_t)d->out.size();
  return VP9R_OK;
}

int vp9r_decoder_frame(const vp9r_decoder *d, uint32_t i, vp9r_frame *out) {
  if (i >= d->out.size()) {
    set_error("no such frame");
    return VP9R_ERR_INVALID_INPUT;
  }
  *out = d->out[i].desc;
  out->data = (uint8_t *)d->out[i].buf->p;
  return VP9R_OK;
}

uint64_t vp9r_decoder_frames_decoded(const vp9r_decoder *d) { return d->frames; }

void vp9r_decoder_test_fail_stores(vp9r_decoder *d, uint32_t n) { d->test_fail_stores = n; }

}  // extern "C"
