// The multi-stream server of include/vp9r.h: many independent VP9 streams
// decoded on one GPU with continuous batching.
//
//   submit -> [parse workers] -> per-stream ready queue -> [dispatcher] ->
//   one batch per launch sequence (one frame of each ready stream) ->
//   device-to-host copies of the shown frames -> [finisher] -> poll
//
// Threads and what they own:
//   * parse workers: a stream's parser (one worker per stream at a time),
//     the worker's work-list builder and upload stream.  Each frame's inputs
//     go to an input slot of its stream (pinned staging + HBM copy), uploaded
//     on the worker's stream; the batch waits for that upload on the device.
//   * the dispatcher: the reconstruction context (its compute and copy
//     streams), every stream's reference slots and frame stores, the
//     frame-store cache and the output buffers' allocation.
//   * the finisher: waits for each batch in order, releases its input slots,
//     stamps the latency of its frames and ends streams that are done.
// Everything shared is under one mutex; device work happens outside it.
//
// A stream's reference slots and frame stores are those of the crate's
// decoder (src/decoder/mod.rs:125, 398-415): a store is reused once no slot
// holds it; successive frames of a stream are in successive batches, which
// the one compute stream orders, so a frame always sees its references
// complete.  Frames in one batch belong to different streams and touch
// disjoint stores.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "parser.h"
#include "recon.h"

using namespace vp9r;

namespace {

int hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return VP9R_ERR_DEVICE;
}

#define HIP_TRY(expr, what)                          \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int kStoresPerStream = 10;  // 8 reference slots, the current frame, one in flight

struct Stream;

// One frame's inputs: pinned staging, its HBM copy and the upload's event.
struct InputSlot {
  HostBuf host;
  DevBuf dev;
  hipEvent_t uploaded = nullptr;
  bool busy = false;
  ~InputSlot() {
    if (uploaded) (void)hipEventDestroy(uploaded);
  }
};

enum JobKind { kDecode, kShowExisting, kFail };

// One frame of a stream, in decode order.
struct Job {
  Stream *st = nullptr;
  double t_submit = 0;
  JobKind kind = kDecode;
  int status = VP9R_OK;
  std::string error;
  uint32_t show_idx = 0;
  Prepared p;
  InputSlot *in = nullptr;
};

struct Store {
  vp9r_gpu_frame *f = nullptr;
  int holds = 0;  // reference slots holding it
  uint32_t color_space = 0, color_range = 0, render_w = 0, render_h = 0;
  bool fits(uint32_t w, uint32_t h, uint32_t bd, uint32_t sx, uint32_t sy) const {
    return f->width == w && f->height == h && f->bit_depth == bd && f->ss_x == sx && f->ss_y == sy;
  }
};

struct OutBuf {
  HostBuf buf;
};

struct OutRec {
  vp9r_server_frame fr{};
  OutBuf *buf = nullptr;
  double t_submit = 0;
};

struct Stream {
  uint32_t id = 0;
  uint32_t max_w = 0, max_h = 0, bd = 8, ss_x = 1, ss_y = 1;
  uint64_t footprint = 0;
  bool admitted = false, closed = false, failed = false, finished = false;
  int status = VP9R_OK;
  std::string error;
  vp9r_parser parser;
  std::deque<std::pair<std::vector<uint8_t>, double>> packets;
  bool parsing = false;
  std::deque<Job *> ready;  // parsed, waiting for the device
  uint32_t outstanding = 0;  // parsed and not yet finished (ready + dispatched)
  std::vector<std::unique_ptr<InputSlot>> inputs;
  // dispatcher only
  Store *slots[8] = {};
  std::vector<Store *> stores;
  uint32_t shown = 0;
};

struct Batch {
  std::vector<Job *> jobs;
  std::vector<OutRec> outs;
  uint32_t n_frames = 0;  // frames reconstructed
  hipEvent_t done = nullptr;
};

}  // namespace

struct vp9r_server {
  int device = 0;
  vp9r_server_config cfg{};
  vp9r_gpu *g = nullptr;
  hipEvent_t ev_batch = nullptr;  // end of a batch's kernels, for the copy stream
  std::mutex m;
  std::condition_variable cv_work, cv_disp, cv_fin, cv_out;
  std::vector<std::unique_ptr<Stream>> streams;
  std::deque<Stream *> waiting;
  uint64_t reserved = 0;       // footprints of admitted, unfinished streams
  uint64_t store_bytes = 0;    // frame stores allocated (streams + cache)
  std::vector<Store *> cache;  // stores of finished streams
  std::vector<std::unique_ptr<InputSlot>> spare_inputs;  // input slots of finished streams
  std::deque<Batch *> inflight;
  std::vector<hipEvent_t> free_events;
  std::deque<OutRec> done;
  std::unordered_map<uint64_t, OutBuf *> handed_out;
  std::vector<OutBuf *> free_out;
  uint64_t next_handle = 1;
  bool stop = false;
  int fatal = VP9R_OK;
  std::string fatal_msg;
  uint64_t n_shown = 0, n_decoded = 0, n_batches = 0, max_batch_seen = 0, n_admitted = 0;
  std::vector<std::thread> workers;
  std::vector<hipStream_t> upload_streams;
  std::thread dispatcher, finisher;

  // ---- admission (under m)
  void admit_waiting() {
    while (!waiting.empty() && reserved + waiting.front()->footprint <= cfg.hbm_budget) {
      Stream *st = waiting.front();
      waiting.pop_front();
      st->admitted = true;
      reserved += st->footprint;
      n_admitted++;
    }
    cv_work.notify_all();
  }

  bool stream_done(const Stream *st) const {
    return (st->closed || st->failed) && st->packets.empty() && !st->parsing && st->ready.empty() &&
           st->outstanding == 0;
  }

  // Ends a stream whose work is all finished: an end record, its stores to
  // the cache, its footprint released (under m; the dispatcher no longer
  // touches a stream with no work).
  void finish_stream(Stream *st) {
    if (st->finished || !stream_done(st)) return;
    if (!st->admitted) {
      // closed before admission: nothing was parsed
      auto it = std::find(waiting.begin(), waiting.end(), st);
      if (it != waiting.end()) waiting.erase(it);
    }
    st->finished = true;
    OutRec o;
    o.fr.stream = st->id;
    o.fr.index = st->shown;
    o.fr.eos = 1;
    o.fr.status = st->status;
    done.push_back(o);
    for (auto &s : st->slots) s = nullptr;
    for (Store *s : st->stores) {
      s->holds = 0;
      cache.push_back(s);
    }
    st->stores.clear();
    // input slots are kept for other streams (freeing device memory would
    // wait for the device)
    for (auto &in : st->inputs) spare_inputs.push_back(std::move(in));
    st->inputs.clear();
    if (st->admitted) reserved -= st->footprint;
    admit_waiting();
    cv_out.notify_all();
  }

  void fail_stream(Stream *st, int status, const std::string &msg) {
    if (st->failed) return;
    st->failed = true;
    st->status = status ? status : VP9R_ERR_BITSTREAM;
    st->error = msg;
    st->packets.clear();
  }

  // ---- parse workers
  Stream *pick_for_parse() {
    // after a device failure nothing more is parsed (the input slots of
    // the failed batch may still be read by its queued work)
    if (fatal != VP9R_OK) return nullptr;
    Stream *best = nullptr;
    for (auto &up : streams) {
      Stream *st = up.get();
      if (!st->admitted || st->finished || st->failed || st->parsing || st->packets.empty()) continue;
      if (st->ready.size() >= cfg.parse_ahead) continue;
      if (!best || st->packets.front().second < best->packets.front().second) best = st;
    }
    return best;
  }

  InputSlot *take_input(Stream *st) {
    for (auto &in : st->inputs)
      if (!in->busy) {
        in->busy = true;
        return in.get();
      }
    std::unique_ptr<InputSlot> in;
    if (!spare_inputs.empty()) {
      in = std::move(spare_inputs.back());
      spare_inputs.pop_back();
    } else {
      in = std::make_unique<InputSlot>();
      if (hipEventCreateWithFlags(&in->uploaded, hipEventDisableTiming) != hipSuccess) return nullptr;
    }
    in->busy = true;
    st->inputs.push_back(std::move(in));
    return st->inputs.back().get();
  }

  void worker_main(int wi) {
    (void)hipSetDevice(device);
    hipStream_t us = upload_streams[wi];
    Builder work;
    std::unique_lock<std::mutex> lk(m);
    while (!stop) {
      Stream *st = pick_for_parse();
      if (!st) {
        cv_work.wait(lk);
        continue;
      }
      auto pkt = std::move(st->packets.front());
      st->packets.pop_front();
      st->parsing = true;
      // Streams with packets share the parse threads: with at most a
      // quarter as many such streams as workers, a stream's tile columns
      // (up to 4) are parsed in parallel by spare threads (the record is
      // the same either way; two settings only, so that a stream's thread
      // pool is rarely rebuilt).
      size_t busy_streams = 0;
      for (auto &up : streams)
        busy_streams += up->admitted && !up->finished && !up->failed && (!up->packets.empty() || up->parsing);
      st->parser.p.tile_threads = 4 * busy_streams <= cfg.parse_threads ? 4 : 1;
      lk.unlock();
      size_t offs[8], sizes[8];
      const std::vector<uint8_t> &data = pkt.first;
      size_t nf = vp9r_superframe_split(data.data(), data.size(), offs, sizes, 8);
      std::vector<Job *> jobs;
      bool failed = false;
      for (size_t i = 0; i < nf && i < 8 && !failed; i++) {
        Job *j = new Job();
        j->st = st;
        j->t_submit = pkt.second;
        vp9r_record rec;
        int s = vp9r_parser_parse_frame(&st->parser, data.data() + offs[i], sizes[i], &rec);
        const vp9r_frame_hdr &h = rec.hdr;
        if (s == VP9R_OK && h.show_existing_frame) {
          j->kind = kShowExisting;
          j->show_idx = h.frame_to_show_map_idx;
        } else if (s == VP9R_OK) {
          if (h.width > st->max_w || h.height > st->max_h || h.bit_depth != st->bd || h.ss_x != st->ss_x ||
              h.ss_y != st->ss_y) {
            s = VP9R_ERR_INVALID_INPUT;
            set_error("a frame outside the stream's declared size or format");
          } else {
            lk.lock();
            InputSlot *in = take_input(st);
            lk.unlock();
            if (!in) {
              s = VP9R_ERR_DEVICE;
              set_error("hipEventCreate failed");
            } else {
              j->in = in;
              s = build_record(work, &rec, j->p, in->host);
              if (s == VP9R_OK) {
                hipError_t e = in->dev.ensure(j->p.bytes);
                if (e == hipSuccess)
                  e = hipMemcpyAsync(in->dev.p, in->host.p, j->p.bytes, hipMemcpyHostToDevice, us);
                if (e == hipSuccess) e = hipEventRecord(in->uploaded, us);
                if (e != hipSuccess) {
                  s = VP9R_ERR_DEVICE;
                  set_error(std::string("input upload: ") + hipGetErrorString(e));
                }
              }
            }
          }
        }
        if (s != VP9R_OK) {
          j->kind = kFail;
          j->status = s;
          j->error = vp9r_last_error();
          failed = true;
        }
        jobs.push_back(j);
      }
      lk.lock();
      for (Job *j : jobs) {
        st->ready.push_back(j);
        st->outstanding++;
      }
      if (failed) st->packets.clear();
      st->parsing = false;
      cv_disp.notify_one();
      cv_work.notify_all();
    }
  }

  // ---- dispatcher (frame stores: dispatcher thread only, cache under m)
  Store *take_store(Stream *st, const vp9r_frame_hdr &h, int *status) {
    for (Store *s : st->stores)
      if (s->holds == 0 && s->fits(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)) return s;
    Store *got = nullptr;
    {
      std::lock_guard<std::mutex> lk(m);
      // stores of other geometries that no slot holds go back to the cache
      std::vector<Store *> keep;
      for (Store *s : st->stores) {
        if (s->holds == 0 && !s->fits(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y))
          cache.push_back(s);
        else
          keep.push_back(s);
      }
      st->stores.swap(keep);
      for (size_t i = 0; i < cache.size(); i++)
        if (cache[i]->fits(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y)) {
          got = cache[i];
          cache.erase(cache.begin() + i);
          break;
        }
    }
    if (!got) {
      if ((int)st->stores.size() >= kStoresPerStream) {
        *status = VP9R_ERR_INVALID_INPUT;
        set_error("stream needs more frame stores than its footprint");
        return nullptr;
      }
      const uint64_t need = vp9r_frame_store_bytes(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y);
      // make room: drop cached stores (freeing waits for the device)
      for (;;) {
        Store *victim = nullptr;
        {
          std::lock_guard<std::mutex> lk(m);
          if (store_bytes + need <= cfg.hbm_budget || cache.empty()) break;
          victim = cache.back();
          cache.pop_back();
          store_bytes -= victim->f->bytes;
        }
        vp9r_gpu_frame_free(victim->f);
        delete victim;
      }
      got = new Store();
      int s = vp9r_gpu_frame_new(g, h.width, h.height, h.bit_depth, h.ss_x, h.ss_y, &got->f);
      if (s) {
        delete got;
        *status = s;
        return nullptr;
      }
      std::lock_guard<std::mutex> lk(m);
      store_bytes += got->f->bytes;
    }
    got->holds = 0;
    st->stores.push_back(got);
    return got;
  }

  OutBuf *take_out(size_t n) {
    {
      std::lock_guard<std::mutex> lk(m);
      for (size_t i = 0; i < free_out.size(); i++)
        if (free_out[i]->buf.cap >= n) {
          OutBuf *o = free_out[i];
          free_out.erase(free_out.begin() + i);
          return o;
        }
    }
    OutBuf *o = new OutBuf();
    if (o->buf.ensure(n) != hipSuccess) {
      delete o;
      return nullptr;
    }
    return o;
  }

  static void set_slot(Stream *st, int k, Store *s) {
    if (st->slots[k]) st->slots[k]->holds--;
    st->slots[k] = s;
    if (s) s->holds++;
  }

  // Adds the output copy of store s (shown frame `index` of st) to the batch.
  bool add_output(Batch *b, Job *j, Store *s) {
    vp9r_gpu_frame *f = s->f;
    const size_t n = vp9r_packed_size(f->width, f->height, f->bit_depth, f->ss_x, f->ss_y);
    OutBuf *ob = take_out(n);
    if (!ob) return false;
    OutRec o;
    o.buf = ob;
    o.t_submit = j->t_submit;
    o.fr.stream = j->st->id;
    o.fr.index = j->st->shown++;
    vp9r_frame &d = o.fr.frame;
    d.width = f->width;
    d.height = f->height;
    d.bit_depth = f->bit_depth;
    d.ss_x = f->ss_x;
    d.ss_y = f->ss_y;
    d.color_space = s->color_space;
    d.full_range = s->color_range;
    d.render_width = s->render_w;
    d.render_height = s->render_h;
    d.data = (uint8_t *)ob->buf.p;
    d.size = n;
    size_t bps = f->bit_depth > 8 ? 2 : 1, off = 0;
    for (int p = 0; p < 3; p++) {
      d.plane_width[p] = p ? (f->width + f->ss_x) >> f->ss_x : f->width;
      d.plane_height[p] = p ? (f->height + f->ss_y) >> f->ss_y : f->height;
      d.plane_offset[p] = off;
      off += (size_t)d.plane_width[p] * d.plane_height[p] * bps;
    }
    b->outs.push_back(o);
    outs_src.push_back(f);
    return true;
  }
  std::vector<vp9r_gpu_frame *> outs_src;  // dispatcher scratch: the store of each output

  // Forms the next batch (under m): per stream in order of its oldest ready
  // frame, the jobs up to and including one frame to decode.
  Batch *form_batch() {
    std::vector<Stream *> order;
    for (auto &up : streams)
      if (up->admitted && !up->finished && !up->ready.empty()) order.push_back(up.get());
    if (order.empty()) return nullptr;
    std::sort(order.begin(), order.end(),
              [](const Stream *a, const Stream *b) { return a->ready.front()->t_submit < b->ready.front()->t_submit; });
    Batch *b = new Batch();
    uint32_t frames = 0;
    for (Stream *st : order) {
      if (frames >= cfg.max_batch) break;
      bool took = false;
      while (!st->ready.empty()) {
        Job *j = st->ready.front();
        if (j->kind == kDecode && took) break;
        st->ready.pop_front();
        b->jobs.push_back(j);
        if (j->kind == kDecode) {
          took = true;
          frames++;
        }
        if (j->kind == kFail) break;
      }
    }
    cv_work.notify_all();  // room in the ready queues
    return b;
  }

  int run_batch(Batch *b) {
    std::vector<BatchFrame> frames;
    outs_src.clear();
    for (Job *j : b->jobs) {
      Stream *st = j->st;
      if (st->failed) continue;  // dropped: the stream already ended in error
      if (j->kind == kFail) {
        std::lock_guard<std::mutex> lk(m);
        fail_stream(st, j->status, j->error);
        continue;
      }
      if (j->kind == kShowExisting) {
        Store *s = st->slots[j->show_idx];
        if (!s) {
          std::lock_guard<std::mutex> lk(m);
          fail_stream(st, VP9R_ERR_BITSTREAM, "show_existing_frame names a slot that holds no frame");
          continue;
        }
        if (!add_output(b, j, s)) return VP9R_ERR_DEVICE;
        continue;
      }
      const vp9r_frame_hdr &h = j->p.hdr;
      BatchFrame bf{&j->p, (const uint8_t *)j->in->dev.p, {nullptr, nullptr, nullptr}, nullptr};
      int status = VP9R_OK;
      bool ok = true;
      if (!h.frame_is_intra)
        for (int k = 0; k < 3; k++) {
          Store *r = st->slots[h.ref_frame_idx[k]];
          if (!r || !r->fits(h.ref_width[k], h.ref_height[k], h.bit_depth, h.ss_x, h.ss_y)) ok = false;
          bf.refs[k] = r ? r->f : nullptr;
        }
      if (!ok) {
        std::lock_guard<std::mutex> lk(m);
        fail_stream(st, VP9R_ERR_BITSTREAM, "missing reference, or a reference of another size");
        continue;
      }
      Store *cur = take_store(st, h, &status);
      if (!cur) {
        if (status == VP9R_ERR_DEVICE) return status;
        std::lock_guard<std::mutex> lk(m);
        fail_stream(st, status, vp9r_last_error());
        continue;
      }
      cur->color_space = h.color_space;
      cur->color_range = h.color_range;
      cur->render_w = h.render_width;
      cur->render_h = h.render_height;
      bf.cur = cur->f;
      HIP_TRY(hipStreamWaitEvent(g->stream, j->in->uploaded, 0), "hipStreamWaitEvent");
      frames.push_back(bf);
      for (int k = 0; k < 8; k++)
        if ((h.refresh_frame_flags >> k) & 1) set_slot(st, k, cur);
      if (h.show_frame && !add_output(b, j, cur)) return VP9R_ERR_DEVICE;
    }
    b->n_frames = (uint32_t)frames.size();
    if (!frames.empty()) {
      int s = launch_batch(g, frames.data(), (int)frames.size(), true, true);
      if (s) return s;
    }
    // the shown frames to pinned host memory, behind the batch
    if (b->done == nullptr) HIP_TRY(hipEventCreateWithFlags(&b->done, hipEventDisableTiming), "hipEventCreate");
    HIP_TRY(hipEventRecord(ev_batch, g->stream), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(g->copy_stream, ev_batch, 0), "hipStreamWaitEvent");
    for (size_t i = 0; i < b->outs.size(); i++) {
      vp9r_gpu_frame *f = outs_src[i];
      uint8_t *dst = (uint8_t *)b->outs[i].buf->buf.p;
      size_t bps = f->bit_depth > 8 ? 2 : 1, off = 0;
      for (int p = 0; p < 3; p++) {
        size_t w = p ? (f->width + f->ss_x) >> f->ss_x : f->width;
        size_t hh = p ? (f->height + f->ss_y) >> f->ss_y : f->height;
        HIP_TRY(hipMemcpy2DAsync(dst + off, w * bps, f->planes.p[p], (size_t)f->planes.stride[p] * bps, w * bps, hh,
                                 hipMemcpyDeviceToHost, g->copy_stream),
                "download frame");
        off += w * hh * bps;
      }
      HIP_TRY(hipEventRecord(f->copy_done, g->copy_stream), "hipEventRecord");
      f->copy_pending = true;
    }
    HIP_TRY(hipEventRecord(b->done, g->copy_stream), "hipEventRecord");
    return VP9R_OK;
  }

  void dispatcher_main() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(m);
    while (!stop && fatal == VP9R_OK) {
      Batch *b = nullptr;
      if (inflight.size() < cfg.max_inflight) b = form_batch();
      if (!b) {
        cv_disp.wait(lk);
        continue;
      }
      if (!free_events.empty()) {
        b->done = free_events.back();
        free_events.pop_back();
      }
      lk.unlock();
      int s = run_batch(b);
      lk.lock();
      if (s) {
        fatal = s;
        fatal_msg = vp9r_last_error();
        // The batch's work may be partly queued, and its done event was not
        // recorded (a recycled event reads as complete): wait for both
        // streams here, so that the finisher releases the batch's input
        // slots only once nothing reads them.  Its frames are not handed out.
        (void)hipStreamSynchronize(g->stream);
        (void)hipStreamSynchronize(g->copy_stream);
        if (b->done) free_events.push_back(b->done);
        b->done = nullptr;
        for (OutRec &o : b->outs) free_out.push_back(o.buf);
        b->outs.clear();
      }
      n_decoded += b->n_frames;
      n_batches++;
      max_batch_seen = std::max<uint64_t>(max_batch_seen, b->n_frames);
      inflight.push_back(b);
      cv_fin.notify_one();
    }
    cv_out.notify_all();
  }

  // ---- finisher
  void finisher_main() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      if (inflight.empty()) {
        if (stop) break;
        cv_fin.wait(lk);
        continue;
      }
      Batch *b = inflight.front();
      lk.unlock();
      hipError_t e = b->done ? hipEventSynchronize(b->done) : hipSuccess;
      const double t = now_ms();
      lk.lock();
      inflight.pop_front();
      if (e != hipSuccess && fatal == VP9R_OK) {
        fatal = VP9R_ERR_DEVICE;
        fatal_msg = std::string("batch: ") + hipGetErrorString(e);
      }
      std::vector<Stream *> touched;
      for (Job *j : b->jobs) {
        if (j->in) j->in->busy = false;
        j->st->outstanding--;
        touched.push_back(j->st);
        delete j;
      }
      for (OutRec &o : b->outs) {
        o.fr.latency_ms = t - o.t_submit;
        o.fr.handle = next_handle++;
        handed_out[o.fr.handle] = o.buf;
        done.push_back(o);
        n_shown++;
      }
      if (b->done) free_events.push_back(b->done);
      delete b;
      for (Stream *st : touched) finish_stream(st);
      cv_out.notify_all();
      cv_disp.notify_one();
      cv_work.notify_all();
    }
  }
};

extern "C" {

uint64_t vp9r_frame_store_bytes(uint32_t width, uint32_t height, uint32_t bd, uint32_t ss_x, uint32_t ss_y) {
  const uint64_t sbc = ((width + 7) / 8 + 7) / 8, sbr = ((height + 7) / 8 + 7) / 8;
  const uint64_t bps = bd > 8 ? 2 : 1;
  uint64_t total = 0;
  for (int p = 0; p < 3; p++) {
    const uint64_t w = (sbc * 64) >> (p ? ss_x : 0), h = (sbr * 64) >> (p ? ss_y : 0);
    total = ((total + w * h * bps) + 255) & ~(uint64_t)255;
  }
  return total;
}

int vp9r_server_new(int device, const vp9r_server_config *cfg, vp9r_server **out) {
  *out = nullptr;
  vp9r_server *s = new (std::nothrow) vp9r_server();
  if (!s) return VP9R_ERR_INVALID_INPUT;
  s->device = device;
  if (cfg) s->cfg = *cfg;
  int st = vp9r_gpu_new(device, &s->g);
  if (st) {
    delete s;
    return st;
  }
  if (hipEventCreateWithFlags(&s->ev_batch, hipEventDisableTiming) != hipSuccess) {
    vp9r_gpu_free(s->g);
    delete s;
    set_error("hipEventCreate failed");
    return VP9R_ERR_DEVICE;
  }
  if (!s->cfg.hbm_budget) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      vp9r_gpu_free(s->g);
      delete s;
      set_error("hipMemGetInfo failed");
      return VP9R_ERR_DEVICE;
    }
    s->cfg.hbm_budget = (uint64_t)(free_b * 0.9);
  }
  if (!s->cfg.parse_threads) {
    unsigned n = std::thread::hardware_concurrency();
    s->cfg.parse_threads = n > 3 ? n - 2 : 1;
  }
  if (!s->cfg.max_batch) s->cfg.max_batch = 64;
  if (!s->cfg.max_inflight) s->cfg.max_inflight = 2;
  if (!s->cfg.parse_ahead) s->cfg.parse_ahead = 4;
  for (uint32_t i = 0; i < s->cfg.parse_threads; i++) {
    hipStream_t us = nullptr;
    if (hipStreamCreateWithFlags(&us, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t x : s->upload_streams) (void)hipStreamDestroy(x);
      vp9r_gpu_free(s->g);
      delete s;
      set_error("hipStreamCreate failed");
      return VP9R_ERR_DEVICE;
    }
    s->upload_streams.push_back(us);
  }
  for (uint32_t i = 0; i < s->cfg.parse_threads; i++) s->workers.emplace_back([s, i] { s->worker_main((int)i); });
  s->dispatcher = std::thread([s] { s->dispatcher_main(); });
  s->finisher = std::thread([s] { s->finisher_main(); });
  *out = s;
  return VP9R_OK;
}

void vp9r_server_free(vp9r_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_disp.notify_all();
  s->cv_fin.notify_all();
  s->cv_out.notify_all();
  for (auto &t : s->workers) t.join();
  s->dispatcher.join();
  s->finisher.join();
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  for (auto &up : s->streams) {
    for (Job *j : up->ready) delete j;
    for (Store *st : up->stores) {
      vp9r_gpu_frame_free(st->f);
      delete st;
    }
    up->inputs.clear();
  }
  for (Store *st : s->cache) {
    vp9r_gpu_frame_free(st->f);
    delete st;
  }
  for (auto &kv : s->handed_out) delete kv.second;
  for (OutRec &o : s->done) delete o.buf;
  for (OutBuf *o : s->free_out) delete o;
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  for (hipStream_t x : s->upload_streams) (void)hipStreamDestroy(x);
  s->streams.clear();
  s->spare_inputs.clear();
  (void)hipEventDestroy(s->ev_batch);
  vp9r_gpu_free(s->g);
  delete s;
}

int vp9r_server_open(vp9r_server *s, uint32_t max_w, uint32_t max_h, uint32_t bd, uint32_t ss_x, uint32_t ss_y,
                     uint32_t *id) {
  if (max_w == 0 || max_h == 0 || max_w > 65536 || max_h > 65536 || (bd != 8 && bd != 10 && bd != 12) || ss_x > 1 ||
      ss_y > 1) {
    set_error("invalid stream geometry");
    return VP9R_ERR_INVALID_INPUT;
  }
  const uint64_t fp = kStoresPerStream * vp9r_frame_store_bytes(max_w, max_h, bd, ss_x, ss_y);
  std::lock_guard<std::mutex> lk(s->m);
  if (fp > s->cfg.hbm_budget) {
    set_error("the stream's footprint exceeds the server's HBM budget");
    return VP9R_ERR_INVALID_INPUT;
  }
  if (s->fatal) {
    set_error(s->fatal_msg);
    return s->fatal;
  }
  auto st = std::make_unique<Stream>();
  st->id = (uint32_t)s->streams.size();
  st->max_w = max_w;
  st->max_h = max_h;
  st->bd = bd;
  st->ss_x = ss_x;
  st->ss_y = ss_y;
  st->footprint = fp;
  *id = st->id;
  s->waiting.push_back(st.get());
  s->streams.push_back(std::move(st));
  s->admit_waiting();
  return VP9R_OK;
}

int vp9r_server_submit(vp9r_server *s, uint32_t id, const uint8_t *packet, size_t len) {
  const double t = now_ms();
  std::lock_guard<std::mutex> lk(s->m);
  if (id >= s->streams.size() || s->streams[id]->closed) {
    set_error("no such open stream");
    return VP9R_ERR_INVALID_INPUT;
  }
  if (s->fatal) {
    set_error(s->fatal_msg);
    return s->fatal;
  }
  Stream *st = s->streams[id].get();
  if (st->failed) return VP9R_OK;  // its end record carries the error
  st->packets.emplace_back(std::vector<uint8_t>(packet, packet + len), t);
  s->cv_work.notify_one();
  return VP9R_OK;
}

int vp9r_server_close(vp9r_server *s, uint32_t id) {
  std::lock_guard<std::mutex> lk(s->m);
  if (id >= s->streams.size() || s->streams[id]->closed) {
    set_error("no such open stream");
    return VP9R_ERR_INVALID_INPUT;
  }
  s->streams[id]->closed = true;
  s->finish_stream(s->streams[id].get());
  return VP9R_OK;
}

int vp9r_server_poll(vp9r_server *s, vp9r_server_frame *out, int timeout_ms) {
  std::unique_lock<std::mutex> lk(s->m);
  auto ready = [&] { return !s->done.empty() || s->fatal != VP9R_OK; };
  if (timeout_ms < 0)
    s->cv_out.wait(lk, ready);
  else
    s->cv_out.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
  if (!s->done.empty()) {
    *out = s->done.front().fr;
    s->done.pop_front();
    return VP9R_OK;
  }
  if (s->fatal) {
    set_error(s->fatal_msg);
    return s->fatal;
  }
  return VP9R_AGAIN;
}

void vp9r_server_release(vp9r_server *s, uint64_t handle) {
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->handed_out.find(handle);
  if (it == s->handed_out.end()) return;
  s->free_out.push_back(it->second);
  s->handed_out.erase(it);
}

int vp9r_server_stats(vp9r_server *s, uint64_t out[8]) {
  std::lock_guard<std::mutex> lk(s->m);
  out[0] = s->n_shown;
  out[1] = s->n_decoded;
  out[2] = s->n_batches;
  out[3] = s->max_batch_seen;
  out[4] = s->n_admitted;
  out[5] = s->waiting.size();
  out[6] = s->store_bytes;
  out[7] = s->cfg.hbm_budget;
  return VP9R_OK;
}

}  // extern "C"
