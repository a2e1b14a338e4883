// C ABI of the CPU side: the parser, superframe splitting and the error
// message (include/vp9r.h).
#include <cstring>
#include <new>
#include <string>

#include "parser.h"
#include "vp9r.h"

namespace vp9r {
thread_local std::string g_last_error;
void set_error(const std::string &m) { g_last_error = m; }
}  // namespace vp9r

extern "C" {

const char *vp9r_last_error(void) { return vp9r::g_last_error.c_str(); }

// Annex B (src/superframe.rs:6-52): the frames of a superframe, or the whole
// packet when it has no valid index; zero-size entries are skipped.
size_t vp9r_superframe_split(const uint8_t *packet, size_t len, size_t *offsets, size_t *sizes,
                             size_t max) {
  size_t n = 0;
  if (len > 0) {
    uint8_t last = packet[len - 1];
    if ((last & 0xe0) == 0xc0) {
      size_t sz_bytes = ((last >> 3) & 3) + 1, frames = (last & 7) + 1;
      size_t index_len = 2 + frames * sz_bytes;
      if (len >= index_len && packet[len - index_len] == last) {
        size_t p = len - index_len + 1, total = 0;
        size_t fs[8];
        for (size_t i = 0; i < frames; i++) {
          size_t s = 0;
          for (size_t b = 0; b < sz_bytes; b++) s |= (size_t)packet[p + b] << (8 * b);
          p += sz_bytes;
          total += s;
          fs[i] = s;
        }
        if (total <= len - index_len) {
          size_t pos = 0;
          for (size_t i = 0; i < frames; i++) {
            if (fs[i] == 0) continue;
            if (n < max) {
              offsets[n] = pos;
              sizes[n] = fs[i];
            }
            n++;
            pos += fs[i];
          }
          return n;
        }
      }
    }
  }
  if (max > 0) {
    offsets[0] = 0;
    sizes[0] = len;
  }
  return 1;
}

vp9r_parser *vp9r_parser_new(void) { return new (std::nothrow) vp9r_parser(); }
void vp9r_parser_free(vp9r_parser *p) { delete p; }
void vp9r_parser_set_max_pixels(vp9r_parser *p, uint64_t pixels) { p->p.max_pixels = pixels; }

void vp9r_parser_set_threads(vp9r_parser *p, uint32_t threads) { p->p.tile_threads = threads ? threads : 1; }

int vp9r_parser_parse_frame(vp9r_parser *p, const uint8_t *data, size_t len, vp9r_record *out) {
  try {
    p->p.parse_frame(data, len, out);
    vp9r::g_last_error.clear();
    return VP9R_OK;
  } catch (const vp9r::ParseError &e) {
    vp9r::g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    vp9r::g_last_error = "out of memory";
    return VP9R_ERR_INVALID_INPUT;
  }
}

}  // extern "C"

extern "C" size_t vp9r_packed_size(uint32_t width, uint32_t height, uint32_t bit_depth, uint32_t ss_x,
                                   uint32_t ss_y) {
  size_t bps = bit_depth > 8 ? 2 : 1;
  size_t cw = (width + ss_x) >> ss_x, ch = (height + ss_y) >> ss_y;
  return bps * ((size_t)width * height + 2 * cw * ch);
}
