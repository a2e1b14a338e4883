// Malformed-input driver for the CPU side of the decoder: the product parser
// and the oracle's reconstruction, built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/fuzz/Makefile).  Test infrastructure.
//
// Follows the reference's tests/fuzz.rs:28-64: arbitrary bytes (alone, and
// as the second frame after a valid key frame), and real streams with bits
// flipped, a packet cut and garbage spliced in, decoded with a pixel limit
// of 1 << 20.  Every frame must end in a status, never a crash; a frame the
// parser accepts is reconstructed and loop filtered by the oracle.  Each
// frame is parsed twice, with one thread and with four (tile columns in
// parallel), and the two must agree on the status, the error message and,
// for accepted frames, the record.
//
//   parser_fuzz PACKETS SEED CASES
//
// PACKETS: the committed streams' packets, as a sequence of
//   u32 stream count, then per stream: u32 packet count, then per packet
//   u32 length and the bytes (little endian).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "vp9r.h"
#include "../../oracle/vp9_oracle.h"

typedef std::vector<uint8_t> Bytes;

static uint64_t rng_state;
static uint64_t next_u64() {  // splitmix64
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static size_t below(size_t n) { return n ? (size_t)(next_u64() % n) : 0; }

static bool read_u32(FILE *f, uint32_t *v) {
  uint8_t b[4];
  if (fread(b, 1, 4, f) != 4) return false;
  *v = b[0] | b[1] << 8 | b[2] << 16 | (uint32_t)b[3] << 24;
  return true;
}

static bool same_bytes(const void *a, const void *b, size_t n) { return n == 0 || !memcmp(a, b, n); }

// The records of the two parses must be the same.
static bool same_record(const vp9r_record &a, const vp9r_record &b) {
  if (memcmp(&a.hdr, &b.hdr, sizeof a.hdr) || a.n_blocks != b.n_blocks || a.n_txbs != b.n_txbs ||
      a.n_coefs != b.n_coefs)
    return false;
  if (a.hdr.show_existing_frame) return true;
  size_t n_mi = (size_t)a.hdr.mi_rows * a.hdr.mi_cols;
  return same_bytes(a.blocks, b.blocks, sizeof(vp9r_block) * a.n_blocks) &&
         same_bytes(a.txbs, b.txbs, sizeof(vp9r_txb) * a.n_txbs) &&
         same_bytes(a.coef_pos, b.coef_pos, sizeof(uint16_t) * a.n_coefs) &&
         same_bytes(a.coef_val, b.coef_val, sizeof(int32_t) * a.n_coefs) &&
         same_bytes(a.mi, b.mi, sizeof(vp9r_mi) * n_mi);
}

struct FrameDel {
  void operator()(orc_frame *f) const { orc_frame_free(f); }
};
typedef std::shared_ptr<orc_frame> OFrame;

// The CPU decoder of the tests: parser + oracle, 8 slots.
struct Dec {
  vp9r_parser *p1 = vp9r_parser_new(), *p4 = vp9r_parser_new();
  OFrame slots[8];
  uint64_t ok = 0, errors = 0;
  Dec() {
    vp9r_parser_set_max_pixels(p1, 1 << 20);
    vp9r_parser_set_max_pixels(p4, 1 << 20);
    vp9r_parser_set_threads(p4, 4);
  }
  ~Dec() {
    vp9r_parser_free(p1);
    vp9r_parser_free(p4);
  }
  void frame(const uint8_t *d, size_t n) {
    vp9r_record r1, r4;
    int s1 = vp9r_parser_parse_frame(p1, d, n, &r1);
    std::string m1 = s1 ? vp9r_last_error() : "";
    int s4 = vp9r_parser_parse_frame(p4, d, n, &r4);
    std::string m4 = s4 ? vp9r_last_error() : "";
    if (s1 != s4 || m1 != m4 || (!s1 && !same_record(r1, r4))) {
      fprintf(stderr, "parse with 1 and 4 threads differ: %d '%s' / %d '%s'\n", s1, m1.c_str(), s4, m4.c_str());
      exit(3);
    }
    if (s1) {
      errors++;
      return;
    }
    ok++;
    const vp9r_frame_hdr &h = r1.hdr;
    if (h.show_existing_frame) return;
    OFrame cur(orc_frame_new(h.width, h.height, h.bit_depth, h.ss_x, h.ss_y), FrameDel());
    const orc_frame *refs[3] = {nullptr, nullptr, nullptr};
    if (!h.frame_is_intra)
      for (int i = 0; i < 3; i++) refs[i] = slots[h.ref_frame_idx[i]].get();
    if (orc_reconstruct(&r1, refs, cur.get())) return;
    orc_loop_filter(&r1, cur.get());
    for (int i = 0; i < 8; i++)
      if ((h.refresh_frame_flags >> i) & 1) slots[i] = cur;
  }
  void packet(const Bytes &pk) {
    size_t offs[8], sizes[8];
    size_t n = vp9r_superframe_split(pk.data(), pk.size(), offs, sizes, 8);
    for (size_t i = 0; i < n && i < 8; i++) frame(pk.data() + offs[i], sizes[i]);
  }
};

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: parser_fuzz PACKETS SEED CASES\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<std::vector<Bytes>> streams;
  uint32_t ns;
  if (!read_u32(f, &ns)) return 2;
  for (uint32_t s = 0; s < ns; s++) {
    uint32_t np;
    if (!read_u32(f, &np)) return 2;
    std::vector<Bytes> pk(np);
    for (auto &p : pk) {
      uint32_t len;
      if (!read_u32(f, &len)) return 2;
      p.resize(len);
      if (len && fread(p.data(), 1, len, f) != len) return 2;
    }
    streams.push_back(pk);
  }
  fclose(f);
  if (streams.empty()) return 2;
  rng_state = strtoull(argv[2], nullptr, 10);
  const long cases = strtol(argv[3], nullptr, 10);
  uint64_t ok = 0, errors = 0;
  for (long c = 0; c < cases; c++) {
    // arbitrary bytes, alone and after a valid key frame (fuzz.rs:28-44)
    Bytes junk(below(4096));
    for (auto &b : junk) b = (uint8_t)next_u64();
    {
      Dec d;
      d.packet(junk);
      ok += d.ok;
      errors += d.errors;
    }
    {
      Dec d;
      d.packet(streams[0][0]);
      d.packet(junk);
      ok += d.ok;
      errors += d.errors;
    }
    // a mutated stream (fuzz.rs:46-64): 1..7 bit flips, one packet cut,
    // up to 63 garbage bytes spliced into another
    std::vector<Bytes> s = streams[below(streams.size())];
    const size_t n = s.size();
    for (size_t k = 1 + below(7); k--;) {
      Bytes &p = s[below(n)];
      if (!p.empty()) p[below(p.size())] ^= (uint8_t)(1u << below(8));
    }
    Bytes &cut = s[below(n)];
    cut.resize(below(cut.size() + 1));
    Bytes &sp = s[below(n)];
    Bytes garbage(below(64));
    for (auto &b : garbage) b = (uint8_t)next_u64();
    sp.insert(sp.begin() + below(sp.size() + 1), garbage.begin(), garbage.end());
    Dec d;
    for (const Bytes &p : s) d.packet(p);
    ok += d.ok;
    errors += d.errors;
  }
  printf("parser_fuzz: %ld cases, %llu frames parsed, %llu errors\n", cases, (unsigned long long)ok,
         (unsigned long long)errors);
  return 0;
}
