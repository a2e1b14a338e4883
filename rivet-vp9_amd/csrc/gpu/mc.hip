// Inter prediction (8.5.2): motion vector selection, clamping and scaling
// (src/decoder/recon.rs:106-217, src/decoder/mod.rs:495-541), the separable
// 8-tap filter with per-pass rounding and clamping (src/dsp/inter.rs:289-395)
// and compound averaging (src/dsp/pixel.rs:22-50).
//
// One 64-lane workgroup predicts one tile of at most 16 x 16 samples of a
// prediction unit: its reference footprint (up to 38 x 38 for a 2:1 scaled
// reference) is staged in LDS with the reference's edge samples repeated
// (the specification's clamped coordinates), the horizontal pass writes an
// LDS intermediate, the vertical pass gives each lane up to 4 samples.  One code
// path serves every step: the specification's formula, of which the
// reference's unscaled fast path is a special case (its own test,
// inter.rs:486-564, checks they agree).
//
// 8-bit tiles at steps of 16 (every unscaled reference) whose sides are
// multiples of 4 take a second path: the footprint is staged as dwords, and
// each 8-tap sum is two v_dot4_i32_i8 on samples biased by -128 (they then
// fit signed bytes; sum f (p - 128) + 128 * 128 = sum f p, the taps of every
// fractional phase fitting signed bytes and summing to 128), a lane producing
// four samples per pass: four rows of one column in the horizontal pass
// (stored column-major), four columns of one row in the vertical pass (one
// dword of the output).
#include <type_traits>

#include <mutex>

#include "dev_common.h"
#include "frame_dev.h"
#include "launch.h"
#include "lf_mask.h"
#include "mc_pos.h"
#include "vp9_consts.h"
#include "vp9_tables.h"

namespace vp9g {

__constant__ int16_t c_filters[4][16][8];
// the taps as signed bytes, taps 0..3 then 4..7 (phase 0 unused)
__constant__ uint32_t c_taps8[4][16][2];
// the taps as signed 16-bit pairs (taps 2k, 2k + 1 in dword k)
__constant__ uint32_t c_taps16[4][16][4];

// The tap tables go to each device's constant memory once, on the first MC
// launch on that device (under a lock: contexts of several threads may launch
// at once).  A failed copy leaves the device unmarked, and its error is the
// one hipGetLastError() reports to the launch's caller.
static std::mutex filters_mu;
static uint64_t filters_ready = 0;  // bit d: device d holds the tables
static void upload_filters() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
  std::lock_guard<std::mutex> lock(filters_mu);
  if ((filters_ready >> dev) & 1) return;
  int16_t f[4][16][8];
  uint32_t t8[4][16][2] = {}, t16[4][16][4] = {};
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 16; b++)
      for (int c = 0; c < 8; c++) {
        f[a][b][c] = (int16_t)VP9_SUBPEL_FILTERS[a][b][c];
        t8[a][b][c >> 2] |= (uint32_t)(uint8_t)(int8_t)VP9_SUBPEL_FILTERS[a][b][c] << (8 * (c & 3));
        t16[a][b][c >> 1] |= (uint32_t)(uint16_t)(int16_t)VP9_SUBPEL_FILTERS[a][b][c] << (16 * (c & 1));
      }
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_filters), f, sizeof f) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_taps8), t8, sizeof t8) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_taps16), t16, sizeof t16) != hipSuccess)
    return;
  filters_ready |= 1ull << dev;
}

// A tile is at most 16 x 16 samples; its footprint at most 38 x 38 (a 2:1
// step: ((16 - 1) * 32 + 15) / 16 + 8).
constexpr int TILE = 16;
constexpr int FPW = 40;  // footprint row pitch in LDS
constexpr int FPH = 38;

struct McShared {
  uint16_t fp[FPH * FPW];  // reference footprint, edge samples repeated
  int16_t hp[FPH * TILE];  // after the horizontal pass
};

// The 8-bit unscaled path's staging for tiles of up to MAXH rows: footprint
// rows of FP8 dwords (bytes from a dword-aligned column; the pitch of 10
// puts the row groups a read touches in distinct banks), and the horizontal
// pass column-major (column c holds rows 0 .. MAXH + 7 in HP dwords).
// Each wave of an MC workgroup works on a tile of its own, in LDS of its
// own: the tile's LDS exchanges need the wave's own accesses ordered, not a
// workgroup barrier (wave_lds_sync).

constexpr int FP8 = 10;
template <int MAXH>
struct Mc8Shared {
  static constexpr int ROWS = MAXH + 8, HP = ROWS / 4;
  uint32_t fp[ROWS * FP8];
  uint32_t hp[16 * HP];
};

// The 10/12-bit unscaled path's staging: footprint rows of FP16 dwords (two
// samples each, from an even column), the horizontal pass column-major in
// sample pairs (column c holds rows 0 .. MAXH + 7 in HW dwords).
constexpr int FP16 = 13;
template <int MAXH>
struct Mc16Shared {
  static constexpr int ROWS = MAXH + 8, HW = ROWS / 2;
  uint32_t fp[ROWS * FP16];
  uint32_t hp[16 * HW];
};

union McLds {
  McShared g;
  Mc8Shared<16> f16;
  Mc8Shared<32> f32;
  Mc16Shared<16> h16;
  Mc16Shared<32> h32;
};

// Byte-wise rounding average of four packed samples: (a + b + 1) >> 1
// (pixel.rs:22-50).
DEV uint32_t avg4(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) >> 1) & 0x7f7f7f7fu); }

// One 8-tap sum over the biased window of bytes s .. s + 7 of (w0, w1, w2)
// (samples p - 128): since the taps sum to 128,
// Round2(sum f p, 7) - 128 = (sum f (p - 128) + 64) >> 7, returned clamped to
// [-128, 127] (the 8-bit range, still biased).  The clamp is an explicit
// v_med3_i32: left to itself the compiler fuses two such clamps and their
// packing into v_ashr_pk_u8_i32 and takes the result's upper half as zero,
// but the instruction keeps its destination's previous upper half.
DEV uint32_t tap8(uint32_t w0, uint32_t w1, uint32_t w2, int s, uint32_t f0, uint32_t f1) {
  const uint32_t lo = __builtin_amdgcn_alignbyte(w1, w0, s);
  const uint32_t hi = __builtin_amdgcn_alignbyte(w2, w1, s);
  const int sum = __builtin_amdgcn_sdot4((int)lo, (int)f0, __builtin_amdgcn_sdot4((int)hi, (int)f1, 64, false), false);
  uint32_t v;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(v) : "v"(sum >> 7), "v"(-128), "v"(127));
  return v;
}

// The low bytes of four values packed, a first (three byte permutes).
DEV uint32_t pack4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  const uint32_t ab = __builtin_amdgcn_perm(b, a, 0x0c0c0400u), cd = __builtin_amdgcn_perm(d, c, 0x0c0c0400u);
  return __builtin_amdgcn_perm(cd, ab, 0x05040100u);
}

// Four samples packed: filtered (taps f0, f1; biased) or the plain bytes at
// byte offset s of rows w, w + pitch, w + 2 pitch, w + 3 pitch.
template <int PITCH>
DEV uint32_t four_samples(const uint32_t *w, int s, bool filter, uint32_t f0, uint32_t f1) {
  uint32_t v[4];
  if (filter) {
    uint32_t a[4][3];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int m = 0; m < 3; m++) a[j][m] = w[j * PITCH + m];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = tap8(a[j][0], a[j][1], a[j][2], s, f0, f1);
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = w[j * PITCH] >> (8 * s);
  }
  return pack4(v[0], v[1], v[2], v[3]);
}

// The 8-bit tile of tw x th samples (multiples of 4; tw at most 16, th at
// most MAXH, 16 or 32) whose first sample is at (x0, y0) (1/16 sample, steps
// 16) from `ref` (bottom-right sample (last_x, last_y)).  res[h] is the
// lane's output dword of row 16 h + lane / 4, columns 4 (lane % 4) .. + 3
// (meaningful when inside the tile).  A footprint wholly inside the plane is
// read as dwords; otherwise sample by sample at clamped coordinates.  All 64
// lanes call it (it holds barriers).
template <int MAXH>
DEV void mc_tile8(const uint8_t *ref, uint32_t stride, int last_x, int last_y, int x0, int y0, int tw, int th,
                  int filter, Mc8Shared<MAXH> &sh, uint32_t (&res)[MAXH / 16]) {
  constexpr int HP = Mc8Shared<MAXH>::HP, LR = (MAXH + 7 + 7) / 8;  // load rounds of 8 rows
  const int lane = threadIdx.x & 63;
  const int px = x0 & 15, py = y0 & 15;
  const bool fx = px != 0, fy = py != 0;
  const int fx0 = (x0 >> 4) - (fx ? 3 : 0), fy0 = (y0 >> 4) - (fy ? 3 : 0);
  const int iw = tw + (fx ? 7 : 0), ih = th + (fy ? 7 : 0);
  const int ax = fx0 & ~3, off = fx0 - ax, nd = (off + iw + 3) >> 2;
  const bool inside = fx0 >= 0 && fy0 >= 0 && ax + 4 * nd - 1 <= last_x && fy0 + ih - 1 <= last_y &&
                      ((stride | (uint32_t)(uintptr_t)ref) & 3) == 0;
  // footprint: lane (row lane / 8 + 8 i, dword lane % 8), all loads issued
  // before the first store; the horizontal pass's input is biased by -128
  const int dd = lane & 7, r0 = lane >> 3;
  uint32_t w[LR];
#pragma unroll
  for (int i = 0; i < LR; i++) w[i] = 0;
  if (inside) {
#pragma unroll
    for (int i = 0; i < LR; i++)
      if (dd < nd && r0 + 8 * i < ih)
        w[i] = gld((const uint32_t *)(ref + (size_t)(fy0 + r0 + 8 * i) * stride + ax + 4 * dd));
  } else {
#pragma unroll
    for (int i = 0; i < LR; i++)
      if (dd < nd && r0 + 8 * i < ih) {
        const uint8_t *row = ref + (size_t)iclamp(fy0 + r0 + 8 * i, 0, last_y) * stride;
#pragma unroll
        for (int b = 0; b < 4; b++) w[i] |= (uint32_t)gld(row + iclamp(ax + 4 * dd + b, 0, last_x)) << (8 * b);
      }
  }
  const uint32_t flip = fx ? 0x80808080u : 0u;
#pragma unroll
  for (int i = 0; i < LR; i++)
    if (dd < nd && r0 + 8 * i < ih) sh.fp[(r0 + 8 * i) * FP8 + dd] = w[i] ^ flip;
  wave_lds_sync();
  // horizontal: lane task t (column t % 16, rows 4 (t / 16) .. + 3) for
  // t = lane + 64 h.  Its results are biased when filtered (fx); the
  // vertical pass wants them biased when it filters (fy)
  const uint32_t f0 = c_taps8[filter][px][0], f1 = c_taps8[filter][px][1];
  const uint32_t vflip = fx != fy ? 0x80808080u : 0u;
#pragma unroll
  for (int h = 0; h < (16 * HP + 63) / 64; h++) {
    const int t = lane + 64 * h, c = t & 15, rq = t >> 4;
    if (t < 16 * HP && c < tw && 4 * rq < ih) {
      const int o = off + c;
      sh.hp[c * HP + rq] = four_samples<FP8>(sh.fp + 4 * rq * FP8 + (o >> 2), o & 3, fx, f0, f1) ^ vflip;
    }
  }
  wave_lds_sync();
  // vertical: lane task (row 16 h + lane / 4, columns 4 q .. 4 q + 3)
  const int q = lane & 3;
  const uint32_t g0 = c_taps8[filter][py][0], g1 = c_taps8[filter][py][1];
#pragma unroll
  for (int h = 0; h < MAXH / 16; h++) {
    const int r = 16 * h + (lane >> 2);
    res[h] = 0;
    // four_samples reads columns 4 q .. 4 q + 3 as its "rows" (pitch HP);
    // filtered results are biased
    if (r < th && 4 * q < tw)
      res[h] = four_samples<HP>(sh.hp + 4 * q * HP + (r >> 2), r & 3, fy, g0, g1) ^ (fy ? 0x80808080u : 0u);
  }
  wave_lds_sync();
}

// Per-sample rounding average of two packed pairs: (a + b + 1) >> 1.
DEV uint32_t avg2(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) >> 1) & 0x7fff7fffu); }

typedef short McV2 __attribute__((ext_vector_type(2)));

// One 8-tap sum over samples s .. s + 7 (s = 0 or 1) of the pairs
// w0 .. w4, rounded by 7 and clamped to [0, max].
DEV uint32_t tap8w(const uint32_t (&w)[5], int s, const uint32_t (&f)[4], int max) {
  int sum = 64;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t d = __builtin_amdgcn_alignbyte(w[k + 1], w[k], 2 * s);
    sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(McV2, d), __builtin_bit_cast(McV2, f[k]), sum, false);
  }
  uint32_t v;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(v) : "v"(sum >> 7), "v"(max));
  return v;
}

// Four samples as two dwords (first pair, second pair): filtered (taps f)
// or plain, at sample offset s of rows w, w + pitch, ... (rows j: dwords
// w[j * pitch] ..).
template <int PITCH>
DEV void four_samples16(const uint32_t *w, int s, bool filter, const uint32_t (&f)[4], int max, uint32_t (&out)[2]) {
  uint32_t v[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (filter) {
      uint32_t a[5];
#pragma unroll
      for (int m = 0; m < 5; m++) a[m] = w[j * PITCH + m];
      v[j] = tap8w(a, s, f, max);
    } else {
      v[j] = (__builtin_amdgcn_alignbyte(w[j * PITCH + 1], w[j * PITCH], 2 * s)) & 0xffffu;
    }
  }
  out[0] = v[0] | v[1] << 16;
  out[1] = v[2] | v[3] << 16;
}

// The 10/12-bit tile of tw x th samples (multiples of 4; tw at most 16, th
// at most MAXH) at (x0, y0) (1/16 sample, steps 16): the 8-bit path's
// structure on sample pairs, each 8-tap sum four v_dot2_i32_i16 (samples
// and taps fit signed 16 bits, the sums 32).  res[h][0..1] are the lane's
// four samples of row 16 h + lane / 4, columns 4 (lane % 4) .. + 3.
template <int MAXH>
DEV void mc_tile16(const uint16_t *ref, uint32_t stride, int last_x, int last_y, int x0, int y0, int tw, int th,
                   int filter, int max, Mc16Shared<MAXH> &sh, uint32_t (&res)[MAXH / 16][2]) {
  constexpr int HW = Mc16Shared<MAXH>::HW, LR = (MAXH + 7 + 3) / 4;  // load rounds of 4 rows
  const int lane = threadIdx.x & 63;
  const int px = x0 & 15, py = y0 & 15;
  const bool fx = px != 0, fy = py != 0;
  const int fx0 = (x0 >> 4) - (fx ? 3 : 0), fy0 = (y0 >> 4) - (fy ? 3 : 0);
  const int iw = tw + (fx ? 7 : 0), ih = th + (fy ? 7 : 0);
  const int ax = fx0 & ~1, off = fx0 - ax, nd = (off + iw + 1) >> 1;
  const bool inside = fx0 >= 0 && fy0 >= 0 && ax + 2 * nd - 1 <= last_x && fy0 + ih - 1 <= last_y &&
                      ((uint32_t)(uintptr_t)ref & 3) == 0 && (stride & 1) == 0;
  // footprint: lane (row lane / 16 + 4 i, dword lane % 16)
  const int dd = lane & 15, r0 = lane >> 4;
  uint32_t w[LR];
#pragma unroll
  for (int i = 0; i < LR; i++) w[i] = 0;
  if (inside) {
#pragma unroll
    for (int i = 0; i < LR; i++)
      if (dd < nd && r0 + 4 * i < ih) w[i] = gld((const uint32_t *)(ref + (size_t)(fy0 + r0 + 4 * i) * stride + ax + 2 * dd));
  } else {
#pragma unroll
    for (int i = 0; i < LR; i++)
      if (dd < nd && r0 + 4 * i < ih) {
        const uint16_t *row = ref + (size_t)iclamp(fy0 + r0 + 4 * i, 0, last_y) * stride;
        w[i] = (uint32_t)gld(row + iclamp(ax + 2 * dd, 0, last_x)) |
               (uint32_t)gld(row + iclamp(ax + 2 * dd + 1, 0, last_x)) << 16;
      }
  }
#pragma unroll
  for (int i = 0; i < LR; i++)
    if (dd < nd && r0 + 4 * i < ih) sh.fp[(r0 + 4 * i) * FP16 + dd] = w[i];
  wave_lds_sync();
  uint32_t f[4], g[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    f[k] = c_taps16[filter][px][k];
    g[k] = c_taps16[filter][py][k];
  }
  // horizontal: lane task t (column t % 16, rows 4 (t / 16) .. + 3)
#pragma unroll
  for (int h = 0; h < (16 * (HW / 2) + 63) / 64; h++) {
    const int t = lane + 64 * h, c = t & 15, rq = t >> 4;
    if (t < 16 * (HW / 2) && c < tw && 4 * rq < ih) {
      const int o = off + c;
      uint32_t two[2];
      four_samples16<FP16>(sh.fp + 4 * rq * FP16 + (o >> 1), o & 1, fx, f, max, two);
      sh.hp[c * HW + 2 * rq] = two[0];
      sh.hp[c * HW + 2 * rq + 1] = two[1];
    }
  }
  wave_lds_sync();
  // vertical: lane task (row 16 h + lane / 4, columns 4 q .. 4 q + 3)
  const int q = lane & 3;
#pragma unroll
  for (int h = 0; h < MAXH / 16; h++) {
    const int r = 16 * h + (lane >> 2);
    res[h][0] = res[h][1] = 0;
    if (r < th && 4 * q < tw) four_samples16<HW>(sh.hp + 4 * q * HW + (r >> 1), r & 1, fy, g, max, res[h]);
  }
  wave_lds_sync();
}

// Predicts the tw x th (<= 16 x 16) tile whose first sample is at (x0, y0)
// (1/16 sample) with steps (xs, ys) from `ref`, whose bottom-right sample is
// (last_x, last_y).  Output o = lane + 64 i (row o / tw, column o % tw) goes
// to out[i].  The specification's formula (8.5.2.3): each pass rounds by 7
// and clamps; a pass whose positions are all whole samples (step 16, phase
// 0) is the identity and is skipped, with its 7 extra footprint samples.
// All 64 lanes call it (it holds barriers).
template <class P>
DEV void mc_tile(const P *ref, uint32_t stride, int last_x, int last_y, int x0, int y0, int xs, int ys,
                 int tw, int th, int filter, int max, McShared &sh, int out[4]) {
  const int lane = threadIdx.x & 63;
  const bool fx = xs != 16 || (x0 & 15), fy = ys != 16 || (y0 & 15);
  const int fx0 = (x0 >> 4) - (fx ? 3 : 0), fy0 = (y0 >> 4) - (fy ? 3 : 0);
  const int iw = fx ? ((((tw - 1) * xs + (x0 & 15)) >> 4) + 8) : tw;
  const int ih = fy ? ((((th - 1) * ys + (y0 & 15)) >> 4) + 8) : th;
  for (int k = lane; k < ih * iw; k += 64) {
    const int rr = k / iw, cc = k - rr * iw;
    const int sy = iclamp(fy0 + rr, 0, last_y), sx = iclamp(fx0 + cc, 0, last_x);
    sh.fp[rr * FPW + cc] = ref[(size_t)sy * stride + sx];
  }
  wave_lds_sync();
  for (int k = lane; k < ih * tw; k += 64) {
    const int rr = k / tw, cc = k - rr * tw;
    int v;
    if (fx) {
      const int p = x0 + xs * cc;
      const int16_t *f = c_filters[filter][p & 15];
      const uint16_t *src = sh.fp + rr * FPW + (p >> 4) - 3 - fx0;
      int sum = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) sum += f[i] * (int)src[i];
      v = iclamp((sum + 64) >> 7, 0, max);
    } else {
      v = sh.fp[rr * FPW + cc];
    }
    sh.hp[rr * TILE + cc] = (int16_t)v;
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int o = lane + 64 * i;
    int v = 0;
    if (o < th * tw) {
      const int r = o / tw, c = o - r * tw;
      if (fy) {
        const int p = (y0 & 15) + ys * r;
        const int16_t *f = c_filters[filter][p & 15];
        const int16_t *src = sh.hp + (p >> 4) * TILE + c;
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) sum += f[k] * (int)src[k * TILE];
        v = iclamp((sum + 64) >> 7, 0, max);
      } else {
        v = sh.hp[r * TILE + c];
      }
    }
    out[i] = v;
  }
  wave_lds_sync();
}

template <int TH>
DEV Mc8Shared<TH> &mc8_lds(McLds &sh) {
  if constexpr (TH == 32)
    return sh.f32;
  else
    return sh.f16;
}
template <int TH>
DEV Mc16Shared<TH> &mc16_lds(McLds &sh) {
  if constexpr (TH == 32)
    return sh.h32;
  else
    return sh.h16;
}

// The frame kernel's unscaled fast paths on a tile of TH rows (16, or 32
// for the tiles the host makes tall: 8-bit or high bit depth, both
// references unscaled), the same arithmetic as the per-kernel seam's.
template <int TH>
DEV void mc_frame8(const McArgs &a, const McTile &t, const RefPos &rp0, const RefPos &rp1, int r0, int r1,
                   bool compound, int filter, uint8_t *dst, uint32_t stride, McLds &sh) {
  constexpr int NH = TH / 16;
  uint32_t v[NH], v2[NH];
  mc_tile8<TH>((const uint8_t *)a.ref[r0].p[t.plane], a.ref[r0].stride[t.plane], rp0.last_x, rp0.last_y,
               rp0.x + 16 * t.tx, rp0.y + 16 * t.ty, t.tw, t.th, filter, mc8_lds<TH>(sh), v);
  if (compound) {
    mc_tile8<TH>((const uint8_t *)a.ref[r1].p[t.plane], a.ref[r1].stride[t.plane], rp1.last_x, rp1.last_y,
                 rp1.x + 16 * t.tx, rp1.y + 16 * t.ty, t.tw, t.th, filter, mc8_lds<TH>(sh), v2);
#pragma unroll
    for (int h = 0; h < NH; h++) v[h] = avg4(v[h], v2[h]);
  }
  const int q = threadIdx.x & 3;
#pragma unroll
  for (int h = 0; h < NH; h++) {
    const int r = 16 * h + ((threadIdx.x & 63) >> 2);
    if (r < t.th && 4 * q < t.tw)
      gst((uint32_t *)(dst + (size_t)(t.y + t.ty + r) * stride + t.x + t.tx + 4 * q), v[h]);
  }
}

template <int TH>
DEV void mc_frame16(const McArgs &a, const McTile &t, const RefPos &rp0, const RefPos &rp1, int r0, int r1,
                    bool compound, int filter, int max, uint16_t *dst, uint32_t stride, McLds &sh) {
  constexpr int NH = TH / 16;
  uint32_t v[NH][2], v2[NH][2];
  mc_tile16<TH>((const uint16_t *)a.ref[r0].p[t.plane], a.ref[r0].stride[t.plane], rp0.last_x, rp0.last_y,
                rp0.x + 16 * t.tx, rp0.y + 16 * t.ty, t.tw, t.th, filter, max, mc16_lds<TH>(sh), v);
  if (compound) {
    mc_tile16<TH>((const uint16_t *)a.ref[r1].p[t.plane], a.ref[r1].stride[t.plane], rp1.last_x, rp1.last_y,
                  rp1.x + 16 * t.tx, rp1.y + 16 * t.ty, t.tw, t.th, filter, max, mc16_lds<TH>(sh), v2);
#pragma unroll
    for (int h = 0; h < NH; h++) {
      v[h][0] = avg2(v[h][0], v2[h][0]);
      v[h][1] = avg2(v[h][1], v2[h][1]);
    }
  }
  const int q = threadIdx.x & 3;
#pragma unroll
  for (int h = 0; h < NH; h++) {
    const int r = 16 * h + ((threadIdx.x & 63) >> 2);
    if (r < t.th && 4 * q < t.tw) {
      uint32_t *o = (uint32_t *)(dst + (size_t)(t.y + t.ty + r) * stride + t.x + t.tx + 4 * q);
      gst(o, v[h][0]);
      gst(o + 1, v[h][1]);
    }
  }
}

// W waves per workgroup, each on a tile of its own (fewer workgroups to
// dispatch).  A launch of one frame (ONE) reads its tiles straight from the
// tiles argument (n1 of them), beside the frame's arguments; a batch finds
// its frame and tiles through the range table (W tiles per workgroup).
// Workgroup b of nb (the launch's MC part; b % 8 is the workgroup's XCD).
template <bool HBD, bool ONE, int W>
DEV void mc_wg(const McTile *tiles1, uint32_t n1, const McArgs *fa, const Range *rg, uint32_t nf, uint32_t b,
               uint32_t nb) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ McLds shw[W];
  const int wave = W == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  McLds &sh = shw[wave];
  const uint32_t wg = xcd_order(b, nb);
  const uint32_t fi = ONE ? 0 : range_frame(rg, nf, wg);
  const uint32_t local = (wg - (ONE ? 0 : rg[fi].wg)) * W + wave;
  if (local >= (ONE ? n1 : rg[fi].count)) return;
  const McArgs &a = fa[fi];
  const McTile t = ONE ? tiles1[local] : a.tiles[rg[fi].first + local];
  const int max = (1 << a.g.bit_depth) - 1;
  const bool compound = t.ref_frame[1] > INTRA_FRAME;
  P *dst = (P *)a.cur.p[t.plane];
  const uint32_t stride = a.cur.stride[t.plane];
  const RefPos rp0 = mc_position_mv(a.g, a.ref_w, a.ref_h, t.mv[0][0], t.mv[0][1], t.mi_row, t.mi_col, t.mi_size,
                                    t.ref_frame[0], t.plane, t.x, t.y);
  const RefPos rp1 = compound ? mc_position_mv(a.g, a.ref_w, a.ref_h, t.mv[1][0], t.mv[1][1], t.mi_row, t.mi_col,
                                               t.mi_size, t.ref_frame[1], t.plane, t.x, t.y)
                              : rp0;
  const int r0 = t.ref_frame[0] - LAST_FRAME, r1 = compound ? t.ref_frame[1] - LAST_FRAME : r0;
  const int filter = t.interp_filter;
  if (!HBD && ((t.tw | t.th | stride | (t.x + t.tx)) & 3) == 0 && rp0.xs == 16 && rp0.ys == 16 &&
      rp1.xs == 16 && rp1.ys == 16) {
    if (t.th > 16)
      mc_frame8<32>(a, t, rp0, rp1, r0, r1, compound, filter, (uint8_t *)dst, stride, sh);
    else
      mc_frame8<16>(a, t, rp0, rp1, r0, r1, compound, filter, (uint8_t *)dst, stride, sh);
    return;
  }
  if (HBD && ((t.tw | t.th | (t.x + t.tx)) & 3) == 0 && (stride & 1) == 0 && ((uintptr_t)dst & 3) == 0 &&
      rp0.xs == 16 && rp0.ys == 16 && rp1.xs == 16 && rp1.ys == 16) {
    if (t.th > 16)
      mc_frame16<32>(a, t, rp0, rp1, r0, r1, compound, filter, max, (uint16_t *)dst, stride, sh);
    else
      mc_frame16<16>(a, t, rp0, rp1, r0, r1, compound, filter, max, (uint16_t *)dst, stride, sh);
    return;
  }
  int out[2][4];
  mc_tile<P>((const P *)a.ref[r0].p[t.plane], a.ref[r0].stride[t.plane], rp0.last_x, rp0.last_y,
             rp0.x + rp0.xs * t.tx, rp0.y + rp0.ys * t.ty, rp0.xs, rp0.ys, t.tw, t.th, filter, max, sh.g,
             out[0]);
  if (compound)
    mc_tile<P>((const P *)a.ref[r1].p[t.plane], a.ref[r1].stride[t.plane], rp1.last_x, rp1.last_y,
               rp1.x + rp1.xs * t.tx, rp1.y + rp1.ys * t.ty, rp1.xs, rp1.ys, t.tw, t.th, filter, max,
               sh.g, out[1]);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int o = (threadIdx.x & 63) + 64 * i;
    if (o < t.th * t.tw) {
      const int r = o / t.tw, c = o - r * t.tw;
      const int v = compound ? (out[0][i] + out[1][i] + 1) >> 1 : out[0][i];
      dst[(size_t)(t.y + t.ty + r) * stride + t.x + t.tx + c] = (P)v;
    }
  }
}

template <bool HBD, bool ONE, int W>
__global__ __launch_bounds__(64 * W) void k_mc(const McTile *tiles1, uint32_t n1, const McArgs *fa, const Range *rg,
                                               uint32_t nf) {
  mc_wg<HBD, ONE, W>(tiles1, n1, fa, rg, nf, blockIdx.x, gridDim.x);
}

// The first launch of a single frame: the inter prediction tiles
// (workgroups below n_mc, a multiple of 8, so that the XCD order of the
// tiles holds), then the margin clear (n_clear workgroups) and the loop
// filter decisions (n_mask; a workgroup per superblock entry: measured at
// 1080p, the launch takes 16.7 us, 22.3 with a workgroup per superblock,
// 17.9 with the clear and the decisions dispatched first).  The clear writes samples outside the
// 8x8-aligned area, the prediction inside it, the decisions their own
// buffer; the prediction reads only the reference frames.
template <bool HBD, int K>
__global__ __launch_bounds__(64 * kMcWavesOne) void k_prep_mc(const ClearArgs *fc, const Range *rc, uint32_t n_clear,
                                                              const LfArgs *fl, const Range *rl, uint32_t n_mask,
                                                              uint32_t n_mc, const McTile *tiles1, uint32_t n1,
                                                              const McArgs *fa, const Range *rg) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  const uint32_t b = blockIdx.x;
  if (b < n_mc) {
    mc_wg<HBD, true, kMcWavesOne>(tiles1, n1, fa, rg, 1, b, n_mc);
  } else if (b - n_mc < n_clear) {
    clear_margins_wg<P, 64 * kMcWavesOne>(fc, rc, 1, b - n_mc);
  } else if (b - n_mc - n_clear < n_mask) {
    lf_mask_wg<K, 64 * kMcWavesOne>(fl, rl, 1, b - n_mc - n_clear);
  }
}

void launch_prep_mc(hipStream_t s, const ClearArgs *fc, const Range *rc, uint32_t n_clear, const LfArgs *fl,
                    const Range *rl, uint32_t n_mask, int per_sb, const McArgs *fa, const Range *rg,
                    const McTile *tiles, uint32_t n_tiles, bool hbd) {
  upload_filters();
  const uint32_t n_mc = ((n_tiles + kMcWavesOne - 1) / kMcWavesOne + 7) & ~7u;
  const uint32_t total = n_mc + n_clear + n_mask;
  if (!total) return;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(total), dim3(64 * kMcWavesOne), 0, s, fc, rc, n_clear, fl, rl, n_mask, n_mc,
                       tiles, n_tiles, fa, rg);
  };
  if (hbd)
    per_sb == 6 ? go(k_prep_mc<true, 6>) : go(k_prep_mc<true, 1>);
  else
    per_sb == 6 ? go(k_prep_mc<false, 6>) : go(k_prep_mc<false, 1>);
}

// Per-kernel seam: explicit positions (vp9r_mc_desc), tiles expanded on
// the host as (desc index, tile x, tile y).
struct DescTile {
  uint32_t desc;
  uint16_t tx, ty;
};

// Rows ty .. ty + th - 1 (th <= 16) of the tile at column tx of a
// descriptor through the general path, written at dst (the tile's first
// sample of those rows).
template <class P>
DEV void mc_desc_half(const vp9r_mc_desc &d, const vp9r_ref_plane &r0, const vp9r_ref_plane &r1, P *dst,
                      uint32_t out_stride, int tx, int ty, int tw, int th, int max, McShared &sh) {
  const bool compound = d.ref2_idx >= 0;
  int o0[4], o1[4];
  mc_tile<P>((const P *)r0.data, r0.stride, r0.last_x, r0.last_y, d.x + d.xs * tx, d.y + d.ys * ty, d.xs, d.ys, tw,
             th, d.filter, max, sh, o0);
  if (compound)
    mc_tile<P>((const P *)r1.data, r1.stride, r1.last_x, r1.last_y, d.x2 + d.xs * tx, d.y2 + d.ys * ty, d.xs, d.ys,
               tw, th, d.filter, max, sh, o1);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int o = (threadIdx.x & 63) + 64 * i;
    if (o < th * tw) {
      const int r = o / tw, c = o - r * tw;
      dst[(size_t)r * out_stride + c] = (P)(compound ? (o0[i] + o1[i] + 1) >> 1 : o0[i]);
    }
  }
}

// EXPAND: no tile list; tile 8 i + k is tile k of descriptor i (16 columns
// by 32 rows: column (k & 3), row half (k >> 2)), and tiles outside the
// block exit at once (their wave; no wave waits on another).  TH: the
// tile's rows, 16 or 32 (the autotuner's choice per bit depth class).
template <bool HBD, bool EXPAND, int TH, int W>
__global__ __launch_bounds__(64 * W) void k_mc_desc(const vp9r_mc_desc *descs, const DescTile *tiles,
                                                    uint32_t n_tiles, const vp9r_ref_plane *planes, void *out,
                                                    uint32_t out_stride, int bd) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  constexpr int NH = TH / 16;
  __shared__ McLds shw[W];
  const int wave = W == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  McLds &sh = shw[wave];
  DescTile t;
  // W consecutive tiles per workgroup, a wave each
  const uint32_t id = xcd_order(blockIdx.x, gridDim.x) * W + wave;
  if (id >= n_tiles) return;
  if (EXPAND) {
    t.desc = id >> 3;
    t.tx = (uint16_t)(16 * (id & 3));
    t.ty = (uint16_t)(32 * ((id >> 2) & 1));
    const vp9r_mc_desc dd = descs[t.desc];
    if (t.tx >= dd.w || t.ty >= dd.h) return;
  } else {
    t = tiles[id];
  }
  const vp9r_mc_desc d = descs[t.desc];
  // each reference its own plane (inter.rs:21-27 RefPlane)
  const vp9r_ref_plane r0 = planes[d.ref_idx], r1 = planes[d.ref2_idx >= 0 ? d.ref2_idx : d.ref_idx];
  const int max = (1 << bd) - 1;
  const int th = d.h - t.ty < TH ? d.h - t.ty : TH, tw = d.w - t.tx < TILE ? d.w - t.tx : TILE;
  const bool compound = d.ref2_idx >= 0;
  P *dst = (P *)out + d.out_off + (size_t)t.ty * out_stride + t.tx;
  if (!HBD && d.xs == 16 && d.ys == 16 && ((tw | th | out_stride | (uint32_t)(uintptr_t)dst) & 3) == 0) {
    uint32_t v[NH], v2[NH];
    mc_tile8<TH>((const uint8_t *)r0.data, r0.stride, r0.last_x, r0.last_y, d.x + 16 * t.tx, d.y + 16 * t.ty, tw,
                 th, d.filter, mc8_lds<TH>(sh), v);
    if (compound) {
      mc_tile8<TH>((const uint8_t *)r1.data, r1.stride, r1.last_x, r1.last_y, d.x2 + 16 * t.tx, d.y2 + 16 * t.ty,
                   tw, th, d.filter, mc8_lds<TH>(sh), v2);
#pragma unroll
      for (int h = 0; h < NH; h++) v[h] = avg4(v[h], v2[h]);
    }
    const int q = threadIdx.x & 3;
#pragma unroll
    for (int h = 0; h < NH; h++) {
      const int r = 16 * h + ((threadIdx.x & 63) >> 2);
      if (r < th && 4 * q < tw) gst((uint32_t *)(dst + (size_t)r * out_stride + 4 * q), v[h]);
    }
    return;
  }
  if (HBD && d.xs == 16 && d.ys == 16 && ((tw | th) & 3) == 0 && (out_stride & 1) == 0 &&
      ((uintptr_t)dst & 3) == 0) {
    uint32_t v[NH][2], v2[NH][2];
    mc_tile16<TH>((const uint16_t *)r0.data, r0.stride, r0.last_x, r0.last_y, d.x + 16 * t.tx, d.y + 16 * t.ty, tw,
                  th, d.filter, max, mc16_lds<TH>(sh), v);
    if (compound) {
      mc_tile16<TH>((const uint16_t *)r1.data, r1.stride, r1.last_x, r1.last_y, d.x2 + 16 * t.tx, d.y2 + 16 * t.ty,
                    tw, th, d.filter, max, mc16_lds<TH>(sh), v2);
#pragma unroll
      for (int h = 0; h < NH; h++) {
        v[h][0] = avg2(v[h][0], v2[h][0]);
        v[h][1] = avg2(v[h][1], v2[h][1]);
      }
    }
    const int q = threadIdx.x & 3;
#pragma unroll
    for (int h = 0; h < NH; h++) {
      const int r = 16 * h + ((threadIdx.x & 63) >> 2);
      if (r < th && 4 * q < tw) {
        uint32_t *o = (uint32_t *)(dst + (size_t)r * out_stride + 4 * q);
        gst(o, v[h][0]);
        gst(o + 1, v[h][1]);
      }
    }
    return;
  }
  // the general path: the tile's rows in halves of at most 16
  for (int hh = 0; hh < th; hh += TILE)
    mc_desc_half<P>(d, r0, r1, dst + (size_t)hh * out_stride, out_stride, t.tx, t.ty + hh, tw,
                    th - hh < TILE ? th - hh : TILE, max, sh.g);
}

void mc_prepare() { upload_filters(); }

void launch_mc(hipStream_t s, const McArgs *fa, const Range *r, uint32_t n, uint32_t wgs, bool hbd,
               const McTile *one_tiles, uint32_t one_count) {
  upload_filters();
  if (!wgs) return;
  auto go = [&](auto kernel, int waves) {
    hipLaunchKernelGGL(kernel, dim3(wgs), dim3(64 * waves), 0, s, one_tiles, one_count, fa, r, n);
  };
  if (one_tiles) {
    if (hbd)
      go(k_mc<true, true, kMcWavesOne>, kMcWavesOne);
    else
      go(k_mc<false, true, kMcWavesOne>, kMcWavesOne);
  } else {
    if (hbd)
      go(k_mc<true, false, 1>, 1);
    else
      go(k_mc<false, false, 1>, 1);
  }
}

void launch_mc_desc_tiles(hipStream_t s, const vp9r_mc_desc *desc, const void *tiles, uint32_t n_tiles,
                          const vp9r_ref_plane *planes, void *out, uint32_t out_stride, uint32_t bd, int tile_rows) {
  upload_filters();
  if (!n_tiles) return;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(n_tiles), dim3(64), 0, s, desc, (const DescTile *)tiles, n_tiles, planes,
                       out, out_stride, (int)bd);
  };
  if (bd > 8) {
    if (tile_rows == 16)
      go(k_mc_desc<true, false, 16, 1>);
    else
      go(k_mc_desc<true, false, 32, 1>);
  } else {
    if (tile_rows == 16)
      go(k_mc_desc<false, false, 16, 1>);
    else
      go(k_mc_desc<false, false, 32, 1>);
  }
}

// The same without a tile list: 8 tiles per descriptor (those of a 64 x 64
// block), those outside a smaller block exit.
void launch_mc_desc_expand(hipStream_t s, const vp9r_mc_desc *desc, uint32_t count, const vp9r_ref_plane *planes,
                           void *out, uint32_t out_stride, uint32_t bd) {
  upload_filters();
  if (!count) return;
  const uint32_t n_tiles = 8 * count;
  if (bd > 8)
    hipLaunchKernelGGL((k_mc_desc<true, true, 32, 1>), dim3(n_tiles), dim3(64), 0, s, desc, nullptr,
                       n_tiles, planes, out, out_stride, (int)bd);
  else
    hipLaunchKernelGGL((k_mc_desc<false, true, 32, 1>), dim3(n_tiles), dim3(64), 0, s, desc, nullptr,
                       n_tiles, planes, out, out_stride, (int)bd);
}

}  // namespace vp9g
