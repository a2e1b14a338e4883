// The loop filter's unit levels D at a time (D = 2 .. 5; the levels
// after the last intra level of a frame).
//
// The unit wavefront of lf.hip runs one level per launch: unit U(r, c) =
// H(r, c) V(r, c + 1) of level r + c follows U(r, c - 1) and U(r - 1, c) of
// level r + c - 1.  A launch here covers levels L .. L + D - 1.  Each item V
// = U(r, c) of the last level is a job, whose workgroup filters, in LDS, V's
// backward cone inside the launch: on superblock row r - k (wave k), the
// units U(r - k, c - D + 1 + k + m) of level L + m, m = 0 .. D - 1 - k (a
// unit of column -1 is the lone vertical pass V(r - k, 0), which no unit
// includes; columns below -1 hold nothing).  Step m runs level L + m on
// every wave that has it, each unit after its left neighbour (the same
// wave's previous step) and its upper one (wave k + 1's previous step):
// D steps with a workgroup barrier between them.  The lone item V(L + D, 0)
// of the last level is a job of its own.
//
// Units filtered by several jobs are filtered alike (the same inputs), and
// each sample is written by one job.  A unit U(r', c') of a level before the
// launch's last has final samples E = the 8 rows above its superblock and
// its superblock (without its bottom 8 rows when row r' + 1 exists, which
// the next level rewrites): those are written by the job of the unit's own
// superblock row (wave 0), or, when V is in the last superblock column, by
// every wave (no job of the rows between exists).  V writes all it touches:
// the 8 rows above superblock (r, c), superblocks (r, c) and (r, c + 1).
//
// Jobs read the frame as it was after level L - 1 while others write, so
// each launch reads one copy of the planes (in) and writes the other (out),
// and the launches alternate between the two.  The remaining workgroups
// copy from in to out the samples that no level of the launch writes and
// that out lacks: all of them in a frame's first such launch, else those
// that the previous launch, levels Lp .. L - 1, changed.  Superblock (r, c)
// of a plane (SW x SH samples) is written by the unit levels r + c - 1 and
// r + c, and its bottom 8 rows also by level r + c + 1 (U(r + 1, c)'s
// horizontal pass reaches 8 rows above its superblock).
#include <type_traits>

#include "dev_common.h"
#include "launch.h"
#include "lf_dev.h"

namespace vp9g {

// LDS of a job: plane rows from 8 above superblock row r - D + 1 to the
// end of row r (8 + D SH), plane columns of superblocks c - D + 1 .. c + 1
// after 8 margin columns (LDS column 8 + x holds plane column
// (c - D + 1) SW + x); an odd number of dwords per row.  lf_stream reads up
// to 11 samples past a line's end (never into a result): 4 slack rows
// below, and the vertical lines of superblock c + 1 read up to LDS column
// (D + 1) 64 + 11 (in the row at 8 bits, the next row's first samples
// above).  Decisions: two passes per cone unit (U(0, 0)'s third apart).
template <class T, int D>
struct __align__(16) LfConeShared {
  static constexpr int BP = sizeof(T) == 1 ? (D + 1) * 64 + 12 : (D + 1) * 64 + 10, ROWS = 8 + D * 64 + 4;
  static constexpr int UNITS = D * (D + 1) / 2, EXTRA = 2 * UNITS;
  T buf[ROWS * BP];
  // two passes per unit (slots 2 u, 2 u + 1), and the first vertical pass
  // of unit (0, 0), its third, in slot EXTRA; 256 entries and the 4 that
  // lf_stream reads ahead past the last run's
  uint32_t dq[2 * UNITS + 1][260];
};

// A rectangle of whole 16-byte vectors (16 samples at 8 bits, 8 above):
// plane rows [y0, y0 + rows), plane columns from x0, 1 << lq vectors per
// row (a superblock's width); `on` false for a rectangle the job lacks.
struct LfRect {
  int y0, rows, x0;
  bool on;
};

// Vectors t + NT i (i < N) of rectangle `rc`, for thread t of NT: loads,
// then stores into LDS, and LDS back to a plane.  LDS offsets are
// dword-aligned (the pitch and the column offsets are multiples of 4
// bytes), so each vector goes in and out as four dwords.
template <class P, int NT, int N>
DEV void cone_load(const P *pl, uint32_t stride, const LfRect &rc, int lq, uint4 (&v)[N], int t) {
  constexpr int QB = 16 / (int)sizeof(P);
  const int kend = rc.rows << lq;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int k = t + NT * i;
    if (rc.on && k < kend)
      v[i] = gld((const uint4 *)(pl + (size_t)(rc.y0 + (k >> lq)) * stride + rc.x0 + QB * (k & ((1 << lq) - 1))));
  }
}

template <class P, int D, int NT, int N>
DEV void cone_to_lds(const LfRect &rc, int lq, const uint4 (&v)[N], P *buf, int ly0, int lx0, int t) {
  constexpr int BP = LfConeShared<P, D>::BP, QB = 16 / (int)sizeof(P);
  const int kend = rc.rows << lq;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int k = t + NT * i;
    if (rc.on && k < kend) {
      uint32_t *l = (uint32_t *)(buf + (rc.y0 - ly0 + (k >> lq)) * BP + rc.x0 - lx0 + 8 + QB * (k & ((1 << lq) - 1)));
      l[0] = v[i].x;
      l[1] = v[i].y;
      l[2] = v[i].z;
      l[3] = v[i].w;
    }
  }
}

// Every LDS read is issued before the first store.
template <class P, int D, int NT, int N>
DEV void cone_store(P *pl, uint32_t stride, const LfRect &rc, int lq, const P *buf, int ly0, int lx0, int t) {
  constexpr int BP = LfConeShared<P, D>::BP, QB = 16 / (int)sizeof(P);
  uint4 v[N];
  const int kend = rc.rows << lq;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int k = t + NT * i;
    const bool ok = rc.on && k < kend;
    const uint32_t *l = (const uint32_t *)(buf + (ok ? (rc.y0 - ly0 + (k >> lq)) * BP + rc.x0 - lx0 + 8 +
                                                          QB * (k & ((1 << lq) - 1))
                                                    : 8));
    v[i] = make_uint4(l[0], l[1], l[2], l[3]);
  }
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int k = t + NT * i;
    if (rc.on && k < kend)
      gst((uint4 *)(pl + (size_t)(rc.y0 + (k >> lq)) * stride + rc.x0 + QB * (k & ((1 << lq) - 1))), v[i]);
  }
}

// Pass k of a unit (ur, uc) (lone: the vertical pass of superblock (ur, 0)
// alone), in the order the reference filters them: V(0, 0) first for unit
// (0, 0), then H(ur, uc), then V(ur, uc + 1) when uc + 1 < C.  False past
// the unit's passes; else its superblock and direction.
DEV bool cone_pass(bool lone, int ur, int uc, int C, int k, int &sb, bool &vert) {
  if (lone) {
    sb = ur * C;
    vert = true;
    return k == 0;
  }
  const int kk = k - (ur == 0 && uc == 0 ? 1 : 0);
  sb = kk < 0 ? 0 : ur * C + uc + kk;
  vert = kk != 0;
  return kk < 1 || (kk == 1 && uc + 1 < C);
}

// The passes of one unit on the job's LDS region, their decisions in
// sh.dq[slot ..]: one wave, whose LDS accesses stay in order between
// passes.  (r, c): the job's unit, which places the region.
template <class P, int D>
DEV void cone_unit(LfConeShared<P, D> &sh, int slot, bool lone, int ur, int uc, int r, int c, int C, int SW, int SH,
                   int lane, int bd) {
  constexpr int BP = LfConeShared<P, D>::BP;
  const bool three = !lone && ur == 0 && uc == 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int sb;
    bool vert;
    if (!cone_pass(lone, ur, uc, C, k, sb, vert)) break;
    const int rr = sb / C, cc = sb - rr * C;
    const int q = three ? (k == 0 ? LfConeShared<P, D>::EXTRA : slot + k - 1) : slot + k;
    if (vert) {
      // lines: the superblock's rows; edges from 8 samples into the line
      // (each line starts 4-byte aligned: the pitch and SW are multiples of
      // 4 bytes)
      if (lane < SH)
        lf_stream<P, true>(sh.buf, ((rr - r + D - 1) * SH + 8 + lane) * BP + (cc - c + D - 1) * SW, 1, SW + 8,
                           SW / 4, sh.dq[q], lane >> 2, bd);
    } else {
      if (lane < SW)
        lf_stream(sh.buf, (rr - r + D - 1) * SH * BP + 8 + (cc - c + D - 1) * SW + lane, BP, SH + 8, SH / 4,
                  sh.dq[q], lane >> 2, bd);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// The copy workgroups' part: parts of superblock (r, c), t = r + c, as set
// out above.  prev = 0: every part that levels L .. L + D - 1 leave alone;
// else the body (all but the bottom 8 rows) for t in [Lp, L - 1], the
// bottom rows for t in [Lp - 1, L - 2] (in the last superblock row, as the
// body).
DEV void cone_copy_parts(int r, int c, int R, int L, int D, int prev, bool &body, bool &bottom) {
  const int t = r + c;
  if (!prev) {
    body = t < L || t - 1 > L + D - 1;
    bottom = body && !(r + 1 < R && t + 1 >= L && t + 1 <= L + D - 1);
  } else {
    body = t >= L - prev && t <= L - 1;
    bottom = r + 1 < R ? t >= L - prev - 1 && t <= L - 2 : body;
  }
}

template <bool HBD, int D>
__global__ __launch_bounds__(64 * D) void k_lf_cone(uint8_t *i0, uint8_t *i1, uint8_t *i2, uint8_t *o0, uint8_t *o1,
                                                    uint8_t *o2, uint32_t s0, uint32_t s1, uint32_t s2, uint32_t *dq,
                                                    uint32_t geo, uint32_t L, uint32_t n_jobs, uint32_t prev) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  constexpr int NT = 64 * D, BP = LfConeShared<P, D>::BP;
  __shared__ LfConeShared<P, D> sh;
  const int C = (int)(geo & 0xfff), R = (int)((geo >> 12) & 0xfff), M = min(R, C);
  const int bit_depth = (int)(geo >> 26), bd = HBD ? bit_depth : 8;
  const int tid = (int)threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // copy workgroups per plane: 4 superblocks of a row each (first launch),
  // else one superblock of the prev + 1 diagonals Lp - 1 .. L - 1 each
  const int G = (C + 3) / 4;
  const uint32_t b = blockIdx.x, n_copy = prev ? (uint32_t)((prev + 1) * M) : (uint32_t)(R * G);
  const int plane = b < 3 * n_jobs ? (int)(b / n_jobs) : (int)((b - 3 * n_jobs) / n_copy);
  const int sx = plane ? (int)((geo >> 24) & 1) : 0, sy = plane ? (int)((geo >> 25) & 1) : 0;
  const int SW = 64 >> sx, SH = 64 >> sy;
  const P *in = (const P *)(plane == 0 ? i0 : plane == 1 ? i1 : i2);
  P *out = (P *)(plane == 0 ? o0 : plane == 1 ? o1 : o2);
  const uint32_t stride = plane == 0 ? s0 : plane == 1 ? s1 : s2;
  const int Li = (int)L, Pi = (int)prev;
  // 16-byte vectors: 1 << lq per superblock row
  const int lq = (sizeof(P) == 1 ? 2 : 3) - sx;
  constexpr int QB = 16 / (int)sizeof(P);
  if (b >= 3 * n_jobs) {
    const int j = (int)((b - 3 * n_jobs) % n_copy);
    constexpr int NC = (sizeof(P) == 1 ? 256 : 512) / NT + 1;  // vectors of a superblock per thread, rounded up
    uint4 v[4][NC];
    int y0[4], x0[4], rows[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, c;
      bool ok;
      if (!prev) {
        r = j / G;
        c = 4 * (j % G) + q;
        ok = c < C;
      } else {
        const int t = Li - Pi - 1 + j / M;
        r = (t >= C ? t - C + 1 : 0) + j % M;
        c = t - r;
        ok = q == 0 && t >= 0 && r <= min(R - 1, t);
      }
      bool body = false, bottom = false;
      if (ok) cone_copy_parts(r, c, R, Li, D, Pi, body, bottom);
      rows[q] = (body ? SH - 8 : 0) + (bottom ? 8 : 0);
      y0[q] = r * SH + (body ? 0 : SH - 8);
      x0[q] = c * SW;
#pragma unroll
      for (int i = 0; i < NC; i++) {
        const int k = tid + NT * i;
        if (k < (rows[q] << lq))
          v[q][i] = gld((const uint4 *)(in + (size_t)(y0[q] + (k >> lq)) * stride + x0[q] + QB * (k & ((1 << lq) - 1))));
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int i = 0; i < NC; i++) {
        const int k = tid + NT * i;
        if (k < (rows[q] << lq))
          gst((uint4 *)(out + (size_t)(y0[q] + (k >> lq)) * stride + x0[q] + QB * (k & ((1 << lq) - 1))), v[q][i]);
      }
    return;
  }
  // the job: item `job` of level L + D - 1 (units (r, t - r), rows
  // ascending, then the lone vertical pass of superblock (t + 1, 0))
  const int job = (int)(b - (uint32_t)plane * n_jobs), t = Li + D - 1;
  const int r_lo = t >= C ? t - C + 1 : 0, n_units = min(R - 1, t) - r_lo + 1;
  const bool lone = job >= n_units;
  const int r = lone ? t + 1 : r_lo + job, c = lone ? 0 : t - r;
  const size_t n_sb = (size_t)R * C;
  const uint32_t *dv = dq + (size_t)(plane * 2 + 0) * n_sb * 256, *dh = dq + (size_t)(plane * 2 + 1) * n_sb * 256;
  const int ly0 = (r - D + 1) * SH - 8, lx0 = (c - D + 1) * SW;
  constexpr int NS = (sizeof(P) == 1 ? 256 : 512) / 64;  // vectors of a luma superblock per lane
  if (lone) {
    // V(t + 1, 0) alone, on wave 0 (the others only take the barrier)
    const LfRect sb = {r * SH, SH, 0, wave == 0};
    uint4 d = make_uint4(0, 0, 0, 0), v[NS];
    if (wave == 0) d = gld((const uint4 *)(dv + (size_t)r * C * 256) + lane);
    cone_load<P, 64, NS>(in, stride, sb, lq, v, lane);
    if (wave == 0) ((uint4 *)sh.dq[0])[lane] = d;
    cone_to_lds<P, D, 64, NS>(sb, lq, v, sh.buf, ly0, lx0, lane);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (wave == 0) cone_unit<P, D>(sh, 0, true, r, 0, r, c, C, SW, SH, lane, bd);
    __syncthreads();
    cone_store<P, D, NT, NS / D + 1>(out, stride, LfRect{r * SH, SH, 0, true}, lq, sh.buf, ly0, lx0, tid);
    return;
  }
  // Wave k: superblock row rr = r - k, its units at columns cl + m (m = 0
  // .. D - 1 - k, cl = c - D + 1 + k).  It stages the 8 rows above its
  // leftmost unit and superblocks cl .. c + 1 of its row (the rows above
  // its other units are the next wave's).
  const int k = wave, rr = r - k, cl = c - D + 1 + k;
  const bool row_on = rr >= 0;
  const LfRect top = {rr * SH - 8, 8, cl * SW, row_on && rr >= 1 && cl >= 0};
  uint4 vt[1], vs[D + 1][NS];
  cone_load<P, 64, 1>(in, stride, top, lq, vt, lane);
#pragma unroll
  for (int q = 0; q <= D; q++) {
    const int cc = cl + q;
    cone_load<P, 64, NS>(in, stride, LfRect{rr * SH, SH, cc * SW, row_on && q <= D - k && cc >= 0 && cc < C}, lq,
                         vs[q], lane);
  }
  // the decisions of the wave's units: unit m in slots 2 (idx(k, m)) and
  // the next (U(0, 0)'s first vertical pass in slot EXTRA)
  const int base = 2 * (k * D - k * (k - 1) / 2);
  uint4 d[D][3];
#pragma unroll
  for (int m = 0; m < D; m++) {
    const int uc = cl + m;
    const bool unit = row_on && m <= D - 1 - k && uc >= 0, lone_u = row_on && m <= D - 1 - k && uc == -1 && rr >= 1;
#pragma unroll
    for (int p = 0; p < 3; p++) {
      int sb = 0;
      bool vert = false;
      d[m][p] = make_uint4(0, 0, 0, 0);
      if ((unit || lone_u) && cone_pass(lone_u, rr, lone_u ? 0 : uc, C, p, sb, vert))
        d[m][p] = gld((const uint4 *)((vert ? dv : dh) + (size_t)sb * 256) + lane);
    }
  }
#pragma unroll
  for (int m = 0; m < D; m++) {
    const bool three = rr == 0 && cl + m == 0;
#pragma unroll
    for (int p = 0; p < 3; p++) {
      // a unit's passes beyond its second exist only for U(0, 0)
      if (m > D - 1 - k || (p == 2 && !three)) continue;
      const int q = three ? (p == 0 ? LfConeShared<P, D>::EXTRA : base + 2 * m + p - 1) : base + 2 * m + p;
      ((uint4 *)sh.dq[q])[lane] = d[m][p];
    }
  }
  cone_to_lds<P, D, 64, 1>(top, lq, vt, sh.buf, ly0, lx0, lane);
#pragma unroll
  for (int q = 0; q <= D; q++) {
    const int cc = cl + q;
    cone_to_lds<P, D, 64, NS>(LfRect{rr * SH, SH, cc * SW, row_on && q <= D - k && cc >= 0 && cc < C}, lq, vs[q],
                              sh.buf, ly0, lx0, lane);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // Step m: level L + m, on the waves that have it.  A unit before the last
  // level stores its final samples at once, when this job is their writer.
  const bool writer = k == 0 || c == C - 1;
#pragma unroll 1
  for (int m = 0; m < D; m++) {
    const int uc = cl + m;
    if (row_on && m <= D - 1 - k && uc >= -1 && !(uc == -1 && rr < 1)) {
      cone_unit<P, D>(sh, base + 2 * m, uc == -1, rr, uc == -1 ? 0 : uc, r, c, C, SW, SH, lane, bd);
      if (m < D - 1 && uc >= 0 && writer) {
        cone_store<P, D, 64, 1>(out, stride, LfRect{rr * SH - 8, 8, uc * SW, rr >= 1}, lq, sh.buf, ly0, lx0, lane);
        cone_store<P, D, 64, NS>(out, stride, LfRect{rr * SH, SH - (rr + 1 < R ? 8 : 0), uc * SW, true}, lq, sh.buf,
                                 ly0, lx0, lane);
      }
    }
    __syncthreads();
  }
  // V's samples, by every wave: the 8 rows above superblock (r, c),
  // superblocks (r, c) and (r, c + 1)
  cone_store<P, D, NT, 1>(out, stride, LfRect{r * SH - 8, 8, c * SW, r >= 1}, lq, sh.buf, ly0, lx0, tid);
  cone_store<P, D, NT, NS / D + 1>(out, stride, LfRect{r * SH, SH, c * SW, true}, lq, sh.buf, ly0, lx0, tid);
  cone_store<P, D, NT, NS / D + 1>(out, stride, LfRect{r * SH, SH, (c + 1) * SW, c + 1 < C}, lq, sh.buf, ly0, lx0,
                                   tid);
}

// level t's items: its units, and the lone vertical pass of superblock
// (t + 1, 0) when that row exists
static uint32_t lf_level_items(uint32_t rows, uint32_t cols, uint32_t t_level) {
  const int R = (int)rows, C = (int)cols, t = (int)t_level;
  const int r_lo = t >= C ? t - C + 1 : 0, n_units = std::min(R - 1, t) - r_lo + 1;
  return (uint32_t)(n_units + (t + 1 < R ? 1 : 0));
}

void launch_lf_cone(hipStream_t s, const LfOne &in, const Planes &out, int depth, int prev, bool hbd) {
  const uint32_t C = in.geo & 0xfff, R = (in.geo >> 12) & 0xfff, L = in.t, M = std::min(R, C);
  const uint32_t n_jobs = lf_level_items(R, C, L + depth - 1);
  const uint32_t wgs = 3 * n_jobs + 3 * (prev ? (prev + 1) * M : R * ((C + 3) / 4));
  auto go = [&](auto kernel, int d) {
    hipLaunchKernelGGL(kernel, dim3(wgs), dim3(64 * d), 0, s, in.p[0], in.p[1], in.p[2], (uint8_t *)out.p[0],
                       (uint8_t *)out.p[1], (uint8_t *)out.p[2], in.stride[0], in.stride[1], in.stride[2], in.dq,
                       in.geo, L, n_jobs, (uint32_t)prev);
  };
  if (hbd) {
    if (depth == 2)
      go(k_lf_cone<true, 2>, 2);
    else
      go(k_lf_cone<true, 3>, 3);
  } else {
    if (depth == 2)
      go(k_lf_cone<false, 2>, 2);
    else if (depth == 3)
      go(k_lf_cone<false, 3>, 3);
    else if (depth == 4)
      go(k_lf_cone<false, 4>, 4);
    else
      go(k_lf_cone<false, 5>, 5);
  }
}

}  // namespace vp9g
