// The loop filter's unit levels two at a time (the levels after the last
// intra level of a frame).
//
// The unit wavefront of lf.hip runs one level per launch: unit U(r, c) =
// H(r, c) V(r, c + 1) of level r + c follows U(r, c - 1) and U(r - 1, c) of
// level r + c - 1.  A launch here covers levels L and L + 1.  Each item V of
// level L + 1 is a job: its workgroup stages the region of V and of its two
// predecessors (the left one, U(r, c - 1) or the lone vertical pass of
// superblock (r, 0), and the upper one, U(r - 1, c)) in LDS, filters the two
// predecessors on two waves (they touch disjoint samples), then V.  A
// predecessor is filtered by both jobs that need it, so the samples a
// launch writes must each have one writer: V's job writes V's samples, and a
// level-L unit's samples that no level-(L + 1) unit rewrites are written by
// the job where it is the left predecessor (else, in the last superblock
// column, by the job where it is the upper one).  Jobs read the frame as it
// was after level L - 1 while others write, so each launch reads one copy
// of the planes (in) and writes the other (out), and the launches alternate
// between the two.  The remaining workgroups copy from in to out the
// samples that no level of the launch writes and that out lacks (all of
// them in a frame's first such launch, else those that the previous launch
// changed), so that out holds the whole frame after level L + 1.
//
// Superblock (r, c) of a plane (SW x SH samples) is written by the unit
// levels r + c - 1 and r + c, and its bottom 8 rows also by level r + c + 1
// (U(r + 1, c)'s horizontal pass reaches 8 rows above its superblock).
#include <type_traits>

#include "dev_common.h"
#include "launch.h"
#include "lf_dev.h"

namespace vp9g {

// LDS of a job: plane rows from 8 above superblock row r - 1 to the end of
// row r (8 + 2 SH), plane columns of superblocks c - 1 .. c + 1 after 8
// margin columns (LDS column 8 + x holds plane column (c - 1) SW + x); an
// odd number of dwords per row.  lf_stream reads up to 11 samples past a
// line's end (never into a result): 4 slack rows below, and the
// vertical lines of superblock c + 1 read up to LDS column 203 (in the row
// at 8 bits, the next row's first two samples above).  Decisions of up to 8 passes (the predecessors' three
// each at most, V's two).
template <class T>
struct __align__(16) LfPairShared {
  static constexpr int BP = sizeof(T) == 1 ? 204 : 202, ROWS = 8 + 128 + 4;
  T buf[ROWS * BP];
  uint32_t dq[8][272];
};

// A rectangle of whole 16-byte vectors (16 samples at 8 bits, 8 above):
// plane rows [y0, y0 + rows), plane columns from x0, 1 << lq vectors per
// row (a superblock's width); `on` false for a rectangle the job lacks.
struct LfRect {
  int y0, rows, x0;
  bool on;
};

// Vectors k0 + t + NT i < kend (i < N) of rectangle `rc`, for thread t of
// NT: loads, then stores into LDS, and LDS back to a plane.  LDS offsets
// are dword-aligned (the pitch and the column offsets are multiples of 4
// bytes), so each vector goes in and out as four dwords.
template <class P, int NT, int N>
DEV void pair_load(const P *pl, uint32_t stride, const LfRect &rc, int lq, uint4 (&v)[N], int t, int k0, int kend) {
  constexpr int QB = 16 / (int)sizeof(P);
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int k = k0 + t + NT * i;
    if (rc.on && k < kend)
      v[i] = gld((const uint4 *)(pl + (size_t)(rc.y0 + (k >> lq)) * stride + rc.x0 + QB * (k & ((1 << lq) - 1))));
  }
}

template <class P, int NT, int N>
DEV void pair_to_lds(const LfRect &rc, int lq, const uint4 (&v)[N], P *buf, int ly0, int lx0, int t, int k0,
                     int kend) {
  constexpr int BP = LfPairShared<P>::BP, QB = 16 / (int)sizeof(P);
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int k = k0 + t + NT * i;
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
template <class P, int NT, int N>
DEV void pair_store(P *pl, uint32_t stride, const LfRect &rc, int lq, const P *buf, int ly0, int lx0, int t) {
  constexpr int BP = LfPairShared<P>::BP, QB = 16 / (int)sizeof(P);
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
DEV bool pair_pass(bool lone, int ur, int uc, int C, int k, int &sb, bool &vert) {
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
template <class P>
DEV void pair_unit(LfPairShared<P> &sh, int slot, bool lone, int ur, int uc, int r, int c, int C, int SW, int SH,
                   int lane, int bd) {
  constexpr int BP = LfPairShared<P>::BP;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int sb;
    bool vert;
    if (!pair_pass(lone, ur, uc, C, k, sb, vert)) break;
    const int rr = sb / C, cc = sb - rr * C;
    if (vert) {
      // lines: the superblock's rows; edges from 8 samples into the line
      if (lane < SH)
        lf_stream(sh.buf, ((rr - r + 1) * SH + 8 + lane) * BP + (cc - c + 1) * SW, 1, SW + 8, SW / 4,
                  sh.dq[slot + k], lane >> 2, bd);
    } else {
      if (lane < SW)
        lf_stream(sh.buf, (rr - r + 1) * SH * BP + 8 + (cc - c + 1) * SW + lane, BP, SH + 8, SH / 4,
                  sh.dq[slot + k], lane >> 2, bd);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

template <bool HBD>
__global__ __launch_bounds__(128) void k_lf_pair(uint8_t *i0, uint8_t *i1, uint8_t *i2, uint8_t *o0, uint8_t *o1,
                                                 uint8_t *o2, uint32_t s0, uint32_t s1, uint32_t s2, uint32_t *dq,
                                                 uint32_t geo, uint32_t L, uint32_t n_jobs, uint32_t copy_all) {
  typedef typename std::conditional<HBD, uint16_t, uint8_t>::type P;
  __shared__ LfPairShared<P> sh;
  const int C = (int)(geo & 0xfff), R = (int)((geo >> 12) & 0xfff), M = min(R, C);
  const int bit_depth = (int)(geo >> 26), bd = HBD ? bit_depth : 8;
  const int tid = (int)threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t b = blockIdx.x, n_copy = copy_all ? (uint32_t)(R * C) : (uint32_t)(3 * M);
  const int plane = b < 3 * n_jobs ? (int)(b / n_jobs) : (int)((b - 3 * n_jobs) / n_copy);
  const int sx = plane ? (int)((geo >> 24) & 1) : 0, sy = plane ? (int)((geo >> 25) & 1) : 0;
  const int SW = 64 >> sx, SH = 64 >> sy;
  const P *in = (const P *)(plane == 0 ? i0 : plane == 1 ? i1 : i2);
  P *out = (P *)(plane == 0 ? o0 : plane == 1 ? o1 : o2);
  const uint32_t stride = plane == 0 ? s0 : plane == 1 ? s1 : s2;
  const int Li = (int)L;
  if (b >= 3 * n_jobs) {
    // The parts of superblock (r, c), t = r + c, whose samples in `out` are
    // not those of the frame after level L + 1 and that no job writes.
    // copy_all (out holds nothing yet): every part that the levels L and
    // L + 1 leave alone, workgroup j taking superblock j.  Else out holds the
    // frame after level L - 3, and the parts to copy are those the levels
    // L - 2 and L - 1 wrote and L and L + 1 leave alone: the body (all but
    // the bottom 8 rows) for t = L - 2 and L - 1, the bottom rows for t =
    // L - 3 and L - 2 (L - 2 and L - 1 in the last superblock row), workgroup
    // j taking the superblock of row max(0, t - C + 1) + j % M on diagonal
    // t = L - 3 + j / M, if there is one.
    const int j = (int)((b - 3 * n_jobs) % n_copy);
    int r, c, t;
    bool body, bottom;
    if (copy_all) {
      r = j / C;
      c = j % C;
      t = r + c;
      body = !(t - 1 == Li || t - 1 == Li + 1 || t == Li || t == Li + 1);
      bottom = body && !(r + 1 < R && (t + 1 == Li || t + 1 == Li + 1));
    } else {
      t = Li - 3 + j / M;
      r = (t >= C ? t - C + 1 : 0) + j % M;
      if (t < 0 || r > min(R - 1, t)) return;
      c = t - r;
      body = t == Li - 2 || t == Li - 1;
      bottom = r + 1 < R ? t == Li - 3 || t == Li - 2 : body;
    }
    // rows of 16-byte vectors: qs per row
    constexpr int QB = 16 / (int)sizeof(P);  // samples per vector
    const int lq = sizeof(P) == 1 ? 2 - sx : 3 - sx, y0 = r * SH, x0 = c * SW;
    const int rows = (body ? SH - 8 : 0) + (bottom ? 8 : 0), first = body ? 0 : SH - 8;
    if (!rows) return;
    // at most 64 rows x 8 vectors: 4 slots of 128 threads
    uint4 v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = tid + 128 * i;
      if (k < (rows << lq))
        v[i] = gld((const uint4 *)(in + (size_t)(y0 + first + (k >> lq)) * stride + x0 + QB * (k & ((1 << lq) - 1))));
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = tid + 128 * i;
      if (k < (rows << lq))
        gst((uint4 *)(out + (size_t)(y0 + first + (k >> lq)) * stride + x0 + QB * (k & ((1 << lq) - 1))), v[i]);
    }
    return;
  }
  // the job: item `job` of level L + 1 (units (r, L + 1 - r), rows
  // ascending, then the lone vertical pass of superblock (L + 2, 0))
  const int job = (int)(b - (uint32_t)plane * n_jobs), t = Li + 1;
  const int r_lo = t >= C ? t - C + 1 : 0, n_units = min(R - 1, t) - r_lo + 1;
  const bool lone = job >= n_units;
  const int r = lone ? t + 1 : r_lo + job, c = lone ? 0 : t - r;
  const bool has_a = !lone && (c >= 1 || r >= 1);  // left predecessor (a unit, or the lone pass of (r, 0))
  const bool a_lone = has_a && c == 0;
  const bool has_b = !lone && r >= 1;  // upper predecessor U(r - 1, c)
  // the region's origin: LDS row 0 is plane row (r - 1) SH - 8, LDS column
  // 8 is plane column (c - 1) SW
  const int ly0 = (r - 1) * SH - 8, lx0 = (c - 1) * SW;
  const bool right = c + 1 < C;
  // the rectangles of the region: wave 1 loads B's (0 .. 2), wave 0 A's
  // (3 .. 5; for a lone job, its superblock), and each wave half of
  // superblock (r, c + 1), which only V reads
  const LfRect rc0 = {(r - 1) * SH - 8, 8, c * SW, has_b && r >= 2};          // above U(r - 1, c)
  const LfRect rc1 = {(r - 1) * SH, SH, c * SW, has_b};                      // superblock (r - 1, c)
  const LfRect rc2 = {(r - 1) * SH, SH, (c + 1) * SW, has_b && right};       // superblock (r - 1, c + 1)
  const LfRect rc3 = {r * SH - 8, 8, (c - 1) * SW, has_a && !a_lone && r >= 1};  // above U(r, c - 1)
  const LfRect rc4 = {r * SH, SH, (c - 1) * SW, has_a && !a_lone};           // superblock (r, c - 1)
  const LfRect rc5 = {r * SH, SH, c * SW, true};                             // superblock (r, c)
  const LfRect rc6 = {r * SH, SH, (c + 1) * SW, !lone && right};            // superblock (r, c + 1)
  const LfRect &q0 = wave ? rc0 : rc3, &q1 = wave ? rc1 : rc4, &q2 = wave ? rc2 : rc5;
  // the passes' decisions: predecessor A in slots 0..2, B in 3..5, V in 6..7
  const size_t n_sb = (size_t)R * C;
  const uint32_t *dv = dq + (size_t)(plane * 2 + 0) * n_sb * 256, *dh = dq + (size_t)(plane * 2 + 1) * n_sb * 256;
  const int a_r = r, a_c = a_lone ? 0 : c - 1;
  uint4 d[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    // slots 0..2: A's passes (wave 0), 3..5: B's (wave 1), 6..7: V's (wave 0)
    const int grp = k < 3 ? 0 : k < 6 ? 1 : 2, kk = k - 3 * grp;
    int sb = 0;
    bool vert = false;
    const bool on = grp == 0   ? has_a && pair_pass(a_lone, a_r, a_c, C, kk, sb, vert)
                    : grp == 1 ? has_b && pair_pass(false, r - 1, c, C, kk, sb, vert)
                               : pair_pass(lone, r, c, C, kk, sb, vert);
    d[k] = make_uint4(0, 0, 0, 0);
    if (on && wave == (grp == 1 ? 1 : 0)) d[k] = gld((const uint4 *)((vert ? dv : dh) + (size_t)sb * 256) + lane);
  }
  // 16-byte vectors: 1 << lq per superblock row; NS per lane for a
  // superblock, NS / 2 for half of one, one for 8 rows
  const int lq = (sizeof(P) == 1 ? 2 : 3) - sx, nsb = SH << lq, nb8 = 8 << lq, half = nsb / 2;
  constexpr int NS = sizeof(P) == 1 ? 4 : 8;
  uint4 v0[1], v1[NS], v2[NS], v6[NS / 2];
  pair_load<P, 64, 1>(in, stride, q0, lq, v0, lane, 0, nb8);
  pair_load<P, 64, NS>(in, stride, q1, lq, v1, lane, 0, nsb);
  pair_load<P, 64, NS>(in, stride, q2, lq, v2, lane, 0, nsb);
  pair_load<P, 64, NS / 2>(in, stride, rc6, lq, v6, lane, wave * half, wave * half + half);
  // every decision slot its wave owns, used or not
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (wave == (k >= 3 && k < 6 ? 1 : 0)) ((uint4 *)sh.dq[k])[lane] = d[k];
  pair_to_lds<P, 64, 1>(q0, lq, v0, sh.buf, ly0, lx0, lane, 0, nb8);
  pair_to_lds<P, 64, NS>(q1, lq, v1, sh.buf, ly0, lx0, lane, 0, nsb);
  pair_to_lds<P, 64, NS>(q2, lq, v2, sh.buf, ly0, lx0, lane, 0, nsb);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // Each predecessor on the wave that staged its region: A's passes read
  // rectangles 3 .. 5 only, B's 0 .. 2 (the reads past a line's end that
  // reach the other wave's rectangles never reach a result).  Then the
  // samples of the predecessor that this job writes (A's that no
  // level-(L + 1) unit rewrites: V is A's right successor; B's likewise
  // when B has no right successor, in the last superblock column), which V
  // leaves alone, and the wave's half of superblock (r, c + 1).
  if (wave == 0) {
    if (has_a) pair_unit<P>(sh, 0, a_lone, a_r, a_c, r, c, C, SW, SH, lane, bd);
    pair_store<P, 64, 1>(out, stride, LfRect{r * SH - 8, 8, (c - 1) * SW, has_a && !a_lone && r >= 1}, lq, sh.buf,
                         ly0, lx0, lane);
    pair_store<P, 64, NS>(out, stride, LfRect{r * SH, SH - (r + 1 < R ? 8 : 0), (c - 1) * SW, has_a && !a_lone}, lq,
                          sh.buf, ly0, lx0, lane);
  } else {
    if (has_b) pair_unit<P>(sh, 3, false, r - 1, c, r, c, C, SW, SH, lane, bd);
    pair_store<P, 64, 1>(out, stride, LfRect{(r - 1) * SH - 8, 8, c * SW, has_b && !right && r >= 2}, lq, sh.buf,
                         ly0, lx0, lane);
    pair_store<P, 64, NS>(out, stride, LfRect{(r - 1) * SH, SH - 8, c * SW, has_b && !right}, lq, sh.buf, ly0, lx0,
                          lane);
  }
  pair_to_lds<P, 64, NS / 2>(rc6, lq, v6, sh.buf, ly0, lx0, lane, wave * half, wave * half + half);
  __syncthreads();
  if (wave == 0) pair_unit<P>(sh, 6, lone, r, c, r, c, C, SW, SH, lane, bd);
  __syncthreads();
  // V's samples: the 8 rows above superblock (r, c), superblocks (r, c)
  // and (r, c + 1)
  pair_store<P, 128, 1>(out, stride, LfRect{r * SH - 8, 8, c * SW, !lone && r >= 1}, lq, sh.buf, ly0, lx0, tid);
  pair_store<P, 128, NS / 2>(out, stride, rc5, lq, sh.buf, ly0, lx0, tid);
  pair_store<P, 128, NS / 2>(out, stride, rc6, lq, sh.buf, ly0, lx0, tid);
}

// level L + 1's items: its units, and the lone vertical pass of superblock
// (L + 2, 0) when that row exists
static uint32_t lf_pair_jobs(uint32_t rows, uint32_t cols, uint32_t L) {
  const int R = (int)rows, C = (int)cols, t = (int)L + 1;
  const int r_lo = t >= C ? t - C + 1 : 0, n_units = std::min(R - 1, t) - r_lo + 1;
  return (uint32_t)(n_units + (t + 1 < R ? 1 : 0));
}

void launch_lf_pair(hipStream_t s, const LfOne &in, const Planes &out, bool copy_all, bool hbd) {
  const uint32_t C = in.geo & 0xfff, R = (in.geo >> 12) & 0xfff, L = in.t;
  const uint32_t n_jobs = lf_pair_jobs(R, C, L), wgs = 3 * n_jobs + 3 * (copy_all ? R * C : 3 * std::min(R, C));
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(wgs), dim3(128), 0, s, in.p[0], in.p[1], in.p[2], (uint8_t *)out.p[0],
                       (uint8_t *)out.p[1], (uint8_t *)out.p[2], in.stride[0], in.stride[1], in.stride[2], in.dq,
                       in.geo, L, n_jobs, (uint32_t)copy_all);
  };
  if (hbd)
    go(k_lf_pair<true>);
  else
    go(k_lf_pair<false>);
}

}  // namespace vp9g
