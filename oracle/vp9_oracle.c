/* CPU oracle of the VP9 reconstruction path.  TEST INFRASTRUCTURE ONLY: see
 * vp9_oracle.h.  File:line citations are the reference crate's
 * (safewords/rivet-vp9); section numbers are the VP9 specification's.
 *
 * Pinning (what checks each part; see tests/test_oracle.py):
 *   - orc_itx_add: against the 64-bit specification process at conforming
 *     input (the reference's itx.rs:1252-1301, 1344-1378 tests), the real
 *     DCT/sine transforms (itx.rs:1121-1205), the WHT DC spread, the DC-only
 *     shortcut over every conforming DC (itx.rs:1431-1450), and end to end by
 *     the 15 committed stream MD5s.  Non-conforming 8-bit wrapping and the
 *     +-65536 residual clamp when it fires: parity unpinned (restatement only).
 *   - orc_inter_predict: the integer-position copy and filter taps summing to
 *     128 (inter.rs:450-481, 566-573), and the MD5s (regular, smooth,
 *     bilinear, switchable, scaled references).
 *   - orc_intra_predict: known answers and constant edges (intra.rs:187-209),
 *     and the MD5s.
 *   - loop filter: flat-step monotone / strong edge untouched (lf.rs:570-593),
 *     and the MD5s for the decisions.  Sharpness > 0: parity unpinned.
 *   - frame process: the MD5s only.
 *   - orc_itx_add_batch, orc_mc_batch: equal to the per-block calls above
 *     (tests/test_microbench.py). */
#include "vp9_oracle.h"

#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>

#include "vp9_consts.h"
#include "vp9_tables.h"

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static inline int64_t lclamp(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : v > hi ? hi : v; }

/* ======================================================================== */
/* Inverse transforms (8.7; src/dsp/itx.rs).                                */

/* cos64_lookup (8.7.1.1) */
static const int32_t COS64[33] = {16384, 16364, 16305, 16207, 16069, 15893, 15679, 15426, 15137,
                                  14811, 14449, 14053, 13623, 13160, 12665, 12140, 11585, 11003,
                                  10394, 9760,  9102,  8423,  7723,  7005,  6270,  5520,  4756,
                                  3981,  3196,  2404,  1606,  804,   0};
enum { SINPI_1_9 = 5283, SINPI_2_9 = 9929, SINPI_3_9 = 13377, SINPI_4_9 = 15212 };

static int32_t cos64(int angle) {
  int a = angle & 127;
  if (a <= 32) return COS64[a];
  if (a <= 64) return -COS64[64 - a];
  if (a <= 96) return -COS64[a - 64];
  return COS64[128 - a];
}
static int32_t sin64(int angle) { return cos64(angle - 32); }

static int brev(int bits, int x) {
  int t = 0;
  for (int i = 0; i < bits; i++) t |= ((x >> i) & 1) << (bits - 1 - i);
  return t;
}

/* Lane arithmetic.  w32 = 1: the 8-bit lane S32 (itx.rs:573-614), every
 * operation wrapping at 32 bits.  w32 = 0: the high-bit-depth lane S64
 * (itx.rs:626-666), 64-bit, whose multiply takes the low 32 bits of its
 * operand sign-extended. */
static inline int64_t L_add(int w32, int64_t a, int64_t b) {
  return w32 == 1 ? (int64_t)(int32_t)((uint32_t)a + (uint32_t)b) : (int64_t)((uint64_t)a + (uint64_t)b);
}
static inline int64_t L_sub(int w32, int64_t a, int64_t b) {
  return w32 == 1 ? (int64_t)(int32_t)((uint32_t)a - (uint32_t)b) : (int64_t)((uint64_t)a - (uint64_t)b);
}
/* w32 = 2: the specification's plain 64-bit product (for the restatement
 * that pins the lanes; the other operations are those of S64). */
static inline int64_t L_mul(int w32, int64_t a, int32_t c) {
  if (w32 == 2) return a * c;
  return w32 ? (int64_t)(int32_t)((uint32_t)a * (uint32_t)c)
             : (int64_t)((uint64_t)(int64_t)(int32_t)a * (uint64_t)(int64_t)c);
}
static inline int64_t L_rnd(int w32, int64_t a, int n) {
  return w32 == 1 ? (int64_t)((int32_t)((uint32_t)a + (1u << (n - 1))) >> n)
             : (int64_t)((uint64_t)a + ((uint64_t)1 << (n - 1))) >> n;
}

/* B(a, b, angle, flip) */
static inline void B(int w32, int64_t *t, int a, int b, int angle, int flip) {
  int32_t c = cos64(angle), s = sin64(angle);
  int64_t x = L_rnd(w32, L_sub(w32, L_mul(w32, t[a], c), L_mul(w32, t[b], s)), 14);
  int64_t y = L_rnd(w32, L_add(w32, L_mul(w32, t[a], s), L_mul(w32, t[b], c)), 14);
  t[a] = flip ? y : x;
  t[b] = flip ? x : y;
}
/* H(a, b, flip) */
static inline void H(int w32, int64_t *t, int a, int b, int flip) {
  if (flip) {
    int k = a;
    a = b;
    b = k;
  }
  int64_t x = t[a], y = t[b];
  t[a] = L_add(w32, x, y);
  t[b] = L_sub(w32, x, y);
}
/* SB(a, b, angle, flip): products kept unrounded in s */
static inline void SB(int w32, const int64_t *t, int64_t *s, int a, int b, int angle, int flip) {
  int32_t c = cos64(angle), sn = sin64(angle);
  int64_t x = L_sub(w32, L_mul(w32, t[a], c), L_mul(w32, t[b], sn));
  int64_t y = L_add(w32, L_mul(w32, t[a], sn), L_mul(w32, t[b], c));
  s[a] = flip ? y : x;
  s[b] = flip ? x : y;
}
/* SH(a, b) */
static inline void SH(int w32, int64_t *t, const int64_t *s, int a, int b) {
  t[a] = L_rnd(w32, L_add(w32, s[a], s[b]), 14);
  t[b] = L_rnd(w32, L_sub(w32, s[a], s[b]), 14);
}

/* The steps of the inverse DCT process (8.7.1.3) for size 2^n, after the
 * recursive half. */
static inline void idct_stage(int w32, int64_t *t, int n) {
  int n0 = 1 << n, n1 = n0 >> 1, n2 = n0 >> 2, n3 = n0 >> 3;
  for (int i = 0; i < n2; i++) B(w32, t, n1 + i, n0 - 1 - i, 32 - brev(5, n1 + i), 0);
  if (n >= 3)
    for (int i = 0; i < n3; i++)
      for (int j = 0; j < 2; j++) H(w32, t, n1 + 4 * i + 2 * j, n1 + 1 + 4 * i + 2 * j, j);
  if (n == 5) {
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        B(w32, t, n0 - n + 3 - n2 * j - 4 * i, n1 + n - 4 + n2 * j + 4 * i, 28 - 16 * i + 56 * j, 1);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++) H(w32, t, n1 + n3 * j + i, n1 + n2 - 5 + n3 * j - i, j & 1);
  }
  if (n >= 4) {
    for (int i = 0; i <= (n == 5 ? 1 : 0); i++)
      for (int j = 0; j < 2; j++)
        B(w32, t, n0 - n + 2 - i - n2 * j, n1 + n - 3 + i + n2 * j, 24 + 48 * j, 1);
    for (int i = 0; i < 2 * n - 6; i++)
      for (int j = 0; j < 2; j++) H(w32, t, n1 + n2 * j + i, n1 + n2 - 1 + n2 * j - i, j & 1);
  }
  if (n >= 3)
    for (int i = 0; i < n3; i++) B(w32, t, n0 - n3 - 1 - i, n1 + n3 + i, 16, 1);
  for (int i = 0; i < n1; i++) H(w32, t, i, n0 - 1 - i, 0);
}

/* inverse DCT array permutation (8.7.1.2) then the inverse DCT */
static inline void idct(int w32, int64_t *t, int n) {
  int n0 = 1 << n;
  int64_t c[32];
  memcpy(c, t, n0 * sizeof(int64_t));
  for (int i = 0; i < n0; i++) t[i] = c[brev(n, i)];
  B(w32, t, 0, 1, 16, 1);
  for (int k = 2; k <= n; k++) idct_stage(w32, t, k);
}

static inline void adst_in_perm(int64_t *t, int n) {
  int n0 = 1 << n, n1 = n0 >> 1;
  int64_t c[16];
  memcpy(c, t, n0 * sizeof(int64_t));
  for (int i = 0; i < n1; i++) {
    t[2 * i] = c[n0 - 1 - 2 * i];
    t[2 * i + 1] = c[2 * i];
  }
}

static inline void adst_out_perm(int64_t *t, int n) {
  int64_t c[16];
  memcpy(c, t, (1 << n) * sizeof(int64_t));
  if (n == 4) {
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int cc = 0; cc < 2; cc++)
          for (int d = 0; d < 2; d++)
            t[8 * a + 4 * b + 2 * cc + d] = c[8 * (d ^ cc) + 4 * (cc ^ b) + 2 * (b ^ a) + a];
  } else {
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int cc = 0; cc < 2; cc++) t[4 * a + 2 * b + cc] = c[4 * (cc ^ b) + 2 * (b ^ a) + a];
  }
}

/* inverse ADST4 (8.7.1.5) */
static inline void iadst4(int w32, int64_t *t) {
  int64_t s0 = L_mul(w32, t[0], SINPI_1_9), s1 = L_mul(w32, t[0], SINPI_2_9);
  int64_t s2 = L_mul(w32, t[1], SINPI_3_9), s3 = L_mul(w32, t[2], SINPI_4_9);
  int64_t s4 = L_mul(w32, t[2], SINPI_1_9), s5 = L_mul(w32, t[3], SINPI_2_9);
  int64_t s6 = L_mul(w32, t[3], SINPI_4_9);
  int64_t v = L_add(w32, L_sub(w32, t[0], t[2]), t[3]);
  int64_t s7 = L_mul(w32, v, SINPI_3_9);
  int64_t x0 = L_add(w32, L_add(w32, s0, s3), s5);
  int64_t x1 = L_sub(w32, L_sub(w32, s1, s4), s6);
  int64_t x2 = s7, x3 = s2;
  t[0] = L_rnd(w32, L_add(w32, x0, x3), 14);
  t[1] = L_rnd(w32, L_add(w32, x1, x3), 14);
  t[2] = L_rnd(w32, x2, 14);
  t[3] = L_rnd(w32, L_sub(w32, L_add(w32, x0, x1), x3), 14);
}

/* inverse ADST8 (8.7.1.6) */
static inline void iadst8(int w32, int64_t *t) {
  int64_t s[8] = {0};
  adst_in_perm(t, 3);
  for (int i = 0; i < 4; i++) SB(w32, t, s, 2 * i, 1 + 2 * i, 30 - 8 * i, 1);
  for (int i = 0; i < 4; i++) SH(w32, t, s, i, 4 + i);
  for (int i = 0; i < 2; i++) SB(w32, t, s, 4 + 3 * i, 5 + i, 24 - 16 * i, 1);
  for (int i = 0; i < 2; i++) SH(w32, t, s, 4 + i, 6 + i);
  for (int i = 0; i < 2; i++) H(w32, t, i, 2 + i, 0);
  for (int i = 0; i < 2; i++) B(w32, t, 2 + 4 * i, 3 + 4 * i, 16, 1);
  adst_out_perm(t, 3);
  for (int i = 0; i < 4; i++) t[1 + 2 * i] = L_sub(w32, 0, t[1 + 2 * i]);
}

/* inverse ADST16 (8.7.1.7) */
static inline void iadst16(int w32, int64_t *t) {
  int64_t s[16] = {0};
  adst_in_perm(t, 4);
  for (int i = 0; i < 8; i++) SB(w32, t, s, 2 * i, 1 + 2 * i, 31 - 4 * i, 1);
  for (int i = 0; i < 8; i++) SH(w32, t, s, i, 8 + i);
  for (int i = 0; i < 4; i++) SB(w32, t, s, 8 + 2 * i, 9 + 2 * i, 28 - 16 * i, 1);
  for (int i = 0; i < 4; i++) SH(w32, t, s, 8 + i, 12 + i);
  for (int i = 0; i < 4; i++) H(w32, t, i, 4 + i, 0);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) SB(w32, t, s, 4 + 8 * i + 3 * j, 5 + 8 * i + j, 24 - 16 * j, 1);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) SH(w32, t, s, 4 + 8 * j + i, 6 + 8 * j + i);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) H(w32, t, 8 * j + i, 2 + 8 * j + i, 0);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) B(w32, t, 2 + 4 * j + 8 * i, 3 + 4 * j + 8 * i, 48 + 64 * (i ^ j), 0);
  adst_out_perm(t, 4);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) t[1 + 12 * j + 2 * i] = L_sub(w32, 0, t[1 + 12 * j + 2 * i]);
}

/* the inverse ADST process (8.7.1.9); a 32-point request falls to the
 * 16-point transform of the first half (itx.rs:463-469) */
static inline void iadst(int w32, int64_t *t, int n) {
  if (n == 2)
    iadst4(w32, t);
  else if (n == 3)
    iadst8(w32, t);
  else
    iadst16(w32, t);
}

/* inverse WHT (8.7.1.10) */
static inline void iwht(int w32, int64_t *t, int shift) {
  int64_t a = t[0] >> shift, c = t[1] >> shift, d = t[2] >> shift, b = t[3] >> shift;
  a = L_add(w32, a, c);
  d = L_sub(w32, d, b);
  int64_t e = L_sub(w32, a, d) >> 1;
  b = L_sub(w32, e, b);
  c = L_sub(w32, e, c);
  a = L_sub(w32, a, b);
  d = L_add(w32, d, c);
  t[0] = a;
  t[1] = b;
  t[2] = c;
  t[3] = d;
}

static inline void tx1d(int w32, int64_t *t, int n, int dct, int lossless, int wht_shift) {
  if (lossless)
    iwht(w32, t, wht_shift);
  else if (dct)
    idct(w32, t, n);
  else
    iadst(w32, t, n);
}

/* inverse_2d_add (itx.rs:519-570) with one lane: rows (WHT shift 2), no
 * intermediate rounding, columns (WHT shift 0), Round2 by min(n + 2, 6)
 * unless lossless, residual clamped to +-2^16, added and clamped. */
static inline void inverse_2d_add(int w32, const int32_t *coefs, int n, int tx_type, int lossless,
                                  uint16_t *dst, size_t stride, int max) {
  int n0 = 1 << n;
  int row_dct = tx_type == DCT_DCT || tx_type == ADST_DCT;
  int col_dct = tx_type == DCT_DCT || tx_type == DCT_ADST;
  static __thread int64_t tmp[1024];
  int64_t t[32];
  for (int i = 0; i < n0; i++) {
    const int32_t *row = coefs + i * n0;
    int zero = 1;
    for (int j = 0; j < n0; j++) zero &= row[j] == 0;
    if (zero && !lossless) {
      for (int j = 0; j < n0; j++) tmp[i * n0 + j] = 0;
      continue;
    }
    for (int j = 0; j < n0; j++) t[j] = row[j];
    tx1d(w32, t, n, row_dct, lossless, 2);
    for (int j = 0; j < n0; j++) tmp[i * n0 + j] = t[j];
  }
  int shift = imin(n + 2, 6);
  for (int j = 0; j < n0; j++) {
    for (int i = 0; i < n0; i++) t[i] = tmp[i * n0 + j];
    tx1d(w32, t, n, col_dct, lossless, 0);
    for (int i = 0; i < n0; i++) {
      int64_t r = lossless ? t[i] : L_rnd(w32, t[i], shift);
      r = lclamp(r, -(1 << 16), 1 << 16);
      uint16_t *p = dst + (size_t)i * stride + j;
      *p = (uint16_t)lclamp((int64_t)*p + r, 0, max);
    }
  }
}

__attribute__((flatten)) static void itx32(const int32_t *c, int n, int tt, int ll, uint16_t *d,
                                           size_t s) {
  inverse_2d_add(1, c, n, tt, ll, d, s, 255);
}
__attribute__((flatten)) static void itx64(const int32_t *c, int n, int tt, int ll, uint16_t *d,
                                           size_t s, int max) {
  inverse_2d_add(0, c, n, tt, ll, d, s, max);
}

/* dc_only_value (itx.rs:689-694): the constant a DC-only DCT_DCT block adds */
static int32_t dc_only_value(int32_t dc, int n) {
  int32_t c = cos64(16);
  int64_t v = L_rnd(1, L_mul(1, dc, c), 14);
  int64_t w = L_rnd(1, L_mul(1, v, c), 14);
  return (int32_t)L_rnd(1, w, imin(n + 2, 6));
}

void orc_add_const(uint16_t *dst, size_t stride, int n, int v) {
  v = iclamp(v, -(1 << 16), 1 << 16);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      uint16_t *p = dst + (size_t)i * stride + j;
      *p = (uint16_t)iclamp((int)*p + v, 0, 255);
    }
}

void orc_itx_add(const int32_t *coefs, int n, int tx_type, int lossless, int eob, int bd,
                 uint16_t *dst, size_t stride) {
  if (bd > 8) {
    itx64(coefs, n, tx_type, lossless, dst, stride, (1 << bd) - 1);
    return;
  }
  if (eob == 1 && tx_type == DCT_DCT && !lossless) {
    orc_add_const(dst, stride, 1 << n, dc_only_value(coefs[0], n));
    return;
  }
  itx32(coefs, n, tx_type, lossless, dst, stride);
}

/* The 64-bit specification process (the reference's test module spec,
 * itx.rs:776-1112): plain 64-bit arithmetic, no truncation. */
static int64_t sp_r14(int64_t x) { return (x + (1 << 13)) >> 14; }
static void sp_B(int64_t *t, int a, int b, int ang, int flip) {
  int64_t x = t[a] * cos64(ang) - t[b] * sin64(ang), y = t[a] * sin64(ang) + t[b] * cos64(ang);
  t[a] = sp_r14(flip ? y : x);
  t[b] = sp_r14(flip ? x : y);
}
static void sp_H(int64_t *t, int a, int b, int flip) {
  if (flip) {
    int k = a;
    a = b;
    b = k;
  }
  int64_t x = t[a], y = t[b];
  t[a] = x + y;
  t[b] = x - y;
}
static void sp_idct_core(int64_t *t, int n) {
  int n0 = 1 << n, n1 = n0 >> 1, n2 = n0 >> 2, n3 = n0 >> 3;
  if (n == 2)
    sp_B(t, 0, 1, 16, 1);
  else
    sp_idct_core(t, n - 1);
  for (int i = 0; i < n2; i++) sp_B(t, n1 + i, n0 - 1 - i, 32 - brev(5, n1 + i), 0);
  if (n >= 3)
    for (int i = 0; i < n3; i++)
      for (int j = 0; j < 2; j++) sp_H(t, n1 + 4 * i + 2 * j, n1 + 1 + 4 * i + 2 * j, j);
  if (n == 5) {
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        sp_B(t, n0 - n + 3 - n2 * j - 4 * i, n1 + n - 4 + n2 * j + 4 * i, 28 - 16 * i + 56 * j, 1);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++) sp_H(t, n1 + n3 * j + i, n1 + n2 - 5 + n3 * j - i, j & 1);
  }
  if (n >= 4) {
    for (int i = 0; i <= (n == 5 ? 1 : 0); i++)
      for (int j = 0; j < 2; j++) sp_B(t, n0 - n + 2 - i - n2 * j, n1 + n - 3 + i + n2 * j, 24 + 48 * j, 1);
    for (int i = 0; i < 2 * n - 6; i++)
      for (int j = 0; j < 2; j++) sp_H(t, n1 + n2 * j + i, n1 + n2 - 1 + n2 * j - i, j & 1);
  }
  if (n >= 3)
    for (int i = 0; i < n3; i++) sp_B(t, n0 - n3 - 1 - i, n1 + n3 + i, 16, 1);
  for (int i = 0; i < n1; i++) sp_H(t, i, n0 - 1 - i, 0);
}
void orc_spec_idct(int64_t *t, int n) {
  int n0 = 1 << n;
  int64_t c[32];
  memcpy(c, t, n0 * sizeof(int64_t));
  for (int i = 0; i < n0; i++) t[i] = c[brev(n, i)];
  sp_idct_core(t, n);
}
void orc_spec_iadst(int64_t *t, int n) {
  /* the 64-bit lane has no truncation for values within 32 bits; here plain
   * 64-bit arithmetic is used throughout */
  if (n == 2) {
    int64_t s0 = SINPI_1_9 * t[0], s1 = SINPI_2_9 * t[0], s2 = SINPI_3_9 * t[1];
    int64_t s3 = SINPI_4_9 * t[2], s4 = SINPI_1_9 * t[2], s5 = SINPI_2_9 * t[3];
    int64_t s6 = SINPI_4_9 * t[3], s7 = SINPI_3_9 * (t[0] - t[2] + t[3]);
    int64_t x0 = s0 + s3 + s5, x1 = s1 - s4 - s6, x2 = s7, x3 = s2;
    t[0] = sp_r14(x0 + x3);
    t[1] = sp_r14(x1 + x3);
    t[2] = sp_r14(x2);
    t[3] = sp_r14(x0 + x1 - x3);
    return;
  }
  if (n == 3)
    iadst8(2, t);
  else
    iadst16(2, t);
}
void orc_itx_spec_reconstruct(const int32_t *block, int n, int tx_type, int lossless,
                              uint16_t *pred, int max) {
  int n0 = 1 << n;
  int64_t t[32], tmp[1024];
  for (int i = 0; i < n0; i++) {
    int zero = 1;
    for (int j = 0; j < n0; j++) zero &= block[i * n0 + j] == 0;
    if (zero && !lossless) {
      for (int j = 0; j < n0; j++) tmp[i * n0 + j] = 0;
      continue;
    }
    for (int j = 0; j < n0; j++) t[j] = block[i * n0 + j];
    if (lossless)
      iwht(0, t, 2);
    else if (tx_type == DCT_DCT || tx_type == ADST_DCT)
      orc_spec_idct(t, n);
    else
      orc_spec_iadst(t, n);
    for (int j = 0; j < n0; j++) tmp[i * n0 + j] = t[j];
  }
  int shift = imin(n + 2, 6);
  for (int j = 0; j < n0; j++) {
    for (int i = 0; i < n0; i++) t[i] = tmp[i * n0 + j];
    if (lossless)
      iwht(0, t, 0);
    else if (tx_type == DCT_DCT || tx_type == DCT_ADST)
      orc_spec_idct(t, n);
    else
      orc_spec_iadst(t, n);
    for (int i = 0; i < n0; i++) {
      int64_t r = lossless ? t[i] : (t[i] + (1 << (shift - 1))) >> shift;
      uint16_t *p = pred + i * n0 + j;
      *p = (uint16_t)lclamp((int64_t)*p + r, 0, max);
    }
  }
}

/* ======================================================================== */
/* Inter prediction (8.5.2.4; src/dsp/inter.rs).                            */

/* The specification's clamped-coordinate formula, which the reference's
 * predict() equals for every step, size and position (inter.rs:406-448,
 * 486-564): a horizontal pass over ih rows of the reference (rounded and
 * clamped), then a vertical pass (rounded and clamped). */
void orc_inter_predict(const uint16_t *ref, size_t ref_stride, int last_x, int last_y, int x, int y,
                       int x_step, int y_step, int w, int h, int filter, int bd, uint16_t *out,
                       size_t out_stride) {
  int max = (1 << bd) - 1;
  int ih = ((((h - 1) * y_step + 15) >> 4) + 8);
  int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * ih * w);
  const int32_t(*F)[8] = VP9_SUBPEL_FILTERS[filter];
  for (int r = 0; r < ih; r++) {
    const uint16_t *row = ref + (size_t)iclamp((y >> 4) + r - 3, 0, last_y) * ref_stride;
    for (int c = 0; c < w; c++) {
      int p = x + x_step * c;
      const int32_t *f = F[p & 15];
      int32_t s = 0;
      for (int t = 0; t < 8; t++) s += f[t] * row[iclamp((p >> 4) + t - 3, 0, last_x)];
      tmp[r * w + c] = iclamp((s + 64) >> 7, 0, max);
    }
  }
  for (int r = 0; r < h; r++) {
    int p = (y & 15) + y_step * r;
    const int32_t *f = F[p & 15];
    const int32_t *base = tmp + (p >> 4) * w;
    for (int c = 0; c < w; c++) {
      int32_t s = 0;
      for (int t = 0; t < 8; t++) s += f[t] * base[t * w + c];
      out[(size_t)r * out_stride + c] = (uint16_t)iclamp((s + 64) >> 7, 0, max);
    }
  }
  free(tmp);
}

void orc_avg(const uint16_t *a, const uint16_t *b, int w, int h, uint16_t *dst, size_t stride) {
  for (int i = 0; i < h; i++)
    for (int j = 0; j < w; j++)
      dst[(size_t)i * stride + j] = (uint16_t)(((uint32_t)a[i * w + j] + b[i * w + j] + 1) >> 1);
}

void orc_itx_add_batch(uint32_t count, const int32_t *coefs, const uint32_t *coef_off, const uint8_t *tx,
                       const uint8_t *tx_type, const uint16_t *eob, const uint64_t *dst_off, uint16_t *plane,
                       size_t stride, int bd, int lossless) {
  for (uint32_t i = 0; i < count; i++)
    orc_itx_add(coefs + coef_off[i], 2 + tx[i], tx_type[i], lossless, eob[i], bd, plane + dst_off[i], stride);
}

void orc_mc_batch(uint32_t count, const orc_mc_desc *desc, const uint16_t *const *refs, size_t ref_stride,
                  int last_x, int last_y, uint16_t *out, size_t out_stride, int bd) {
  uint16_t a[64 * 64], b[64 * 64];
  for (uint32_t i = 0; i < count; i++) {
    const orc_mc_desc *d = &desc[i];
    uint16_t *o = out + d->out_off;
    if (d->ref2_idx < 0) {
      orc_inter_predict(refs[d->ref_idx], ref_stride, last_x, last_y, d->x, d->y, d->xs, d->ys, d->w, d->h,
                        d->filter, bd, o, out_stride);
      continue;
    }
    orc_inter_predict(refs[d->ref_idx], ref_stride, last_x, last_y, d->x, d->y, d->xs, d->ys, d->w, d->h,
                      d->filter, bd, a, d->w);
    orc_inter_predict(refs[d->ref2_idx], ref_stride, last_x, last_y, d->x2, d->y2, d->xs, d->ys, d->w, d->h,
                      d->filter, bd, b, d->w);
    orc_avg(a, b, d->w, d->h, o, out_stride);
  }
}

/* ======================================================================== */
/* Intra prediction (8.5.1; src/dsp/intra.rs:19-175).                        */

static inline int r2(int x, int n) { return (x + (1 << (n - 1))) >> n; }

void orc_intra_predict(int mode, int log2, const int32_t *above, const int32_t *left, int have_left,
                       int have_above, int bd, uint16_t *dst, size_t stride) {
  int size = 1 << log2;
#define A(i) above[(i) + 1]
#define L(i) left[i]
#define D(i, j) dst[(size_t)(i) * stride + (j)]
  switch (mode) {
    case V_PRED:
      for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) D(i, j) = (uint16_t)A(j);
      break;
    case H_PRED:
      for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) D(i, j) = (uint16_t)L(i);
      break;
    case D207_PRED:
      for (int j = 0; j < size; j++) D(size - 1, j) = (uint16_t)L(size - 1);
      for (int i = 0; i < size - 1; i++) D(i, 0) = (uint16_t)r2(L(i) + L(i + 1), 1);
      for (int i = 0; i < size - 2; i++) D(i, 1) = (uint16_t)r2(L(i) + 2 * L(i + 1) + L(i + 2), 2);
      D(size - 2, 1) = (uint16_t)r2(L(size - 2) + 3 * L(size - 1), 2);
      for (int j = 2; j < size; j++)
        for (int i = size - 2; i >= 0; i--) D(i, j) = D(i + 1, j - 2);
      break;
    case D45_PRED:
      for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) {
          int k = i + j;
          D(i, j) = (uint16_t)(k + 2 < 2 * size ? r2(A(k) + 2 * A(k + 1) + A(k + 2), 2) : A(2 * size - 1));
        }
      break;
    case D63_PRED:
      for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) {
          int k = i / 2 + j;
          D(i, j) = (uint16_t)((i & 1) ? r2(A(k) + 2 * A(k + 1) + A(k + 2), 2) : r2(A(k) + A(k + 1), 1));
        }
      break;
    case D117_PRED:
      for (int j = 0; j < size; j++) D(0, j) = (uint16_t)r2(A(j - 1) + A(j), 1);
      D(1, 0) = (uint16_t)r2(L(0) + 2 * A(-1) + A(0), 2);
      for (int j = 1; j < size; j++) D(1, j) = (uint16_t)r2(A(j - 2) + 2 * A(j - 1) + A(j), 2);
      D(2, 0) = (uint16_t)r2(A(-1) + 2 * L(0) + L(1), 2);
      for (int i = 3; i < size; i++) D(i, 0) = (uint16_t)r2(L(i - 3) + 2 * L(i - 2) + L(i - 1), 2);
      for (int i = 2; i < size; i++)
        for (int j = 1; j < size; j++) D(i, j) = D(i - 2, j - 1);
      break;
    case D135_PRED:
      D(0, 0) = (uint16_t)r2(L(0) + 2 * A(-1) + A(0), 2);
      for (int j = 1; j < size; j++) D(0, j) = (uint16_t)r2(A(j - 2) + 2 * A(j - 1) + A(j), 2);
      D(1, 0) = (uint16_t)r2(A(-1) + 2 * L(0) + L(1), 2);
      for (int i = 2; i < size; i++) D(i, 0) = (uint16_t)r2(L(i - 2) + 2 * L(i - 1) + L(i), 2);
      for (int i = 1; i < size; i++)
        for (int j = 1; j < size; j++) D(i, j) = D(i - 1, j - 1);
      break;
    case D153_PRED:
      D(0, 0) = (uint16_t)r2(L(0) + A(-1), 1);
      for (int i = 1; i < size; i++) D(i, 0) = (uint16_t)r2(L(i - 1) + L(i), 1);
      D(0, 1) = (uint16_t)r2(L(0) + 2 * A(-1) + A(0), 2);
      D(1, 1) = (uint16_t)r2(A(-1) + 2 * L(0) + L(1), 2);
      for (int i = 2; i < size; i++) D(i, 1) = (uint16_t)r2(L(i - 2) + 2 * L(i - 1) + L(i), 2);
      for (int j = 2; j < size; j++) D(0, j) = (uint16_t)r2(A(j - 3) + 2 * A(j - 2) + A(j - 1), 2);
      for (int i = 1; i < size; i++)
        for (int j = 2; j < size; j++) D(i, j) = D(i - 1, j - 2);
      break;
    case TM_PRED: {
      int max = (1 << bd) - 1;
      for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) D(i, j) = (uint16_t)iclamp(A(j) + L(i) - A(-1), 0, max);
      break;
    }
    default: { /* DC_PRED */
      int v;
      if (have_left && have_above) {
        int s = 0;
        for (int k = 0; k < size; k++) s += L(k) + A(k);
        v = (s + size) >> (log2 + 1);
      } else if (have_left) {
        int s = 0;
        for (int k = 0; k < size; k++) s += L(k);
        v = (s + (1 << (log2 - 1))) >> log2;
      } else if (have_above) {
        int s = 0;
        for (int k = 0; k < size; k++) s += A(k);
        v = (s + (1 << (log2 - 1))) >> log2;
      } else {
        v = 1 << (bd - 1);
      }
      for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) D(i, j) = (uint16_t)v;
    }
  }
#undef A
#undef L
#undef D
}

/* ======================================================================== */
/* Loop filter samples (8.8.5; src/dsp/lf.rs).                               */

static void lf_narrow(uint16_t *buf, ptrdiff_t pos, ptrdiff_t step, int hev, int bd, int p1, int p0,
                      int q0, int q1) {
  int lo = -(1 << (bd - 1)), hi = (1 << (bd - 1)) - 1;
  int off = 0x80 << (bd - 8);
  int ps1 = p1 - off, ps0 = p0 - off, qs0 = q0 - off, qs1 = q1 - off;
  int f = hev ? iclamp(ps1 - qs1, lo, hi) : 0;
  f = iclamp(f + 3 * (qs0 - ps0), lo, hi);
  int f1 = iclamp(f + 4, lo, hi) >> 3;
  int f2 = iclamp(f + 3, lo, hi) >> 3;
  buf[pos] = (uint16_t)(iclamp(qs0 - f1, lo, hi) + off);
  buf[pos - step] = (uint16_t)(iclamp(ps0 + f2, lo, hi) + off);
  if (!hev) {
    int g = (f1 + 1) >> 1;
    buf[pos + step] = (uint16_t)(iclamp(qs1 - g, lo, hi) + off);
    buf[pos - 2 * step] = (uint16_t)(iclamp(ps1 + g, lo, hi) + off);
  }
}

/* the wide filter (8.8.5.3), log2_size 3 (8 taps) or 4 (16 taps) */
static void lf_wide(uint16_t *buf, ptrdiff_t pos, ptrdiff_t step, int log2_size) {
  int n = (1 << (log2_size - 1)) - 1;
  int s[16], f[16];
  for (int k = -(n + 1); k <= n; k++) s[k + 8] = buf[pos + k * step];
  for (int i = -n; i < n; i++) {
    int t = s[i + 8];
    for (int j = -n; j <= n; j++) t += s[iclamp(i + j, -(n + 1), n) + 8];
    f[i + 8] = (t + (1 << (log2_size - 1))) >> log2_size;
  }
  for (int i = -n; i < n; i++) buf[pos + i * step] = (uint16_t)f[i + 8];
}

static inline int iabs(int v) { return v < 0 ? -v : v; }

void orc_lf_filter(uint16_t *buf, ptrdiff_t pos, ptrdiff_t step, int fs, int limit, int blimit,
                   int thresh, int bd) {
#define AT(k) ((int)buf[pos + (k) * step])
  int q0 = AT(0), q1 = AT(1), q2 = AT(2), q3 = AT(3);
  int p0 = AT(-1), p1 = AT(-2), p2 = AT(-3), p3 = AT(-4);
  int sh = bd - 8;
  int lim = limit << sh, blim = blimit << sh;
  if (iabs(p3 - p2) > lim || iabs(p2 - p1) > lim || iabs(p1 - p0) > lim || iabs(q1 - q0) > lim ||
      iabs(q2 - q1) > lim || iabs(q3 - q2) > lim || iabs(p0 - q0) * 2 + iabs(p1 - q1) / 2 > blim)
    return;
  int th = thresh << sh;
  int hev = iabs(p1 - p0) > th || iabs(q1 - q0) > th;
  int one = 1 << sh;
  int flat = fs >= 1 && iabs(p1 - p0) <= one && iabs(q1 - q0) <= one && iabs(p2 - p0) <= one &&
             iabs(q2 - q0) <= one && iabs(p3 - p0) <= one && iabs(q3 - q0) <= one;
  if (fs == 0 || !flat) {
    lf_narrow(buf, pos, step, hev, bd, p1, p0, q0, q1);
    return;
  }
  int flat2 = 0;
  if (fs >= 2) {
    flat2 = 1;
    for (int k = 4; k < 8; k++)
      if (iabs(AT(-k - 1) - p0) > one || iabs(AT(k) - q0) > one) {
        flat2 = 0;
        break;
      }
  }
  lf_wide(buf, pos, step, (fs == 1 || !flat2) ? 3 : 4);
#undef AT
}

void orc_lf_filter_edges(uint16_t *buf, size_t stride, int x0, int y0, int vertical, int first,
                         int n_edges, int n_runs, const orc_edges *e, int bd) {
  ptrdiff_t step = vertical ? 1 : (ptrdiff_t)stride, along = vertical ? (ptrdiff_t)stride : 1;
  for (int ed = first; ed < n_edges; ed++)
    for (int r = 0; r < n_runs; r++) {
      int fs = e->fs[ed][r];
      if (fs < 0) continue;
      ptrdiff_t pos = vertical ? (ptrdiff_t)(y0 + 4 * r) * stride + x0 + 4 * ed
                               : (ptrdiff_t)(y0 + 4 * ed) * stride + x0 + 4 * r;
      for (int k = 0; k < 4; k++)
        orc_lf_filter(buf, pos + k * along, step, fs, e->limit[ed][r], e->blimit[ed][r],
                      e->thresh[ed][r], bd);
    }
}

/* ======================================================================== */
/* Frames.                                                                   */

orc_frame *orc_frame_new(uint32_t width, uint32_t height, uint32_t bd, uint32_t ssx, uint32_t ssy) {
  orc_frame *f = (orc_frame *)calloc(1, sizeof *f);
  uint32_t mi_cols = (width + 7) >> 3, mi_rows = (height + 7) >> 3;
  uint32_t w = ((mi_cols + 7) >> 3) * 64, h = ((mi_rows + 7) >> 3) * 64;
  f->width = width;
  f->height = height;
  f->bit_depth = bd;
  f->ss_x = ssx;
  f->ss_y = ssy;
  for (int p = 0; p < 3; p++) {
    f->stride[p] = p ? w >> ssx : w;
    f->rows[p] = p ? h >> ssy : h;
    f->plane[p] = (uint16_t *)calloc((size_t)f->stride[p] * f->rows[p], sizeof(uint16_t));
  }
  return f;
}

void orc_frame_free(orc_frame *f) {
  if (!f) return;
  for (int p = 0; p < 3; p++) free(f->plane[p]);
  free(f);
}

size_t orc_packed_size(const orc_frame *f) {
  size_t bps = f->bit_depth > 8 ? 2 : 1;
  size_t cw = (f->width + f->ss_x) >> f->ss_x, ch = (f->height + f->ss_y) >> f->ss_y;
  return bps * ((size_t)f->width * f->height + 2 * cw * ch);
}

void orc_pack(const orc_frame *f, uint8_t *dst) {
  int high = f->bit_depth > 8;
  for (int p = 0; p < 3; p++) {
    uint32_t w = p ? (f->width + f->ss_x) >> f->ss_x : f->width;
    uint32_t h = p ? (f->height + f->ss_y) >> f->ss_y : f->height;
    for (uint32_t y = 0; y < h; y++) {
      const uint16_t *row = f->plane[p] + (size_t)y * f->stride[p];
      for (uint32_t x = 0; x < w; x++) {
        if (high) {
          *dst++ = (uint8_t)(row[x] & 0xff);
          *dst++ = (uint8_t)(row[x] >> 8);
        } else {
          *dst++ = (uint8_t)row[x];
        }
      }
    }
  }
}

/* ---- reconstruction of a record (block.rs:1405-1498; recon.rs) ---- */

typedef struct recon_ctx {
  const vp9r_record *rec;
  const vp9r_frame_hdr *h;
  const orc_frame *const *refs;
  orc_frame *cur;
  int32_t coefs[1024];
  uint16_t pred[2][64 * 64];
} recon_ctx;

static void uv_subsampling(const vp9r_frame_hdr *h, int *sx, int *sy) {
  if (h->legacy_uv) {
    *sx = *sy = 1;
  } else {
    *sx = h->ss_x;
    *sy = h->ss_y;
  }
}

/* get_uv_tx_size (block.rs:1501-1508) */
static int uv_tx_size(const vp9r_frame_hdr *h, const vp9r_block *b) {
  if (b->mi_size < BLOCK_8X8) return TX_4X4;
  int sx, sy;
  uv_subsampling(h, &sx, &sy);
  return imin(b->tx_size, VP9_MAX_TXSIZE_LOOKUP[VP9_SS_SIZE_LOOKUP[b->mi_size][sx][sy]]);
}

static inline int32_t q2(int32_t v) { return (v < 0 ? v - 1 : v + 1) / 2; }
static inline int32_t q4(int32_t v) { return (v < 0 ? v - 2 : v + 2) / 4; }

/* predict_intra (recon.rs:17-103) */
static void predict_intra(recon_ctx *R, const vp9r_block *b, int plane, int x, int y, int have_left,
                          int have_above, int not_on_right, int tx_sz, int block_idx) {
  const vp9r_frame_hdr *h = R->h;
  int mode = plane > 0 ? b->uv_mode : (b->mi_size >= BLOCK_8X8 ? b->y_mode : b->sub_modes[block_idx]);
  int log2 = tx_sz + 2, size = 1 << log2;
  int sx = plane ? h->ss_x : 0, sy = plane ? h->ss_y : 0;
  int max_x = (int)((h->mi_cols * 8) >> sx) - 1, max_y = (int)((h->mi_rows * 8) >> sy) - 1;
  int base = 1 << (h->bit_depth - 1);
  uint16_t *pl = R->cur->plane[plane];
  size_t stride = R->cur->stride[plane];
  int32_t above[65], left[32];
  if (have_above) {
    const uint16_t *row = pl + (size_t)(y - 1) * stride;
    for (int i = 0; i < size; i++) above[1 + i] = row[imin(max_x, x + i)];
    if (not_on_right && tx_sz == TX_4X4)
      for (int i = size; i < 2 * size; i++) above[1 + i] = row[imin(max_x, x + i)];
    else
      for (int i = size; i < 2 * size; i++) above[1 + i] = above[size];
    above[0] = have_left ? row[imin(max_x, x - 1)] : base + 1;
  } else {
    for (int i = 0; i <= 2 * size; i++) above[i] = base - 1;
  }
  if (have_left)
    for (int i = 0; i < size; i++) left[i] = pl[(size_t)imin(max_y, y + i) * stride + x - 1];
  else
    for (int i = 0; i < size; i++) left[i] = base + 1;
  orc_intra_predict(mode, log2, above, left, have_left, have_above, h->bit_depth,
                    pl + (size_t)y * stride + x, stride);
}

/* predict_inter (recon.rs:106-217) with inter_scale / Scale::position
 * (decoder/mod.rs:495-541) */
static int predict_inter(recon_ctx *R, const vp9r_block *b, int plane, int x, int y, int w, int h_,
                         int block_idx) {
  const vp9r_frame_hdr *h = R->h;
  int is_compound = b->ref_frame[1] > INTRA_FRAME;
  for (int rl = 0; rl < 1 + is_compound; rl++) {
    const int32_t(*bm)[2] = b->mv[rl];
    int32_t mv[2];
    if (plane == 0 || b->mi_size >= BLOCK_8X8) {
      mv[0] = bm[block_idx][0];
      mv[1] = bm[block_idx][1];
    } else {
      int usx, usy;
      uv_subsampling(h, &usx, &usy);
      if (usx == 0 && usy == 0) {
        mv[0] = bm[block_idx][0];
        mv[1] = bm[block_idx][1];
      } else if (usx == 0) {
        mv[0] = q2(bm[block_idx][0] + bm[block_idx + 2][0]);
        mv[1] = q2(bm[block_idx][1] + bm[block_idx + 2][1]);
      } else if (usy == 0) {
        mv[0] = q2(bm[block_idx][0] + bm[block_idx + 1][0]);
        mv[1] = q2(bm[block_idx][1] + bm[block_idx + 1][1]);
      } else {
        mv[0] = q4(bm[0][0] + bm[1][0] + bm[2][0] + bm[3][0]);
        mv[1] = q4(bm[0][1] + bm[1][1] + bm[2][1] + bm[3][1]);
      }
    }
    /* motion vector clamping (8.5.2.2) */
    int sx = plane ? h->ss_x : 0, sy = plane ? h->ss_y : 0;
    int bh = VP9_NUM_8X8_HIGH[b->mi_size], bw = VP9_NUM_8X8_WIDE[b->mi_size];
    int mi_row = b->mi_row, mi_col = b->mi_col;
    int to_top = -((mi_row * 8) * 16) >> sy;
    int to_bottom = ((((int)h->mi_rows - bh - mi_row) * 8) * 16) >> sy;
    int to_left = -((mi_col * 8) * 16) >> sx;
    int to_right = ((((int)h->mi_cols - bw - mi_col) * 8) * 16) >> sx;
    int spel_left = (INTERP_EXTEND + ((bw * 8) >> sx)) << 4, spel_right = spel_left - 16;
    int spel_top = (INTERP_EXTEND + ((bh * 8) >> sy)) << 4, spel_bottom = spel_top - 16;
    int32_t cmv0 = iclamp((2 * mv[0]) >> sy, to_top - spel_top, to_bottom + spel_bottom);
    int32_t cmv1 = iclamp((2 * mv[1]) >> sx, to_left - spel_left, to_right + spel_right);
    /* motion vector scaling (8.5.2.3) */
    int rf = b->ref_frame[rl];
    const orc_frame *r = R->refs[rf - LAST_FRAME];
    if (!r) return VP9R_ERR_BITSTREAM;
    uint32_t rw = r->width, rh = r->height, fw = h->width, fh = h->height;
    if (2 * fw < rw || 2 * fh < rh || fw > 16 * rw || fh > 16 * rh) return VP9R_ERR_BITSTREAM;
    int64_t xs = ((int64_t)rw << 14) / fw, ys = ((int64_t)rh << 14) / fh;
    int64_t base_x = ((int64_t)x * xs) >> 14, base_y = ((int64_t)y * ys) >> 14;
    int64_t luma_x = plane ? (int64_t)x << h->ss_x : x, luma_y = plane ? (int64_t)y << h->ss_y : y;
    int64_t frac_x = ((16 * luma_x * xs) >> 14) & 15, frac_y = ((16 * luma_y * ys) >> 14) & 15;
    int64_t dx = (((int64_t)cmv1 * xs) >> 14) + frac_x, dy = (((int64_t)cmv0 * ys) >> 14) + frac_y;
    int step_x = (int)((16 * xs) >> 14), step_y = (int)((16 * ys) >> 14);
    int start_x = (int)((base_x << 4) + dx), start_y = (int)((base_y << 4) + dy);
    int last_x = (int)((rw + sx) >> sx) - 1, last_y = (int)((rh + sy) >> sy) - 1;
    uint16_t *out;
    size_t out_stride;
    if (is_compound) {
      out = R->pred[rl];
      out_stride = w;
    } else {
      out = R->cur->plane[plane] + (size_t)y * R->cur->stride[plane] + x;
      out_stride = R->cur->stride[plane];
    }
    orc_inter_predict(r->plane[plane], r->stride[plane], last_x, last_y, start_x, start_y, step_x,
                      step_y, w, h_, b->interp_filter, h->bit_depth, out, out_stride);
  }
  if (is_compound)
    orc_avg(R->pred[0], R->pred[1], w, h_, R->cur->plane[plane] + (size_t)y * R->cur->stride[plane] + x,
            R->cur->stride[plane]);
  return 0;
}

/* qindex (recon.rs:219-230) */
static int qindex(const vp9r_frame_hdr *h, int seg) {
  if (h->seg_enabled && h->seg_feature_enabled[seg][SEG_LVL_ALT_Q]) {
    int data = h->seg_feature_data[seg][SEG_LVL_ALT_Q];
    if (!h->seg_abs_or_delta_update) data += h->base_q_idx;
    return iclamp(data, 0, 255);
  }
  return h->base_q_idx;
}

/* reconstruct (recon.rs:233-281): dequantisation then the inverse transform */
static void reconstruct(recon_ctx *R, const vp9r_block *b, const vp9r_txb *t) {
  const vp9r_frame_hdr *h = R->h;
  int bd_idx = (h->bit_depth - 8) >> 1;
  int q = qindex(h, b->segment_id);
  int dc_delta = t->plane == 0 ? h->delta_q_y_dc : h->delta_q_uv_dc;
  int ac_delta = t->plane == 0 ? 0 : h->delta_q_uv_ac;
  int64_t dc_q = VP9_DC_QLOOKUP[bd_idx][iclamp(q + dc_delta, 0, 255)];
  int64_t ac_q = VP9_AC_QLOOKUP[bd_idx][iclamp(q + ac_delta, 0, 255)];
  int n = 2 + t->tx_size, n0 = 1 << n;
  memset(R->coefs, 0, sizeof(int32_t) * n0 * n0);
  for (uint32_t k = 0; k < t->ncoef; k++) {
    int pos = R->rec->coef_pos[t->coef_start + k];
    int64_t v = R->rec->coef_val[t->coef_start + k];
    int64_t d = v * (pos == 0 ? dc_q : ac_q);
    if (t->tx_size == TX_32X32) d /= 2;
    R->coefs[pos] = (int32_t)lclamp(d, INT32_MIN, INT32_MAX);
  }
  uint16_t *pl = R->cur->plane[t->plane];
  size_t stride = R->cur->stride[t->plane];
  orc_itx_add(R->coefs, n, t->tx_type, h->lossless, t->eob, h->bit_depth, pl + (size_t)t->y * stride + t->x,
              stride);
}

/* One block of the replay (block.rs:1405-1498): for each plane, inter
 * prediction of the whole block (or each 4x4 below 8x8), then its transform
 * blocks from *cursor on, in raster order: intra prediction, then the
 * residual when eob > 0. */
static int recon_block(recon_ctx *R, uint32_t bi, uint32_t *cursor) {
  const vp9r_record *rec = R->rec;
  const vp9r_frame_hdr *h = R->h;
  const vp9r_block *b = &rec->blocks[bi];
  int bsize = b->mi_size < BLOCK_8X8 ? BLOCK_8X8 : b->mi_size;
  int status = 0;
  for (int plane = 0; plane < 3 && !status; plane++) {
    int sx = plane ? h->ss_x : 0, sy = plane ? h->ss_y : 0;
    int tx_sz = plane ? uv_tx_size(h, b) : b->tx_size;
    int step = 1 << tx_sz;
    int plane_sz = VP9_SS_SIZE_LOOKUP[bsize][sx][sy];
    int n4w = VP9_NUM_4X4_WIDE[plane_sz], n4h = VP9_NUM_4X4_HIGH[plane_sz];
    int base_x = (b->mi_col * 8) >> sx, base_y = (b->mi_row * 8) >> sy;
    if (b->is_inter) {
      if (b->mi_size < BLOCK_8X8) {
        for (int y = 0; y < n4h && !status; y++)
          for (int x = 0; x < n4w && !status; x++)
            status = predict_inter(R, b, plane, base_x + 4 * x, base_y + 4 * y, 4, 4, y * n4w + x);
      } else {
        status = predict_inter(R, b, plane, base_x, base_y, n4w * 4, n4h * 4, 0);
      }
    }
    int per_row = (n4w + step - 1) / step;
    while (!status && *cursor < rec->n_txbs && rec->txbs[*cursor].block == bi && rec->txbs[*cursor].plane == plane) {
      const vp9r_txb *t = &rec->txbs[(*cursor)++];
      int xi = (t->x - base_x) >> 2, yi = (t->y - base_y) >> 2;
      int block_idx = (yi / step) * per_row + xi / step;
      if (!b->is_inter)
        predict_intra(R, b, plane, t->x, t->y, b->avail_l || xi > 0, b->avail_u || yi > 0, xi + step < n4w, tx_sz,
                      block_idx);
      if (t->eob > 0) reconstruct(R, b, t);
    }
  }
  return status;
}

static recon_ctx *recon_new(const vp9r_record *rec, const orc_frame *const refs[3], orc_frame *cur) {
  recon_ctx *R = (recon_ctx *)malloc(sizeof *R);
  R->rec = rec;
  R->h = &rec->hdr;
  R->refs = refs;
  R->cur = cur;
  return R;
}

int orc_reconstruct(const vp9r_record *rec, const orc_frame *const refs[3], orc_frame *cur) {
  recon_ctx *R = recon_new(rec, refs, cur);
  uint32_t cursor = 0;
  int status = 0;
  for (uint32_t bi = 0; bi < rec->n_blocks && !status; bi++) status = recon_block(R, bi, &cursor);
  free(R);
  return status;
}

/* ---- the threaded forms (the all-core CPU baseline) ----
 *
 * Reconstruction with the tile columns on several threads, as the
 * reference's decode_tiles_parallel (src/decoder/block.rs:1769-1866): a tile
 * column reads nothing of another (intra edges stop at the tile's left edge
 * and never reach right of the block; inter prediction reads references
 * only), so each column's blocks, in record order, are one task.  The loop
 * filter as the reference's filter_bands (src/decoder/loopfilter.rs:195-265):
 * superblock rows of each plane as tasks in order, superblock (r, c) after
 * (r, c - 1) and after (r - 1, c + 1) of the same plane, whose regions are
 * the only earlier ones in raster order that overlap it; the result is the
 * serial filter's. */

typedef struct mt_pool {
  int n_tasks;
  int next; /* atomic */
  void (*run)(struct mt_pool *, int);
} mt_pool;

static void *mt_worker(void *arg) {
  mt_pool *p = (mt_pool *)arg;
  for (;;) {
    int k = __atomic_fetch_add(&p->next, 1, __ATOMIC_RELAXED);
    if (k >= p->n_tasks) return NULL;
    p->run(p, k);
  }
}

/* Tasks are taken in index order, so a task only ever waits for tasks that
 * were taken before it. */
static void mt_run(mt_pool *p, int threads) {
  if (threads > 64) threads = 64;
  if (threads > p->n_tasks) threads = p->n_tasks;
  pthread_t tid[64];
  int started = 0;
  for (int i = 1; i < threads; i++)
    if (pthread_create(&tid[started], NULL, mt_worker, p) == 0) started++;
  mt_worker(p);
  for (int i = 0; i < started; i++) pthread_join(tid[i], NULL);
}

typedef struct recon_mt {
  mt_pool pool;
  const vp9r_record *rec;
  const orc_frame *const *refs;
  orc_frame *cur;
  const uint32_t *txb_first; /* per block: its first transform block */
  const uint8_t *col_of;     /* per block: its tile column */
  int status[64];
} recon_mt;

static void recon_column(mt_pool *p, int col) {
  recon_mt *M = (recon_mt *)p;
  recon_ctx *R = recon_new(M->rec, M->refs, M->cur);
  int status = 0;
  for (uint32_t bi = 0; bi < M->rec->n_blocks && !status; bi++) {
    if (M->col_of[bi] != col) continue;
    uint32_t cursor = M->txb_first[bi];
    status = recon_block(R, bi, &cursor);
  }
  M->status[col] = status;
  free(R);
}

int orc_reconstruct_mt(const vp9r_record *rec, const orc_frame *const refs[3], orc_frame *cur, int threads) {
  const vp9r_frame_hdr *h = &rec->hdr;
  const int cols = 1 << h->tile_cols_log2;
  if (threads <= 1 || cols == 1 || cols > 64) return orc_reconstruct(rec, refs, cur);
  recon_mt M;
  memset(&M, 0, sizeof M);
  uint32_t *txb_first = (uint32_t *)malloc(sizeof(uint32_t) * (rec->n_blocks + 1));
  uint8_t *col_of = (uint8_t *)malloc(rec->n_blocks + 1);
  uint32_t t = 0;
  for (uint32_t bi = 0; bi < rec->n_blocks; bi++) {
    while (t < rec->n_txbs && rec->txbs[t].block < bi) t++;
    txb_first[bi] = t;
    /* the tile column: MiColStarts (6.4), in whole superblocks */
    int c = 0;
    while (c + 1 < cols && (uint32_t)((((c + 1) * h->sb64_cols) >> h->tile_cols_log2) * 8) <= rec->blocks[bi].mi_col) c++;
    col_of[bi] = (uint8_t)c;
  }
  M.pool.n_tasks = cols;
  M.pool.run = recon_column;
  M.rec = rec;
  M.refs = refs;
  M.cur = cur;
  M.txb_first = txb_first;
  M.col_of = col_of;
  mt_run(&M.pool, threads);
  free(txb_first);
  free(col_of);
  for (int c = 0; c < cols; c++)
    if (M.status[c]) return M.status[c];
  return 0;
}

/* ---- loop filter frame process (src/decoder/loopfilter.rs) ---- */

typedef struct lf_ctx {
  const vp9r_record *rec;
  uint8_t lvl[MAX_SEGMENTS][4][2];
  int params[64][3];
} lf_ctx;

/* lvl_lookup (loopfilter.rs:15-47) and the limits of Ctx::new (135-170) */
static void lf_setup(const vp9r_record *rec, lf_ctx *c) {
  const vp9r_frame_hdr *h = &rec->hdr;
  c->rec = rec;
  int n_shift = h->lf_level >> 5;
  for (int s = 0; s < MAX_SEGMENTS; s++) {
    int lvl_seg = h->lf_level;
    if (h->seg_enabled && h->seg_feature_enabled[s][SEG_LVL_ALT_L]) {
      int data = h->seg_feature_data[s][SEG_LVL_ALT_L];
      lvl_seg = h->seg_abs_or_delta_update ? data : data + h->lf_level;
      lvl_seg = iclamp(lvl_seg, 0, MAX_LOOP_FILTER);
    }
    if (!h->lf_delta_enabled) {
      for (int r = 0; r < 4; r++) c->lvl[s][r][0] = c->lvl[s][r][1] = (uint8_t)lvl_seg;
    } else {
      int intra = lvl_seg + (h->lf_ref_deltas[INTRA_FRAME] * (1 << n_shift));
      c->lvl[s][0][0] = c->lvl[s][0][1] = (uint8_t)iclamp(intra, 0, MAX_LOOP_FILTER);
      for (int rf = 1; rf < 4; rf++)
        for (int m = 0; m < 2; m++) {
          int inter = lvl_seg + h->lf_ref_deltas[rf] * (1 << n_shift) + h->lf_mode_deltas[m] * (1 << n_shift);
          c->lvl[s][rf][m] = (uint8_t)iclamp(inter, 0, MAX_LOOP_FILTER);
        }
    }
  }
  int sharp = h->lf_sharpness;
  int shift = sharp > 4 ? 2 : sharp > 0 ? 1 : 0;
  for (int l = 0; l < 64; l++) {
    int limit = sharp > 0 ? iclamp(l >> shift, 1, 9 - sharp) : imax(l >> shift, 1);
    c->params[l][0] = limit;
    c->params[l][1] = 2 * (l + 2) + limit;
    c->params[l][2] = l >> 4;
  }
}

/* superblock (loopfilter.rs:466-620) */
static int lf_superblock(const lf_ctx *C, int plane, int pass, uint32_t row, uint32_t col,
                         orc_edges *e, int *x0, int *y0) {
  const vp9r_frame_hdr *h = &C->rec->hdr;
  const vp9r_mi *mi = C->rec->mi;
  int sub_x = plane ? h->ss_x : 0, sub_y = plane ? h->ss_y : 0;
  int sub = pass == 0 ? sub_x : sub_y;
  int edge_len = pass == 0 ? 64 >> sub_y : 64 >> sub_x;
  uint32_t mi_rows = h->mi_rows, mi_cols = h->mi_cols;
  int n_edges = 16 >> sub, n_runs = edge_len / 4;
  for (int i = 0; i < n_edges; i++)
    for (int j = 0; j < n_runs; j++) e->fs[i][j] = -1;
  int ux, uy;
  uv_subsampling(h, &ux, &uy);
#define AT_X(ed, rn) (pass == 0 ? col * 8 + (ed) * (4u << sub_x) : col * 8 + ((4u * (rn)) << sub_x))
#define AT_Y(ed, rn) (pass == 0 ? row * 8 + ((4u * (rn)) << sub_y) : row * 8 + (ed) * (4u << sub_y))
  for (int ca = 0; ca < n_edges / 2; ca++)
    for (int cb = 0; cb < n_runs / 2; cb++) {
      uint32_t x = AT_X(2 * ca, 2 * cb), y = AT_Y(2 * ca, 2 * cb);
      if (x >= 8 * mi_cols || y >= 8 * mi_rows) continue;
      uint32_t loop_col = ((x >> 3) >> sub_x) << sub_x, loop_row = ((y >> 3) >> sub_y) << sub_y;
      const vp9r_mi *m = &mi[loop_row * mi_cols + loop_col];
      int mode = m->y_mode;
      int mode_type = mode == NEARESTMV || mode == NEARMV || mode == NEWMV;
      int rf = imax(m->ref_frame0, 0);
      int l = C->lvl[m->segment_id][rf][mode_type];
      if (l == 0) continue;
      int limit = C->params[l][0], blimit = C->params[l][1], thresh = C->params[l][2];
      int mi_size = m->mi_size;
      int tx_sz;
      if (plane > 0) {
        tx_sz = mi_size < BLOCK_8X8 ? TX_4X4
                                    : imin(m->tx_size, VP9_MAX_TXSIZE_LOOKUP[VP9_SS_SIZE_LOOKUP[mi_size][ux][uy]]);
      } else {
        tx_sz = m->tx_size;
      }
      int sb_size = sub == 0 ? mi_size : imax(mi_size, BLOCK_16X16);
      uint32_t block_mask = pass == 0 ? 8u * VP9_NUM_8X8_WIDE[sb_size] - 1 : 8u * VP9_NUM_8X8_HIGH[sb_size] - 1;
      int coded = m->ref_frame0 <= INTRA_FRAME || !m->skip;
      for (int eo = 0; eo < 2; eo++) {
        int ed = 2 * ca + eo;
        x = AT_X(ed, 2 * cb);
        y = AT_Y(ed, 2 * cb);
        if ((pass == 0 && x == 0) || (pass == 1 && y == 0)) continue;
        uint32_t across = pass == 0 ? x : y;
        int is_block_edge = (across & block_mask) == 0;
        int tx_edge = (ed & ((1 << tx_sz) - 1)) == 0;
        if (!is_block_edge && !(tx_edge && coded)) continue;
        int is_32_edge = ed % 8 == 0;
        int base_size = (tx_sz == TX_4X4 && is_32_edge) ? TX_8X8 : imin(tx_sz, TX_16X16);
        int filter_size =
            ((pass == 0 && sub_x == 1 && base_size == TX_16X16 && (x >> 3) == mi_cols - 1) ||
             (pass == 1 && sub_y == 1 && base_size == TX_16X16 && (y >> 3) == mi_rows - 1))
                ? TX_8X8
                : base_size;
        for (int k = 0; k < 2; k++) {
          int rn = 2 * cb + k;
          uint32_t xx = AT_X(ed, rn), yy = AT_Y(ed, rn);
          if (k == 1 && (xx >= 8 * mi_cols || yy >= 8 * mi_rows)) break;
          int is_tx_edge = (pass == 1 && sub_x == 1 && (mi_cols & 1) && (ed & 1) && (xx + 8) >= mi_cols * 8)
                               ? 0
                               : tx_edge;
          if (!(is_block_edge || (is_tx_edge && coded))) continue;
          e->fs[ed][rn] = (int8_t)filter_size;
          e->limit[ed][rn] = (uint8_t)limit;
          e->blimit[ed][rn] = (uint8_t)blimit;
          e->thresh[ed][rn] = (uint8_t)thresh;
        }
      }
    }
#undef AT_X
#undef AT_Y
  *x0 = (int)((col * 8) >> sub_x);
  *y0 = (int)((row * 8) >> sub_y);
  return n_edges | (n_runs << 8);
}

int orc_lf_superblock(const vp9r_record *rec, int plane, int pass, uint32_t row, uint32_t col,
                      orc_edges *e, int *x0, int *y0) {
  lf_ctx C;
  lf_setup(rec, &C);
  return lf_superblock(&C, plane, pass, row, col, e, x0, y0);
}

void orc_loop_filter(const vp9r_record *rec, orc_frame *cur) {
  const vp9r_frame_hdr *h = &rec->hdr;
  if (h->lf_level == 0) return;
  lf_ctx C;
  lf_setup(rec, &C);
  orc_edges e;
  for (int plane = 0; plane < 3; plane++)
    for (uint32_t row = 0; row < h->mi_rows; row += 8)
      for (uint32_t col = 0; col < h->mi_cols; col += 8)
        for (int pass = 0; pass < 2; pass++) {
          int x0, y0;
          int g = lf_superblock(&C, plane, pass, row, col, &e, &x0, &y0);
          orc_lf_filter_edges(cur->plane[plane], cur->stride[plane], x0, y0, pass == 0, 0, g & 0xff,
                              g >> 8, &e, h->bit_depth);
        }
}

typedef struct lf_mt {
  mt_pool pool;
  const vp9r_record *rec;
  orc_frame *cur;
  lf_ctx C;
  uint32_t rows, cols;
  int *done; /* per (plane, superblock row): superblocks finished (atomic) */
} lf_mt;

static void lf_band(mt_pool *p, int k) {
  lf_mt *M = (lf_mt *)p;
  const vp9r_frame_hdr *h = &M->rec->hdr;
  const int plane = k / (int)M->rows, r = k % (int)M->rows;
  int *above = r ? &M->done[plane * M->rows + r - 1] : NULL;
  orc_edges e;
  for (uint32_t c = 0; c < M->cols; c++) {
    /* (r - 1, c + 1) done, or the whole row above */
    const int need = (int)(c + 2 < M->cols ? c + 2 : M->cols);
    if (above)
      while (__atomic_load_n(above, __ATOMIC_ACQUIRE) < need) sched_yield();
    for (int pass = 0; pass < 2; pass++) {
      int x0, y0;
      int g = lf_superblock(&M->C, plane, pass, (uint32_t)r * 8, c * 8, &e, &x0, &y0);
      orc_lf_filter_edges(M->cur->plane[plane], M->cur->stride[plane], x0, y0, pass == 0, 0, g & 0xff, g >> 8, &e,
                          h->bit_depth);
    }
    __atomic_store_n(&M->done[plane * M->rows + r], (int)c + 1, __ATOMIC_RELEASE);
  }
}

void orc_loop_filter_mt(const vp9r_record *rec, orc_frame *cur, int threads) {
  const vp9r_frame_hdr *h = &rec->hdr;
  if (h->lf_level == 0) return;
  if (threads <= 1) {
    orc_loop_filter(rec, cur);
    return;
  }
  lf_mt M;
  memset(&M, 0, sizeof M);
  lf_setup(rec, &M.C);
  M.rec = rec;
  M.cur = cur;
  M.rows = (h->mi_rows + 7) / 8;
  M.cols = (h->mi_cols + 7) / 8;
  M.done = (int *)calloc(3 * M.rows, sizeof(int));
  M.pool.n_tasks = (int)(3 * M.rows);
  M.pool.run = lf_band;
  mt_run(&M.pool, threads);
  free(M.done);
}
