// Device-side building blocks shared by the gfx950 kernels: integer lanes of
// the inverse transforms, the 1D DCT/ADST/WHT butterflies, sample types.
//
// The transforms are the specification's processes (8.7.1) in the exact
// integer arithmetic the reference crate uses (src/dsp/itx.rs): at 8 bits a
// 32-bit lane whose every operation wraps (itx.rs:573-614), above 8 bits a
// 64-bit lane whose multiply takes the low 32 bits of its operand
// (itx.rs:626-666).  Each 1D transform runs in one lane's registers, fully
// unrolled for its size, so every index and constant folds at compile time.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEV __device__ __forceinline__
#define HDEV __host__ __device__ __forceinline__

// Loads and stores through global-memory pointers.  Pointers that reach a
// kernel inside argument structs are generic, and generic (flat) accesses
// count against the LDS wait counter too: an LDS wait would also wait for
// every frame load or store in flight.
// (through a native type of the same size: 4, 8 or 16 bytes)
template <int N>
struct GWord;
template <>
struct GWord<1> {
  typedef unsigned char T;
};
template <>
struct GWord<2> {
  typedef unsigned short T;
};
template <>
struct GWord<4> {
  typedef unsigned T;
};
template <>
struct GWord<8> {
  typedef unsigned __attribute__((ext_vector_type(2))) T;
};
template <>
struct GWord<16> {
  typedef unsigned __attribute__((ext_vector_type(4))) T;
};
template <class T>
DEV T gld(const T *p) {
  typedef typename GWord<sizeof(T)>::T W;
  const W w = *(const __attribute__((address_space(1))) W *)p;
  T r;
  __builtin_memcpy(&r, &w, sizeof(T));
  return r;
}
template <class T>
DEV void gst(T *p, const T &v) {
  typedef typename GWord<sizeof(T)>::T W;
  W w;
  __builtin_memcpy(&w, &v, sizeof(T));
  *(__attribute__((address_space(1))) W *)p = w;
}

namespace vp9g {

// Orders a wave's own LDS accesses across lanes without a workgroup barrier,
// for LDS that one wave alone uses: a wave's LDS instructions execute in
// program order, and the fence keeps the compiler from moving accesses
// across the point.
DEV void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---- integer lanes

struct Lane32 {
  typedef int32_t T;
  static DEV T add(T a, T b) { return (T)((uint32_t)a + (uint32_t)b); }
  static DEV T sub(T a, T b) { return (T)((uint32_t)a - (uint32_t)b); }
  static DEV T mul(T a, int32_t c) { return (T)((uint32_t)a * (uint32_t)c); }
  static DEV T rnd(T a, int n) { return (T)((uint32_t)a + (1u << (n - 1))) >> n; }
  static DEV T sar(T a, int n) { return a >> n; }
  static DEV T from32(int32_t v) { return v; }
};

struct Lane64 {
  typedef int64_t T;
  static DEV T add(T a, T b) { return (T)((uint64_t)a + (uint64_t)b); }
  static DEV T sub(T a, T b) { return (T)((uint64_t)a - (uint64_t)b); }
  static DEV T mul(T a, int32_t c) { return (int64_t)(int32_t)a * (int64_t)c; }
  static DEV T rnd(T a, int n) { return (T)((uint64_t)a + (1ull << (n - 1))) >> n; }
  static DEV T sar(T a, int n) { return a >> n; }
  static DEV T from32(int32_t v) { return v; }
};

// ---- constants of 8.7.1.1

constexpr int32_t kCos64[33] = {16384, 16364, 16305, 16207, 16069, 15893, 15679, 15426, 15137,
                                14811, 14449, 14053, 13623, 13160, 12665, 12140, 11585, 11003,
                                10394, 9760,  9102,  8423,  7723,  7005,  6270,  5520,  4756,
                                3981,  3196,  2404,  1606,  804,   0};
constexpr int32_t cos64(int angle) {
  return (angle & 127) <= 32   ? kCos64[angle & 127]
         : (angle & 127) <= 64 ? -kCos64[64 - (angle & 127)]
         : (angle & 127) <= 96 ? -kCos64[(angle & 127) - 64]
                               : kCos64[128 - (angle & 127)];
}
constexpr int32_t sin64(int angle) { return cos64(angle - 32); }
constexpr int brev(int bits, int x) {
  int t = 0;
  for (int i = 0; i < bits; i++) t |= ((x >> i) & 1) << (bits - 1 - i);
  return t;
}
constexpr int32_t kSinpi19 = 5283, kSinpi29 = 9929, kSinpi39 = 13377, kSinpi49 = 15212;

// ---- butterflies: B, H, SB, SH of 8.7.1.1

template <class L>
DEV void bf_B(typename L::T *t, int a, int b, int angle, bool flip) {
  const int32_t c = cos64(angle), s = sin64(angle);
  typename L::T x = L::rnd(L::sub(L::mul(t[a], c), L::mul(t[b], s)), 14);
  typename L::T y = L::rnd(L::add(L::mul(t[a], s), L::mul(t[b], c)), 14);
  t[a] = flip ? y : x;
  t[b] = flip ? x : y;
}
template <class L>
DEV void bf_H(typename L::T *t, int a, int b, bool flip) {
  int p = flip ? b : a, q = flip ? a : b;
  typename L::T x = t[p], y = t[q];
  t[p] = L::add(x, y);
  t[q] = L::sub(x, y);
}
template <class L>
DEV void bf_SB(const typename L::T *t, typename L::T *s, int a, int b, int angle, bool flip) {
  const int32_t c = cos64(angle), sn = sin64(angle);
  typename L::T x = L::sub(L::mul(t[a], c), L::mul(t[b], sn));
  typename L::T y = L::add(L::mul(t[a], sn), L::mul(t[b], c));
  s[a] = flip ? y : x;
  s[b] = flip ? x : y;
}
template <class L>
DEV void bf_SH(typename L::T *t, const typename L::T *s, int a, int b) {
  t[a] = L::rnd(L::add(s[a], s[b]), 14);
  t[b] = L::rnd(L::sub(s[a], s[b]), 14);
}

// The steps of the inverse DCT process (8.7.1.3) for size 2^n after the
// recursive half.
template <class L, int n>
DEV void idct_stage(typename L::T *t) {
  constexpr int n0 = 1 << n, n1 = n0 >> 1, n2 = n0 >> 2, n3 = n0 >> 3;
#pragma unroll
  for (int i = 0; i < n2; i++) bf_B<L>(t, n1 + i, n0 - 1 - i, 32 - brev(5, n1 + i), false);
  if constexpr (n >= 3) {
#pragma unroll
    for (int i = 0; i < n3; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) bf_H<L>(t, n1 + 4 * i + 2 * j, n1 + 1 + 4 * i + 2 * j, j == 1);
  }
  if constexpr (n == 5) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        bf_B<L>(t, n0 - n + 3 - n2 * j - 4 * i, n1 + n - 4 + n2 * j + 4 * i, 28 - 16 * i + 56 * j, true);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) bf_H<L>(t, n1 + n3 * j + i, n1 + n2 - 5 + n3 * j - i, (j & 1) == 1);
  }
  if constexpr (n >= 4) {
#pragma unroll
    for (int i = 0; i <= (n == 5 ? 1 : 0); i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        bf_B<L>(t, n0 - n + 2 - i - n2 * j, n1 + n - 3 + i + n2 * j, 24 + 48 * j, true);
#pragma unroll
    for (int i = 0; i < 2 * n - 6; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) bf_H<L>(t, n1 + n2 * j + i, n1 + n2 - 1 + n2 * j - i, (j & 1) == 1);
  }
  if constexpr (n >= 3) {
#pragma unroll
    for (int i = 0; i < n3; i++) bf_B<L>(t, n0 - n3 - 1 - i, n1 + n3 + i, 16, true);
  }
#pragma unroll
  for (int i = 0; i < n1; i++) bf_H<L>(t, i, n0 - 1 - i, false);
}

// inverse DCT array permutation (8.7.1.2), then the inverse DCT
template <class L, int n>
DEV void idct(typename L::T *t) {
  constexpr int n0 = 1 << n;
  typename L::T c[n0];
#pragma unroll
  for (int i = 0; i < n0; i++) c[i] = t[i];
#pragma unroll
  for (int i = 0; i < n0; i++) t[i] = c[brev(n, i)];
  bf_B<L>(t, 0, 1, 16, true);
  idct_stage<L, 2>(t);
  if constexpr (n >= 3) idct_stage<L, 3>(t);
  if constexpr (n >= 4) idct_stage<L, 4>(t);
  if constexpr (n >= 5) idct_stage<L, 5>(t);
}

template <class L, int n>
DEV void adst_in_perm(typename L::T *t) {
  constexpr int n0 = 1 << n, n1 = n0 >> 1;
  typename L::T c[n0];
#pragma unroll
  for (int i = 0; i < n0; i++) c[i] = t[i];
#pragma unroll
  for (int i = 0; i < n1; i++) {
    t[2 * i] = c[n0 - 1 - 2 * i];
    t[2 * i + 1] = c[2 * i];
  }
}

template <class L, int n>
DEV void adst_out_perm(typename L::T *t) {
  constexpr int n0 = 1 << n;
  typename L::T c[n0];
#pragma unroll
  for (int i = 0; i < n0; i++) c[i] = t[i];
  if constexpr (n == 4) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int cc = 0; cc < 2; cc++)
#pragma unroll
          for (int d = 0; d < 2; d++)
            t[8 * a + 4 * b + 2 * cc + d] = c[8 * (d ^ cc) + 4 * (cc ^ b) + 2 * (b ^ a) + a];
  } else {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int cc = 0; cc < 2; cc++) t[4 * a + 2 * b + cc] = c[4 * (cc ^ b) + 2 * (b ^ a) + a];
  }
}

template <class L>
DEV void iadst4(typename L::T *t) {
  typedef typename L::T T;
  T s0 = L::mul(t[0], kSinpi19), s1 = L::mul(t[0], kSinpi29), s2 = L::mul(t[1], kSinpi39);
  T s3 = L::mul(t[2], kSinpi49), s4 = L::mul(t[2], kSinpi19), s5 = L::mul(t[3], kSinpi29);
  T s6 = L::mul(t[3], kSinpi49);
  T v = L::add(L::sub(t[0], t[2]), t[3]);
  T s7 = L::mul(v, kSinpi39);
  T x0 = L::add(L::add(s0, s3), s5), x1 = L::sub(L::sub(s1, s4), s6), x2 = s7, x3 = s2;
  t[0] = L::rnd(L::add(x0, x3), 14);
  t[1] = L::rnd(L::add(x1, x3), 14);
  t[2] = L::rnd(x2, 14);
  t[3] = L::rnd(L::sub(L::add(x0, x1), x3), 14);
}

template <class L>
DEV void iadst8(typename L::T *t) {
  typename L::T s[8];
  adst_in_perm<L, 3>(t);
#pragma unroll
  for (int i = 0; i < 4; i++) bf_SB<L>(t, s, 2 * i, 1 + 2 * i, 30 - 8 * i, true);
#pragma unroll
  for (int i = 0; i < 4; i++) bf_SH<L>(t, s, i, 4 + i);
#pragma unroll
  for (int i = 0; i < 2; i++) bf_SB<L>(t, s, 4 + 3 * i, 5 + i, 24 - 16 * i, true);
#pragma unroll
  for (int i = 0; i < 2; i++) bf_SH<L>(t, s, 4 + i, 6 + i);
#pragma unroll
  for (int i = 0; i < 2; i++) bf_H<L>(t, i, 2 + i, false);
#pragma unroll
  for (int i = 0; i < 2; i++) bf_B<L>(t, 2 + 4 * i, 3 + 4 * i, 16, true);
  adst_out_perm<L, 3>(t);
#pragma unroll
  for (int i = 0; i < 4; i++) t[1 + 2 * i] = L::sub(0, t[1 + 2 * i]);
}

template <class L>
DEV void iadst16(typename L::T *t) {
  typename L::T s[16];
  adst_in_perm<L, 4>(t);
#pragma unroll
  for (int i = 0; i < 8; i++) bf_SB<L>(t, s, 2 * i, 1 + 2 * i, 31 - 4 * i, true);
#pragma unroll
  for (int i = 0; i < 8; i++) bf_SH<L>(t, s, i, 8 + i);
#pragma unroll
  for (int i = 0; i < 4; i++) bf_SB<L>(t, s, 8 + 2 * i, 9 + 2 * i, 28 - 16 * i, true);
#pragma unroll
  for (int i = 0; i < 4; i++) bf_SH<L>(t, s, 8 + i, 12 + i);
#pragma unroll
  for (int i = 0; i < 4; i++) bf_H<L>(t, i, 4 + i, false);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) bf_SB<L>(t, s, 4 + 8 * i + 3 * j, 5 + 8 * i + j, 24 - 16 * j, true);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) bf_SH<L>(t, s, 4 + 8 * j + i, 6 + 8 * j + i);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) bf_H<L>(t, 8 * j + i, 2 + 8 * j + i, false);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) bf_B<L>(t, 2 + 4 * j + 8 * i, 3 + 4 * j + 8 * i, 48 + 64 * (i ^ j), false);
  adst_out_perm<L, 4>(t);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) t[1 + 12 * j + 2 * i] = L::sub(0, t[1 + 12 * j + 2 * i]);
}

// the inverse ADST process (8.7.1.9); a 32-point request transforms the
// first 16 values only (itx.rs:463-469)
template <class L, int n>
DEV void iadst(typename L::T *t) {
  if constexpr (n == 2)
    iadst4<L>(t);
  else if constexpr (n == 3)
    iadst8<L>(t);
  else
    iadst16<L>(t);
}

// inverse WHT (8.7.1.10)
template <class L>
DEV void iwht(typename L::T *t, int shift) {
  typedef typename L::T T;
  T a = L::sar(t[0], shift), c = L::sar(t[1], shift), d = L::sar(t[2], shift), b = L::sar(t[3], shift);
  a = L::add(a, c);
  d = L::sub(d, b);
  T e = L::sar(L::sub(a, d), 1);
  b = L::sub(e, b);
  c = L::sub(e, c);
  a = L::sub(a, b);
  d = L::add(d, c);
  t[0] = a;
  t[1] = b;
  t[2] = c;
  t[3] = d;
}

template <class L, int n>
DEV void tx1d(typename L::T *t, bool dct, bool lossless, int wht_shift) {
  if constexpr (n == 2) {
    if (lossless) {
      iwht<L>(t, wht_shift);
      return;
    }
  }
  if (dct)
    idct<L, n>(t);
  else
    iadst<L, n>(t);
}

// ---- samples

template <int BD8>
struct Pix;
template <>
struct Pix<1> {
  typedef uint8_t T;
};
template <>
struct Pix<0> {
  typedef uint16_t T;
};

HDEV int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
DEV int64_t lclamp(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// per-frame dequantisation factors: [segment][plane > 0][dc, ac]
struct QTable {
  int32_t q[8][2][2];
};


// ---- block-size lookups (10.2) as packed constants usable in device code

constexpr int kW8Log2[13] = {0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3};  // num_8x8_wide, log2
constexpr int kH8Log2[13] = {0, 0, 0, 0, 1, 0, 1, 2, 1, 2, 3, 2, 3};  // num_8x8_high, log2
constexpr int kW4Log2[13] = {0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4};  // num_4x4_wide, log2
constexpr int kH4Log2[13] = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4};  // num_4x4_high, log2
constexpr uint64_t pack3(const int (&v)[13]) {
  uint64_t k = 0;
  for (int i = 0; i < 13; i++) k |= (uint64_t)v[i] << (3 * i);
  return k;
}
constexpr uint64_t kW8P = pack3(kW8Log2), kH8P = pack3(kH8Log2), kW4P = pack3(kW4Log2),
                   kH4P = pack3(kH4Log2);
HDEV int num8x8_wide(int s) { return 1 << ((kW8P >> (3 * s)) & 7); }
HDEV int num8x8_high(int s) { return 1 << ((kH8P >> (3 * s)) & 7); }
HDEV int num4x4_wide(int s) { return 1 << ((kW4P >> (3 * s)) & 7); }
HDEV int num4x4_high(int s) { return 1 << ((kH4P >> (3 * s)) & 7); }

}  // namespace vp9g
