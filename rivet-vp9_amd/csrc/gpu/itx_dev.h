// The 2D inverse transform of one transform block by a team of n0 = 2^n
// lanes (one 1D transform per lane: rows, then columns through an LDS tile
// transpose), plus the dequantisation scatter.  Shared by the inter residual
// kernel (destination: the frame in HBM) and the intra wavefront kernel
// (destination: the superblock staged in LDS).
//
// Follows inverse_2d_add (src/dsp/itx.rs:519-570), the DC-only shortcut
// (itx.rs:689-694, 729-733) and reconstruct's dequantisation
// (src/decoder/recon.rs:233-281).
#pragma once

#include "dev_common.h"
#include "vp9_consts.h"

namespace vp9g {

// dc_only_value: the constant every sample of a DC-only DCT_DCT block gets
// at 8 bits (32-bit wrapping arithmetic).
DEV int32_t dc_only_value(int32_t dc, int n) {
  const int32_t c = cos64(16);
  int32_t v = Lane32::rnd(Lane32::mul(dc, c), 14);
  int32_t w = Lane32::rnd(Lane32::mul(v, c), 14);
  return Lane32::rnd(w, n + 2 < 6 ? n + 2 : 6);
}

// The dequantised value of quantised coefficient v at raster position pos.
DEV int32_t dequant(int32_t v, int pos, int32_t dcq, int32_t acq, bool half) {
  int64_t d = (int64_t)v * (pos == 0 ? dcq : acq);
  if (half) d /= 2;  // 32x32: truncating division
  return (int32_t)lclamp(d, INT32_MIN, INT32_MAX);
}

// Adds residual r to sample (i, j) of a destination, clamped to [0, max].
template <class P>
struct GlobalDst {
  P *base;
  uint32_t stride;
  DEV void add(int i, int j, int r, int max) {
    P *p = base + (size_t)i * stride + j;
    *p = (P)iclamp((int)*p + r, 0, max);
  }
};

// The inter residual kernel's destination: added into the frame, or, for a
// transform block of an intra block (res set), stored as int16.  Beyond
// +-32767 the clip to the sample range after the prediction is added gives 0
// or the maximum whatever the prediction, as it does for the full value.
template <class P>
struct ItxDst {
  P *base;
  int16_t *res;
  uint32_t stride;
  DEV void add(int i, int j, int r, int max) {
    if (res) {
      res[(size_t)i * stride + j] = (int16_t)iclamp(r, -32768, 32767);
    } else {
      P *p = base + (size_t)i * stride + j;
      *p = (P)iclamp((int)*p + r, 0, max);
    }
  }
};

// Rows, then columns, then Round2 by min(n + 2, 6) (none when lossless),
// residual clamp +-2^16, add.  `tile` holds the dequantised n0 x n0 block
// (row stride n0 + 1).  Every lane of the workgroup calls this (lanes with
// lane >= n0 idle) since it holds a barrier.  dc_only: the 8-bit shortcut,
// adding dc_v to every sample.
template <class L, int n, class Dst>
DEV void itx_2d(typename L::T *tile, int lane, int tx_type, bool lossless, bool dc_only, int32_t dc_v,
                Dst &dst, int max) {
  typedef typename L::T T;
  constexpr int n0 = 1 << n, S = n0 + 1;
  T t[n0];
  if (lane < n0 && !dc_only) {
    T *row = tile + lane * S;
    bool zero = true;
#pragma unroll
    for (int k = 0; k < n0; k++) {
      t[k] = row[k];
      zero = zero && (t[k] == 0);
    }
    if (!zero || lossless) {
      tx1d<L, n>(t, (tx_type & 2) == 0, lossless, 2);
#pragma unroll
      for (int k = 0; k < n0; k++) row[k] = t[k];
    }
  }
  __syncthreads();
  if (lane < n0) {
    if (dc_only) {
      int v = iclamp(dc_v, -65536, 65536);
#pragma unroll
      for (int i = 0; i < n0; i++) dst.add(i, lane, v, 255);
    } else {
#pragma unroll
      for (int i = 0; i < n0; i++) t[i] = tile[i * S + lane];
      tx1d<L, n>(t, (tx_type & 1) == 0, lossless, 0);
      constexpr int shift = n + 2 < 6 ? n + 2 : 6;
#pragma unroll
      for (int i = 0; i < n0; i++) {
        T r = lossless ? t[i] : L::rnd(t[i], shift);
        dst.add(i, lane, (int)lclamp((int64_t)r, -65536, 65536), max);
      }
    }
  }
}

}  // namespace vp9g
