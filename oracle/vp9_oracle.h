/* CPU oracle of the VP9 reconstruction path: a plain-C restatement of the
 * reference crate's scalar kernels and frame process.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * use it, as the checker.
 *
 * Each function names the reference file:line it follows (paths relative to
 * the safewords/rivet-vp9 checkout).  Samples are u16 at every bit depth, as
 * the reference stores them. */
#ifndef VP9_ORACLE_H
#define VP9_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "vp9r.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- dsp kernels ---- */

/* itx::inverse_transform_add (src/dsp/itx.rs:701-749) with its S32 (8-bit,
 * wrapping 32-bit) and S64 (10/12-bit) lanes (itx.rs:573-666); n_log2 2..5. */
void orc_itx_add(const int32_t *coefs, int n_log2, int tx_type, int lossless, int eob, int bit_depth,
                 uint16_t *dst, size_t stride);
/* The 64-bit specification process of 8.7.2 (the reference's test restatement,
 * itx.rs:776-1112), on the given 1D size: used to pin orc_itx_add. */
void orc_itx_spec_reconstruct(const int32_t *block, int n_log2, int tx_type, int lossless,
                              uint16_t *pred, int max);
/* 1D 64-bit spec transforms (in place on t[0..2^n)). */
void orc_spec_idct(int64_t *t, int n_log2);
void orc_spec_iadst(int64_t *t, int n_log2);

/* inter::predict (src/dsp/inter.rs:289-395): w x h samples at 1/16-sample
 * position (x, y), steps (x_step, y_step), from a plane whose bottom-right
 * sample is (last_x, last_y). */
void orc_inter_predict(const uint16_t *ref, size_t ref_stride, int last_x, int last_y, int x, int y,
                       int x_step, int y_step, int w, int h, int filter, int bit_depth,
                       uint16_t *out, size_t out_stride);
/* pixel::avg_scalar (src/dsp/pixel.rs:43-50) and pixel::add_const (9-16). */
void orc_avg(const uint16_t *a, const uint16_t *b, int w, int h, uint16_t *dst, size_t stride);
void orc_add_const(uint16_t *dst, size_t stride, int n, int v);

/* Batched forms for the kernel microbenchmark (the loops of
 * src/dsp/bench.rs:26-185 over many blocks).  orc_itx_add_batch: block i is
 * the dense n x n coefficients at coefs + coef_off[i], n = 4 << tx[i], added
 * to plane + dst_off[i].  orc_mc_batch: descriptor i (the layout of
 * vp9r_mc_desc) predicts into out + out_off, averaging with a second
 * prediction when ref2_idx >= 0. */
void orc_itx_add_batch(uint32_t count, const int32_t *coefs, const uint32_t *coef_off, const uint8_t *tx,
                       const uint8_t *tx_type, const uint16_t *eob, const uint64_t *dst_off, uint16_t *plane,
                       size_t stride, int bit_depth, int lossless);
typedef struct orc_mc_desc {
  int32_t x, y, x2, y2;
  int16_t xs, ys;
  uint8_t w, h, filter;
  int8_t ref_idx, ref2_idx;
  uint8_t pad;
  uint64_t out_off;
} orc_mc_desc;
void orc_mc_batch(uint32_t count, const orc_mc_desc *desc, const uint16_t *const *refs, size_t ref_stride,
                  int last_x, int last_y, uint16_t *out, size_t out_stride, int bit_depth);

/* intra::predict (src/dsp/intra.rs:19-175): above[0] is aboveRow[-1]. */
void orc_intra_predict(int mode, int log2, const int32_t *above, const int32_t *left,
                       int have_left, int have_above, int bit_depth, uint16_t *dst, size_t stride);

/* lf::filter (src/dsp/lf.rs:24-86) at buf[pos] across step. */
void orc_lf_filter(uint16_t *buf, ptrdiff_t pos, ptrdiff_t step, int filter_size, int limit,
                   int blimit, int thresh, int bit_depth);
/* lf::Edges (lf.rs:153-169): fs SKIP = -1. */
typedef struct orc_edges {
  int8_t fs[16][16];
  uint8_t limit[16][16], blimit[16][16], thresh[16][16];
} orc_edges;
/* lf::edges_scalar (lf.rs:221-259) */
void orc_lf_filter_edges(uint16_t *buf, size_t stride, int x0, int y0, int vertical, int first,
                         int n_edges, int n_runs, const orc_edges *e, int bit_depth);

/* ---- frames ---- */

typedef struct orc_frame {
  uint32_t width, height, bit_depth, ss_x, ss_y;
  uint32_t stride[3], rows[3];
  uint16_t *plane[3];
} orc_frame;

/* A zeroed frame store of whole superblocks (decoder/mod.rs:27-61). */
orc_frame *orc_frame_new(uint32_t width, uint32_t height, uint32_t bit_depth, uint32_t ss_x,
                         uint32_t ss_y);
void orc_frame_free(orc_frame *f);

/* The reconstruction of one record (block.rs:1405-1498 replay order;
 * recon.rs:17-281): prediction, dequantisation, inverse transforms.
 * refs[i]: the LAST/GOLDEN/ALTREF frames.  Returns 0, or VP9R_ERR_* on a
 * missing or badly scaled reference. */
int orc_reconstruct(const vp9r_record *rec, const orc_frame *const refs[3], orc_frame *cur);
/* loopfilter::filter_frame (src/decoder/loopfilter.rs:51-81), serial order. */
void orc_loop_filter(const vp9r_record *rec, orc_frame *cur);
/* The same with the tile columns on up to `threads` threads (the
 * reference's decode_tiles_parallel, block.rs:1769-1866), and the loop
 * filter with the superblock rows of each plane as a wavefront on up to
 * `threads` threads (filter_bands, loopfilter.rs:195-265).  Results equal
 * the serial forms'.  The all-core CPU baseline. */
int orc_reconstruct_mt(const vp9r_record *rec, const orc_frame *const refs[3], orc_frame *cur, int threads);
void orc_loop_filter_mt(const vp9r_record *rec, orc_frame *cur, int threads);
/* The edge decisions of one superblock pass (loopfilter.rs:466-620); returns
 * n_edges | n_runs << 8, writes the origin to x0/y0. */
int orc_lf_superblock(const vp9r_record *rec, int plane, int pass, uint32_t row, uint32_t col,
                      orc_edges *e, int *x0, int *y0);
/* RefFrame::to_frame (decoder/mod.rs:79-106): the packed Frame bytes. */
size_t orc_packed_size(const orc_frame *f);
void orc_pack(const orc_frame *f, uint8_t *dst);

#ifdef __cplusplus
}
#endif

#endif
