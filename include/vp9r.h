/* vp9r.h: the C ABI of the MI355X reconstruction backend for rivet-vp9's VP9
 * decoder.
 *
 * Three layers, all plain C types and pointers:
 *   1. the decoder API, mirroring the crate's public surface
 *      (Decoder::new/set_threads/threads/set_max_pixels/decode/decode_all/
 *      frames_decoded and Frame, src/decoder/mod.rs:150-216, src/frame.rs);
 *   2. the per-frame seam: the CPU parser's record (blocks, transform blocks,
 *      quantised coefficients, mode info; src/decoder/block.rs:130-152) and
 *      the GPU reconstruction of one record into an HBM-resident frame;
 *   3. the per-kernel seam mirroring src/dsp's scalar kernel signatures, as
 *      batched launches over device buffers (and one-call host forms for
 *      tests).
 *
 * Status codes map onto the crate's Error kinds (src/lib.rs:67-81).
 * The library fails loudly (VP9R_ERR_DEVICE) when no GPU is usable: there is
 * no CPU fallback behind this ABI.
 */
#ifndef VP9R_H
#define VP9R_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status (src/lib.rs:67-81: Error::{Bitstream, Unsupported, InvalidInput}) */

enum {
  VP9R_OK = 0,
  VP9R_ERR_BITSTREAM = 1,
  VP9R_ERR_UNSUPPORTED = 2,
  VP9R_ERR_INVALID_INPUT = 3,
  VP9R_ERR_DEVICE = 4, /* HIP error, or no GPU */
  VP9R_AGAIN = 5,      /* vp9r_server_poll: no frame within the time given */
};

/* The message of the last failure on this thread ("" if none). */
const char *vp9r_last_error(void);

/* ------------------------------------------------------------------------ */
/* The per-frame record (what the CPU parser emits, what recon consumes).   */

/* Header fields reconstruction reads (src/header.rs:88-146, 16-56). */
typedef struct vp9r_frame_hdr {
  uint32_t width, height;             /* FrameWidth, FrameHeight */
  uint32_t mi_cols, mi_rows;          /* 8x8 units */
  uint32_t sb64_cols, sb64_rows;      /* 64x64 units */
  uint32_t render_width, render_height;
  uint8_t profile, bit_depth, ss_x, ss_y;
  uint8_t frame_type, show_frame, show_existing_frame, frame_to_show_map_idx;
  uint8_t intra_only, frame_is_intra, error_resilient_mode, lossless;
  uint8_t refresh_frame_flags, legacy_uv, interp_filter, tx_mode;
  uint8_t ref_frame_idx[3], reference_mode;
  uint8_t color_space, color_range, tile_cols_log2, tile_rows_log2;
  int32_t base_q_idx, delta_q_y_dc, delta_q_uv_dc, delta_q_uv_ac;
  /* loop filter (6.2.8) */
  uint8_t lf_level, lf_sharpness, lf_delta_enabled, _pad0;
  int8_t lf_ref_deltas[4];
  int8_t lf_mode_deltas[2];
  /* segmentation (6.2.11): the parts reconstruction reads */
  uint8_t seg_enabled, seg_abs_or_delta_update;
  uint8_t seg_feature_enabled[8][4];
  int16_t seg_feature_data[8][4];
  /* sizes of the active LAST/GOLDEN/ALTREF references (inter frames) */
  uint32_t ref_width[3], ref_height[3];
} vp9r_frame_hdr;

/* One coded block as the residual syntax sees it (block.rs:51-76). */
typedef struct vp9r_block {
  uint16_t mi_row, mi_col;
  uint8_t mi_size, skip, tx_size, is_inter;
  uint8_t y_mode, uv_mode, interp_filter, segment_id;
  uint8_t sub_modes[4];
  int8_t ref_frame[2];
  uint8_t avail_u, avail_l;
  int32_t mv[2][4][2]; /* block_mvs[list][sub-block][row, col], 1/8 sample */
} vp9r_block;

/* One transform block in reconstruction order: every transform block of an
 * intra block that starts inside the frame, and every transform block of an
 * inter block with eob > 0 (block.rs:1314-1364). */
typedef struct vp9r_txb {
  uint16_t x, y;        /* plane sample position of the top-left */
  uint8_t plane, tx_size, tx_type, _pad;
  uint16_t eob;         /* 0: prediction only */
  uint16_t ncoef;       /* nonzero quantised coefficients at coef_start */
  uint32_t coef_start;
  uint32_t block;       /* index of the owning vp9r_block */
} vp9r_txb;

/* What the loop filter reads of each 8x8 position (block.rs:19-31). */
typedef struct vp9r_mi {
  uint8_t mi_size, skip, tx_size, y_mode;
  uint8_t segment_id;
  int8_t ref_frame0;
  uint8_t _pad[2];
} vp9r_mi;

typedef struct vp9r_record {
  vp9r_frame_hdr hdr;
  const vp9r_block *blocks;
  uint32_t n_blocks;
  const vp9r_txb *txbs;
  uint32_t n_txbs;
  const uint16_t *coef_pos; /* raster position in the n x n block */
  const int32_t *coef_val;  /* quantised value (dequantised by recon) */
  uint32_t n_coefs;
  const vp9r_mi *mi;        /* mi_rows * mi_cols, row-major */
} vp9r_record;

/* ------------------------------------------------------------------------ */
/* Containers (src/superframe.rs:6-52, src/ivf.rs).                          */

/* Splits a packet into frames (Annex B superframe index).  Writes up to
 * `max` (offset, size) pairs; returns the frame count. */
size_t vp9r_superframe_split(const uint8_t *packet, size_t len, size_t *offsets, size_t *sizes,
                             size_t max);

/* ------------------------------------------------------------------------ */
/* CPU parser (headers, bool decoder, probabilities, block syntax, motion   */
/* vector prediction, tokens).  Stays on the CPU; touches no GPU.           */

typedef struct vp9r_parser vp9r_parser;

vp9r_parser *vp9r_parser_new(void);
void vp9r_parser_free(vp9r_parser *p);
/* Parses one frame (not a superframe).  On VP9R_OK `out` points into the
 * parser's buffers, valid until the next call.  A show_existing_frame frame
 * gives hdr.show_existing_frame = 1 and no blocks. */
int vp9r_parser_parse_frame(vp9r_parser *p, const uint8_t *data, size_t len, vp9r_record *out);
/* decoder/mod.rs:183-188 */
void vp9r_parser_set_max_pixels(vp9r_parser *p, uint64_t pixels);
/* Threads that parse a frame's tile columns at once (decoder/mod.rs:171-181
 * set_threads, as decode_tiles_parallel uses them, block.rs:1769-1866);
 * 0 or 1: the calling thread only.  The record is the same with any number. */
void vp9r_parser_set_threads(vp9r_parser *p, uint32_t threads);

/* ------------------------------------------------------------------------ */
/* GPU reconstruction (one device per context).                              */

typedef struct vp9r_gpu vp9r_gpu;
typedef struct vp9r_gpu_frame vp9r_gpu_frame; /* an HBM-resident frame */

/* Creates a context on HIP device `device`.  Fails with VP9R_ERR_DEVICE when
 * there is no usable GPU. */
int vp9r_gpu_new(int device, vp9r_gpu **out);
/* Releases the caller's context.  Frame stores and device buffers made from
 * it keep it alive until they are freed too, so objects may be freed in any
 * order (the context is destroyed with the last of them). */
void vp9r_gpu_free(vp9r_gpu *g);

/* A frame store of the record's geometry (planes of whole superblocks:
 * decoder/mod.rs:27-61), u8 samples at 8 bits, u16 above. */
int vp9r_gpu_frame_new(vp9r_gpu *g, uint32_t width, uint32_t height, uint32_t bit_depth,
                       uint32_t ss_x, uint32_t ss_y, vp9r_gpu_frame **out);
void vp9r_gpu_frame_free(vp9r_gpu_frame *f);
/* Uploads / downloads whole planes (row-major, stride = plane stride, u16
 * samples regardless of bit depth) for tests. */
int vp9r_gpu_frame_upload_u16(vp9r_gpu_frame *f, int plane, const uint16_t *src);
int vp9r_gpu_frame_download_u16(vp9r_gpu_frame *f, int plane, uint16_t *dst);
void vp9r_gpu_frame_geometry(const vp9r_gpu_frame *f, int plane, uint32_t *stride, uint32_t *rows);

/* Reconstructs the record into `cur`: inter prediction, dequantisation and
 * inverse transforms, the intra wavefront (the replay of
 * src/decoder/block.rs:1405-1498 with src/decoder/recon.rs).  refs[i] is the
 * LAST/GOLDEN/ALTREF frame (NULL for intra frames).  Asynchronous on the
 * context's stream; the record's host buffers may be reused on return. */
int vp9r_gpu_reconstruct(vp9r_gpu *g, const vp9r_record *rec, vp9r_gpu_frame *const refs[3],
                         vp9r_gpu_frame *cur);
/* The loop filter of the reconstructed frame (loopfilter::filter_frame,
 * src/decoder/loopfilter.rs:51-81) from the record's mode info and header.
 * `rec` is the record last given to vp9r_gpu_reconstruct on this context.
 * Asynchronous, like vp9r_gpu_reconstruct. */
int vp9r_gpu_loop_filter(vp9r_gpu *g, const vp9r_record *rec, vp9r_gpu_frame *cur);

/* The same two steps split at the host/device boundary: vp9r_gpu_prepare
 * checks a record, forms its work lists and copies everything to HBM (and
 * waits); vp9r_gpu_run then reconstructs and loop filters the frame from
 * those HBM-resident inputs alone, asynchronously.  A prepared record can
 * be run any number of times, with reference frames of the sizes it names. */
typedef struct vp9r_prepared vp9r_prepared;
int vp9r_gpu_prepare(vp9r_gpu *g, const vp9r_record *rec, vp9r_prepared **out);
void vp9r_prepared_free(vp9r_prepared *p);
int vp9r_gpu_run(vp9r_gpu *g, const vp9r_prepared *p, vp9r_gpu_frame *const refs[3],
                 vp9r_gpu_frame *cur);
/* Runs n prepared records of independent frames (of any sizes and formats)
 * together: each launch of the schedule covers every frame of the batch,
 * so n frames cost one launch sequence instead of n.  refs holds 3 frame
 * stores per record (refs[3 * i + k]; unused for intra frames), cur one.
 * No frame of the batch may be a reference or the output of another frame
 * in it; the results are those of n vp9r_gpu_run calls.  Frames that depend
 * on each other go into successive batches.  This is the cross-stream
 * batching of a multi-stream server (one ready frame per stream); the crate
 * decodes one frame at a time (src/decoder/mod.rs:218-422) and has no
 * counterpart. */
int vp9r_gpu_run_batch(vp9r_gpu *g, uint32_t n, const vp9r_prepared *const *p,
                       vp9r_gpu_frame *const *refs, vp9r_gpu_frame *const *cur);
/* The loop filter alone (loopfilter::filter_frame, loopfilter.rs:51-81) of n
 * frames, each launch covering every frame: cur[i] must hold the
 * reconstruction of record p[i] (by vp9r_gpu_run with a filter level of 0,
 * or an earlier copy).  Waits for completion; the device time of its
 * launches is read with vp9r_gpu_seam_ms.  Not while stage timing is on.
 * Batches the single-frame vp9r_gpu_loop_filter. */
int vp9r_gpu_filter_batch(vp9r_gpu *g, uint32_t n, const vp9r_prepared *const *p,
                          vp9r_gpu_frame *const *cur);
/* A prepared frame's launch sequence captured as a HIP graph, replayed by
 * one hipGraphLaunch (no counterpart in the reference, whose decoder calls
 * its kernels directly).  The graph is made for the frame stores given (its
 * batch tables and kernel parameters name them); vp9r_gpu_graph_update
 * re-targets it to other stores of the same geometry by capturing again and
 * updating the instantiated graph's kernel parameters in place.  Runs on the
 * context's stream like vp9r_gpu_run, with the same results; the prepared
 * record and the stores must outlive the graph's last run.  Stage timing
 * must be off. */
typedef struct vp9r_graph vp9r_graph;
int vp9r_gpu_graph_new(vp9r_gpu *g, const vp9r_prepared *p, vp9r_gpu_frame *const refs[3], vp9r_gpu_frame *cur,
                       vp9r_graph **out);
int vp9r_gpu_graph_update(vp9r_graph *gr, vp9r_gpu_frame *const refs[3], vp9r_gpu_frame *cur);
int vp9r_gpu_graph_run(vp9r_graph *gr);
/* The kernel launches the graph holds. */
uint32_t vp9r_gpu_graph_nodes(const vp9r_graph *gr);
void vp9r_gpu_graph_free(vp9r_graph *gr);
/* Work of a prepared record: [0] inter prediction tiles (<= 16x16), [1..4]
 * inter transform blocks of 4x4 .. 32x32, [5] intra blocks, [6] intra
 * dependency levels (one launch each), [7] loop filter wavefront launches,
 * [8..11] algorithmic HBM bytes of inter prediction (reference footprints
 * read + prediction written), inter residual (n^2 x (4 + 2b) per transform
 * block with coefficients), intra (edges + block + coefficients) and the loop
 * filter (2 x frame bytes), b = bytes per sample. */
int vp9r_gpu_prepared_stats(const vp9r_prepared *p, uint64_t out[12]);
/* The launch schedule of a prepared record run alone (vp9r_gpu_run): [0]
 * fused wavefront launches holding an intra level (with the loop filter
 * levels placed beside them), [1] loop filter launches after the last intra
 * level (two levels per launch, the default; [3] of them under
 * VP9R_LF=levels), [2] the loop filter levels, [3] the loop filter levels
 * after the last intra level.  Measurement aid; no counterpart in the
 * reference (its filter runs after reconstruction, decoder/mod.rs:355-358). */
int vp9r_gpu_prepared_schedule(const vp9r_prepared *p, uint32_t out[4]);
/* Writes the visible frame in the crate's Frame::packed layout (Y, U, V; u8
 * at 8 bits, little-endian u16 above; frame.rs:112-229) to `dst`, which
 * holds vp9r_packed_size() bytes.  Waits for the frame. */
int vp9r_gpu_download_packed(vp9r_gpu *g, vp9r_gpu_frame *f, uint8_t *dst);
/* The same download queued on the context's copy stream behind the work
 * queued so far, so that it overlaps the reconstruction of the next frames.
 * `dst` (best pinned: vp9r_host_alloc) must stay valid until
 * vp9r_gpu_frame_wait(f) returns; later work that writes f waits for the
 * copy on the device. */
int vp9r_gpu_download_packed_async(vp9r_gpu *g, vp9r_gpu_frame *f, uint8_t *dst);
int vp9r_gpu_frame_wait(vp9r_gpu_frame *f);
/* Page-locked host memory for the asynchronous downloads. */
int vp9r_host_alloc(size_t bytes, void **out);
void vp9r_host_free(void *p);
size_t vp9r_packed_size(uint32_t width, uint32_t height, uint32_t bit_depth, uint32_t ss_x,
                        uint32_t ss_y);
int vp9r_gpu_synchronize(vp9r_gpu *g);
/* Per-stage device time, in ms (HIP events on the context's stream, where
 * the kernels run), summed over the frames reconstructed since timing was
 * enabled or last read: [0] inter prediction (with the record upload in
 * vp9r_gpu_reconstruct), [1] inter residual, [2] intra wavefront, [3] loop
 * filter, [4] whole frame.  *frames (may be NULL) gets the frame count.
 * Waits for the last timed frame.  Read it after a frame's loop filter. */
void vp9r_gpu_set_timing(vp9r_gpu *g, int enable);
int vp9r_gpu_stage_times(vp9r_gpu *g, float out[5], uint32_t *frames);

/* ------------------------------------------------------------------------ */
/* Per-kernel seam: src/dsp signatures as batched device launches.          */

/* Dequantise-free inverse transform + add (itx::inverse_transform_add,
 * src/dsp/itx.rs:701-749), batched: block i has its n x n dequantised
 * coefficients at coefs + coef_off[i] (row-major, int32), size n = 4 << tx[i],
 * tx_type[i], eob[i]; its destination at dst + dst_off[i] samples (u8 when
 * bit_depth is 8, else u16) with row stride `stride`.  `tx` is a host array
 * (the size classes are formed on the host); every other pointer is a device
 * pointer.  lossless selects the WHT and needs tx[i] == 0.  Runs on the
 * context's stream. */
int vp9r_itx_add_batch(vp9r_gpu *g, uint32_t count, const int32_t *coefs, const uint32_t *coef_off,
                       const uint8_t *tx, const uint8_t *tx_type, const uint16_t *eob,
                       const uint64_t *dst_off, void *dst, uint32_t stride, uint32_t bit_depth,
                       int lossless);

/* Motion compensation (inter::predict, src/dsp/inter.rs:289-395, and
 * pixel::avg, src/dsp/pixel.rs:22-50), batched.  Block i predicts w x h
 * samples (w, h in 4..64) at 1/16-sample position (x, y) with steps
 * (xs, ys) in 1..32 from reference plane ref_idx of the `refs` array, and,
 * when ref2_idx >= 0, averages with a second prediction at (x2, y2) from
 * plane ref2_idx.  Writes to out + out_off[i] samples with stride
 * out_stride.  `desc` is a host array (the wrapper expands it into tiles of
 * 16 columns by up to 32 rows); `out` is a device pointer. */
typedef struct vp9r_mc_desc {
  int32_t x, y, x2, y2;
  int16_t xs, ys;
  uint8_t w, h;
  uint8_t filter;
  int8_t ref_idx, ref2_idx;
  uint8_t _pad;
  uint64_t out_off;
} vp9r_mc_desc;
/* One reference plane (inter.rs:21-27 RefPlane: data, stride, last_x,
 * last_y): a device pointer to samples (u8 at 8 bits, else u16), its row
 * stride in samples, and its bottom-right sample (the reference's visible
 * size minus one, recon.rs:169-170). */
typedef struct vp9r_ref_plane {
  const void *data;
  uint32_t stride;
  int32_t last_x, last_y;
  uint32_t _pad;
} vp9r_ref_plane;
/* Each reference with its own plane geometry: `planes` is a host array of
 * n_refs planes. */
int vp9r_mc_batch_planes(vp9r_gpu *g, uint32_t count, const vp9r_mc_desc *desc, const vp9r_ref_plane *planes,
                         uint32_t n_refs, void *out, uint32_t out_stride, uint32_t bit_depth);
/* vp9r_mc_batch_planes without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The descriptors and planes are
 * staged in a fixed device array of the library (about 2 MB per GPU, held
 * until the process ends): at most 65536 descriptors and 64 planes
 * (VP9R_ERR_INVALID_INPUT beyond).  The calls on one GPU share that array:
 * calls that overlap on one GPU race.  The same output as
 * vp9r_mc_batch_planes; waits for completion, and is not timed
 * (vp9r_gpu_seam_ms). */
int vp9r_mc_batch_planes_noalloc(vp9r_gpu *g, uint32_t count, const vp9r_mc_desc *desc,
                                 const vp9r_ref_plane *planes, uint32_t n_refs, void *out, uint32_t out_stride,
                                 uint32_t bit_depth);
/* The same with every reference plane of one geometry: `refs` is a host
 * array of n_refs device pointers, all of row stride ref_stride and
 * bottom-right sample (last_x, last_y). */
int vp9r_mc_batch(vp9r_gpu *g, uint32_t count, const vp9r_mc_desc *desc,
                  const void *const *refs, uint32_t n_refs, uint32_t ref_stride, int32_t last_x,
                  int32_t last_y, void *out, uint32_t out_stride, uint32_t bit_depth);

/* pixel::avg (src/dsp/pixel.rs:22-50), batched: block i writes
 * (a + b + 1) >> 1 over w x h samples to dst + dst_off (row stride
 * `stride`), a and b at a + a_off and b + b_off with row stride w.  `desc`
 * is a host array; a, b and dst are device pointers to samples (u8 at 8
 * bits, else u16). */
typedef struct vp9r_avg_desc {
  uint64_t a_off, b_off, dst_off;
  uint8_t w, h, _pad[6];
} vp9r_avg_desc;
int vp9r_avg_batch(vp9r_gpu *g, uint32_t count, const vp9r_avg_desc *desc, const void *a, const void *b,
                   void *dst, uint32_t stride, uint32_t bit_depth);
/* pixel::add_const (src/dsp/pixel.rs:9-16), batched: v (clamped to
 * +-65536) added to the n x n samples at dst + dst_off (u8, row stride
 * `stride`), each clamped to 0..255; 8 bits only, as the reference. */
typedef struct vp9r_add_const_desc {
  uint64_t dst_off;
  int32_t v;
  uint8_t n, _pad[3];
} vp9r_add_const_desc;
int vp9r_add_const_batch(vp9r_gpu *g, uint32_t count, const vp9r_add_const_desc *desc, void *dst,
                         uint32_t stride);

/* intra::predict (src/dsp/intra.rs:19-29), batched: block i predicts mode
 * `mode` at size 1 << log2 (log2 2..5) from the int32 edges at
 * edges + edge_off (aboveRow[-1 .. 2 size - 1], then leftCol[0 .. size -
 * 1]: intra.rs's above and left) into dst + dst_off (row stride `stride`).
 * `desc` is a host array; edges and dst are device pointers (dst: u8 at 8
 * bits, else u16). */
typedef struct vp9r_intra_desc {
  uint64_t dst_off;
  uint32_t edge_off;
  uint8_t mode, log2, have_left, have_above;
} vp9r_intra_desc;
int vp9r_intra_predict_batch(vp9r_gpu *g, uint32_t count, const vp9r_intra_desc *desc, const int32_t *edges,
                             void *dst, uint32_t stride, uint32_t bit_depth);

/* Loop filter sample kernels (lf::filter_edges, src/dsp/lf.rs:178-218) for a
 * sequence of superblock passes over one plane: pass i filters edges
 * [first[i], n_edges[i]) (n_runs[i] runs) of the region at (x0[i], y0[i])
 * with the Edges tables at edges + 1024 * i (fs[16][16] int8, -1 = none,
 * then limit, blimit, thresh [16][16] uint8, unshifted).  Passes run in
 * order, as the reference calls them, in one launch.  `buf` is a device
 * plane of `rows` rows of `stride` samples (u8 at 8 bits, else u16); every
 * pass's reach (8 samples past its last edge) must lie inside it.  The
 * other arrays are host arrays; `first` may be NULL (every pass from edge
 * 0). */
int vp9r_lf_edges_batch(vp9r_gpu *g, uint32_t count, const uint8_t *edges, const uint16_t *x0,
                        const uint16_t *y0, const uint8_t *vertical, const uint8_t *first,
                        const uint8_t *n_edges, const uint8_t *n_runs, void *buf, uint32_t stride,
                        uint32_t rows, uint32_t bit_depth);

/* Autotuner of the batched forms' launch shapes (no counterpart in the
 * reference: src/dsp/mod.rs:32-96 chooses its kernels by CPU feature).  It
 * times, on synthetic batches, every transform launch shape of each size
 * (4x4 .. 32x32) and bit depth class (8-bit, 10/12-bit): 1, 2 or 4 waves per
 * workgroup, the row-to-column transposes through LDS or in registers
 * (DPP / ds_swizzle lane exchanges, 4x4 .. 16x16), and one workgroup per
 * 64 W / n blocks or persistent (resident) workgroups looping over them;
 * and tiles of 16 and 32 rows for motion compensation per bit depth class.
 * It keeps the fastest of each for later calls on this context (the
 * arithmetic, and so every output, is the same for every shape).  times_us
 * (may be NULL) gets VP9R_TUNE_TIMES candidate times, -1 for a shape a size
 * lacks: transforms [class][size][persistent 0/1][registers 0/1][waves 1, 2,
 * 4] (96), then motion compensation [class][rows 32, 16].  Takes under a
 * second; allocates its own inputs. */
#define VP9R_TUNE_TIMES 100
int vp9r_gpu_autotune(vp9r_gpu *g, float times_us[VP9R_TUNE_TIMES]);
/* The shapes in use: [0..3] the transform shape at 8 bits for 4x4 ..
 * 32x32, [4..7] the same at 10/12 bits, [8..9] motion compensation tile
 * rows at 8 and at 10/12 bits, [10] 1 if vp9r_gpu_autotune has run.  A
 * transform shape is its waves per workgroup (1, 2 or 4) + 16 for register
 * transposes (4x4 .. 16x16 only) + 32 for persistent workgroups.
 * vp9r_gpu_set_tuning sets [0..9]. */
int vp9r_gpu_tuning(const vp9r_gpu *g, uint32_t out[11]);
int vp9r_gpu_set_tuning(vp9r_gpu *g, const uint32_t in[10]);

/* Device memory for callers of the batched forms (no counterpart in the
 * reference: a plain-pointer ABI needs its own allocation and copies).  The
 * copies are ordered on the context's stream and wait for completion. */
int vp9r_gpu_malloc(vp9r_gpu *g, size_t bytes, void **out);
void vp9r_gpu_mfree(vp9r_gpu *g, void *p);
int vp9r_gpu_to_device(vp9r_gpu *g, void *dst, const void *src, size_t bytes);
int vp9r_gpu_to_host(vp9r_gpu *g, void *dst, const void *src, size_t bytes);
/* The elapsed device time (ms) of the kernels of the last successful
 * vp9r_itx_add_batch, vp9r_mc_batch or vp9r_gpu_filter_batch call, between events recorded on the
 * context's stream around its launches (uploads excluded).  No counterpart in
 * the reference (src/dsp/bench.rs:9-20 times the scalar calls with the host
 * clock): a measurement aid for the kernel microbenchmarks. */
int vp9r_gpu_seam_ms(vp9r_gpu *g, float *ms);
/* The device address of a frame store's plane (for the batched forms). */
void *vp9r_gpu_frame_plane(vp9r_gpu_frame *f, int plane);

/* One-call host forms of the dsp kernels, for parity tests (each copies to
 * the device, launches, and copies back). */
int vp9r_itx_add_host(vp9r_gpu *g, const int32_t *coefs, uint32_t n, uint32_t tx_type, int lossless,
                      uint32_t eob, uint32_t bit_depth, uint16_t *dst, uint32_t stride);
int vp9r_intra_predict_host(vp9r_gpu *g, uint32_t mode, uint32_t log2, const int32_t *above,
                            const int32_t *left, int have_left, int have_above, uint32_t bit_depth,
                            uint16_t *dst, uint32_t stride);

/* ------------------------------------------------------------------------ */
/* Decoder API (src/decoder/mod.rs:150-216).                                 */

typedef struct vp9r_decoder vp9r_decoder;

/* A decoded, shown frame (frame.rs:112-137): packed Y, U, V. */
typedef struct vp9r_frame {
  uint32_t width, height, bit_depth;
  uint32_t ss_x, ss_y;
  uint32_t color_space, full_range, render_width, render_height;
  uint8_t *data; /* owned by the decoder, valid until the next decode call */
  size_t size;
  size_t plane_offset[3];
  uint32_t plane_width[3], plane_height[3];
} vp9r_frame;

int vp9r_decoder_new(int device, vp9r_decoder **out);
void vp9r_decoder_free(vp9r_decoder *d);
/* Threads for the CPU parse (tile columns in parallel, as the crate's
 * decode_tiles_parallel); 0 (the default) is one per hardware thread.  The
 * output is the same with any number (decoder/mod.rs:171-181). */
void vp9r_decoder_set_threads(vp9r_decoder *d, uint32_t threads);
uint32_t vp9r_decoder_threads(const vp9r_decoder *d);
void vp9r_decoder_set_max_pixels(vp9r_decoder *d, uint64_t pixels);
/* decode_all: decodes one packet; returns the number of shown frames in
 * *n_frames and their descriptions through vp9r_decoder_frame(). */
int vp9r_decoder_decode(vp9r_decoder *d, const uint8_t *packet, size_t len, uint32_t *n_frames);
int vp9r_decoder_frame(const vp9r_decoder *d, uint32_t i, vp9r_frame *out);
uint64_t vp9r_decoder_frames_decoded(const vp9r_decoder *d);
/* Test hook, no counterpart in the reference: the next n frame stores the
 * decoder asks for are refused (VP9R_ERR_DEVICE), to test the recovery of
 * a reconstruction that fails after its frame was parsed. */
void vp9r_decoder_test_fail_stores(vp9r_decoder *d, uint32_t n);

/* ------------------------------------------------------------------------ */
/* Multi-stream server: many independent streams decoded on one GPU.        */
/*                                                                          */
/* A stream is admitted when its HBM footprint fits the budget: 10 frame    */
/* stores (8 reference slots, the frame being written, one in flight) of   */
/* its declared largest frame, superblock-aligned planes.  Streams wait in  */
/* open order until they fit.  Parse worker threads parse submitted packets */
/* (one worker per stream at a time) and upload each frame's inputs; a      */
/* dispatcher thread puts the next ready frame of every stream into one     */
/* batch (vp9r_gpu_run_batch) while earlier batches run, and copies shown   */
/* frames to pinned host memory; a finisher thread hands them to            */
/* vp9r_server_poll with their latency.  Every stream's frames are those of */
/* decoding it alone.  The crate decodes one stream per Decoder             */
/* (src/decoder/mod.rs:150-216) and has no counterpart.                     */

typedef struct vp9r_server vp9r_server;

typedef struct vp9r_server_config {
  uint64_t hbm_budget;    /* bytes of frame stores; 0: 90% of the device's free memory */
  uint32_t parse_threads; /* 0: hardware threads - 2 (at least 1) */
  uint32_t max_batch;     /* frames per batch; 0: 64 */
  uint32_t max_inflight;  /* batches queued on the device at once; 0: 2 */
  uint32_t parse_ahead;   /* parsed frames per stream waiting for the device; 0: 4 */
} vp9r_server_config;

typedef struct vp9r_server_frame {
  uint32_t stream;   /* the id vp9r_server_open gave */
  uint32_t index;    /* shown-frame number within the stream */
  uint32_t eos;      /* 1: the stream is over and this record holds no frame */
  int32_t status;    /* with eos: VP9R_OK, or the error that ended the stream */
  double latency_ms; /* submit of the frame's packet to the frame complete in host memory */
  uint64_t handle;   /* for vp9r_server_release (0 with eos) */
  vp9r_frame frame;  /* packed Y, U, V; data valid until released */
} vp9r_server_frame;

int vp9r_server_new(int device, const vp9r_server_config *cfg, vp9r_server **out);
/* Stops the threads, waits for the device and frees everything (frames not
 * yet released included). */
void vp9r_server_free(vp9r_server *s);
/* Bytes of one frame store (superblock-aligned planes) and a stream's
 * footprint (10 stores). */
uint64_t vp9r_frame_store_bytes(uint32_t width, uint32_t height, uint32_t bit_depth, uint32_t ss_x,
                                uint32_t ss_y);
/* Opens a stream whose frames are at most max_width x max_height, of this
 * bit depth and subsampling (a frame outside them ends the stream with
 * VP9R_ERR_INVALID_INPUT).  Fails when its footprint exceeds the budget. */
int vp9r_server_open(vp9r_server *s, uint32_t max_width, uint32_t max_height, uint32_t bit_depth,
                     uint32_t ss_x, uint32_t ss_y, uint32_t *id);
/* Queues one packet (a frame or a superframe); the data is copied. */
int vp9r_server_submit(vp9r_server *s, uint32_t id, const uint8_t *packet, size_t len);
/* No more packets: once its frames are out, the stream ends (an eos record)
 * and its footprint is released. */
int vp9r_server_close(vp9r_server *s, uint32_t id);
/* The next finished frame or end of stream, waiting up to timeout_ms
 * (negative: no limit); VP9R_AGAIN when none came. */
int vp9r_server_poll(vp9r_server *s, vp9r_server_frame *out, int timeout_ms);
void vp9r_server_release(vp9r_server *s, uint64_t handle);
/* [0] frames shown, [1] frames decoded, [2] batches, [3] largest batch,
 * [4] streams admitted so far, [5] streams waiting for admission,
 * [6] frame-store bytes allocated, [7] the budget. */
int vp9r_server_stats(vp9r_server *s, uint64_t out[8]);

#ifdef __cplusplus
}
#endif

#endif
