/* Symbolic constants, block-size lookups and decode trees of the VP9
 * bitstream specification (sections 3, 6.4, 7, 9.3.1, 10.2).  Spec data only:
 * the parser, the reconstruction host code, the HIP kernels and the CPU
 * oracle all read these.  The numeric probability/scan/filter tables are in
 * vp9_tables.h. */
#ifndef VP9R_CONSTS_H
#define VP9R_CONSTS_H
#include <stdint.h>

#ifdef __cplusplus
#define VP9R_CONST static constexpr
#else
#define VP9R_CONST static const
#endif

enum {
  KEY_FRAME = 0,
  /* block sizes */
  BLOCK_4X4 = 0, BLOCK_4X8, BLOCK_8X4, BLOCK_8X8, BLOCK_8X16, BLOCK_16X8, BLOCK_16X16,
  BLOCK_16X32, BLOCK_32X16, BLOCK_32X32, BLOCK_32X64, BLOCK_64X32, BLOCK_64X64,
  BLOCK_INVALID = 14,
  /* partitions */
  PARTITION_NONE = 0, PARTITION_HORZ, PARTITION_VERT, PARTITION_SPLIT,
  /* intra modes then inter modes: one numbering (y_mode) */
  DC_PRED = 0, V_PRED, H_PRED, D45_PRED, D135_PRED, D117_PRED, D153_PRED, D207_PRED,
  D63_PRED, TM_PRED, NEARESTMV, NEARMV, ZEROMV, NEWMV,
  /* transform sizes and modes */
  TX_4X4 = 0, TX_8X8, TX_16X16, TX_32X32,
  ONLY_4X4 = 0, ALLOW_32X32 = 3, TX_MODE_SELECT = 4,
  /* transform types: first word is the column (vertical) transform */
  DCT_DCT = 0, ADST_DCT = 1, DCT_ADST = 2, ADST_ADST = 3,
  /* interpolation filters */
  EIGHTTAP = 0, EIGHTTAP_SMOOTH = 1, EIGHTTAP_SHARP = 2, BILINEAR = 3, SWITCHABLE = 4,
  /* reference modes */
  SINGLE_REFERENCE = 0, COMPOUND_REFERENCE = 1, REFERENCE_MODE_SELECT = 2,
  /* motion vector joints */
  MV_JOINT_ZERO = 0, MV_JOINT_HNZVZ, MV_JOINT_HZVNZ, MV_JOINT_HNZVNZ,
  /* tokens */
  ZERO_TOKEN = 0, DCT_VAL_CAT6 = 10,
  /* segmentation features */
  SEG_LVL_ALT_Q = 0, SEG_LVL_ALT_L = 1, SEG_LVL_REF_FRAME = 2, SEG_LVL_SKIP = 3,
  MAX_SEGMENTS = 8,
};

/* reference frames; NONE sorts below INTRA_FRAME */
enum { REF_NONE = -1, INTRA_FRAME = 0, LAST_FRAME = 1, GOLDEN_FRAME = 2, ALTREF_FRAME = 3 };

enum {
  MV_BORDER = 128,
  INTERP_EXTEND = 4,
  BORDERINPIXELS = 160,
  COMPANDED_MVREF_THRESH = 8,
  MAX_LOOP_FILTER = 63,
  MIN_TILE_WIDTH_B64 = 4,
  MAX_TILE_WIDTH_B64 = 64,
};

/* inter mode contexts */
enum {
  BOTH_ZERO = 0, ZERO_PLUS_PREDICTED = 1, BOTH_PREDICTED = 2, NEW_PLUS_NON_INTRA = 3,
  BOTH_NEW = 4, INTRA_PLUS_NON_INTRA = 5, BOTH_INTRA = 6, INVALID_CASE = 9,
};

VP9R_CONST uint8_t VP9_COUNTER_TO_CONTEXT[19] = {
    BOTH_PREDICTED, NEW_PLUS_NON_INTRA, BOTH_NEW, ZERO_PLUS_PREDICTED, NEW_PLUS_NON_INTRA,
    INVALID_CASE, BOTH_ZERO, INVALID_CASE, INVALID_CASE, INTRA_PLUS_NON_INTRA,
    INTRA_PLUS_NON_INTRA, INVALID_CASE, INTRA_PLUS_NON_INTRA, INVALID_CASE, INVALID_CASE,
    INVALID_CASE, INVALID_CASE, INVALID_CASE, BOTH_INTRA};

VP9R_CONST uint8_t VP9_MAX_TXSIZE_LOOKUP[13] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3};
VP9R_CONST uint8_t VP9_TX_MODE_TO_BIGGEST_TX_SIZE[5] = {0, 1, 2, 3, 3};

/* ss_size_lookup[bsize][subsampling_x][subsampling_y] */
#define VP9R_I 14
VP9R_CONST uint8_t VP9_SS_SIZE_LOOKUP[13][2][2] = {
    {{BLOCK_4X4, VP9R_I}, {VP9R_I, VP9R_I}},
    {{BLOCK_4X8, BLOCK_4X4}, {VP9R_I, VP9R_I}},
    {{BLOCK_8X4, VP9R_I}, {BLOCK_4X4, VP9R_I}},
    {{BLOCK_8X8, BLOCK_8X4}, {BLOCK_4X8, BLOCK_4X4}},
    {{BLOCK_8X16, BLOCK_8X8}, {VP9R_I, BLOCK_4X8}},
    {{BLOCK_16X8, VP9R_I}, {BLOCK_8X8, BLOCK_8X4}},
    {{BLOCK_16X16, BLOCK_16X8}, {BLOCK_8X16, BLOCK_8X8}},
    {{BLOCK_16X32, BLOCK_16X16}, {VP9R_I, BLOCK_8X16}},
    {{BLOCK_32X16, VP9R_I}, {BLOCK_16X16, BLOCK_16X8}},
    {{BLOCK_32X32, BLOCK_32X16}, {BLOCK_16X32, BLOCK_16X16}},
    {{BLOCK_32X64, BLOCK_32X32}, {VP9R_I, BLOCK_16X32}},
    {{BLOCK_64X32, VP9R_I}, {BLOCK_32X32, BLOCK_32X16}},
    {{BLOCK_64X64, BLOCK_64X32}, {BLOCK_32X64, BLOCK_32X32}},
};

/* subsize_lookup[partition][bsize] */
VP9R_CONST uint8_t VP9_SUBSIZE_LOOKUP[4][13] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12},
    {VP9R_I, VP9R_I, VP9R_I, BLOCK_8X4, VP9R_I, VP9R_I, BLOCK_16X8, VP9R_I, VP9R_I, BLOCK_32X16,
     VP9R_I, VP9R_I, BLOCK_64X32},
    {VP9R_I, VP9R_I, VP9R_I, BLOCK_4X8, VP9R_I, VP9R_I, BLOCK_8X16, VP9R_I, VP9R_I, BLOCK_16X32,
     VP9R_I, VP9R_I, BLOCK_32X64},
    {VP9R_I, VP9R_I, VP9R_I, BLOCK_4X4, VP9R_I, VP9R_I, BLOCK_8X8, VP9R_I, VP9R_I, BLOCK_16X16,
     VP9R_I, VP9R_I, BLOCK_32X32},
};
#undef VP9R_I

/* mode2txfm_map: the transform type of an intra mode (inter modes: DCT_DCT) */
VP9R_CONST uint8_t VP9_MODE2TXFM_MAP[14] = {
    DCT_DCT, ADST_DCT, DCT_ADST, DCT_DCT, ADST_ADST, ADST_DCT, DCT_ADST,
    DCT_ADST, ADST_DCT, ADST_ADST, DCT_DCT, DCT_DCT, DCT_DCT, DCT_DCT};

/* extra_bits[token]: category, number of extra bits, base value */
VP9R_CONST int32_t VP9_EXTRA_BITS[11][3] = {
    {0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 0, 3}, {0, 0, 4}, {1, 1, 5},
    {2, 2, 7}, {3, 3, 11}, {4, 4, 19}, {5, 5, 35}, {6, 14, 67}};

/* cat_probs[cat] (first entries of each row) */
VP9R_CONST uint8_t VP9_CAT_PROBS[7][14] = {
    {0},
    {159},
    {165, 145},
    {173, 148, 140},
    {176, 155, 140, 135},
    {180, 157, 141, 134, 130},
    {254, 254, 254, 252, 249, 243, 230, 196, 177, 153, 140, 133, 130, 129}};

/* literal_to_type: the interp_filter literal of the uncompressed header */
VP9R_CONST uint8_t VP9_LITERAL_TO_TYPE[4] = {EIGHTTAP_SMOOTH, EIGHTTAP, EIGHTTAP_SHARP, BILINEAR};

/* Decode trees (9.3.1): leaves are negated values (a leaf of value 0 is 0,
 * which no branch targets). */
VP9R_CONST int8_t VP9_PARTITION_TREE[6] = {-PARTITION_NONE, 2, -PARTITION_HORZ, 4,
                                          -PARTITION_VERT, -PARTITION_SPLIT};
VP9R_CONST int8_t VP9_INTRA_MODE_TREE[18] = {
    -DC_PRED, 2, -TM_PRED, 4, -V_PRED, 6, 8, 12, -H_PRED, 10,
    -D135_PRED, -D117_PRED, -D45_PRED, 14, -D63_PRED, 16, -D153_PRED, -D207_PRED};
VP9R_CONST int8_t VP9_SEGMENT_TREE[14] = {2, 4, 6, 8, 10, 12, 0, -1, -2, -3, -4, -5, -6, -7};
VP9R_CONST int8_t VP9_TX_SIZE_32_TREE[6] = {-TX_4X4, 2, -TX_8X8, 4, -TX_16X16, -TX_32X32};
VP9R_CONST int8_t VP9_TX_SIZE_16_TREE[4] = {-TX_4X4, 2, -TX_8X8, -TX_16X16};
VP9R_CONST int8_t VP9_TX_SIZE_8_TREE[2] = {-TX_4X4, -TX_8X8};
/* inter_mode_tree, values y_mode - NEARESTMV */
VP9R_CONST int8_t VP9_INTER_MODE_TREE[6] = {-(ZEROMV - NEARESTMV), 2, -(NEARESTMV - NEARESTMV), 4,
                                           -(NEARMV - NEARESTMV), -(NEWMV - NEARESTMV)};
VP9R_CONST int8_t VP9_INTERP_FILTER_TREE[4] = {-EIGHTTAP, 2, -EIGHTTAP_SMOOTH, -EIGHTTAP_SHARP};
VP9R_CONST int8_t VP9_MV_JOINT_TREE[6] = {-MV_JOINT_ZERO, 2, -MV_JOINT_HNZVZ, 4,
                                         -MV_JOINT_HZVNZ, -MV_JOINT_HNZVNZ};
VP9R_CONST int8_t VP9_MV_CLASS_TREE[20] = {0, 2, -1, 4, 6, 8, -2, -3, 10, 12,
                                          -4, -5, -6, 14, 16, 18, -7, -8, -9, -10};
VP9R_CONST int8_t VP9_MV_FR_TREE[6] = {0, 2, -1, 4, -2, -3};
VP9R_CONST int8_t VP9_TOKEN_TREE[20] = {0, 2, -1, 4, 6, 10, -2, 8, -3, -4,
                                       12, 14, -5, -6, 16, 18, -7, -8, -9, -10};
/* small_token_tree (8.4.3), adapted from index 2 */
VP9R_CONST int8_t VP9_SMALL_TOKEN_TREE[6] = {0, 0, 0, 4, -1, -2};
VP9R_CONST int8_t VP9_BINARY_TREE[2] = {0, -1};

#endif
