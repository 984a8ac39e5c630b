/* AV1 enumerations and constants used by the analysis backend (names and
 * values from the AV1 specification, section 6.10 and section 3), plus the
 * index layout of the flattened tables in av1_tables.h.
 * Plain C so that the host code, the HIP kernels and the oracle share it. */
#ifndef MAV1_AV1_DEFS_H
#define MAV1_AV1_DEFS_H

enum { AV1_KEY_FRAME = 0, AV1_INTER_FRAME = 1 };
enum {
  INTRA_FRAME = 0, LAST_FRAME = 1, LAST2_FRAME = 2, LAST3_FRAME = 3, GOLDEN_FRAME = 4,
  BWDREF_FRAME = 5, ALTREF2_FRAME = 6, ALTREF_FRAME = 7
};
enum {
  BLOCK_4X4, BLOCK_4X8, BLOCK_8X4, BLOCK_8X8, BLOCK_8X16, BLOCK_16X8, BLOCK_16X16, BLOCK_16X32,
  BLOCK_32X16, BLOCK_32X32, BLOCK_32X64, BLOCK_64X32, BLOCK_64X64, BLOCK_64X128, BLOCK_128X64,
  BLOCK_128X128, BLOCK_4X16, BLOCK_16X4, BLOCK_8X32, BLOCK_32X8, BLOCK_16X64, BLOCK_64X16,
  BLOCK_SIZES_ALL
};
enum {
  PARTITION_NONE, PARTITION_HORZ, PARTITION_VERT, PARTITION_SPLIT, PARTITION_HORZ_A,
  PARTITION_HORZ_B, PARTITION_VERT_A, PARTITION_VERT_B, PARTITION_HORZ_4, PARTITION_VERT_4
};
enum {
  DC_PRED, V_PRED, H_PRED, D45_PRED, D135_PRED, D113_PRED, D157_PRED, D203_PRED, D67_PRED,
  SMOOTH_PRED, SMOOTH_V_PRED, SMOOTH_H_PRED, PAETH_PRED, UV_CFL_PRED
};
enum { NEWMV = 16 }; /* YMode numbering: NEARESTMV 13, NEARMV 14, GLOBALMV 15, NEWMV 16 */
enum {
  TX_4X4, TX_8X8, TX_16X16, TX_32X32, TX_64X64, TX_4X8, TX_8X4, TX_8X16, TX_16X8, TX_16X32,
  TX_32X16, TX_32X64, TX_64X32, TX_4X16, TX_16X4, TX_8X32, TX_32X8, TX_16X64, TX_64X16
};
enum {
  DCT_DCT, ADST_DCT, DCT_ADST, ADST_ADST, FLIPADST_DCT, DCT_FLIPADST, FLIPADST_FLIPADST,
  ADST_FLIPADST, FLIPADST_ADST, IDTX, V_DCT, H_DCT, V_ADST, H_ADST, V_FLIPADST, H_FLIPADST
};
enum { TX_SET_DCTONLY = 0, TX_SET_INTRA_1 = 1, TX_SET_INTRA_2 = 2, TX_SET_INTER_1 = 1, TX_SET_INTER_2 = 2, TX_SET_INTER_3 = 3 };
enum { EIGHTTAP = 0, EIGHTTAP_SMOOTH = 1, EIGHTTAP_SHARP = 2, BILINEAR = 3 };
enum { RESTORE_NONE = 0, RESTORE_WIENER = 1, RESTORE_SGRPROJ = 2, RESTORE_SWITCHABLE = 3 };

#define MI_SIZE 4
#define MI_SIZE_LOG2 2
#define MAX_LOOP_FILTER 63
#define SUBPEL_BITS 4
#define SUBPEL_MASK 15
#define SCALE_SUBPEL_BITS 10
#define REF_SCALE_SHIFT 14
#define MAX_ANGLE_DELTA 3
#define ANGLE_STEP 3
#define MAX_TX_DEPTH 2
#define INTRA_EDGE_TAPS 5
#define FILTER_BITS 7
#define SGRPROJ_PARAMS_BITS 4
#define SGRPROJ_PRJ_BITS 7
#define SGRPROJ_RST_BITS 4
#define SGRPROJ_MTABLE_BITS 20
#define SGRPROJ_RECIP_BITS 12
#define SGRPROJ_SGR_BITS 8
#define SGRPROJ_PRJ_SUBEXP_K 4
#define RESTORATION_TILESIZE_MAX 256
#define MAX_TILE_WIDTH 4096
#define MAX_TILE_COLS 64

/* Flattened table layouts (av1_tables.h). */
#define SUBPEL_FILTER(f, p) (Subpel_Filters + ((f) * 16 + (p)) * 8) /* [6][16][8] */
#define DC_Q(q) (Dc_Qlookup[(q)])                                    /* [3][256], 8-bit row */
#define AC_Q(q) (Ac_Qlookup[(q)])
#define PARTITION_SUBSIZE(p, bs) (Partition_Subsize[(p) * 22 + (bs)]) /* [10][22] */
#define SUBSAMPLED_SIZE(bs, sx, sy) (Subsampled_Size[(bs) * 4 + (sx) * 2 + (sy)])
#define TX_IN_SET_INTRA(set, t) (Tx_Type_In_Set_Intra[(set) * 16 + (t)]) /* [3][16] */
#define TX_IN_SET_INTER(set, t) (Tx_Type_In_Set_Inter[(set) * 16 + (t)]) /* [4][16] */
#define INTRA_EDGE_KERNEL(s, j) (Intra_Edge_Kernel[(s) * 5 + (j)])      /* [3][5] */
#define CDEF_DIRECTION(d, k, c) (Cdef_Directions[(d) * 4 + (k) * 2 + (c)]) /* [8][2][2] */
#define CDEF_PRI_TAP(s, k) (Cdef_Pri_Taps[(s) * 2 + (k)])
#define CDEF_SEC_TAP(s, k) (Cdef_Sec_Taps[(s) * 2 + (k)])
#define CDEF_UV_DIR(sx, sy, d) (Cdef_Uv_Dir[(sx) * 16 + (sy) * 8 + (d)])
#define SGR_PARAM(set, i) (Sgr_Params[(set) * 4 + (i)]) /* [16][4] */
#define BLOCK_WIDTH(bs) (4 * Num_4x4_Blocks_Wide[bs])
#define BLOCK_HEIGHT(bs) (4 * Num_4x4_Blocks_High[bs])

#endif
