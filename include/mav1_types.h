/* Data records exchanged across the mav1 analysis-backend boundary.
 *
 * These are the encoder's shared-buffer records, so their layout is the
 * data contract between the frame driver and any backend: every field is a
 * 32-bit integer and the order, count and size match the mav1 encoder's
 * structs field for field (a backend receives the encoder's buffers as they
 * are, no repacking):
 *   Mav1FrameParams  == FrameParams  (src/kernels/k_common.h:52-85,  392 B)
 *   Mav1BlockInfo    == BlockInfo    (src/kernels/k_common.h:98-115, 108 B)
 *   Mav1SbInfo       == SbInfo       (src/kernels/k_common.h:118-122, 356 B)
 *   Mav1LrUnit       == LrUnit       (src/kernels/k_lr.h:12-18,       48 B)
 */
#ifndef MAV1_TYPES_H
#define MAV1_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-frame constants (filled by the driver's setup_params,
 * src/encoder.c:407-510). */
typedef struct {
  int32_t mi_cols, mi_rows;
  int32_t sb_cols, sb_rows;
  int32_t plane_w[3], plane_h[3];   /* allocated plane size */
  int32_t stride[3], offset[3];     /* frame buffer layout: plane p at offset[p], row stride stride[p] */
  int32_t base_q_idx;
  int32_t dc_q[3], ac_q[3];
  int32_t frame_is_intra;
  int32_t reduced_tx_set;
  int32_t enable_intra_edge_filter;
  int32_t tx_search;                /* closed-loop luma tx type search */
  int32_t lambda;                   /* RD multiplier: ac_q^2 * 256 / 430 */
  int32_t wave;
  int32_t num_blocks;
  int32_t frame_w, frame_h;         /* real (cropped) luma size */
  int32_t me_range;                 /* full-pel motion search range */
  int32_t lossless;
  int32_t tx_mode_select;
  int32_t tile_w_sb;                /* uniform tile column width in superblocks */
  int32_t fast;
  int32_t leaf;                     /* non-reference frame: inter only */
  int32_t depth_thr;                /* depth early-exit threshold (SATD per sample x16) */
  int32_t switchable_interp;
  int32_t golden;                   /* a second reference is searched */
  int32_t ref2_frame;               /* GOLDEN_FRAME or ALTREF_FRAME */
  int32_t compound;                 /* LAST + second reference compound allowed */
  int32_t lf_level[4];
  int32_t lf_sharpness;
  int32_t cdef_damping, cdef_bits;
  int32_t cdef_y_pri[8], cdef_y_sec[8], cdef_uv_pri[8], cdef_uv_sec[8];
  int32_t lr_type[3], lr_unit_size[3], lr_unit_rows[3], lr_unit_cols[3], lr_unit_off[3];
  int32_t turbo;                    /* speed-mode bit mask (MAV1_T_* below) */
} Mav1FrameParams;

/* One coded block in decode order: the mode decision plus the
 * reconstruction outputs. */
typedef struct {
  int32_t mi_row, mi_col, bsize;
  int32_t is_inter;
  int32_t y_mode, uv_mode, angle_y, angle_uv;
  int32_t mv_row, mv_col;           /* Mv[0], 1/8 pel */
  int32_t tx_type;                  /* luma tx type of the first transform */
  int32_t eob[3];                   /* per plane: any non-zero coefficient */
  int32_t coef_off[3];              /* offsets into the coefficient buffer */
  int32_t cfl_alpha_u, cfl_alpha_v;
  int32_t tx_depth;
  int32_t tx_types_lo, tx_types_hi; /* luma tx type of each transform, 4 bits each */
  int32_t interp_filter;            /* filter[0] | filter[1] << 2; bit 4/5: intra edge filter type */
  int32_t ref_frame;                /* RefFrame[0] of inter blocks */
  int32_t ref_frame1;               /* RefFrame[1] of compound blocks, else 0 */
  int32_t mv1_row, mv1_col;         /* Mv[1] of compound blocks */
} Mav1BlockInfo;

/* Index range of a superblock's blocks and its partition tree. */
typedef struct {
  int32_t first_block, num_blocks;
  int32_t sb_row, sb_col;
  int32_t part[85];                 /* partition per square node (-1 outside the frame) */
} Mav1SbInfo;

/* Loop-restoration parameters of one unit. */
typedef struct {
  int32_t type;                     /* RESTORE_NONE / RESTORE_WIENER / RESTORE_SGRPROJ */
  int32_t set;                      /* self-guided parameter set */
  int32_t xqd[2];
  int32_t wiener[2][3];             /* [0] vertical, [1] horizontal taps 0..2 */
  int32_t pad[2];
} Mav1LrUnit;

/* Layout constants of the shared buffers (src/kernels/k_analysis.h:22-32,
 * src/kernels/k_me.h:102, src/kernels/k_recon.h:24). */
#define MAV1_NUM_NODES 85          /* square nodes per 64x64 superblock: 1 + 4 + 16 + 64 */
#define MAV1_A_SLOTS 13            /* analysis table slots per node */
#define MAV1_A_ITEMS (MAV1_NUM_NODES * MAV1_A_SLOTS)
#define MAV1_MAX_BLOCKS_PER_SB 64
#define MAV1_COEFS_PER_SB 6144
#define MAV1_QP_ENTRIES 3          /* quarter-pel plane entries of the turbo motion search */

/* Speed-mode bits (FrameParams.turbo); default mask 0x47CBF1
 * (src/encoder.c:449, src/kernels/k_me.h:79-92, k_analysis.h:38-48, :600-603). */
#define MAV1_T_SQUARE      0x000001
#define MAV1_T_ME_CR1      0x000002
#define MAV1_T_ME_L3COPY   0x000004
#define MAV1_T_ME_KEEP     0x000008
#define MAV1_T_ME_HPEL     0x000010
#define MAV1_T_ME_DEC      0x000020
#define MAV1_T_ME_CR2      0x000040
#define MAV1_T_INTRA       0x000080
#define MAV1_T_KEY_SQUARE  0x000100
#define MAV1_T_ME_REF2     0x000200
#define MAV1_T_LEAF_NO8    0x000400
#define MAV1_T_CDEF2       0x000800
#define MAV1_T_ME_L3SUB    0x001000
#define MAV1_T_ME_CROSS    0x002000
#define MAV1_T_ME_G3COPY   0x004000
#define MAV1_T_ME_HALF01   0x008000
#define MAV1_T_ME_RSUB     0x010000
#define MAV1_T_ME_HME3     0x020000
#define MAV1_T_ME_USUB     0x040000
#define MAV1_T_ANA_USUB    0x400000
#define MAV1_TURBO_DEFAULT 0x47CBF1

#ifdef __cplusplus
}
#endif

#endif
