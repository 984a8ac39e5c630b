/* ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the mav1 analysis path
 * (the reference's src/kernels/k_*.h, backend_cpu.c and the host searches
 * of encoder.c / lr_search.c), written for this project to check the HIP
 * backend against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker: the product path
 * (libmav1hip.so) never links or calls it.
 *
 * Independent superblocks, rows and units run on OpenMP threads; the
 * results do not depend on the thread count (every parallel loop writes
 * disjoint outputs, and the serial dependencies stay serial).
 *
 * Every function cites the reference file:line whose behaviour it follows.
 * The restatement is pinned against the reference itself: oracle/Makefile
 * also compiles the reference's kernel headers, unchanged, into
 * oracle/_ref/libmav1ref.so, and tests/test_oracle_vs_ref.py compares the
 * two on seeded synthetic frames; closed forms (Hadamard SATD, SAD, the
 * quantiser's division formula) are checked independently.
 *
 * Buffers use the encoder's padded planar layout (Mav1FrameParams.stride /
 * offset); records are the ABI structs of include/mav1_types.h.
 */
#ifndef MAV1_ORACLE_H
#define MAV1_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/mav1_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- motion search (src/kernels/k_me.h, src/backend_cpu.c:23-56) --- */
/* Bytes of one quarter-pel plane entry: 16 W x H planes + the decimated plane. */
long oracle_qp_entry_bytes(const Mav1FrameParams *fp);
/* The 16 EIGHTTAP quarter-pel planes of a reference (k_me.h:289-313). */
void oracle_qpel_build(const Mav1FrameParams *fp, const uint8_t *ref, uint8_t *planes16);
/* 4x4-mean decimated luma plane, ceil(W/4) x ceil(H/4) (k_me.h:112-118). */
void oracle_dec_build(const Mav1FrameParams *fp, const uint8_t *frame, uint8_t *dec);
/* Whole-frame motion search of every node of every level, LAST into
 * mvfield and (fp->golden) the second reference into mvgold; both
 * NUM_SBS * 85 * 2 int32 (k_me.h:343-379 per node). */
void oracle_motion_search(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref, const uint8_t *gold,
                          int32_t *mvfield, int32_t *mvgold);

/* --- analysis (src/kernels/k_analysis.h, src/backend_cpu.c:58-74) --- */
/* Pass 1 over every (SB, level, node, slot) item into costs / infos
 * (NUM_SBS * 1105 entries each; cost -1: not evaluated), then pass 2 into
 * blocks (NUM_SBS * 64) and sbs (NUM_SBS). */
void oracle_analyze(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref, const uint8_t *gold,
                    const int32_t *mvfield, const int32_t *mvgold, int64_t *costs, Mav1BlockInfo *infos,
                    Mav1BlockInfo *blocks, Mav1SbInfo *sbs);

/* --- reconstruction (src/kernels/k_recon.h, src/backend_cpu.c:81-156) --- */
/* Edge filter types of the blocks (src/encoder.c:319-346), in place. */
void oracle_edge_types(const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs);
/* Every superblock in raster order: prediction, transform, quantisation,
 * reconstruction into rec, levels into coefs (NUM_SBS * 6144 int32), the
 * block outputs into blocks. mats: the Q20 forward matrices (15 * 4096). */
void oracle_recon(const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs, const uint8_t *src,
                  uint8_t *rec, const uint8_t *ref, const uint8_t *gold, int32_t *coefs, const int32_t *mats);
/* Single transforms (row stride 64 for coefficient / residual blocks). */
void oracle_itx_2d(const int32_t *dq, int32_t *res, int txSz, int txType, int lossless);
void oracle_fwd_2d(const int32_t *resid, int32_t *out, int txSz, int txType, const int32_t *mats);
void oracle_quantize(const int32_t *coef, int32_t *quant, int txSz, int dcQ, int acQ);
void oracle_dequant(const int32_t *quant, int32_t *dq, int txSz, int dcQ, int acQ);

/* --- in-loop filters --- */
/* Per mi (mi_rows x mi_cols bytes): bsize | skip << 5 | tx_depth << 6. */
void oracle_filter_info(const Mav1FrameParams *fp, const Mav1BlockInfo *blocks, const Mav1SbInfo *sbs, uint8_t *info);
/* Deblocking of rec in place. */
void oracle_loop_filter(const Mav1FrameParams *fp, uint8_t *rec, const uint8_t *info);
/* CDEF of cur with parameter sets 0 .. n-1 into out + k * frameBytes. */
void oracle_cdef(const Mav1FrameParams *fp, const uint8_t *cur, uint8_t *out, int64_t frameBytes, const uint8_t *info,
                 int n);
/* Per superblock the set of smallest SSE against src, copied into rec;
 * choice: one byte per superblock. */
void oracle_cdef_select(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdefbuf, int64_t frameBytes,
                        int n, uint8_t *rec, int8_t *choice);
/* Loop restoration of the CDEF output (cdef) with the deblocked frame
 * (dbk) above / below stripes into out. */
void oracle_lr_apply(const Mav1FrameParams *fp, const uint8_t *cdef, const uint8_t *dbk, uint8_t *out,
                     const Mav1LrUnit *units);
/* Per-unit restoration parameter search; writes units and fp->lr_type.
 * tmp: frame-sized scratch. */
void oracle_lr_select(Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdef, const uint8_t *dbk, uint8_t *tmp,
                      Mav1LrUnit *units, int lrTest);

/* --- decision record --- */
/* Per-8x8 grid ((mi_rows+1)/2 x (mi_cols+1)/2 cells of 8 bytes). */
void oracle_decision_grid(const Mav1FrameParams *fp, const Mav1BlockInfo *blocks, const Mav1SbInfo *sbs, uint8_t *grid);

/* --- shared prediction primitives --- */
/* Intra prediction arguments (the IntraArgs record of src/kernels/k_intra.h:79-85). */
typedef struct {
  int32_t haveLeft, haveAbove, haveAboveRight, haveBelowLeft;
  int32_t mode, angleDelta, log2W, log2H;
  int32_t maxX, maxY;
  int32_t edgeFilter, filterType;
} OracleIntraArgs;
void oracle_predict_intra(const uint8_t *px, int stride, int x, int y, const OracleIntraArgs *a, uint8_t *dst,
                          int dstStride);
int oracle_satd4x4(const int32_t *d);
void oracle_predict_inter(const uint8_t *ref, int refStride, int refW, int refH, uint8_t *dst, int dstStride, int x,
                          int y, int w, int h, int mvRow, int mvCol, int subX, int subY, int filter);

#ifdef __cplusplus
}
#endif

#endif
