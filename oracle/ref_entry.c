/* ORACLE — test infrastructure only.
 *
 * Entry points into the mav1 reference's own kernel code, compiled from the
 * reference sources where they lie (-I <reference>/src) into
 * oracle/_ref/libmav1ref.so by oracle/Makefile. Nothing here restates an
 * algorithm: each function is the serial loop of the reference's CPU
 * backend (src/backend_cpu.c, which itself needs libdispatch and cannot be
 * built here) around the reference's unchanged kernel functions, so the
 * outputs are the reference's. Used to pin the oracle restatement and as
 * the reference CPU baseline.
 */
#include <stdlib.h>
#include <string.h>

#include "kernels/k_analysis.h"
#include "kernels/k_cdef.h"
#include "kernels/k_lf.h"
#include "kernels/k_lr.h"
#include "kernels/k_me.h"
#include "kernels/k_recon.h"
#include "fwd_matrices.h"

_Static_assert(sizeof(FrameParams) == 392, "FrameParams layout");
_Static_assert(sizeof(BlockInfo) == 108, "BlockInfo layout");
_Static_assert(sizeof(SbInfo) == 356, "SbInfo layout");
_Static_assert(sizeof(LrUnit) == 48, "LrUnit layout");

long ref_qp_entry_bytes(const FrameParams *fp) { return qp_entry_bytes(fp); }

void ref_qpel_build(const FrameParams *fp, const uint8_t *ref, uint8_t *planes16) {
  for (int y = 0; y < fp->frame_h; y++)
    for (int x = 0; x < fp->frame_w; x++) qpel_build(fp, ref, planes16, y, x);
}

void ref_dec_build(const FrameParams *fp, const uint8_t *frame, uint8_t *dec) {
  for (int y = 0; y < dec_h(fp); y++)
    for (int x = 0; x < dec_w(fp); x++) dec_build(fp, frame, dec, y, x);
}

/* The serial form of cpu_motion_search (src/backend_cpu.c:23-56). */
void ref_motion_search(const FrameParams *fp, const uint8_t *src, const uint8_t *ref, const uint8_t *gold,
                       int32_t *mvfield, int32_t *mvgold) {
  int nref = fp->golden ? 2 : 1, nsb = fp->sb_cols * fp->sb_rows;
  long eb = qp_entry_bytes(fp);
  uint8_t *qp = calloc((size_t)(ME_QP_ENTRIES * eb + eb), 1);
  uint8_t *srcDec = qp + ME_QP_ENTRIES * eb;
  int32_t *scratch = calloc((size_t)nsb * SCRATCH_INTS, sizeof(int32_t));
  if (fp->turbo & T_ME_HPEL) {
    int W = fp->frame_w, H = fp->frame_h;
    for (int g = 0; g < H * (nref + 1); g++) {
      int r = g / H, y = g % H;
      for (int x = 0; x < W; x++) {
        int dec = (fp->turbo & T_ME_DEC) && y < dec_h(fp) && x < dec_w(fp);
        if (r == nref) {
          if (dec) dec_build(fp, src, srcDec, y, x);
          continue;
        }
        uint8_t *ent = qp + r * eb;
        qpel_build(fp, r ? gold : ref, ent, y, x);
        if (dec) dec_build(fp, r ? gold : ref, ent + 16L * W * H, y, x);
      }
    }
  }
  for (int level = 0; level < 4; level++) {
    int nn = 1 << (2 * level), nodes = nsb * nn;
    for (int g = 0; g < nodes * nref; g++) {
      int r = g / nodes, sb = (g % nodes) / nn, i = g % nn;
      me_node_serial(fp, sb, level, i, src, r ? gold : ref, qp + r * eb, srcDec, qp_dec(fp, qp + r * eb),
                     r ? mvgold : mvfield, scratch + (size_t)sb * SCRATCH_INTS + r * (SCRATCH_INTS / 2), r);
    }
  }
  free(scratch);
  free(qp);
}

/* The serial form of cpu_backend_analyze (src/backend_cpu.c:58-74). */
void ref_analyze(const FrameParams *fp, const uint8_t *src, const uint8_t *ref, const uint8_t *gold,
                 const int32_t *mvfield, const int32_t *mvgold, long *costs, BlockInfo *infos, BlockInfo *blocks,
                 SbInfo *sbs) {
  int nsb = fp->sb_cols * fp->sb_rows;
  int32_t *scratch = calloc(SCRATCH_INTS, sizeof(int32_t));
  for (int i = 0; i < nsb; i++) {
    for (int level = 0; level < 4; level++) {
      int ns = level_slots(fp, level);
      for (int node = 0; node < (1 << (2 * level)); node++)
        for (int slot = 0; slot < ns; slot++)
          analyze_item(fp, i, level, node, slot, src, ref, gold, mvfield, mvgold, costs, infos, scratch);
    }
    analyze_sb_table(fp, i, costs, infos, blocks, sbs);
  }
  free(scratch);
}

void ref_predict_intra(const uint8_t *px, int stride, int x, int y, const IntraArgs *a, uint8_t *dst, int dstStride) {
  predict_intra(px, stride, x, y, *a, dst, dstStride);
}

/* The serial form of cpu_backend_recon (src/backend_cpu.c:153-156):
 * superblocks in raster order satisfy every left / above-right dependency. */
void ref_recon(const FrameParams *fp, BlockInfo *blocks, const SbInfo *sbs, const uint8_t *src, uint8_t *rec,
               const uint8_t *ref, const uint8_t *gold, int32_t *coefs, const int32_t *mats) {
  int nsb = fp->sb_cols * fp->sb_rows;
  int32_t *scratch = calloc(SCRATCH_INTS, sizeof(int32_t));
  for (int i = 0; i < nsb; i++) recon_sb(fp, sbs[i], blocks, src, rec, ref, gold, coefs, mats, scratch);
  free(scratch);
}

void ref_itx_2d(int32_t *dq, int32_t *res, int txSz, int txType, int lossless) { itx_2d(dq, res, txSz, txType, lossless); }
void ref_fwd_2d(const int32_t *resid, int32_t *out, int txSz, int txType, const int32_t *mats) {
  static int32_t tmp[64 * 64];
  fwd_2d(resid, out, txSz, txType, mats, tmp);
}
void ref_quantize(const int32_t *coef, int32_t *quant, int txSz, int dcQ, int acQ) {
  quantize_block(coef, quant, txSz, dcQ, acQ, 26, 22);
}
void ref_dequant(const int32_t *quant, int32_t *dq, int txSz, int dcQ, int acQ) { dequant_block(quant, dq, txSz, dcQ, acQ); }
void ref_forward_matrices(int32_t *mats) { fwd_build_matrices(mats); }

/* The serial form of cpu_filter's deblocking (src/backend_cpu.c:158-170). */
void ref_loop_filter(const FrameParams *fp, uint8_t *rec, const uint8_t *info) {
  for (int pass = 0; pass < 2; pass++) {
    int lines = pass == 0 ? fp->mi_rows : fp->mi_cols;
    for (int i = 0; i < 3 * lines; i++) {
      int plane = i / lines, line = i % lines;
      if (plane > 0 && (line & 1)) continue;
      lf_line(fp, rec, info, plane, pass, line);
    }
  }
}

/* The serial form of cpu_filter's CDEF (src/backend_cpu.c:171-176). */
void ref_cdef(const FrameParams *fp, const uint8_t *rec, uint8_t *out, long frameBytes, const uint8_t *info, int n) {
  int bc = (fp->mi_cols + 1) >> 1, br = (fp->mi_rows + 1) >> 1;
  for (int i = 0; i < bc * br; i++) cdef_block_all(fp, rec, out, frameBytes, info, (i / bc) * 2, (i % bc) * 2, n);
}

/* The serial form of cpu_restore (src/backend_cpu.c:179-188). */
void ref_lr_apply(const FrameParams *fp, const uint8_t *cdef, const uint8_t *dbk, uint8_t *out, const LrUnit *units) {
  int rows4 = (fp->frame_h + 3) >> 2, cols4 = (fp->frame_w + 3) >> 2, n = rows4 * cols4;
  for (int g = 0; g < 3 * n; g++) lr_block(fp, g / n, (g % n) / cols4, (g % n) % cols4, cdef, dbk, out, units);
}
