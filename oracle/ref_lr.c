/* ORACLE — test infrastructure only.
 *
 * Calls the reference's own loop-restoration parameter search
 * (src/lr_search.c, compiled unchanged from where it lies) on buffers given
 * by the caller: a zeroed encoder context carries exactly the fields
 * lr_select reads (geometry, the frame buffers, the per-unit records, the
 * frame parameters and the lr_test flag). */
#include <stdlib.h>

#include "enc_internal.h"

void lr_select(Mav1Encoder *e);

void ref_lr_select(FrameParams *fp, const uint8_t *src, const uint8_t *cdef, const uint8_t *dbk, uint8_t *tmp,
                   void *units, int lrTest) {
  Mav1Encoder *e = calloc(1, sizeof(*e));
  e->w = fp->frame_w;
  e->h = fp->frame_h;
  for (int p = 0; p < 3; p++) {
    e->stride[p] = fp->stride[p];
    e->offset[p] = fp->offset[p];
  }
  e->params.base = (uint8_t *)fp;
  e->src.base = (uint8_t *)src;
  e->rec.base = (uint8_t *)cdef;
  e->dbk.base = (uint8_t *)dbk;
  e->lrout.base = tmp;
  e->lrunits.base = (uint8_t *)units;
  e->cfg.lr_test = lrTest;
  lr_select(e);
  free(e);
}
