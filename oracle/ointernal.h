/* ORACLE — test infrastructure only. Functions shared between the oracle's
 * translation units. */
#ifndef MAV1_ORACLE_INTERNAL_H
#define MAV1_ORACLE_INTERNAL_H

#include <stdint.h>

typedef struct {
  int haveLeft, haveAbove, haveAboveRight, haveBelowLeft;
  int mode, angleDelta, log2W, log2H;
  int maxX, maxY;
  int edgeFilter, filterType;
} OIntraArgs;

void o_predict_intra(const uint8_t *px, int stride, int x, int y, const OIntraArgs *a, uint8_t *dst, int dstStride);
void o_cfl_predict(const uint8_t *luma, int ls, int sx, int sy, int maxW, int maxH, int log2W, int log2H, int alpha,
                   const int *dc, uint8_t *dst, int dstStride);
int o_cfl_best_alpha(const uint8_t *luma, int ls, int sx, int sy, int maxW, int maxH, int log2W, int log2H,
                     const int *dc, const uint8_t *src, int srcStride, int excludeZero);
/* Compound single-reference prediction (InterRound1 = 7, no clip). */
void o_predict_inter_raw(const uint8_t *ref, int refStride, int refW, int refH, int32_t *dst, int dstStride, int x,
                         int y, int w, int h, int mvRow, int mvCol, int subX, int subY, int filter);

/* Partition geometry (src/kernels/k_analysis.h:408-443). */
int o_part_block(int p, int bSize, int k, int halfMi, int *dr, int *dc);
int o_part_count(const Mav1FrameParams *fp, int p, int r, int c, int halfMi);

#endif
