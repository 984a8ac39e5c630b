/* ORACLE — test infrastructure only. Prediction primitives. */
#include "ocommon.h"

/* Sum of absolute values of the 2-D 4x4 Hadamard transform of d (row-major),
 * unnormalised (src/kernels/k_analysis.h:58-73). Horizontal butterflies on
 * each row, then vertical ones on each column. */
int oracle_satd4x4(const int32_t *d) {
  int m[16];
  for (int r = 0; r < 4; r++) {
    const int32_t *v = d + 4 * r;
    int s01 = v[0] + v[1], d01 = v[0] - v[1], s23 = v[2] + v[3], d23 = v[2] - v[3];
    m[4 * r + 0] = s01 + s23;
    m[4 * r + 1] = d01 + d23;
    m[4 * r + 2] = s01 - s23;
    m[4 * r + 3] = d01 - d23;
  }
  int sum = 0;
  for (int c = 0; c < 4; c++) {
    int s01 = m[c] + m[4 + c], d01 = m[c] - m[4 + c], s23 = m[8 + c] + m[12 + c], d23 = m[8 + c] - m[12 + c];
    sum += o_abs(s01 + s23) + o_abs(d01 + d23) + o_abs(s01 - s23) + o_abs(d01 - d23);
  }
  return sum;
}

/* Block inter prediction from an unscaled reference (AV1 spec 7.11.3.3 with
 * xScale = yScale = 1 << 14, then 7.11.3.4), 8-bit, non-compound rounding
 * (InterRound0 3, InterRound1 11), as src/kernels/k_inter.h:136-189.
 * filter = InterpFilter[0] (vertical) | InterpFilter[1] (horizontal) << 2;
 * blocks 4 wide (high) use the 4-tap kernels 4 / 5. */
void oracle_predict_inter(const uint8_t *ref, int refStride, int refW, int refH, uint8_t *dst, int dstStride, int x,
                          int y, int w, int h, int mvRow, int mvCol, int subX, int subY, int filter) {
  /* Unscaled 7.11.3.3: the start position in 1/1024 sample units. */
  int posX = ((x << SUBPEL_BITS) + ((2 * mvCol) >> subX)) * 64 + 32;
  int posY = ((y << SUBPEL_BITS) + ((2 * mvRow) >> subY)) * 64 + 32;
  int lastX = ((refW + subX) >> subX) - 1, lastY = ((refH + subY) >> subY) - 1;
  int fv = filter & 3, fh = (filter >> 2) & 3;
  if (w <= 4) fh = (fh == EIGHTTAP_SMOOTH) ? 5 : (fh == EIGHTTAP || fh == EIGHTTAP_SHARP) ? 4 : fh;
  if (h <= 4) fv = (fv == EIGHTTAP_SMOOTH) ? 5 : (fv == EIGHTTAP || fv == EIGHTTAP_SHARP) ? 4 : fv;
  const int16_t *kh = SUBPEL_FILTER(fh, (posX >> 6) & SUBPEL_MASK);
  const int16_t *kv = SUBPEL_FILTER(fv, (posY >> 6) & SUBPEL_MASK);
  int ix = posX >> 10, iy = posY >> 10;
  int *mid = (int *)malloc(sizeof(int) * (size_t)(h + 7) * w);
  for (int r = 0; r < h + 7; r++) {
    const uint8_t *row = ref + o_clip(0, lastY, iy + r - 3) * refStride;
    for (int c = 0; c < w; c++) {
      int acc = 0;
      for (int t = 0; t < 8; t++) acc += kh[t] * row[o_clip(0, lastX, ix + c + t - 3)];
      mid[r * w + c] = o_round2(acc, 3);
    }
  }
  for (int r = 0; r < h; r++)
    for (int c = 0; c < w; c++) {
      int acc = 0;
      for (int t = 0; t < 8; t++) acc += kv[t] * mid[(r + t) * w + c];
      dst[r * dstStride + c] = (uint8_t)o_clip1(o_round2(acc, 11));
    }
  free(mid);
}

/* Compound single-reference prediction (7.11.3.1 with isCompound: vertical
 * InterRound1 = 7, not clipped), as src/kernels/k_inter.h:195-248. */
void o_predict_inter_raw(const uint8_t *ref, int refStride, int refW, int refH, int32_t *dst, int dstStride, int x,
                         int y, int w, int h, int mvRow, int mvCol, int subX, int subY, int filter) {
  int posX = ((x << SUBPEL_BITS) + ((2 * mvCol) >> subX)) * 64 + 32;
  int posY = ((y << SUBPEL_BITS) + ((2 * mvRow) >> subY)) * 64 + 32;
  int lastX = ((refW + subX) >> subX) - 1, lastY = ((refH + subY) >> subY) - 1;
  int fv = filter & 3, fh = (filter >> 2) & 3;
  if (w <= 4) fh = (fh == EIGHTTAP_SMOOTH) ? 5 : (fh == EIGHTTAP || fh == EIGHTTAP_SHARP) ? 4 : fh;
  if (h <= 4) fv = (fv == EIGHTTAP_SMOOTH) ? 5 : (fv == EIGHTTAP || fv == EIGHTTAP_SHARP) ? 4 : fv;
  const int16_t *kh = SUBPEL_FILTER(fh, (posX >> 6) & SUBPEL_MASK), *kv = SUBPEL_FILTER(fv, (posY >> 6) & SUBPEL_MASK);
  int ix = posX >> 10, iy = posY >> 10;
  int *mid = (int *)malloc(sizeof(int) * (size_t)(h + 7) * w);
  for (int r = 0; r < h + 7; r++) {
    const uint8_t *row = ref + o_clip(0, lastY, iy + r - 3) * refStride;
    for (int c = 0; c < w; c++) {
      int acc = 0;
      for (int t = 0; t < 8; t++) acc += kh[t] * row[o_clip(0, lastX, ix + c + t - 3)];
      mid[r * w + c] = o_round2(acc, 3);
    }
  }
  for (int r = 0; r < h; r++)
    for (int c = 0; c < w; c++) {
      int acc = 0;
      for (int t = 0; t < 8; t++) acc += kv[t] * mid[(r + t) * w + c];
      dst[r * dstStride + c] = o_round2(acc, 7);
    }
  free(mid);
}
