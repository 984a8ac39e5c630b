/* ORACLE — test infrastructure only. Intra prediction (AV1 spec 7.11.2)
 * and chroma-from-luma (7.11.5), as src/kernels/k_intra.h. */
#include "ocommon.h"
#include "ointernal.h"

static const uint8_t *sm_weights(int log2) {
  switch (log2) {
    case 2: return Sm_Weights_Tx_4x4;
    case 3: return Sm_Weights_Tx_8x8;
    case 4: return Sm_Weights_Tx_16x16;
    case 5: return Sm_Weights_Tx_32x32;
    default: return Sm_Weights_Tx_64x64;
  }
}

/* Edge filter strength (7.11.2.9), k_intra.h:25-41. */
static int edge_strength(int w, int h, int filterType, int delta) {
  int d = o_abs(delta), wh = w + h;
  if (filterType == 0) {
    if (wh <= 8) return d >= 56 ? 1 : 0;
    if (wh <= 16) return d >= 40 ? 1 : 0;
    if (wh <= 24) return d >= 32 ? 3 : d >= 16 ? 2 : d >= 8 ? 1 : 0;
    if (wh <= 32) return d >= 32 ? 3 : d >= 4 ? 2 : 1;
    return 3;
  }
  if (wh <= 8) return d >= 64 ? 2 : d >= 40 ? 1 : 0;
  if (wh <= 16) return d >= 48 ? 2 : d >= 20 ? 1 : 0;
  if (wh <= 24) return d >= 4 ? 3 : 0;
  return 3;
}

static int use_upsample(int w, int h, int filterType, int delta) {
  int d = o_abs(delta);
  if (d <= 0 || d >= 40) return 0;
  return filterType == 0 ? (w + h <= 16) : (w + h <= 8);
}

/* 7.11.2.12 on e[-1 .. sz-2] (e points at element 0). */
static void edge_filter(int *e, int sz, int strength) {
  if (!strength) return;
  int cp[160];
  for (int i = 0; i < sz; i++) cp[i] = e[i - 1];
  for (int i = 1; i < sz; i++) {
    int s = 0;
    for (int j = 0; j < 5; j++) s += INTRA_EDGE_KERNEL(strength - 1, j) * cp[o_clip(0, sz - 1, i - 2 + j)];
    e[i - 1] = (s + 8) >> 4;
  }
}

/* 7.11.2.11: doubles e[-1 .. numPx-1] into e[-2 .. 2 numPx - 2]. */
static void edge_upsample(int *e, int numPx) {
  int dup[80];
  dup[0] = e[-1];
  for (int i = -1; i < numPx; i++) dup[i + 2] = e[i];
  dup[numPx + 2] = e[numPx - 1];
  e[-2] = dup[0];
  for (int i = 0; i < numPx; i++) {
    int s = -dup[i] + 9 * dup[i + 1] + 9 * dup[i + 2] - dup[i + 3];
    e[2 * i - 1] = o_clip1(o_round2(s, 4));
    e[2 * i] = dup[i + 2];
  }
}

void o_predict_intra(const uint8_t *px, int stride, int x, int y, const OIntraArgs *a, uint8_t *dst, int dstStride) {
  int w = 1 << a->log2W, h = 1 << a->log2H;
  int abuf[16 + 300], lbuf[16 + 300];
  int *above = abuf + 16, *left = lbuf + 16;
  int hl = a->haveLeft, ha = a->haveAbove;
  /* neighbouring samples (7.11.2 steps 1-3) */
  int aLim = o_min(a->maxX, x + (a->haveAboveRight ? 2 * w : w) - 1);
  int lLim = o_min(a->maxY, y + (a->haveBelowLeft ? 2 * h : h) - 1);
  for (int i = 0; i < w + h; i++) {
    above[i] = ha ? px[(y - 1) * stride + o_min(aLim, x + i)] : hl ? px[y * stride + x - 1] : 127;
    left[i] = hl ? px[o_min(lLim, y + i) * stride + x - 1] : ha ? px[(y - 1) * stride + x] : 129;
  }
  int corner = (ha && hl) ? px[(y - 1) * stride + x - 1] : ha ? px[(y - 1) * stride + x] : hl ? px[y * stride + x - 1] : 128;
  above[-1] = left[-1] = corner;
  int mode = a->mode;
  if (mode >= V_PRED && mode <= D67_PRED) {
    int pAngle = Mode_To_Angle[mode] + a->angleDelta * ANGLE_STEP;
    int upA = 0, upL = 0;
    if (a->edgeFilter) {
      if (pAngle != 90 && pAngle != 180) {
        if (pAngle > 90 && pAngle < 180 && w + h >= 24)
          above[-1] = left[-1] = o_round2(left[0] * 5 + above[-1] * 6 + above[0] * 5, 4);
        if (ha)
          edge_filter(above, o_min(w, a->maxX - x + 1) + (pAngle < 90 ? h : 0) + 1,
                      edge_strength(w, h, a->filterType, pAngle - 90));
        if (hl)
          edge_filter(left, o_min(h, a->maxY - y + 1) + (pAngle > 180 ? w : 0) + 1,
                      edge_strength(w, h, a->filterType, pAngle - 180));
      }
      upA = use_upsample(w, h, a->filterType, pAngle - 90);
      if (upA) edge_upsample(above, w + (pAngle < 90 ? h : 0));
      upL = use_upsample(w, h, a->filterType, pAngle - 180);
      if (upL) edge_upsample(left, h + (pAngle > 180 ? w : 0));
    }
    int dx = 0, dy = 0;
    if (pAngle < 90) dx = Dr_Intra_Derivative[pAngle];
    else if (pAngle > 90 && pAngle < 180) dx = Dr_Intra_Derivative[180 - pAngle];
    if (pAngle > 90 && pAngle < 180) dy = Dr_Intra_Derivative[pAngle - 90];
    else if (pAngle > 180) dy = Dr_Intra_Derivative[270 - pAngle];
    for (int i = 0; i < h; i++)
      for (int j = 0; j < w; j++) {
        int v;
        if (pAngle == 90) {
          v = above[j];
        } else if (pAngle == 180) {
          v = left[i];
        } else if (pAngle < 90) {
          int idx = (i + 1) * dx, base = (idx >> (6 - upA)) + (j << upA), sh = ((idx * (1 << upA)) >> 1) & 31;
          int maxBase = (w + h - 1) << upA;
          v = base < maxBase ? o_round2(above[base] * (32 - sh) + above[base + 1] * sh, 5) : above[maxBase];
        } else if (pAngle < 180) {
          int idx = (j << 6) - (i + 1) * dx, base = idx >> (6 - upA);
          if (base >= -(1 << upA)) {
            int sh = ((idx * (1 << upA)) >> 1) & 31;
            v = o_round2(above[base] * (32 - sh) + above[base + 1] * sh, 5);
          } else {
            idx = (i << 6) - (j + 1) * dy;
            base = idx >> (6 - upL);
            int sh = ((idx * (1 << upL)) >> 1) & 31;
            v = o_round2(left[base] * (32 - sh) + left[base + 1] * sh, 5);
          }
        } else {
          int idx = (j + 1) * dy, base = (idx >> (6 - upL)) + (i << upL), sh = ((idx * (1 << upL)) >> 1) & 31;
          v = o_round2(left[base] * (32 - sh) + left[base + 1] * sh, 5);
        }
        dst[i * dstStride + j] = (uint8_t)v;
      }
    return;
  }
  const uint8_t *wx = sm_weights(a->log2W), *wy = sm_weights(a->log2H);
  if (mode == SMOOTH_PRED || mode == SMOOTH_V_PRED || mode == SMOOTH_H_PRED) {
    for (int i = 0; i < h; i++)
      for (int j = 0; j < w; j++) {
        int v;
        if (mode == SMOOTH_PRED)
          v = o_round2(wy[i] * above[j] + (256 - wy[i]) * left[h - 1] + wx[j] * left[i] + (256 - wx[j]) * above[w - 1], 9);
        else if (mode == SMOOTH_V_PRED)
          v = o_round2(wy[i] * above[j] + (256 - wy[i]) * left[h - 1], 8);
        else
          v = o_round2(wx[j] * left[i] + (256 - wx[j]) * above[w - 1], 8);
        dst[i * dstStride + j] = (uint8_t)v;
      }
    return;
  }
  if (mode == DC_PRED) {
    int avg = 128, s = 0;
    if (ha && hl) {
      for (int k = 0; k < h; k++) s += left[k];
      for (int k = 0; k < w; k++) s += above[k];
      avg = (s + ((w + h) >> 1)) / (w + h);
    } else if (hl) {
      for (int k = 0; k < h; k++) s += left[k];
      avg = o_clip1((s + (h >> 1)) >> a->log2H);
    } else if (ha) {
      for (int k = 0; k < w; k++) s += above[k];
      avg = o_clip1((s + (w >> 1)) >> a->log2W);
    }
    for (int i = 0; i < h; i++) memset(dst + i * dstStride, avg, (size_t)w);
    return;
  }
  /* PAETH */
  for (int i = 0; i < h; i++)
    for (int j = 0; j < w; j++) {
      int base = above[j] + left[i] - corner;
      int pl = o_abs(base - left[i]), pt = o_abs(base - above[j]), ptl = o_abs(base - corner);
      dst[i * dstStride + j] = (uint8_t)((pl <= pt && pl <= ptl) ? left[i] : pt <= ptl ? above[j] : corner);
    }
}

/* Luma average term of CfL: 2x2 sums << 1 of the co-located luma, with the
 * right / bottom replication of 7.11.5 (positions clamped to maxW-2 / maxH-2). */
static int cfl_luma(const uint8_t *luma, int ls, int sx, int sy, int maxW, int maxH, int i, int j) {
  int ly = o_min((sy + i) << 1, maxH - 2), lx = o_min((sx + j) << 1, maxW - 2);
  const uint8_t *l = luma + ly * ls + lx;
  return ((int)l[0] + l[1] + l[ls] + l[ls + 1]) << 1;
}

static int cfl_avg(const uint8_t *luma, int ls, int sx, int sy, int maxW, int maxH, int log2W, int log2H) {
  int s = 0;
  for (int i = 0; i < (1 << log2H); i++)
    for (int j = 0; j < (1 << log2W); j++) s += cfl_luma(luma, ls, sx, sy, maxW, maxH, i, j);
  return o_round2(s, log2W + log2H);
}

void o_cfl_predict(const uint8_t *luma, int ls, int sx, int sy, int maxW, int maxH, int log2W, int log2H, int alpha,
                   const int *dc, uint8_t *dst, int dstStride) {
  int w = 1 << log2W, h = 1 << log2H, avg = cfl_avg(luma, ls, sx, sy, maxW, maxH, log2W, log2H);
  for (int i = 0; i < h; i++)
    for (int j = 0; j < w; j++) {
      int v = cfl_luma(luma, ls, sx, sy, maxW, maxH, i, j) - avg;
      dst[i * dstStride + j] = (uint8_t)o_clip1(dc[i * w + j] + o_round2signed(alpha * v, 6));
    }
}

/* k_intra.h:277-313: alpha in [-16, 16] with the smallest SSE, first minimum. */
int o_cfl_best_alpha(const uint8_t *luma, int ls, int sx, int sy, int maxW, int maxH, int log2W, int log2H,
                     const int *dc, const uint8_t *src, int srcStride, int excludeZero) {
  int w = 1 << log2W, h = 1 << log2H, avg = cfl_avg(luma, ls, sx, sy, maxW, maxH, log2W, log2H);
  int64_t sse[33] = {0};
  for (int i = 0; i < h; i++)
    for (int j = 0; j < w; j++) {
      int v = cfl_luma(luma, ls, sx, sy, maxW, maxH, i, j) - avg, s0 = src[i * srcStride + j];
      for (int a = -16; a <= 16; a++) {
        int d = o_clip1(dc[i * w + j] + o_round2signed(a * v, 6)) - s0;
        sse[a + 16] += d * d;
      }
    }
  int64_t best = -1;
  int bestA = 0;
  for (int a = -16; a <= 16; a++) {
    if (a == 0 && excludeZero) continue;
    if (best < 0 || sse[a + 16] < best) {
      best = sse[a + 16];
      bestA = a;
    }
  }
  return bestA;
}

void oracle_predict_intra(const uint8_t *px, int stride, int x, int y, const OracleIntraArgs *a, uint8_t *dst,
                          int dstStride) {
  OIntraArgs b = {a->haveLeft, a->haveAbove, a->haveAboveRight, a->haveBelowLeft, a->mode, a->angleDelta,
                  a->log2W, a->log2H, a->maxX, a->maxY, a->edgeFilter, a->filterType};
  o_predict_intra(px, stride, x, y, &b, dst, dstStride);
}
