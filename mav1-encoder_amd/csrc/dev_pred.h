/* Wave-cooperative prediction helpers for the analysis and reconstruction
 * kernels: unscaled inter prediction (AV1 spec 7.11.3.3-7.11.3.4) through an
 * LDS-staged reference window, intra edge preparation (7.11.2 steps 1-3,
 * the edge filter and upsampling of 7.11.2.9-12) and the per-sample intra
 * predictor, and 4x4 Hadamard SATD over the units of a block.
 * All functions are called by every thread of a workgroup of NT threads
 * (a wave64 by default); `lane` is the thread's index in the workgroup. */
#ifndef MAV1_DEV_PRED_H
#define MAV1_DEV_PRED_H

#include "dev_common.h"

/* Sum over the NT threads of the workgroup; every thread gets it. Beyond
 * one wave, the waves' sums meet in LDS between two barriers (the first
 * keeps a previous sum's readers ahead of the writes). */
template <int NT>
DEVI int group_sum(int v) {
  v = wave_sum(v);
  if constexpr (NT == 64) {
    return v;
  } else {
    __shared__ int part[NT / 64];
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) t += part[i];
    return t;
  }
}

/* Inter prediction of a w x h block of one plane into dst (row stride ds).
 * win / mid are LDS scratch of (h + 7) x (w + 8) bytes and (h + 7) x w
 * int16; round1 11 writes clipped bytes to dst, round1 7 (compound) writes
 * the unclipped int32 values to dst32 instead. */
template <int NT = 64>
DEVI void wave_predict_inter(const uint8_t *__restrict__ plane, int S, int refW, int refH, int x, int y, int w, int h,
                             int mvr, int mvc, int sub, int filter, uint8_t *win, int16_t *mid, uint8_t *dst, int ds,
                             int32_t *dst32, int round1, int lane) {
  const int posX = ((x << SUBPEL_BITS) + ((2 * mvc) >> sub)) * 64 + 32;
  const int posY = ((y << SUBPEL_BITS) + ((2 * mvr) >> sub)) * 64 + 32;
  const int ix = posX >> 10, iy = posY >> 10;
  const int lastX = ((refW + sub) >> sub) - 1, lastY = ((refH + sub) >> sub) - 1;
  int fv = filter & 3, fh = (filter >> 2) & 3;
  if (w <= 4) fh = (fh == EIGHTTAP_SMOOTH) ? 5 : (fh == EIGHTTAP || fh == EIGHTTAP_SHARP) ? 4 : fh;
  if (h <= 4) fv = (fv == EIGHTTAP_SMOOTH) ? 5 : (fv == EIGHTTAP || fv == EIGHTTAP_SHARP) ? 4 : fv;
  int kh[8], kv[8];
  const int16_t *th = SUBPEL_FILTER(fh, (posX >> 6) & SUBPEL_MASK), *tv = SUBPEL_FILTER(fv, (posY >> 6) & SUBPEL_MASK);
#pragma unroll
  for (int t = 0; t < 8; t++) {
    kh[t] = th[t];
    kv[t] = tv[t];
  }
  const int WW = w + 8, nr = h + 7;
  for (int k = lane; k < nr * (w + 7); k += NT) {
    int r = k / (w + 7), c = k - r * (w + 7);
    win[r * WW + c] = plane[d_clip(0, lastY, iy - 3 + r) * S + d_clip(0, lastX, ix - 3 + c)];
  }
  __syncthreads();
  for (int k = lane; k < nr * w; k += NT) {
    int r = k / w, c = k - r * w;
    const uint8_t *q = win + r * WW + c;
    int acc = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) acc += kh[t] * q[t];
    mid[k] = (int16_t)d_round2(acc, 3);
  }
  __syncthreads();
  for (int k = lane; k < h * w; k += NT) {
    int r = k / w, c = k - r * w;
    int acc = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) acc += kv[t] * mid[(r + t) * w + c];
    if (round1 == 11)
      dst[r * ds + c] = (uint8_t)d_clip1(d_round2(acc, 11));
    else
      dst32[r * ds + c] = d_round2(acc, round1);
  }
  __syncthreads();
}

/* SATD of the 4x4 units of s against p (both row stride st); `sub` keeps the
 * units with even (row + column) parity and doubles the sum. */
template <int NT = 64>
DEVI int wave_satd(const uint8_t *s, const uint8_t *p, int st, int w, int h, int sub, int lane) {
  const int nx = w >> 2, nu = nx * (h >> 2);
  int acc = 0;
  for (int u = lane; u < nu; u += NT) {
    int uy = u / nx, ux = u - uy * nx;
    if (sub && ((ux + uy) & 1)) continue;
    const uint8_t *a = s + uy * 4 * st + ux * 4, *b = p + uy * 4 * st + ux * 4;
    int d[16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) d[4 * i + j] = (int)a[i * st + j] - (int)b[i * st + j];
    acc += satd4x4_arr(d);
  }
  return group_sum<NT>(acc) * (sub ? 2 : 1);
}

/* SATD of s against pa, pb and their rounded average, unit by unit. */
template <int NT = 64>
DEVI void wave_satd3(const uint8_t *s, const uint8_t *pa, const uint8_t *pb, int st, int w, int h, int sub, int want,
                     int lane, int *r0, int *r1, int *r2) {
  const int nx = w >> 2, nu = nx * (h >> 2);
  int a0 = 0, a1 = 0, a2 = 0;
  for (int u = lane; u < nu; u += NT) {
    int uy = u / nx, ux = u - uy * nx;
    if (sub && ((ux + uy) & 1)) continue;
    int o = uy * 4 * st + ux * 4;
    int d0[16], d1[16], d2[16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        int sv = s[o + i * st + j], x0 = pa[o + i * st + j], x1 = pb[o + i * st + j];
        d0[4 * i + j] = sv - x0;
        d1[4 * i + j] = sv - x1;
        d2[4 * i + j] = sv - ((x0 + x1 + 1) >> 1);
      }
    a0 += satd4x4_arr(d0);
    if (want > 1) a1 += satd4x4_arr(d1);
    if (want > 2) a2 += satd4x4_arr(d2);
  }
  int m = sub ? 2 : 1;
  *r0 = group_sum<NT>(a0) * m;
  *r1 = want > 1 ? group_sum<NT>(a1) * m : 0;
  *r2 = want > 2 ? group_sum<NT>(a2) * m : 0;
}

/* ---- intra ---- */
#define EDGE_NEG 16 /* edge arrays are addressed from index -EDGE_NEG */

/* Neighbouring samples (7.11.2 steps 1-3) of a w x h block at plane (x, y):
 * A[0 .. w+h-1] above, L[0 .. w+h-1] left, A[-1] = L[-1] = the corner. */
template <int NT = 64>
DEVI void wave_intra_edges(const uint8_t *px, int S, int x, int y, int w, int h, int hl, int ha, int har, int hbl,
                           int maxX, int maxY, int *A, int *L, int lane) {
  const int aLim = d_min(maxX, x + (har ? 2 * w : w) - 1), lLim = d_min(maxY, y + (hbl ? 2 * h : h) - 1);
  for (int i = lane; i < w + h; i += NT) {
    A[i] = ha ? px[(y - 1) * S + d_min(aLim, x + i)] : hl ? px[y * S + x - 1] : 127;
    L[i] = hl ? px[d_min(lLim, y + i) * S + x - 1] : ha ? px[(y - 1) * S + x] : 129;
  }
  if (lane == 0) {
    int corner = (ha && hl) ? px[(y - 1) * S + x - 1] : ha ? px[(y - 1) * S + x] : hl ? px[y * S + x - 1] : 128;
    A[-1] = L[-1] = corner;
  }
  __syncthreads();
}

DEVI const uint8_t *sm_weights(int lg) {
  return lg == 2 ? Sm_Weights_Tx_4x4 : lg == 3 ? Sm_Weights_Tx_8x8 : lg == 4 ? Sm_Weights_Tx_16x16
         : lg == 5 ? Sm_Weights_Tx_32x32 : Sm_Weights_Tx_64x64;
}

/* Per-block constants of one intra mode (uniform over the wave). */
struct IntraMode {
  int mode, pAngle, dx, dy, upA, upL, dcv, lw, lh, corner;
};

DEVI void intra_mode_setup(IntraMode &m, int mode, int delta, int lw, int lh) {
  m.mode = mode;
  m.lw = lw;
  m.lh = lh;
  m.upA = m.upL = 0;
  m.dx = m.dy = 0;
  m.pAngle = 0;
  if (mode >= V_PRED && mode <= D67_PRED) {
    int pa = Mode_To_Angle[mode] + delta * ANGLE_STEP;
    m.pAngle = pa;
    if (pa < 90) m.dx = Dr_Intra_Derivative[pa];
    else if (pa > 90 && pa < 180) m.dx = Dr_Intra_Derivative[180 - pa];
    if (pa > 90 && pa < 180) m.dy = Dr_Intra_Derivative[pa - 90];
    else if (pa > 180) m.dy = Dr_Intra_Derivative[270 - pa];
  }
}

/* DC value (7.11.2.5 DC_PRED) from the edges; every lane gets it. */
template <int NT = 64>
DEVI int wave_intra_dc(const int *A, const int *L, int w, int h, int lw, int lh, int hl, int ha, int lane) {
  int s = 0;
  for (int i = lane; i < w; i += NT) s += ha ? A[i] : 0;
  for (int i = lane; i < h; i += NT) s += hl ? L[i] : 0;
  s = group_sum<NT>(s);
  if (ha && hl) return (s + ((w + h) >> 1)) / (w + h);
  if (hl) return d_clip1((s + (h >> 1)) >> lh);
  if (ha) return d_clip1((s + (w >> 1)) >> lw);
  return 128;
}

/* Predicted sample (i, j) of the block (7.11.2.4 - 7.11.2.7). */
DEVI int intra_px(const IntraMode &m, const int *A, const int *L, int i, int j) {
  const int w = 1 << m.lw, h = 1 << m.lh;
  switch (m.mode) {
    case DC_PRED: return m.dcv;
    case SMOOTH_PRED: {
      const uint8_t *wx = sm_weights(m.lw), *wy = sm_weights(m.lh);
      return d_round2(wy[i] * A[j] + (256 - wy[i]) * L[h - 1] + wx[j] * L[i] + (256 - wx[j]) * A[w - 1], 9);
    }
    case SMOOTH_V_PRED: {
      const uint8_t *wy = sm_weights(m.lh);
      return d_round2(wy[i] * A[j] + (256 - wy[i]) * L[h - 1], 8);
    }
    case SMOOTH_H_PRED: {
      const uint8_t *wx = sm_weights(m.lw);
      return d_round2(wx[j] * L[i] + (256 - wx[j]) * A[w - 1], 8);
    }
    case PAETH_PRED: {
      int base = A[j] + L[i] - m.corner;
      int pl = d_abs(base - L[i]), pt = d_abs(base - A[j]), ptl = d_abs(base - m.corner);
      return (pl <= pt && pl <= ptl) ? L[i] : pt <= ptl ? A[j] : m.corner;
    }
    default: break;
  }
  const int pa = m.pAngle, upA = m.upA, upL = m.upL;
  if (pa == 90) return A[j];
  if (pa == 180) return L[i];
  if (pa < 90) {
    int idx = (i + 1) * m.dx, base = (idx >> (6 - upA)) + (j << upA), sh = ((idx * (1 << upA)) >> 1) & 31;
    int maxBase = (w + h - 1) << upA;
    return base < maxBase ? d_round2(A[base] * (32 - sh) + A[base + 1] * sh, 5) : A[maxBase];
  }
  if (pa < 180) {
    int idx = (j << 6) - (i + 1) * m.dx, base = idx >> (6 - upA);
    if (base >= -(1 << upA)) {
      int sh = ((idx * (1 << upA)) >> 1) & 31;
      return d_round2(A[base] * (32 - sh) + A[base + 1] * sh, 5);
    }
    idx = (i << 6) - (j + 1) * m.dy;
    base = idx >> (6 - upL);
    int sh = ((idx * (1 << upL)) >> 1) & 31;
    return d_round2(L[base] * (32 - sh) + L[base + 1] * sh, 5);
  }
  int idx = (j + 1) * m.dy, base = (idx >> (6 - upL)) + (i << upL), sh = ((idx * (1 << upL)) >> 1) & 31;
  return d_round2(L[base] * (32 - sh) + L[base + 1] * sh, 5);
}

/* SATD of an intra mode's prediction against the source block s (stride st). */
template <int NT = 64>
DEVI int wave_intra_satd(const IntraMode &m, const int *A, const int *L, const uint8_t *s, int st, int lane) {
  const int w = 1 << m.lw, h = 1 << m.lh, nx = w >> 2, nu = nx * (h >> 2);
  int acc = 0;
  for (int u = lane; u < nu; u += NT) {
    int uy = u / nx, ux = u - uy * nx;
    int d[16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        int yy = uy * 4 + i, xx = ux * 4 + j;
        d[4 * i + j] = (int)s[yy * st + xx] - intra_px(m, A, L, yy, xx);
      }
    acc += satd4x4_arr(d);
  }
  return group_sum<NT>(acc);
}

#endif
