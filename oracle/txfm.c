/* ORACLE — test infrastructure only. Transforms and quantisation:
 * inverse transforms of AV1 spec 7.13 (src/kernels/k_itx.h), the forward
 * transform through the Q20 matrices and the dead-zone quantiser
 * (src/kernels/k_fwd.h). */
#include "ocommon.h"
#include "otx.h"

static int cos128(int angle) {
  int a = angle & 255;
  if (a <= 64) return Cos128_Lookup[a];
  if (a <= 128) return -Cos128_Lookup[128 - a];
  if (a <= 192) return -Cos128_Lookup[a - 128];
  return Cos128_Lookup[256 - a];
}
static int sin128(int angle) { return cos128(angle - 64); }

static int brev(int bits, int x) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

/* butterfly rotation B(a, b, angle, flip) (7.13.2.2) */
static void B(int32_t *T, int a, int b, int angle, int flip) {
  int64_t x = (int64_t)T[a] * cos128(angle) - (int64_t)T[b] * sin128(angle);
  int64_t y = (int64_t)T[a] * sin128(angle) + (int64_t)T[b] * cos128(angle);
  T[a] = (int32_t)o_round2l(x, 12);
  T[b] = (int32_t)o_round2l(y, 12);
  if (flip) {
    int32_t t = T[a];
    T[a] = T[b];
    T[b] = t;
  }
}

/* Hadamard rotation H(a, b, flip, r) with r-bit clamping */
static void Hr(int32_t *T, int a, int b, int flip, int r) {
  if (flip) {
    int t = a;
    a = b;
    b = t;
  }
  int32_t x = T[a], y = T[b];
  T[a] = o_clip(-(1 << (r - 1)), (1 << (r - 1)) - 1, x + y);
  T[b] = o_clip(-(1 << (r - 1)), (1 << (r - 1)) - 1, x - y);
}

/* inverse DCT array process (7.13.2.3), length 1 << n */
static void idct(int32_t *T, int n, int r) {
  int32_t c[64];
  int n0 = 1 << n;
  for (int i = 0; i < n0; i++) c[i] = T[i];
  for (int i = 0; i < n0; i++) T[i] = c[brev(n, i)];
  if (n == 6) for (int i = 0; i < 16; i++) B(T, 32 + i, 63 - i, 63 - 4 * brev(4, i), 0);
  if (n >= 5) for (int i = 0; i < 8; i++) B(T, 16 + i, 31 - i, 6 + (brev(3, 7 - i) << 3), 0);
  if (n == 6) for (int i = 0; i < 16; i++) Hr(T, 32 + i * 2, 33 + i * 2, i & 1, r);
  if (n >= 4) for (int i = 0; i < 4; i++) B(T, 8 + i, 15 - i, 12 + (brev(2, 3 - i) << 4), 0);
  if (n >= 5) for (int i = 0; i < 8; i++) Hr(T, 16 + 2 * i, 17 + 2 * i, i & 1, r);
  if (n == 6)
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++) B(T, 62 - i * 4 - j, 33 + i * 4 + j, 60 - 16 * brev(2, i) + 64 * j, 1);
  if (n >= 3) for (int i = 0; i < 2; i++) B(T, 4 + i, 7 - i, 56 - 32 * i, 0);
  if (n >= 4) for (int i = 0; i < 4; i++) Hr(T, 8 + 2 * i, 9 + 2 * i, i & 1, r);
  if (n >= 5)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) B(T, 30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1);
  if (n == 6)
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 2; j++) Hr(T, 32 + i * 4 + j, 35 + i * 4 - j, i & 1, r);
  for (int i = 0; i < 2; i++) B(T, 2 * i, 2 * i + 1, 32 + 16 * i, 1 - i);
  if (n >= 3) for (int i = 0; i < 2; i++) Hr(T, 4 + 2 * i, 5 + 2 * i, i, r);
  if (n >= 4) for (int i = 0; i < 2; i++) B(T, 14 - i, 9 + i, 48 + 64 * i, 1);
  if (n >= 5)
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++) Hr(T, 16 + 4 * i + j, 19 + 4 * i - j, i & 1, r);
  if (n == 6)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++) B(T, 61 - i * 8 - j, 34 + i * 8 + j, 56 - i * 32 + (j >> 1) * 64, 1);
  for (int i = 0; i < 2; i++) Hr(T, i, 3 - i, 0, r);
  if (n >= 3) B(T, 6, 5, 32, 1);
  if (n >= 4)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) Hr(T, 8 + 4 * i + j, 11 + 4 * i - j, i, r);
  if (n >= 5) for (int i = 0; i < 4; i++) B(T, 29 - i, 18 + i, 48 + (i >> 1) * 64, 1);
  if (n == 6)
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) Hr(T, 32 + 8 * i + j, 39 + 8 * i - j, i & 1, r);
  if (n >= 3) for (int i = 0; i < 4; i++) Hr(T, i, 7 - i, 0, r);
  if (n >= 4) for (int i = 0; i < 2; i++) B(T, 13 - i, 10 + i, 32, 1);
  if (n >= 5)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++) Hr(T, 16 + i * 8 + j, 23 + i * 8 - j, i, r);
  if (n == 6) for (int i = 0; i < 8; i++) B(T, 59 - i, 36 + i, i < 4 ? 48 : 112, 1);
  if (n >= 4) for (int i = 0; i < 8; i++) Hr(T, i, 15 - i, 0, r);
  if (n >= 5) for (int i = 0; i < 4; i++) B(T, 27 - i, 20 + i, 32, 1);
  if (n == 6)
    for (int i = 0; i < 8; i++) {
      Hr(T, 32 + i, 47 - i, 0, r);
      Hr(T, 48 + i, 63 - i, 1, r);
    }
  if (n >= 5) for (int i = 0; i < 16; i++) Hr(T, i, 31 - i, 0, r);
  if (n == 6) for (int i = 0; i < 8; i++) B(T, 55 - i, 40 + i, 32, 1);
  if (n == 6) for (int i = 0; i < 32; i++) Hr(T, i, 63 - i, 0, r);
}

static void adst_in(int32_t *T, int n) {
  int32_t c[16];
  int n0 = 1 << n;
  for (int i = 0; i < n0; i++) c[i] = T[i];
  for (int i = 0; i < n0; i++) T[i] = c[(i & 1) ? i - 1 : n0 - i - 1];
}

static void adst_out(int32_t *T, int n) {
  int32_t c[16];
  int n0 = 1 << n;
  for (int i = 0; i < n0; i++) c[i] = T[i];
  for (int i = 0; i < n0; i++) {
    int a = (i >> 3) & 1, b = ((i >> 2) & 1) ^ ((i >> 3) & 1), cc = ((i >> 1) & 1) ^ ((i >> 2) & 1),
        d = (i & 1) ^ ((i >> 1) & 1);
    int idx = ((d << 3) | (cc << 2) | (b << 1) | a) >> (4 - n);
    T[i] = (i & 1) ? -c[idx] : c[idx];
  }
}

/* inverse ADST4 (7.13.2.6) */
static void iadst4(int32_t *T) {
  const int64_t S1 = 1321, S2 = 2482, S3 = 3344, S4 = 3803;
  int64_t s0 = S1 * T[0], s1 = S2 * T[0], s2 = S3 * T[1], s3 = S4 * T[2], s4 = S1 * T[2], s5 = S2 * T[3],
          s6 = S4 * T[3];
  int64_t b7 = (int64_t)T[0] - T[2] + T[3];
  s0 += s3;
  s1 -= s4;
  s3 = s2;
  s2 = S3 * b7;
  s0 += s5;
  s1 -= s6;
  int64_t x0 = s0 + s3, x1 = s1 + s3, x2 = s2, x3 = s0 + s1 - s3;
  T[0] = (int32_t)o_round2l(x0, 12);
  T[1] = (int32_t)o_round2l(x1, 12);
  T[2] = (int32_t)o_round2l(x2, 12);
  T[3] = (int32_t)o_round2l(x3, 12);
}

static void iadst8(int32_t *T, int r) {
  adst_in(T, 3);
  for (int i = 0; i < 4; i++) B(T, 2 * i, 2 * i + 1, 60 - 16 * i, 1);
  for (int i = 0; i < 4; i++) Hr(T, i, 4 + i, 0, r);
  for (int i = 0; i < 2; i++) B(T, 4 + 3 * i, 5 + i, 48 - 32 * i, 1);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) Hr(T, 4 * j + i, 2 + 4 * j + i, 0, r);
  for (int i = 0; i < 2; i++) B(T, 2 + 4 * i, 3 + 4 * i, 32, 1);
  adst_out(T, 3);
}

static void iadst16(int32_t *T, int r) {
  adst_in(T, 4);
  for (int i = 0; i < 8; i++) B(T, 2 * i, 2 * i + 1, 62 - 8 * i, 1);
  for (int i = 0; i < 8; i++) Hr(T, i, 8 + i, 0, r);
  for (int i = 0; i < 2; i++) {
    B(T, 8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1);
    B(T, 13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1);
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 2; j++) Hr(T, 8 * j + i, 4 + 8 * j + i, 0, r);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) B(T, 4 + 8 * j + 3 * i, 5 + 8 * j + i, 48 - 32 * i, 1);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 4; j++) Hr(T, 4 * j + i, 2 + 4 * j + i, 0, r);
  for (int i = 0; i < 4; i++) B(T, 2 + 4 * i, 3 + 4 * i, 32, 1);
  adst_out(T, 4);
}

static void iidentity(int32_t *T, int n) {
  for (int i = 0; i < (1 << n); i++) {
    if (n == 2) T[i] = (int32_t)o_round2l((int64_t)T[i] * 5793, 12);
    else if (n == 3) T[i] = T[i] * 2;
    else if (n == 4) T[i] = (int32_t)o_round2l((int64_t)T[i] * 11586, 12);
    else T[i] = T[i] * 4;
  }
}

/* inverse WHT (7.13.2.10) */
static void iwht(int32_t *T, int shift) {
  int32_t a = T[0] >> shift, c = T[1] >> shift, d = T[2] >> shift, b = T[3] >> shift;
  a += c;
  d -= b;
  int32_t e = (a - d) >> 1;
  b = e - b;
  c = e - c;
  a -= b;
  d += c;
  T[0] = a;
  T[1] = b;
  T[2] = c;
  T[3] = d;
}

int o_row_kind(int t) {
  if (t == DCT_DCT || t == ADST_DCT || t == FLIPADST_DCT || t == H_DCT) return 0;
  if (t == DCT_ADST || t == ADST_ADST || t == DCT_FLIPADST || t == FLIPADST_FLIPADST || t == ADST_FLIPADST ||
      t == FLIPADST_ADST || t == H_ADST || t == H_FLIPADST)
    return 1;
  return 2;
}
int o_col_kind(int t) {
  if (t == DCT_DCT || t == DCT_ADST || t == DCT_FLIPADST || t == V_DCT) return 0;
  if (t == ADST_DCT || t == ADST_ADST || t == FLIPADST_DCT || t == FLIPADST_FLIPADST || t == ADST_FLIPADST ||
      t == FLIPADST_ADST || t == V_ADST || t == V_FLIPADST)
    return 1;
  return 2;
}
int o_flip_ud(int t) { return t == FLIPADST_DCT || t == FLIPADST_ADST || t == V_FLIPADST || t == FLIPADST_FLIPADST; }
int o_flip_lr(int t) { return t == DCT_FLIPADST || t == ADST_FLIPADST || t == H_FLIPADST || t == FLIPADST_FLIPADST; }

void o_itx_1d(int32_t *T, int kind, int n, int r) {
  if (kind == 0) idct(T, n, r);
  else if (kind == 1) {
    if (n == 2) iadst4(T);
    else if (n == 3) iadst8(T, r);
    else iadst16(T, r);
  } else
    iidentity(T, n);
}

/* 2-D inverse transform (7.13.3) as src/kernels/k_itx.h:262-296: dq holds
 * the dequantised Min(32,w) x Min(32,h) block (row stride 64); res gets h
 * rows of w residuals (row stride 64). */
void o_itx_2d(const int32_t *dq, int32_t *res, int txSz, int txType, int lossless) {
  int lw = Tx_Width_Log2[txSz], lh = Tx_Height_Log2[txSz], w = 1 << lw, h = 1 << lh;
  int rowShift = lossless ? 0 : Transform_Row_Shift[txSz], colShift = lossless ? 0 : 4;
  int rk = o_row_kind(txType), ck = o_col_kind(txType);
  int32_t T[64];
  for (int i = 0; i < h; i++) {
    for (int j = 0; j < w; j++) T[j] = (i < 32 && j < 32) ? dq[i * 64 + j] : 0;
    if (o_abs(lw - lh) == 1)
      for (int j = 0; j < w; j++) T[j] = (int32_t)o_round2l((int64_t)T[j] * 2896, 12);
    if (lossless) iwht(T, 2);
    else o_itx_1d(T, rk, lw, 16);
    for (int j = 0; j < w; j++) res[i * 64 + j] = o_round2(T[j], rowShift);
  }
  for (int i = 0; i < h; i++)
    for (int j = 0; j < w; j++) res[i * 64 + j] = o_clip(-(1 << 15), (1 << 15) - 1, res[i * 64 + j]);
  for (int j = 0; j < w; j++) {
    for (int i = 0; i < h; i++) T[i] = res[i * 64 + j];
    if (lossless) iwht(T, 0);
    else o_itx_1d(T, ck, lh, 16);
    for (int i = 0; i < h; i++) res[i * 64 + j] = o_round2(T[i], colShift);
  }
}

/* Forward 2-D transform (k_fwd.h:35-59): row pass through the row kind's
 * Q20 matrix (int64 sums, Round2 by 16), column pass through the column
 * kind's, scaled by 2^(row shift + 4), the 5793 / 4096 rectangular
 * correction, signed Round2 by 20. Output Min(32,h) x Min(32,w), stride 64. */
void o_fwd_2d(const int32_t *resid, int32_t *out, int txSz, int txType, const int32_t *mats) {
  int lw = Tx_Width_Log2[txSz], lh = Tx_Height_Log2[txSz], w = 1 << lw, h = 1 << lh;
  int tw = o_min(32, w), th = o_min(32, h);
  const int32_t *Mh = mats + ((o_row_kind(txType) * 5 + lw - 2) * 4096);
  const int32_t *Mv = mats + ((o_col_kind(txType) * 5 + lh - 2) * 4096);
  int32_t tmp[64 * 64];
  for (int i = 0; i < h; i++)
    for (int k = 0; k < tw; k++) {
      int64_t s = 0;
      for (int j = 0; j < w; j++) s += (int64_t)resid[i * 64 + j] * Mh[k * w + j];
      tmp[i * 64 + k] = (int32_t)o_round2l(s, 16);
    }
  int sh = Transform_Row_Shift[txSz] + 4, rect = o_abs(lw - lh) == 1;
  for (int k2 = 0; k2 < th; k2++)
    for (int k = 0; k < tw; k++) {
      int64_t s = 0;
      for (int i = 0; i < h; i++) s += (int64_t)Mv[k2 * h + i] * tmp[i * 64 + k];
      s *= (int64_t)1 << sh;
      if (rect) s = o_round2l(s, 12) * 5793;
      out[k2 * 64 + k] = (int32_t)(s >= 0 ? o_round2l(s, 20) : -o_round2l(-s, 20));
    }
}

int o_dq_denom(int txSz) {
  if (txSz == TX_32X32 || txSz == TX_16X32 || txSz == TX_32X16 || txSz == TX_16X64 || txSz == TX_64X16) return 2;
  if (txSz == TX_64X64 || txSz == TX_32X64 || txSz == TX_64X32) return 4;
  return 1;
}

/* Dead-zone quantiser (k_fwd.h:66-90): level = (|v| * denom + q * rnd / 4) / 16 / q
 * (rnd 26 DC, 22 AC), clamped to 0xFFFF; written in decoder layout (stride tw).
 * Computed with a plain division, which the reference's reciprocal equals. */
void o_quantize(const int32_t *coef, int32_t *quant, int txSz, int dcQ, int acQ) {
  int tw = o_min(32, Tx_Width[txSz]), th = o_min(32, Tx_Height[txSz]), denom = o_dq_denom(txSz);
  for (int i = 0; i < th; i++)
    for (int j = 0; j < tw; j++) {
      int dc = i == 0 && j == 0, q = dc ? dcQ : acQ, rnd = dc ? 26 : 22, v = coef[i * 64 + j];
      int64_t t = ((int64_t)o_abs(v) * denom + (((int64_t)q * rnd) >> 2)) >> 4;
      int64_t level = t / q;
      if (level > 0xFFFF) level = 0xFFFF;
      quant[i * tw + j] = v < 0 ? -(int32_t)level : (int32_t)level;
    }
}

/* Dequantisation (spec 7.12.3 step 1, no quantizer matrix), k_fwd.h:95-107. */
void o_dequant(const int32_t *quant, int32_t *dq, int txSz, int dcQ, int acQ) {
  int tw = o_min(32, Tx_Width[txSz]), th = o_min(32, Tx_Height[txSz]), denom = o_dq_denom(txSz);
  for (int i = 0; i < th; i++)
    for (int j = 0; j < tw; j++) {
      int q = (i == 0 && j == 0) ? dcQ : acQ;
      int64_t d = (int64_t)quant[i * tw + j] * q, ad = d < 0 ? -d : d;
      int64_t d2 = (ad & 0xFFFFFF) >> (denom >> 1);
      dq[i * 64 + j] = o_clip(-(1 << 15), (1 << 15) - 1, (int)(d < 0 ? -d2 : d2));
    }
}

/* Forward WHT of a lossless 4x4 block (k_fwd.h:112-136). */
void o_fwht4x4(const int32_t *resid, int32_t *out) {
  int t[16];
  for (int j = 0; j < 4; j++) {
    int a = resid[j], b = resid[64 + j], c = resid[128 + j], d = resid[192 + j];
    a += b;
    d -= c;
    int e = (a - d) >> 1;
    b = e - b;
    c = e - c;
    a -= c;
    d += b;
    t[j] = a;
    t[4 + j] = c;
    t[8 + j] = d;
    t[12 + j] = b;
  }
  for (int i = 0; i < 4; i++) {
    int a = t[4 * i], b = t[4 * i + 1], c = t[4 * i + 2], d = t[4 * i + 3];
    a += b;
    d -= c;
    int e = (a - d) >> 1;
    b = e - b;
    c = e - c;
    a -= c;
    d += b;
    out[4 * i] = a;
    out[4 * i + 1] = c;
    out[4 * i + 2] = d;
    out[4 * i + 3] = b;
  }
}

/* exported for the transform tests */
void oracle_itx_2d(const int32_t *dq, int32_t *res, int txSz, int txType, int lossless) {
  o_itx_2d(dq, res, txSz, txType, lossless);
}
void oracle_fwd_2d(const int32_t *resid, int32_t *out, int txSz, int txType, const int32_t *mats) {
  o_fwd_2d(resid, out, txSz, txType, mats);
}
void oracle_quantize(const int32_t *coef, int32_t *quant, int txSz, int dcQ, int acQ) {
  o_quantize(coef, quant, txSz, dcQ, acQ);
}
void oracle_dequant(const int32_t *quant, int32_t *dq, int txSz, int dcQ, int acQ) { o_dequant(quant, dq, txSz, dcQ, acQ); }
