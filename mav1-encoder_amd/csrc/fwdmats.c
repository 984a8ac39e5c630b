/* Forward 1-D transform matrices in Q20 (host, once per context).
 *
 * The forward transform is not normative; to stay consistent with the
 * normative inverse transform, each forward kernel is the least-squares
 * inverse of the integer inverse kernel of the same kind and length:
 * A = responses of the inverse kernel to scaled unit impulses (n outputs x
 * min(32, n) inputs), F = (A^T A)^-1 A^T by Gauss-Jordan elimination with
 * partial pivoting in double precision, rounded to Q20 (behaviour of
 * src/fwd_matrices.c:17-65; layout FWD_SLOT of src/kernels/k_fwd.h:16-21:
 * slot (kind * 5 + log2n - 2) * 4096, kind 0 DCT, 1 ADST, 2 identity).
 * Compiled with -ffp-contract=off: the elimination's rounding sequence is
 * part of the result. */
#include <math.h>
#include <string.h>

#include "../../include/mav1_hip.h"
#include "av1_defs.h"
#include "av1_tables.h"

static int cos128(int angle) {
  int a = angle & 255;
  if (a <= 64) return Cos128_Lookup[a];
  if (a <= 128) return -Cos128_Lookup[128 - a];
  if (a <= 192) return -Cos128_Lookup[a - 128];
  return Cos128_Lookup[256 - a];
}
static int brev(int bits, int x) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}
static long long rnd12(long long x) { return (x + 2048) >> 12; }
static void rot(int32_t *T, int a, int b, int angle, int flip) {
  long long x = (long long)T[a] * cos128(angle) - (long long)T[b] * cos128(angle - 64);
  long long y = (long long)T[a] * cos128(angle - 64) + (long long)T[b] * cos128(angle);
  T[a] = (int32_t)rnd12(x);
  T[b] = (int32_t)rnd12(y);
  if (flip) {
    int32_t t = T[a];
    T[a] = T[b];
    T[b] = t;
  }
}
/* the matrices are derived with a 31-bit range: the Hadamard steps never clamp */
static void had(int32_t *T, int a, int b, int flip) {
  if (flip) {
    int t = a;
    a = b;
    b = t;
  }
  long long x = T[a], y = T[b], lo = -(1LL << 30), hi = (1LL << 30) - 1;
  long long s = x + y, d = x - y;
  T[a] = (int32_t)(s < lo ? lo : s > hi ? hi : s);
  T[b] = (int32_t)(d < lo ? lo : d > hi ? hi : d);
}

/* inverse DCT (AV1 spec 7.13.2.3) */
static void inv_dct(int32_t *T, int n) {
  int32_t c[64];
  int n0 = 1 << n;
  memcpy(c, T, sizeof(int32_t) * n0);
  for (int i = 0; i < n0; i++) T[i] = c[brev(n, i)];
  if (n == 6) for (int i = 0; i < 16; i++) rot(T, 32 + i, 63 - i, 63 - 4 * brev(4, i), 0);
  if (n >= 5) for (int i = 0; i < 8; i++) rot(T, 16 + i, 31 - i, 6 + (brev(3, 7 - i) << 3), 0);
  if (n == 6) for (int i = 0; i < 16; i++) had(T, 32 + i * 2, 33 + i * 2, i & 1);
  if (n >= 4) for (int i = 0; i < 4; i++) rot(T, 8 + i, 15 - i, 12 + (brev(2, 3 - i) << 4), 0);
  if (n >= 5) for (int i = 0; i < 8; i++) had(T, 16 + 2 * i, 17 + 2 * i, i & 1);
  if (n == 6)
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++) rot(T, 62 - i * 4 - j, 33 + i * 4 + j, 60 - 16 * brev(2, i) + 64 * j, 1);
  if (n >= 3) for (int i = 0; i < 2; i++) rot(T, 4 + i, 7 - i, 56 - 32 * i, 0);
  if (n >= 4) for (int i = 0; i < 4; i++) had(T, 8 + 2 * i, 9 + 2 * i, i & 1);
  if (n >= 5)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) rot(T, 30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1);
  if (n == 6)
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 2; j++) had(T, 32 + i * 4 + j, 35 + i * 4 - j, i & 1);
  for (int i = 0; i < 2; i++) rot(T, 2 * i, 2 * i + 1, 32 + 16 * i, 1 - i);
  if (n >= 3) for (int i = 0; i < 2; i++) had(T, 4 + 2 * i, 5 + 2 * i, i);
  if (n >= 4) for (int i = 0; i < 2; i++) rot(T, 14 - i, 9 + i, 48 + 64 * i, 1);
  if (n >= 5)
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++) had(T, 16 + 4 * i + j, 19 + 4 * i - j, i & 1);
  if (n == 6)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++) rot(T, 61 - i * 8 - j, 34 + i * 8 + j, 56 - i * 32 + (j >> 1) * 64, 1);
  for (int i = 0; i < 2; i++) had(T, i, 3 - i, 0);
  if (n >= 3) rot(T, 6, 5, 32, 1);
  if (n >= 4)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) had(T, 8 + 4 * i + j, 11 + 4 * i - j, i);
  if (n >= 5) for (int i = 0; i < 4; i++) rot(T, 29 - i, 18 + i, 48 + (i >> 1) * 64, 1);
  if (n == 6)
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) had(T, 32 + 8 * i + j, 39 + 8 * i - j, i & 1);
  if (n >= 3) for (int i = 0; i < 4; i++) had(T, i, 7 - i, 0);
  if (n >= 4) for (int i = 0; i < 2; i++) rot(T, 13 - i, 10 + i, 32, 1);
  if (n >= 5)
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++) had(T, 16 + i * 8 + j, 23 + i * 8 - j, i);
  if (n == 6) for (int i = 0; i < 8; i++) rot(T, 59 - i, 36 + i, i < 4 ? 48 : 112, 1);
  if (n >= 4) for (int i = 0; i < 8; i++) had(T, i, 15 - i, 0);
  if (n >= 5) for (int i = 0; i < 4; i++) rot(T, 27 - i, 20 + i, 32, 1);
  if (n == 6)
    for (int i = 0; i < 8; i++) {
      had(T, 32 + i, 47 - i, 0);
      had(T, 48 + i, 63 - i, 1);
    }
  if (n >= 5) for (int i = 0; i < 16; i++) had(T, i, 31 - i, 0);
  if (n == 6) for (int i = 0; i < 8; i++) rot(T, 55 - i, 40 + i, 32, 1);
  if (n == 6) for (int i = 0; i < 32; i++) had(T, i, 63 - i, 0);
}

static void adst_perm_in(int32_t *T, int n) {
  int32_t c[16];
  int n0 = 1 << n;
  memcpy(c, T, sizeof(int32_t) * n0);
  for (int i = 0; i < n0; i++) T[i] = c[(i & 1) ? i - 1 : n0 - i - 1];
}
static void adst_perm_out(int32_t *T, int n) {
  int32_t c[16];
  int n0 = 1 << n;
  memcpy(c, T, sizeof(int32_t) * n0);
  for (int i = 0; i < n0; i++) {
    int a = (i >> 3) & 1, b = ((i >> 2) & 1) ^ a, cc = ((i >> 1) & 1) ^ ((i >> 2) & 1), d = (i & 1) ^ ((i >> 1) & 1);
    int idx = ((d << 3) | (cc << 2) | (b << 1) | a) >> (4 - n);
    T[i] = (i & 1) ? -c[idx] : c[idx];
  }
}

/* inverse ADST (7.13.2.5 - 7.13.2.9) */
static void inv_adst(int32_t *T, int n) {
  if (n == 2) {
    long long s0 = 1321LL * T[0], s1 = 2482LL * T[0], s2 = 3344LL * T[1], s3 = 3803LL * T[2], s4 = 1321LL * T[2],
              s5 = 2482LL * T[3], s6 = 3803LL * T[3], b7 = (long long)T[0] - T[2] + T[3];
    s0 += s3;
    s1 -= s4;
    s3 = s2;
    s2 = 3344LL * b7;
    s0 += s5;
    s1 -= s6;
    T[0] = (int32_t)rnd12(s0 + s3);
    T[1] = (int32_t)rnd12(s1 + s3);
    T[2] = (int32_t)rnd12(s2);
    T[3] = (int32_t)rnd12(s0 + s1 - s3);
    return;
  }
  adst_perm_in(T, n);
  if (n == 3) {
    for (int i = 0; i < 4; i++) rot(T, 2 * i, 2 * i + 1, 60 - 16 * i, 1);
    for (int i = 0; i < 4; i++) had(T, i, 4 + i, 0);
    for (int i = 0; i < 2; i++) rot(T, 4 + 3 * i, 5 + i, 48 - 32 * i, 1);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) had(T, 4 * j + i, 2 + 4 * j + i, 0);
    for (int i = 0; i < 2; i++) rot(T, 2 + 4 * i, 3 + 4 * i, 32, 1);
  } else {
    for (int i = 0; i < 8; i++) rot(T, 2 * i, 2 * i + 1, 62 - 8 * i, 1);
    for (int i = 0; i < 8; i++) had(T, i, 8 + i, 0);
    for (int i = 0; i < 2; i++) {
      rot(T, 8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1);
      rot(T, 13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1);
    }
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++) had(T, 8 * j + i, 4 + 8 * j + i, 0);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) rot(T, 4 + 8 * j + 3 * i, 5 + 8 * j + i, 48 - 32 * i, 1);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++) had(T, 4 * j + i, 2 + 4 * j + i, 0);
    for (int i = 0; i < 4; i++) rot(T, 2 + 4 * i, 3 + 4 * i, 32, 1);
  }
  adst_perm_out(T, n);
}

static void inv_identity(int32_t *T, int n) {
  for (int i = 0; i < (1 << n); i++)
    T[i] = n == 2 ? (int32_t)rnd12((long long)T[i] * 5793) : n == 3 ? T[i] * 2
         : n == 4 ? (int32_t)rnd12((long long)T[i] * 11586) : T[i] * 4;
}

static void derive(int kind, int log2n, int32_t *out) {
  const int n = 1 << log2n, m = n > 32 ? 32 : n;
  double A[64][32], M[32][96];
  const double S = 65536.0;
  for (int j = 0; j < m; j++) {
    int32_t T[64] = {0};
    T[j] = (int32_t)S;
    if (kind == 0) inv_dct(T, log2n);
    else if (kind == 1) inv_adst(T, log2n);
    else inv_identity(T, log2n);
    for (int i = 0; i < n; i++) A[i][j] = T[i] / S;
  }
  /* [A^T A | A^T] */
  for (int r = 0; r < m; r++) {
    for (int c = 0; c < m; c++) {
      double s = 0;
      for (int i = 0; i < n; i++) s += A[i][r] * A[i][c];
      M[r][c] = s;
    }
    for (int i = 0; i < n; i++) M[r][m + i] = A[i][r];
  }
  for (int p = 0; p < m; p++) {
    int piv = p;
    for (int r = p + 1; r < m; r++)
      if (fabs(M[r][p]) > fabs(M[piv][p])) piv = r;
    if (piv != p)
      for (int c = 0; c < m + n; c++) {
        double t = M[p][c];
        M[p][c] = M[piv][c];
        M[piv][c] = t;
      }
    double inv = 1.0 / M[p][p];
    for (int c = 0; c < m + n; c++) M[p][c] *= inv;
    for (int r = 0; r < m; r++) {
      if (r == p || M[r][p] == 0) continue;
      double f = M[r][p];
      for (int c = 0; c < m + n; c++) M[r][c] -= f * M[p][c];
    }
  }
  for (int k = 0; k < m; k++)
    for (int i = 0; i < n; i++) out[k * n + i] = (int32_t)lrint(M[k][m + i] * (double)(1 << 20));
}

void mav1hip_forward_matrices(int32_t *mats) {
  memset(mats, 0, sizeof(int32_t) * 4096 * 15);
  for (int kind = 0; kind < 3; kind++)
    for (int log2n = 2; log2n <= 6; log2n++) {
      if (kind == 1 && log2n > 4) continue;
      if (kind == 2 && log2n > 5) continue;
      derive(kind, log2n, mats + (kind * 5 + log2n - 2) * 4096);
    }
}
