/* Device transforms: the AV1 inverse 1-D transforms (spec 7.13.2) as fully
 * unrolled templates whose working array stays in registers, and the
 * wave-cooperative 2-D forward / inverse transforms and quantiser of one
 * transform block (behaviour of src/kernels/k_itx.h and k_fwd.h). */
#ifndef MAV1_DEV_TX_H
#define MAV1_DEV_TX_H

#include "dev_common.h"

DEVI int c128(int angle) {
  int a = angle & 255;
  return a <= 64 ? Cos128_Lookup[a] : a <= 128 ? -Cos128_Lookup[128 - a] : a <= 192 ? -Cos128_Lookup[a - 128]
                                                                                     : Cos128_Lookup[256 - a];
}
__host__ __device__ constexpr int cbrev(int bits, int x) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}
/* The rotation B(a, b, angle, flip) in 32-bit arithmetic, equal to the
 * 64-bit products of the spec: every value a rotation reads is a 1-D
 * transform input (dequantised coefficients are clipped to 16 bits, the
 * column inputs clamped to 16 bits) or a Hadamard output (clamped to r = 16
 * bits) -- no rotation of the DCT / ADST networks reads another rotation's
 * output -- so |T[a]|, |T[b]| <= 2^15, |c|, |s| <= 4096 and every sum stays
 * below 2^28: the 24-bit multiplies (v_mul_i32_i24) and 32-bit sums are
 * exact. */
DEVI void tB(int32_t *T, int a, int b, int angle, int flip) {
  const int c = c128(angle), s = c128(angle - 64);
  const int32_t x = __mul24(T[a], c) - __mul24(T[b], s);
  const int32_t y = __mul24(T[a], s) + __mul24(T[b], c);
  const int32_t xa = (x + 2048) >> 12, yb = (y + 2048) >> 12;
  T[a] = flip ? yb : xa;
  T[b] = flip ? xa : yb;
}
DEVI void tH(int32_t *T, int a, int b, int flip, int r) {
  const int lo = -(1 << (r - 1)), hi = (1 << (r - 1)) - 1;
  int32_t x = T[flip ? b : a], y = T[flip ? a : b];
  int32_t s = d_clip(lo, hi, x + y), d = d_clip(lo, hi, x - y);
  if (flip) {
    T[b] = s;
    T[a] = d;
  } else {
    T[a] = s;
    T[b] = d;
  }
}

template <int n>
DEVI void t_idct(int32_t *T, int r) {
  constexpr int N0 = 1 << n;
  int32_t c[N0];
#pragma unroll
  for (int i = 0; i < N0; i++) c[i] = T[i];
#pragma unroll
  for (int i = 0; i < N0; i++) T[i] = c[cbrev(n, i)];
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 16; i++) tB(T, 32 + i, 63 - i, 63 - 4 * cbrev(4, i), 0);
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 8; i++) tB(T, 16 + i, 31 - i, 6 + (cbrev(3, 7 - i) << 3), 0);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 16; i++) tH(T, 32 + i * 2, 33 + i * 2, i & 1, r);
  }
  if constexpr (n >= 4) {
#pragma unroll
    for (int i = 0; i < 4; i++) tB(T, 8 + i, 15 - i, 12 + (cbrev(2, 3 - i) << 4), 0);
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 8; i++) tH(T, 16 + 2 * i, 17 + 2 * i, i & 1, r);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) tB(T, 62 - i * 4 - j, 33 + i * 4 + j, 60 - 16 * cbrev(2, i) + 64 * j, 1);
  }
  if constexpr (n >= 3) {
#pragma unroll
    for (int i = 0; i < 2; i++) tB(T, 4 + i, 7 - i, 56 - 32 * i, 0);
  }
  if constexpr (n >= 4) {
#pragma unroll
    for (int i = 0; i < 4; i++) tH(T, 8 + 2 * i, 9 + 2 * i, i & 1, r);
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) tB(T, 30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) tH(T, 32 + i * 4 + j, 35 + i * 4 - j, i & 1, r);
  }
#pragma unroll
  for (int i = 0; i < 2; i++) tB(T, 2 * i, 2 * i + 1, 32 + 16 * i, 1 - i);
  if constexpr (n >= 3) {
#pragma unroll
    for (int i = 0; i < 2; i++) tH(T, 4 + 2 * i, 5 + 2 * i, i, r);
  }
  if constexpr (n >= 4) {
#pragma unroll
    for (int i = 0; i < 2; i++) tB(T, 14 - i, 9 + i, 48 + 64 * i, 1);
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) tH(T, 16 + 4 * i + j, 19 + 4 * i - j, i & 1, r);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) tB(T, 61 - i * 8 - j, 34 + i * 8 + j, 56 - i * 32 + (j >> 1) * 64, 1);
  }
#pragma unroll
  for (int i = 0; i < 2; i++) tH(T, i, 3 - i, 0, r);
  if constexpr (n >= 3) tB(T, 6, 5, 32, 1);
  if constexpr (n >= 4) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) tH(T, 8 + 4 * i + j, 11 + 4 * i - j, i, r);
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 4; i++) tB(T, 29 - i, 18 + i, 48 + (i >> 1) * 64, 1);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) tH(T, 32 + 8 * i + j, 39 + 8 * i - j, i & 1, r);
  }
  if constexpr (n >= 3) {
#pragma unroll
    for (int i = 0; i < 4; i++) tH(T, i, 7 - i, 0, r);
  }
  if constexpr (n >= 4) {
#pragma unroll
    for (int i = 0; i < 2; i++) tB(T, 13 - i, 10 + i, 32, 1);
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) tH(T, 16 + i * 8 + j, 23 + i * 8 - j, i, r);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; i++) tB(T, 59 - i, 36 + i, i < 4 ? 48 : 112, 1);
  }
  if constexpr (n >= 4) {
#pragma unroll
    for (int i = 0; i < 8; i++) tH(T, i, 15 - i, 0, r);
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 4; i++) tB(T, 27 - i, 20 + i, 32, 1);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      tH(T, 32 + i, 47 - i, 0, r);
      tH(T, 48 + i, 63 - i, 1, r);
    }
  }
  if constexpr (n >= 5) {
#pragma unroll
    for (int i = 0; i < 16; i++) tH(T, i, 31 - i, 0, r);
  }
  if constexpr (n == 6) {
#pragma unroll
    for (int i = 0; i < 8; i++) tB(T, 55 - i, 40 + i, 32, 1);
#pragma unroll
    for (int i = 0; i < 32; i++) tH(T, i, 63 - i, 0, r);
  }
}

template <int n>
DEVI void t_adst_in(int32_t *T) {
  constexpr int N0 = 1 << n;
  int32_t c[N0];
#pragma unroll
  for (int i = 0; i < N0; i++) c[i] = T[i];
#pragma unroll
  for (int i = 0; i < N0; i++) T[i] = c[(i & 1) ? i - 1 : N0 - i - 1];
}
__host__ __device__ constexpr int adst_out_idx(int i, int n) {
  return ((((i & 1) ^ ((i >> 1) & 1)) << 3) | ((((i >> 1) & 1) ^ ((i >> 2) & 1)) << 2) |
          ((((i >> 2) & 1) ^ ((i >> 3) & 1)) << 1) | ((i >> 3) & 1)) >> (4 - n);
}
template <int n>
DEVI void t_adst_out(int32_t *T) {
  constexpr int N0 = 1 << n;
  int32_t c[N0];
#pragma unroll
  for (int i = 0; i < N0; i++) c[i] = T[i];
#pragma unroll
  for (int i = 0; i < N0; i++) T[i] = (i & 1) ? -c[adst_out_idx(i, n)] : c[adst_out_idx(i, n)];
}

/* 7.13.2.6 in 32-bit arithmetic: inputs within 16 bits, so the largest sum
 * (the fourth output's) is below 10950 * 2^15 < 2^29. */
DEVI void t_adst4(int32_t *T) {
  int32_t s0 = __mul24(1321, T[0]), s1 = __mul24(2482, T[0]), s2 = __mul24(3344, T[1]), s3 = __mul24(3803, T[2]),
          s4 = __mul24(1321, T[2]), s5 = __mul24(2482, T[3]), s6 = __mul24(3803, T[3]), b7 = T[0] - T[2] + T[3];
  s0 += s3;
  s1 -= s4;
  s3 = s2;
  s2 = __mul24(3344, b7);
  s0 += s5;
  s1 -= s6;
  T[0] = (s0 + s3 + 2048) >> 12;
  T[1] = (s1 + s3 + 2048) >> 12;
  T[2] = (s2 + 2048) >> 12;
  T[3] = (s0 + s1 - s3 + 2048) >> 12;
}
DEVI void t_adst8(int32_t *T, int r) {
  t_adst_in<3>(T);
#pragma unroll
  for (int i = 0; i < 4; i++) tB(T, 2 * i, 2 * i + 1, 60 - 16 * i, 1);
#pragma unroll
  for (int i = 0; i < 4; i++) tH(T, i, 4 + i, 0, r);
#pragma unroll
  for (int i = 0; i < 2; i++) tB(T, 4 + 3 * i, 5 + i, 48 - 32 * i, 1);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) tH(T, 4 * j + i, 2 + 4 * j + i, 0, r);
#pragma unroll
  for (int i = 0; i < 2; i++) tB(T, 2 + 4 * i, 3 + 4 * i, 32, 1);
  t_adst_out<3>(T);
}
DEVI void t_adst16(int32_t *T, int r) {
  t_adst_in<4>(T);
#pragma unroll
  for (int i = 0; i < 8; i++) tB(T, 2 * i, 2 * i + 1, 62 - 8 * i, 1);
#pragma unroll
  for (int i = 0; i < 8; i++) tH(T, i, 8 + i, 0, r);
#pragma unroll
  for (int i = 0; i < 2; i++) {
    tB(T, 8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1);
    tB(T, 13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1);
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) tH(T, 8 * j + i, 4 + 8 * j + i, 0, r);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) tB(T, 4 + 8 * j + 3 * i, 5 + 8 * j + i, 48 - 32 * i, 1);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) tH(T, 4 * j + i, 2 + 4 * j + i, 0, r);
#pragma unroll
  for (int i = 0; i < 4; i++) tB(T, 2 + 4 * i, 3 + 4 * i, 32, 1);
  t_adst_out<4>(T);
}
template <int n>
DEVI void t_identity(int32_t *T) {
#pragma unroll
  for (int i = 0; i < (1 << n); i++) {
    if constexpr (n == 2) T[i] = (__mul24(T[i], 5793) + 2048) >> 12; /* 16-bit inputs: below 2^29 */
    else if constexpr (n == 3) T[i] = T[i] * 2;
    else if constexpr (n == 4) T[i] = (__mul24(T[i], 11586) + 2048) >> 12;
    else T[i] = T[i] * 4;
  }
}

/* 1-D inverse transform of kind 0 DCT, 1 ADST, 2 identity, length 1 << n. */
DEVI void t_itx1d(int32_t *T, int kind, int n, int r) {
  if (kind == 0) {
    switch (n) {
      case 2: t_idct<2>(T, r); break;
      case 3: t_idct<3>(T, r); break;
      case 4: t_idct<4>(T, r); break;
      case 5: t_idct<5>(T, r); break;
      default: t_idct<6>(T, r); break;
    }
  } else if (kind == 1) {
    if (n == 2) t_adst4(T);
    else if (n == 3) t_adst8(T, r);
    else t_adst16(T, r);
  } else {
    switch (n) {
      case 2: t_identity<2>(T); break;
      case 3: t_identity<3>(T); break;
      case 4: t_identity<4>(T); break;
      case 5: t_identity<5>(T); break;
      default: t_identity<6>(T); break;
    }
  }
}

DEVI int row_kind(int t) {
  if (t == DCT_DCT || t == ADST_DCT || t == FLIPADST_DCT || t == H_DCT) return 0;
  if (t == DCT_ADST || t == ADST_ADST || t == DCT_FLIPADST || t == FLIPADST_FLIPADST || t == ADST_FLIPADST ||
      t == FLIPADST_ADST || t == H_ADST || t == H_FLIPADST)
    return 1;
  return 2;
}
DEVI int col_kind(int t) {
  if (t == DCT_DCT || t == DCT_ADST || t == DCT_FLIPADST || t == V_DCT) return 0;
  if (t == ADST_DCT || t == ADST_ADST || t == FLIPADST_DCT || t == FLIPADST_FLIPADST || t == ADST_FLIPADST ||
      t == FLIPADST_ADST || t == V_ADST || t == V_FLIPADST)
    return 1;
  return 2;
}
DEVI int flip_ud(int t) { return t == FLIPADST_DCT || t == FLIPADST_ADST || t == V_FLIPADST || t == FLIPADST_FLIPADST; }
DEVI int flip_lr(int t) { return t == DCT_FLIPADST || t == ADST_FLIPADST || t == H_FLIPADST || t == FLIPADST_FLIPADST; }
DEVI int dq_denom(int tx) {
  if (tx == TX_32X32 || tx == TX_16X32 || tx == TX_32X16 || tx == TX_16X64 || tx == TX_64X16) return 2;
  if (tx == TX_64X64 || tx == TX_32X64 || tx == TX_64X32) return 4;
  return 1;
}

/* The transform working arrays live in LDS; the stage functions receive
 * them as generic pointers and address them through 32-bit LDS pointers. */
typedef __attribute__((address_space(3))) int32_t lds_i32;
typedef __attribute__((address_space(3))) int16_t lds_i16;
DEVI lds_i32 *lds(int32_t *p) { return (lds_i32 *)p; }
DEVI const lds_i32 *lds(const int32_t *p) { return (const lds_i32 *)p; }

/* The threads that code one transform block: a whole workgroup (wg) or a
 * single wave. A wave's lanes run in lockstep, so ordering its LDS accesses
 * needs only a compiler fence; a workgroup needs a barrier. red: LDS space
 * for one partial per wave of a workgroup reduction. */
struct TxCtx {
  int tid, nt;
  bool wg;
  long long *red;
};

/* Wave reductions through DPP lane swaps instead of LDS permutes, for a
 * wave whose 64 lanes are all active; the result is a scalar (every lane
 * gets it). Within each row of 16 lanes: swaps across 1 and 2 (quad
 * permutes), the half-row mirror and the row mirror leave the row's total in
 * every lane of the row; the four row totals are then read out. */
#define DPP_QX1 0xB1   /* quad_perm [1, 0, 3, 2] */
#define DPP_QX2 0x4E   /* quad_perm [2, 3, 0, 1] */
#define DPP_HMIRROR 0x141
#define DPP_MIRROR 0x140
template <int CTRL>
DEVI int dpp_t(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
DEVI int wave_max_dpp(int v) {
  v = d_max(v, dpp_t<DPP_QX1>(v));
  v = d_max(v, dpp_t<DPP_QX2>(v));
  v = d_max(v, dpp_t<DPP_HMIRROR>(v));
  v = d_max(v, dpp_t<DPP_MIRROR>(v));
  return d_max(d_max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               d_max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
DEVI int wave_sum_dpp(int v) {
  v += dpp_t<DPP_QX1>(v);
  v += dpp_t<DPP_QX2>(v);
  v += dpp_t<DPP_HMIRROR>(v);
  v += dpp_t<DPP_MIRROR>(v);
  return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
         __builtin_amdgcn_readlane(v, 48);
}
template <int CTRL>
DEVI long long dpp_t64(long long v) {
  const int lo = dpp_t<CTRL>((int)(unsigned)(unsigned long long)v), hi = dpp_t<CTRL>((int)(v >> 32));
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
DEVI long long wave_sum64_dpp(long long v) {
  v += dpp_t64<DPP_QX1>(v);
  v += dpp_t64<DPP_QX2>(v);
  v += dpp_t64<DPP_HMIRROR>(v);
  v += dpp_t64<DPP_MIRROR>(v);
  long long t = 0;
#pragma unroll
  for (int r = 0; r < 64; r += 16) {
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)(unsigned long long)v, r);
    const unsigned hi = __builtin_amdgcn_readlane((int)(v >> 32), r);
    t += (long long)(((unsigned long long)hi << 32) | lo);
  }
  return t;
}

DEVI void tx_sync(const TxCtx &c) {
  if (c.wg) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}
DEVI long long tx_sum64(const TxCtx &c, long long v) {
  v = wave_sum64_dpp(v);
  if (!c.wg) return v;
  if ((c.tid & 63) == 0) c.red[c.tid >> 6] = v;
  __syncthreads();
  long long s = 0;
  for (int i = 0; i < (c.nt >> 6); i++) s += c.red[i];
  __syncthreads();
  return s;
}
DEVI int tx_max(const TxCtx &c, int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = d_max(v, __shfl_xor(v, o, 64));
  if (!c.wg) return v;
  if ((c.tid & 63) == 0) c.red[c.tid >> 6] = v;
  __syncthreads();
  int m = 0;
  for (int i = 0; i < (c.nt >> 6); i++) m = d_max(m, (int)c.red[i]);
  __syncthreads();
  return m;
}

/* 2-D inverse transform (7.13.3): dq holds Min(32,h) x Min(32,w) values with
 * row stride qs; res receives h x w residuals with row stride rs. One
 * thread per row, then one per column; each (kind, length) pass is its own
 * non-inlined function so the unrolled butterflies exist once in the code
 * object. */
template <int KIND, int N>
DEVI void t_1d(int32_t *T, int r) {
  if constexpr (KIND == 0) t_idct<N>(T, r);
  else if constexpr (KIND == 1) {
    if constexpr (N == 2) t_adst4(T);
    else if constexpr (N == 3) t_adst8(T, r);
    else t_adst16(T, r);
  } else
    t_identity<N>(T);
}

template <int KIND, int N>
DEVI void itx_rows(const int32_t *dq_, int qs, int32_t *res_, int rs, int h, int rect,
                                                    int rowShift, int tid, int nt) {
  constexpr int W = 1 << N;
  const lds_i32 *dq = lds(dq_);
  lds_i32 *res = lds(res_);
  for (int i = tid; i < h; i += nt) {
    int32_t T[W];
    int nz = 0;
#pragma unroll
    for (int j = 0; j < W; j++) {
      int32_t v = (i < 32 && j < 32) ? dq[i * qs + j] : 0;
      nz |= v;
      T[j] = rect ? (__mul24(v, 2896) + 2048) >> 12 : v;
    }
    if (nz) { /* every 1-D kernel maps an all-zero row to zeros */
      t_1d<KIND, N>(T, 16);
#pragma unroll
      for (int j = 0; j < W; j++) res[i * rs + j] = d_clip(-(1 << 15), (1 << 15) - 1, d_round2(T[j], rowShift));
    } else {
#pragma unroll
      for (int j = 0; j < W; j++) res[i * rs + j] = 0;
    }
  }
}
template <int KIND, int N>
DEVI void itx_cols(int32_t *res_, int rs, int w, int tid, int nt) {
  constexpr int H = 1 << N;
  lds_i32 *res = lds(res_);
  for (int j = tid; j < w; j += nt) {
    int32_t T[H];
#pragma unroll
    for (int i = 0; i < H; i++) T[i] = res[i * rs + j];
    t_1d<KIND, N>(T, 16);
#pragma unroll
    for (int i = 0; i < H; i++) res[i * rs + j] = d_round2(T[i], 4);
  }
}

#define ITX_DISPATCH(FN, kind, n, ...)                                      \
  do {                                                                      \
    if (kind == 0) {                                                        \
      switch (n) {                                                          \
        case 2: FN<0, 2>(__VA_ARGS__); break;                               \
        case 3: FN<0, 3>(__VA_ARGS__); break;                               \
        case 4: FN<0, 4>(__VA_ARGS__); break;                               \
        case 5: FN<0, 5>(__VA_ARGS__); break;                               \
        default: FN<0, 6>(__VA_ARGS__); break;                              \
      }                                                                     \
    } else if (kind == 1) {                                                 \
      switch (n) {                                                          \
        case 2: FN<1, 2>(__VA_ARGS__); break;                               \
        case 3: FN<1, 3>(__VA_ARGS__); break;                               \
        default: FN<1, 4>(__VA_ARGS__); break;                              \
      }                                                                     \
    } else {                                                                \
      switch (n) {                                                          \
        case 2: FN<2, 2>(__VA_ARGS__); break;                               \
        case 3: FN<2, 3>(__VA_ARGS__); break;                               \
        case 4: FN<2, 4>(__VA_ARGS__); break;                               \
        default: FN<2, 5>(__VA_ARGS__); break;                              \
      }                                                                     \
    }                                                                       \
  } while (0)

/* The dispatch over kind and length of the transforms up to 16 points. */
#define ITX_DISPATCH_SMALL(FN, kind, n, ...)                                \
  do {                                                                      \
    if (n == 2) {                                                           \
      if (kind == 0) FN<0, 2>(__VA_ARGS__);                                 \
      else if (kind == 1) FN<1, 2>(__VA_ARGS__);                            \
      else FN<2, 2>(__VA_ARGS__);                                           \
    } else if (n == 3) {                                                    \
      if (kind == 0) FN<0, 3>(__VA_ARGS__);                                 \
      else if (kind == 1) FN<1, 3>(__VA_ARGS__);                            \
      else FN<2, 3>(__VA_ARGS__);                                           \
    } else {                                                                \
      if (kind == 0) FN<0, 4>(__VA_ARGS__);                                 \
      else if (kind == 1) FN<1, 4>(__VA_ARGS__);                            \
      else FN<2, 4>(__VA_ARGS__);                                           \
    }                                                                       \
  } while (0)

/* SMALL: the caller's transforms are at most 16x16, so only those lengths
 * are compiled in (smaller code where the small-block path runs). */
template <bool SMALL = false>
DEVI void tx_inverse(const TxCtx &c, const int32_t *dq, int qs, int32_t *res, int rs, int txSz, int txType) {
  txSz = d_uni(txSz);
  txType = d_uni(txType);
  const int lw = d_uni(Tx_Width_Log2[txSz]), lh = d_uni(Tx_Height_Log2[txSz]);
  const int rect = (lw - lh == 1) || (lh - lw == 1);
  const int rk = row_kind(txType), ck = col_kind(txType);
  if constexpr (SMALL) {
    ITX_DISPATCH_SMALL(itx_rows, rk, lw, dq, qs, res, rs, 1 << lh, rect, d_uni(Transform_Row_Shift[txSz]), c.tid, c.nt);
    tx_sync(c);
    ITX_DISPATCH_SMALL(itx_cols, ck, lh, res, rs, 1 << lw, c.tid, c.nt);
  } else {
    ITX_DISPATCH(itx_rows, rk, lw, dq, qs, res, rs, 1 << lh, rect, d_uni(Transform_Row_Shift[txSz]), c.tid, c.nt);
    tx_sync(c);
    ITX_DISPATCH(itx_cols, ck, lh, res, rs, 1 << lw, c.tid, c.nt);
  }
  tx_sync(c);
}

/* Forward 2-D transform through the Q20 matrices (k_fwd.h:35-59): resid
 * h x w (row stride rs) -> coef Min(32,h) x Min(32,w) (row stride tw); tmp
 * holds h x Min(32,w) (row stride tw). One thread per output in both
 * passes; the sums of a fixed length are unrolled by up to 16 so the matrix
 * loads issue back to back. smats (may be NULL): the packed short matrices
 * in LDS, used for lengths up to 16.
 * Widths: residuals are within +-255 and no matrix row has an L1 norm above
 * 1.49e6 (entries below 2^20), so a row sum stays below 2^28.5 and takes
 * 24-bit multiplies into 32 bits, exactly; a row output is then below 5800
 * in magnitude and the column sums (up to 2^33) take 64-bit multiply-adds. */
template <int W, class MP>
DEVI void fwd_rows(const TxCtx &c, const int32_t *resid_, int rs, int32_t *tmp_, int h, int ltw, MP Mh) {
  const int tw = 1 << ltw;
  const lds_i32 *resid = lds(resid_);
  lds_i32 *tmp = lds(tmp_);
  for (int k = c.tid; k < h * tw; k += c.nt) {
    const int i = k >> ltw, kk = k & (tw - 1);
    const lds_i32 *rr = resid + i * rs;
    const MP m = Mh + kk * W;
    int s = 0;
#pragma unroll 16
    for (int j = 0; j < W; j++) s += __mul24(rr[j], m[j]);
    tmp[k] = (s + 32768) >> 16;
  }
}
template <int H, class MP>
DEVI void fwd_cols(const TxCtx &c, const int32_t *tmp_, int32_t *coef_, int th, int ltw, MP Mv, int sh, bool rect) {
  const int tw = 1 << ltw;
  const lds_i32 *tmp = lds(tmp_);
  lds_i32 *coef = lds(coef_);
  for (int k = c.tid; k < th * tw; k += c.nt) {
    const int k2 = k >> ltw, kk = k & (tw - 1);
    const MP m = Mv + k2 * H;
    long long s = 0;
#pragma unroll 16
    for (int i = 0; i < H; i++) s += (long long)m[i] * tmp[i * tw + kk];
    s *= (1ll << sh);
    if (rect) s = ((s + 2048) >> 12) * 5793;
    coef[k] = (int32_t)(s >= 0 ? (s + (1ll << 19)) >> 20 : -((-s + (1ll << 19)) >> 20));
  }
}

/* The 4-, 8- and 16-point matrices of the three kinds packed back to back
 * (TX_SMATS ints): kind k, length 1 << n at k * 336 + {0, 16, 80}[n - 2]. */
#define TX_SMATS (3 * 336)
DEVI int smat_off(int kind, int n) { return kind * 336 + (n == 2 ? 0 : n == 3 ? 16 : 80); }

template <class MP, bool SMALL = false>
DEVI void tx_forward_m(const TxCtx &c, const int32_t *resid, int rs, int32_t *tmp, int32_t *coef, int txSz, MP Mh,
                       MP Mv) {
  const int lw = d_uni(Tx_Width_Log2[txSz]), lh = d_uni(Tx_Height_Log2[txSz]), h = 1 << lh;
  const int ltw = d_min(5, lw), th = d_min(32, h);
  if (lw == 2) fwd_rows<4>(c, resid, rs, tmp, h, ltw, Mh);
  else if (lw == 3) fwd_rows<8>(c, resid, rs, tmp, h, ltw, Mh);
  else if (SMALL || lw == 4) fwd_rows<16>(c, resid, rs, tmp, h, ltw, Mh);
  else if constexpr (!SMALL) {
    if (lw == 5) fwd_rows<32>(c, resid, rs, tmp, h, ltw, Mh);
    else fwd_rows<64>(c, resid, rs, tmp, h, ltw, Mh);
  }
  tx_sync(c);
  const int sh = d_uni(Transform_Row_Shift[txSz]) + 4;
  const bool rect = (lw - lh == 1) || (lh - lw == 1);
  if (lh == 2) fwd_cols<4>(c, tmp, coef, th, ltw, Mv, sh, rect);
  else if (lh == 3) fwd_cols<8>(c, tmp, coef, th, ltw, Mv, sh, rect);
  else if (SMALL || lh == 4) fwd_cols<16>(c, tmp, coef, th, ltw, Mv, sh, rect);
  else if constexpr (!SMALL) {
    if (lh == 5) fwd_cols<32>(c, tmp, coef, th, ltw, Mv, sh, rect);
    else fwd_cols<64>(c, tmp, coef, th, ltw, Mv, sh, rect);
  }
  tx_sync(c);
}

/* The quantiser's reciprocal for tx_quantize: 1 / q rounded to float. */
DEVI float tx_qrcp(int q) { return 1.0f / (float)q; }

/* floor(t / q) for 0 <= t < 2^19 and 4 <= q < 2^16 (which the reference
 * computes as t * (2^44 / q + 1) >> 44): t * rq is within 2^-5 of t / q
 * (t and q exact in float, rq and the product one rounding each), so its
 * truncation is off by at most one, which the remainder corrects. */
DEVI int tx_qdiv(int t, int q, float rq) {
  int l = (int)((float)t * rq);
  const int r = t - __mul24(l, q);
  return l + (r >= q) - (r < 0);
}

/* One coefficient v through the dead-zone quantiser (k_fwd.h:66-90): the
 * signed level. As in the reference, t below 2^19 (every coefficient of an
 * 8-bit residual) takes the reciprocal, larger t an integer division. */
DEVI int tx_quant1(int v, bool dc, int dcQ, int acQ, float rDc, float rAc, int denom) {
  const int qv = dc ? dcQ : acQ, rnd = dc ? 26 : 22, av = d_abs(v);
  int level;
  if (av < (1 << 21)) {
    const int t = (av * denom + ((qv * rnd) >> 2)) >> 4;
    level = t < (1 << 19) ? tx_qdiv(t, qv, dc ? rDc : rAc) : t / qv;
  } else {
    const long long t = ((long long)av * denom + ((qv * rnd) >> 2)) >> 4;
    level = t / qv > 0xFFFF ? 0xFFFF : (int)(t / qv);
  }
  if (level > 0xFFFF) level = 0xFFFF;
  return v < 0 ? -level : level;
}

/* One wave's forward transform of a (1 << LH) x (1 << LW) block, both
 * lengths at most 16, with the sizes known at compile time: every LDS
 * access is a fixed offset from a per-lane base, so the loads of a sum issue
 * back to back. The row pass writes tmp transposed (column kk at kk * H) so
 * that the column pass also reads consecutive words. Same sums and
 * roundings as fwd_rows / fwd_cols. */
template <int LW, int LH, bool QUANT = false>
DEVI int sfwd(const lds_i32 *resid, int rs, lds_i32 *tmp, lds_i32 *coef, const lds_i32 *Mh, const lds_i32 *Mv, int sh,
              int lane_, int dcQ = 0, int acQ = 0, float rDc = 0.f, float rAc = 0.f, const lds_i16 *iscan = nullptr) {
  constexpr int W = 1 << LW, H = 1 << LH, NO = W * H, IT = (NO + 63) / 64;
  const int lane = d_local(lane_);
  constexpr bool RECT = LW - LH == 1 || LH - LW == 1;
#pragma unroll 1
  for (int t = 0; t < IT; t++) {
    const int k = lane + 64 * t;
    if (NO % 64 == 0 || k < NO) {
      const int i = k >> LW, kk = k & (W - 1);
      const lds_i32 *rr = resid + i * rs, *m = Mh + kk * W;
      int s = 0;
#pragma unroll
      for (int j = 0; j < W; j++) s += __mul24(rr[j], m[j]);
      tmp[kk * H + i] = (s + 32768) >> 16;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int eob = 0;
#pragma unroll 1
  for (int t = 0; t < IT; t++) {
    const int k = lane + 64 * t;
    if (NO % 64 == 0 || k < NO) {
      const int k2 = k >> LW, kk = k & (W - 1);
      const lds_i32 *col = tmp + kk * H, *m = Mv + k2 * H;
      long long s = 0;
#pragma unroll
      for (int i = 0; i < H; i++) s += (long long)m[i] * col[i];
      s *= (1ll << sh);
      if constexpr (RECT) s = ((s + 2048) >> 12) * 5793;
      const int v = (int32_t)(s >= 0 ? (s + (1ll << 19)) >> 20 : -((-s + (1ll << 19)) >> 20));
      /* QUANT: the level at once (sizes up to 16 have no dequantiser
       * denominator), and the eob: 1 + the largest scan position (iscan)
       * of a non-zero level */
      if constexpr (QUANT) {
        const int lv = tx_quant1(v, k == 0, dcQ, acQ, rDc, rAc, 1);
        coef[k] = lv;
        if (lv) eob = d_max(eob, iscan[k] + 1);
      } else {
        coef[k] = v;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return QUANT ? wave_max_dpp(eob) : 0;
}

/* One wave's forward transform and quantiser of a block of at most 16x16:
 * levels into q (stride 1 << lw); returns the eob for the inverse scan
 * iscan (the scan position of each coefficient). The whole wave calls it. */
DEVI int tx_forward_quant_small(const TxCtx &c, const int32_t *resid, int rs, int32_t *tmp, int32_t *q, int txSz,
                                int txType, const int32_t *smats, int dcQ, int acQ, float rDc, float rAc,
                                const int16_t *iscan) {
  txSz = d_uni(txSz);
  txType = d_uni(txType);
  const int lw = d_uni(Tx_Width_Log2[txSz]), lh = d_uni(Tx_Height_Log2[txSz]);
  const lds_i32 *Mh = lds(smats) + smat_off(row_kind(txType), lw), *Mv = lds(smats) + smat_off(col_kind(txType), lh);
  const int sh = d_uni(Transform_Row_Shift[txSz]) + 4;
  const lds_i32 *r = lds(resid);
  lds_i32 *tp = lds(tmp), *qo = lds(q);
  const lds_i16 *is = (const lds_i16 *)iscan;
#define SFWDQ(a, b) return sfwd<a, b, true>(r, rs, tp, qo, Mh, Mv, sh, c.tid, dcQ, acQ, rDc, rAc, is)
  switch (lw * 4 + lh) {
    case 10: SFWDQ(2, 2);
    case 11: SFWDQ(2, 3);
    case 12: SFWDQ(2, 4);
    case 14: SFWDQ(3, 2);
    case 15: SFWDQ(3, 3);
    case 16: SFWDQ(3, 4);
    case 18: SFWDQ(4, 2);
    case 19: SFWDQ(4, 3);
    default: SFWDQ(4, 4);
  }
#undef SFWDQ
}

/* smats (may be NULL): the packed short matrices in LDS, used when both
 * lengths are at most 16; otherwise the full matrices in global memory.
 * SMALL: both lengths are at most 16, smats is given and c is one wave. */
template <bool SMALL = false>
DEVI void tx_forward(const TxCtx &c, const int32_t *resid, int rs, int32_t *tmp, int32_t *coef, int txSz, int txType,
                     const int32_t *__restrict__ mats, const int32_t *smats) {
  txSz = d_uni(txSz);
  txType = d_uni(txType);
  const int lw = d_uni(Tx_Width_Log2[txSz]), lh = d_uni(Tx_Height_Log2[txSz]);
  const int rk = row_kind(txType), ck = col_kind(txType);
  if constexpr (SMALL) {
    /* one wave (c.nt == 64) */
    const lds_i32 *Mh = lds(smats) + smat_off(rk, lw), *Mv = lds(smats) + smat_off(ck, lh);
    const int sh = d_uni(Transform_Row_Shift[txSz]) + 4;
    const lds_i32 *r = lds(resid);
    lds_i32 *tp = lds(tmp), *co = lds(coef);
    switch (lw * 4 + lh) {
      case 10: sfwd<2, 2>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      case 11: sfwd<2, 3>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      case 12: sfwd<2, 4>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      case 14: sfwd<3, 2>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      case 15: sfwd<3, 3>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      case 16: sfwd<3, 4>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      case 18: sfwd<4, 2>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      case 19: sfwd<4, 3>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
      default: sfwd<4, 4>(r, rs, tp, co, Mh, Mv, sh, c.tid); break;
    }
  } else {
    if (smats && lw <= 4 && lh <= 4)
      tx_forward_m(c, resid, rs, tmp, coef, txSz, lds(smats) + smat_off(rk, lw), lds(smats) + smat_off(ck, lh));
    else
      tx_forward_m(c, resid, rs, tmp, coef, txSz, mats + (rk * 5 + lw - 2) * 4096, mats + (ck * 5 + lh - 2) * 4096);
  }
}

/* The quantiser's reciprocal of q: floor(2^44 / q) + 1. */
/* Dead-zone quantiser (k_fwd.h:66-90) into the decoder layout (stride tw);
 * rDc / rAc: tx_qrcp of dcQ / acQ. */
DEVI void tx_quantize(const TxCtx &c, const int32_t *coef_, int32_t *q_, int txSz, int dcQ, int acQ, float rDc,
                      float rAc) {
  const lds_i32 *coef = lds(coef_);
  lds_i32 *q = lds(q_);
  txSz = d_uni(txSz);
  const int tw = d_min(32, d_uni(Tx_Width[txSz])), th = d_min(32, d_uni(Tx_Height[txSz])), denom = dq_denom(txSz);
  for (int k = c.tid; k < th * tw; k += c.nt) q[k] = tx_quant1(coef[k], k == 0, dcQ, acQ, rDc, rAc, denom);
  tx_sync(c);
}

/* Dequantisation (spec 7.12.3), output row stride qs. */
DEVI void tx_dequant(const TxCtx &c, const int32_t *q_, int32_t *dq_, int qs, int txSz, int dcQ, int acQ) {
  const lds_i32 *q = lds(q_);
  lds_i32 *dq = lds(dq_);
  txSz = d_uni(txSz);
  const int tw = d_min(32, d_uni(Tx_Width[txSz])), th = d_min(32, d_uni(Tx_Height[txSz])), denom = dq_denom(txSz);
  const int ltw = d_min(5, d_uni(Tx_Width_Log2[txSz]));
  for (int k = c.tid; k < th * tw; k += c.nt) {
    const int i = k >> ltw, j = k & (tw - 1);
    /* levels are at most 0xFFFF and quantisers at most 1828: the product
     * stays below 2^27 */
    const int d = __mul24(q[k], k == 0 ? dcQ : acQ), ad = d < 0 ? -d : d;
    const int d2 = (ad & 0xFFFFFF) >> (denom >> 1);
    dq[i * qs + j] = d_clip(-(1 << 15), (1 << 15) - 1, d < 0 ? -d2 : d2);
  }
  tx_sync(c);
}

/* One wave's dequantisation, inverse transform and reconstruction of a
 * block of at most 16x16 (tx_dequant, tx_inverse and the addition into the
 * prediction, fused): the row pass dequantises the levels q (stride
 * 1 << lw) as it loads them, the column pass adds its outputs to px in the
 * type's flipped order, clipped. res: scratch for the row outputs (row
 * stride rs). Same values as the three passes. */
template <int KIND, int N>
DEVI void sitx_rows_dq(const lds_i32 *q, lds_i32 *res, int rs, int h, int rect, int rowShift, int dcQ, int acQ,
                       int tid) {
  constexpr int W = 1 << N;
  for (int i = tid; i < h; i += 64) {
    int32_t T[W];
    int nz = 0;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const int k = i * W + j, d = __mul24(q[k], k == 0 ? dcQ : acQ), ad = (d < 0 ? -d : d) & 0xFFFFFF;
      const int v = d_clip(-(1 << 15), (1 << 15) - 1, d < 0 ? -ad : ad);
      nz |= v;
      T[j] = rect ? (__mul24(v, 2896) + 2048) >> 12 : v;
    }
    if (nz) {
      t_1d<KIND, N>(T, 16);
#pragma unroll
      for (int j = 0; j < W; j++) res[i * rs + j] = d_clip(-(1 << 15), (1 << 15) - 1, d_round2(T[j], rowShift));
    } else {
#pragma unroll
      for (int j = 0; j < W; j++) res[i * rs + j] = 0;
    }
  }
}
template <int KIND, int N>
DEVI void sitx_cols_add(const lds_i32 *res, int rs, int w, uint8_t *px, int S, int x, int y, int fud, int flr,
                        int tid) {
  constexpr int H = 1 << N;
  for (int j = tid; j < w; j += 64) {
    int32_t T[H];
#pragma unroll
    for (int i = 0; i < H; i++) T[i] = res[i * rs + j];
    t_1d<KIND, N>(T, 16);
    uint8_t *col = px + y * S + x + (flr ? w - 1 - j : j);
#pragma unroll
    for (int i = 0; i < H; i++) {
      uint8_t *o = col + (fud ? H - 1 - i : i) * S;
      *o = (uint8_t)d_clip1((int)*o + d_round2(T[i], 4));
    }
  }
}
DEVI void tx_recon_small(const TxCtx &c, const int32_t *q, int32_t *res, int rs, int txSz, int txType, int dcQ,
                         int acQ, uint8_t *px, int S, int x, int y) {
  txSz = d_uni(txSz);
  txType = d_uni(txType);
  const int lw = d_uni(Tx_Width_Log2[txSz]), lh = d_uni(Tx_Height_Log2[txSz]);
  const int rect = (lw - lh == 1) || (lh - lw == 1);
  const int rk = row_kind(txType), ck = col_kind(txType);
  const lds_i32 *ql = lds(q);
  lds_i32 *rl = lds(res);
  ITX_DISPATCH_SMALL(sitx_rows_dq, rk, lw, ql, rl, rs, 1 << lh, rect, d_uni(Transform_Row_Shift[txSz]), dcQ, acQ,
                     c.tid);
  tx_sync(c);
  ITX_DISPATCH_SMALL(sitx_cols_add, ck, lh, rl, rs, 1 << lw, px, S, x, y, flip_ud(txType), flip_lr(txType), c.tid);
  tx_sync(c);
}

#endif
