/* Device-side helpers shared by the gfx950 kernels: integer rounding
 * conventions of the mav1 kernels (src/kernels/k_common.h:39-48), wave64
 * reductions, and the AV1 tables in constant memory. */
#ifndef MAV1_DEV_COMMON_H
#define MAV1_DEV_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mav1_types.h"
#include "av1_defs.h"

#define MAV1_TBL static __constant__ const
#include "av1_tables.h"

#define DEVI __device__ __forceinline__

DEVI int d_min(int a, int b) { return a < b ? a : b; }
/* v, which the caller knows to be equal in every active lane, as a scalar
 * (v_readfirstlane): branches on it become scalar branches and values
 * derived from it stay in scalar registers. */
DEVI int d_uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
/* The value unchanged, opaque to the optimiser: per-lane values derived
 * from it are computed where they are used instead of being hoisted out of
 * the enclosing loops and held (or spilled) across the whole kernel. */
DEVI int d_local(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
DEVI int d_max(int a, int b) { return a > b ? a : b; }
DEVI int d_abs(int a) { return a < 0 ? -a : a; }
DEVI int d_clip(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
DEVI int d_clip1(int v) { return d_clip(0, 255, v); }
DEVI int d_round2(int x, int n) { return n ? (x + (1 << (n - 1))) >> n : x; }
DEVI long long d_round2l(long long x, int n) { return n ? (x + (1ll << (n - 1))) >> n : x; }
DEVI int d_round2signed(int x, int n) { return x >= 0 ? d_round2(x, n) : -d_round2(-x, n); }
/* floor(log2(x)) for x >= 1 */
DEVI int d_floorlog2(unsigned x) { return 31 - __clz(x); }
DEVI int d_node_base(int level) { return ((1 << (2 * level)) - 1) / 3; }

/* XCD-aware work order: the GPU hands workgroup b of a launch to XCD b % 8
 * (round robin), and each XCD has its own L2. Workgroup b works on item
 * xcd_item(b, n, k): the items are dealt to the XCDs in runs of k
 * consecutive items (a superblock's nodes), so the items that share
 * reference rows meet in one L2 while the runs stay spread over the frame
 * (balanced). A bijection of [0, n); items past the last whole round of 8
 * runs keep their index. */
DEVI int xcd_item(int b, int n, int k) {
  const int full = n / (8 * k) * (8 * k);
  if (b >= full) return b;
  const int x = b & 7, i = b >> 3;
  return ((i / k) * 8 + x) * k + i % k;
}

/* wave64 reductions (every lane gets the result) */
DEVI int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
DEVI long long wave_sum64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
/* lexicographic minimum of (cost, index) packed as cost << 32 | index,
 * cost >= 0: the serial loop's first minimum */
DEVI unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
DEVI unsigned long long pack_key(int cost, int idx) {
  return ((unsigned long long)(unsigned)cost << 32) | (unsigned)idx;
}

/* Sum of |4x4 Hadamard transform| of a difference block (rows of 4). */
DEVI int satd4x4_rows(int a0, int a1, int a2, int a3, int b0, int b1, int b2, int b3, int c0, int c1, int c2, int c3,
                      int e0, int e1, int e2, int e3) {
  int r[4][4] = {{a0, a1, a2, a3}, {b0, b1, b2, b3}, {c0, c1, c2, c3}, {e0, e1, e2, e3}};
  int m[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int s01 = r[i][0] + r[i][1], d01 = r[i][0] - r[i][1], s23 = r[i][2] + r[i][3], d23 = r[i][2] - r[i][3];
    m[i][0] = s01 + s23;
    m[i][1] = d01 + d23;
    m[i][2] = s01 - s23;
    m[i][3] = d01 - d23;
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    int s01 = m[0][j] + m[1][j], d01 = m[0][j] - m[1][j], s23 = m[2][j] + m[3][j], d23 = m[2][j] - m[3][j];
    s += d_abs(s01 + s23) + d_abs(d01 + d23) + d_abs(s01 - s23) + d_abs(d01 - d23);
  }
  return s;
}
DEVI int satd4x4_arr(const int *d) {
  return satd4x4_rows(d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11], d[12], d[13], d[14],
                      d[15]);
}

#endif
