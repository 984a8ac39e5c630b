/* ORACLE — test infrastructure only. Small integer helpers shared by the
 * oracle's translation units (the arithmetic conventions of
 * src/kernels/k_common.h:39-48: Round2 rounds half up with an arithmetic
 * shift, FloorLog2(x) for x >= 1). */
#ifndef MAV1_ORACLE_COMMON_H
#define MAV1_ORACLE_COMMON_H

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../mav1-encoder_amd/csrc/av1_defs.h"
#include "../mav1-encoder_amd/csrc/av1_tables.h"
#include "oracle.h"

static inline int o_min(int a, int b) { return a < b ? a : b; }
static inline int o_max(int a, int b) { return a > b ? a : b; }
static inline int o_abs(int a) { return a < 0 ? -a : a; }
static inline int o_clip(int lo, int hi, int v) { return v < lo ? lo : v > hi ? hi : v; }
static inline int o_clip1(int v) { return o_clip(0, 255, v); }
static inline int o_round2(int x, int n) { return n ? (x + (1 << (n - 1))) >> n : x; }
static inline int64_t o_round2l(int64_t x, int n) { return n ? (x + ((int64_t)1 << (n - 1))) >> n : x; }
static inline int o_round2signed(int x, int n) { return x >= 0 ? o_round2(x, n) : -o_round2(-x, n); }
static inline int o_floorlog2(unsigned x) {
  int s = -1;
  while (x) { x >>= 1; s++; }
  return s < 0 ? 0 : s;
}

/* Square-node numbering of a superblock: level l (64 >> l wide) starts at
 * node 0, 1, 5, 21 (src/kernels/k_analysis.h:31). */
static inline int o_node_base(int level) { return (((1 << (2 * level)) - 1) / 3); }

#endif
