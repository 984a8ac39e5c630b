/* Motion search on gfx950: quarter-pel / decimated plane build and the
 * 4-level node search of the mav1 encoder (behaviour of
 * src/kernels/k_me.h:343-379 per node, src/backend_cpu.c:23-56 per frame).
 *
 * Plane build: one workgroup per 64 x 16 tile of a reference; the 23 x 71
 * source window (8-tap support, edge-clamped) is staged in LDS, the three
 * non-trivial horizontal phases are filtered once per row into LDS, and
 * each thread produces 4 adjacent pixels of all 16 phase planes.
 *
 * Node search: one wave64 per (node, reference). Full-pel candidates are
 * spread over the lanes (lane k evaluates candidates k, k + 64, ...), each
 * scored against the source block and the +-range search window staged in
 * LDS (edge clamping applied while staging, which reproduces the
 * per-coordinate clamping of the serial search); the winner is the
 * lexicographic minimum of (cost, candidate index) over the wave, i.e. the
 * serial loop's first minimum. Sub-pel steps spread the 4x4 units of one
 * candidate over the lanes and sum with a wave reduction; candidates are
 * compared in ring order with a strict <, as the serial refinement does.
 * Levels depend on their parent level and run as separate launches.
 */
#include "dev_common.h"
#include "internal.h"

#define TB(fp, bit) (((fp).turbo & (bit)) != 0)

long me_entry_bytes(const Mav1FrameParams &fp) {
  long W = fp.frame_w, H = fp.frame_h;
  return 16 * W * H + ((W + 3) >> 2) * ((H + 3) >> 2);
}
size_t me_dec_bytes(const Mav1FrameParams &fp) {
  return (size_t)((fp.frame_w + 3) >> 2) * ((fp.frame_h + 3) >> 2);
}

/* ------------------------------------------------------------------------ */
/* Quarter-pel planes                                                        */
/* ------------------------------------------------------------------------ */
#define QT_W 64
#define QT_H 16

DEVI void qpel_build_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ ref, uint8_t *__restrict__ out, const int blk_x, const int blk_y, const int blk_z) {
  __shared__ uint8_t tile[QT_H + 7][QT_W + 8];
  __shared__ int16_t hz[3][QT_H + 7][QT_W];
  const int W = fp.frame_w, H = fp.frame_h, S = fp.stride[0];
  const uint8_t *luma = ref + fp.offset[0];
  const int x0 = blk_x * QT_W, y0 = blk_y * QT_H, tid = threadIdx.x;
  for (int i = tid; i < (QT_H + 7) * (QT_W + 7); i += 256) {
    int r = i / (QT_W + 7), c = i - r * (QT_W + 7);
    tile[r][c] = luma[d_clip(0, H - 1, y0 + r - 3) * S + d_clip(0, W - 1, x0 + c - 3)];
  }
  int kf[4][8];
#pragma unroll
  for (int f = 0; f < 4; f++)
#pragma unroll
    for (int t = 0; t < 8; t++) kf[f][t] = SUBPEL_FILTER(EIGHTTAP, 4 * f)[t];
  __syncthreads();
  /* horizontal phases 1..3 (phase 0 is the sample itself: Round2(128 p, 3) = 16 p) */
  for (int i = tid; i < (QT_H + 7) * QT_W; i += 256) {
    int r = i / QT_W, c = i - r * QT_W;
#pragma unroll
    for (int fx = 1; fx < 4; fx++) {
      int acc = 0;
#pragma unroll
      for (int u = 0; u < 8; u++) acc += kf[fx][u] * tile[r][c + u];
      hz[fx - 1][r][c] = (int16_t)d_round2(acc, 3);
    }
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  const int y = y0 + ty, xb = x0 + tx * 4;
  if (y >= H || xb >= W) return;
  const size_t n = (size_t)W * H, o = (size_t)y * W + xb;
  const bool wide = ((W & 3) == 0) && xb + 4 <= W;
#pragma unroll
  for (int fx = 0; fx < 4; fx++) {
    uint32_t word[4] = {0, 0, 0, 0};
#pragma unroll
    for (int p = 0; p < 4; p++) {
      int c = tx * 4 + p, col[8];
#pragma unroll
      for (int t = 0; t < 8; t++) col[t] = fx ? (int)hz[fx - 1][ty + t][c] : 16 * (int)tile[ty + t][c + 3];
#pragma unroll
      for (int fy = 0; fy < 4; fy++) {
        int acc = 0;
#pragma unroll
        for (int t = 0; t < 8; t++) acc += kf[fy][t] * col[t];
        word[fy] |= (uint32_t)d_clip1(d_round2(acc, 11)) << (8 * p);
      }
    }
#pragma unroll
    for (int fy = 0; fy < 4; fy++) {
      uint8_t *dst = out + (size_t)(fy * 4 + fx) * n + o;
      if (wide) {
        *(uint32_t *)dst = word[fy];
      } else {
        for (int p = 0; p < 4 && xb + p < W; p++) dst[p] = (uint8_t)(word[fy] >> (8 * p));
      }
    }
  }
}
__global__ __launch_bounds__(256) void k_qpel_build(const Mav1FrameParams fp, const uint8_t *__restrict__ ref,
                                                    uint8_t *__restrict__ out) {
  qpel_build_body(fp, ref, out, blockIdx.x, blockIdx.y, blockIdx.z);
}

hipError_t launch_qpel_build(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *ref, uint8_t *planes16) {
  dim3 grid((fp.frame_w + QT_W - 1) / QT_W, (fp.frame_h + QT_H - 1) / QT_H);
  hipLaunchKernelGGL(k_qpel_build, grid, dim3(256), 0, st, fp, ref, planes16);
  return hipGetLastError();
}

/* 4x4-mean decimated plane (k_me.h:112-118 per sample). */
DEVI void dec_build_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ frame, uint8_t *__restrict__ dec, const int blk_x, const int blk_y, const int blk_z) {
  const int W = fp.frame_w, H = fp.frame_h, dw = (W + 3) >> 2, dh = (H + 3) >> 2;
  const int x = blk_x * 64 + (threadIdx.x & 63), y = blk_y * 4 + (threadIdx.x >> 6);
  if (x >= dw || y >= dh) return;
  const uint8_t *luma = frame + fp.offset[0];
  int s = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint8_t *row = luma + d_clip(0, H - 1, 4 * y + i) * fp.stride[0];
#pragma unroll
    for (int j = 0; j < 4; j++) s += row[d_clip(0, W - 1, 4 * x + j)];
  }
  dec[(size_t)y * dw + x] = (uint8_t)((s + 8) >> 4);
}
__global__ __launch_bounds__(256) void k_dec_build(const Mav1FrameParams fp, const uint8_t *__restrict__ frame,
                                                   uint8_t *__restrict__ dec) {
  dec_build_body(fp, frame, dec, blockIdx.x, blockIdx.y, blockIdx.z);
}

hipError_t launch_dec_build(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *frame, uint8_t *dec) {
  int dw = (fp.frame_w + 3) >> 2, dh = (fp.frame_h + 3) >> 2;
  hipLaunchKernelGGL(k_dec_build, dim3((dw + 63) / 64, (dh + 3) / 4), dim3(256), 0, st, fp, frame, dec);
  return hipGetLastError();
}

/* ------------------------------------------------------------------------ */
/* Node search                                                               */
/* ------------------------------------------------------------------------ */
#define WIN_R 4 /* largest full-pel radius served from the LDS window */

struct NodeState {
  int x, y, cy, cx, range, step, nstage, pr, pc, copy, valid;
};

DEVI void me_cand(const NodeState &m, int k, int &dy, int &dx) {
  if (k == 0) {
    dy = m.cy;
    dx = m.cx;
    return;
  }
  int side = 2 * (m.range / m.step) + 1;
  int q = (k - 1) / side;
  dy = m.cy - m.range + q * m.step;
  dx = m.cx - m.range + (k - 1 - q * side) * m.step;
}

DEVI int me_bits(int dy, int dx) {
  return 2 + 2 * d_floorlog2((unsigned)(d_abs(dy) * 8 + 1)) + 2 * d_floorlog2((unsigned)(d_abs(dx) * 8 + 1));
}

template <int BW>
DEVI int sad_lds(const uint8_t *srcL, const uint8_t *win, int WW, int oy, int ox, int rs) {
  int s = 0;
  for (int i = 0; i < BW; i += rs) {
    const uint8_t *a = srcL + i * BW, *b = win + (i + oy) * WW + ox;
#pragma unroll 8
    for (int j = 0; j < BW; j++) s += d_abs((int)a[j] - (int)b[j]);
  }
  return s * rs;
}

template <int BW>
DEVI int sad_global(const Mav1FrameParams &fp, const uint8_t *srcL, const uint8_t *refY, int x, int y, int dy, int dx,
                    int rs) {
  int s = 0, S = fp.stride[0];
  for (int i = 0; i < BW; i += rs) {
    const uint8_t *b = refY + d_clip(0, fp.frame_h - 1, y + i + dy) * S;
    for (int j = 0; j < BW; j++) s += d_abs((int)srcL[i * BW + j] - (int)b[d_clip(0, fp.frame_w - 1, x + j + dx)]);
  }
  return s * rs;
}

template <int BW>
DEVI int sad_dec(const Mav1FrameParams &fp, const uint8_t *ds, const uint8_t *dr, int x, int y, int dy, int dx) {
  const int W = (fp.frame_w + 3) >> 2, H = (fp.frame_h + 3) >> 2, n = BW >> 2;
  const int x0 = x >> 2, y0 = y >> 2, oy = dy >> 2, ox = dx >> 2;
  int s = 0;
  for (int i = 0; i < n; i++) {
    const uint8_t *a = ds + d_min(y0 + i, H - 1) * W, *b = dr + d_clip(0, H - 1, y0 + i + oy) * W;
    for (int j = 0; j < n; j++) s += d_abs((int)a[d_min(x0 + j, W - 1)] - (int)b[d_clip(0, W - 1, x0 + j + ox)]);
  }
  return s * 16;
}

/* Sub-pel cost at an even 1/8-pel vector from the quarter-pel planes: the
 * SATD of every st-th 4x4 unit (raster order), summed over the wave. */
template <int BW>
DEVI int subpel_cost_hp(const Mav1FrameParams &fp, const uint8_t *srcL, const uint8_t *planes, int x, int y, int mvr,
                        int mvc, int st, int lane) {
  constexpr int NR = BW / 4, NU = NR * NR;
  const int W = fp.frame_w, H = fp.frame_h;
  const uint8_t *pl = planes + (size_t)(((mvr >> 1) & 3) * 4 + ((mvc >> 1) & 3)) * W * H;
  int s = 0;
  for (int u = lane * st; u < NU; u += 64 * st) {
    int ucol = u % NR, urow = u / NR;
    int iy = y + urow * 4 + (mvr >> 3), ix = x + ucol * 4 + (mvc >> 3);
    int d[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint8_t *prow = pl + (size_t)d_clip(0, H - 1, iy + i) * W;
      const uint8_t *srow = srcL + (urow * 4 + i) * BW + ucol * 4;
#pragma unroll
      for (int j = 0; j < 4; j++) d[4 * i + j] = (int)srow[j] - (int)prow[d_clip(0, W - 1, ix + j)];
    }
    s += satd4x4_arr(d);
  }
  return wave_sum(s) * st;
}

/* Sub-pel costs of up to 9 candidates at once from the quarter-pel planes:
 * the lanes take (candidate, 4x4 unit) pairs, and each candidate's cost is
 * the same sum as subpel_cost_hp's (every st-th unit, times st), added up
 * in LDS. Integer sums, so the order does not matter. */
template <int BW>
DEVI void subpel_costs_hp(const Mav1FrameParams &fp, const uint8_t *srcL, const uint8_t *planes, int x, int y,
                          const int *cr, const int *cc, int ncand, int st, int lane, int *cost) {
  constexpr int NR = BW / 4, NU = NR * NR;
  const int W = fp.frame_w, H = fp.frame_h, nue = NU / st;
  if (lane < ncand) cost[lane] = 0;
  __syncthreads();
  for (int p = lane; p < ncand * nue; p += 64) {
    const int c = p / nue, u = (p - c * nue) * st;
    const int mvr = cr[c], mvc = cc[c];
    const uint8_t *pl = planes + (size_t)(((mvr >> 1) & 3) * 4 + ((mvc >> 1) & 3)) * W * H;
    const int ucol = u % NR, urow = u / NR;
    const int iy = y + urow * 4 + (mvr >> 3), ix = x + ucol * 4 + (mvc >> 3);
    int d[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint8_t *prow = pl + (size_t)d_clip(0, H - 1, iy + i) * W;
      const uint8_t *srow = srcL + (urow * 4 + i) * BW + ucol * 4;
#pragma unroll
      for (int j = 0; j < 4; j++) d[4 * i + j] = (int)srow[j] - (int)prow[d_clip(0, W - 1, ix + j)];
    }
    atomicAdd(&cost[c], satd4x4_arr(d) * st);
  }
  __syncthreads();
}

/* Sub-pel cost at an even 1/8-pel vector without the quarter-pel planes
 * (k_me.h:247-258): the normative EIGHTTAP prediction of the block
 * (7.11.3, staged through LDS) and the SATD of all its 4x4 units. */
template <int BW>
DEVI int subpel_cost_pred(const Mav1FrameParams &fp, const uint8_t *srcL, const uint8_t *refY, int x, int y, int mvr,
                          int mvc, uint8_t *pwin, int16_t *pmid, int lane) {
  constexpr int WW = BW + 8, NR = BW / 4;
  const int S = fp.stride[0], lastX = fp.frame_w - 1, lastY = fp.frame_h - 1;
  const int posX = ((x << SUBPEL_BITS) + 2 * mvc) * 64 + 32, posY = ((y << SUBPEL_BITS) + 2 * mvr) * 64 + 32;
  const int ix = posX >> 10, iy = posY >> 10;
  const int16_t *kh = SUBPEL_FILTER(EIGHTTAP, (posX >> 6) & SUBPEL_MASK), *kv = SUBPEL_FILTER(EIGHTTAP, (posY >> 6) & SUBPEL_MASK);
  __syncthreads();
  for (int k = lane; k < (BW + 7) * (BW + 7); k += 64) {
    const int r = k / (BW + 7), c = k - r * (BW + 7);
    pwin[r * WW + c] = refY[d_clip(0, lastY, iy - 3 + r) * S + d_clip(0, lastX, ix - 3 + c)];
  }
  __syncthreads();
  for (int k = lane; k < (BW + 7) * BW; k += 64) {
    const int r = k / BW, c = k - r * BW;
    const uint8_t *q = pwin + r * WW + c;
    int acc = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) acc += kh[t] * q[t];
    pmid[k] = (int16_t)d_round2(acc, 3);
  }
  __syncthreads();
  int s = 0;
  for (int u = lane; u < NR * NR; u += 64) {
    const int ucol = u % NR, urow = u / NR;
    int d[16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int r = urow * 4 + i, c = ucol * 4 + j;
        int acc = 0;
#pragma unroll
        for (int t = 0; t < 8; t++) acc += kv[t] * pmid[(r + t) * BW + c];
        d[4 * i + j] = (int)srcL[r * BW + c] - d_clip1(d_round2(acc, 11));
      }
    s += satd4x4_arr(d);
  }
  return wave_sum(s);
}

/* HP: sub-pel costs from the quarter-pel planes, else from predictions. */
template <int LEVEL, bool HP>
DEVI void me_level_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, const uint8_t *__restrict__ ref0, const uint8_t *__restrict__ ref1, const uint8_t *__restrict__ planes, long entryBytes, const uint8_t *__restrict__ srcDec, int32_t *__restrict__ field0, int32_t *__restrict__ field1, int nodesPerRef, const int blk_x, const int blk_y, const int blk_z) {
  constexpr int BW = 64 >> LEVEL;
  constexpr int N = 1 << LEVEL;
  constexpr int WW = BW + 2 * WIN_R;
  __shared__ uint8_t srcL[BW * BW];
  __shared__ uint8_t win[WW * WW];
  __shared__ uint8_t pwin[HP ? 1 : (BW + 7) * (BW + 8)];
  __shared__ int16_t pmid[HP ? 1 : (BW + 7) * BW];
  __shared__ int scost[9];
  const int lane = threadIdx.x;
  const int r = blk_x / nodesPerRef, node = blk_x - r * nodesPerRef;
  const int sb = node >> (2 * LEVEL), i = node & (N * N - 1);
  const uint8_t *ref = r ? ref1 : ref0;
  int32_t *field = r ? field1 : field0;
  const uint8_t *pl = planes + r * entryBytes;
  const uint8_t *refDec = pl + (size_t)16 * fp.frame_w * fp.frame_h;

  NodeState m;
  m.x = (sb % fp.sb_cols) * 64 + (i % N) * BW;
  m.y = (sb / fp.sb_cols) * 64 + (i / N) * BW;
  int32_t *out = field + (sb * MAV1_NUM_NODES + d_node_base(LEVEL) + i) * 2;
  if (m.x >= fp.frame_w || m.y >= fp.frame_h) {
    if (lane == 0) out[0] = out[1] = 0;
    return;
  }
  m.range = fp.me_range;
  m.step = 1;
  m.nstage = 1;
  m.pr = m.pc = m.cy = m.cx = 0;
  m.copy = 0;
  if (LEVEL > 0) {
    int parent = ((i / N) >> 1) * (N >> 1) + ((i % N) >> 1);
    const int32_t *pv = field + (sb * MAV1_NUM_NODES + d_node_base(LEVEL - 1) + parent) * 2;
    m.pr = pv[0];
    m.pc = pv[1];
    m.cy = m.pr >> 3;
    m.cx = m.pc >> 3;
    m.range = d_min(m.range, TB(fp, MAV1_T_ME_CR1) ? 1 : TB(fp, MAV1_T_ME_CR2) ? 2 : 4);
    bool leafNo8 = TB(fp, MAV1_T_LEAF_NO8) && fp.leaf && !fp.frame_is_intra;
    m.copy = LEVEL == 3 && (TB(fp, MAV1_T_ME_L3COPY) || leafNo8);
  } else {
    m.step = TB(fp, MAV1_T_ME_HME3) ? 8 : 4;
    m.range = (m.range / m.step) * m.step;
    m.nstage = TB(fp, MAV1_T_ME_HME3) ? 3 : 2;
  }
  if (LEVEL == 3 && TB(fp, MAV1_T_ME_L3SUB)) m.nstage = 0;
  if (r && LEVEL == 3 && TB(fp, MAV1_T_ME_G3COPY) && fp.leaf && !fp.frame_is_intra) m.copy = 1;
  if (m.copy) {
    if (lane == 0) {
      out[0] = m.pr;
      out[1] = m.pc;
    }
    return;
  }
  const int S = fp.stride[0];
  const uint8_t *srcY = src + fp.offset[0], *refY = ref + fp.offset[0];
  for (int k = lane; k < BW * BW; k += 64) srcL[k] = srcY[(m.y + k / BW) * S + m.x + (k % BW)];
  __syncthreads();

  int by = 0, bx = 0;
  for (int stg = 0; stg < m.nstage; stg++) {
    if (stg > 0) {
      m.cy = by;
      m.cx = bx;
      if (stg == 1 && TB(fp, MAV1_T_ME_HME3)) {
        m.range = 4;
        m.step = 4;
      } else {
        m.range = TB(fp, MAV1_T_ME_REF2) ? 2 : 3;
        m.step = 1;
      }
    }
    const int side = 2 * (m.range / m.step) + 1, count = 1 + side * side;
    const bool dec = TB(fp, MAV1_T_ME_DEC) && m.step >= 4;
    const int rs = (TB(fp, MAV1_T_ME_RSUB) && m.step == 1) ? 2 : 1;
    const bool lds = !dec && m.range <= WIN_R;
    const int oy0 = m.cy - m.range, ox0 = m.cx - m.range, wdim = BW + 2 * m.range;
    if (lds) {
      for (int k = lane; k < wdim * wdim; k += 64) {
        int wr = k / wdim, wc = k - wr * wdim;
        win[wr * WW + wc] = refY[d_clip(0, fp.frame_h - 1, m.y + oy0 + wr) * S + d_clip(0, fp.frame_w - 1, m.x + ox0 + wc)];
      }
      __syncthreads();
    }
    unsigned long long best = ~0ull;
    for (int k = lane; k < count; k += 64) {
      int dy, dx;
      me_cand(m, k, dy, dx);
      int rate = fp.ac_q[0] * me_bits(dy, dx) / 64, c;
      if (dec)
        c = sad_dec<BW>(fp, srcDec, refDec, m.x, m.y, dy, dx);
      else if (lds)
        c = sad_lds<BW>(srcL, win, WW, dy - oy0, dx - ox0, rs);
      else
        c = sad_global<BW>(fp, srcL, refY, m.x, m.y, dy, dx, rs);
      unsigned long long key = pack_key(c + rate, k);
      best = key < best ? key : best;
    }
    best = wave_min_u64(best);
    me_cand(m, (int)(best & 0xffffffffu), by, bx);
    __syncthreads();
  }
  if (LEVEL > 0 && TB(fp, MAV1_T_ME_KEEP) && by == m.cy && bx == m.cx) {
    if (lane == 0) {
      out[0] = m.pr;
      out[1] = m.pc;
    }
    return;
  }
  /* sub-pel refinement: half-pel ring, then quarter-pel ring unless
   * T_ME_HALF01 stops nodes >= 32 wide at half-pel */
  int mvr = m.nstage ? by * 8 : m.pr, mvc = m.nstage ? bx * 8 : m.pc;
  const int st = (TB(fp, MAV1_T_ME_USUB) && BW >= 32) ? 2 : 1;
  auto subpel = [&](int r_, int c_) {
    if constexpr (HP) return subpel_cost_hp<BW>(fp, srcL, pl, m.x, m.y, r_, c_, st, lane);
    else return subpel_cost_pred<BW>(fp, srcL, refY, m.x, m.y, r_, c_, pwin, pmid, lane);
  };
  const int minStep = (BW >= 32 && TB(fp, MAV1_T_ME_HALF01)) ? 4 : 2;
  const bool cross = TB(fp, MAV1_T_ME_CROSS);
  if constexpr (HP) {
    /* The ring's costs do not depend on each other: all of a ring at once
     * (the centre with the first ring), then the serial walk over them in
     * ring order with the strict <, which keeps its first minimum. */
    int bestS = 0;
    for (int step = 4; step >= minStep; step >>= 1) {
      const int cr0 = mvr, cc0 = mvc, first = step == 4;
      int vr[9], vc[9], n = 0;
      if (first) {
        vr[0] = cr0;
        vc[0] = cc0;
        n = 1;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (cross && !(j == 1 || j == 3 || j == 4 || j == 6)) continue;
        vr[n] = cr0 + (j < 3 ? -1 : j < 5 ? 0 : 1) * step;
        vc[n] = cc0 + ((j == 0 || j == 3 || j == 5) ? -1 : (j == 1 || j == 6) ? 0 : 1) * step;
        n++;
      }
      subpel_costs_hp<BW>(fp, srcL, pl, m.x, m.y, vr, vc, n, st, lane, scost);
      if (first) bestS = scost[0];
      for (int k = first; k < n; k++) {
        const int c = scost[k];
        if (c < bestS) {
          bestS = c;
          mvr = vr[k];
          mvc = vc[k];
        }
      }
      __syncthreads();
    }
    if (lane == 0) {
      out[0] = mvr;
      out[1] = mvc;
    }
    return;
  }
  int bestS = subpel(mvr, mvc);
  for (int step = 4; step >= minStep; step >>= 1) {
    const int cr = mvr, cc = mvc;
    for (int j = 0; j < (cross ? 4 : 8); j++) {
      int k = cross ? (j == 0 ? 1 : j == 1 ? 3 : j == 2 ? 4 : 6) : j;
      int dr = (k < 3 ? -1 : k < 5 ? 0 : 1) * step;
      int dc = ((k == 0 || k == 3 || k == 5) ? -1 : (k == 1 || k == 6) ? 0 : 1) * step;
      int c = subpel(cr + dr, cc + dc);
      if (c < bestS) {
        bestS = c;
        mvr = cr + dr;
        mvc = cc + dc;
      }
    }
  }
  if (lane == 0) {
    out[0] = mvr;
    out[1] = mvc;
  }
}
template <int LEVEL, bool HP>
__global__ __launch_bounds__(64) void k_me_level(const Mav1FrameParams fp, const uint8_t *__restrict__ src,
                                                 const uint8_t *__restrict__ ref0, const uint8_t *__restrict__ ref1,
                                                 const uint8_t *__restrict__ planes, long entryBytes,
                                                 const uint8_t *__restrict__ srcDec, int32_t *__restrict__ field0,
                                                 int32_t *__restrict__ field1, int nodesPerRef) {
  me_level_body<LEVEL, HP>(fp, src, ref0, ref1, planes, entryBytes, srcDec, field0, field1, nodesPerRef,
                           xcd_item(blockIdx.x, gridDim.x, 1 << (2 * LEVEL)), blockIdx.y, blockIdx.z);
}

template <int LEVEL>
static void launch_level(hipStream_t st, bool hp, dim3 grid, const Mav1FrameParams &fp, const uint8_t *src,
                         const uint8_t *ref0, const uint8_t *ref1, const uint8_t *planes, long entryBytes,
                         const uint8_t *srcDec, int32_t *field0, int32_t *field1, int nodesPerRef) {
  if (hp)
    hipLaunchKernelGGL((k_me_level<LEVEL, true>), grid, dim3(64), 0, st, fp, src, ref0, ref1, planes, entryBytes, srcDec,
                       field0, field1, nodesPerRef);
  else
    hipLaunchKernelGGL((k_me_level<LEVEL, false>), grid, dim3(64), 0, st, fp, src, ref0, ref1, planes, entryBytes,
                       srcDec, field0, field1, nodesPerRef);
}

hipError_t launch_me_level(hipStream_t st, int level, const Mav1FrameParams &fp, const uint8_t *src,
                           const uint8_t *ref0, const uint8_t *ref1, const uint8_t *planes, long entryBytes,
                           const uint8_t *srcDec, int32_t *field0, int32_t *field1, int nref) {
  int nodesPerRef = fp.sb_cols * fp.sb_rows << (2 * level);
  dim3 grid(nodesPerRef * nref);
  const bool hp = (fp.turbo & MAV1_T_ME_HPEL) != 0;
  switch (level) {
    case 0: launch_level<0>(st, hp, grid, fp, src, ref0, ref1, planes, entryBytes, srcDec, field0, field1, nodesPerRef); break;
    case 1: launch_level<1>(st, hp, grid, fp, src, ref0, ref1, planes, entryBytes, srcDec, field0, field1, nodesPerRef); break;
    case 2: launch_level<2>(st, hp, grid, fp, src, ref0, ref1, planes, entryBytes, srcDec, field0, field1, nodesPerRef); break;
    default: launch_level<3>(st, hp, grid, fp, src, ref0, ref1, planes, entryBytes, srcDec, field0, field1, nodesPerRef); break;
  }
  return hipGetLastError();
}

/* ---- batched launches (batch.h): the same kernels over several frames ---- */

/* which 0: LAST's planes, 1: the second reference's. */
__global__ __launch_bounds__(256) void k_qpel_build_b(const BatchFrame *__restrict__ T, const BatchGrid g, int which) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  qpel_build_body(F.fp, which ? gptr(F.gold) : gptr(F.ref), gptr(F.planes) + which * F.entry_bytes, b.x, b.y, b.z);
}

hipError_t launch_qpel_build_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int which) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++)
    batch_grid_add(g, (H[f].fp.frame_w + QT_W - 1) / QT_W, (H[f].fp.frame_h + QT_H - 1) / QT_H, 1);
  hipLaunchKernelGGL(k_qpel_build_b, dim3(g.start[n]), dim3(256), 0, st, T, g, which);
  return hipGetLastError();
}

/* which 0 / 1: the decimated plane of LAST / the second reference behind
 * its 16 quarter-pel planes; 2: the source's, after both entries. */
__global__ __launch_bounds__(256) void k_dec_build_b(const BatchFrame *__restrict__ T, const BatchGrid g, int which) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  const size_t n16 = (size_t)16 * F.fp.frame_w * F.fp.frame_h;
  const uint8_t *frame = which == 2 ? gptr(F.src) : which ? gptr(F.gold) : gptr(F.ref);
  uint8_t *dec = which == 2 ? gptr(F.planes) + 2 * F.entry_bytes : gptr(F.planes) + which * F.entry_bytes + n16;
  dec_build_body(F.fp, frame, dec, b.x, b.y, b.z);
}

hipError_t launch_dec_build_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int which) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) {
    const int dw = (H[f].fp.frame_w + 3) >> 2, dh = (H[f].fp.frame_h + 3) >> 2;
    batch_grid_add(g, (dw + 63) / 64, (dh + 3) / 4, 1);
  }
  hipLaunchKernelGGL(k_dec_build_b, dim3(g.start[n]), dim3(256), 0, st, T, g, which);
  return hipGetLastError();
}

template <int LEVEL, bool HP>
__global__ __launch_bounds__(64) void k_me_level_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  const BatchIdx b = batch_idx(g, xcd_item(blockIdx.x, gridDim.x, 1 << (2 * LEVEL)));
  const BatchFrame &F = T[b.f];
  me_level_body<LEVEL, HP>(F.fp, gptr(F.src), gptr(F.ref), gptr(F.gold), gptr(F.planes), (long)F.entry_bytes, gptr(F.planes) + 2 * F.entry_bytes,
                           gptr(F.mvf), gptr(F.mvg), g.aux[b.f], b.x, b.y, b.z);
}

template <int LEVEL>
static void launch_level_b(hipStream_t st, bool hp, const BatchFrame *T, const BatchGrid &g) {
  if (hp)
    hipLaunchKernelGGL((k_me_level_b<LEVEL, true>), dim3(g.start[g.n]), dim3(64), 0, st, T, g);
  else
    hipLaunchKernelGGL((k_me_level_b<LEVEL, false>), dim3(g.start[g.n]), dim3(64), 0, st, T, g);
}

hipError_t launch_me_level_b(hipStream_t st, int level, const BatchFrame *T, const BatchFrame *H, int n, int nref) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) {
    const int nodesPerRef = H[f].fp.sb_cols * H[f].fp.sb_rows << (2 * level);
    batch_grid_add(g, nodesPerRef * nref, 1, 1, nodesPerRef);
  }
  const bool hp = (H[0].fp.turbo & MAV1_T_ME_HPEL) != 0;
  switch (level) {
    case 0: launch_level_b<0>(st, hp, T, g); break;
    case 1: launch_level_b<1>(st, hp, T, g); break;
    case 2: launch_level_b<2>(st, hp, T, g); break;
    default: launch_level_b<3>(st, hp, T, g); break;
  }
  return hipGetLastError();
}
