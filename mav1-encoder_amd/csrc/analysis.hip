/* Open-loop mode and partition analysis on gfx950 (behaviour of
 * src/kernels/k_analysis.h; per-SB driver src/backend_cpu.c:58-74).
 *
 * Pass 1: one workgroup per table item (superblock, node, slot), one launch
 * per level (a node's depth pruning reads its ancestors' costs): four waves
 * for the 64x64 items, two for the 32x32 ones, a wave64 for the smaller ones. The item's
 * source block (luma + both chroma planes) is staged in LDS; inter
 * predictions run through an LDS reference window (8-tap horizontal pass
 * into int16 LDS rows, vertical pass into LDS bytes); intra predictions are
 * produced per sample inside the SATD loop from LDS edge arrays. 4x4 units
 * are spread over the threads and every SATD is a workgroup reduction, so
 * each cost is a uniform value and the mode decisions (first minimum, ties
 * as the serial code) are taken identically by all threads.
 *
 * Pass 2: one wave per superblock runs the bottom-up partition choice over
 * the 85 nodes (lanes over the nodes of a level) and writes the blocks in
 * decode order (lanes over the 64 Z-order positions, prefix-scanned).
 */
#include "dev_pred.h"
#include "internal.h"

#define TB(fp, bit) (((fp).turbo & (bit)) != 0)

DEVI int mode_bits8(int m) {
  if (m == DC_PRED) return 12;
  if (m == V_PRED || m == H_PRED || m == SMOOTH_PRED || m == PAETH_PRED) return 26;
  if (m == SMOOTH_V_PRED || m == SMOOTH_H_PRED) return 34;
  return 40;
}
DEVI int mv_bits8(int r, int c) {
  return 8 * (2 + 2 * d_floorlog2((unsigned)d_abs(r) + 1) + 2 * d_floorlog2((unsigned)d_abs(c) + 1));
}
DEVI int level_bsize(int l) { return l == 0 ? BLOCK_64X64 : l == 1 ? BLOCK_32X32 : l == 2 ? BLOCK_16X16 : BLOCK_8X8; }
DEVI bool turbo_square(const Mav1FrameParams &fp) {
  return TB(fp, MAV1_T_SQUARE) && (!fp.frame_is_intra || TB(fp, MAV1_T_KEY_SQUARE));
}
DEVI bool part_allowed(const Mav1FrameParams &fp, int level, int p, int hasRows, int hasCols) {
  if (p == PARTITION_SPLIT) return false;
  if (level == 3 && p != PARTITION_NONE) return false;
  if (fp.fast && p > PARTITION_SPLIT) return false;
  if (turbo_square(fp) && p != PARTITION_NONE) return false;
  if (level == 2 && (p == PARTITION_HORZ_4 || p == PARTITION_VERT_4)) return false;
  if (!(hasRows && hasCols)) return (p == PARTITION_HORZ && hasCols && !hasRows) || (p == PARTITION_VERT && hasRows && !hasCols);
  return true;
}
DEVI int part_block(int p, int bSize, int k, int halfMi, int &dr, int &dc) {
  int sub = PARTITION_SUBSIZE(p, bSize), split = PARTITION_SUBSIZE(PARTITION_SPLIT, bSize), q = halfMi >> 1;
  dr = dc = 0;
  switch (p) {
    case PARTITION_HORZ: dr = k * halfMi; return sub;
    case PARTITION_VERT: dc = k * halfMi; return sub;
    case PARTITION_HORZ_A:
      if (k < 2) { dc = k * halfMi; return split; }
      dr = halfMi;
      return sub;
    case PARTITION_HORZ_B:
      if (k == 0) return sub;
      dr = halfMi;
      dc = (k - 1) * halfMi;
      return split;
    case PARTITION_VERT_A:
      if (k < 2) { dr = k * halfMi; return split; }
      dc = halfMi;
      return sub;
    case PARTITION_VERT_B:
      if (k == 0) return sub;
      dc = halfMi;
      dr = (k - 1) * halfMi;
      return split;
    case PARTITION_HORZ_4: dr = k * q; return sub;
    case PARTITION_VERT_4: dc = k * q; return sub;
    default: return sub;
  }
}
DEVI int part_count(const Mav1FrameParams &fp, int p, int r, int c, int halfMi) {
  int hasRows = r + halfMi < fp.mi_rows, hasCols = c + halfMi < fp.mi_cols;
  switch (p) {
    case PARTITION_NONE: return 1;
    case PARTITION_HORZ: return hasRows ? 2 : 1;
    case PARTITION_VERT: return hasCols ? 2 : 1;
    case PARTITION_HORZ_4: return r + (halfMi >> 1) * 3 < fp.mi_rows ? 4 : 3;
    case PARTITION_VERT_4: return c + (halfMi >> 1) * 3 < fp.mi_cols ? 4 : 3;
    default: return 3;
  }
}
DEVI int part_slot(int p, int k) {
  switch (p) {
    case PARTITION_NONE: return 0;
    case PARTITION_HORZ: return 1 + k;
    case PARTITION_VERT: return 3 + k;
    case PARTITION_HORZ_4: return 5 + k;
    case PARTITION_VERT_4: return 9 + k;
    case PARTITION_HORZ_A: return k < 2 ? -1 : 2;
    case PARTITION_HORZ_B: return k == 0 ? 1 : -1;
    case PARTITION_VERT_A: return k < 2 ? -1 : 4;
    default: return k == 0 ? 3 : -1;
  }
}

int analysis_level_slots(const Mav1FrameParams &fp, int level) {
  bool sq = (fp.turbo & MAV1_T_SQUARE) && (!fp.frame_is_intra || (fp.turbo & MAV1_T_KEY_SQUARE));
  if (level == 3 || sq) return 1;
  if (level == 2 || fp.fast) return 5;
  return MAV1_A_SLOTS;
}

template <int BW>
struct AnaLds {
  uint8_t srcY[BW * BW];
  uint8_t srcC[2][(BW / 2) * (BW / 2)];
  uint8_t pa[BW * BW];
  uint8_t pb[BW * BW];
  uint8_t win[(BW + 7) * (BW + 8)];
  int16_t mid[(BW + 7) * BW];
  int edge[3][2][EDGE_NEG + 2 * BW + 8]; /* [plane][above, left] */
  int dcp[BW / 2 * BW / 2];
};

template <int LEVEL>
DEVI void analyze_level_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, const uint8_t *__restrict__ ref, const uint8_t *__restrict__ gold, const int32_t *__restrict__ mvf, const int32_t *__restrict__ mvg, long long *__restrict__ costs, Mav1BlockInfo *__restrict__ infos, int nslots, const int8_t *__restrict__ prune, const int blk_x, const int blk_y, const int blk_z) {
  constexpr int BW = 64 >> LEVEL, N = 1 << LEVEL;
  /* threads per item: four waves for the 64x64 level (510 items at 1080p),
   * two for the 32x32 level, one below */
  constexpr int NT = LEVEL == 0 ? 256 : LEVEL == 1 ? 128 : 64;
  __shared__ AnaLds<BW> L;
  const int lane = threadIdx.x;
  const int perSb = N * N * nslots;
  const int sb = blk_x / perSb, rem = blk_x - sb * perSb;
  const int i = rem / nslots, slot = rem - i * nslots;
  const int sbRow = sb / fp.sb_cols, sbCol = sb - sbRow * fp.sb_cols;
  const int size4 = 16 >> LEVEL, half = size4 >> 1, node = d_node_base(LEVEL) + i;
  const int r = sbRow * 16 + (i / N) * size4, c = sbCol * 16 + (i % N) * size4;
  const long long idx = (long long)sb * MAV1_A_ITEMS + node * MAV1_A_SLOTS + slot;

  /* ---- item validity (analyze_item) ---- */
  bool skip = r >= fp.mi_rows || c >= fp.mi_cols;
  /* the learned pruner's deepest level for this superblock (batched form
   * only; an item it cuts is "not evaluated", as depth pruning leaves it) */
  if (!skip && prune && LEVEL > prune[sb]) skip = true;
  if (!skip && fp.fast && LEVEL > 0 && fp.depth_thr) {
    int yy = i >> LEVEL, xx = i & (N - 1);
    for (int a = LEVEL - 1; a >= 0 && !skip; a--) {
      yy >>= 1;
      xx >>= 1;
      long long pc = costs[(long long)sb * MAV1_A_ITEMS + (d_node_base(a) + (yy << a) + xx) * MAV1_A_SLOTS];
      int pw = 64 >> a;
      if (pc >= 0 && pc * 16 < (long long)pw * pw * fp.depth_thr) skip = true;
    }
  }
  if (!skip && LEVEL == 3 && TB(fp, MAV1_T_LEAF_NO8) && fp.leaf && !fp.frame_is_intra) {
    int pr = sbRow * 16 + ((i >> 3) & ~1) * 2, pcn = sbCol * 16 + ((i & 7) & ~1) * 2;
    if (!(pr + 4 > fp.mi_rows || pcn + 4 > fp.mi_cols)) skip = true;
  }
  const int hasRows = r + half < fp.mi_rows, hasCols = c + half < fp.mi_cols;
  /* slot -> (partition, block): 0 NONE, 1-2 HORZ, 3-4 VERT, 5-8 HORZ_4, 9-12 VERT_4 */
  const int p = slot == 0 ? PARTITION_NONE : slot < 3 ? PARTITION_HORZ : slot < 5 ? PARTITION_VERT
                : slot < 9 ? PARTITION_HORZ_4 : PARTITION_VERT_4;
  const int k = slot == 0 ? 0 : slot < 3 ? slot - 1 : slot < 5 ? slot - 3 : slot < 9 ? slot - 5 : slot - 9;
  if (!skip && !part_allowed(fp, LEVEL, p, hasRows, hasCols)) skip = true;
  if (!skip && k >= part_count(fp, p, r, c, half)) skip = true;
  if (skip) {
    if (lane == 0) costs[idx] = -1;
    return;
  }
  int dr, dcol;
  const int bs = part_block(p, level_bsize(LEVEL), k, half, dr, dcol);
  const int miRow = r + dr, miCol = c + dcol;

  /* ---- eval_block ---- */
  const int lw = Mi_Width_Log2[bs] + 2, lh = Mi_Height_Log2[bs] + 2, w = 1 << lw, h = 1 << lh;
  const int x = miCol * 4, y = miRow * 4, lam = fp.ac_q[0] / 16;
  const int cx = x >> 1, cy = y >> 1, clw = lw - 1, clh = lh - 1, cw = w >> 1, ch = h >> 1;
  for (int q = lane; q < w * h; q += NT) {
    int yy = q >> lw, xx = q & (w - 1);
    L.srcY[yy * BW + xx] = src[fp.offset[0] + (y + yy) * fp.stride[0] + x + xx];
  }
  for (int q = lane; q < 2 * cw * ch; q += NT) {
    int pl = q / (cw * ch), e = q - pl * cw * ch, yy = e >> clw, xx = e & (cw - 1);
    L.srcC[pl][yy * (BW / 2) + xx] = src[fp.offset[1 + pl] + (cy + yy) * fp.stride[1] + cx + xx];
  }
  __syncthreads();

  long long interCost = -1;
  int mvr = 0, mvc = 0, refFrame = LAST_FRAME, refFrame1 = 0, mvr1 = 0, mvc1 = 0;
  if (!fp.frame_is_intra) {
    const int sbIdx = (y >> 6) * fp.sb_cols + (x >> 6);
    const int lr = mvf[(sbIdx * MAV1_NUM_NODES + node) * 2], lc = mvf[(sbIdx * MAV1_NUM_NODES + node) * 2 + 1];
    int gr = 0, gc = 0;
    if (fp.golden) {
      gr = mvg[(sbIdx * MAV1_NUM_NODES + node) * 2];
      gc = mvg[(sbIdx * MAV1_NUM_NODES + node) * 2 + 1];
    }
    const int want = fp.golden ? (fp.compound ? 3 : 2) : 1;
    long long s0 = 0, s1 = 0, s2 = 0;
    for (int pl = 0; pl < 3; pl++) {
      const int sub = pl > 0, bx = x >> sub, by = y >> sub, bw = w >> sub, bh = h >> sub, S = fp.stride[pl];
      const int st = sub ? BW / 2 : BW;
      const uint8_t *sblk = sub ? L.srcC[pl - 1] : L.srcY;
      const int usub = TB(fp, MAV1_T_ANA_USUB) && bw >= 32;
      wave_predict_inter<NT>(ref + fp.offset[pl], S, fp.frame_w, fp.frame_h, bx, by, bw, bh, lr, lc, sub, EIGHTTAP, L.win,
                         L.mid, L.pa, st, nullptr, 11, lane);
      if (want > 1)
        wave_predict_inter<NT>(gold + fp.offset[pl], S, fp.frame_w, fp.frame_h, bx, by, bw, bh, gr, gc, sub, EIGHTTAP,
                           L.win, L.mid, L.pb, st, nullptr, 11, lane);
      int a0, a1, a2;
      wave_satd3<NT>(sblk, L.pa, L.pb, st, bw, bh, usub, want, lane, &a0, &a1, &a2);
      s0 += a0;
      s1 += a1;
      s2 += a2;
    }
    interCost = s0 + (long long)lam * (4 + mv_bits8(lr, lc)) / 8;
    mvr = lr;
    mvc = lc;
    if (fp.golden) {
      long long g = s1 + (long long)lam * (4 + mv_bits8(gr, gc)) / 8 + (long long)lam * 8 / 8;
      if (g < interCost) {
        interCost = g;
        mvr = gr;
        mvc = gc;
        refFrame = fp.ref2_frame;
      }
      if (fp.compound) {
        long long cc = s2 + (long long)lam * (8 + mv_bits8(lr, lc) + mv_bits8(gr, gc) + 16) / 8;
        if (cc < interCost) {
          interCost = cc;
          mvr = lr;
          mvc = lc;
          mvr1 = gr;
          mvc1 = gc;
          refFrame = LAST_FRAME;
          refFrame1 = fp.ref2_frame;
        }
      }
    }
  }

  /* luma intra (source neighbours, no edge filter, no above-right / below-left) */
  const bool tIntra = TB(fp, MAV1_T_INTRA) && !fp.frame_is_intra;
  long long best = -1;
  int bestMode = DC_PRED, bestDelta = 0;
  const bool skipIntra = interCost >= 0 && (fp.leaf || (fp.fast && interCost * 16 < (long long)w * h * lam * 16));
  int *AY = L.edge[0][0] + EDGE_NEG, *LY = L.edge[0][1] + EDGE_NEG;
  const int maxXY = fp.mi_cols * 4 - 1, maxYY = fp.mi_rows * 4 - 1;
  if (!skipIntra) {
    wave_intra_edges<NT>(src + fp.offset[0], fp.stride[0], x, y, w, h, x > 0, y > 0, 0, 0, maxXY, maxYY, AY, LY, lane);
    const int dcY = wave_intra_dc<NT>(AY, LY, w, h, lw, lh, x > 0, y > 0, lane);
    for (int m = DC_PRED; m <= PAETH_PRED; m++) {
      if (tIntra && m != DC_PRED && m != V_PRED && m != H_PRED && m != SMOOTH_PRED) continue;
      IntraMode im;
      intra_mode_setup(im, m, 0, lw, lh);
      im.dcv = dcY;
      im.corner = AY[-1];
      long long cst = wave_intra_satd<NT>(im, AY, LY, L.srcY, BW, lane) + (long long)lam * mode_bits8(m) / 8;
      if (best < 0 || cst < best) {
        best = cst;
        bestMode = m;
      }
    }
  }
  if (skipIntra) best = 1ll << 40;
  if (bestMode >= V_PRED && bestMode <= D67_PRED && !fp.fast) {
    for (int d = -MAX_ANGLE_DELTA; d <= MAX_ANGLE_DELTA; d++) {
      if (!d) continue;
      IntraMode im;
      intra_mode_setup(im, bestMode, d, lw, lh);
      im.corner = AY[-1];
      im.dcv = 0;
      long long cst = wave_intra_satd<NT>(im, AY, LY, L.srcY, BW, lane) + (long long)lam * (mode_bits8(bestMode) + 16) / 8;
      if (cst < best) {
        best = cst;
        bestDelta = d;
      }
    }
  }

  /* chroma intra */
  const bool skipUv = fp.fast && interCost >= 0 && best + (long long)lam * 16 / 8 >= interCost;
  long long bestUv = skipUv ? 0 : -1;
  int bestUvMode = DC_PRED;
  const int maxXC = ((fp.mi_cols * 4) >> 1) - 1, maxYC = ((fp.mi_rows * 4) >> 1) - 1;
  int dcC[2] = {128, 128};
  if (!skipUv) {
    for (int pl = 0; pl < 2; pl++) {
      int *A = L.edge[1 + pl][0] + EDGE_NEG, *Lc = L.edge[1 + pl][1] + EDGE_NEG;
      wave_intra_edges<NT>(src + fp.offset[1 + pl], fp.stride[1], cx, cy, cw, ch, cx > 0, cy > 0, 0, 0, maxXC, maxYC, A, Lc,
                       lane);
      dcC[pl] = wave_intra_dc<NT>(A, Lc, cw, ch, clw, clh, cx > 0, cy > 0, lane);
    }
    for (int m = DC_PRED; m <= PAETH_PRED; m++) {
      if (m >= V_PRED && m <= D67_PRED && m != V_PRED && m != H_PRED && m != bestMode) continue;
      if (tIntra && m != DC_PRED && m != V_PRED && m != H_PRED && m != SMOOTH_PRED) continue;
      const int d = m == bestMode ? bestDelta : 0;
      long long cst = (long long)lam * (m == DC_PRED ? 8 : 24) / 8;
      for (int pl = 0; pl < 2; pl++) {
        IntraMode im;
        intra_mode_setup(im, m, d, clw, clh);
        im.dcv = dcC[pl];
        im.corner = L.edge[1 + pl][0][EDGE_NEG - 1];
        cst += wave_intra_satd<NT>(im, L.edge[1 + pl][0] + EDGE_NEG, L.edge[1 + pl][1] + EDGE_NEG, L.srcC[pl], BW / 2, lane);
      }
      if (bestUv < 0 || cst < bestUv) {
        bestUv = cst;
        bestUvMode = m;
      }
    }
  }
  const bool cflOk = fp.lossless ? SUBSAMPLED_SIZE(bs, 1, 1) == BLOCK_4X4 : (w <= 32 && h <= 32);
  if (!skipUv && cflOk && !tIntra) {
    /* chroma-from-luma on the source luma: DC prediction (a constant) plus
     * alpha times the zero-mean luma, alpha by the smallest SSE */
    long long cst = (long long)lam * 40 / 8;
    const int nC = cw * ch;
    int lsum = 0;
    for (int q = lane; q < nC; q += NT) {
      int yy = q >> clw, xx = q & (cw - 1);
      int ly = d_min(yy << 1, h - 2), lx = d_min(xx << 1, w - 2);
      const uint8_t *l = L.srcY + ly * BW + lx;
      int v = ((int)l[0] + l[1] + l[BW] + l[BW + 1]) << 1;
      L.dcp[q] = v;
      lsum += v;
    }
    lsum = group_sum<NT>(lsum);
    const int avg = d_round2(lsum, clw + clh);
    __syncthreads();
    for (int pl = 0; pl < 2; pl++) {
      const uint8_t *sC = L.srcC[pl];
      const int dcv = dcC[pl];
      long long bestS = -1;
      int bestA = 0;
      for (int a = -16; a <= 16; a++) {
        int e = 0;
        for (int q = lane; q < nC; q += NT) {
          int yy = q >> clw, xx = q & (cw - 1);
          int dd = d_clip1(dcv + d_round2signed(a * (L.dcp[q] - avg), 6)) - sC[yy * (BW / 2) + xx];
          e += dd * dd;
        }
        long long sse = group_sum<NT>(e);
        if (bestS < 0 || sse < bestS) {
          bestS = sse;
          bestA = a;
        }
      }
      int acc = 0;
      for (int u = lane; u < nC / 16; u += NT) {
        int nx = cw >> 2, uy = u / nx, ux = u - uy * nx, dd[16];
#pragma unroll
        for (int ii = 0; ii < 4; ii++)
#pragma unroll
          for (int jj = 0; jj < 4; jj++) {
            int q = (uy * 4 + ii) * cw + ux * 4 + jj;
            int pv = d_clip1(dcv + d_round2signed(bestA * (L.dcp[q] - avg), 6));
            dd[4 * ii + jj] = (int)sC[(uy * 4 + ii) * (BW / 2) + ux * 4 + jj] - pv;
          }
        acc += satd4x4_arr(dd);
      }
      cst += group_sum<NT>(acc);
    }
    if (cst < bestUv) {
      bestUv = cst;
      bestUvMode = UV_CFL_PRED;
    }
  }

  if (lane == 0) {
    Mav1BlockInfo b = {};
    b.mi_row = miRow;
    b.mi_col = miCol;
    b.bsize = bs;
    b.y_mode = bestMode;
    b.angle_y = bestDelta;
    b.uv_mode = bestUvMode;
    b.angle_uv = bestUvMode == bestMode ? bestDelta : 0;
    b.ref_frame = INTRA_FRAME;
    b.tx_type = DCT_DCT;
    long long cost = best + bestUv + (long long)lam * (fp.frame_is_intra ? 0 : 16) / 8;
    if (interCost >= 0 && interCost < cost) {
      cost = interCost;
      b.is_inter = 1;
      b.y_mode = NEWMV;
      b.uv_mode = DC_PRED;
      b.angle_y = b.angle_uv = 0;
      b.mv_row = mvr;
      b.mv_col = mvc;
      b.ref_frame = refFrame;
      b.ref_frame1 = refFrame1;
      b.mv1_row = mvr1;
      b.mv1_col = mvc1;
    }
    costs[idx] = cost;
    infos[idx] = b;
  }
}
template <int LEVEL>
__global__ __launch_bounds__(LEVEL == 0 ? 256 : LEVEL == 1 ? 128 : 64) void k_analyze_level(const Mav1FrameParams fp, const uint8_t *__restrict__ src,
                                                      const uint8_t *__restrict__ ref, const uint8_t *__restrict__ gold,
                                                      const int32_t *__restrict__ mvf, const int32_t *__restrict__ mvg,
                                                      long long *__restrict__ costs, Mav1BlockInfo *__restrict__ infos,
                                                      int nslots) {
  analyze_level_body<LEVEL>(fp, src, ref, gold, mvf, mvg, costs, infos, nslots, nullptr,
                            xcd_item(blockIdx.x, gridDim.x, (1 << (2 * LEVEL)) * nslots), blockIdx.y, blockIdx.z);
}

/* Pass 2 (analyze_sb_table): one wave per superblock. The bottom-up DP runs
 * level by level with the lanes over the level's nodes (a node only reads
 * its children, decided by the previous level) and the per-node state in
 * LDS; the blocks are emitted with one lane per 8x8 position in Z order
 * and an exclusive prefix scan over the lanes giving each position's block
 * index and coefficient offset. */
DEVI int wave_excl_scan(int v, int lane) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

DEVI void analyze_table_body(const Mav1FrameParams &fp, const long long *__restrict__ costs, const Mav1BlockInfo *__restrict__ infos, Mav1BlockInfo *__restrict__ blocks, Mav1SbInfo *__restrict__ sbs, const int blk_x, const int blk_y, const int blk_z) {
  __shared__ long long cost[MAV1_NUM_NODES], noneCost[MAV1_NUM_NODES];
  __shared__ int part[MAV1_NUM_NODES];
  const int sb = blk_x, lane = threadIdx.x;
  const int sbRow = sb / fp.sb_cols, sbCol = sb - sbRow * fp.sb_cols, lam = fp.ac_q[0] / 16;
  const long long *tc = costs + (long long)sb * MAV1_A_ITEMS;
  const Mav1BlockInfo *ti = infos + (long long)sb * MAV1_A_ITEMS;
  for (int level = 3; level >= 0; level--) {
    const int n = 1 << level, size4 = 16 >> level, bSize = level_bsize(level);
    for (int i = lane; i < n * n; i += 64) {
      const int node = d_node_base(level) + i;
      const int r = sbRow * 16 + (i / n) * size4, c = sbCol * 16 + (i % n) * size4;
      noneCost[node] = -1;
      if (r >= fp.mi_rows || c >= fp.mi_cols) {
        cost[node] = 0;
        part[node] = -1;
        continue;
      }
      const int half = size4 >> 1, hasRows = r + half < fp.mi_rows, hasCols = c + half < fp.mi_cols;
      const int cn = 2 * n, ci = (i / n) * 2 * cn + (i % n) * 2, cb = level < 3 ? d_node_base(level + 1) : 0;
      long long bestC = -1, myNone = -1;
      int bestP = PARTITION_SPLIT;
      if (level < 3) {
        long long c0 = cost[cb + ci], c1 = cost[cb + ci + 1], c2 = cost[cb + ci + cn], c3 = cost[cb + ci + cn + 1];
        if (c0 >= 0 && c1 >= 0 && c2 >= 0 && c3 >= 0) bestC = c0 + c1 + c2 + c3 + (long long)lam * 16 / 8;
      }
      for (int p = PARTITION_NONE; p <= PARTITION_VERT_4; p++) {
        if (!part_allowed(fp, level, p, hasRows, hasCols)) continue;
        long long pc = 0;
        const int nb = part_count(fp, p, r, c, half);
        for (int k = 0; k < nb && pc >= 0; k++) {
          int s = part_slot(p, k), dr, dc;
          long long cc;
          if (s < 0) {
            part_block(p, bSize, k, half, dr, dc);
            cc = noneCost[cb + ci + (dr ? cn : 0) + (dc ? 1 : 0)];
          } else {
            cc = tc[node * MAV1_A_SLOTS + s];
          }
          pc = cc < 0 ? -1 : pc + cc;
        }
        if (pc < 0) continue;
        pc += (long long)lam * (p == PARTITION_NONE ? 8 : 24) / 8;
        if (p == PARTITION_NONE) myNone = pc;
        if (bestC < 0 || pc <= bestC) {
          bestC = pc;
          bestP = p;
        }
      }
      noneCost[node] = myNone;
      cost[node] = bestC;
      part[node] = bestP;
    }
    __syncthreads();
  }
  Mav1SbInfo *S = &sbs[sb];
  for (int q = lane; q < MAV1_NUM_NODES; q += 64) S->part[q] = part[q];
  /* emission: lane = Z-order position of an 8x8 unit */
  const int z = lane;
  int zx = 0, zy = 0;
#pragma unroll
  for (int b = 0; b < 3; b++) {
    zx |= ((z >> (2 * b)) & 1) << b;
    zy |= ((z >> (2 * b + 1)) & 1) << b;
  }
  int nb = 0, lvl = 0, p = 0, node = 0;
  for (int level = 0; level <= 3; level++) {
    const int shift = 3 - level, n = 1 << level;
    node = d_node_base(level) + (zy >> shift) * n + (zx >> shift);
    p = part[node];
    if (p < 0) break;
    if (p == PARTITION_SPLIT) continue;
    if ((zx & ((1 << shift) - 1)) == 0 && (zy & ((1 << shift) - 1)) == 0) {
      const int size4 = 16 >> level;
      nb = part_count(fp, p, sbRow * 16 + (zy >> shift) * size4, sbCol * 16 + (zx >> shift) * size4, size4 >> 1);
      lvl = level;
    }
    break;
  }
  Mav1BlockInfo mine[4];
  int ncoef = 0;
  {
    const int n = 1 << lvl, shift = 3 - lvl, half = (16 >> lvl) >> 1;
    const int cn = 2 * n, ci = (zy >> shift) * 2 * cn + (zx >> shift) * 2, cb = lvl < 3 ? d_node_base(lvl + 1) : 0;
    for (int k = 0; k < nb; k++) {
      int s = part_slot(p, k), dr, dc;
      if (s < 0) {
        part_block(p, level_bsize(lvl), k, half, dr, dc);
        mine[k] = ti[(cb + ci + (dr ? cn : 0) + (dc ? 1 : 0)) * MAV1_A_SLOTS];
      } else {
        mine[k] = ti[node * MAV1_A_SLOTS + s];
      }
      const int w = BLOCK_WIDTH(mine[k].bsize), h = BLOCK_HEIGHT(mine[k].bsize);
      ncoef += w * h + 2 * ((w >> 1) * (h >> 1));
    }
  }
  const int first = wave_excl_scan(nb, lane);
  int coefOff = sb * MAV1_COEFS_PER_SB + wave_excl_scan(ncoef, lane);
  for (int k = 0; k < nb; k++) {
    Mav1BlockInfo b = mine[k];
    const int w = BLOCK_WIDTH(b.bsize), h = BLOCK_HEIGHT(b.bsize);
    b.coef_off[0] = coefOff;
    coefOff += w * h;
    b.coef_off[1] = coefOff;
    coefOff += (w >> 1) * (h >> 1);
    b.coef_off[2] = coefOff;
    coefOff += (w >> 1) * (h >> 1);
    b.eob[0] = b.eob[1] = b.eob[2] = 0;
    blocks[sb * MAV1_MAX_BLOCKS_PER_SB + first + k] = b;
  }
  const int count = wave_sum(nb);
  if (lane == 0) {
    S->first_block = sb * MAV1_MAX_BLOCKS_PER_SB;
    S->num_blocks = count;
    S->sb_row = sbRow;
    S->sb_col = sbCol;
  }
}
__global__ __launch_bounds__(64) void k_analyze_table(const Mav1FrameParams fp, const long long *__restrict__ costs,
                                                      const Mav1BlockInfo *__restrict__ infos,
                                                      Mav1BlockInfo *__restrict__ blocks, Mav1SbInfo *__restrict__ sbs) {
  analyze_table_body(fp, costs, infos, blocks, sbs, blockIdx.x, blockIdx.y, blockIdx.z);
}

hipError_t launch_analyze(hipStream_t st, int level, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *ref,
                          const uint8_t *gold, const int32_t *mvf, const int32_t *mvg, long long *costs,
                          Mav1BlockInfo *infos) {
  const int nslots = analysis_level_slots(fp, level);
  const int items = fp.sb_cols * fp.sb_rows * (1 << (2 * level)) * nslots;
  switch (level) {
    case 0: hipLaunchKernelGGL(k_analyze_level<0>, dim3(items), dim3(256), 0, st, fp, src, ref, gold, mvf, mvg, costs, infos, nslots); break;
    case 1: hipLaunchKernelGGL(k_analyze_level<1>, dim3(items), dim3(128), 0, st, fp, src, ref, gold, mvf, mvg, costs, infos, nslots); break;
    case 2: hipLaunchKernelGGL(k_analyze_level<2>, dim3(items), dim3(64), 0, st, fp, src, ref, gold, mvf, mvg, costs, infos, nslots); break;
    default: hipLaunchKernelGGL(k_analyze_level<3>, dim3(items), dim3(64), 0, st, fp, src, ref, gold, mvf, mvg, costs, infos, nslots); break;
  }
  return hipGetLastError();
}

hipError_t launch_analyze_table(hipStream_t st, const Mav1FrameParams &fp, const long long *costs,
                                const Mav1BlockInfo *infos, Mav1BlockInfo *blocks, Mav1SbInfo *sbs) {
  const int nsb = fp.sb_cols * fp.sb_rows;
  hipLaunchKernelGGL(k_analyze_table, dim3(nsb), dim3(64), 0, st, fp, costs, infos, blocks, sbs);
  return hipGetLastError();
}

/* ---- batched launches (batch.h): the same kernels over several frames ---- */

template <int LEVEL>
__global__ __launch_bounds__(LEVEL == 0 ? 256 : LEVEL == 1 ? 128 : 64) void k_analyze_level_b(
    const BatchFrame *__restrict__ T, const BatchGrid g) {
  /* runs of 64 items: the frames' slot counts may differ */
  const BatchIdx b = batch_idx(g, xcd_item(blockIdx.x, gridDim.x, 64));
  const BatchFrame &F = T[b.f];
  analyze_level_body<LEVEL>(F.fp, gptr(F.src), gptr(F.ref), gptr(F.gold), gptr(F.mvf), gptr(F.mvg), gptr(F.acosts), gptr(F.ainfos), g.aux[b.f], F.prune, b.x, b.y,
                            b.z);
}

hipError_t launch_analyze_b(hipStream_t st, int level, const BatchFrame *T, const BatchFrame *H, int n) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) {
    const int nslots = analysis_level_slots(H[f].fp, level);
    batch_grid_add(g, H[f].fp.sb_cols * H[f].fp.sb_rows * (1 << (2 * level)) * nslots, 1, 1, nslots);
  }
  const dim3 grid(g.start[n]);
  switch (level) {
    case 0: hipLaunchKernelGGL(k_analyze_level_b<0>, grid, dim3(256), 0, st, T, g); break;
    case 1: hipLaunchKernelGGL(k_analyze_level_b<1>, grid, dim3(128), 0, st, T, g); break;
    case 2: hipLaunchKernelGGL(k_analyze_level_b<2>, grid, dim3(64), 0, st, T, g); break;
    default: hipLaunchKernelGGL(k_analyze_level_b<3>, grid, dim3(64), 0, st, T, g); break;
  }
  return hipGetLastError();
}

__global__ __launch_bounds__(64) void k_analyze_table_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  analyze_table_body(F.fp, gptr(F.acosts), gptr(F.ainfos), gptr(F.blocks), gptr(F.sbs), b.x, b.y, b.z);
}

hipError_t launch_analyze_table_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) batch_grid_add(g, H[f].fp.sb_cols * H[f].fp.sb_rows, 1, 1);
  hipLaunchKernelGGL(k_analyze_table_b, dim3(g.start[n]), dim3(64), 0, st, T, g);
  return hipGetLastError();
}
