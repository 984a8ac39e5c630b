/* ORACLE — test infrastructure only. Open-loop mode and partition
 * analysis restated from src/kernels/k_analysis.h (eval_block :241-404,
 * pass 1 analyze_item :656-679, pass 2 analyze_sb_table :697-808) and the
 * CPU driver src/backend_cpu.c:58-74. */
#include "ocommon.h"
#include "ointernal.h"

#define TURBO(fp, bit) (((fp)->turbo & (bit)) != 0)

/* Rate estimate (1/8 bit) of a luma intra mode, k_analysis.h:51-56. */
static int mode_bits8(int m) {
  if (m == DC_PRED) return 12;
  if (m == V_PRED || m == H_PRED || m == SMOOTH_PRED || m == PAETH_PRED) return 26;
  if (m == SMOOTH_V_PRED || m == SMOOTH_H_PRED) return 34;
  return 40;
}

/* k_analysis.h:146-148 */
static int mv_bits8(int r, int c) {
  return 8 * (2 + 2 * o_floorlog2((unsigned)o_abs(r) + 1) + 2 * o_floorlog2((unsigned)o_abs(c) + 1));
}

/* SATD of src against pred over a w x h block; `sub` keeps only the 4x4
 * units with even (row + column) unit parity and doubles the sum
 * (k_analysis.h:84-119). */
static int64_t blk_satd(const uint8_t *s, int ss, const uint8_t *p, int ps, int w, int h, int sub) {
  int64_t tot = 0;
  for (int uy = 0; uy < h; uy += 4)
    for (int ux = 0; ux < w; ux += 4) {
      if (sub && (((uy >> 2) + (ux >> 2)) & 1)) continue;
      int32_t d[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) d[4 * i + j] = (int)s[(uy + i) * ss + ux + j] - (int)p[(uy + i) * ps + ux + j];
      tot += oracle_satd4x4(d);
    }
  return sub ? tot * 2 : tot;
}

static int satd_sub(const Mav1FrameParams *fp, int w) { return TURBO(fp, MAV1_T_ANA_USUB) && w >= 32; }

/* Intra prediction from the source's own neighbours (no edge filter, no
 * above-right / below-left) and its SATD (k_analysis.h:126-144). */
static int64_t intra_satd(const Mav1FrameParams *fp, const uint8_t *src, int p, int x, int y, int lw, int lh, int mode,
                          int delta, uint8_t *pred) {
  int sub = p > 0;
  const uint8_t *pl = src + fp->offset[p];
  int S = fp->stride[p];
  OIntraArgs a = {x > 0, y > 0, 0, 0, mode, delta, lw, lh, ((fp->mi_cols * 4) >> sub) - 1,
                  ((fp->mi_rows * 4) >> sub) - 1, 0, 0};
  o_predict_intra(pl, S, x, y, &a, pred, 64);
  return blk_satd(pl + y * S + x, S, pred, 64, 1 << lw, 1 << lh, 0);
}

static int turbo_mode_off(const Mav1FrameParams *fp, int m) {
  return TURBO(fp, MAV1_T_INTRA) && !fp->frame_is_intra && m != DC_PRED && m != V_PRED && m != H_PRED &&
         m != SMOOTH_PRED;
}

static int cfl_allowed(int bs, int lossless) {
  if (lossless) return SUBSAMPLED_SIZE(bs, 1, 1) == BLOCK_4X4;
  return o_max(BLOCK_WIDTH(bs), BLOCK_HEIGHT(bs)) <= 32;
}

/* eval_block (k_analysis.h:241-404): the block's best intra / inter
 * decision and its cost. */
static int64_t eval_block(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref, const uint8_t *gold,
                          const int32_t *mvf, const int32_t *mvg, int mvNode, int miRow, int miCol, int bs,
                          Mav1BlockInfo *out) {
  int lw = Mi_Width_Log2[bs] + 2, lh = Mi_Height_Log2[bs] + 2, w = 1 << lw, h = 1 << lh;
  int x = miCol * 4, y = miRow * 4, lam = fp->ac_q[0] / 16;
  int cx = x >> 1, cy = y >> 1, clw = lw - 1, clh = lh - 1;
  uint8_t pa[64 * 64], pb[64 * 64], pc[64 * 64];
  int64_t interCost = -1;
  int mvr = 0, mvc = 0, refFrame = LAST_FRAME, refFrame1 = 0, mvr1 = 0, mvc1 = 0;
  if (!fp->frame_is_intra) {
    int sbIdx = (y >> 6) * fp->sb_cols + (x >> 6);
    const int32_t *m0 = mvf + (sbIdx * MAV1_NUM_NODES + mvNode) * 2;
    int lr = m0[0], lc = m0[1], gr = 0, gc = 0;
    if (fp->golden) {
      const int32_t *m1 = mvg + (sbIdx * MAV1_NUM_NODES + mvNode) * 2;
      gr = m1[0];
      gc = m1[1];
    }
    /* per plane: LAST, second reference and the rounded average (the
     * compound prediction) against the source */
    int64_t s[3] = {0, 0, 0};
    for (int p = 0; p < 3; p++) {
      int sub = p > 0, bx = x >> sub, by = y >> sub, bw = w >> sub, bh = h >> sub, S = fp->stride[p];
      const uint8_t *sp = src + fp->offset[p] + by * S + bx;
      int usub = satd_sub(fp, bw);
      oracle_predict_inter(ref + fp->offset[p], S, fp->frame_w, fp->frame_h, pa, 64, bx, by, bw, bh, lr, lc, sub, sub,
                           EIGHTTAP);
      s[0] += blk_satd(sp, S, pa, 64, bw, bh, usub);
      if (fp->golden) {
        oracle_predict_inter(gold + fp->offset[p], S, fp->frame_w, fp->frame_h, pb, 64, bx, by, bw, bh, gr, gc, sub,
                             sub, EIGHTTAP);
        s[1] += blk_satd(sp, S, pb, 64, bw, bh, usub);
        if (fp->compound) {
          for (int i = 0; i < bh; i++)
            for (int j = 0; j < bw; j++) pc[i * 64 + j] = (uint8_t)((pa[i * 64 + j] + pb[i * 64 + j] + 1) >> 1);
          s[2] += blk_satd(sp, S, pc, 64, bw, bh, usub);
        }
      }
    }
    interCost = s[0] + (int64_t)lam * (4 + mv_bits8(lr, lc)) / 8;
    mvr = lr;
    mvc = lc;
    if (fp->golden) {
      int64_t g = s[1] + (int64_t)lam * (4 + mv_bits8(gr, gc)) / 8 + (int64_t)lam * 8 / 8;
      if (g < interCost) { /* ties stay on LAST */
        interCost = g;
        mvr = gr;
        mvc = gc;
        refFrame = fp->ref2_frame;
      }
      if (fp->compound) {
        int64_t cc = s[2] + (int64_t)lam * (8 + mv_bits8(lr, lc) + mv_bits8(gr, gc) + 16) / 8;
        if (cc < interCost) {
          interCost = cc;
          mvr = lr;
          mvc = lc;
          mvr1 = gr;
          mvc1 = gc;
          refFrame = LAST_FRAME;
          refFrame1 = fp->ref2_frame;
        }
      }
    }
  }
  /* luma intra */
  int64_t best = -1;
  int bestMode = DC_PRED, bestDelta = 0;
  int skipIntra = interCost >= 0 && (fp->leaf || (fp->fast && interCost * 16 < (int64_t)w * h * lam * 16));
  for (int m = DC_PRED; m <= PAETH_PRED && !skipIntra; m++) {
    if (turbo_mode_off(fp, m)) continue;
    int64_t c = intra_satd(fp, src, 0, x, y, lw, lh, m, 0, pa) + (int64_t)lam * mode_bits8(m) / 8;
    if (best < 0 || c < best) {
      best = c;
      bestMode = m;
    }
  }
  if (skipIntra) best = (int64_t)1 << 40;
  if (bestMode >= V_PRED && bestMode <= D67_PRED && !fp->fast)
    for (int d = -MAX_ANGLE_DELTA; d <= MAX_ANGLE_DELTA; d++) {
      if (!d) continue;
      int64_t c = intra_satd(fp, src, 0, x, y, lw, lh, bestMode, d, pa) + (int64_t)lam * (mode_bits8(bestMode) + 16) / 8;
      if (c < best) {
        best = c;
        bestDelta = d;
      }
    }
  /* chroma intra */
  int skipUv = fp->fast && interCost >= 0 && best + (int64_t)lam * 16 / 8 >= interCost;
  int64_t bestUv = skipUv ? 0 : -1;
  int bestUvMode = DC_PRED;
  for (int m = DC_PRED; m <= PAETH_PRED && !skipUv; m++) {
    if (m >= V_PRED && m <= D67_PRED && m != V_PRED && m != H_PRED && m != bestMode) continue;
    if (turbo_mode_off(fp, m)) continue;
    int d = m == bestMode ? bestDelta : 0;
    int64_t c = intra_satd(fp, src, 1, cx, cy, clw, clh, m, d, pa) + intra_satd(fp, src, 2, cx, cy, clw, clh, m, d, pa) +
                (int64_t)lam * (m == DC_PRED ? 8 : 24) / 8;
    if (bestUv < 0 || c < bestUv) {
      bestUv = c;
      bestUvMode = m;
    }
  }
  if (!skipUv && cfl_allowed(bs, fp->lossless) && !(TURBO(fp, MAV1_T_INTRA) && !fp->frame_is_intra)) {
    int cw = 1 << clw, ch = 1 << clh;
    int64_t c = (int64_t)lam * 40 / 8;
    int dc[32 * 32];
    for (int p = 1; p < 3; p++) {
      intra_satd(fp, src, p, cx, cy, clw, clh, DC_PRED, 0, pa);
      for (int i = 0; i < ch; i++)
        for (int j = 0; j < cw; j++) dc[i * cw + j] = pa[i * 64 + j];
      const uint8_t *sp = src + fp->offset[p] + cy * fp->stride[p] + cx;
      int al = o_cfl_best_alpha(src + fp->offset[0], fp->stride[0], cx, cy, x + w, y + h, clw, clh, dc, sp,
                                fp->stride[p], 0);
      o_cfl_predict(src + fp->offset[0], fp->stride[0], cx, cy, x + w, y + h, clw, clh, al, dc, pa, 64);
      c += blk_satd(sp, fp->stride[p], pa, 64, cw, ch, 0);
    }
    if (c < bestUv) {
      bestUv = c;
      bestUvMode = UV_CFL_PRED;
    }
  }
  memset(out, 0, sizeof *out);
  out->mi_row = miRow;
  out->mi_col = miCol;
  out->bsize = bs;
  out->y_mode = bestMode;
  out->angle_y = bestDelta;
  out->uv_mode = bestUvMode;
  out->angle_uv = bestUvMode == bestMode ? bestDelta : 0;
  out->ref_frame = INTRA_FRAME;
  out->tx_type = DCT_DCT;
  int64_t cost = best + bestUv + (int64_t)lam * (fp->frame_is_intra ? 0 : 16) / 8;
  if (interCost >= 0 && interCost < cost) {
    cost = interCost;
    out->is_inter = 1;
    out->y_mode = NEWMV;
    out->uv_mode = DC_PRED;
    out->angle_y = out->angle_uv = 0;
    out->mv_row = mvr;
    out->mv_col = mvc;
    out->ref_frame = refFrame;
    out->ref_frame1 = refFrame1;
    out->mv1_row = mvr1;
    out->mv1_col = mvc1;
  }
  return cost;
}

/* k_analysis.h:408-432 */
int o_part_block(int p, int bSize, int k, int halfMi, int *dr, int *dc) {
  int sub = PARTITION_SUBSIZE(p, bSize), split = PARTITION_SUBSIZE(PARTITION_SPLIT, bSize), q = halfMi >> 1;
  *dr = *dc = 0;
  switch (p) {
    case PARTITION_HORZ: *dr = k * halfMi; return sub;
    case PARTITION_VERT: *dc = k * halfMi; return sub;
    case PARTITION_HORZ_A:
      if (k < 2) { *dc = k * halfMi; return split; }
      *dr = halfMi;
      return sub;
    case PARTITION_HORZ_B:
      if (k == 0) return sub;
      *dr = halfMi;
      *dc = (k - 1) * halfMi;
      return split;
    case PARTITION_VERT_A:
      if (k < 2) { *dr = k * halfMi; return split; }
      *dc = halfMi;
      return sub;
    case PARTITION_VERT_B:
      if (k == 0) return sub;
      *dc = halfMi;
      *dr = (k - 1) * halfMi;
      return split;
    case PARTITION_HORZ_4: *dr = k * q; return sub;
    case PARTITION_VERT_4: *dc = k * q; return sub;
    default: return sub;
  }
}

/* k_analysis.h:435-443 */
int o_part_count(const Mav1FrameParams *fp, int p, int r, int c, int halfMi) {
  int hasRows = r + halfMi < fp->mi_rows, hasCols = c + halfMi < fp->mi_cols;
  switch (p) {
    case PARTITION_NONE: return 1;
    case PARTITION_HORZ: return hasRows ? 2 : 1;
    case PARTITION_VERT: return hasCols ? 2 : 1;
    case PARTITION_HORZ_4: return r + (halfMi >> 1) * 3 < fp->mi_rows ? 4 : 3;
    case PARTITION_VERT_4: return c + (halfMi >> 1) * 3 < fp->mi_cols ? 4 : 3;
    default: return 3;
  }
}

static int turbo_square(const Mav1FrameParams *fp) {
  return TURBO(fp, MAV1_T_SQUARE) && (!fp->frame_is_intra || TURBO(fp, MAV1_T_KEY_SQUARE));
}

/* k_analysis.h:615-631 */
static int part_allowed(const Mav1FrameParams *fp, int level, int p, int hasRows, int hasCols) {
  if (p == PARTITION_SPLIT) return 0;
  if (level == 3 && p != PARTITION_NONE) return 0;
  if (fp->fast && p > PARTITION_SPLIT) return 0;
  if (turbo_square(fp) && p != PARTITION_NONE) return 0;
  if (level == 2 && (p == PARTITION_HORZ_4 || p == PARTITION_VERT_4)) return 0;
  if (!(hasRows && hasCols)) return (p == PARTITION_HORZ && hasCols && !hasRows) || (p == PARTITION_VERT && hasRows && !hasCols);
  return 1;
}

static int level_slots(const Mav1FrameParams *fp, int level) {
  if (level == 3 || turbo_square(fp)) return 1;
  if (level == 2 || fp->fast) return 5;
  return MAV1_A_SLOTS;
}

/* slot -> (partition, block index), k_analysis.h:590-596 */
static void slot_part(int slot, int *p, int *k) {
  static const int P[13] = {0, 1, 1, 2, 2, 8, 8, 8, 8, 9, 9, 9, 9}, K[13] = {0, 0, 1, 0, 1, 0, 1, 2, 3, 0, 1, 2, 3};
  *p = P[slot];
  *k = K[slot];
}

/* k_analysis.h:682-694: table slot of block k of partition p (-1: the child
 * square's NONE entry). */
static int part_slot(int p, int k) {
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

static int level_bsize(int l) { return l == 0 ? BLOCK_64X64 : l == 1 ? BLOCK_32X32 : l == 2 ? BLOCK_16X16 : BLOCK_8X8; }

/* k_analysis.h:642-654: with the fast preset a node is skipped when any
 * ancestor's NONE cost is below depth_thr / 16 per sample. */
static int depth_pruned(const Mav1FrameParams *fp, const int64_t *costs, int sb, int level, int i) {
  if (!fp->fast || level == 0 || fp->depth_thr == 0) return 0;
  int y = i >> level, x = i & ((1 << level) - 1);
  for (int a = level - 1; a >= 0; a--) {
    y >>= 1;
    x >>= 1;
    int64_t pc = costs[(int64_t)sb * MAV1_A_ITEMS + (o_node_base(a) + (y << a) + x) * MAV1_A_SLOTS];
    int pw = 64 >> a;
    if (pc >= 0 && pc * 16 < (int64_t)pw * pw * fp->depth_thr) return 1;
  }
  return 0;
}

static void analyze_item(const Mav1FrameParams *fp, int sb, int level, int i, int slot, const uint8_t *src,
                         const uint8_t *ref, const uint8_t *gold, const int32_t *mvf, const int32_t *mvg,
                         int64_t *costs, Mav1BlockInfo *infos) {
  int n = 1 << level, size4 = 16 >> level, half = size4 >> 1, node = o_node_base(level) + i;
  int r = (sb / fp->sb_cols) * 16 + (i / n) * size4, c = (sb % fp->sb_cols) * 16 + (i % n) * size4;
  int64_t idx = (int64_t)sb * MAV1_A_ITEMS + node * MAV1_A_SLOTS + slot;
  costs[idx] = -1;
  if (r >= fp->mi_rows || c >= fp->mi_cols) return;
  if (depth_pruned(fp, costs, sb, level, i)) return;
  if (level == 3 && TURBO(fp, MAV1_T_LEAF_NO8) && fp->leaf && !fp->frame_is_intra) {
    /* the 8x8 node is coded only where its 16x16 parent crosses the frame edge */
    int pr = (sb / fp->sb_cols) * 16 + ((i >> 3) & ~1) * 2, pc = (sb % fp->sb_cols) * 16 + ((i & 7) & ~1) * 2;
    if (!(pr + 4 > fp->mi_rows || pc + 4 > fp->mi_cols)) return;
  }
  int hasRows = r + half < fp->mi_rows, hasCols = c + half < fp->mi_cols, p, k, dr, dc;
  slot_part(slot, &p, &k);
  if (!part_allowed(fp, level, p, hasRows, hasCols)) return;
  if (k >= o_part_count(fp, p, r, c, half)) return;
  int bs = o_part_block(p, level_bsize(level), k, half, &dr, &dc);
  costs[idx] = eval_block(fp, src, ref, gold, mvf, mvg, node, r + dr, c + dc, bs, &infos[idx]);
}

/* Pass 2 (k_analysis.h:697-808): bottom-up partition choice from the
 * table, then the blocks in decode (Z) order with coefficient offsets. */
static void sb_table(const Mav1FrameParams *fp, int sb, const int64_t *costs, const Mav1BlockInfo *infos,
                     Mav1BlockInfo *blocks, Mav1SbInfo *sbs) {
  int sbRow = sb / fp->sb_cols, sbCol = sb % fp->sb_cols, lam = fp->ac_q[0] / 16;
  int64_t cost[MAV1_NUM_NODES], noneCost[MAV1_NUM_NODES];
  Mav1SbInfo *S = &sbs[sb];
  const int64_t *tc = costs + (int64_t)sb * MAV1_A_ITEMS;
  const Mav1BlockInfo *ti = infos + (int64_t)sb * MAV1_A_ITEMS;
  for (int level = 3; level >= 0; level--) {
    int n = 1 << level, size4 = 16 >> level, bSize = level_bsize(level);
    for (int i = 0; i < n * n; i++) {
      int node = o_node_base(level) + i;
      int r = sbRow * 16 + (i / n) * size4, c = sbCol * 16 + (i % n) * size4;
      noneCost[node] = -1;
      if (r >= fp->mi_rows || c >= fp->mi_cols) {
        cost[node] = 0;
        S->part[node] = -1;
        continue;
      }
      int half = size4 >> 1, hasRows = r + half < fp->mi_rows, hasCols = c + half < fp->mi_cols;
      int cn = 2 * n, ci = (i / n) * 2 * cn + (i % n) * 2, cb = level < 3 ? o_node_base(level + 1) : 0;
      int64_t bestC = -1;
      int bestP = PARTITION_SPLIT;
      if (level < 3) {
        int64_t c0 = cost[cb + ci], c1 = cost[cb + ci + 1], c2 = cost[cb + ci + cn], c3 = cost[cb + ci + cn + 1];
        if (c0 >= 0 && c1 >= 0 && c2 >= 0 && c3 >= 0) bestC = c0 + c1 + c2 + c3 + (int64_t)lam * 16 / 8;
      }
      for (int p = PARTITION_NONE; p <= PARTITION_VERT_4; p++) {
        if (!part_allowed(fp, level, p, hasRows, hasCols)) continue;
        int64_t pc = 0;
        int nb = o_part_count(fp, p, r, c, half);
        for (int k = 0; k < nb && pc >= 0; k++) {
          int s = part_slot(p, k), dr, dc;
          int64_t cc;
          if (s < 0) {
            o_part_block(p, bSize, k, half, &dr, &dc);
            cc = noneCost[cb + ci + (dr ? cn : 0) + (dc ? 1 : 0)];
          } else {
            cc = tc[node * MAV1_A_SLOTS + s];
          }
          pc = cc < 0 ? -1 : pc + cc;
        }
        if (pc < 0) continue;
        pc += (int64_t)lam * (p == PARTITION_NONE ? 8 : 24) / 8;
        if (p == PARTITION_NONE) noneCost[node] = pc;
        if (bestC < 0 || pc <= bestC) { /* ties: the later partition */
          bestC = pc;
          bestP = p;
        }
      }
      cost[node] = bestC;
      S->part[node] = bestP;
    }
  }
  int count = 0, coefOff = sb * MAV1_COEFS_PER_SB;
  for (int z = 0; z < 64; z++) {
    int zx = 0, zy = 0;
    for (int b = 0; b < 3; b++) {
      zx |= ((z >> (2 * b)) & 1) << b;
      zy |= ((z >> (2 * b + 1)) & 1) << b;
    }
    for (int level = 0; level <= 3; level++) {
      int shift = 3 - level, n = 1 << level;
      int node = o_node_base(level) + (zy >> shift) * n + (zx >> shift);
      int p = S->part[node];
      if (p < 0) break;
      if (p == PARTITION_SPLIT) continue;
      if ((zx & ((1 << shift) - 1)) == 0 && (zy & ((1 << shift) - 1)) == 0) {
        int size4 = 16 >> level, half = size4 >> 1;
        int r = sbRow * 16 + (zy >> shift) * size4, c = sbCol * 16 + (zx >> shift) * size4;
        int nb = o_part_count(fp, p, r, c, half);
        int cn = 2 * n, ci = (zy >> shift) * 2 * cn + (zx >> shift) * 2, cb = level < 3 ? o_node_base(level + 1) : 0;
        for (int k = 0; k < nb; k++) {
          int s = part_slot(p, k), dr, dc;
          Mav1BlockInfo b;
          if (s < 0) {
            o_part_block(p, level_bsize(level), k, half, &dr, &dc);
            b = ti[(cb + ci + (dr ? cn : 0) + (dc ? 1 : 0)) * MAV1_A_SLOTS];
          } else {
            b = ti[node * MAV1_A_SLOTS + s];
          }
          int w = BLOCK_WIDTH(b.bsize), h = BLOCK_HEIGHT(b.bsize);
          b.coef_off[0] = coefOff;
          coefOff += w * h;
          b.coef_off[1] = coefOff;
          coefOff += (w >> 1) * (h >> 1);
          b.coef_off[2] = coefOff;
          coefOff += (w >> 1) * (h >> 1);
          b.eob[0] = b.eob[1] = b.eob[2] = 0;
          blocks[sb * MAV1_MAX_BLOCKS_PER_SB + count++] = b;
        }
      }
      break;
    }
  }
  S->first_block = sb * MAV1_MAX_BLOCKS_PER_SB;
  S->num_blocks = count;
  S->sb_row = sbRow;
  S->sb_col = sbCol;
}

void oracle_analyze(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref, const uint8_t *gold,
                    const int32_t *mvf, const int32_t *mvg, int64_t *costs, Mav1BlockInfo *infos,
                    Mav1BlockInfo *blocks, Mav1SbInfo *sbs) {
  int nsb = fp->sb_cols * fp->sb_rows;
  /* a superblock's items and table read only its own entries */
#pragma omp parallel for schedule(dynamic, 2)
  for (int sb = 0; sb < nsb; sb++) {
    for (int level = 0; level < 4; level++) {
      int ns = level_slots(fp, level);
      for (int i = 0; i < (1 << (2 * level)); i++)
        for (int s = 0; s < ns; s++) analyze_item(fp, sb, level, i, s, src, ref, gold, mvf, mvg, costs, infos);
    }
    sb_table(fp, sb, costs, infos, blocks, sbs);
  }
}
