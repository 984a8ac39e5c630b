/* ORACLE — test infrastructure only. The in-loop filter stages of a coded
 * frame: per-mi filter info, deblocking (AV1 spec 7.14), CDEF (7.15) with
 * the per-superblock strength choice, and loop restoration (7.17) apply.
 * Serial C restatement of the mav1 reference's behaviour, each function
 * citing the reference lines it follows. */
#include "ocommon.h"

/* build_filter_info (src/encoder.c:513-526): per mi,
 * bsize | skip << 5 | tx_depth << 6 over the visible part of each block. */
void oracle_filter_info(const Mav1FrameParams *fp, const Mav1BlockInfo *blocks, const Mav1SbInfo *sbs, uint8_t *info) {
  const int nsb = fp->sb_cols * fp->sb_rows;
  for (int s = 0; s < nsb; s++)
    for (int k = 0; k < sbs[s].num_blocks; k++) {
      const Mav1BlockInfo *b = &blocks[sbs[s].first_block + k];
      const int skip = !b->eob[0] && !b->eob[1] && !b->eob[2];
      const uint8_t v = (uint8_t)(b->bsize | skip << 5 | b->tx_depth << 6);
      const int r1 = o_min(fp->mi_rows, b->mi_row + Num_4x4_Blocks_High[b->bsize]);
      const int c1 = o_min(fp->mi_cols, b->mi_col + Num_4x4_Blocks_Wide[b->bsize]);
      for (int r = b->mi_row; r < r1; r++)
        for (int c = b->mi_col; c < c1; c++) info[r * fp->mi_cols + c] = v;
    }
}

/* ---- deblocking ---- */

/* Transform size of an mi for the filter (src/kernels/k_lf.h:12-15): the
 * luma size follows from MiSize and tx_depth (one transform per block), the
 * chroma size is the largest for the subsampled block. */
static int lf_txsz(int info, int plane) {
  const int bs = info & 31;
  if (plane) {
    int t = Max_Tx_Size_Rect[SUBSAMPLED_SIZE(bs, 1, 1)];
    if (Tx_Width[t] == 64 || Tx_Height[t] == 64) t = Tx_Width[t] == 16 ? TX_16X32 : Tx_Height[t] == 16 ? TX_32X16 : TX_32X32;
    return t;
  }
  int t = Max_Tx_Size_Rect[bs];
  for (int d = (info >> 6) & 3; d > 0; d--) t = Split_Tx_Size[t];
  return t;
}

/* One sample position across an edge (spec 7.14.6; k_lf.h:17-104).
 * s points at q0, step crosses the edge. */
static void lf_filter_sample(uint8_t *s, int step, int plane, int lvl, int sharp, int size) {
  int p[7], q[7];
  for (int k = 0; k < 4; k++) {
    q[k] = s[k * step];
    p[k] = s[-(k + 1) * step];
  }
  /* 7.14.6.2 filter mask */
  const int shift = sharp > 4 ? 2 : sharp > 0 ? 1 : 0;
  const int limit = sharp > 0 ? o_clip(1, 9 - sharp, lvl >> shift) : o_max(1, lvl >> shift);
  const int blimit = 2 * (lvl + 2) + limit, thresh = lvl >> 4;
  const int len = size == 4 ? 4 : plane ? 6 : size == 8 ? 8 : 16;
  const int hev = o_abs(p[1] - p[0]) > thresh || o_abs(q[1] - q[0]) > thresh;
  int bad = o_abs(p[1] - p[0]) > limit || o_abs(q[1] - q[0]) > limit ||
            o_abs(p[0] - q[0]) * 2 + o_abs(p[1] - q[1]) / 2 > blimit;
  if (len >= 6) bad |= o_abs(p[2] - p[1]) > limit || o_abs(q[2] - q[1]) > limit;
  if (len >= 8) bad |= o_abs(p[3] - p[2]) > limit || o_abs(q[3] - q[2]) > limit;
  if (bad) return;
  int flat = 0, flat2 = 0;
  if (size >= 8) {
    int f = o_abs(p[1] - p[0]) <= 1 && o_abs(q[1] - q[0]) <= 1 && o_abs(p[2] - p[0]) <= 1 && o_abs(q[2] - q[0]) <= 1;
    if (len >= 8) f = f && o_abs(p[3] - p[0]) <= 1 && o_abs(q[3] - q[0]) <= 1;
    flat = f;
  }
  if (size >= 16) {
    for (int k = 4; k < 7; k++) {
      q[k] = s[k * step];
      p[k] = s[-(k + 1) * step];
    }
    flat2 = 1;
    for (int k = 4; k < 7; k++) flat2 = flat2 && o_abs(p[k] - p[0]) <= 1 && o_abs(q[k] - q[0]) <= 1;
  }
  if (size == 4 || !flat) {
    /* 7.14.6.3 narrow filter */
    const int ps1 = p[1] - 128, ps0 = p[0] - 128, qs0 = q[0] - 128, qs1 = q[1] - 128;
    int f = hev ? o_clip(-128, 127, ps1 - qs1) : 0;
    f = o_clip(-128, 127, f + 3 * (qs0 - ps0));
    const int f1 = o_clip(-128, 127, f + 4) >> 3, f2 = o_clip(-128, 127, f + 3) >> 3;
    s[0] = (uint8_t)(o_clip(-128, 127, qs0 - f1) + 128);
    s[-step] = (uint8_t)(o_clip(-128, 127, ps0 + f2) + 128);
    if (!hev) {
      const int f3 = o_round2(f1, 1);
      s[step] = (uint8_t)(o_clip(-128, 127, qs1 - f3) + 128);
      s[-2 * step] = (uint8_t)(o_clip(-128, 127, ps1 + f3) + 128);
    }
    return;
  }
  /* 7.14.6.4 wide filter */
  const int log2 = (size == 8 || !flat2) ? 3 : 4;
  const int n = log2 == 4 ? 6 : plane == 0 ? 3 : 2;
  const int n2 = (log2 == 3 && plane == 0) ? 0 : 1;
  int v[14], out[12];
  for (int k = -(n + 1); k <= n; k++) v[k + 7] = k < 0 ? p[-k - 1] : q[k];
  for (int i = -n; i < n; i++) {
    int t = 0;
    for (int j = -n; j <= n; j++) t += v[o_clip(-(n + 1), n, i + j) + 7] * (o_abs(j) <= n2 ? 2 : 1);
    out[i + n] = o_round2(t, log2);
  }
  for (int i = -n; i < n; i++) s[i * step] = (uint8_t)out[i + n];
}

/* Edge loop of one plane and pass (7.14.2; k_lf.h:106-136 with the work
 * split of src/backend_cpu.c:158-170): vertical edges (pass 0) at every mi
 * column of every mi row, then horizontal edges; chroma mi positions step
 * by two. An edge is filtered where it is a transform edge; its length is
 * the smaller transform dimension across it, capped at 16 (luma) / 8. */
static void lf_pass(const Mav1FrameParams *fp, uint8_t *rec, const uint8_t *info, int plane, int pass) {
  const int lvl = o_clip(0, MAX_LOOP_FILTER, fp->lf_level[plane == 0 ? pass : plane + 1]);
  if (!lvl) return;
  const int sub = plane > 0, S = fp->stride[plane], step = 1 << sub;
  uint8_t *base = rec + fp->offset[plane];
  /* edges of one pass touch disjoint samples (an edge filters at most half
   * the transform size on either side), so mi rows are independent */
#pragma omp parallel for schedule(static)
  for (int r0 = 0; r0 < fp->mi_rows; r0 += step)
    for (int c0 = 0; c0 < fp->mi_cols; c0 += step) {
      const int x = c0 * 4, y = r0 * 4;
      if (x >= fp->frame_w || y >= fp->frame_h) continue;
      if ((pass == 0 && x == 0) || (pass == 1 && y == 0)) continue;
      const int row = r0 | sub, col = c0 | sub, xp = x >> sub, yp = y >> sub;
      const int pr = pass == 1 ? row - step : row, pc = pass == 0 ? col - step : col;
      const int t = lf_txsz(info[row * fp->mi_cols + col], plane);
      const int tp = lf_txsz(info[pr * fp->mi_cols + pc], plane);
      const int onEdge = pass == 0 ? (xp % Tx_Width[t]) == 0 : (yp % Tx_Height[t]) == 0;
      if (!onEdge) continue;
      const int sz = pass == 0 ? o_min(Tx_Width[t], Tx_Width[tp]) : o_min(Tx_Height[t], Tx_Height[tp]);
      const int size = o_min(plane ? 8 : 16, sz);
      for (int i = 0; i < 4; i++) {
        uint8_t *s = pass == 0 ? base + (yp + i) * S + xp : base + yp * S + xp + i;
        lf_filter_sample(s, pass == 0 ? 1 : S, plane, lvl, fp->lf_sharpness, size);
      }
    }
}

void oracle_loop_filter(const Mav1FrameParams *fp, uint8_t *rec, const uint8_t *info) {
  for (int pass = 0; pass < 2; pass++)
    for (int plane = 0; plane < 3; plane++) lf_pass(fp, rec, info, plane, pass);
}

/* ---- CDEF ---- */

static int cdef_constrain(int diff, int thr, int damping) {
  if (!thr) return 0;
  const int adj = o_max(0, damping - o_floorlog2((unsigned)thr));
  const int m = o_clip(0, o_abs(diff), thr - (o_abs(diff) >> adj));
  return diff < 0 ? -m : m;
}

/* 7.15.2 cdef direction process over an 8x8 luma block (k_cdef.h:16-57). */
static void cdef_dir(const uint8_t *px, int S, int *dirOut, int *varOut) {
  int part[8][15];
  memset(part, 0, sizeof(part));
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) {
      const int x = (int)px[i * S + j] - 128;
      part[0][i + j] += x;
      part[1][i + j / 2] += x;
      part[2][i] += x;
      part[3][3 + i - j / 2] += x;
      part[4][7 + i - j] += x;
      part[5][3 - i / 2 + j] += x;
      part[6][j] += x;
      part[7][i / 2 + j] += x;
    }
  int cost[8] = {0};
  for (int i = 0; i < 8; i++) {
    cost[2] += part[2][i] * part[2][i];
    cost[6] += part[6][i] * part[6][i];
  }
  cost[2] *= Div_Table[8];
  cost[6] *= Div_Table[8];
  for (int i = 0; i < 7; i++) {
    cost[0] += (part[0][i] * part[0][i] + part[0][14 - i] * part[0][14 - i]) * Div_Table[i + 1];
    cost[4] += (part[4][i] * part[4][i] + part[4][14 - i] * part[4][14 - i]) * Div_Table[i + 1];
  }
  cost[0] += part[0][7] * part[0][7] * Div_Table[8];
  cost[4] += part[4][7] * part[4][7] * Div_Table[8];
  for (int d = 1; d < 8; d += 2) {
    for (int j = 3; j < 8; j++) cost[d] += part[d][j] * part[d][j];
    cost[d] *= Div_Table[8];
    for (int j = 0; j < 3; j++) cost[d] += (part[d][j] * part[d][j] + part[d][10 - j] * part[d][10 - j]) * Div_Table[2 * j + 2];
  }
  int best = 0, bestCost = 0;
  for (int d = 0; d < 8; d++)
    if (cost[d] > bestCost) {
      bestCost = cost[d];
      best = d;
    }
  *dirOut = best;
  *varOut = (bestCost - cost[(best + 4) & 7]) >> 10;
}

/* 7.15.3 cdef filter process for one plane of the 8x8 block at mi (r, c)
 * (k_cdef.h:59-95): taps outside the frame's mi area are skipped. */
static void cdef_plane(const Mav1FrameParams *fp, const uint8_t *cur, uint8_t *out, int plane, int r, int c, int pri,
                       int sec, int damping, int dir) {
  const int sub = plane > 0, S = fp->stride[plane], n = 8 >> sub;
  const uint8_t *in = cur + fp->offset[plane];
  uint8_t *dst = out + fp->offset[plane];
  const int x0 = (c * 4) >> sub, y0 = (r * 4) >> sub;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const int y = y0 + i, x = x0 + j, v = in[y * S + x];
      int sum = 0, hi = v, lo = v;
      for (int k = 0; k < 2; k++)
        for (int sg = -1; sg <= 1; sg += 2)
          for (int t = 0; t < 3; t++) {
            const int d = t == 0 ? dir : (dir + (t == 1 ? 6 : 2)) & 7;
            const int yy = y + sg * CDEF_DIRECTION(d, k, 0), xx = x + sg * CDEF_DIRECTION(d, k, 1);
            const int cr = (yy * (1 << sub)) >> 2, cc = (xx * (1 << sub)) >> 2;
            if (cr < 0 || cr >= fp->mi_rows || cc < 0 || cc >= fp->mi_cols) continue;
            const int pv = in[yy * S + xx];
            sum += t == 0 ? CDEF_PRI_TAP(pri & 1, k) * cdef_constrain(pv - v, pri, damping)
                          : CDEF_SEC_TAP(pri & 1, k) * cdef_constrain(pv - v, sec, damping);
            hi = o_max(hi, pv);
            lo = o_min(lo, pv);
          }
      dst[y * S + x] = (uint8_t)o_clip(lo, hi, v + ((8 + sum - (sum < 0)) >> 4));
    }
}

/* Every 8x8 block for parameter sets 0 .. n-1 into out + k * frameBytes
 * (cdef_block_all, k_cdef.h:124-150, over the block grid of
 * src/backend_cpu.c:171-176): skip blocks and all-zero sets are copies. */
void oracle_cdef(const Mav1FrameParams *fp, const uint8_t *cur, uint8_t *out, int64_t frameBytes, const uint8_t *info,
                 int n) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < fp->mi_rows; r += 2)
    for (int c = 0; c < fp->mi_cols; c += 2) {
      const int skip = (info[r * fp->mi_cols + c] >> 5) & 1;
      int dir = 0, var = 0;
      if (!skip) cdef_dir(cur + fp->offset[0] + r * 4 * fp->stride[0] + c * 4, fp->stride[0], &dir, &var);
      for (int k = 0; k < n; k++) {
        uint8_t *o = out + k * frameBytes;
        for (int p = 0; p < 3; p++) {
          const int sub = p > 0, S = fp->stride[p], m = 8 >> sub, x0 = (c * 4) >> sub, y0 = (r * 4) >> sub;
          for (int i = 0; i < m; i++)
            memcpy(o + fp->offset[p] + (y0 + i) * S + x0, cur + fp->offset[p] + (y0 + i) * S + x0, (size_t)m);
        }
        if (skip) continue;
        const int yp = fp->cdef_y_pri[k], ys = fp->cdef_y_sec[k], up = fp->cdef_uv_pri[k], us = fp->cdef_uv_sec[k];
        if (!yp && !ys && !up && !us) continue;
        const int vs = (var >> 6) ? o_min(o_floorlog2((unsigned)(var >> 6)), 12) : 0;
        const int ypAdj = var ? (yp * (4 + vs) + 8) >> 4 : 0;
        cdef_plane(fp, cur, o, 0, r, c, ypAdj, ys, fp->cdef_damping, yp ? dir : 0);
        const int uvDir = up ? CDEF_UV_DIR(1, 1, dir) : 0;
        cdef_plane(fp, cur, o, 1, r, c, up, us, fp->cdef_damping - 1, uvDir);
        cdef_plane(fp, cur, o, 2, r, c, up, us, fp->cdef_damping - 1, uvDir);
      }
    }
}

/* cdef_select (src/encoder.c:530-563): per superblock the parameter set
 * with the smallest SSE against the source over the visible area (first
 * minimum), copied into rec over the mi-aligned area. */
void oracle_cdef_select(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdefbuf, int64_t frameBytes,
                        int n, uint8_t *rec, int8_t *choice) {
#pragma omp parallel for schedule(static)
  for (int sr = 0; sr < fp->sb_rows; sr++)
    for (int sc = 0; sc < fp->sb_cols; sc++) {
      int64_t best = -1;
      int bestK = 0;
      for (int k = 0; k < n; k++) {
        const uint8_t *o = cdefbuf + k * frameBytes;
        int64_t sse = 0;
        for (int p = 0; p < 3; p++) {
          const int sub = p > 0, m = 64 >> sub, S = fp->stride[p];
          const int pw = p ? (fp->frame_w + 1) >> 1 : fp->frame_w, ph = p ? (fp->frame_h + 1) >> 1 : fp->frame_h;
          for (int y = sr * m; y < o_min((sr + 1) * m, ph); y++)
            for (int x = sc * m; x < o_min((sc + 1) * m, pw); x++) {
              const int d = (int)o[fp->offset[p] + y * S + x] - (int)src[fp->offset[p] + y * S + x];
              sse += d * d;
            }
        }
        if (best < 0 || sse < best) {
          best = sse;
          bestK = k;
        }
      }
      choice[sr * fp->sb_cols + sc] = (int8_t)bestK;
      const uint8_t *o = cdefbuf + bestK * frameBytes;
      for (int p = 0; p < 3; p++) {
        const int sub = p > 0, m = 64 >> sub, S = fp->stride[p];
        const int mw = (fp->mi_cols * 4) >> sub, mh = (fp->mi_rows * 4) >> sub;
        const int x0 = sc * m, x1 = o_min((sc + 1) * m, mw);
        for (int y = sr * m; y < o_min((sr + 1) * m, mh); y++)
          memcpy(rec + fp->offset[p] + y * S + x0, o + fp->offset[p] + y * S + x0, (size_t)(x1 - x0));
      }
    }
}
