/* In-loop filters of a coded frame on gfx950: per-mi filter info
 * (behaviour of src/encoder.c:513-526), deblocking (AV1 spec 7.14;
 * src/kernels/k_lf.h), CDEF (7.15; src/kernels/k_cdef.h) with the
 * per-superblock strength choice (src/encoder.c:530-563) and loop
 * restoration (7.17; src/kernels/k_lr.h).
 *
 * Deblocking: within one pass the edges never share a sample that one of
 * them writes and another one uses (a filter is no longer than the
 * transforms on either side), so every (edge, sample line) is one thread;
 * consecutive threads walk along a sample row (vertical edges) or along the
 * samples of a horizontal edge, so their byte loads coalesce. CDEF: the
 * direction of each 8x8 block once, then one thread per output sample
 * filters it for every strength set. */
#include "dev_common.h"
#include "internal.h"

/* ---- filter info ---- */

DEVI void filter_info_body(const Mav1FrameParams &fp, const Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo *__restrict__ sbs, uint8_t *__restrict__ info, const int blk_x, const int blk_y, const int blk_z) {
  const int sb = blk_x, lane = threadIdx.x;
  if (lane >= sbs[sb].num_blocks) return;
  const Mav1BlockInfo &b = blocks[sbs[sb].first_block + lane];
  const int skip = !b.eob[0] && !b.eob[1] && !b.eob[2];
  const uint8_t v = (uint8_t)(b.bsize | skip << 5 | b.tx_depth << 6);
  const int r1 = d_min(fp.mi_rows, b.mi_row + Num_4x4_Blocks_High[b.bsize]);
  const int c1 = d_min(fp.mi_cols, b.mi_col + Num_4x4_Blocks_Wide[b.bsize]);
  for (int r = b.mi_row; r < r1; r++)
    for (int c = b.mi_col; c < c1; c++) info[r * fp.mi_cols + c] = v;
}
__global__ __launch_bounds__(64) void k_filter_info(const Mav1FrameParams fp, const Mav1BlockInfo *__restrict__ blocks,
                                                    const Mav1SbInfo *__restrict__ sbs, uint8_t *__restrict__ info) {
  filter_info_body(fp, blocks, sbs, info, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* ---- deblocking ---- */

DEVI int lf_txsz(int info, int plane) {
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

/* 7.14.6 at one sample position: s points at q0, step crosses the edge.
 * Only the samples the filter length uses are read. */
DEVI void lf_sample(uint8_t *s, int step, int plane, int lvl, int sharp, int size) {
  const int len = size == 4 ? 4 : plane ? 6 : size == 8 ? 8 : 16;
  int p[7], q[7];
  const int nr = len == 4 ? 2 : len == 6 ? 3 : len == 8 ? 4 : 7;
#pragma unroll
  for (int k = 0; k < 7; k++)
    if (k < nr) {
      q[k] = s[k * step];
      p[k] = s[-(k + 1) * step];
    }
  const int shift = sharp > 4 ? 2 : sharp > 0 ? 1 : 0;
  const int limit = sharp > 0 ? d_clip(1, 9 - sharp, lvl >> shift) : d_max(1, lvl >> shift);
  const int blimit = 2 * (lvl + 2) + limit, thresh = lvl >> 4;
  const int hev = d_abs(p[1] - p[0]) > thresh || d_abs(q[1] - q[0]) > thresh;
  int bad = d_abs(p[1] - p[0]) > limit || d_abs(q[1] - q[0]) > limit ||
            d_abs(p[0] - q[0]) * 2 + d_abs(p[1] - q[1]) / 2 > blimit;
  if (len >= 6) bad |= d_abs(p[2] - p[1]) > limit || d_abs(q[2] - q[1]) > limit;
  if (len >= 8) bad |= d_abs(p[3] - p[2]) > limit || d_abs(q[3] - q[2]) > limit;
  if (bad) return;
  int flat = 0, flat2 = 0;
  if (size >= 8) {
    int f = d_abs(p[1] - p[0]) <= 1 && d_abs(q[1] - q[0]) <= 1 && d_abs(p[2] - p[0]) <= 1 && d_abs(q[2] - q[0]) <= 1;
    if (len >= 8) f = f && d_abs(p[3] - p[0]) <= 1 && d_abs(q[3] - q[0]) <= 1;
    flat = f;
  }
  if (size >= 16) {
    flat2 = 1;
#pragma unroll
    for (int k = 4; k < 7; k++) flat2 = flat2 && d_abs(p[k] - p[0]) <= 1 && d_abs(q[k] - q[0]) <= 1;
  }
  if (size == 4 || !flat) {
    const int ps1 = p[1] - 128, ps0 = p[0] - 128, qs0 = q[0] - 128, qs1 = q[1] - 128;
    int f = hev ? d_clip(-128, 127, ps1 - qs1) : 0;
    f = d_clip(-128, 127, f + 3 * (qs0 - ps0));
    const int f1 = d_clip(-128, 127, f + 4) >> 3, f2 = d_clip(-128, 127, f + 3) >> 3;
    s[0] = (uint8_t)(d_clip(-128, 127, qs0 - f1) + 128);
    s[-step] = (uint8_t)(d_clip(-128, 127, ps0 + f2) + 128);
    if (!hev) {
      const int f3 = d_round2(f1, 1);
      s[step] = (uint8_t)(d_clip(-128, 127, qs1 - f3) + 128);
      s[-2 * step] = (uint8_t)(d_clip(-128, 127, ps1 + f3) + 128);
    }
    return;
  }
  const int log2 = (size == 8 || !flat2) ? 3 : 4;
  const int n = log2 == 4 ? 6 : plane == 0 ? 3 : 2;
  const int n2 = (log2 == 3 && plane == 0) ? 0 : 1;
  int v[14], out[12];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    v[6 - k] = p[k];
    v[7 + k] = q[k];
  }
  for (int i = -n; i < n; i++) {
    int t = 0;
    for (int j = -n; j <= n; j++) t += v[d_clip(-(n + 1), n, i + j) + 7] * (d_abs(j) <= n2 ? 2 : 1);
    out[i + n] = d_round2(t, log2);
  }
  for (int i = -n; i < n; i++) s[i * step] = (uint8_t)out[i + n];
}

/* pass 0: x = vertical edge unit (4 samples), y = sample row;
 * pass 1: x = sample column, y = horizontal edge unit; z = plane. */
DEVI void deblock_body(const Mav1FrameParams &fp, uint8_t *rec, const uint8_t *__restrict__ info, int pass, const int blk_x, const int blk_y, const int blk_z) {
  const int plane = blk_z, sub = plane > 0;
  const int lvl = d_clip(0, MAX_LOOP_FILTER, fp.lf_level[plane == 0 ? pass : plane + 1]);
  if (!lvl) return;
  const int gx = blk_x * 64 + threadIdx.x, gy = blk_y;
  int c0, r0, sx, sy;
  if (pass == 0) {
    c0 = gx << sub;
    r0 = (gy >> 2) << sub;
    sx = gx * 4;
    sy = gy;
  } else {
    c0 = (gx >> 2) << sub;
    r0 = gy << sub;
    sx = gx;
    sy = gy * 4;
  }
  if (c0 >= fp.mi_cols || r0 >= fp.mi_rows) return;
  const int x = c0 * 4, y = r0 * 4;
  if (x >= fp.frame_w || y >= fp.frame_h) return;
  if (pass == 0 ? x == 0 : y == 0) return;
  const int row = r0 | sub, col = c0 | sub, step = 1 << sub;
  const int pr = pass == 1 ? row - step : row, pc = pass == 0 ? col - step : col;
  const int t = lf_txsz(info[row * fp.mi_cols + col], plane), tp = lf_txsz(info[pr * fp.mi_cols + pc], plane);
  const int onEdge = pass == 0 ? ((x >> sub) & (Tx_Width[t] - 1)) == 0 : ((y >> sub) & (Tx_Height[t] - 1)) == 0;
  if (!onEdge) return;
  const int sz = pass == 0 ? d_min(Tx_Width[t], Tx_Width[tp]) : d_min(Tx_Height[t], Tx_Height[tp]);
  const int S = fp.stride[plane];
  lf_sample(rec + fp.offset[plane] + sy * S + sx, pass == 0 ? 1 : S, plane, lvl, fp.lf_sharpness,
            d_min(plane ? 8 : 16, sz));
}
__global__ __launch_bounds__(64) void k_deblock(const Mav1FrameParams fp, uint8_t *rec, const uint8_t *__restrict__ info,
                                                int pass) {
  deblock_body(fp, rec, info, pass, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* ---- CDEF ---- */

/* Direction and variance of every 8x8 block (7.15.2), one thread each;
 * dirvar = -1 for skip blocks, else dir | var << 3. */
DEVI void cdef_dir_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ cur, const uint8_t *__restrict__ info, int *__restrict__ dirvar, const int blk_x, const int blk_y, const int blk_z) {
  const int bc = fp.mi_cols >> 1, br = fp.mi_rows >> 1;
  const int bi = blk_x * 64 + threadIdx.x;
  if (bi >= bc * br) return;
  const int r = (bi / bc) * 2, c = (bi % bc) * 2;
  if ((info[r * fp.mi_cols + c] >> 5) & 1) {
    dirvar[bi] = -1;
    return;
  }
  const int S = fp.stride[0];
  const uint8_t *px = cur + fp.offset[0] + r * 4 * S + c * 4;
  int part[8][15];
#pragma unroll
  for (int d = 0; d < 8; d++)
#pragma unroll
    for (int k = 0; k < 15; k++) part[d][k] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint2 row = *(const uint2 *)(px + i * S);
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int x = (int)(((j < 4 ? row.x : row.y) >> (8 * (j & 3))) & 255) - 128;
      part[0][i + j] += x;
      part[1][i + j / 2] += x;
      part[2][i] += x;
      part[3][3 + i - j / 2] += x;
      part[4][7 + i - j] += x;
      part[5][3 - i / 2 + j] += x;
      part[6][j] += x;
      part[7][i / 2 + j] += x;
    }
  }
  int cost[8];
#pragma unroll
  for (int d = 0; d < 8; d++) cost[d] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    cost[2] += part[2][i] * part[2][i];
    cost[6] += part[6][i] * part[6][i];
  }
  cost[2] *= Div_Table[8];
  cost[6] *= Div_Table[8];
#pragma unroll
  for (int i = 0; i < 7; i++) {
    cost[0] += (part[0][i] * part[0][i] + part[0][14 - i] * part[0][14 - i]) * Div_Table[i + 1];
    cost[4] += (part[4][i] * part[4][i] + part[4][14 - i] * part[4][14 - i]) * Div_Table[i + 1];
  }
  cost[0] += part[0][7] * part[0][7] * Div_Table[8];
  cost[4] += part[4][7] * part[4][7] * Div_Table[8];
#pragma unroll
  for (int d = 1; d < 8; d += 2) {
#pragma unroll
    for (int j = 3; j < 8; j++) cost[d] += part[d][j] * part[d][j];
    cost[d] *= Div_Table[8];
#pragma unroll
    for (int j = 0; j < 3; j++) cost[d] += (part[d][j] * part[d][j] + part[d][10 - j] * part[d][10 - j]) * Div_Table[2 * j + 2];
  }
  int best = 0, bestCost = 0;
#pragma unroll
  for (int d = 0; d < 8; d++)
    if (cost[d] > bestCost) {
      bestCost = cost[d];
      best = d;
    }
  int opp = 0;
#pragma unroll
  for (int d = 0; d < 8; d++)
    if (d == ((best + 4) & 7)) opp = cost[d];
  dirvar[bi] = best | ((bestCost - opp) >> 10) << 3;
}
__global__ __launch_bounds__(64) void k_cdef_dir(const Mav1FrameParams fp, const uint8_t *__restrict__ cur,
                                                 const uint8_t *__restrict__ info, int *__restrict__ dirvar) {
  cdef_dir_body(fp, cur, info, dirvar, blockIdx.x, blockIdx.y, blockIdx.z);
}

DEVI int cdef_constrain(int diff, int thr, int damping) {
  if (!thr) return 0;
  const int adj = d_max(0, damping - d_floorlog2((unsigned)thr));
  const int m = d_clip(0, d_abs(diff), thr - (d_abs(diff) >> adj));
  return diff < 0 ? -m : m;
}

/* One output sample per thread for every strength set (7.15.3 with the
 * preset handling of src/kernels/k_cdef.h:124-150): x = sample column,
 * y = sample row, z = plane, over the mi-covered area. */
DEVI void cdef_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ cur, uint8_t *__restrict__ out, long long frameBytes, const int *__restrict__ dirvar, int n, const int blk_x, const int blk_y, const int blk_z) {
  const int plane = blk_z, sub = plane > 0;
  const int x = blk_x * 64 + threadIdx.x, y = blk_y;
  const int W = (fp.mi_cols * 4) >> sub, H = (fp.mi_rows * 4) >> sub;
  if (x >= W || y >= H) return;
  const int S = fp.stride[plane], o = fp.offset[plane] + y * S + x;
  const int bsh = 3 - sub, bc = fp.mi_cols >> 1;
  const int dv = dirvar[(y >> bsh) * bc + (x >> bsh)];
  const uint8_t *in = cur + fp.offset[plane];
  const int v = in[y * S + x];
  for (int k = 0; k < n; k++) {
    uint8_t *dst = out + k * frameBytes;
    const int yp = fp.cdef_y_pri[k], ys = fp.cdef_y_sec[k], up = fp.cdef_uv_pri[k], us = fp.cdef_uv_sec[k];
    if (dv < 0 || (!yp && !ys && !up && !us)) {
      dst[o] = (uint8_t)v;
      continue;
    }
    const int dir = dv & 7, var = dv >> 3;
    int pri, sec, damping, fdir;
    if (plane == 0) {
      const int vs = (var >> 6) ? d_min(d_floorlog2((unsigned)(var >> 6)), 12) : 0;
      pri = var ? (yp * (4 + vs) + 8) >> 4 : 0;
      sec = ys;
      damping = fp.cdef_damping;
      fdir = yp ? dir : 0;
    } else {
      pri = up;
      sec = us;
      damping = fp.cdef_damping - 1;
      fdir = up ? CDEF_UV_DIR(1, 1, dir) : 0;
    }
    int sum = 0, hi = v, lo = v;
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
      for (int sg = -1; sg <= 1; sg += 2)
#pragma unroll
        for (int t = 0; t < 3; t++) {
          const int d = t == 0 ? fdir : (fdir + (t == 1 ? 6 : 2)) & 7;
          const int yy = y + sg * CDEF_DIRECTION(d, kk, 0), xx = x + sg * CDEF_DIRECTION(d, kk, 1);
          const int cr = (yy * (1 << sub)) >> 2, cc = (xx * (1 << sub)) >> 2;
          if (cr < 0 || cr >= fp.mi_rows || cc < 0 || cc >= fp.mi_cols) continue;
          const int pv = in[yy * S + xx];
          sum += t == 0 ? CDEF_PRI_TAP(pri & 1, kk) * cdef_constrain(pv - v, pri, damping)
                        : CDEF_SEC_TAP(pri & 1, kk) * cdef_constrain(pv - v, sec, damping);
          hi = d_max(hi, pv);
          lo = d_min(lo, pv);
        }
    dst[o] = (uint8_t)d_clip(lo, hi, v + ((8 + sum - (sum < 0)) >> 4));
  }
}
__global__ __launch_bounds__(64) void k_cdef(const Mav1FrameParams fp, const uint8_t *__restrict__ cur,
                                             uint8_t *__restrict__ out, long long frameBytes,
                                             const int *__restrict__ dirvar, int n) {
  cdef_body(fp, cur, out, frameBytes, dirvar, n, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* Per superblock (256 threads): visible-area SSE of each set against the
 * source, first minimum, then the chosen set over the mi-aligned area
 * into rec. */
DEVI void cdef_select_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, const uint8_t *__restrict__ cdefbuf, long long frameBytes, int n, uint8_t *__restrict__ rec, int8_t *__restrict__ choice, const int blk_x, const int blk_y, const int blk_z) {
  __shared__ long long part[4];
  __shared__ int bestK;
  const int sb = blk_x, sr = sb / fp.sb_cols, sc = sb % fp.sb_cols, tid = threadIdx.x;
  long long best = -1;
  int bk = 0;
  for (int k = 0; k < n; k++) {
    const uint8_t *o = cdefbuf + k * frameBytes;
    long long e = 0;
    for (int p = 0; p < 3; p++) {
      const int sub = p > 0, m = 64 >> sub, S = fp.stride[p];
      const int pw = p ? (fp.frame_w + 1) >> 1 : fp.frame_w, ph = p ? (fp.frame_h + 1) >> 1 : fp.frame_h;
      const int w = d_min(m, pw - sc * m), h = d_min(m, ph - sr * m);
      for (int q = tid; q < w * h; q += 256) {
        const int i = q / w, j = q - i * w, a = fp.offset[p] + (sr * m + i) * S + sc * m + j;
        const int d = (int)o[a] - (int)src[a];
        e += d * d;
      }
    }
    e = wave_sum64(e);
    if ((tid & 63) == 0) part[tid >> 6] = e;
    __syncthreads();
    const long long sse = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    if (best < 0 || sse < best) {
      best = sse;
      bk = k;
    }
  }
  if (tid == 0) {
    bestK = bk;
    choice[sb] = (int8_t)bk;
  }
  __syncthreads();
  const uint8_t *o = cdefbuf + bestK * frameBytes;
  for (int p = 0; p < 3; p++) {
    const int sub = p > 0, m = 64 >> sub, S = fp.stride[p];
    const int mw = (fp.mi_cols * 4) >> sub, mh = (fp.mi_rows * 4) >> sub;
    const int w = d_min(m, mw - sc * m), h = d_min(m, mh - sr * m);
    for (int q = tid; q < w * h; q += 256) {
      const int i = q / w, j = q - i * w, a = fp.offset[p] + (sr * m + i) * S + sc * m + j;
      rec[a] = o[a];
    }
  }
}
__global__ __launch_bounds__(256) void k_cdef_select(const Mav1FrameParams fp, const uint8_t *__restrict__ src,
                                                     const uint8_t *__restrict__ cdefbuf, long long frameBytes, int n,
                                                     uint8_t *__restrict__ rec, int8_t *__restrict__ choice) {
  cdef_select_body(fp, src, cdefbuf, frameBytes, n, rec, choice, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* ---- loop restoration ---- */

struct LrBlk {
  int x, y, w, h, s0, s1, endX, endY, ur, uc;
};

DEVI int lr_geom(const Mav1FrameParams &fp, int p, int row, int col, LrBlk &b) {
  const int sub = p > 0, ly = row * 4, us = fp.lr_unit_size[p];
  b.s0 = (((ly + 8) / 64) * 64 - 8) >> sub;
  b.s1 = b.s0 + (64 >> sub) - 1;
  b.ur = d_min(fp.lr_unit_rows[p] - 1, ((ly + 8) >> sub) / us);
  b.uc = d_min(fp.lr_unit_cols[p] - 1, ((col * 4) >> sub) / us);
  b.endX = ((fp.frame_w + sub) >> sub) - 1;
  b.endY = ((fp.frame_h + sub) >> sub) - 1;
  b.x = (col * 4) >> sub;
  b.y = ly >> sub;
  b.w = d_min(4 >> sub, b.endX - b.x + 1);
  b.h = d_min(4 >> sub, b.endY - b.y + 1);
  return b.w > 0 && b.h > 0;
}

/* 7.17.6: clamped to the plane; rows beyond the stripe from the deblocked
 * frame, at most two rows out. */
DEVI int lr_px(const Mav1FrameParams &fp, int p, const LrBlk &b, const uint8_t *cdef, const uint8_t *dbk, int x, int y) {
  x = d_clip(0, b.endX, x);
  y = d_clip(0, b.endY, y);
  const int o = fp.offset[p] + x;
  if (y < b.s0) return dbk[o + d_max(b.s0 - 2, y) * fp.stride[p]];
  if (y > b.s1) return dbk[o + d_min(b.s1 + 2, y) * fp.stride[p]];
  return cdef[o + y * fp.stride[p]];
}

DEVI void lr_wiener(const Mav1FrameParams &fp, int p, const LrBlk &b, const uint8_t *cdef, const uint8_t *dbk,
                    uint8_t *out, const Mav1LrUnit &u) {
  int vf[7], hf[7];
  vf[3] = hf[3] = 128;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    vf[k] = vf[6 - k] = u.wiener[0][k];
    hf[k] = hf[6 - k] = u.wiener[1][k];
    vf[3] -= 2 * u.wiener[0][k];
    hf[3] -= 2 * u.wiener[1][k];
  }
  int mid[10][4];
  for (int r = 0; r < b.h + 6; r++)
    for (int c = 0; c < b.w; c++) {
      int s = 0;
#pragma unroll
      for (int t = 0; t < 7; t++) s += hf[t] * lr_px(fp, p, b, cdef, dbk, b.x + c + t - 3, b.y + r - 3);
      mid[r][c] = d_clip(-2048, 8191 - 2048, d_round2(s, 3));
    }
  for (int r = 0; r < b.h; r++)
    for (int c = 0; c < b.w; c++) {
      int s = 0;
#pragma unroll
      for (int t = 0; t < 7; t++) s += vf[t] * mid[r + t][c];
      out[fp.offset[p] + (b.y + r) * fp.stride[p] + b.x + c] = (uint8_t)d_clip1(d_round2(s, 11));
    }
}

DEVI int lr_box(const Mav1FrameParams &fp, int p, const LrBlk &b, const uint8_t *cdef, const uint8_t *dbk, int set,
                int pass, int *F) {
  const int r = SGR_PARAM(set, pass * 2);
  if (!r) return 0;
  const int eps = SGR_PARAM(set, pass * 2 + 1), n = (2 * r + 1) * (2 * r + 1), n2e = n * n * eps;
  const int s = ((1 << SGRPROJ_MTABLE_BITS) + n2e / 2) / n2e;
  const int one = ((1 << SGRPROJ_RECIP_BITS) + n / 2) / n;
  int A[6][6], B[6][6];
  for (int i = -1; i <= b.h; i++)
    for (int j = -1; j <= b.w; j++) {
      int sq = 0, sum = 0;
      for (int dy = -r; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++) {
          const int c = lr_px(fp, p, b, cdef, dbk, b.x + j + dx, b.y + i + dy);
          sq += c * c;
          sum += c;
        }
      long long pv = (long long)sq * n - (long long)sum * sum;
      if (pv < 0) pv = 0;
      const long long z = (pv * s + (1 << (SGRPROJ_MTABLE_BITS - 1))) >> SGRPROJ_MTABLE_BITS;
      const int zi = (int)d_min(255, (int)z), a2 = z >= 255 ? 256 : z == 0 ? 1 : ((zi << SGRPROJ_SGR_BITS) + zi / 2) / (zi + 1);
      const long long b2 = (long long)((1 << SGRPROJ_SGR_BITS) - a2) * sum * one;
      A[i + 1][j + 1] = a2;
      B[i + 1][j + 1] = (int)((b2 + (1 << (SGRPROJ_RECIP_BITS - 1))) >> SGRPROJ_RECIP_BITS);
    }
  for (int i = 0; i < b.h; i++) {
    const int sh = SGRPROJ_SGR_BITS + ((pass == 0 && (i & 1)) ? 4 : 5) - SGRPROJ_RST_BITS;
    for (int j = 0; j < b.w; j++) {
      long long a = 0, bb = 0;
#pragma unroll
      for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          const int wgt = pass == 0 ? (((i + dy) & 1) ? (dx == 0 ? 6 : 5) : 0) : (dx == 0 || dy == 0) ? 4 : 3;
          a += wgt * A[i + dy + 1][j + dx + 1];
          bb += wgt * B[i + dy + 1][j + dx + 1];
        }
      const long long v = a * cdef[fp.offset[p] + (b.y + i) * fp.stride[p] + b.x + j] + bb;
      F[i * 4 + j] = (int)((v + (1ll << (sh - 1))) >> sh);
    }
  }
  return 1;
}

DEVI void lr_sgr(const Mav1FrameParams &fp, int p, const LrBlk &b, const uint8_t *cdef, const uint8_t *dbk, uint8_t *out,
                 const Mav1LrUnit &u) {
  int f0[16], f1[16];
  const int on0 = lr_box(fp, p, b, cdef, dbk, u.set, 0, f0), on1 = lr_box(fp, p, b, cdef, dbk, u.set, 1, f1);
  const int w0 = u.xqd[0], w1 = u.xqd[1], w2 = (1 << SGRPROJ_PRJ_BITS) - w0 - w1;
  for (int i = 0; i < b.h; i++)
    for (int j = 0; j < b.w; j++) {
      const int o = fp.offset[p] + (b.y + i) * fp.stride[p] + b.x + j;
      const int uu = cdef[o] << SGRPROJ_RST_BITS;
      const int v = w1 * uu + (on0 ? w0 * f0[i * 4 + j] : w0 * uu) + (on1 ? w2 * f1[i * 4 + j] : w2 * uu);
      out[o] = (uint8_t)d_clip1(d_round2(v, SGRPROJ_RST_BITS + SGRPROJ_PRJ_BITS));
    }
}

/* One 4x4 luma block (2x2 chroma) per thread: x = block column, y = block
 * row, z = plane; lrType: the frame's restoration type per plane. */
DEVI void lr_apply_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ cdef, const uint8_t *__restrict__ dbk, uint8_t *__restrict__ out, const Mav1LrUnit *__restrict__ units, const int *__restrict__ lrType, const int blk_x, const int blk_y, const int blk_z) {
  const int p = blk_z, col = blk_x * 64 + threadIdx.x, row = blk_y;
  if (col >= ((fp.frame_w + 3) >> 2)) return;
  LrBlk b;
  if (!lr_geom(fp, p, row, col, b)) return;
  Mav1LrUnit u;
  u.type = RESTORE_NONE;
  if (lrType[p] != RESTORE_NONE) u = units[fp.lr_unit_off[p] + b.ur * fp.lr_unit_cols[p] + b.uc];
  if (u.type == RESTORE_WIENER) lr_wiener(fp, p, b, cdef, dbk, out, u);
  else if (u.type == RESTORE_SGRPROJ) lr_sgr(fp, p, b, cdef, dbk, out, u);
  else
    for (int i = 0; i < b.h; i++)
      for (int j = 0; j < b.w; j++) {
        const int o = fp.offset[p] + (b.y + i) * fp.stride[p] + b.x + j;
        out[o] = cdef[o];
      }
}
__global__ __launch_bounds__(64) void k_lr_apply(const Mav1FrameParams fp, const uint8_t *__restrict__ cdef,
                                                 const uint8_t *__restrict__ dbk, uint8_t *__restrict__ out,
                                                 const Mav1LrUnit *__restrict__ units, const int *__restrict__ lrType) {
  lr_apply_body(fp, cdef, dbk, out, units, lrType, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* ---- launchers ---- */

hipError_t launch_filter_info(hipStream_t st, const Mav1FrameParams &fp, const Mav1BlockInfo *blocks,
                              const Mav1SbInfo *sbs, uint8_t *info) {
  hipLaunchKernelGGL(k_filter_info, dim3(fp.sb_cols * fp.sb_rows), dim3(64), 0, st, fp, blocks, sbs, info);
  return hipGetLastError();
}

hipError_t launch_deblock(hipStream_t st, const Mav1FrameParams &fp, uint8_t *rec, const uint8_t *info) {
  hipLaunchKernelGGL(k_deblock, dim3((fp.mi_cols + 63) / 64, fp.mi_rows * 4, 3), dim3(64), 0, st, fp, rec, info, 0);
  hipLaunchKernelGGL(k_deblock, dim3((fp.mi_cols * 4 + 63) / 64, fp.mi_rows, 3), dim3(64), 0, st, fp, rec, info, 1);
  return hipGetLastError();
}

hipError_t launch_cdef(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *cur, uint8_t *out,
                       long long frameBytes, const uint8_t *info, int *dirvar, int n) {
  const int nb = (fp.mi_cols >> 1) * (fp.mi_rows >> 1);
  hipLaunchKernelGGL(k_cdef_dir, dim3((nb + 63) / 64), dim3(64), 0, st, fp, cur, info, dirvar);
  hipLaunchKernelGGL(k_cdef, dim3((fp.mi_cols * 4 + 63) / 64, fp.mi_rows * 4, 3), dim3(64), 0, st, fp, cur, out,
                     frameBytes, (const int *)dirvar, n);
  return hipGetLastError();
}

hipError_t launch_cdef_select(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdefbuf,
                              long long frameBytes, int n, uint8_t *rec, int8_t *choice) {
  hipLaunchKernelGGL(k_cdef_select, dim3(fp.sb_cols * fp.sb_rows), dim3(256), 0, st, fp, src, cdefbuf, frameBytes, n,
                     rec, choice);
  return hipGetLastError();
}

hipError_t launch_lr_apply(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *cdef, const uint8_t *dbk,
                           uint8_t *out, const Mav1LrUnit *units, const int *lrType) {
  hipLaunchKernelGGL(k_lr_apply, dim3((((fp.frame_w + 3) >> 2) + 63) / 64, (fp.frame_h + 3) >> 2, 3), dim3(64), 0, st,
                     fp, cdef, dbk, out, units, lrType);
  return hipGetLastError();
}

/* ---- decision record (src/encoder.c:659-704) ---- */

/* Per superblock, lane per block: the 8-byte cells of the 8x8 grid the
 * block covers, where no later block of the superblock (in decode order)
 * covers them too, so each cell holds the last block as the serial dump
 * leaves it. Cells outside every block are cleared first. */
DEVI void decision_grid_body(const Mav1FrameParams &fp, const Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo *__restrict__ sbs, uint8_t *__restrict__ grid, const int blk_x, const int blk_y, const int blk_z) {
  const int sb = blk_x, lane = threadIdx.x;
  const int rows8 = (fp.mi_rows + 1) >> 1, cols8 = (fp.mi_cols + 1) >> 1;
  const int sr = sb / fp.sb_cols, sc = sb % fp.sb_cols;
  /* the superblock's 8x8 cells */
  for (int q = lane; q < 64; q += 64) {
    const int r = sr * 8 + (q >> 3), c = sc * 8 + (q & 7);
    if (r < rows8 && c < cols8) *(uint2 *)(grid + ((size_t)r * cols8 + c) * 8) = make_uint2(0, 0);
  }
  __syncthreads();
  const int n = sbs[sb].num_blocks;
  if (lane >= n) return;
  const Mav1BlockInfo *base = blocks + sbs[sb].first_block;
  const Mav1BlockInfo &b = base[lane];
  const int w8 = (Num_4x4_Blocks_Wide[b.bsize] + 1) >> 1, h8 = (Num_4x4_Blocks_High[b.bsize] + 1) >> 1;
  const int r0 = b.mi_row >> 1, c0 = b.mi_col >> 1;
  const uint32_t lo = (uint32_t)(uint8_t)b.bsize | (uint32_t)(!b.is_inter ? 0 : b.ref_frame1 ? 2 : 1) << 8 |
                      (uint32_t)(uint8_t)b.ref_frame << 16;
  const uint32_t hi = (uint32_t)(uint16_t)(int16_t)b.mv_row | (uint32_t)(uint16_t)(int16_t)b.mv_col << 16;
  for (int r = r0; r < d_min(rows8, r0 + h8); r++)
    for (int c = c0; c < d_min(cols8, c0 + w8); c++) {
      bool later = false;
      for (int j = lane + 1; j < n && !later; j++) {
        const Mav1BlockInfo &o = base[j];
        const int or0 = o.mi_row >> 1, oc0 = o.mi_col >> 1;
        later = r >= or0 && r < or0 + ((Num_4x4_Blocks_High[o.bsize] + 1) >> 1) && c >= oc0 &&
                c < oc0 + ((Num_4x4_Blocks_Wide[o.bsize] + 1) >> 1);
      }
      if (!later) *(uint2 *)(grid + ((size_t)r * cols8 + c) * 8) = make_uint2(lo, hi);
    }
}
__global__ __launch_bounds__(64) void k_decision_grid(const Mav1FrameParams fp, const Mav1BlockInfo *__restrict__ blocks,
                                                      const Mav1SbInfo *__restrict__ sbs, uint8_t *__restrict__ grid) {
  decision_grid_body(fp, blocks, sbs, grid, blockIdx.x, blockIdx.y, blockIdx.z);
}

hipError_t launch_decision_grid(hipStream_t st, const Mav1FrameParams &fp, const Mav1BlockInfo *blocks,
                                const Mav1SbInfo *sbs, uint8_t *grid) {
  hipLaunchKernelGGL(k_decision_grid, dim3(fp.sb_cols * fp.sb_rows), dim3(64), 0, st, fp, blocks, sbs, grid);
  return hipGetLastError();
}

/* ---- batched launches (batch.h): the same kernels over several frames ---- */

#define BATCH_FRAME()              \
  const BatchIdx b = batch_idx(g); \
  const BatchFrame &F = T[b.f]

__global__ __launch_bounds__(64) void k_filter_info_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  BATCH_FRAME();
  filter_info_body(F.fp, gptr(F.blocks), gptr(F.sbs), gptr(F.lfmask), b.x, b.y, b.z);
}
__global__ __launch_bounds__(64) void k_deblock_b(const BatchFrame *__restrict__ T, const BatchGrid g, int pass) {
  BATCH_FRAME();
  deblock_body(F.fp, gptr(F.rec), gptr(F.lfmask), pass, b.x, b.y, b.z);
}
__global__ __launch_bounds__(64) void k_cdef_dir_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  BATCH_FRAME();
  cdef_dir_body(F.fp, gptr(F.rec), gptr(F.lfmask), gptr(F.dirvar), b.x, b.y, b.z);
}
__global__ __launch_bounds__(64) void k_cdef_b(const BatchFrame *__restrict__ T, const BatchGrid g, int n) {
  BATCH_FRAME();
  cdef_body(F.fp, gptr(F.rec), gptr(F.cdefbuf), F.frame_bytes, gptr(F.dirvar), n, b.x, b.y, b.z);
}
__global__ __launch_bounds__(256) void k_cdef_select_b(const BatchFrame *__restrict__ T, const BatchGrid g, int n) {
  BATCH_FRAME();
  cdef_select_body(F.fp, gptr(F.src), gptr(F.cdefbuf), F.frame_bytes, n, gptr(F.rec), gptr(F.choice), b.x, b.y, b.z);
}
__global__ __launch_bounds__(64) void k_lr_apply_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  BATCH_FRAME();
  lr_apply_body(F.fp, gptr(F.rec), gptr(F.dbk), gptr(F.lrout), gptr(F.lrunits), gptr(F.lrtype), b.x, b.y, b.z);
}
__global__ __launch_bounds__(64) void k_decision_grid_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  BATCH_FRAME();
  decision_grid_body(F.fp, gptr(F.blocks), gptr(F.sbs), gptr(F.decgrid), b.x, b.y, b.z);
}

/* The driver's frame copies and clears, 16 bytes per thread (the buffers
 * are 256-byte aligned allocations; a tail shorter than 16 bytes is copied
 * bytewise). */
__global__ __launch_bounds__(256) void k_frame_op_b(const BatchFrame *__restrict__ T, const BatchGrid g, int which) {
  BATCH_FRAME();
  const long long o = ((long long)b.x * 256 + threadIdx.x) * 16;
  if (which == 2) {
    const long long nsb = (long long)F.fp.sb_cols * F.fp.sb_rows;
    for (long long q = o; q < o + 16 && q < nsb; q++) gptr(F.choice)[q] = 0;
    return;
  }
  const uint8_t *src = gptr(F.rec);
  uint8_t *dst = which == 0 ? gptr(F.dbk) : gptr(F.lrout);
  const long long n = F.frame_bytes;
  if (o + 16 <= n) {
    *(uint4 *)(dst + o) = *(const uint4 *)(src + o);
  } else {
    for (long long q = o; q < n; q++) dst[q] = src[q];
  }
}

static BatchGrid sb_grid(const BatchFrame *H, int n) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) batch_grid_add(g, H[f].fp.sb_cols * H[f].fp.sb_rows, 1, 1);
  return g;
}

hipError_t launch_filter_info_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  const BatchGrid g = sb_grid(H, n);
  hipLaunchKernelGGL(k_filter_info_b, dim3(g.start[n]), dim3(64), 0, st, T, g);
  return hipGetLastError();
}

hipError_t launch_deblock_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  BatchGrid g0 = batch_grid(), g1 = batch_grid();
  for (int f = 0; f < n; f++) {
    const Mav1FrameParams &fp = H[f].fp;
    batch_grid_add(g0, (fp.mi_cols + 63) / 64, fp.mi_rows * 4, 3);
    batch_grid_add(g1, (fp.mi_cols * 4 + 63) / 64, fp.mi_rows, 3);
  }
  hipLaunchKernelGGL(k_deblock_b, dim3(g0.start[n]), dim3(64), 0, st, T, g0, 0);
  hipLaunchKernelGGL(k_deblock_b, dim3(g1.start[n]), dim3(64), 0, st, T, g1, 1);
  return hipGetLastError();
}

hipError_t launch_cdef_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int nsets) {
  BatchGrid g0 = batch_grid(), g1 = batch_grid();
  for (int f = 0; f < n; f++) {
    const Mav1FrameParams &fp = H[f].fp;
    const int nb = (fp.mi_cols >> 1) * (fp.mi_rows >> 1);
    batch_grid_add(g0, (nb + 63) / 64, 1, 1);
    batch_grid_add(g1, (fp.mi_cols * 4 + 63) / 64, fp.mi_rows * 4, 3);
  }
  hipLaunchKernelGGL(k_cdef_dir_b, dim3(g0.start[n]), dim3(64), 0, st, T, g0);
  hipLaunchKernelGGL(k_cdef_b, dim3(g1.start[n]), dim3(64), 0, st, T, g1, nsets);
  return hipGetLastError();
}

hipError_t launch_cdef_select_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int nsets) {
  const BatchGrid g = sb_grid(H, n);
  hipLaunchKernelGGL(k_cdef_select_b, dim3(g.start[n]), dim3(256), 0, st, T, g, nsets);
  return hipGetLastError();
}

hipError_t launch_lr_apply_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) {
    const Mav1FrameParams &fp = H[f].fp;
    batch_grid_add(g, (((fp.frame_w + 3) >> 2) + 63) / 64, (fp.frame_h + 3) >> 2, 3);
  }
  hipLaunchKernelGGL(k_lr_apply_b, dim3(g.start[n]), dim3(64), 0, st, T, g);
  return hipGetLastError();
}

hipError_t launch_decision_grid_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  const BatchGrid g = sb_grid(H, n);
  hipLaunchKernelGGL(k_decision_grid_b, dim3(g.start[n]), dim3(64), 0, st, T, g);
  return hipGetLastError();
}

hipError_t launch_frame_op_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int which) {
  if (which < 0 || which > 2) return hipErrorInvalidValue;
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) {
    const long long bytes = which == 2 ? (long long)H[f].fp.sb_cols * H[f].fp.sb_rows : H[f].frame_bytes;
    batch_grid_add(g, (int)((bytes + 16 * 256 - 1) / (16 * 256)), 1, 1);
  }
  hipLaunchKernelGGL(k_frame_op_b, dim3(g.start[n]), dim3(256), 0, st, T, g, which);
  return hipGetLastError();
}
