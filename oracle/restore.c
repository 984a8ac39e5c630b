/* ORACLE — test infrastructure only. Loop restoration (AV1 spec 7.17):
 * the per-4x4 apply of src/kernels/k_lr.h and the per-unit parameter search
 * of src/lr_search.c. The search fits in FP64; every floating-point
 * expression below keeps the reference's operand order (the file is built
 * with -ffp-contract=off), because the rounded fits decide the coded taps.
 * Serial C restatement; each function cites the reference lines it follows. */
#include <math.h>

#include "ocommon.h"

typedef struct {
  int x, y, w, h;            /* block in plane samples */
  int s0, s1;                /* stripe rows */
  int endX, endY;            /* last plane sample */
  int ur, uc;                /* restoration unit */
} LrBlk;

/* Geometry of the 4x4 luma block (row, col) in plane p (k_lr.h:28-45):
 * stripes of 64 luma rows offset by 8; units 64x64 with the +8 row offset. */
static int lr_geom(const Mav1FrameParams *fp, int p, int row, int col, LrBlk *b) {
  const int sub = p > 0, ly = row * 4, us = fp->lr_unit_size[p];
  b->s0 = (((ly + 8) / 64) * 64 - 8) >> sub;
  b->s1 = b->s0 + (64 >> sub) - 1;
  b->ur = o_min(fp->lr_unit_rows[p] - 1, ((ly + 8) >> sub) / us);
  b->uc = o_min(fp->lr_unit_cols[p] - 1, ((col * 4) >> sub) / us);
  b->endX = ((fp->frame_w + sub) >> sub) - 1;
  b->endY = ((fp->frame_h + sub) >> sub) - 1;
  b->x = (col * 4) >> sub;
  b->y = ly >> sub;
  b->w = o_min(4 >> sub, b->endX - b->x + 1);
  b->h = o_min(4 >> sub, b->endY - b->y + 1);
  return b->w > 0 && b->h > 0;
}

/* 7.17.6 get_source_sample: clamp to the plane, rows outside the stripe
 * come from the deblocked frame (at most 2 rows beyond it). */
static int lr_px(const Mav1FrameParams *fp, int p, const LrBlk *b, const uint8_t *cdef, const uint8_t *dbk, int x, int y) {
  x = o_clip(0, b->endX, x);
  y = o_clip(0, b->endY, y);
  const int o = fp->offset[p] + x;
  if (y < b->s0) return dbk[o + o_max(b->s0 - 2, y) * fp->stride[p]];
  if (y > b->s1) return dbk[o + o_min(b->s1 + 2, y) * fp->stride[p]];
  return cdef[o + y * fp->stride[p]];
}

/* 7.17.4 Wiener filter, 8-bit rounding 3 / 11 (k_lr.h:67-100). */
static void lr_wiener(const Mav1FrameParams *fp, int p, const LrBlk *b, const uint8_t *cdef, const uint8_t *dbk,
                      uint8_t *out, const Mav1LrUnit *u) {
  int vf[7], hf[7];
  vf[3] = hf[3] = 128;
  for (int k = 0; k < 3; k++) {
    vf[k] = vf[6 - k] = u->wiener[0][k];
    hf[k] = hf[6 - k] = u->wiener[1][k];
    vf[3] -= 2 * u->wiener[0][k];
    hf[3] -= 2 * u->wiener[1][k];
  }
  int mid[10][4];
  for (int r = 0; r < b->h + 6; r++)
    for (int c = 0; c < b->w; c++) {
      int s = 0;
      for (int t = 0; t < 7; t++) s += hf[t] * lr_px(fp, p, b, cdef, dbk, b->x + c + t - 3, b->y + r - 3);
      mid[r][c] = o_clip(-2048, 8191 - 2048, o_round2(s, 3));
    }
  for (int r = 0; r < b->h; r++)
    for (int c = 0; c < b->w; c++) {
      int s = 0;
      for (int t = 0; t < 7; t++) s += vf[t] * mid[r + t][c];
      out[fp->offset[p] + (b->y + r) * fp->stride[p] + b->x + c] = (uint8_t)o_clip1(o_round2(s, 11));
    }
}

/* 7.17.3 box filter of one pass (radius from the set; k_lr.h:103-154);
 * F receives h x w values, row stride 4. Returns 0 when the pass is off. */
static int lr_box(const Mav1FrameParams *fp, int p, const LrBlk *b, const uint8_t *cdef, const uint8_t *dbk, int set,
                  int pass, int *F) {
  const int r = SGR_PARAM(set, pass * 2);
  if (!r) return 0;
  const int eps = SGR_PARAM(set, pass * 2 + 1), n = (2 * r + 1) * (2 * r + 1), n2e = n * n * eps;
  const int s = ((1 << SGRPROJ_MTABLE_BITS) + n2e / 2) / n2e;
  const int one = ((1 << SGRPROJ_RECIP_BITS) + n / 2) / n;
  int A[6][6], B[6][6];
  for (int i = -1; i <= b->h; i++)
    for (int j = -1; j <= b->w; j++) {
      int64_t sq = 0, sum = 0;
      for (int dy = -r; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++) {
          const int64_t c = lr_px(fp, p, b, cdef, dbk, b->x + j + dx, b->y + i + dy);
          sq += c * c;
          sum += c;
        }
      int64_t pv = sq * n - sum * sum;
      if (pv < 0) pv = 0;
      const int64_t z = (pv * s + (1 << (SGRPROJ_MTABLE_BITS - 1))) >> SGRPROJ_MTABLE_BITS;
      const int64_t a2 = z >= 255 ? 256 : z == 0 ? 1 : ((z << SGRPROJ_SGR_BITS) + z / 2) / (z + 1);
      const int64_t b2 = ((1 << SGRPROJ_SGR_BITS) - a2) * sum * one;
      A[i + 1][j + 1] = (int)a2;
      B[i + 1][j + 1] = (int)((b2 + (1 << (SGRPROJ_RECIP_BITS - 1))) >> SGRPROJ_RECIP_BITS);
    }
  for (int i = 0; i < b->h; i++) {
    const int sh = SGRPROJ_SGR_BITS + ((pass == 0 && (i & 1)) ? 4 : 5) - SGRPROJ_RST_BITS;
    for (int j = 0; j < b->w; j++) {
      int64_t a = 0, bb = 0;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int wgt = pass == 0 ? (((i + dy) & 1) ? (dx == 0 ? 6 : 5) : 0) : (dx == 0 || dy == 0) ? 4 : 3;
          a += wgt * A[i + dy + 1][j + dx + 1];
          bb += wgt * B[i + dy + 1][j + dx + 1];
        }
      const int64_t v = a * cdef[fp->offset[p] + (b->y + i) * fp->stride[p] + b->x + j] + bb;
      F[i * 4 + j] = (int)((v + ((int64_t)1 << (sh - 1))) >> sh);
    }
  }
  return 1;
}

/* 7.17.2 self-guided filter with projection (k_lr.h:157-178). */
static void lr_sgr(const Mav1FrameParams *fp, int p, const LrBlk *b, const uint8_t *cdef, const uint8_t *dbk,
                   uint8_t *out, const Mav1LrUnit *u) {
  int f0[16], f1[16];
  const int on0 = lr_box(fp, p, b, cdef, dbk, u->set, 0, f0), on1 = lr_box(fp, p, b, cdef, dbk, u->set, 1, f1);
  const int w0 = u->xqd[0], w1 = u->xqd[1], w2 = (1 << SGRPROJ_PRJ_BITS) - w0 - w1;
  for (int i = 0; i < b->h; i++)
    for (int j = 0; j < b->w; j++) {
      const int o = fp->offset[p] + (b->y + i) * fp->stride[p] + b->x + j;
      const int uu = cdef[o] << SGRPROJ_RST_BITS;
      const int v = w1 * uu + (on0 ? w0 * f0[i * 4 + j] : w0 * uu) + (on1 ? w2 * f1[i * 4 + j] : w2 * uu);
      out[o] = (uint8_t)o_clip1(o_round2(v, SGRPROJ_RST_BITS + SGRPROJ_PRJ_BITS));
    }
}

/* lr_apply (k_lr.h:182-197): NONE copies the CDEF output. */
static void lr_apply(const Mav1FrameParams *fp, int p, const LrBlk *b, const uint8_t *cdef, const uint8_t *dbk,
                     uint8_t *out, const Mav1LrUnit *u) {
  if (u->type == RESTORE_WIENER) lr_wiener(fp, p, b, cdef, dbk, out, u);
  else if (u->type == RESTORE_SGRPROJ) lr_sgr(fp, p, b, cdef, dbk, out, u);
  else
    for (int i = 0; i < b->h; i++) {
      const int o = fp->offset[p] + (b->y + i) * fp->stride[p] + b->x;
      memcpy(out + o, cdef + o, (size_t)b->w);
    }
}

/* The restore stage (src/backend_cpu.c:179-188 over lr_block, k_lr.h:
 * 200-208): every plane and 4x4 luma block; planes whose lr_type is NONE
 * copy. */
void oracle_lr_apply(const Mav1FrameParams *fp, const uint8_t *cdef, const uint8_t *dbk, uint8_t *out,
                     const Mav1LrUnit *units) {
  const int rows4 = (fp->frame_h + 3) >> 2, cols4 = (fp->frame_w + 3) >> 2;
  for (int p = 0; p < 3; p++)
    for (int r = 0; r < rows4; r++)
      for (int c = 0; c < cols4; c++) {
        LrBlk b;
        if (!lr_geom(fp, p, r, c, &b)) continue;
        Mav1LrUnit u;
        memset(&u, 0, sizeof(u));
        if (fp->lr_type[p] != RESTORE_NONE) u = units[fp->lr_unit_off[p] + b.ur * fp->lr_unit_cols[p] + b.uc];
        lr_apply(fp, p, &b, cdef, dbk, out, &u);
      }
}

/* ---- parameter search (src/lr_search.c) ---- */

static int recenter(int r, int v) {
  if (v > 2 * r) return v;
  return v >= r ? (v - r) << 1 : ((r - v) << 1) - 1;
}

/* Bit count of v coded as decode_signed_subexp_with_ref_bool(low, high,
 * k, r) reads it (lr_search.c:31-56). */
static int subexp_bits(int low, int high, int k, int r, int v) {
  const int mx = high - low;
  r -= low;
  v -= low;
  const int x = 2 * r <= mx ? recenter(r, v) : recenter(mx - 1 - r, mx - 1 - v);
  int bits = 0, mk = 0;
  for (int i = 0;; i++) {
    const int b2 = i ? k + i - 1 : k, a = 1 << b2;
    if (mx <= mk + 3 * a) {
      const int n = mx - mk, w = o_floorlog2((unsigned)n) + 1, m = (1 << w) - n;
      return bits + (x - mk < m ? w - 1 : w);
    }
    bits++;
    if (x < mk + a) return bits + b2;
    mk += a;
  }
}

/* Parameter bits of a unit against the running references, which it
 * updates (lr_search.c:60-79). */
static int unit_bits(int p, const Mav1LrUnit *u, int refW[2][3], int refX[2]) {
  int bits = 0;
  if (u->type == RESTORE_WIENER) {
    for (int pass = 0; pass < 2; pass++)
      for (int j = p ? 1 : 0; j < 3; j++) {
        bits += subexp_bits(Wiener_Taps_Min[j], Wiener_Taps_Max[j] + 1, Wiener_Taps_K[j], refW[pass][j], u->wiener[pass][j]);
        refW[pass][j] = u->wiener[pass][j];
      }
  } else if (u->type == RESTORE_SGRPROJ) {
    bits += SGRPROJ_PARAMS_BITS;
    for (int i = 0; i < 2; i++) {
      if (SGR_PARAM(u->set, i * 2))
        bits += subexp_bits(Sgrproj_Xqd_Min[i], Sgrproj_Xqd_Max[i] + 1, SGRPROJ_PRJ_SUBEXP_K, refX[i], u->xqd[i]);
      refX[i] = u->xqd[i];
    }
  }
  return bits;
}

typedef struct {
  const Mav1FrameParams *fp;
  const uint8_t *src, *cdef, *dbk;
  uint8_t *tmp;
  int p, n;
  const int *rc; /* (row, col) of the unit's 4x4 blocks in raster order */
} Unit;

static int64_t unit_sse(const Unit *U, const Mav1LrUnit *u) {
  int64_t sse = 0;
  for (int k = 0; k < U->n; k++) {
    LrBlk b;
    lr_geom(U->fp, U->p, U->rc[2 * k], U->rc[2 * k + 1], &b);
    lr_apply(U->fp, U->p, &b, U->cdef, U->dbk, U->tmp, u);
    for (int i = 0; i < b.h; i++)
      for (int j = 0; j < b.w; j++) {
        const int o = U->fp->offset[U->p] + (b.y + i) * U->fp->stride[U->p] + b.x + j;
        const int d = (int)U->tmp[o] - (int)U->src[o];
        sse += d * d;
      }
  }
  return sse;
}

/* Gaussian elimination with partial pivoting (lr_search.c:98-129). */
static int solve3(int n, double M[3][3], double v[3], double *x) {
  for (int i = 0; i < n; i++) {
    int piv = i;
    for (int r = i + 1; r < n; r++)
      if (fabs(M[r][i]) > fabs(M[piv][i])) piv = r;
    if (fabs(M[piv][i]) < 1e-9) return 0;
    if (piv != i) {
      for (int k = 0; k < n; k++) {
        const double t = M[i][k];
        M[i][k] = M[piv][k];
        M[piv][k] = t;
      }
      const double t = v[i];
      v[i] = v[piv];
      v[piv] = t;
    }
    for (int r = i + 1; r < n; r++) {
      const double f = M[r][i] / M[i][i];
      for (int k = i; k < n; k++) M[r][k] -= f * M[i][k];
      v[r] -= f * v[i];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = v[i];
    for (int k = i + 1; k < n; k++) s -= M[i][k] * x[k];
    x[i] = s / M[i][i];
  }
  return 1;
}

/* Separable least-squares Wiener fit, four alternating solves starting
 * with the vertical filter (lr_search.c:132-180). */
static void wiener_fit(const Unit *U, int out[2][3]) {
  const int first = U->p ? 1 : 0, nc = 3 - first;
  double fv[7] = {0, 0, 0, 1, 0, 0, 0}, fh[7] = {0, 0, 0, 1, 0, 0, 0};
  for (int it = 0; it < 4; it++) {
    const int vert = !(it & 1);
    double M[3][3] = {{0}}, V[3] = {0};
    for (int k = 0; k < U->n; k++) {
      LrBlk b;
      lr_geom(U->fp, U->p, U->rc[2 * k], U->rc[2 * k + 1], &b);
      for (int i = 0; i < b.h; i++)
        for (int j = 0; j < b.w; j++) {
          double S[7][7], Z[7], phi[3];
          for (int a = 0; a < 7; a++)
            for (int c = 0; c < 7; c++) S[a][c] = lr_px(U->fp, U->p, &b, U->cdef, U->dbk, b.x + j + c - 3, b.y + i + a - 3);
          for (int a = 0; a < 7; a++) {
            double s = 0;
            for (int t = 0; t < 7; t++) s += vert ? fh[t] * S[a][t] : fv[t] * S[t][a];
            Z[a] = s;
          }
          for (int a = first; a < 3; a++) phi[a - first] = Z[a] + Z[6 - a] - 2 * Z[3];
          const double tgt = U->src[U->fp->offset[U->p] + (b.y + i) * U->fp->stride[U->p] + b.x + j] - Z[3];
          for (int a = 0; a < nc; a++) {
            V[a] += phi[a] * tgt;
            for (int c = 0; c < nc; c++) M[a][c] += phi[a] * phi[c];
          }
        }
    }
    double x[3] = {0, 0, 0};
    for (int a = 0; a < nc; a++) M[a][a] += 1e-3;
    if (!solve3(nc, M, V, x)) break;
    double *dst = vert ? fv : fh, sum = 0;
    for (int a = 0; a < 3; a++) {
      const double v = a < first ? 0 : x[a - first];
      dst[a] = dst[6 - a] = v;
      sum += v;
    }
    dst[3] = 1 - 2 * sum;
  }
  for (int pass = 0; pass < 2; pass++)
    for (int j = 0; j < 3; j++)
      out[pass][j] = j < first ? 0
                               : o_clip(Wiener_Taps_Min[j], Wiener_Taps_Max[j], (int)lround((pass ? fh : fv)[j] * 128));
}

/* Least-squares projection weights of self-guided set `set`
 * (lr_search.c:183-231). Every accumulated term is an integer, so the sums
 * are exact. */
static void sgr_fit(const Unit *U, int set, int xqd[2]) {
  const int r0 = SGR_PARAM(set, 0), r1 = SGR_PARAM(set, 2);
  double H00 = 0, H01 = 0, H11 = 0, C0 = 0, C1 = 0;
  for (int k = 0; k < U->n; k++) {
    LrBlk b;
    lr_geom(U->fp, U->p, U->rc[2 * k], U->rc[2 * k + 1], &b);
    int f0[16], f1[16];
    lr_box(U->fp, U->p, &b, U->cdef, U->dbk, set, 0, f0);
    lr_box(U->fp, U->p, &b, U->cdef, U->dbk, set, 1, f1);
    for (int i = 0; i < b.h; i++)
      for (int j = 0; j < b.w; j++) {
        const int o = U->fp->offset[U->p] + (b.y + i) * U->fp->stride[U->p] + b.x + j;
        const double u = U->cdef[o] << SGRPROJ_RST_BITS;
        const double s = (U->src[o] << SGRPROJ_RST_BITS) - u;
        const double e0 = r0 ? f0[i * 4 + j] - u : 0, e1 = r1 ? f1[i * 4 + j] - u : 0;
        H00 += e0 * e0;
        H01 += e0 * e1;
        H11 += e1 * e1;
        C0 += e0 * s;
        C1 += e1 * s;
      }
  }
  double a = 0, bb = 0;
  if (r0 && r1) {
    const double det = H00 * H11 - H01 * H01;
    if (fabs(det) > 1e-9) {
      a = (C0 * H11 - C1 * H01) / det;
      bb = (C1 * H00 - C0 * H01) / det;
    }
  } else if (r0) {
    if (H00 > 0) a = C0 / H00;
  } else if (r1) {
    if (H11 > 0) bb = C1 / H11;
  }
  const int w0 = r0 ? (int)lround(a * (1 << SGRPROJ_PRJ_BITS)) : 0, w2 = (int)lround(bb * (1 << SGRPROJ_PRJ_BITS));
  xqd[0] = o_clip(Sgrproj_Xqd_Min[0], Sgrproj_Xqd_Max[0], w0);
  xqd[1] = o_clip(Sgrproj_Xqd_Min[1], Sgrproj_Xqd_Max[1], (1 << SGRPROJ_PRJ_BITS) - xqd[0] - (r1 ? w2 : 0));
}

/* The reference-independent results of one unit: the SSE of NONE, the
 * fitted Wiener filter and its SSE, each self-guided set's fitted weights
 * and SSE. */
typedef struct {
  int64_t none_sse, w_sse, sg_sse[1 << SGRPROJ_PARAMS_BITS];
  Mav1LrUnit w, sg[1 << SGRPROJ_PARAMS_BITS];
} Fit;

static int64_t rd_cost(const Mav1FrameParams *fp, int64_t sse, int bits) { return sse * 16 + (int64_t)fp->lambda * bits; }

/* lr_select (lr_search.c:235-359): per plane, units in raster order with
 * running subexp references; per unit NONE, the fitted Wiener filter with
 * a greedy +-1 refinement of each tap, and the 16 self-guided sets, the
 * smallest 16 * SSE + lambda * bits winning (ties keep the earlier
 * candidate). lrTest forces Wiener on even units and the self-guided set
 * of smallest SSE on odd ones. tmp is a frame-sized scratch buffer.
 * Writes units and fp->lr_type. */
void oracle_lr_select(Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdef, const uint8_t *dbk, uint8_t *tmp,
                      Mav1LrUnit *units, int lrTest) {
  const int rows4 = (fp->frame_h + 3) >> 2, cols4 = (fp->frame_w + 3) >> 2;
  int *rc = malloc(sizeof(int) * 2 * rows4 * cols4);
  for (int p = 0; p < 3; p++) {
    const int nu = fp->lr_unit_rows[p] * fp->lr_unit_cols[p];
    int *start = calloc((size_t)nu + 1, sizeof(int)), *fill = calloc((size_t)nu, sizeof(int));
    /* group the 4x4 blocks by unit, raster order kept (counting sort) */
    for (int r = 0; r < rows4; r++)
      for (int c = 0; c < cols4; c++) {
        LrBlk b;
        if (lr_geom(fp, p, r, c, &b)) start[b.ur * fp->lr_unit_cols[p] + b.uc + 1]++;
      }
    for (int u = 0; u < nu; u++) start[u + 1] += start[u];
    for (int r = 0; r < rows4; r++)
      for (int c = 0; c < cols4; c++) {
        LrBlk b;
        if (!lr_geom(fp, p, r, c, &b)) continue;
        const int u = b.ur * fp->lr_unit_cols[p] + b.uc, k = start[u] + fill[u]++;
        rc[2 * k] = r;
        rc[2 * k + 1] = c;
      }
    /* The fits and the SSEs of the fitted candidates do not depend on the
     * running subexp references, so they are computed for every unit first
     * (in parallel: units cover disjoint samples, also of tmp); the
     * reference-dependent part (bit costs, the greedy tap refinement) then
     * runs in raster order exactly as the serial search does. */
    Fit *fit = malloc(sizeof(Fit) * (size_t)(nu > 0 ? nu : 1));
#pragma omp parallel for schedule(dynamic, 1)
    for (int u = 0; u < nu; u++) {
      const Unit U = {fp, src, cdef, dbk, tmp, p, start[u + 1] - start[u], rc + 2 * start[u]};
      Fit *F = &fit[u];
      Mav1LrUnit t;
      memset(&t, 0, sizeof(t));
      t.type = RESTORE_NONE;
      F->none_sse = unit_sse(&U, &t);
      memset(&F->w, 0, sizeof(F->w));
      F->w.type = RESTORE_WIENER;
      wiener_fit(&U, F->w.wiener);
      F->w_sse = unit_sse(&U, &F->w);
      for (int set = 0; set < (1 << SGRPROJ_PARAMS_BITS); set++) {
        Mav1LrUnit *sg = &F->sg[set];
        memset(sg, 0, sizeof(*sg));
        sg->type = RESTORE_SGRPROJ;
        sg->set = set;
        sgr_fit(&U, set, sg->xqd);
        F->sg_sse[set] = unit_sse(&U, sg);
      }
    }
    int refW[2][3], refX[2], used = 0;
    for (int i = 0; i < 2; i++) {
      refX[i] = Sgrproj_Xqd_Mid[i];
      for (int j = 0; j < 3; j++) refW[i][j] = Wiener_Taps_Mid[j];
    }
    for (int u = 0; u < nu; u++) {
      const Unit U = {fp, src, cdef, dbk, tmp, p, start[u + 1] - start[u], rc + 2 * start[u]};
      const Fit *F = &fit[u];
      int tw[2][3], tx[2];
      Mav1LrUnit best, w = F->w;
      memset(&best, 0, sizeof(best));
      best.type = RESTORE_NONE;
      int64_t bestCost = rd_cost(fp, F->none_sse, 1);
      memcpy(tw, refW, sizeof(tw));
      memcpy(tx, refX, sizeof(tx));
      int64_t wc = rd_cost(fp, F->w_sse, 2 + unit_bits(p, &w, tw, tx));
      for (int pass = 0; pass < 2; pass++)
        for (int j = p ? 1 : 0; j < 3; j++)
          for (int d = -1; d <= 1; d += 2) {
            Mav1LrUnit t = w;
            t.wiener[pass][j] += d;
            if (t.wiener[pass][j] < Wiener_Taps_Min[j] || t.wiener[pass][j] > Wiener_Taps_Max[j]) continue;
            memcpy(tw, refW, sizeof(tw));
            memcpy(tx, refX, sizeof(tx));
            const int64_t tc = rd_cost(fp, unit_sse(&U, &t), 2 + unit_bits(p, &t, tw, tx));
            if (tc < wc) {
              wc = tc;
              w = t;
            }
          }
      if (wc < bestCost) {
        bestCost = wc;
        best = w;
      }
      for (int set = 0; set < (1 << SGRPROJ_PARAMS_BITS); set++) {
        const Mav1LrUnit sg = F->sg[set];
        memcpy(tw, refW, sizeof(tw));
        memcpy(tx, refX, sizeof(tx));
        const int64_t sc = rd_cost(fp, F->sg_sse[set], 2 + unit_bits(p, &sg, tw, tx));
        if (sc < bestCost) {
          bestCost = sc;
          best = sg;
        }
      }
      if (lrTest) {
        if (!(u & 1)) best = w;
        else {
          int64_t bs = -1;
          for (int set = 0; set < (1 << SGRPROJ_PARAMS_BITS); set++) {
            const int64_t v = F->sg_sse[set];
            if (bs < 0 || v < bs) {
              bs = v;
              best = F->sg[set];
            }
          }
        }
      }
      unit_bits(p, &best, refW, refX);
      units[fp->lr_unit_off[p] + u] = best;
      used |= best.type != RESTORE_NONE;
    }
    fp->lr_type[p] = used ? RESTORE_SWITCHABLE : RESTORE_NONE;
    free(fit);
    free(start);
    free(fill);
  }
  free(rc);
}
