/* Loop-restoration parameter search on gfx950 (behaviour of lr_select,
 * src/lr_search.c:235-359): per restoration unit NONE, a separable
 * least-squares Wiener fit refined greedily tap by tap, and the 16
 * self-guided sets with least-squares projection weights; the smallest
 * 16 * SSE + lambda * bits wins.
 *
 * The choice of a unit depends on the previous unit of its plane only
 * through the subexp references of the coded parameters (bit counts), so:
 *   k_lr_fit    one workgroup per unit, all units at once: SSE of NONE, the
 *               Wiener fit, each self-guided set's weights and SSE;
 *   k_lr_path   one workgroup per unit: the greedy tap refinement under an
 *               assumed incoming Wiener reference;
 *   k_lr_chain  one workgroup per plane walks its units in order with the
 *               true references, reruns a unit's refinement where the
 *               assumption was wrong, and writes the chosen parameters.
 * The Wiener fit is FP64 and its rounded result decides the coded taps: it
 * accumulates in the reference's order (block by block, sample by sample,
 * one lane per accumulator) with the reference's operand order, and this
 * file is compiled without floating-point contraction. Every other sum is
 * an exact integer. */
#include "dev_common.h"
#include "internal.h"

#include <climits>
#pragma clang fp contract(off)

#define LR_PITCH 104   /* unit width (<= 95) + 6 halo columns, rounded up */
#define LR_ROWS 128    /* unit rows (<= 103) + 6 halo rows per stripe segment (<= 4) */
#define LR_MAXSEG 4
#define LR_MAXPX (104 * 96)
#define LR_T 256       /* threads per workgroup */
#define LR_PH (4 * 95 + 8) /* a block row's samples (<= 4 rows x 95), + 8 read ahead by the ordered sums */

struct LrGeo {
  int p, sub, x0, x1, y0, y1, endX, endY;
  int nseg;
  int segY0[LR_MAXSEG], segY1[LR_MAXSEG], segS0[LR_MAXSEG], segS1[LR_MAXSEG], segBase[LR_MAXSEG];
  int bw, bh, nbr, nbc;
};

/* Working set of the workgroup searching one unit. */
struct LrLds {
  uint8_t img[LR_ROWS * LR_PITCH]; /* stripe-resolved source samples (get_source_sample) per segment */
  uint8_t srcu[LR_MAXPX];          /* the unit's samples of the source frame, row stride = unit width */
  union {
    struct {
      int16_t f0[LR_MAXPX], f1[LR_MAXPX]; /* self-guided filter outputs */
    } sg;
    struct {
      double h[2][1024];           /* a block row's 7-tap filtered samples (two rows in flight) */
      double ph[2][4][LR_PH];      /* a block row's fit terms, phi[0..2] then target, in sample order */
    } wf;
    struct {
      int16_t mid[2][LR_ROWS * 96]; /* Wiener horizontal pass: current taps, candidate taps */
    } wp;
  } u;
  long long red[LR_T / 64 * 12];
  double acc[12];
  int itmp[16];
  int xtab[256]; /* a of the self-guided filter by z (7.17.3) */
};
__shared__ LrLds gLr;

/* Per-unit search state between the kernels. */
struct LrUnitStat {
  long long sseNone;
  long long sseSgr[16];
  int xqd[16][2];
  int w0[6];          /* fitted Wiener taps: [pass * 3 + j] */
  int refAssumed[6];  /* incoming Wiener reference the path was computed with */
  int wStar[6];       /* refined taps */
  long long sseStar;  /* SSE of the refined taps */
  long long costStar; /* their cost under refAssumed */
  int fitGen;         /* generation of the launch whose fit w0 holds */
  int pad;
};

/* Per plane: where the unit chain stands between its launches. */
struct LrChain {
  int start;     /* next unit (plane index) to decide */
  int done;      /* all units decided */
  int used, nrerun;
  int ref[6];    /* running Wiener reference */
  int refX[2];   /* running self-guided reference */
};

DEVI int lr_stripe_row(int k, int sub) { return (64 * k - 8) >> sub; }

/* Pixel bounds of unit u and its stripe segments (the 4x4 blocks of
 * lr_block_setup, src/kernels/k_lr.h:28-45, whose unit is u). */
DEVI void lr_geo(const Mav1FrameParams &fp, int u, LrGeo &g) {
  const int p = u >= fp.lr_unit_off[2] ? 2 : u >= fp.lr_unit_off[1] ? 1 : 0;
  const int l = u - fp.lr_unit_off[p], ucols = fp.lr_unit_cols[p], urows = fp.lr_unit_rows[p];
  const int ur = l / ucols, uc = l - (l / ucols) * ucols, sub = p > 0;
  g.p = p;
  g.sub = sub;
  g.endX = ((fp.frame_w + sub) >> sub) - 1;
  g.endY = ((fp.frame_h + sub) >> sub) - 1;
  g.x0 = uc * 64;
  g.x1 = uc == ucols - 1 ? g.endX + 1 : d_min((uc + 1) * 64, g.endX + 1);
  const int off = 8 >> sub;
  g.y0 = d_max(0, ur * 64 - off);
  g.y1 = ur == urows - 1 ? g.endY + 1 : d_min((ur + 1) * 64 - off, g.endY + 1);
  g.bw = 4 >> sub;
  g.bh = 4 >> sub;
  g.nbr = (g.y1 - g.y0 + g.bh - 1) / g.bh;
  g.nbc = (g.x1 - g.x0 + g.bw - 1) / g.bw;
  /* unrolled, every index a constant: the arrays stay in registers (the
   * fields of segments past nseg are never read) */
  int k = ((g.y0 << sub) + 8) / 64, base = 0, y = g.y0;
  g.nseg = 0;
#pragma unroll
  for (int i = 0; i < LR_MAXSEG; i++, k++) {
    const int s0 = lr_stripe_row(k, sub), s1 = lr_stripe_row(k + 1, sub) - 1;
    const int a = d_max(y, s0), b = d_min(g.y1, s1 + 1);
    g.segY0[i] = a;
    g.segY1[i] = b;
    g.segS0[i] = s0;
    g.segS1[i] = s1;
    g.segBase[i] = base;
    if (y < g.y1) {
      base += (b - a + 6) * LR_PITCH;
      g.nseg++;
      y = b;
    }
  }
}

/* v[s] by selects over values, so that a run-time segment index does not
 * move the arrays to scratch memory (a select over the array's own
 * elements would be folded back into an indexed load). */
DEVI int lr_pick(int v0, int v1, int v2, int v3, int s) { return s == 0 ? v0 : s == 1 ? v1 : s == 2 ? v2 : v3; }
#define lr_sel(v, s) lr_pick((v)[0], (v)[1], (v)[2], (v)[3], (s))

DEVI int lr_seg_of(const LrGeo &g, int y) { return y < g.segY1[0] ? 0 : y < g.segY1[1] ? 1 : y < g.segY1[2] ? 2 : 3; }

/* Sample (x, y) of segment s, -3 <= x - x0 < w + 3, -3 <= y - segY0 < h + 3. */
DEVI int lr_img(const LrGeo &g, int s, int x, int y) {
  return gLr.img[lr_sel(g.segBase, s) + (y - lr_sel(g.segY0, s) + 3) * LR_PITCH + (x - g.x0 + 3)];
}

/* get_source_sample (7.17.6) of every sample the unit's filters read, and
 * the unit's source samples. */
DEVI void lr_stage(const Mav1FrameParams &fp, const LrGeo &g, const uint8_t *cdef, const uint8_t *dbk,
                   const uint8_t *src, int tid) {
  const int S = fp.stride[g.p], o = fp.offset[g.p], w = g.x1 - g.x0 + 6, uw = g.x1 - g.x0;
  for (int q = tid; q < uw * (g.y1 - g.y0); q += LR_T) {
    const int r = q / uw;
    gLr.srcu[q] = src[o + (g.y0 + r) * S + g.x0 + q - r * uw];
  }
  for (int s = 0; s < g.nseg; s++) {
    const int sy0 = lr_sel(g.segY0, s), ss0 = lr_sel(g.segS0, s), ss1 = lr_sel(g.segS1, s);
    const int h = lr_sel(g.segY1, s) - sy0 + 6, sb = lr_sel(g.segBase, s);
    for (int q = tid; q < w * h; q += LR_T) {
      const int r = q / w, c = q - r * w;
      const int x = d_clip(0, g.endX, g.x0 - 3 + c);
      int y = d_clip(0, g.endY, sy0 - 3 + r);
      const uint8_t *pl = cdef;
      if (y < ss0) {
        y = d_max(ss0 - 2, y);
        pl = dbk;
      } else if (y > ss1) {
        y = d_min(ss1 + 2, y);
        pl = dbk;
      }
      gLr.img[sb + r * LR_PITCH + c] = pl[o + y * S + x];
    }
  }
  __syncthreads();
}

/* Block-row / sample-order helpers: block row br holds h rows of nbc
 * blocks; sample q of the row is in block q / (h * bw) (the last block may
 * be narrower). */
DEVI void lr_row_sample(const LrGeo &g, int br, int q, int &x, int &y) {
  const int y0 = g.y0 + br * g.bh, h = d_min(g.bh, g.endY - y0 + 1);
  const int full = h * g.bw, j = q / full;
  const int bx = g.x0 + j * g.bw, w = d_min(g.bw, g.x1 - bx);
  const int rem = q - j * full;
  y = y0 + rem / w;
  x = bx + rem - (rem / w) * w;
}
DEVI int lr_row_count(const LrGeo &g, int br) {
  const int y0 = g.y0 + br * g.bh;
  return d_min(g.bh, g.endY - y0 + 1) * (g.x1 - g.x0);
}

DEVI long long block_sum64(long long v, int tid) {
  v = wave_sum64(v);
  if ((tid & 63) == 0) gLr.red[tid >> 6] = v;
  __syncthreads();
  long long s = 0;
#pragma unroll
  for (int i = 0; i < LR_T / 64; i++) s += gLr.red[i];
  __syncthreads();
  return s;
}

/* ---- subexp bit counts (lr_search.c:18-79) ---- */
DEVI int lr_recenter(int r, int v) {
  if (v > 2 * r) return v;
  return v >= r ? (v - r) << 1 : ((r - v) << 1) - 1;
}
DEVI int lr_subexp_bits(int low, int high, int k, int r, int v) {
  const int mx = high - low;
  r -= low;
  v -= low;
  const int x = 2 * r <= mx ? lr_recenter(r, v) : lr_recenter(mx - 1 - r, mx - 1 - v);
  int bits = 0, mk = 0;
  for (int i = 0;; i++) {
    const int b2 = i ? k + i - 1 : k, a = 1 << b2;
    if (mx <= mk + 3 * a) {
      const int n = mx - mk, w = d_floorlog2((unsigned)n) + 1, m = (1 << w) - n;
      return bits + (x - mk < m ? w - 1 : w);
    }
    bits++;
    if (x < mk + a) return bits + b2;
    mk += a;
  }
}
DEVI int lr_wiener_bits(int p, const int *w, const int *ref) {
  int bits = 0;
  for (int pass = 0; pass < 2; pass++)
    for (int j = p ? 1 : 0; j < 3; j++)
      bits += lr_subexp_bits(Wiener_Taps_Min[j], Wiener_Taps_Max[j] + 1, Wiener_Taps_K[j], ref[pass * 3 + j],
                             w[pass * 3 + j]);
  return bits;
}
DEVI int lr_sgr_bits(int set, const int *xqd, const int *refX) {
  int bits = SGRPROJ_PARAMS_BITS;
  for (int i = 0; i < 2; i++)
    if (SGR_PARAM(set, i * 2))
      bits += lr_subexp_bits(Sgrproj_Xqd_Min[i], Sgrproj_Xqd_Max[i] + 1, SGRPROJ_PRJ_SUBEXP_K, refX[i], xqd[i]);
  return bits;
}
DEVI long long lr_rd(const Mav1FrameParams &fp, long long sse, int bits) { return sse * 16 + (long long)fp.lambda * bits; }

/* ---- self-guided filter of one 4x4 block (7.17.3) from the staged samples ---- */
/* The filter's a by z: 1 at z = 0, 256 from z = 255, else round(256 z / (z + 1)). */
DEVI void lr_xtab(int tid) {
  for (int z = tid; z < 256; z += LR_T)
    gLr.xtab[z] = z == 0 ? 1 : z == 255 ? 256 : ((z << SGRPROJ_SGR_BITS) + z / 2) / (z + 1);
  __syncthreads();
}
/* A and B at the (BS+2)^2 positions (rows by-1 .. by+BS, columns bx-1 ..
 * bx+BS) around the BS x BS block at (bx, by) of segment s, box radius R
 * (blocks: 4x4 luma, 2x2 chroma), by sliding
 * window sums (integer, so any summation order gives the same values).
 * Positions outside a partial block read clamped samples; they are never
 * used. */
template <int R, int BS>
DEVI void lr_ab(const LrGeo &g, int s, int bx, int by, int sc, int one, int A[6][6], int B[6][6]) {
  constexpr int P = BS + 2, N = P + 2 * R, D = 2 * R + 1, n = D * D;
  const int xmax = g.x1 + 2, ymax = lr_sel(g.segY1, s) + 2;
  const uint8_t *img = gLr.img + lr_sel(g.segBase, s) + (3 - lr_sel(g.segY0, s)) * LR_PITCH + (3 - g.x0); /* img[y * P + x] */
  int xs[N], cs[N], cq[N];
#pragma unroll
  for (int c = 0; c < N; c++) {
    xs[c] = d_min(bx - 1 - R + c, xmax);
    cs[c] = cq[c] = 0;
  }
#pragma unroll
  for (int t = 0; t < D; t++) {
    const uint8_t *row = img + d_min(by - 1 - R + t, ymax) * LR_PITCH;
#pragma unroll
    for (int c = 0; c < N; c++) {
      const int v = row[xs[c]];
      cs[c] += v;
      cq[c] += v * v;
    }
  }
#pragma unroll
  for (int i = 0; i < P; i++) {
    if (i > 0) {
      const uint8_t *ro = img + d_min(by - 2 - R + i, ymax) * LR_PITCH;
      const uint8_t *rn = img + d_min(by - 2 + R + i + 1, ymax) * LR_PITCH;
#pragma unroll
      for (int c = 0; c < N; c++) {
        const int vo = ro[xs[c]], vn = rn[xs[c]];
        cs[c] += vn - vo;
        cq[c] += vn * vn - vo * vo;
      }
    }
    int sum = 0, sq = 0;
#pragma unroll
    for (int c = 0; c < D; c++) {
      sum += cs[c];
      sq += cq[c];
    }
#pragma unroll
    for (int j = 0; j < P; j++) {
      if (j > 0) {
        sum += cs[j + D - 1] - cs[j - 1];
        sq += cq[j + D - 1] - cq[j - 1];
      }
      const int pv = d_max(0, sq * n - sum * sum);
      const long long z = ((long long)pv * sc + (1 << (SGRPROJ_MTABLE_BITS - 1))) >> SGRPROJ_MTABLE_BITS;
      const int a2 = gLr.xtab[(int)d_min(255ll, z)];
      const int b2 = ((1 << SGRPROJ_SGR_BITS) - a2) * sum * one;
      A[i][j] = a2;
      B[i][j] = (b2 + (1 << (SGRPROJ_RECIP_BITS - 1))) >> SGRPROJ_RECIP_BITS;
    }
  }
}
/* The filtered BS x BS block (pass 0: radius 2, every other row of A / B;
 * pass 1: radius 1) into F (row stride 4). */
template <int PASS, int BS>
DEVI void lr_f(const LrGeo &g, int s, int bx, int by, const int A[6][6], const int B[6][6], int *F) {
  const uint8_t *img = gLr.img + lr_sel(g.segBase, s) + (3 - lr_sel(g.segY0, s)) * LR_PITCH + (3 - g.x0);
  const int xmax = g.x1 + 2, ymax = lr_sel(g.segY1, s) + 2;
#pragma unroll
  for (int i = 0; i < BS; i++) {
    const int sh = SGRPROJ_SGR_BITS + ((PASS == 0 && (i & 1)) ? 4 : 5) - SGRPROJ_RST_BITS;
#pragma unroll
    for (int j = 0; j < BS; j++) {
      int a = 0, bb = 0;
#pragma unroll
      for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          const int wgt = PASS == 0 ? (((i + dy) & 1) ? (dx == 0 ? 6 : 5) : 0) : (dx == 0 || dy == 0) ? 4 : 3;
          a += wgt * A[i + dy + 1][j + dx + 1];
          bb += wgt * B[i + dy + 1][j + dx + 1];
        }
      const int v = a * img[d_min(by + i, ymax) * LR_PITCH + d_min(bx + j, xmax)] + bb;
      F[i * 4 + j] = (v + (1 << (sh - 1))) >> sh;
    }
  }
}
/* Both passes of one block; a pass with radius 0 is off. */
template <int BS>
DEVI void lr_sgr_block(const LrGeo &g, int s, int bx, int by, int r0, int r1, const int *sc, const int *one, int *F0,
                       int *F1) {
  int A[6][6], B[6][6];
  if (r0) {
    lr_ab<2, BS>(g, s, bx, by, sc[0], one[0], A, B);
    lr_f<0, BS>(g, s, bx, by, A, B, F0);
  }
  if (r1) {
    lr_ab<1, BS>(g, s, bx, by, sc[1], one[1], A, B);
    lr_f<1, BS>(g, s, bx, by, A, B, F1);
  }
}

/* ---- Wiener SSE of a unit for taps w (pass-major) ---- */
/* Horizontal pass of taps hf into mid buffer m: every staged row, the
 * unit's columns. */
DEVI void lr_mid(const LrGeo &g, const int *hw, int m, int tid) {
  int hf[7];
  hf[3] = 128;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    hf[k] = hf[6 - k] = hw[k];
    hf[3] -= 2 * hw[k];
  }
  const int uw = g.x1 - g.x0;
  int rowBase = 0;
  for (int s = 0; s < g.nseg; s++) {
    const int rows = lr_sel(g.segY1, s) - lr_sel(g.segY0, s) + 6, sb = lr_sel(g.segBase, s);
    for (int q = tid; q < rows * uw; q += LR_T) {
      const int r = q / uw, c = q - r * uw;
      const uint8_t *ip = gLr.img + sb + r * LR_PITCH + c;
      int acc = 0;
#pragma unroll
      for (int t = 0; t < 7; t++) acc += hf[t] * ip[t];
      gLr.u.wp.mid[m][(rowBase + r) * 96 + c] = (int16_t)d_clip(-2048, 8191 - 2048, d_round2(acc, 3));
    }
    rowBase += rows;
  }
  __syncthreads();
}
/* Vertical pass of taps vw over mid buffer m and SSE against the source. */
DEVI long long lr_wiener_sse(const Mav1FrameParams &fp, const LrGeo &g, const int *vw, int m, const uint8_t *src,
                             int tid) {
  int vf[7];
  vf[3] = 128;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    vf[k] = vf[6 - k] = vw[k];
    vf[3] -= 2 * vw[k];
  }
  const int uw = g.x1 - g.x0, S = fp.stride[g.p], o = fp.offset[g.p];
  long long e = 0;
  int rowBase = 0;
  for (int s = 0; s < g.nseg; s++) {
    const int sy0 = lr_sel(g.segY0, s), h = lr_sel(g.segY1, s) - sy0;
    for (int q = tid; q < h * uw; q += LR_T) {
      const int r = q / uw, c = q - r * uw;
      const int16_t *mp = gLr.u.wp.mid[m] + (rowBase + r) * 96 + c;
      int acc = 0;
#pragma unroll
      for (int t = 0; t < 7; t++) acc += vf[t] * mp[t * 96];
      const int d = d_clip1(d_round2(acc, 11)) - (int)gLr.srcu[(sy0 + r - g.y0) * uw + c];
      e += d * d;
    }
    rowBase += h + 6;
  }
  return block_sum64(e, tid);
}

/* Greedy +-1 refinement of each coded tap (lr_search.c:283-301) from the
 * taps in w (updated in place) with incoming reference ref; returns the
 * final cost, *sseOut its SSE. Needs the staged samples. */
DEVI long long lr_refine(const Mav1FrameParams &fp, const LrGeo &g, int *w, const int *ref, const uint8_t *src,
                         long long *sseOut, int tid) {
  int cur = 0;
  lr_mid(g, w + 3, cur, tid);
  long long sse = lr_wiener_sse(fp, g, w, cur, src, tid);
  long long wc = lr_rd(fp, sse, 2 + lr_wiener_bits(g.p, w, ref));
  for (int pass = 0; pass < 2; pass++)
    for (int j = g.p ? 1 : 0; j < 3; j++)
      for (int d = -1; d <= 1; d += 2) {
        int t[6];
#pragma unroll
        for (int k = 0; k < 6; k++) t[k] = w[k];
        t[pass * 3 + j] += d;
        if (t[pass * 3 + j] < Wiener_Taps_Min[j] || t[pass * 3 + j] > Wiener_Taps_Max[j]) continue;
        long long ts;
        if (pass == 0) {
          ts = lr_wiener_sse(fp, g, t, cur, src, tid);
        } else {
          lr_mid(g, t + 3, cur ^ 1, tid);
          ts = lr_wiener_sse(fp, g, t, cur ^ 1, src, tid);
        }
        const long long tc = lr_rd(fp, ts, 2 + lr_wiener_bits(g.p, t, ref));
        if (tc < wc) {
          wc = tc;
          sse = ts;
#pragma unroll
          for (int k = 0; k < 6; k++) w[k] = t[k];
          if (pass == 1) cur ^= 1;
        }
      }
  *sseOut = sse;
  return wc;
}

/* Gaussian elimination with partial pivoting (lr_search.c:98-129). */
DEVI int lr_solve(int n, double M[3][3], double v[3], double *x) {
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

/* Separable least-squares Wiener fit (lr_search.c:132-180): four
 * alternating solves; the sums run over the unit's samples in the
 * reference's order, one lane per accumulator. */
/* The 7-tap sums of block row br under filter f, each in the reference's
 * order (t = 0 .. 6): vert, H(x, yy) = sum_t f[t] img(x + t - 3, yy) for the
 * row's samples' rows -3 .. +3 (row stride uw); else V(xx, y) = sum_t f[t]
 * img(xx, y + t - 3) for their columns -3 .. +3 (row stride uw + 6). */
DEVI void lr_wf_filter(const LrGeo &g, int br, int vert, const double *f, double *h, int t0, int nt) {
  const int y0 = g.y0 + br * g.bh, hh = d_min(g.bh, g.endY - y0 + 1), uw = g.x1 - g.x0, s = lr_seg_of(g, y0);
  const int n = vert ? (hh + 6) * uw : hh * (uw + 6), rw = vert ? uw : uw + 6;
  for (int k = t0; k < n; k += nt) {
    const int r = k / rw, c = k - r * rw;
    double sum = 0;
    if (vert) {
#pragma unroll
      for (int t = 0; t < 7; t++) sum += f[t] * (double)lr_img(g, s, g.x0 + c + t - 3, y0 - 3 + r);
    } else {
#pragma unroll
      for (int t = 0; t < 7; t++) sum += f[t] * (double)lr_img(g, s, g.x0 - 3 + c, y0 + r + t - 3);
    }
    h[k] = sum;
  }
}
/* The fit terms of block row br's samples in the reference's sample order
 * (lr_search.c:150-160): phi[a - first] = Z[a] + Z[6 - a] - 2 Z[3] and the
 * target src - Z[3], Z from the row's 7-tap sums. */
DEVI void lr_wf_phi(const LrGeo &g, int br, int vert, int first, const double *h, double (*ph)[LR_PH], int t0,
                    int nt) {
  const int y0 = g.y0 + br * g.bh, uw = g.x1 - g.x0, cnt = lr_row_count(g, br);
  for (int q = t0; q < cnt; q += nt) {
    int x, y;
    lr_row_sample(g, br, q, x, y);
    double Z[7];
#pragma unroll
    for (int a = 0; a < 7; a++)
      Z[a] = vert ? h[(y - y0 + a) * uw + x - g.x0] : h[(y - y0) * (uw + 6) + x - g.x0 + a];
    for (int a = first; a < 3; a++) ph[a - first][q] = Z[a] + Z[6 - a] - 2 * Z[3];
    ph[3][q] = (double)gLr.srcu[(y - g.y0) * uw + x - g.x0] - Z[3];
  }
}

/* The first solve's sums (a vertical fit under the identity horizontal
 * filter): every fit term is then an integer (|phi| <= 510, |target| <=
 * 255), so every product and partial sum of the reference's FP64
 * accumulation is an integer below 2^53 and exact, whatever the order.
 * All lanes sum their samples in int32 (<= 39 samples x 510^2 each), the
 * waves in int64; the NC * NC + NC sums land in gLr.acc as doubles. */
template <int NC>
DEVI void lr_wf_exact(const LrGeo &g, int tid) {
  constexpr int FIRST = 3 - NC, NACC = NC * NC + NC;
  const int uw = g.x1 - g.x0, n = uw * (g.y1 - g.y0);
  int sum[NACC];
#pragma unroll
  for (int i = 0; i < NACC; i++) sum[i] = 0;
  for (int q = tid; q < n; q += LR_T) {
    const int r = q / uw, y = g.y0 + r, x = g.x0 + q - r * uw, s = lr_seg_of(g, y);
    int Z[7];
#pragma unroll
    for (int a = 0; a < 7; a++) Z[a] = lr_img(g, s, x, y + a - 3);
    int ph[NC + 1];
#pragma unroll
    for (int a = FIRST; a < 3; a++) ph[a - FIRST] = Z[a] + Z[6 - a] - 2 * Z[3];
    ph[NC] = (int)gLr.srcu[q] - Z[3];
#pragma unroll
    for (int a = 0; a < NC; a++) {
#pragma unroll
      for (int b = 0; b < NC; b++) sum[a * NC + b] += ph[a] * ph[b];
      sum[NC * NC + a] += ph[a] * ph[NC];
    }
  }
  const int wave = tid >> 6;
#pragma unroll
  for (int i = 0; i < NACC; i++) {
    const int v = wave_sum(sum[i]);
    if ((tid & 63) == 0) gLr.red[wave * 12 + i] = v;
  }
  __syncthreads();
  if (tid < NACC) {
    long long t = 0;
#pragma unroll
    for (int w = 0; w < LR_T / 64; w++) t += gLr.red[w * 12 + tid];
    gLr.acc[tid] = (double)t;
  }
}

/* Separable least-squares Wiener fit (lr_search.c:132-180): four
 * alternating solves. The first solve's sums are exact integers
 * (lr_wf_exact). Each later solve's sums run over the unit's samples in
 * the reference's order, one lane of the last wave per accumulator; the
 * other waves prepare the fit terms two block rows ahead (filter a row,
 * then its terms, then its sums: a three-step pipeline with one barrier a
 * step). */
DEVI void lr_wiener_fit(const Mav1FrameParams &fp, const LrGeo &g, const uint8_t *src, int *out, int tid) {
  (void)fp;
  (void)src;
  const int first = g.p ? 1 : 0, nc = 3 - first, nacc = nc * nc + nc;
  const int wave = tid >> 6, lane = tid & 63, NCT = LR_T - 64;
  double fv[7] = {0, 0, 0, 1, 0, 0, 0}, fh[7] = {0, 0, 0, 1, 0, 0, 0};
  for (int it = 0; it < 4; it++) {
    const int vert = !(it & 1);
    if (it == 0) {
      if (nc == 3) lr_wf_exact<3>(g, tid);
      else lr_wf_exact<2>(g, tid);
    } else {
      double f[7];
#pragma unroll
      for (int t = 0; t < 7; t++) f[t] = vert ? fh[t] : fv[t];
      double acc = 0; /* lane < nacc of the last wave: M[a][b] (a * nc + b), then V[a] */
      for (int st = 0; st < g.nbr + 2; st++) {
        if (wave < LR_T / 64 - 1) {
          if (st < g.nbr) lr_wf_filter(g, st, vert, f, gLr.u.wf.h[st & 1], tid, NCT);
          if (st >= 1 && st - 1 < g.nbr)
            lr_wf_phi(g, st - 1, vert, first, gLr.u.wf.h[(st - 1) & 1], gLr.u.wf.ph[(st - 1) & 1], tid, NCT);
        } else if (st >= 2 && lane < nacc) {
          const int br = st - 2, n = lr_row_count(g, br), n16 = n & ~15;
          const int a = lane < nc * nc ? lane / nc : lane - nc * nc, b = lane < nc * nc ? lane - (lane / nc) * nc : 3;
          typedef __attribute__((address_space(3))) const double lds_f64;
          const lds_f64 *pa = (lds_f64 *)gLr.u.wf.ph[br & 1][a], *pb = (lds_f64 *)gLr.u.wf.ph[br & 1][b];
          /* in order, sample by sample, in batches of 8 that alternate
           * between two register sets: one batch adds while the next loads.
           * The scheduling barriers keep each batch's loads ahead of the
           * previous batch's adds, and the empty asm statements make each
           * set opaque once its batch is needed (the compiler would
           * otherwise move the loads back next to their use). */
          double xa[8], xb[8], ya[8], yb[8];
#pragma unroll
          for (int j = 0; j < 8; j++) {
            xa[j] = pa[j];
            xb[j] = pb[j];
          }
          for (int k = 0; k < n16; k += 16) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
              ya[j] = pa[k + 8 + j];
              yb[j] = pb[k + 8 + j];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < 8; j++) acc += xa[j] * xb[j];
#pragma unroll
            for (int j = 0; j < 8; j++) asm volatile("" : "+v"(ya[j]), "+v"(yb[j]));
#pragma unroll
            for (int j = 0; j < 8; j++) {
              xa[j] = pa[k + 16 + j];
              xb[j] = pb[k + 16 + j];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < 8; j++) acc += ya[j] * yb[j];
#pragma unroll
            for (int j = 0; j < 8; j++) asm volatile("" : "+v"(xa[j]), "+v"(xb[j]));
          }
          for (int k = n16; k < n; k++) acc += pa[k] * pb[k];
        }
        __syncthreads();
      }
      if (wave == LR_T / 64 - 1 && lane < nacc) gLr.acc[lane] = acc;
    }
    __syncthreads();
    if (tid == 0) {
      double M[3][3], V[3], x[3] = {0, 0, 0};
      for (int a = 0; a < nc; a++) {
        V[a] = gLr.acc[nc * nc + a];
        for (int b = 0; b < nc; b++) M[a][b] = gLr.acc[a * nc + b];
      }
      for (int a = 0; a < nc; a++) M[a][a] += 1e-3;
      const int ok = lr_solve(nc, M, V, x);
      gLr.itmp[0] = ok;
      for (int a = 0; a < 3; a++) gLr.acc[a] = a < first ? 0 : x[a - first];
    }
    __syncthreads();
    if (!gLr.itmp[0]) break;
    double sum = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double v = gLr.acc[a];
      if (vert) fv[a] = fv[6 - a] = v;
      else fh[a] = fh[6 - a] = v;
      sum += v;
    }
    if (vert) fv[3] = 1 - 2 * sum;
    else fh[3] = 1 - 2 * sum;
    __syncthreads();
  }
#pragma unroll
  for (int pass = 0; pass < 2; pass++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      out[pass * 3 + j] =
          j < first ? 0 : d_clip(Wiener_Taps_Min[j], Wiener_Taps_Max[j], (int)lround((pass ? fh[j] : fv[j]) * 128));
}

/* Per unit: SSE of NONE, the Wiener fit, and each self-guided set's
 * weights (sgr_fit, lr_search.c:183-231) and SSE. Four workgroups per unit
 * (blockIdx.y): NONE and the Wiener fit, the sets 0-7, the sets 8-15, and
 * the greedy refinement of the fitted taps under the default reference
 * (what k_lr_path computes before the first chain launch). The refinement
 * waits for its unit's fit, published with generation `gen`; every fit
 * workgroup is dispatched before any refinement workgroup (blockIdx.y
 * order) and waits for nothing, so the wait ends. Should it run out after
 * maxSpin polls, the refinement records an assumed reference no chain
 * holds, and the chain refines the unit again: the result never depends on
 * the wait. */
DEVI void lr_fit_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, const uint8_t *__restrict__ cdef, const uint8_t *__restrict__ dbk, LrUnitStat *__restrict__ stats, int gen, int maxSpin, const int blk_x, const int blk_y, const int blk_z) {
  const int u = blk_x, part = blk_y, tid = threadIdx.x;
  LrGeo g;
  lr_geo(fp, u, g);
  lr_stage(fp, g, cdef, dbk, src, tid);
  LrUnitStat &st = stats[u];
  if (part == 3) {
    if (tid == 0) {
      int ok = 0;
      for (int it = 0; it < maxSpin && !ok; it++) {
        ok = __hip_atomic_load(&st.fitGen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen;
        if (!ok) __builtin_amdgcn_s_sleep(4);
      }
      gLr.itmp[0] = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int ok = gLr.itmp[0];
    int w[6], ref[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      w[k] = st.w0[k];
      ref[k] = Wiener_Taps_Mid[k % 3];
    }
    long long sse = 0, cost = 0;
    if (ok) cost = lr_refine(fp, g, w, ref, src, &sse, tid);
    if (tid == 0) {
      for (int k = 0; k < 6; k++) {
        st.wStar[k] = w[k];
        st.refAssumed[k] = ok ? ref[k] : INT_MIN;
      }
      st.sseStar = sse;
      st.costStar = cost;
    }
    return;
  }
  const int S = fp.stride[g.p], o = fp.offset[g.p], uw = g.x1 - g.x0, uh = g.y1 - g.y0;
  if (part == 0) {
    /* NONE and the Wiener fit */
    long long e = 0;
    for (int q = tid; q < uw * uh; q += LR_T) {
      const int a = o + (g.y0 + q / uw) * S + g.x0 + q % uw;
      const int d = (int)cdef[a] - (int)gLr.srcu[q];
      e += d * d;
    }
    e = block_sum64(e, tid);
    if (tid == 0) st.sseNone = e;
    int w0[6];
    lr_wiener_fit(fp, g, src, w0, tid);
    if (tid < 6) st.w0[tid] = w0[tid];
    /* publish the taps to this unit's refinement workgroup */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&st.fitGen, gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  /* half of the self-guided sets */
  lr_xtab(tid);
  const int nblk = g.nbr * g.nbc;
  const int r_t = tid / uw, c_t = tid - r_t * uw, dr_t = LR_T / uw, dc_t = LR_T - dr_t * uw;
  for (int set = (part - 1) * 8; set < part * 8; set++) {
    /* the radii are 2 (pass 0) and 1 (pass 1), or 0: pass off */
    const int r0 = SGR_PARAM(set, 0), r1 = SGR_PARAM(set, 2);
    int sc[2], one[2];
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
      const int n = pass ? 9 : 25, n2e = n * n * SGR_PARAM(set, pass * 2 + 1);
      sc[pass] = n2e ? ((1 << SGRPROJ_MTABLE_BITS) + n2e / 2) / n2e : 0;
      one[pass] = ((1 << SGRPROJ_RECIP_BITS) + n / 2) / n;
    }
    long long h00 = 0, h01 = 0, h11 = 0, c0 = 0, c1 = 0;
    for (int bi = tid; bi < nblk; bi += LR_T) {
      const int br = bi / g.nbc, bc = bi - br * g.nbc;
      const int by = g.y0 + br * g.bh, bx = g.x0 + bc * g.bw;
      const int h = d_min(g.bh, g.endY - by + 1), w = d_min(g.bw, g.x1 - bx), s = lr_seg_of(g, by);
      int F0[16], F1[16];
      if (g.sub) lr_sgr_block<2>(g, s, bx, by, r0, r1, sc, one, F0, F1);
      else lr_sgr_block<4>(g, s, bx, by, r0, r1, sc, one, F0, F1);
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (i >= h || j >= w) continue;
          const int uu = lr_img(g, s, bx + j, by + i) << SGRPROJ_RST_BITS;
          const int sv = ((int)gLr.srcu[(by + i - g.y0) * uw + bx + j - g.x0] << SGRPROJ_RST_BITS) - uu;
          const int e0 = r0 ? F0[i * 4 + j] - uu : 0, e1 = r1 ? F1[i * 4 + j] - uu : 0;
          h00 += (long long)e0 * e0;
          h01 += (long long)e0 * e1;
          h11 += (long long)e1 * e1;
          c0 += (long long)e0 * sv;
          c1 += (long long)e1 * sv;
          const int pi = (by + i - g.y0) * uw + bx + j - g.x0;
          gLr.u.sg.f0[pi] = (int16_t)(r0 ? F0[i * 4 + j] : 0);
          gLr.u.sg.f1[pi] = (int16_t)(r1 ? F1[i * 4 + j] : 0);
        }
    }
    /* the five block sums in one exchange: each wave's sums into red, one
     * barrier, lane 0 of the workgroup adds them (red is read again only
     * after the barrier below) */
    {
      const long long part[5] = {h00, h01, h11, c0, c1};
#pragma unroll
      for (int i = 0; i < 5; i++) {
        const long long v = wave_sum64(part[i]);
        if ((tid & 63) == 0) gLr.red[(tid >> 6) * 5 + i] = v;
      }
    }
    __syncthreads();
    if (tid == 0) {
      long long tot[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int w = 0; w < LR_T / 64; w++)
#pragma unroll
        for (int i = 0; i < 5; i++) tot[i] += gLr.red[w * 5 + i];
      h00 = tot[0];
      h01 = tot[1];
      h11 = tot[2];
      c0 = tot[3];
      c1 = tot[4];
      /* the reference's double sums of integer terms are exact: the same
       * values as these integer sums */
      const double H00 = (double)h00, H01 = (double)h01, H11 = (double)h11, C0 = (double)c0, C1 = (double)c1;
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
      const int w0q = r0 ? (int)lround(a * (1 << SGRPROJ_PRJ_BITS)) : 0, w2q = (int)lround(bb * (1 << SGRPROJ_PRJ_BITS));
      const int x0 = d_clip(Sgrproj_Xqd_Min[0], Sgrproj_Xqd_Max[0], w0q);
      const int x1 = d_clip(Sgrproj_Xqd_Min[1], Sgrproj_Xqd_Max[1], (1 << SGRPROJ_PRJ_BITS) - x0 - (r1 ? w2q : 0));
      gLr.itmp[0] = x0;
      gLr.itmp[1] = x1;
      st.xqd[set][0] = x0;
      st.xqd[set][1] = x1;
    }
    __syncthreads();
    const int xw0 = gLr.itmp[0], xw1 = gLr.itmp[1], xw2 = (1 << SGRPROJ_PRJ_BITS) - xw0 - xw1;
    /* each lane's samples step by LR_T in row-major order (row r, column
     * c, no division in the loop); at most 39 squared differences of
     * 8-bit samples per lane, so the lane's sum fits an int */
    int es = 0;
    for (int q = tid, r = r_t, c = c_t; q < uw * uh; q += LR_T) {
      const int y = g.y0 + r, x = g.x0 + c;
      const int uu = lr_img(g, lr_seg_of(g, y), x, y) << SGRPROJ_RST_BITS;
      const int v = xw1 * uu + (r0 ? xw0 * gLr.u.sg.f0[q] : xw0 * uu) + (r1 ? xw2 * gLr.u.sg.f1[q] : xw2 * uu);
      const int d = d_clip1(d_round2(v, SGRPROJ_RST_BITS + SGRPROJ_PRJ_BITS)) - (int)gLr.srcu[q];
      es += d * d;
      c += dc_t;
      r += dr_t;
      if (c >= uw) {
        c -= uw;
        r++;
      }
    }
    const long long ess = block_sum64(es, tid);
    if (tid == 0) st.sseSgr[set] = ess;
    __syncthreads();
  }
}
__global__ __launch_bounds__(LR_T) void k_lr_fit(const Mav1FrameParams fp, const uint8_t *__restrict__ src,
                                                 const uint8_t *__restrict__ cdef, const uint8_t *__restrict__ dbk,
                                                 LrUnitStat *__restrict__ stats, int gen, int maxSpin) {
  lr_fit_body(fp, src, cdef, dbk, stats, gen, maxSpin, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* Per unit: the greedy refinement of the fitted taps under the plane's
 * running reference, for the units from the chain's restart point on (all
 * of them in the first pass, where the reference is the default). */
DEVI void lr_path_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, const uint8_t *__restrict__ cdef, const uint8_t *__restrict__ dbk, LrUnitStat *__restrict__ stats, const LrChain *__restrict__ chain, const int blk_x, const int blk_y, const int blk_z) {
  const int u = blk_x, tid = threadIdx.x;
  const int p = u < fp.lr_unit_off[1] ? 0 : u < fp.lr_unit_off[2] ? 1 : 2;
  const LrChain &cs = chain[p];
  if (cs.done || u - fp.lr_unit_off[p] < cs.start) return;
  LrGeo g;
  lr_geo(fp, u, g);
  lr_stage(fp, g, cdef, dbk, src, tid);
  LrUnitStat &st = stats[u];
  int w[6], ref[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    w[k] = st.w0[k];
    ref[k] = cs.ref[k];
  }
  long long sse;
  const long long cost = lr_refine(fp, g, w, ref, src, &sse, tid);
  if (tid == 0) {
    for (int k = 0; k < 6; k++) {
      st.wStar[k] = w[k];
      st.refAssumed[k] = ref[k];
    }
    st.sseStar = sse;
    st.costStar = cost;
  }
}
__global__ __launch_bounds__(LR_T) void k_lr_path(const Mav1FrameParams fp, const uint8_t *__restrict__ src,
                                                  const uint8_t *__restrict__ cdef, const uint8_t *__restrict__ dbk,
                                                  LrUnitStat *__restrict__ stats, const LrChain *__restrict__ chain) {
  lr_path_body(fp, src, cdef, dbk, stats, chain, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* Per plane, units in raster order with the true running references: the
 * choice of lr_select (lr_search.c:262-349). A unit whose refinement
 * k_lr_path ran under another reference than the running one either ends
 * the launch (stop != 0: the chain records where it stands, and the next
 * k_lr_path pass refines the remaining units under the running reference
 * in parallel) or is refined again here (stop == 0, the last launch).
 * units[] and lrType[plane] receive the result, reruns[plane] the number of
 * refinements rerun here. The unit statistics are staged into LDS LR_CB
 * units at a time; the sixteen self-guided costs of a unit are scored one
 * per lane. */
#define LR_CB 32

/* The unit record of a choice: kind (RESTORE_*), self-guided set. */
DEVI Mav1LrUnit lr_unit_of(const LrUnitStat &st, int kind, int set, const int *w) {
  Mav1LrUnit un = {};
  un.type = kind;
  if (kind == RESTORE_WIENER)
    for (int k = 0; k < 6; k++) un.wiener[k / 3][k % 3] = w[k];
  if (kind == RESTORE_SGRPROJ) {
    un.set = set;
    un.xqd[0] = st.xqd[set][0];
    un.xqd[1] = st.xqd[set][1];
  }
  return un;
}

/* Speculative choices of staged units li .. nb-1 under the current
 * references sRef / sRefX: sMis[u] = the unit's refinement assumed another
 * Wiener reference, sDec[u] = kind | set << 4 (NONE, then WIENER, then the
 * sets in order, first strict minimum; the lr_test pattern). Sixteen lanes
 * per unit, one per self-guided set. */
DEVI void lr_spec(const Mav1FrameParams &fp, const LrUnitStat *sSt, int li, int nb, int b0, int lrTest,
                  const int *sRef, const int *sRefX, int *sDec, int *sMis, int tid) {
  const int set = tid & 15;
  const int has0 = SGR_PARAM(set, 0) != 0, has1 = SGR_PARAM(set, 2) != 0;
  const int xlo0 = Sgrproj_Xqd_Min[0], xhi0 = Sgrproj_Xqd_Max[0] + 1;
  const int xlo1 = Sgrproj_Xqd_Min[1], xhi1 = Sgrproj_Xqd_Max[1] + 1;
  for (int q = tid; q < (nb - li) * 16; q += LR_T) {
    const int u = li + (q >> 4);
    const LrUnitStat &st = sSt[u];
    int bits = 2 + SGRPROJ_PARAMS_BITS;
    if (has0) bits += lr_subexp_bits(xlo0, xhi0, SGRPROJ_PRJ_SUBEXP_K, sRefX[0], st.xqd[set][0]);
    if (has1) bits += lr_subexp_bits(xlo1, xhi1, SGRPROJ_PRJ_SUBEXP_K, sRefX[1], st.xqd[set][1]);
    long long sc = lr_rd(fp, st.sseSgr[set], bits);
    int si = set;
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) {
      const long long oc = __shfl_xor(sc, m);
      const int oi = __shfl_xor(si, m);
      if (oc < sc || (oc == sc && oi < si)) {
        sc = oc;
        si = oi;
      }
    }
    if (set == 0) {
      int mis = 0;
      for (int k = 0; k < 6; k++) mis |= st.refAssumed[k] != sRef[k];
      long long best = lr_rd(fp, st.sseNone, 1);
      int kind = RESTORE_NONE, bs = 0;
      if (st.costStar < best) {
        best = st.costStar;
        kind = RESTORE_WIENER;
      }
      if (sc < best) {
        kind = RESTORE_SGRPROJ;
        bs = si;
      }
      if (lrTest) {
        if (!((b0 + u) & 1)) {
          kind = RESTORE_WIENER;
        } else {
          long long e = -1;
          kind = RESTORE_SGRPROJ;
          for (int k = 0; k < 16; k++)
            if (e < 0 || st.sseSgr[k] < e) {
              e = st.sseSgr[k];
              bs = k;
            }
        }
      }
      sDec[u] = kind | bs << 4;
      sMis[u] = mis;
    }
  }
}

DEVI void lr_chain_body(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, const uint8_t *__restrict__ cdef, const uint8_t *__restrict__ dbk, const LrUnitStat *__restrict__ stats, Mav1LrUnit *__restrict__ units, int *__restrict__ lrType, int lrTest, int *__restrict__ reruns, LrChain *__restrict__ chain, int stop, const int blk_x, const int blk_y, const int blk_z) {
  const int p = blk_x, tid = threadIdx.x;
  LrChain &cs = chain[p];
  if (cs.done) return;
  __shared__ LrUnitStat sSt[LR_CB];
  __shared__ Mav1LrUnit sUn[LR_CB];
  __shared__ int sRef[6], sRefX[2], sW[6];
  __shared__ long long sCost, sSse;
  __shared__ int sRerun, sStop, sDec[LR_CB], sMis[LR_CB], sNext, sReason;
  if (tid < 6) sRef[tid] = cs.ref[tid];
  if (tid < 2) sRefX[tid] = cs.refX[tid];
  if (tid == 0) sStop = -1;
  const int nu = fp.lr_unit_rows[p] * fp.lr_unit_cols[p], off = fp.lr_unit_off[p];
  int used = cs.used, nrerun = cs.nrerun;
  /* lane `set` (< 16): which of its passes are on; the projection ranges */
  const int has0 = tid < 16 && SGR_PARAM(tid, 0) != 0, has1 = tid < 16 && SGR_PARAM(tid, 2) != 0;
  const int xlo0 = Sgrproj_Xqd_Min[0], xhi0 = Sgrproj_Xqd_Max[0] + 1;
  const int xlo1 = Sgrproj_Xqd_Min[1], xhi1 = Sgrproj_Xqd_Max[1] + 1;
  for (int b0 = cs.start; b0 < nu; b0 += LR_CB) {
    int nb = d_min(LR_CB, nu - b0);
    const int words = nb * (int)(sizeof(LrUnitStat) / 4);
    __syncthreads();
    for (int q = tid; q < words; q += LR_T) ((int *)sSt)[q] = ((const int *)(stats + off + b0))[q];
    __syncthreads();
    for (int li = 0; li < nb; li++) {
      /* the units from li on under the current references, accepted up to
       * the first that changes a reference or whose refinement assumed
       * another one */
      lr_spec(fp, sSt, li, nb, b0, lrTest, sRef, sRefX, sDec, sMis, tid);
      __syncthreads();
      if (tid == 0) {
        int k = li, reason = 0;
        for (; k < nb; k++) {
          if (sMis[k]) {
            reason = 2;
            break;
          }
          const LrUnitStat &st = sSt[k];
          const int kind = sDec[k] & 15, set = sDec[k] >> 4;
          int change = 0;
          if (kind == RESTORE_WIENER)
            for (int j = 0; j < 6; j++)
              if (p == 0 || j % 3) {
                change |= sRef[j] != st.wStar[j];
                sRef[j] = st.wStar[j];
              }
          if (kind == RESTORE_SGRPROJ)
            for (int j = 0; j < 2; j++) {
              change |= sRefX[j] != st.xqd[set][j];
              sRefX[j] = st.xqd[set][j];
            }
          sUn[k] = lr_unit_of(st, kind, set, st.wStar);
          used |= kind != RESTORE_NONE;
          if (change) {
            k++;
            reason = 1;
            break;
          }
        }
        sNext = k;
        sReason = reason;
      }
      __syncthreads();
      li = sNext;
      if (sReason != 2) {
        li--; /* the loop's increment resumes at sNext */
        continue;
      }
      /* unit li: its refinement assumed another reference */
      const int l = b0 + li, u = off + l;
      const LrUnitStat &st = sSt[li];
      if (tid == 0) {
        int same = 1;
        for (int k = 0; k < 6; k++) same &= st.refAssumed[k] == sRef[k];
        sRerun = !same;
        if (!same && stop) sStop = li;
        for (int k = 0; k < 6; k++) sW[k] = st.wStar[k];
        sCost = st.costStar;
        sSse = st.sseStar;
      }
      __syncthreads();
      if (sStop >= 0) {
        nb = li;
        break;
      }
      if (sRerun) {
        LrGeo g;
        lr_geo(fp, u, g);
        lr_stage(fp, g, cdef, dbk, src, tid);
        int w[6], ref[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
          w[k] = st.w0[k];
          ref[k] = sRef[k];
        }
        long long sse;
        const long long cost = lr_refine(fp, g, w, ref, src, &sse, tid);
        if (tid == 0) {
          for (int k = 0; k < 6; k++) sW[k] = w[k];
          sCost = cost;
          sSse = sse;
        }
        nrerun++;
        __syncthreads();
      }
      if (tid < 64) {
        /* NONE, then WIENER, then the sets in order: first strict minimum */
        long long bestCost = lr_rd(fp, st.sseNone, 1);
        int bestKind = RESTORE_NONE, bestSet = 0;
        if (sCost < bestCost) {
          bestCost = sCost;
          bestKind = RESTORE_WIENER;
        }
        long long sc = LLONG_MAX;
        if (tid < 16) {
          int bits = 2 + SGRPROJ_PARAMS_BITS;
          if (has0) bits += lr_subexp_bits(xlo0, xhi0, SGRPROJ_PRJ_SUBEXP_K, sRefX[0], st.xqd[tid][0]);
          if (has1) bits += lr_subexp_bits(xlo1, xhi1, SGRPROJ_PRJ_SUBEXP_K, sRefX[1], st.xqd[tid][1]);
          sc = lr_rd(fp, st.sseSgr[tid], bits);
        }
        int si = tid;
#pragma unroll
        for (int m = 8; m >= 1; m >>= 1) {
          const long long oc = __shfl_xor(sc, m);
          const int oi = __shfl_xor(si, m);
          if (oc < sc || (oc == sc && oi < si)) {
            sc = oc;
            si = oi;
          }
        }
        if (sc < bestCost) {
          bestKind = RESTORE_SGRPROJ;
          bestSet = si;
        }
        if (tid == 0) {
          Mav1LrUnit wien = {};
          wien.type = RESTORE_WIENER;
          for (int k = 0; k < 6; k++) wien.wiener[k / 3][k % 3] = sW[k];
          Mav1LrUnit best = {};
          best.type = RESTORE_NONE;
          if (bestKind == RESTORE_WIENER) best = wien;
          if (bestKind == RESTORE_SGRPROJ) {
            best.type = RESTORE_SGRPROJ;
            best.set = bestSet;
            best.xqd[0] = st.xqd[bestSet][0];
            best.xqd[1] = st.xqd[bestSet][1];
          }
          if (lrTest) {
            if (!(l & 1)) best = wien;
            else {
              long long bs = -1;
              for (int set = 0; set < 16; set++)
                if (bs < 0 || st.sseSgr[set] < bs) {
                  bs = st.sseSgr[set];
                  best = Mav1LrUnit{};
                  best.type = RESTORE_SGRPROJ;
                  best.set = set;
                  best.xqd[0] = st.xqd[set][0];
                  best.xqd[1] = st.xqd[set][1];
                }
            }
          }
          if (best.type == RESTORE_WIENER)
            for (int k = 0; k < 6; k++)
              if (p == 0 || k % 3) sRef[k] = best.wiener[k / 3][k % 3];
          if (best.type == RESTORE_SGRPROJ)
            for (int i = 0; i < 2; i++) sRefX[i] = best.xqd[i];
          sUn[li] = best;
          used |= best.type != RESTORE_NONE;
        }
      }
      __syncthreads();
    }
    for (int q = tid; q < nb * (int)(sizeof(Mav1LrUnit) / 4); q += LR_T)
      ((int *)(units + off + b0))[q] = ((const int *)sUn)[q];
    if (sStop >= 0) {
      /* the state for the refinement pass and the next launch */
      if (tid == 0) {
        cs.start = b0 + sStop;
        cs.used = used;
        cs.nrerun = nrerun;
        for (int k = 0; k < 6; k++) cs.ref[k] = sRef[k];
        for (int k = 0; k < 2; k++) cs.refX[k] = sRefX[k];
      }
      return;
    }
  }
  if (tid == 0) {
    cs.done = 1;
    lrType[p] = used ? RESTORE_SWITCHABLE : RESTORE_NONE;
    reruns[p] = nrerun;
  }
}
__global__ __launch_bounds__(LR_T) void k_lr_chain(const Mav1FrameParams fp, const uint8_t *__restrict__ src,
                                                   const uint8_t *__restrict__ cdef, const uint8_t *__restrict__ dbk,
                                                   const LrUnitStat *__restrict__ stats, Mav1LrUnit *__restrict__ units,
                                                   int *__restrict__ lrType, int lrTest, int *__restrict__ reruns,
                                                   LrChain *__restrict__ chain, int stop) {
  lr_chain_body(fp, src, cdef, dbk, stats, units, lrType, lrTest, reruns, chain, stop, blockIdx.x, blockIdx.y, blockIdx.z);
}

__global__ void k_lr_chain_init(LrChain *__restrict__ chain) {
  const int p = threadIdx.x;
  if (p >= 3) return;
  LrChain c = {};
  for (int k = 0; k < 6; k++) c.ref[k] = Wiener_Taps_Mid[k % 3];
  for (int k = 0; k < 2; k++) c.refX[k] = Sgrproj_Xqd_Mid[k];
  chain[p] = c;
}

static_assert(sizeof(LrUnitStat) <= LR_STAT_MAX_BYTES, "LR_STAT_MAX_BYTES too small");
static_assert(3 * sizeof(LrChain) <= LR_CHAIN_MAX_BYTES, "LR_CHAIN_MAX_BYTES too small");
size_t lr_stat_bytes() { return sizeof(LrUnitStat); }
size_t lr_chain_bytes() { return 3 * sizeof(LrChain); }

hipError_t launch_lr_fit(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdef,
                         const uint8_t *dbk, void *stats, void *chain, int nunits, int gen) {
  hipLaunchKernelGGL(k_lr_fit, dim3(nunits, 4), dim3(LR_T), 0, st, fp, src, cdef, dbk, (LrUnitStat *)stats, gen,
                     1 << 20);
  hipLaunchKernelGGL(k_lr_chain_init, dim3(1), dim3(64), 0, st, (LrChain *)chain);
  return hipGetLastError();
}

hipError_t launch_lr_path(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdef,
                          const uint8_t *dbk, void *stats, const void *chain, int nunits) {
  hipLaunchKernelGGL(k_lr_path, dim3(nunits), dim3(LR_T), 0, st, fp, src, cdef, dbk, (LrUnitStat *)stats,
                     (const LrChain *)chain);
  return hipGetLastError();
}

hipError_t launch_lr_chain(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdef,
                           const uint8_t *dbk, const void *stats, Mav1LrUnit *units, int *lrType, int lrTest,
                           int *reruns, void *chain, int stop) {
  hipLaunchKernelGGL(k_lr_chain, dim3(3), dim3(LR_T), 0, st, fp, src, cdef, dbk, (const LrUnitStat *)stats, units,
                     lrType, lrTest, reruns, (LrChain *)chain, stop);
  return hipGetLastError();
}

/* ---- batched launches (batch.h): the same kernels over several frames ---- */

/* The fit over every frame's units: the workgroups in part-major order
 * (g.start over units; part = block / units), so every fit workgroup of
 * every frame is dispatched before any refinement workgroup, as in the
 * single-frame launch. */
__global__ __launch_bounds__(LR_T) void k_lr_fit_b(const BatchFrame *__restrict__ T, const BatchGrid g, int maxSpin) {
  const int units = g.start[g.n], part = blockIdx.x / units;
  const BatchIdx b = batch_idx(g, blockIdx.x - part * units);
  const BatchFrame &F = T[b.f];
  lr_fit_body(F.fp, gptr(F.src), gptr(F.rec), gptr(F.dbk), (LrUnitStat *)gptr(F.lrstats), F.lr_gen, maxSpin, b.x, part, 0);
}
__global__ void k_lr_chain_init_b(const BatchFrame *__restrict__ T) {
  const int p = threadIdx.x;
  if (p >= 3) return;
  LrChain c = {};
  for (int k = 0; k < 6; k++) c.ref[k] = Wiener_Taps_Mid[k % 3];
  for (int k = 0; k < 2; k++) c.refX[k] = Sgrproj_Xqd_Mid[k];
  ((LrChain *)T[blockIdx.x].lrchain)[p] = c;
}
__global__ __launch_bounds__(LR_T) void k_lr_path_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  lr_path_body(F.fp, gptr(F.src), gptr(F.rec), gptr(F.dbk), (LrUnitStat *)gptr(F.lrstats), (const LrChain *)gptr(F.lrchain), b.x, b.y, b.z);
}
__global__ __launch_bounds__(LR_T) void k_lr_chain_b(const BatchFrame *__restrict__ T, const BatchGrid g, int lrTest,
                                                     int stop) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  lr_chain_body(F.fp, gptr(F.src), gptr(F.rec), gptr(F.dbk), (const LrUnitStat *)gptr(F.lrstats), gptr(F.lrunits), gptr(F.lrtype), lrTest, gptr(F.lrreruns),
                (LrChain *)gptr(F.lrchain), stop, b.x, b.y, b.z);
}

static BatchGrid unit_grid(const BatchFrame *H, int n) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) batch_grid_add(g, H[f].lr_units, 1, 1);
  return g;
}

hipError_t launch_lr_fit_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  const BatchGrid g = unit_grid(H, n);
  hipLaunchKernelGGL(k_lr_fit_b, dim3(g.start[n] * 4), dim3(LR_T), 0, st, T, g, 1 << 20);
  hipLaunchKernelGGL(k_lr_chain_init_b, dim3(n), dim3(64), 0, st, T);
  return hipGetLastError();
}

hipError_t launch_lr_path_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  const BatchGrid g = unit_grid(H, n);
  hipLaunchKernelGGL(k_lr_path_b, dim3(g.start[n]), dim3(LR_T), 0, st, T, g);
  return hipGetLastError();
}

hipError_t launch_lr_chain_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int lrTest, int stop) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) batch_grid_add(g, 3, 1, 1);
  hipLaunchKernelGGL(k_lr_chain_b, dim3(g.start[n]), dim3(LR_T), 0, st, T, g, lrTest, stop);
  return hipGetLastError();
}
