/* Closed-loop reconstruction on gfx950 (behaviour of src/kernels/k_recon.h:
 * recon_sb / recon_block / code_plane / code_tx_block, in decode order, with
 * the dependencies of src/backend_cpu.c:81-86) and the intra edge filter
 * types of the frame driver (src/encoder.c:319-346).
 *
 * Inter blocks never read reconstructed neighbours: phase A codes every
 * inter block of the frame at once (one wave per block). Intra blocks read
 * the reconstruction of earlier blocks: phase B runs one launch per
 * anti-diagonal t = column-in-tile + 2 * row of the superblock grid (left,
 * above and above-right superblocks are on earlier diagonals, tile columns
 * are independent), one wave per superblock walking its blocks in decode
 * order with the BlockDecoded flags in LDS. Within a transform block the
 * lanes share every step: residual, forward transform (matrix form, int64
 * sums), quantiser, eob / rate / SSE reductions, inverse transform (one
 * lane per row, then per column, in registers), reconstruction. */
#include "dev_pred.h"
#include "dev_tx.h"
#include "internal.h"

#define BDD 18
#define BDI(p, y, x) ((p) * BDD * BDD + ((y) + 1) * BDD + ((x) + 1))

struct RLds {
  union {
    struct {
      uint8_t win[71 * 72];
      int16_t mid[71 * 64];
      int32_t p0[64 * 64];
    } pr;
    struct {
      int32_t resid[64 * RS]; /* residual, then the inverse transform's output */
      int32_t tmp[64 * 32];
    } fw;
  } u;
  int32_t coef[32 * 32];
  int32_t qtry[32 * 32];
  int32_t qbest[32 * 32];
  int32_t dq[32 * QS];
  int32_t cfl[32 * 32];
  int edgeA[EDGE_NEG + 2 * 64 + 32], edgeL[EDGE_NEG + 2 * 64 + 32], ecopy[2 * 64 + 40];
  uint8_t bd[3 * BDD * BDD];
  uint8_t bdSave[BDD * BDD];
  uint8_t save[64 * 64];
  Mav1BlockInfo blk;
};

/* The working set of the wave that reconstructs a block / superblock
 * (one per workgroup). File scope so that the non-inlined stage functions
 * address it as LDS. */
__shared__ RLds gL;
#define NOINL __device__ __attribute__((noinline))

/* ---- scan / tx-set helpers (spec 5.11.47 / 5.11.48) ---- */
DEVI const int16_t *scan_of(int txSz, int kind) {
  switch (txSz) {
    case TX_4X4: return kind == 1 ? Mrow_Scan_4x4 : kind == 2 ? Mcol_Scan_4x4 : Default_Scan_4x4;
    case TX_4X8: return kind == 1 ? Mrow_Scan_4x8 : kind == 2 ? Mcol_Scan_4x8 : Default_Scan_4x8;
    case TX_8X4: return kind == 1 ? Mrow_Scan_8x4 : kind == 2 ? Mcol_Scan_8x4 : Default_Scan_8x4;
    case TX_8X8: return kind == 1 ? Mrow_Scan_8x8 : kind == 2 ? Mcol_Scan_8x8 : Default_Scan_8x8;
    case TX_8X16: return kind == 1 ? Mrow_Scan_8x16 : kind == 2 ? Mcol_Scan_8x16 : Default_Scan_8x16;
    case TX_16X8: return kind == 1 ? Mrow_Scan_16x8 : kind == 2 ? Mcol_Scan_16x8 : Default_Scan_16x8;
    case TX_16X16: return kind == 1 ? Mrow_Scan_16x16 : kind == 2 ? Mcol_Scan_16x16 : Default_Scan_16x16;
    case TX_4X16: return kind == 1 ? Mrow_Scan_4x16 : kind == 2 ? Mcol_Scan_4x16 : Default_Scan_4x16;
    case TX_16X4: return kind == 1 ? Mrow_Scan_16x4 : kind == 2 ? Mcol_Scan_16x4 : Default_Scan_16x4;
    case TX_16X32: return Default_Scan_16x32;
    case TX_32X16: return Default_Scan_32x16;
    case TX_8X32: return Default_Scan_8x32;
    case TX_32X8: return Default_Scan_32x8;
    default: return Default_Scan_32x32;
  }
}
DEVI const int16_t *get_scan(int txSz, int t) {
  if (txSz == TX_16X64) return Default_Scan_16x32;
  if (txSz == TX_64X16) return Default_Scan_32x16;
  if (Tx_Size_Sqr_Up[txSz] == TX_64X64) return Default_Scan_32x32;
  if (t == IDTX) return scan_of(txSz, 0);
  if (t == V_DCT || t == V_ADST || t == V_FLIPADST) return scan_of(txSz, 1);
  if (t == H_DCT || t == H_ADST || t == H_FLIPADST) return scan_of(txSz, 2);
  return scan_of(txSz, 0);
}
DEVI int get_tx_set(int txSz, int isInter, int reduced) {
  const int sq = Tx_Size_Sqr[txSz], up = Tx_Size_Sqr_Up[txSz];
  if (up > TX_32X32) return TX_SET_DCTONLY;
  if (isInter) return (reduced || up == TX_32X32) ? TX_SET_INTER_3 : sq == TX_16X16 ? TX_SET_INTER_2 : TX_SET_INTER_1;
  if (up == TX_32X32) return TX_SET_DCTONLY;
  return (reduced || sq == TX_16X16) ? TX_SET_INTRA_2 : TX_SET_INTRA_1;
}
DEVI int in_set(int set, int t, int isInter) { return isInter ? TX_IN_SET_INTER(set, t) : TX_IN_SET_INTRA(set, t); }
DEVI int seg_eob(int txSz) {
  if (txSz == TX_16X64 || txSz == TX_64X16) return 512;
  return d_min(1024, Tx_Width[txSz] * Tx_Height[txSz]);
}
DEVI int uv_tx_size(int bs) {
  const int t = Max_Tx_Size_Rect[SUBSAMPLED_SIZE(bs, 1, 1)];
  if (Tx_Width[t] == 64 || Tx_Height[t] == 64) return Tx_Width[t] == 16 ? TX_16X32 : Tx_Height[t] == 16 ? TX_32X16 : TX_32X32;
  return t;
}
DEVI int tile_start(const Mav1FrameParams &fp, int miCol) { return ((miCol >> 4) / fp.tile_w_sb) * fp.tile_w_sb * 16; }

DEVI int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = d_max(v, __shfl_xor(v, o, 64));
  return v;
}

/* find_eob (k_recon.h:50-54): 1 + last scan position with a non-zero level. */
DEVI int wave_eob(const int32_t *q, const int16_t *scan, int n, int lane) {
  int m = 0;
  for (int c = lane; c < n; c += 64)
    if (q[scan[c]]) m = c + 1;
  return wave_max(m);
}
/* est_coef_bits (k_recon.h:39-48) */
DEVI int wave_coef_bits(const int32_t *q, const int16_t *scan, int eob, int lane) {
  if (!eob) return 8;
  int b = 0;
  for (int c = lane; c < eob; c += 64) {
    int l = d_abs(q[scan[c]]);
    b += l ? 8 * (2 + 2 * d_floorlog2((unsigned)l)) : 6;
  }
  return 48 + 16 * d_floorlog2((unsigned)eob + 1) + wave_sum(b);
}

/* Lossless 4x4: forward WHT (k_fwd.h:112-136) and inverse WHT (7.13.2.10),
 * one lane. */
DEVI void wht_fwd(const int32_t *r, int32_t *out) {
  int t[16];
  for (int j = 0; j < 4; j++) {
    int a = r[j], b = r[RS + j], c = r[2 * RS + j], d = r[3 * RS + j];
    a += b;
    d -= c;
    int e = (a - d) >> 1;
    b = e - b;
    c = e - c;
    a -= c;
    d += b;
    t[j] = a;
    t[4 + j] = c;
    t[8 + j] = d;
    t[12 + j] = b;
  }
  for (int i = 0; i < 4; i++) {
    int a = t[4 * i], b = t[4 * i + 1], c = t[4 * i + 2], d = t[4 * i + 3];
    a += b;
    d -= c;
    int e = (a - d) >> 1;
    b = e - b;
    c = e - c;
    a -= c;
    d += b;
    out[4 * i] = a;
    out[4 * i + 1] = c;
    out[4 * i + 2] = d;
    out[4 * i + 3] = b;
  }
}
DEVI void wht_inv1(int32_t *T, int shift) {
  int32_t a = T[0] >> shift, c = T[1] >> shift, d = T[2] >> shift, b = T[3] >> shift;
  a += c;
  d -= b;
  int32_t e = (a - d) >> 1;
  b = e - b;
  c = e - c;
  a -= b;
  d += c;
  T[0] = a;
  T[1] = b;
  T[2] = c;
  T[3] = d;
}

/* One transform block whose prediction is in px (k_recon.h:60-173).
 * Returns the eob; the chosen type and RD cost in *outType / *outCost. */
NOINL int code_tx_block(const Mav1FrameParams &fp, const uint8_t *__restrict__ spx, uint8_t *px, int S, int x, int y,
                       int txSz, int plane, const int *cands, int nc, int32_t *coefOut,
                       const int32_t *__restrict__ mats, int lane, int *outType, long long *outCost) {
  RLds &L = gL;
  const int w = Tx_Width[txSz], h = Tx_Height[txSz], tw = d_min(32, w), th = d_min(32, h);
  const int dcQ = fp.dc_q[plane], acQ = fp.ac_q[plane];
  int32_t *resid = L.u.fw.resid;
  if (fp.lossless) {
    int eob = 0;
    if (lane == 0) {
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) resid[i * RS + j] = (int)spx[(y + i) * S + x + j] - (int)px[(y + i) * S + x + j];
      int32_t q16[16];
      wht_fwd(resid, q16);
      for (int c = 15; c >= 0; c--)
        if (q16[Default_Scan_4x4[c]]) {
          eob = c + 1;
          break;
        }
      for (int k = 0; k < 16; k++) coefOut[k] = q16[k];
      if (eob) {
        int32_t r4[16], T[4];
        for (int k = 0; k < 16; k++) {
          long long d = (long long)q16[k] * (k == 0 ? dcQ : acQ), ad = d < 0 ? -d : d;
          long long d2 = ad & 0xFFFFFF;
          r4[k] = d_clip(-(1 << 15), (1 << 15) - 1, (int)(d < 0 ? -d2 : d2));
        }
        for (int i = 0; i < 4; i++) {
          for (int j = 0; j < 4; j++) T[j] = r4[i * 4 + j];
          wht_inv1(T, 2);
          for (int j = 0; j < 4; j++) r4[i * 4 + j] = d_clip(-(1 << 15), (1 << 15) - 1, T[j]);
        }
        for (int j = 0; j < 4; j++) {
          for (int i = 0; i < 4; i++) T[i] = r4[i * 4 + j];
          wht_inv1(T, 0);
          for (int i = 0; i < 4; i++) r4[i * 4 + j] = T[i];
        }
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++) {
            int o = (y + i) * S + x + j;
            px[o] = (uint8_t)d_clip1((int)px[o] + r4[i * 4 + j]);
          }
      }
    }
    __syncthreads();
    *outType = DCT_DCT;
    *outCost = 0;
    return __shfl(eob, 0, 64);
  }
  const bool direct = nc == 1 && !fp.tx_mode_select;
  long long bestCost = -1;
  int bestType = DCT_DCT, bestEob = 0;
  for (int ci = 0; ci < nc; ci++) {
    const int t = cands[ci], fud = flip_ud(t), flr = flip_lr(t);
    for (int k = lane; k < w * h; k += 64) {
      int i = k / w, j = k - i * w;
      int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
      resid[i * RS + j] = (int)spx[o] - (int)px[o];
    }
    __syncthreads();
    wave_fwd2d(resid, L.u.fw.tmp, L.coef, txSz, t, mats, lane);
    int32_t *qo = direct ? L.qbest : L.qtry;
    wave_quantize(L.coef, qo, txSz, dcQ, acQ, lane);
    const int16_t *scan = get_scan(txSz, t);
    const int eob = wave_eob(qo, scan, seg_eob(txSz), lane);
    if (direct) {
      if (eob > 0) {
        wave_dequant(qo, L.dq, txSz, dcQ, acQ, lane);
        wave_itx2d(L.dq, resid, txSz, t, lane);
      }
      bestCost = 0;
      bestType = eob > 0 ? t : DCT_DCT;
      bestEob = eob;
      continue;
    }
    long long sse = 0;
    if (eob > 0) {
      wave_dequant(L.qtry, L.dq, txSz, dcQ, acQ, lane);
      wave_itx2d(L.dq, resid, txSz, t, lane);
      long long e = 0;
      for (int k = lane; k < w * h; k += 64) {
        int i = k / w, j = k - i * w;
        int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
        int d = (int)spx[o] - d_clip1((int)px[o] + resid[i * RS + j]);
        e += d * d;
      }
      sse = wave_sum64(e);
    } else {
      long long e = 0;
      for (int k = lane; k < w * h; k += 64) {
        int v = resid[(k / w) * RS + (k % w)];
        e += v * v;
      }
      sse = wave_sum64(e);
    }
    const long long cost = sse * 16 + (long long)fp.lambda * wave_coef_bits(L.qtry, scan, eob, lane) / 8;
    if (bestCost < 0 || cost < bestCost) {
      bestCost = cost;
      bestType = eob > 0 ? t : DCT_DCT;
      bestEob = eob;
      for (int k = lane; k < tw * th; k += 64) L.qbest[k] = L.qtry[k];
    }
    __syncthreads();
  }
  for (int k = lane; k < tw * th; k += 64) coefOut[k] = L.qbest[k];
  if (bestEob > 0) {
    const int fud = flip_ud(bestType), flr = flip_lr(bestType);
    if (nc > 1) {
      wave_dequant(L.qbest, L.dq, txSz, dcQ, acQ, lane);
      wave_itx2d(L.dq, resid, txSz, bestType, lane);
    }
    for (int k = lane; k < w * h; k += 64) {
      int i = k / w, j = k - i * w;
      int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
      px[o] = (uint8_t)d_clip1((int)px[o] + resid[i * RS + j]);
    }
  }
  __syncthreads();
  *outType = bestType;
  *outCost = bestCost;
  return bestEob;
}

/* ---- intra prediction of one transform block with the normative edge
 * preparation (7.11.2 incl. the edge filter and upsampling) ---- */
DEVI int edge_strength(int w, int h, int filterType, int delta) {
  int d = d_abs(delta), wh = w + h;
  if (filterType == 0) {
    if (wh <= 8) return d >= 56 ? 1 : 0;
    if (wh <= 16) return d >= 40 ? 1 : 0;
    if (wh <= 24) return d >= 32 ? 3 : d >= 16 ? 2 : d >= 8 ? 1 : 0;
    if (wh <= 32) return d >= 32 ? 3 : d >= 4 ? 2 : 1;
    return 3;
  }
  if (wh <= 8) return d >= 64 ? 2 : d >= 40 ? 1 : 0;
  if (wh <= 16) return d >= 48 ? 2 : d >= 20 ? 1 : 0;
  if (wh <= 24) return d >= 4 ? 3 : 0;
  return 3;
}
DEVI int use_upsample(int w, int h, int filterType, int delta) {
  int d = d_abs(delta);
  if (d <= 0 || d >= 40) return 0;
  return filterType == 0 ? (w + h <= 16) : (w + h <= 8);
}
/* 7.11.2.12 on e[-1 .. sz-2] */
DEVI void wave_edge_filter(int *e, int *cp, int sz, int strength, int lane) {
  if (!strength) return;
  for (int i = lane; i < sz; i += 64) cp[i] = e[i - 1];
  __syncthreads();
  for (int i = 1 + lane; i < sz; i += 64) {
    int s = 0;
#pragma unroll
    for (int j = 0; j < 5; j++) s += INTRA_EDGE_KERNEL(strength - 1, j) * cp[d_clip(0, sz - 1, i - 2 + j)];
    e[i - 1] = (s + 8) >> 4;
  }
  __syncthreads();
}
/* 7.11.2.11 */
DEVI void wave_edge_upsample(int *e, int *dup, int numPx, int lane) {
  for (int i = lane - 1; i < numPx; i += 64) dup[i + 2] = e[i];
  if (lane == 0) {
    dup[0] = e[-1];
    dup[numPx + 2] = e[numPx - 1];
  }
  __syncthreads();
  if (lane == 0) e[-2] = dup[0];
  for (int i = lane; i < numPx; i += 64) {
    int s = -dup[i] + 9 * dup[i + 1] + 9 * dup[i + 2] - dup[i + 3];
    e[2 * i - 1] = d_clip1(d_round2(s, 4));
    e[2 * i] = dup[i + 2];
  }
  __syncthreads();
}

NOINL void intra_tx_predict(const Mav1FrameParams &fp, uint8_t *px, int S, int sx, int sy, int lw, int lh, int mode,
                            int delta, int hl, int ha, int har, int hbl, int maxX, int maxY, int edgeFilter,
                            int filterType, int lane) {
  RLds &L = gL;
  const int w = 1 << lw, h = 1 << lh;
  int *A = L.edgeA + EDGE_NEG, *Lc = L.edgeL + EDGE_NEG;
  wave_intra_edges(px, S, sx, sy, w, h, hl, ha, har, hbl, maxX, maxY, A, Lc, lane);
  IntraMode im;
  intra_mode_setup(im, mode, delta, lw, lh);
  im.corner = A[-1];
  if (mode >= V_PRED && mode <= D67_PRED && edgeFilter) {
    const int pa = im.pAngle;
    if (pa != 90 && pa != 180) {
      if (pa > 90 && pa < 180 && w + h >= 24) {
        if (lane == 0) A[-1] = Lc[-1] = d_round2(Lc[0] * 5 + A[-1] * 6 + A[0] * 5, 4);
        __syncthreads();
      }
      if (ha)
        wave_edge_filter(A, L.ecopy, d_min(w, maxX - sx + 1) + (pa < 90 ? h : 0) + 1,
                         edge_strength(w, h, filterType, pa - 90), lane);
      if (hl)
        wave_edge_filter(Lc, L.ecopy, d_min(h, maxY - sy + 1) + (pa > 180 ? w : 0) + 1,
                         edge_strength(w, h, filterType, pa - 180), lane);
    }
    im.upA = use_upsample(w, h, filterType, pa - 90);
    if (im.upA) wave_edge_upsample(A, L.ecopy, w + (pa < 90 ? h : 0), lane);
    im.upL = use_upsample(w, h, filterType, pa - 180);
    if (im.upL) wave_edge_upsample(Lc, L.ecopy, h + (pa > 180 ? w : 0), lane);
  }
  im.dcv = mode == DC_PRED ? wave_intra_dc(A, Lc, w, h, lw, lh, hl, ha, lane) : 0;
  for (int k = lane; k < w * h; k += 64) {
    int i = k >> lw, j = k & (w - 1);
    px[(sy + i) * S + sx + j] = (uint8_t)intra_px(im, A, Lc, i, j);
  }
  __syncthreads();
}

/* CfL (7.11.5) over the DC prediction in px: the alpha with the smallest
 * SSE against the source (first minimum over -16..16) unless given. */
NOINL int cfl_tx(const uint8_t *luma, int LS, int sx, int sy, int maxLW, int maxLH, int lw, int lh, uint8_t *px, int S,
                 const uint8_t *spx, int alphaIn, int search, int excludeZero, int lane) {
  RLds &L = gL;
  const int w = 1 << lw, h = 1 << lh, n = w * h;
  const int dcv = px[sy * S + sx]; /* DC prediction: one value over the block */
  int s = 0;
  for (int k = lane; k < n; k += 64) {
    int i = k >> lw, j = k & (w - 1);
    int ly = d_min((sy + i) << 1, maxLH - 2), lx = d_min((sx + j) << 1, maxLW - 2);
    const uint8_t *l = luma + ly * LS + lx;
    int v = ((int)l[0] + l[1] + l[LS] + l[LS + 1]) << 1;
    L.cfl[k] = v;
    s += v;
  }
  const int avg = d_round2(wave_sum(s), lw + lh);
  __syncthreads();
  int alpha = alphaIn;
  if (search) {
    long long best = -1;
    for (int a = -16; a <= 16; a++) {
      int e = 0;
      for (int k = lane; k < n; k += 64) {
        int i = k >> lw, j = k & (w - 1);
        int d = d_clip1(dcv + d_round2signed(a * (L.cfl[k] - avg), 6)) - spx[(sy + i) * S + sx + j];
        e += d * d;
      }
      long long sse = wave_sum64(e);
      if (a == 0 && excludeZero) continue;
      if (best < 0 || sse < best) {
        best = sse;
        alpha = a;
      }
    }
  }
  for (int k = lane; k < n; k += 64) {
    int i = k >> lw, j = k & (w - 1);
    px[(sy + i) * S + sx + j] = (uint8_t)d_clip1(dcv + d_round2signed(alpha * (L.cfl[k] - avg), 6));
  }
  __syncthreads();
  return alpha;
}

/* code_plane (k_recon.h:190-291) for the block in L.blk. markOnly: only the
 * BlockDecoded flags (the block was coded by the inter phase). */
NOINL long long code_plane(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, uint8_t *rec,
                           int32_t *__restrict__ coefs, const int32_t *__restrict__ mats, int plane, int txSz,
                           int &maxLW, int &maxLH, bool useBd, bool markOnly, int lane) {
  RLds &L = gL;
  Mav1BlockInfo &b = L.blk;
  const int bs = b.bsize, sub = plane > 0, S = fp.stride[plane];
  const int availU = b.mi_row > 0, availL = b.mi_col > tile_start(fp, b.mi_col);
  const int psz = SUBSAMPLED_SIZE(bs, sub, sub);
  const int n4w = Num_4x4_Blocks_Wide[psz], n4h = Num_4x4_Blocks_High[psz];
  const int baseX = (b.mi_col >> sub) * 4, baseY = (b.mi_row >> sub) * 4;
  uint8_t *px = rec + fp.offset[plane];
  const uint8_t *spx = src + fp.offset[plane];
  const int stepX = Tx_Width[txSz] >> 2, stepY = Tx_Height[txSz] >> 2;
  int cands[16], nc = 0;
  if (plane == 0) {
    const int set = get_tx_set(txSz, b.is_inter, fp.reduced_tx_set);
    if (!fp.tx_search || set == TX_SET_DCTONLY || fp.lossless) cands[nc++] = DCT_DCT;
    else
      for (int t = 0; t < 16; t++)
        if (in_set(set, t, b.is_inter)) cands[nc++] = t;
  } else {
    const int set = get_tx_set(txSz, b.is_inter, fp.reduced_tx_set);
    int t = b.is_inter ? b.tx_type : Mode_To_Txfm[b.uv_mode];
    if (Tx_Size_Sqr_Up[txSz] > TX_32X32 || !in_set(set, t, b.is_inter)) t = DCT_DCT;
    cands[nc++] = t;
  }
  long long total = 0;
  int txIdx = 0, anyNz = 0;
  for (int y = 0; y < n4h; y += stepY)
    for (int x = 0; x < n4w; x += stepX) {
      const int sx = baseX + 4 * x, sy = baseY + 4 * y;
      if (sx >= ((fp.mi_cols * 4) >> sub) || sy >= ((fp.mi_rows * 4) >> sub)) continue;
      const int row = (sy << sub) >> 2, col = (sx << sub) >> 2, sbr = row & 15, sbc = col & 15;
      if (!markOnly) {
        if (!b.is_inter) {
          const int isCfl = plane > 0 && b.uv_mode == UV_CFL_PRED;
          const int lw = Tx_Width_Log2[txSz], lh = Tx_Height_Log2[txSz];
          intra_tx_predict(fp, px, S, sx, sy, lw, lh, plane == 0 ? b.y_mode : (isCfl ? DC_PRED : b.uv_mode),
                           plane == 0 ? b.angle_y : b.angle_uv, availL || x > 0, availU || y > 0,
                           L.bd[BDI(plane, (sbr >> sub) - 1, (sbc >> sub) + stepX)],
                           L.bd[BDI(plane, (sbr >> sub) + stepY, (sbc >> sub) - 1)], ((fp.mi_cols * 4) >> sub) - 1,
                           ((fp.mi_rows * 4) >> sub) - 1, fp.enable_intra_edge_filter,
                           (b.interp_filter >> (plane == 0 ? 4 : 5)) & 1, lane);
          if (isCfl) {
            int a;
            if (txIdx > 0)
              a = cfl_tx(rec + fp.offset[0], fp.stride[0], sx, sy, maxLW, maxLH, lw, lh, px, S, spx,
                         plane == 1 ? b.cfl_alpha_u : b.cfl_alpha_v, 0, 0, lane);
            else
              a = cfl_tx(rec + fp.offset[0], fp.stride[0], sx, sy, maxLW, maxLH, lw, lh, px, S, spx, 0, 1,
                         plane == 2 && b.cfl_alpha_u == 0, lane);
            if (txIdx == 0 && lane == 0) {
              if (plane == 1) b.cfl_alpha_u = a;
              else b.cfl_alpha_v = a;
            }
            __syncthreads();
          }
        }
        if (plane == 0) {
          maxLW = sx + stepX * 4;
          maxLH = sy + stepY * 4;
        }
        int t = DCT_DCT;
        long long c = 0;
        const int eob = code_tx_block(fp, spx, px, S, sx, sy, txSz, plane, cands, nc,
                                      coefs + b.coef_off[plane] + txIdx * Tx_Width[txSz] * Tx_Height[txSz], mats,
                                      lane, &t, &c);
        total += c;
        if (plane == 0 && lane == 0) {
          if (txIdx < 16) {
            const int sh = 4 * (txIdx & 7);
            uint32_t *f = (uint32_t *)(txIdx < 8 ? &b.tx_types_lo : &b.tx_types_hi);
            *f = (*f & ~(15u << sh)) | ((uint32_t)t << sh);
          }
          if (txIdx == 0) b.tx_type = t;
        }
        anyNz |= eob > 0;
      }
      txIdx++;
      if (useBd && lane == 0)
        for (int i = 0; i < stepY; i++)
          for (int j = 0; j < stepX; j++) L.bd[BDI(plane, (sbr >> sub) + i, (sbc >> sub) + j)] = 1;
      __syncthreads();
    }
  if (!markOnly && lane == 0) b.eob[plane] = anyNz;
  __syncthreads();
  return total;
}

/* recon_block (k_recon.h:293-394) for L.blk. Inter blocks are fully coded
 * by the inter phase; the intra phase only marks their decoded area. */
NOINL void predict_inter_plane(const uint8_t *__restrict__ plane, int S, int refW, int refH, int x, int y, int w,
                              int h, int mvr, int mvc, int sub, int filter, uint8_t *dst, int ds, int32_t *dst32,
                              int round1, int lane) {
  wave_predict_inter(plane, S, refW, refH, x, y, w, h, mvr, mvc, sub, filter, gL.u.pr.win, gL.u.pr.mid, dst, ds, dst32,
                     round1, lane);
}

NOINL void recon_block(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, uint8_t *rec,
                       const uint8_t *__restrict__ ref, const uint8_t *__restrict__ gold, int32_t *__restrict__ coefs,
                       const int32_t *__restrict__ mats, bool useBd, bool markOnly, int lane) {
  RLds &L = gL;
  Mav1BlockInfo &b = L.blk;
  const int bs = b.bsize, comp = b.is_inter && b.ref_frame1 > INTRA_FRAME;
  const uint8_t *r0 = (b.is_inter && b.ref_frame != LAST_FRAME) ? gold : ref;
  int maxLW = 0, maxLH = 0;
  if (!markOnly) {
    if (lane == 0) {
      b.tx_depth = 0;
      b.interp_filter &= 0x30;
    }
    __syncthreads();
    if (b.is_inter && fp.switchable_interp && !comp) {
      const int w = BLOCK_WIDTH(bs), h = BLOCK_HEIGHT(bs), x0 = b.mi_col * 4, y0 = b.mi_row * 4, S = fp.stride[0];
      uint8_t *px = rec + fp.offset[0] + y0 * S + x0;
      const uint8_t *sp = src + fp.offset[0] + y0 * S + x0;
      const int vw = d_min(w, fp.frame_w - x0), vh = d_min(h, fp.frame_h - y0);
      long long best = -1;
      int bestF = 0;
      for (int f = 0; f < 9; f++) {
        const int packed = (f % 3) | ((f / 3) << 2);
        predict_inter_plane(r0 + fp.offset[0], S, fp.frame_w, fp.frame_h, x0, y0, w, h, b.mv_row, b.mv_col, 0, packed,
                            px, S, nullptr, 11, lane);
        long long e = 0;
        for (int k = lane; k < vw * vh; k += 64) {
          int i = k / vw, j = k - i * vw;
          int d = (int)px[i * S + j] - (int)sp[i * S + j];
          e += d * d;
        }
        const long long c = wave_sum64(e) * 16 + (long long)fp.lambda * ((f % 3 != 0) + (f / 3 != 0));
        if (best < 0 || c < best) {
          best = c;
          bestF = packed;
        }
      }
      if (lane == 0) b.interp_filter = bestF;
      __syncthreads();
    }
  }
  for (int plane = 0; plane < 3; plane++) {
    const int sub = plane > 0, S = fp.stride[plane];
    const int baseX = (b.mi_col >> sub) * 4, baseY = (b.mi_row >> sub) * 4;
    const int w = BLOCK_WIDTH(bs) >> sub, h = BLOCK_HEIGHT(bs) >> sub;
    if (!markOnly) {
      if (comp) {
        predict_inter_plane(r0 + fp.offset[plane], S, fp.frame_w, fp.frame_h, baseX, baseY, w, h, b.mv_row, b.mv_col,
                            sub, b.interp_filter, nullptr, w, L.u.pr.p0, 7, lane);
        /* second prediction's vertical pass combined with the first one */
        int32_t *p1 = (int32_t *)L.coef; /* free until code_plane */
        for (int yy = 0; yy < h; yy += 16) {
          const int hh = d_min(16, h - yy);
          predict_inter_plane(gold + fp.offset[plane], S, fp.frame_w, fp.frame_h, baseX, baseY + yy, w, hh, b.mv1_row,
                              b.mv1_col, sub, b.interp_filter, nullptr, w, p1, 7, lane);
          uint8_t *d = rec + fp.offset[plane] + (baseY + yy) * S + baseX;
          for (int k = lane; k < w * hh; k += 64) {
            int i = k / w, j = k - i * w;
            d[i * S + j] = (uint8_t)d_clip1(d_round2(L.u.pr.p0[(yy + i) * w + j] + p1[i * w + j], 5));
          }
          __syncthreads();
        }
      } else if (b.is_inter) {
        predict_inter_plane(r0 + fp.offset[plane], S, fp.frame_w, fp.frame_h, baseX, baseY, w, h, b.mv_row, b.mv_col,
                            sub, b.interp_filter, rec + fp.offset[plane] + baseY * S + baseX, S, nullptr, 11, lane);
      }
    }
    if (fp.lossless) {
      code_plane(fp, src, rec, coefs, mats, plane, TX_4X4, maxLW, maxLH, useBd, markOnly, lane);
      continue;
    }
    if (plane > 0) {
      code_plane(fp, src, rec, coefs, mats, plane, uv_tx_size(bs), maxLW, maxLH, useBd, markOnly, lane);
      continue;
    }
    const int maxDepth = (!b.is_inter && fp.tx_mode_select) ? d_min(MAX_TX_DEPTH, Max_Tx_Depth[bs]) : 0;
    if (maxDepth == 0) {
      code_plane(fp, src, rec, coefs, mats, 0, Max_Tx_Size_Rect[bs], maxLW, maxLH, useBd, markOnly, lane);
      continue;
    }
    /* tx_depth RD search (intra only, so never markOnly) */
    uint8_t *px = rec + fp.offset[0] + baseY * S + baseX;
    for (int k = lane; k < w * h; k += 64) L.save[k] = px[(k / w) * S + (k % w)];
    for (int k = lane; k < BDD * BDD; k += 64) L.bdSave[k] = L.bd[k];
    __syncthreads();
    long long best = -1;
    int bestD = 0;
    for (int d = 0; d <= maxDepth + 1; d++) {
      const int depth = d <= maxDepth ? d : bestD;
      if (d == maxDepth + 1 && bestD == maxDepth) break;
      if (d > 0) {
        for (int k = lane; k < w * h; k += 64) px[(k / w) * S + (k % w)] = L.save[k];
        for (int k = lane; k < BDD * BDD; k += 64) L.bd[k] = L.bdSave[k];
        __syncthreads();
      }
      int t = Max_Tx_Size_Rect[bs];
      for (int k = 0; k < depth; k++) t = Split_Tx_Size[t];
      const long long c =
          code_plane(fp, src, rec, coefs, mats, 0, t, maxLW, maxLH, useBd, false, lane) + (long long)fp.lambda * 8 * depth;
      if (d <= maxDepth && (best < 0 || c < best)) {
        best = c;
        bestD = depth;
      }
      if (d == maxDepth + 1) break;
    }
    if (lane == 0) b.tx_depth = bestD;
    __syncthreads();
  }
}

/* Phase A: every inter block of the frame, one wave per block slot. */
__global__ __launch_bounds__(64) void k_recon_inter(const Mav1FrameParams fp, Mav1BlockInfo *__restrict__ blocks,
                                                    const Mav1SbInfo *__restrict__ sbs, const uint8_t *__restrict__ src,
                                                    uint8_t *rec, const uint8_t *__restrict__ ref,
                                                    const uint8_t *__restrict__ gold, int32_t *__restrict__ coefs,
                                                    const int32_t *__restrict__ mats) {
  RLds &L = gL;
  const int lane = threadIdx.x, sb = blockIdx.x >> 6, k = blockIdx.x & 63;
  if (k >= sbs[sb].num_blocks) return;
  Mav1BlockInfo *g = blocks + sbs[sb].first_block + k;
  if (!g->is_inter) return;
  if (lane < 27) ((int32_t *)&L.blk)[lane] = ((const int32_t *)g)[lane];
  __syncthreads();
  recon_block(fp, src, rec, ref, gold, coefs, mats, false, false, lane);
  if (lane < 27) ((int32_t *)g)[lane] = ((const int32_t *)&L.blk)[lane];
}

/* Phase B: superblocks of anti-diagonal t, one wave each; grid = tile
 * columns x superblock rows. */
__global__ __launch_bounds__(64) void k_recon_intra(const Mav1FrameParams fp, Mav1BlockInfo *__restrict__ blocks,
                                                    const Mav1SbInfo *__restrict__ sbs, const uint8_t *__restrict__ src,
                                                    uint8_t *rec, const uint8_t *__restrict__ ref,
                                                    const uint8_t *__restrict__ gold, int32_t *__restrict__ coefs,
                                                    const int32_t *__restrict__ mats, int t) {
  RLds &L = gL;
  const int lane = threadIdx.x;
  const int tile = blockIdx.x % ((fp.sb_cols + fp.tile_w_sb - 1) / fp.tile_w_sb), r = blockIdx.x / ((fp.sb_cols + fp.tile_w_sb - 1) / fp.tile_w_sb);
  const int c0 = tile * fp.tile_w_sb, cl = t - 2 * r;
  if (cl < 0 || cl >= fp.tile_w_sb || c0 + cl >= fp.sb_cols || r >= fp.sb_rows) return;
  const int sbIdx = r * fp.sb_cols + c0 + cl;
  const Mav1SbInfo &S = sbs[sbIdx];
  int nIntra = 0;
  for (int i = lane; i < S.num_blocks; i += 64) nIntra += !blocks[S.first_block + i].is_inter;
  if (wave_sum(nIntra) == 0) return;
  /* clear_block_decoded_flags (k_recon.h:402-414) */
  const int mr = r * 16, mc = (c0 + cl) * 16;
  const int tileEnd = d_min(fp.mi_cols, tile_start(fp, mc) + fp.tile_w_sb * 16);
  for (int q = lane; q < 3 * BDD * BDD; q += 64) {
    const int p = q / (BDD * BDD), e = q - p * BDD * BDD, y = e / BDD - 1, x = e % BDD - 1, sub = p > 0;
    const int w4 = (tileEnd - mc) >> sub, h4 = (fp.mi_rows - mr) >> sub;
    uint8_t v = (y < 0 && x < w4) ? 1 : (x < 0 && y < h4) ? 1 : 0;
    if (y == (16 >> sub) && x == -1) v = 0;
    if (y <= (16 >> sub) && x <= (16 >> sub)) L.bd[q] = v;
  }
  __syncthreads();
  for (int i = 0; i < S.num_blocks; i++) {
    Mav1BlockInfo *g = blocks + S.first_block + i;
    if (lane < 27) ((int32_t *)&L.blk)[lane] = ((const int32_t *)g)[lane];
    __syncthreads();
    const bool inter = L.blk.is_inter;
    recon_block(fp, src, rec, ref, gold, coefs, mats, true, inter, lane);
    if (!inter && lane < 27) ((int32_t *)g)[lane] = ((const int32_t *)&L.blk)[lane];
    __syncthreads();
  }
}

/* Edge filter types (src/encoder.c:319-346): smooth flags per mi, then per
 * block the above / left neighbours' flags into interp_filter bits 4 / 5. */
__global__ __launch_bounds__(64) void k_edge_flags(const Mav1FrameParams fp, const Mav1BlockInfo *__restrict__ blocks,
                                                   const Mav1SbInfo *__restrict__ sbs, uint8_t *ys, uint8_t *uvs) {
  const int sb = blockIdx.x, lane = threadIdx.x;
  if (lane >= sbs[sb].num_blocks) return;
  const Mav1BlockInfo b = blocks[sbs[sb].first_block + lane];
  const int sy = !b.is_inter && (b.y_mode == SMOOTH_PRED || b.y_mode == SMOOTH_V_PRED || b.y_mode == SMOOTH_H_PRED);
  const int su = !b.is_inter && (b.uv_mode == SMOOTH_PRED || b.uv_mode == SMOOTH_V_PRED || b.uv_mode == SMOOTH_H_PRED);
  for (int r = b.mi_row; r < b.mi_row + Num_4x4_Blocks_High[b.bsize] && r < fp.mi_rows; r++)
    for (int c = b.mi_col; c < b.mi_col + Num_4x4_Blocks_Wide[b.bsize] && c < fp.mi_cols; c++) {
      ys[r * fp.mi_cols + c] = (uint8_t)sy;
      uvs[r * fp.mi_cols + c] = (uint8_t)su;
    }
}
__global__ __launch_bounds__(64) void k_edge_types(const Mav1FrameParams fp, Mav1BlockInfo *__restrict__ blocks,
                                                   const Mav1SbInfo *__restrict__ sbs, const uint8_t *ys,
                                                   const uint8_t *uvs) {
  const int sb = blockIdx.x, lane = threadIdx.x;
  if (lane >= sbs[sb].num_blocks) return;
  Mav1BlockInfo *b = &blocks[sbs[sb].first_block + lane];
  const int cols = fp.mi_cols, rows = fp.mi_rows, r = b->mi_row, c = b->mi_col;
  const int au = r > 0, al = c > tile_start(fp, c);
  const int fy = (au && ys[(r - 1) * cols + c]) || (al && ys[r * cols + c - 1]);
  const int fuv = (au && uvs[(r - 1) * cols + d_min(c + 1, cols - 1)]) || (al && uvs[d_min(r + 1, rows - 1) * cols + c - 1]);
  b->interp_filter = (b->interp_filter & 15) | fy << 4 | fuv << 5;
}

hipError_t launch_edge_types(hipStream_t st, const Mav1FrameParams &fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                             uint8_t *ys, uint8_t *uvs) {
  const int nsb = fp.sb_cols * fp.sb_rows;
  hipLaunchKernelGGL(k_edge_flags, dim3(nsb), dim3(64), 0, st, fp, blocks, sbs, ys, uvs);
  hipLaunchKernelGGL(k_edge_types, dim3(nsb), dim3(64), 0, st, fp, blocks, sbs, ys, uvs);
  return hipGetLastError();
}

hipError_t launch_recon_inter(hipStream_t st, const Mav1FrameParams &fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                              const uint8_t *src, uint8_t *rec, const uint8_t *ref, const uint8_t *gold,
                              int32_t *coefs, const int32_t *mats) {
  const int nsb = fp.sb_cols * fp.sb_rows;
  hipLaunchKernelGGL(k_recon_inter, dim3(nsb * 64), dim3(64), 0, st, fp, blocks, sbs, src, rec, ref,
                     gold, coefs, mats);
  return hipGetLastError();
}

int recon_intra_steps(const Mav1FrameParams &fp) { return fp.tile_w_sb + 2 * (fp.sb_rows - 1); }

hipError_t launch_recon_intra(hipStream_t st, const Mav1FrameParams &fp, int t, Mav1BlockInfo *blocks,
                              const Mav1SbInfo *sbs, const uint8_t *src, uint8_t *rec, const uint8_t *ref,
                              const uint8_t *gold, int32_t *coefs, const int32_t *mats) {
  const int tiles = (fp.sb_cols + fp.tile_w_sb - 1) / fp.tile_w_sb;
  hipLaunchKernelGGL(k_recon_intra, dim3(tiles * fp.sb_rows), dim3(64), 0, st, fp, blocks, sbs, src, rec, ref, gold,
                     coefs, mats, t);
  return hipGetLastError();
}
