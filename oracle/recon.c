/* ORACLE — test infrastructure only. Closed-loop reconstruction restated
 * from src/kernels/k_recon.h (code_tx_block :60-173, code_plane :190-291,
 * recon_block :293-394, recon_sb :396-417), the edge filter types of the
 * frame driver (src/encoder.c:319-346) and the scan helpers of
 * src/kernels/k_scan.h. Superblocks run in raster order, which satisfies
 * the wavefront dependencies of src/backend_cpu.c:81-86. */
#include "ocommon.h"
#include "ointernal.h"
#include "otx.h"

/* BlockDecoded flags of the current superblock: [plane][-1 .. 16][-1 .. 16] */
#define BDD 18
#define BDF(bd, p, y, x) (bd)[(p) * BDD * BDD + ((y) + 1) * BDD + ((x) + 1)]

static const int16_t *scan_of(int txSz, int kind) {
  /* kind 0 default, 1 mrow, 2 mcol (sizes up to 16 wide / high) */
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

/* get_scan (spec 5.11.47), k_scan.h:51-61 */
static const int16_t *get_scan(int txSz, int t) {
  if (txSz == TX_16X64) return Default_Scan_16x32;
  if (txSz == TX_64X16) return Default_Scan_32x16;
  if (Tx_Size_Sqr_Up[txSz] == TX_64X64) return Default_Scan_32x32;
  if (t == IDTX) return scan_of(txSz, 0);
  if (t == V_DCT || t == V_ADST || t == V_FLIPADST) return scan_of(txSz, 1);
  if (t == H_DCT || t == H_ADST || t == H_FLIPADST) return scan_of(txSz, 2);
  return scan_of(txSz, 0);
}

/* get_tx_set (spec 5.11.48) */
static int get_tx_set(int txSz, int isInter, int reduced) {
  int sq = Tx_Size_Sqr[txSz], up = Tx_Size_Sqr_Up[txSz];
  if (up > TX_32X32) return TX_SET_DCTONLY;
  if (isInter) {
    if (reduced || up == TX_32X32) return TX_SET_INTER_3;
    return sq == TX_16X16 ? TX_SET_INTER_2 : TX_SET_INTER_1;
  }
  if (up == TX_32X32) return TX_SET_DCTONLY;
  if (reduced || sq == TX_16X16) return TX_SET_INTRA_2;
  return TX_SET_INTRA_1;
}
static int in_set(int set, int t, int isInter) { return isInter ? TX_IN_SET_INTER(set, t) : TX_IN_SET_INTRA(set, t); }

/* compute_tx_type for chroma (spec 5.11.40), k_scan.h:84-90 */
static int chroma_tx_type(int txSz, int isInter, int uvMode, int lumaType, int reduced) {
  if (Tx_Size_Sqr_Up[txSz] > TX_32X32) return DCT_DCT;
  int set = get_tx_set(txSz, isInter, reduced), t = isInter ? lumaType : Mode_To_Txfm[uvMode];
  return in_set(set, t, isInter) ? t : DCT_DCT;
}

static int seg_eob(int txSz) {
  if (txSz == TX_16X64 || txSz == TX_64X16) return 512;
  return o_min(1024, Tx_Width[txSz] * Tx_Height[txSz]);
}

static int uv_tx_size(int bs) {
  int t = Max_Tx_Size_Rect[SUBSAMPLED_SIZE(bs, 1, 1)];
  if (Tx_Width[t] == 64 || Tx_Height[t] == 64) return Tx_Width[t] == 16 ? TX_16X32 : Tx_Height[t] == 16 ? TX_32X16 : TX_32X32;
  return t;
}

static int find_eob(const int32_t *q, const int16_t *scan, int n) {
  for (int c = n - 1; c >= 0; c--)
    if (q[scan[c]]) return c + 1;
  return 0;
}

/* rate estimate of a coefficient block, 1/8 bit (k_recon.h:39-48) */
static int coef_bits(const int32_t *q, const int16_t *scan, int eob) {
  if (!eob) return 8;
  int bits = 48 + 16 * o_floorlog2((unsigned)eob + 1);
  for (int c = 0; c < eob; c++) {
    int l = o_abs(q[scan[c]]);
    bits += l ? 8 * (2 + 2 * o_floorlog2((unsigned)l)) : 6;
  }
  return bits;
}

/* Transform, quantise and reconstruct one transform block whose prediction
 * is in rec (k_recon.h:60-173). Returns the eob. */
static int code_tx_block(const Mav1FrameParams *fp, const uint8_t *src, uint8_t *rec, int S, int x, int y, int txSz,
                         int plane, const int *cands, int nc, int32_t *coefOut, const int32_t *mats, int *outType,
                         int64_t *outCost) {
  int w = Tx_Width[txSz], h = Tx_Height[txSz], tw = o_min(32, w), th = o_min(32, h);
  int dcQ = fp->dc_q[plane], acQ = fp->ac_q[plane];
  static _Thread_local int32_t resid[64 * 64], coef[64 * 64], dq[64 * 64], res[64 * 64], qtry[1024], qbest[1024];
  if (fp->lossless) {
    int32_t q16[16];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) resid[i * 64 + j] = (int)src[(y + i) * S + x + j] - (int)rec[(y + i) * S + x + j];
    o_fwht4x4(resid, q16);
    int eob = find_eob(q16, Default_Scan_4x4, 16);
    for (int k = 0; k < 16; k++) coefOut[k] = q16[k];
    if (eob > 0) {
      o_dequant(q16, dq, TX_4X4, dcQ, acQ);
      o_itx_2d(dq, res, TX_4X4, DCT_DCT, 1);
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
          int o = (y + i) * S + x + j;
          rec[o] = (uint8_t)o_clip1(rec[o] + res[i * 64 + j]);
        }
    }
    *outType = DCT_DCT;
    *outCost = 0;
    return eob;
  }
  int direct = nc == 1 && !fp->tx_mode_select; /* one candidate and nothing to compare */
  int64_t bestCost = -1;
  int bestType = DCT_DCT, bestEob = 0;
  for (int ci = 0; ci < nc; ci++) {
    int t = cands[ci], fud = o_flip_ud(t), flr = o_flip_lr(t);
    for (int i = 0; i < h; i++)
      for (int j = 0; j < w; j++) {
        int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
        resid[i * 64 + j] = (int)src[o] - (int)rec[o];
      }
    o_fwd_2d(resid, coef, txSz, t, mats);
    int32_t *qo = direct ? coefOut : qtry;
    o_quantize(coef, qo, txSz, dcQ, acQ);
    const int16_t *scan = get_scan(txSz, t);
    int eob = find_eob(qo, scan, seg_eob(txSz));
    if (direct) {
      if (eob > 0) {
        o_dequant(qo, dq, txSz, dcQ, acQ);
        o_itx_2d(dq, res, txSz, t, 0);
      }
      bestCost = 0;
      bestType = eob > 0 ? t : DCT_DCT;
      bestEob = eob;
      continue;
    }
    int64_t sse = 0;
    if (eob > 0) {
      o_dequant(qtry, dq, txSz, dcQ, acQ);
      o_itx_2d(dq, res, txSz, t, 0);
      for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
          int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
          int d = (int)src[o] - o_clip1((int)rec[o] + res[i * 64 + j]);
          sse += d * d;
        }
    } else {
      for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) sse += resid[i * 64 + j] * resid[i * 64 + j];
    }
    int64_t cost = sse * 16 + (int64_t)fp->lambda * coef_bits(qtry, scan, eob) / 8;
    if (bestCost < 0 || cost < bestCost) {
      bestCost = cost;
      bestType = eob > 0 ? t : DCT_DCT;
      bestEob = eob;
      memcpy(qbest, qtry, sizeof(int32_t) * tw * th);
    }
  }
  if (!direct) memcpy(coefOut, qbest, sizeof(int32_t) * tw * th);
  if (bestEob > 0) {
    int fud = o_flip_ud(bestType), flr = o_flip_lr(bestType);
    if (nc > 1) { /* a single candidate's residual is still in res */
      o_dequant(qbest, dq, txSz, dcQ, acQ);
      o_itx_2d(dq, res, txSz, bestType, 0);
    }
    for (int i = 0; i < h; i++)
      for (int j = 0; j < w; j++) {
        int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
        rec[o] = (uint8_t)o_clip1((int)rec[o] + res[i * 64 + j]);
      }
  }
  *outType = bestType;
  *outCost = bestCost;
  return bestEob;
}

static void set_tx_type(Mav1BlockInfo *b, int idx, int t) {
  if (idx >= 16) return;
  int sh = 4 * (idx & 7);
  uint32_t *f = (uint32_t *)(idx < 8 ? &b->tx_types_lo : &b->tx_types_hi);
  *f = (*f & ~(15u << sh)) | ((uint32_t)t << sh);
}

/* Every transform block of one plane in residual order: intra prediction,
 * CfL, coding (k_recon.h:190-291). Returns the summed RD cost. */
static int64_t code_plane(const Mav1FrameParams *fp, Mav1BlockInfo *b, const uint8_t *src, uint8_t *rec,
                          int32_t *coefs, const int32_t *mats, uint8_t *bd, int plane, int txSz, int *maxLW,
                          int *maxLH) {
  int bs = b->bsize, sub = plane > 0, S = fp->stride[plane];
  int availU = b->mi_row > 0, availL = b->mi_col > ((b->mi_col >> 4) / fp->tile_w_sb) * fp->tile_w_sb * 16;
  int psz = SUBSAMPLED_SIZE(bs, sub, sub), n4w = Num_4x4_Blocks_Wide[psz], n4h = Num_4x4_Blocks_High[psz];
  int baseX = (b->mi_col >> sub) * 4, baseY = (b->mi_row >> sub) * 4;
  uint8_t *px = rec + fp->offset[plane];
  const uint8_t *sp = src + fp->offset[plane];
  int stepX = Tx_Width[txSz] >> 2, stepY = Tx_Height[txSz] >> 2;
  int cands[16], nc = 0;
  if (plane == 0) {
    int set = get_tx_set(txSz, b->is_inter, fp->reduced_tx_set);
    if (!fp->tx_search || set == TX_SET_DCTONLY || fp->lossless) cands[nc++] = DCT_DCT;
    else
      for (int t = 0; t < 16; t++)
        if (in_set(set, t, b->is_inter)) cands[nc++] = t;
  } else {
    cands[nc++] = chroma_tx_type(txSz, b->is_inter, b->uv_mode, b->tx_type, fp->reduced_tx_set);
  }
  int64_t total = 0;
  int txIdx = 0, anyNz = 0;
  for (int y = 0; y < n4h; y += stepY)
    for (int x = 0; x < n4w; x += stepX) {
      int sx = baseX + 4 * x, sy = baseY + 4 * y;
      if (sx >= ((fp->mi_cols * 4) >> sub) || sy >= ((fp->mi_rows * 4) >> sub)) continue;
      int row = (sy << sub) >> 2, col = (sx << sub) >> 2, sbr = row & 15, sbc = col & 15;
      if (!b->is_inter) {
        OIntraArgs a;
        a.haveLeft = availL || x > 0;
        a.haveAbove = availU || y > 0;
        a.haveAboveRight = BDF(bd, plane, (sbr >> sub) - 1, (sbc >> sub) + stepX);
        a.haveBelowLeft = BDF(bd, plane, (sbr >> sub) + stepY, (sbc >> sub) - 1);
        int isCfl = plane > 0 && b->uv_mode == UV_CFL_PRED;
        a.mode = plane == 0 ? b->y_mode : (isCfl ? DC_PRED : b->uv_mode);
        a.angleDelta = plane == 0 ? b->angle_y : b->angle_uv;
        a.log2W = Tx_Width_Log2[txSz];
        a.log2H = Tx_Height_Log2[txSz];
        a.maxX = ((fp->mi_cols * 4) >> sub) - 1;
        a.maxY = ((fp->mi_rows * 4) >> sub) - 1;
        a.edgeFilter = fp->enable_intra_edge_filter;
        a.filterType = (b->interp_filter >> (plane == 0 ? 4 : 5)) & 1;
        uint8_t *d = px + sy * S + sx;
        o_predict_intra(px, S, sx, sy, &a, d, S);
        if (isCfl) {
          int tw = Tx_Width[txSz], th = Tx_Height[txSz], dc[32 * 32], alpha;
          for (int i = 0; i < th; i++)
            for (int j = 0; j < tw; j++) dc[i * tw + j] = d[i * S + j];
          const uint8_t *luma = rec + fp->offset[0];
          if (txIdx > 0) alpha = plane == 1 ? b->cfl_alpha_u : b->cfl_alpha_v;
          else if (plane == 1)
            alpha = b->cfl_alpha_u = o_cfl_best_alpha(luma, fp->stride[0], sx, sy, *maxLW, *maxLH, a.log2W, a.log2H,
                                                      dc, sp + sy * S + sx, S, 0);
          else
            alpha = b->cfl_alpha_v = o_cfl_best_alpha(luma, fp->stride[0], sx, sy, *maxLW, *maxLH, a.log2W, a.log2H,
                                                      dc, sp + sy * S + sx, S, b->cfl_alpha_u == 0);
          o_cfl_predict(luma, fp->stride[0], sx, sy, *maxLW, *maxLH, a.log2W, a.log2H, alpha, dc, d, S);
        }
      }
      if (plane == 0) {
        *maxLW = sx + stepX * 4;
        *maxLH = sy + stepY * 4;
      }
      int t = DCT_DCT;
      int64_t c = 0;
      int eob = code_tx_block(fp, sp, px, S, sx, sy, txSz, plane, cands, nc,
                              coefs + b->coef_off[plane] + txIdx * Tx_Width[txSz] * Tx_Height[txSz], mats, &t, &c);
      total += c;
      if (plane == 0) {
        set_tx_type(b, txIdx, t);
        if (txIdx == 0) b->tx_type = t;
      }
      txIdx++;
      anyNz |= eob > 0;
      for (int i = 0; i < stepY; i++)
        for (int j = 0; j < stepX; j++) BDF(bd, plane, (sbr >> sub) + i, (sbc >> sub) + j) = 1;
    }
  b->eob[plane] = anyNz;
  return total;
}

/* k_recon.h:293-394 */
static void recon_block(const Mav1FrameParams *fp, Mav1BlockInfo *b, const uint8_t *src, uint8_t *rec,
                        const uint8_t *ref, const uint8_t *gold, int32_t *coefs, const int32_t *mats, uint8_t *bd) {
  int bs = b->bsize, comp = b->is_inter && b->ref_frame1 > INTRA_FRAME;
  const uint8_t *r0 = (b->is_inter && b->ref_frame != LAST_FRAME) ? gold : ref;
  int maxLW = 0, maxLH = 0;
  b->tx_depth = 0;
  b->interp_filter &= 0x30;
  if (b->is_inter && fp->switchable_interp && !comp) {
    int w = BLOCK_WIDTH(bs), h = BLOCK_HEIGHT(bs), x0 = b->mi_col * 4, y0 = b->mi_row * 4, S = fp->stride[0];
    uint8_t *px = rec + fp->offset[0] + y0 * S + x0;
    const uint8_t *sp = src + fp->offset[0] + y0 * S + x0;
    int vw = o_min(w, fp->frame_w - x0), vh = o_min(h, fp->frame_h - y0);
    int64_t best = -1;
    for (int f = 0; f < 9; f++) {
      int packed = (f % 3) | ((f / 3) << 2);
      oracle_predict_inter(r0 + fp->offset[0], S, fp->frame_w, fp->frame_h, px, S, x0, y0, w, h, b->mv_row, b->mv_col,
                           0, 0, packed);
      int64_t sse = 0;
      for (int i = 0; i < vh; i++)
        for (int j = 0; j < vw; j++) {
          int d = (int)px[i * S + j] - (int)sp[i * S + j];
          sse += d * d;
        }
      int64_t c = sse * 16 + (int64_t)fp->lambda * ((f % 3 != 0) + (f / 3 != 0));
      if (best < 0 || c < best) {
        best = c;
        b->interp_filter = packed;
      }
    }
  }
  for (int plane = 0; plane < 3; plane++) {
    int sub = plane > 0, S = fp->stride[plane];
    int baseX = (b->mi_col >> sub) * 4, baseY = (b->mi_row >> sub) * 4;
    int w = BLOCK_WIDTH(bs) >> sub, h = BLOCK_HEIGHT(bs) >> sub;
    if (comp) {
      static _Thread_local int32_t p0[64 * 64], p1[64 * 64];
      o_predict_inter_raw(r0 + fp->offset[plane], S, fp->frame_w, fp->frame_h, p0, w, baseX, baseY, w, h, b->mv_row,
                          b->mv_col, sub, sub, b->interp_filter);
      o_predict_inter_raw(gold + fp->offset[plane], S, fp->frame_w, fp->frame_h, p1, w, baseX, baseY, w, h, b->mv1_row,
                          b->mv1_col, sub, sub, b->interp_filter);
      uint8_t *d = rec + fp->offset[plane] + baseY * S + baseX;
      for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) d[i * S + j] = (uint8_t)o_clip1(o_round2(p0[i * w + j] + p1[i * w + j], 5));
    } else if (b->is_inter) {
      oracle_predict_inter(r0 + fp->offset[plane], S, fp->frame_w, fp->frame_h, rec + fp->offset[plane] + baseY * S + baseX,
                           S, baseX, baseY, w, h, b->mv_row, b->mv_col, sub, sub, b->interp_filter);
    }
    if (fp->lossless) {
      code_plane(fp, b, src, rec, coefs, mats, bd, plane, TX_4X4, &maxLW, &maxLH);
      continue;
    }
    if (plane > 0) {
      code_plane(fp, b, src, rec, coefs, mats, bd, plane, uv_tx_size(bs), &maxLW, &maxLH);
      continue;
    }
    int maxDepth = (!b->is_inter && fp->tx_mode_select) ? o_min(MAX_TX_DEPTH, Max_Tx_Depth[bs]) : 0;
    if (maxDepth == 0) {
      code_plane(fp, b, src, rec, coefs, mats, bd, 0, Max_Tx_Size_Rect[bs], &maxLW, &maxLH);
      continue;
    }
    /* tx_depth RD: code luma at each depth from the same state, keep the best */
    uint8_t *px = rec + fp->offset[0] + baseY * S + baseX;
    static _Thread_local uint8_t save[64 * 64];
    uint8_t bdSave[BDD * BDD];
    for (int i = 0; i < h; i++) memcpy(save + i * w, px + i * S, (size_t)w);
    memcpy(bdSave, bd, sizeof bdSave);
    int64_t best = -1;
    int bestD = 0;
    for (int d = 0; d <= maxDepth + 1; d++) {
      int depth = d <= maxDepth ? d : bestD;
      if (d == maxDepth + 1 && bestD == maxDepth) break;
      if (d > 0) {
        for (int i = 0; i < h; i++) memcpy(px + i * S, save + i * w, (size_t)w);
        memcpy(bd, bdSave, sizeof bdSave);
      }
      int t = Max_Tx_Size_Rect[bs];
      for (int k = 0; k < depth; k++) t = Split_Tx_Size[t];
      int64_t c = code_plane(fp, b, src, rec, coefs, mats, bd, 0, t, &maxLW, &maxLH) + (int64_t)fp->lambda * 8 * depth;
      if (d <= maxDepth && (best < 0 || c < best)) {
        best = c;
        bestD = depth;
      }
      if (d == maxDepth + 1) break;
    }
    b->tx_depth = bestD;
  }
}

/* k_recon.h:396-417: BlockDecoded flags of the superblock, then its blocks. */
static void recon_sb(const Mav1FrameParams *fp, const Mav1SbInfo *sb, Mav1BlockInfo *blocks, const uint8_t *src,
                     uint8_t *rec, const uint8_t *ref, const uint8_t *gold, int32_t *coefs, const int32_t *mats) {
  uint8_t bd[3 * BDD * BDD];
  int r = sb->sb_row * 16, c = sb->sb_col * 16;
  int tileEnd = o_min(fp->mi_cols, ((c >> 4) / fp->tile_w_sb) * fp->tile_w_sb * 16 + fp->tile_w_sb * 16);
  for (int p = 0; p < 3; p++) {
    int sub = p > 0, w4 = (tileEnd - c) >> sub, h4 = (fp->mi_rows - r) >> sub;
    for (int y = -1; y <= (16 >> sub); y++)
      for (int x = -1; x <= (16 >> sub); x++)
        BDF(bd, p, y, x) = (y < 0 && x < w4) ? 1 : (x < 0 && y < h4) ? 1 : 0;
    BDF(bd, p, 16 >> sub, -1) = 0;
  }
  for (int i = 0; i < sb->num_blocks; i++)
    recon_block(fp, blocks + sb->first_block + i, src, rec, ref, gold, coefs, mats, bd);
}

void oracle_recon(const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs, const uint8_t *src,
                  uint8_t *rec, const uint8_t *ref, const uint8_t *gold, int32_t *coefs, const int32_t *mats) {
  /* Raster order, run as a wavefront: a superblock reads the reconstruction
   * of its left, above-left, above and above-right neighbours only, and each
   * of those has a smaller col + 2 * row, so the superblocks of one such
   * diagonal are independent once the earlier diagonals are done. */
  const int nd = fp->sb_cols + 2 * (fp->sb_rows - 1), nsb = fp->sb_cols * fp->sb_rows;
  for (int d = 0; d < nd; d++) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nsb; s++)
      if (sbs[s].sb_col + 2 * sbs[s].sb_row == d) recon_sb(fp, &sbs[s], blocks, src, rec, ref, gold, coefs, mats);
  }
}

/* Intra edge filter types (get_filter_type, spec 7.11.2.8) of every block
 * into interp_filter bits 4 (luma) and 5 (chroma): a neighbour above or left
 * (within the tile column) is an intra block with a smooth mode
 * (src/encoder.c:319-346). */
void oracle_edge_types(const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs) {
  int cols = fp->mi_cols, rows = fp->mi_rows, nsb = fp->sb_cols * fp->sb_rows;
  uint8_t *ys = calloc((size_t)rows * cols, 1), *uvs = calloc((size_t)rows * cols, 1);
  for (int s = 0; s < nsb; s++)
    for (int i = 0; i < sbs[s].num_blocks; i++) {
      const Mav1BlockInfo *b = &blocks[sbs[s].first_block + i];
      int sm_y = !b->is_inter && (b->y_mode == SMOOTH_PRED || b->y_mode == SMOOTH_V_PRED || b->y_mode == SMOOTH_H_PRED);
      int sm_uv = !b->is_inter && (b->uv_mode == SMOOTH_PRED || b->uv_mode == SMOOTH_V_PRED || b->uv_mode == SMOOTH_H_PRED);
      for (int r = b->mi_row; r < b->mi_row + Num_4x4_Blocks_High[b->bsize] && r < rows; r++)
        for (int c = b->mi_col; c < b->mi_col + Num_4x4_Blocks_Wide[b->bsize] && c < cols; c++) {
          ys[r * cols + c] = (uint8_t)sm_y;
          uvs[r * cols + c] = (uint8_t)sm_uv;
        }
    }
  for (int s = 0; s < nsb; s++)
    for (int i = 0; i < sbs[s].num_blocks; i++) {
      Mav1BlockInfo *b = &blocks[sbs[s].first_block + i];
      int r = b->mi_row, c = b->mi_col;
      int au = r > 0, al = c > ((c >> 4) / fp->tile_w_sb) * fp->tile_w_sb * 16;
      int fy = (au && ys[(r - 1) * cols + c]) || (al && ys[r * cols + c - 1]);
      int fuv = (au && uvs[(r - 1) * cols + o_min(c + 1, cols - 1)]) || (al && uvs[o_min(r + 1, rows - 1) * cols + c - 1]);
      b->interp_filter = (b->interp_filter & 15) | fy << 4 | fuv << 5;
    }
  free(ys);
  free(uvs);
}
