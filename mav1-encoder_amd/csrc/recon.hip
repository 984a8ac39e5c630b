/* Closed-loop reconstruction on gfx950 (behaviour of src/kernels/k_recon.h:
 * recon_sb / recon_block / code_plane / code_tx_block, in decode order, with
 * the dependencies of src/backend_cpu.c:81-86) and the intra edge filter
 * types of the frame driver (src/encoder.c:319-346).
 *
 * Inter blocks never read reconstructed neighbours: k_recon_list gathers
 * every inter block of the frame into a work list, and the persistent
 * 512-thread workgroups of k_recon_inter take blocks off it with an atomic
 * counter. Intra blocks read the reconstruction of earlier blocks: the
 * intra phase is one dataflow launch. k_intra_sched lists the superblocks
 * holding intra blocks in an order where every neighbour a superblock reads
 * (left, above-left, above, above-right within its tile column) comes
 * first; the persistent workgroups of k_recon_intra take list entries and
 * wait (bounded, with an error word) for the done flags of those
 * neighbours, then walk the superblock's blocks in decode order with the
 * BlockDecoded flags and the reconstruction tiles in LDS.
 *
 * Inside a workgroup the candidates of every search run side by side: the
 * transform types of a block of at most 16x16 one per wave (each wave codes
 * its candidate alone: residual, forward transform, quantiser, eob and
 * rate, inverse transform and distortion), the nine dual-filter pairs over
 * three shared horizontal passes, the 33 CfL alphas over the waves. Larger
 * transforms (32- and 64-point) use the whole workgroup per candidate. */
#include <algorithm>
#include <climits>

#include "dev_common.h"

/* The small AV1 tables that these kernels index with per-lane values live
 * in LDS: in constant memory each such lookup is a vector load through the
 * L2, and the stages chain several of them. Every workgroup copies them in
 * first (stage_tabs); the names below then refer to the copies. */
struct RTabs {
  int16_t m2a[13], dr[90], iek[15], n4w[22], n4h[22], maxRect[22], maxDepth[22], split[19], subs[88];
  int16_t txw[19], txh[19], txwl[19], txhl[19], sqr[19], sqrUp[19], rowShift[19], inIntra[48], inInter[64], m2t[14];
  uint8_t sw[124];
};
__shared__ RTabs gT;
template <class T>
DEVI void tab_copy(int16_t *d, const T *s, int n, int tid, int nt) {
  for (int k = tid; k < n; k += nt) d[k] = (int16_t)s[k];
}
DEVI void stage_tabs(int tid, int nt) {
  tab_copy(gT.m2a, Mode_To_Angle, 13, tid, nt);
  tab_copy(gT.dr, Dr_Intra_Derivative, 90, tid, nt);
  tab_copy(gT.iek, Intra_Edge_Kernel, 15, tid, nt);
  tab_copy(gT.n4w, Num_4x4_Blocks_Wide, 22, tid, nt);
  tab_copy(gT.n4h, Num_4x4_Blocks_High, 22, tid, nt);
  tab_copy(gT.maxRect, Max_Tx_Size_Rect, 22, tid, nt);
  tab_copy(gT.maxDepth, Max_Tx_Depth, 22, tid, nt);
  tab_copy(gT.split, Split_Tx_Size, 19, tid, nt);
  tab_copy(gT.subs, Subsampled_Size, 88, tid, nt);
  tab_copy(gT.txw, Tx_Width, 19, tid, nt);
  tab_copy(gT.txh, Tx_Height, 19, tid, nt);
  tab_copy(gT.txwl, Tx_Width_Log2, 19, tid, nt);
  tab_copy(gT.txhl, Tx_Height_Log2, 19, tid, nt);
  tab_copy(gT.sqr, Tx_Size_Sqr, 19, tid, nt);
  tab_copy(gT.sqrUp, Tx_Size_Sqr_Up, 19, tid, nt);
  tab_copy(gT.rowShift, Transform_Row_Shift, 19, tid, nt);
  tab_copy(gT.inIntra, Tx_Type_In_Set_Intra, 48, tid, nt);
  tab_copy(gT.inInter, Tx_Type_In_Set_Inter, 64, tid, nt);
  tab_copy(gT.m2t, Mode_To_Txfm, 14, tid, nt);
  const uint8_t *w[5] = {Sm_Weights_Tx_4x4, Sm_Weights_Tx_8x8, Sm_Weights_Tx_16x16, Sm_Weights_Tx_32x32,
                         Sm_Weights_Tx_64x64};
  for (int k = tid; k < 124; k += nt) {
    const int lg = k < 4 ? 0 : k < 12 ? 1 : k < 28 ? 2 : k < 60 ? 3 : 4, base = (4 << lg) - 4;
    gT.sw[k] = w[lg][k - base];
  }
}
#define Mode_To_Angle gT.m2a
#define Dr_Intra_Derivative gT.dr
#define Intra_Edge_Kernel gT.iek
#define Num_4x4_Blocks_Wide gT.n4w
#define Num_4x4_Blocks_High gT.n4h
#define Max_Tx_Size_Rect gT.maxRect
#define Max_Tx_Depth gT.maxDepth
#define Split_Tx_Size gT.split
#define Subsampled_Size gT.subs
#define Tx_Width gT.txw
#define Tx_Height gT.txh
#define Tx_Width_Log2 gT.txwl
#define Tx_Height_Log2 gT.txhl
#define Tx_Size_Sqr gT.sqr
#define Tx_Size_Sqr_Up gT.sqrUp
#define Transform_Row_Shift gT.rowShift
#define Tx_Type_In_Set_Intra gT.inIntra
#define Tx_Type_In_Set_Inter gT.inInter
#define Mode_To_Txfm gT.m2t
#define Sm_Weights_Tx_4x4 (gT.sw + 0)
#define Sm_Weights_Tx_8x8 (gT.sw + 4)
#define Sm_Weights_Tx_16x16 (gT.sw + 12)
#define Sm_Weights_Tx_32x32 (gT.sw + 28)
#define Sm_Weights_Tx_64x64 (gT.sw + 60)

#include "dev_pred.h"
#include "dev_tx.h"
#include "internal.h"

#define RT 512       /* threads of a reconstruction workgroup */
#define RW (RT / 64) /* its waves */
#define BDD 18
#define BDI(p, y, x) ((p) * BDD * BDD + ((y) + 1) * BDD + ((x) + 1))

#define SRS 17 /* row strides of a small (<= 16x16) block: residual, dequantised */
#define SQS 17
#define BRS 65 /* and of a large one */
#define BQS 33

/* One wave's working set for a transform block of at most 16x16. */
struct SmallTx {
  int32_t resid[16 * SRS]; /* residual, then the inverse transform's output */
  int32_t tmp[256];
  int32_t coef[256];
  int32_t q[256];
  int32_t dq[16 * SQS];
};
/* The workgroup's working set for a larger transform block. */
struct BigTx {
  int32_t resid[64 * BRS];
  int32_t tmp[64 * 32];
  int32_t coef[32 * 32];
  int32_t qtry[32 * 32];
  int32_t qbest[32 * 32];
  int32_t dq[32 * BQS];
};
/* Inter prediction: the reference window and the horizontal pass of up to
 * three filters; compound: the first prediction and a band of the second. */
struct PredLds {
  uint8_t win[71 * 72];
  int16_t mid[3][71 * 64];
};
struct CompLds {
  uint8_t win[71 * 72];
  int16_t mid[71 * 64];
  int32_t p0[64 * 64];
  int32_t p1[16 * 64];
};

#define TS_Y 132 /* tile row strides: 4 left columns + 128 (luma) / 64 (chroma) */
#define TS_C 68
#define TILE_Y (65 * TS_Y)
#define TILE_C (33 * TS_C)
#define TILE_BYTES (TILE_Y + 2 * TILE_C)

/* Per-wave LDS of the small intra blocks' prediction (edges, CfL). */
#define SI_EDGE (EDGE_NEG + 2 * 32 + 16)
struct SiWave {
  int eA[SI_EDGE], eL[SI_EDGE], ecp[2 * 32 + 40];
  int cfl[64], cflAvg, cflDc;
  long long aSse[33];
};

/* The scans of the nine transform sizes up to 16x16, back to back. */
#define SCAN_SMALL 784

struct RLds {
  union {
    SmallTx sm[RW];
    BigTx big;
    PredLds pr;
    CompLds cp;
    int32_t cfl[32 * 32];
  } u;
  int32_t qbest[256];
  long long candCost[16];
  int candEob[16];
  long long aSse[33];
  long long fcost[9];
  long long red[RW];
  int edgeA[EDGE_NEG + 2 * 64 + 32], edgeL[EDGE_NEG + 2 * 64 + 32], ecopy[2 * 64 + 40];
  uint8_t bd[3 * BDD * BDD];
  uint8_t bdSave[BDD * BDD];
  uint8_t save[64 * 64];
  uint8_t tsrc[64 * 64], tpred[64 * 64]; /* the transform block's source and prediction */
  int32_t smats[TX_SMATS];
  int16_t scans[3 * SCAN_SMALL]; /* scans of the transforms up to 16x16: default, mrow, mcol */
  int16_t iscans[3 * SCAN_SMALL]; /* their inverses: the scan position of each coefficient */
  float qrcp[3][2];         /* per plane the quantiser reciprocals of dc_q / ac_q */
  int ival[4];
  Mav1BlockInfo blk;
  /* intra phase: the superblock's reconstruction with the row above it and
   * four columns left of it (luma row above: 128 samples wide, for the
   * above-right edge), and its source in the same layout */
  alignas(16) uint8_t tile[2][TILE_BYTES];
  SiWave si[3];             /* the small-block path: luma wave, U wave, V wave */
  Mav1BlockInfo slot[2];    /* its blocks in flight */
  int mq[64][3], nmq;       /* inter blocks whose chroma flags follow the pending chroma: bsize, mi_row, mi_col */
  /* the level schedule of a superblock whose intra blocks all take the
   * small-block path: its block records, per block the dependency level
   * (0: inter), the BlockDecoded bits its prediction reads (above-right,
   * below-left per plane) and the schedule (intra blocks by level) */
  Mav1BlockInfo sbb[64];
  int plv[64], pav[64], pord[64], pn, pok;
  SiWave psw[12];           /* luma predictions of a step (4), chroma of the step before (8) */
  int32_t pq[4][256];       /* per luma block of a step: the chosen levels */
  long long pcost[4];       /* running best cost, eob, type and candidate wave */
  int peob[4], ptype[4], pwin[4];
  long long rcost[RW];      /* this round's candidates: cost and eob per wave */
  int reob[RW];
  int ufl[64];              /* per block: U's CfL alpha is set (V's CfL waits for it) */
  int lerr;                 /* a wait for ufl ran out */
  /* the dataflow form: per block the earlier intra blocks its prediction
   * reads (bit j: block j), and the blocks whose luma / U / V are done */
  unsigned long long dmask[64];
  unsigned long long ldm, udm, vdm;
  int psingle;              /* every intra block of the superblock has a single transform type */
};

/* The working set of the workgroup that reconstructs a block / superblock.
 * File scope so that the non-inlined stage functions address it as LDS. */
__shared__ RLds gL;
#define NOINL __device__ __forceinline__

/* Pointer p into tile k (0: reconstruction, 1: source) of `plane` with
 * p[y * S + x] = sample (x, y) of the plane, for the superblock holding mi
 * (miRow, miCol); S = the tile's row stride. */
typedef __attribute__((address_space(3))) uint8_t lds_u8;
DEVI uint8_t *tile_ptr(int k, int plane, int miRow, int miCol, int &S) {
  const int sub = plane > 0;
  S = sub ? TS_C : TS_Y;
  lds_u8 *base = (lds_u8 *)(gL.tile[k] + (plane == 0 ? 0 : TILE_Y + (plane - 1) * TILE_C));
  const int oy = (((miRow >> 4) * 64) >> sub) - 1, ox = (((miCol >> 4) * 64) >> sub) - 4;
  /* the offset is taken off the 32-bit LDS address, so the pointer stays in
   * the LDS address space and its accesses are LDS instructions */
  return (uint8_t *)(lds_u8 *)(uintptr_t)((uint32_t)(uintptr_t)base - (uint32_t)(oy * S + ox));
}

/* The lane index is opaque (d_local) so that the address arithmetic derived
 * from it stays inside each transform stage. */
DEVI TxCtx wg_ctx() {
  TxCtx c;
  c.tid = d_local(threadIdx.x);
  c.nt = RT;
  c.wg = true;
  c.red = gL.red;
  return c;
}
DEVI TxCtx wave_ctx() {
  TxCtx c;
  c.tid = d_local(threadIdx.x & 63);
  c.nt = 64;
  c.wg = false;
  c.red = nullptr;
  return c;
}
DEVI long long wg_sum64(long long v) { return tx_sum64(wg_ctx(), v); }
/* A workgroup barrier that orders LDS only: global stores before it (levels,
 * block records) need not have completed, as nothing in the workgroup reads
 * them back before the superblock is published with an agent-scope release. */
DEVI void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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
/* Offset of a transform size's scan among the staged small scans, -1 for
 * the 32- and 64-point sizes. */
DEVI int small_scan_off(int txSz) {
  switch (txSz) {
    case TX_4X4: return 0;
    case TX_8X8: return 16;
    case TX_16X16: return 80;
    case TX_4X8: return 336;
    case TX_8X4: return 368;
    case TX_8X16: return 400;
    case TX_16X8: return 528;
    case TX_4X16: return 656;
    case TX_16X4: return 720;
    default: return -1;
  }
}
DEVI int scan_kind(int t) {
  return (t == V_DCT || t == V_ADST || t == V_FLIPADST) ? 1 : (t == H_DCT || t == H_ADST || t == H_FLIPADST) ? 2 : 0;
}
DEVI const int16_t *get_scan(int txSz, int t) {
  if (txSz == TX_16X64) return Default_Scan_16x32;
  if (txSz == TX_64X16) return Default_Scan_32x16;
  if (Tx_Size_Sqr_Up[txSz] == TX_64X64) return Default_Scan_32x32;
  const int kind = scan_kind(t);
  const int off = small_scan_off(txSz);
  if (off >= 0) return gL.scans + kind * SCAN_SMALL + off;
  return scan_of(txSz, kind);
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

/* find_eob (k_recon.h:50-54): 1 + last scan position with a non-zero level. */
DEVI int tx_eob(const TxCtx &c, const int32_t *q_, const int16_t *scan, int n) {
  const lds_i32 *q = lds(q_);
  if (!c.wg) {
    /* one wave: the last non-zero scan position from a ballot per 64
     * positions (no cross-lane reduction) */
    int e = 0;
    for (int k0 = 0; k0 < n; k0 += 64) {
      const int k = k0 + c.tid;
      const unsigned long long nz = __ballot(k < n && q[scan[k]] != 0);
      if (nz) e = k0 + 64 - __clzll(nz);
    }
    return e;
  }
  int m = 0;
  for (int k = c.tid; k < n; k += c.nt)
    if (q[scan[k]]) m = k + 1;
  return tx_max(c, m);
}
/* est_coef_bits (k_recon.h:39-48) */
DEVI int tx_bits(const TxCtx &c, const int32_t *q_, const int16_t *scan, int eob) {
  if (!eob) return 8;
  const lds_i32 *q = lds(q_);
  long long b = 0;
  for (int k = c.tid; k < eob; k += c.nt) {
    const int l = d_abs(q[scan[k]]);
    b += l ? 8 * (2 + 2 * d_floorlog2((unsigned)l)) : 6;
  }
  return 48 + 16 * d_floorlog2((unsigned)eob + 1) + (int)tx_sum64(c, b);
}

/* Residual of transform block (x, y), stored in the type's flipped order. */
DEVI void tx_residual(const TxCtx &c, const uint8_t *spx, const uint8_t *px, int S, int x, int y, int lw, int lh, int t,
                      int32_t *resid_, int rs) {
  lds_i32 *resid = lds(resid_);
  const int w = 1 << lw, h = 1 << lh, fud = flip_ud(t), flr = flip_lr(t);
  for (int k = c.tid; k < w * h; k += c.nt) {
    const int i = k >> lw, j = k & (w - 1);
    const int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
    resid[i * rs + j] = (int)spx[o] - (int)px[o];
  }
  tx_sync(c);
}
/* px += residual (inverse transform output in flipped order), clipped. */
DEVI void tx_add(const TxCtx &c, uint8_t *px, int S, int x, int y, int lw, int lh, int t, const int32_t *res_, int rs) {
  const lds_i32 *res = lds(res_);
  const int w = 1 << lw, h = 1 << lh, fud = flip_ud(t), flr = flip_lr(t);
  for (int k = c.tid; k < w * h; k += c.nt) {
    const int i = k >> lw, j = k & (w - 1);
    const int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
    px[o] = (uint8_t)d_clip1((int)px[o] + res[i * rs + j]);
  }
  tx_sync(c);
}

/* One candidate type of one transform block (k_recon.h:60-173 loop body):
 * levels into q, returns the eob; the RD cost 16 * SSE + lambda * bits / 8
 * in *costOut (0 when `direct`: a single type and no depth search). */
template <bool SMALL>
NOINL int tx_eval(const TxCtx &c, const Mav1FrameParams &fp, const uint8_t *__restrict__ spx, const uint8_t *px, int S,
                  int x, int y, int txSz, int plane, int t, bool direct, int32_t *resid, int rs, int32_t *tmp,
                  int32_t *coef, int32_t *q, int32_t *dq, int qs, const int32_t *__restrict__ mats,
                  long long *costOut) {
  /* one transform block per call: its size, plane, type and placement are
   * the same in every lane */
  txSz = d_uni(txSz);
  plane = d_uni(plane);
  t = d_uni(t);
  x = d_uni(x);
  y = d_uni(y);
  S = d_uni(S);
  const int lw = d_uni(Tx_Width_Log2[txSz]), lh = d_uni(Tx_Height_Log2[txSz]), w = 1 << lw, h = 1 << lh;
  const int dcQ = fp.dc_q[plane], acQ = fp.ac_q[plane];
  tx_residual(c, spx, px, S, x, y, lw, lh, t, resid, rs);
  /* transforms up to 16x16: the scans staged in LDS, addressed as LDS; the
   * one-wave path takes the eob from the quantiser (the largest inverse-scan
   * position of a non-zero level, plus one: tx_eob's value) */
  const int16_t *scan = SMALL ? gL.scans + scan_kind(t) * SCAN_SMALL + small_scan_off(txSz) : get_scan(txSz, t);
  int eob;
  if constexpr (SMALL) {
    eob = d_uni(tx_forward_quant_small(c, resid, rs, tmp, q, txSz, t, gL.smats, dcQ, acQ, gL.qrcp[plane][0],
                                       gL.qrcp[plane][1], gL.iscans + scan_kind(t) * SCAN_SMALL + small_scan_off(txSz)));
  } else {
    tx_forward<SMALL>(c, resid, rs, tmp, coef, txSz, t, mats, gL.smats);
    tx_quantize(c, coef, q, txSz, dcQ, acQ, gL.qrcp[plane][0], gL.qrcp[plane][1]);
    eob = d_uni(tx_eob(c, q, scan, d_uni(seg_eob(txSz))));
  }
  if (direct) {
    *costOut = 0;
    return eob;
  }
  long long e = 0;
  const lds_i32 *rl = lds(resid);
  if (eob > 0) {
    tx_dequant(c, q, dq, qs, txSz, dcQ, acQ);
    tx_inverse<SMALL>(c, dq, qs, resid, rs, txSz, t);
    const int fud = flip_ud(t), flr = flip_lr(t);
    for (int k = c.tid; k < w * h; k += c.nt) {
      const int i = k >> lw, j = k & (w - 1);
      const int o = (y + (fud ? h - 1 - i : i)) * S + x + (flr ? w - 1 - j : j);
      const int d = (int)spx[o] - d_clip1((int)px[o] + rl[i * rs + j]);
      e += d * d;
    }
  } else {
    for (int k = c.tid; k < w * h; k += c.nt) {
      const int v = rl[(k >> lw) * rs + (k & (w - 1))];
      e += v * v;
    }
  }
  const long long sse = tx_sum64(c, e);
  *costOut = sse * 16 + (long long)fp.lambda * tx_bits(c, q, scan, eob) / 8;
  return eob;
}

/* Lossless 4x4: forward WHT (k_fwd.h:112-136) and inverse WHT (7.13.2.10),
 * one thread. */
DEVI void wht_fwd(const int32_t *r, int32_t *out) {
  int t[16];
  for (int j = 0; j < 4; j++) {
    int a = r[j], b = r[4 + j], c = r[8 + j], d = r[12 + j];
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

/* One transform block whose prediction is in px (k_recon.h:60-173): the
 * cheapest of the candidate types (first minimum), its levels into coefOut,
 * its reconstruction into px. Returns the eob; the type and cost in
 * *outType / *outCost. */
NOINL int code_tx_block(const Mav1FrameParams &fp, const uint8_t *__restrict__ spx, uint8_t *px, int S, int x, int y,
                        int txSz, int plane, const int *cands, int nc, int32_t *coefOut,
                        const int32_t *__restrict__ mats, int *outType, long long *outCost) {
  RLds &L = gL;
  const int tid = threadIdx.x;
  const int lw = Tx_Width_Log2[txSz], lh = Tx_Height_Log2[txSz], w = 1 << lw, h = 1 << lh;
  const int tw = d_min(32, w), th = d_min(32, h);
  const int dcQ = fp.dc_q[plane], acQ = fp.ac_q[plane];
  if (fp.lossless) {
    if (tid == 0) {
      int32_t r16[16], q16[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r16[i * 4 + j] = (int)spx[(y + i) * S + x + j] - (int)px[(y + i) * S + x + j];
      wht_fwd(r16, q16);
      int eob = 0;
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
            const int o = (y + i) * S + x + j;
            px[o] = (uint8_t)d_clip1((int)px[o] + r4[i * 4 + j]);
          }
      }
      L.ival[0] = eob;
    }
    __syncthreads();
    const int eob = L.ival[0];
    __syncthreads();
    *outType = DCT_DCT;
    *outCost = 0;
    return eob;
  }
  const bool direct = nc == 1 && !fp.tx_mode_select;
  /* up to 16x16: one candidate per wave; larger: the whole workgroup per
   * candidate. One code path for both keeps a single inlined copy of the
   * transform stages. */
  const bool small = w <= 16 && h <= 16;
  const int wave = tid >> 6, per = small ? RW : 1;
  const TxCtx c = small ? wave_ctx() : wg_ctx();
  int32_t *resid, *tmp, *coef, *qtry, *dq, *qbest;
  int rs, qs;
  if (small) {
    SmallTx &W = L.u.sm[wave];
    resid = W.resid, tmp = W.tmp, coef = W.coef, qtry = W.q, dq = W.dq, qbest = L.qbest, rs = SRS, qs = SQS;
  } else {
    BigTx &B = L.u.big;
    resid = B.resid, tmp = B.tmp, coef = B.coef, qtry = B.qtry, dq = B.dq, qbest = B.qbest, rs = BRS, qs = BQS;
  }
  long long bestCost = -1;
  int bestType = DCT_DCT, bestEob = 0;
  /* the candidates read the block's source and prediction from LDS */
  for (int k = tid; k < w * h; k += RT) {
    const int o = (y + (k >> lw)) * S + x + (k & (w - 1));
    L.tsrc[k] = spx[o];
    L.tpred[k] = px[o];
  }
  __syncthreads();
  for (int r0 = 0; r0 < nc; r0 += per) {
    const int ci = r0 + (small ? wave : 0);
    if (ci < nc) {
      long long cost;
      const int eob = tx_eval<false>(c, fp, L.tsrc, L.tpred, w, 0, 0, txSz, plane, cands[ci], direct, resid, rs, tmp, coef,
                              qtry, dq, qs, mats, &cost);
      if (c.tid == 0) {
        L.candCost[ci] = cost;
        L.candEob[ci] = eob;
      }
    }
    __syncthreads();
    /* first minimum in candidate order */
    int pick = -1;
    for (int k = r0; k < d_min(nc, r0 + per); k++)
      if (bestCost < 0 || L.candCost[k] < bestCost) {
        bestCost = L.candCost[k];
        pick = k;
      }
    if (pick >= 0) {
      bestEob = L.candEob[pick];
      bestType = bestEob > 0 ? cands[pick] : DCT_DCT;
      const int32_t *from = small ? L.u.sm[pick - r0].q : L.u.big.qtry;
      for (int k = tid; k < tw * th; k += RT) qbest[k] = from[k];
    }
    __syncthreads();
  }
  for (int k = tid; k < tw * th; k += RT) coefOut[k] = qbest[k];
  if (bestEob > 0 && (!small || wave == 0)) {
    tx_dequant(c, qbest, dq, qs, txSz, dcQ, acQ);
    tx_inverse(c, dq, qs, resid, rs, txSz, bestType);
    tx_add(c, px, S, x, y, lw, lh, bestType, resid, rs);
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
/* 7.11.2.12 on e[-1 .. sz-2], by the team c with copy space cp */
DEVI void edge_filter(const TxCtx &c, int *e, int *cp, int sz, int strength) {
  if (!strength) return;
  for (int i = c.tid; i < sz; i += c.nt) cp[i] = e[i - 1];
  tx_sync(c);
  for (int i = 1 + c.tid; i < sz; i += c.nt) {
    int s = 0;
#pragma unroll
    for (int j = 0; j < 5; j++) s += INTRA_EDGE_KERNEL(strength - 1, j) * cp[d_clip(0, sz - 1, i - 2 + j)];
    e[i - 1] = (s + 8) >> 4;
  }
  tx_sync(c);
}
/* 7.11.2.11 */
DEVI void edge_upsample(const TxCtx &c, int *e, int *dup, int numPx) {
  for (int i = c.tid - 1; i < numPx; i += c.nt) dup[i + 2] = e[i];
  if (c.tid == 0) {
    dup[0] = e[-1];
    dup[numPx + 2] = e[numPx - 1];
  }
  tx_sync(c);
  if (c.tid == 0) e[-2] = dup[0];
  for (int i = c.tid; i < numPx; i += c.nt) {
    int s = -dup[i] + 9 * dup[i + 1] + 9 * dup[i + 2] - dup[i + 3];
    e[2 * i - 1] = d_clip1(d_round2(s, 4));
    e[2 * i] = dup[i + 2];
  }
  tx_sync(c);
}

/* Intra prediction of one transform block into px by the team c (the
 * workgroup or one wave), edge arrays eA / eL (addressed from -EDGE_NEG)
 * and copy space ecp. */
DEVI void intra_tx_predict(const TxCtx &c, uint8_t *px, int S, int sx, int sy, int lw, int lh, int mode, int delta,
                           int hl, int ha, int har, int hbl, int maxX, int maxY, int edgeFilter, int filterType,
                           int *eA, int *eL, int *ecp) {
  /* one transform block per team: the block's parameters are the same in
   * every lane */
  S = d_uni(S);
  sx = d_uni(sx);
  sy = d_uni(sy);
  lw = d_uni(lw);
  lh = d_uni(lh);
  mode = d_uni(mode);
  delta = d_uni(delta);
  hl = d_uni(hl);
  ha = d_uni(ha);
  har = d_uni(har);
  hbl = d_uni(hbl);
  filterType = d_uni(filterType);
  const int tid = c.tid, w = 1 << lw, h = 1 << lh;
  int *A = eA + EDGE_NEG, *Lc = eL + EDGE_NEG;
  /* 7.11.2 steps 1-3: the neighbouring samples */
  {
    const int aLim = d_min(maxX, sx + (har ? 2 * w : w) - 1), lLim = d_min(maxY, sy + (hbl ? 2 * h : h) - 1);
    for (int i = tid; i < w + h; i += c.nt) {
      A[i] = ha ? px[(sy - 1) * S + d_min(aLim, sx + i)] : hl ? px[sy * S + sx - 1] : 127;
      Lc[i] = hl ? px[d_min(lLim, sy + i) * S + sx - 1] : ha ? px[(sy - 1) * S + sx] : 129;
    }
    if (tid == 0) {
      const int corner = (ha && hl) ? px[(sy - 1) * S + sx - 1] : ha ? px[(sy - 1) * S + sx] : hl ? px[sy * S + sx - 1] : 128;
      A[-1] = Lc[-1] = corner;
    }
    tx_sync(c);
  }
  IntraMode im;
  intra_mode_setup(im, mode, delta, lw, lh);
  im.pAngle = d_uni(im.pAngle);
  im.dx = d_uni(im.dx);
  im.dy = d_uni(im.dy);
  im.corner = d_uni(A[-1]);
  if (mode >= V_PRED && mode <= D67_PRED && edgeFilter) {
    const int pa = im.pAngle;
    if (pa != 90 && pa != 180) {
      if (pa > 90 && pa < 180 && w + h >= 24) {
        const int v = d_round2(Lc[0] * 5 + A[-1] * 6 + A[0] * 5, 4);
        tx_sync(c);
        if (tid == 0) A[-1] = Lc[-1] = v;
        tx_sync(c);
      }
      if (ha)
        edge_filter(c, A, ecp, d_min(w, maxX - sx + 1) + (pa < 90 ? h : 0) + 1, edge_strength(w, h, filterType, pa - 90));
      if (hl)
        edge_filter(c, Lc, ecp, d_min(h, maxY - sy + 1) + (pa > 180 ? w : 0) + 1,
                    edge_strength(w, h, filterType, pa - 180));
    }
    im.upA = use_upsample(w, h, filterType, pa - 90);
    if (im.upA) edge_upsample(c, A, ecp, w + (pa < 90 ? h : 0));
    im.upL = use_upsample(w, h, filterType, pa - 180);
    if (im.upL) edge_upsample(c, Lc, ecp, h + (pa > 180 ? w : 0));
  }
  if (mode == DC_PRED) {
    long long s = 0;
    for (int i = tid; i < w; i += c.nt) s += ha ? A[i] : 0;
    for (int i = tid; i < h; i += c.nt) s += hl ? Lc[i] : 0;
    const int sum = (int)tx_sum64(c, s);
    im.dcv = (ha && hl) ? (sum + ((w + h) >> 1)) / (w + h) : hl ? d_clip1((sum + (h >> 1)) >> lh)
             : ha ? d_clip1((sum + (w >> 1)) >> lw) : 128;
  } else {
    im.dcv = 0;
  }
  for (int k = tid; k < w * h; k += c.nt) {
    const int i = k >> lw, j = k & (w - 1);
    px[(sy + i) * S + sx + j] = (uint8_t)intra_px(im, A, Lc, i, j);
  }
  tx_sync(c);
}

/* CfL (7.11.5) over the DC prediction in px: the alpha with the smallest
 * SSE against the source (first minimum over -16..16) unless given; the
 * alphas are scored one per wave. */
NOINL int cfl_tx(const uint8_t *luma, int LS, int sx, int sy, int maxLW, int maxLH, int lw, int lh, uint8_t *px, int S,
                 const uint8_t *__restrict__ spx, int alphaIn, int search, int excludeZero) {
  RLds &L = gL;
  const int tid = threadIdx.x, w = 1 << lw, h = 1 << lh, n = w * h;
  const int dcv = px[sy * S + sx]; /* DC prediction: one value over the block */
  long long s = 0;
  for (int k = tid; k < n; k += RT) {
    const int i = k >> lw, j = k & (w - 1);
    const int ly = d_min((sy + i) << 1, maxLH - 2), lx = d_min((sx + j) << 1, maxLW - 2);
    const uint8_t *l = luma + ly * LS + lx;
    const int v = ((int)l[0] + l[1] + l[LS] + l[LS + 1]) << 1;
    L.u.cfl[k] = v;
    s += v;
  }
  const int avg = d_round2((int)wg_sum64(s), lw + lh);
  int alpha = alphaIn;
  if (search) {
    const int wave = tid >> 6, lane = tid & 63;
    for (int a0 = -16; a0 <= 16; a0 += RW) {
      const int a = a0 + wave;
      if (a <= 16) {
        long long e = 0;
        for (int k = lane; k < n; k += 64) {
          const int i = k >> lw, j = k & (w - 1);
          const int d = d_clip1(dcv + d_round2signed(a * (L.u.cfl[k] - avg), 6)) - spx[(sy + i) * S + sx + j];
          e += d * d;
        }
        e = wave_sum64(e);
        if (lane == 0) L.aSse[a + 16] = e;
      }
    }
    __syncthreads();
    long long best = -1;
    for (int a = -16; a <= 16; a++) {
      if (a == 0 && excludeZero) continue;
      if (best < 0 || L.aSse[a + 16] < best) {
        best = L.aSse[a + 16];
        alpha = a;
      }
    }
  }
  for (int k = tid; k < n; k += RT) {
    const int i = k >> lw, j = k & (w - 1);
    px[(sy + i) * S + sx + j] = (uint8_t)d_clip1(dcv + d_round2signed(alpha * (L.u.cfl[k] - avg), 6));
  }
  __syncthreads();
  return alpha;
}

/* code_plane (k_recon.h:190-291) for the block in L.blk. markOnly: only the
 * BlockDecoded flags (the block was coded by the inter phase). */
NOINL long long code_plane(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, uint8_t *rec, bool tiled,
                           int32_t *__restrict__ coefs, const int32_t *__restrict__ mats, int plane, int txSz,
                           int &maxLW, int &maxLH, bool useBd, bool markOnly) {
  RLds &L = gL;
  Mav1BlockInfo &b = L.blk;
  const int tid = threadIdx.x;
  const int bs = b.bsize, sub = plane > 0;
  const int availU = b.mi_row > 0, availL = b.mi_col > tile_start(fp, b.mi_col);
  const int psz = SUBSAMPLED_SIZE(bs, sub, sub);
  const int n4w = Num_4x4_Blocks_Wide[psz], n4h = Num_4x4_Blocks_High[psz];
  const int baseX = (b.mi_col >> sub) * 4, baseY = (b.mi_row >> sub) * 4;
  /* the reconstruction, source and luma (CfL) samples: the superblock's
   * LDS tiles in the intra phase, else the frame */
  int S, LS;
  uint8_t *px, *luma;
  const uint8_t *spx;
  if (tiled) {
    px = tile_ptr(0, plane, b.mi_row, b.mi_col, S);
    spx = tile_ptr(1, plane, b.mi_row, b.mi_col, S);
    luma = tile_ptr(0, 0, b.mi_row, b.mi_col, LS);
  } else {
    S = fp.stride[plane], LS = fp.stride[0];
    px = rec + fp.offset[plane], spx = src + fp.offset[plane], luma = rec + fp.offset[0];
  }
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
          intra_tx_predict(wg_ctx(), px, S, sx, sy, lw, lh, plane == 0 ? b.y_mode : (isCfl ? DC_PRED : b.uv_mode),
                           plane == 0 ? b.angle_y : b.angle_uv, availL || x > 0, availU || y > 0,
                           L.bd[BDI(plane, (sbr >> sub) - 1, (sbc >> sub) + stepX)],
                           L.bd[BDI(plane, (sbr >> sub) + stepY, (sbc >> sub) - 1)], ((fp.mi_cols * 4) >> sub) - 1,
                           ((fp.mi_rows * 4) >> sub) - 1, fp.enable_intra_edge_filter,
                           (b.interp_filter >> (plane == 0 ? 4 : 5)) & 1, L.edgeA, L.edgeL, L.ecopy);
          if (isCfl) {
            /* the first transform block searches the alpha, the others reuse it */
            const int search = txIdx == 0;
            const int a = cfl_tx(luma, LS, sx, sy, maxLW, maxLH, lw, lh, px, S, spx,
                                 search ? 0 : plane == 1 ? b.cfl_alpha_u : b.cfl_alpha_v, search,
                                 search && plane == 2 && b.cfl_alpha_u == 0);
            if (txIdx == 0 && tid == 0) {
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
                                      coefs + b.coef_off[plane] + txIdx * Tx_Width[txSz] * Tx_Height[txSz], mats, &t,
                                      &c);
        total += c;
        if (plane == 0 && tid == 0) {
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
      if (useBd && tid == 0)
        for (int i = 0; i < stepY; i++)
          for (int j = 0; j < stepX; j++) L.bd[BDI(plane, (sbr >> sub) + i, (sbc >> sub) + j)] = 1;
      __syncthreads();
    }
  if (!markOnly && tid == 0) b.eob[plane] = anyNz;
  __syncthreads();
  return total;
}

/* ---- inter prediction by the workgroup ---- */
DEVI void filter_taps(int w, int h, int filter, int posX, int posY, int *kh, int *kv) {
  int fv = filter & 3, fh = (filter >> 2) & 3;
  if (w <= 4) fh = (fh == EIGHTTAP_SMOOTH) ? 5 : (fh == EIGHTTAP || fh == EIGHTTAP_SHARP) ? 4 : fh;
  if (h <= 4) fv = (fv == EIGHTTAP_SMOOTH) ? 5 : (fv == EIGHTTAP || fv == EIGHTTAP_SHARP) ? 4 : fv;
  const int16_t *th = SUBPEL_FILTER(fh, (posX >> 6) & SUBPEL_MASK), *tv = SUBPEL_FILTER(fv, (posY >> 6) & SUBPEL_MASK);
#pragma unroll
  for (int t = 0; t < 8; t++) {
    kh[t] = th[t];
    kv[t] = tv[t];
  }
}
/* The (h + 7) x (w + 7) reference samples of the block (row stride w + 8),
 * clamped to the reference plane (7.11.3.3). */
DEVI void load_window(const uint8_t *__restrict__ plane, int S, int refW, int refH, int ix, int iy, int w, int h, int sub,
                      uint8_t *win, int tid) {
  const int lastX = ((refW + sub) >> sub) - 1, lastY = ((refH + sub) >> sub) - 1;
  const int WW = w + 8, nr = h + 7;
  for (int k = tid; k < nr * (w + 7); k += RT) {
    const int r = k / (w + 7), c = k - r * (w + 7);
    win[r * WW + c] = plane[d_clip(0, lastY, iy - 3 + r) * S + d_clip(0, lastX, ix - 3 + c)];
  }
}
DEVI void hpass(const uint8_t *win, int w, int h, const int *kh, int16_t *mid, int tid) {
  const int WW = w + 8;
  for (int k = tid; k < (h + 7) * w; k += RT) {
    const int r = k / w, c = k - r * w;
    const uint8_t *q = win + r * WW + c;
    int acc = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) acc += kh[t] * q[t];
    mid[k] = (int16_t)d_round2(acc, 3);
  }
}

/* Inter prediction of a w x h block of one plane (7.11.3): round1 11 writes
 * clipped bytes to dst (row stride ds), round1 7 (compound) the unclipped
 * int32 values to dst32. */
NOINL void predict_inter_plane(const uint8_t *__restrict__ plane, int S, int refW, int refH, int x, int y, int w, int h,
                               int mvr, int mvc, int sub, int filter, uint8_t *dst, int ds, int32_t *dst32, int round1,
                               uint8_t *win, int16_t *mid) {
  const int tid = threadIdx.x;
  const int posX = ((x << SUBPEL_BITS) + ((2 * mvc) >> sub)) * 64 + 32;
  const int posY = ((y << SUBPEL_BITS) + ((2 * mvr) >> sub)) * 64 + 32;
  int kh[8], kv[8];
  filter_taps(w, h, filter, posX, posY, kh, kv);
  load_window(plane, S, refW, refH, posX >> 10, posY >> 10, w, h, sub, win, tid);
  __syncthreads();
  hpass(win, w, h, kh, mid, tid);
  __syncthreads();
  for (int k = tid; k < h * w; k += RT) {
    const int r = k / w, c = k - r * w;
    int acc = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) acc += kv[t] * mid[(r + t) * w + c];
    if (round1 == 11) dst[r * ds + c] = (uint8_t)d_clip1(d_round2(acc, 11));
    else dst32[r * ds + c] = d_round2(acc, round1);
  }
  __syncthreads();
}

/* The dual-filter choice of a single-reference inter block (k_recon.h:
 * 312-334): SSE of the luma prediction of each of the 9 filter pairs over
 * the visible area, + lambda per non-regular filter, first minimum. The
 * window is shared, the three horizontal passes are computed once, each
 * wave scores pairs. */
NOINL int dual_filter_search(const Mav1FrameParams &fp, const uint8_t *__restrict__ refp,
                             const uint8_t *__restrict__ src) {
  RLds &L = gL;
  const Mav1BlockInfo &b = L.blk;
  const int tid = threadIdx.x, bs = b.bsize;
  const int w = BLOCK_WIDTH(bs), h = BLOCK_HEIGHT(bs), x0 = b.mi_col * 4, y0 = b.mi_row * 4, S = fp.stride[0];
  const int vw = d_min(w, fp.frame_w - x0), vh = d_min(h, fp.frame_h - y0);
  const int posX = ((x0 << SUBPEL_BITS) + 2 * b.mv_col) * 64 + 32, posY = ((y0 << SUBPEL_BITS) + 2 * b.mv_row) * 64 + 32;
  load_window(refp + fp.offset[0], S, fp.frame_w, fp.frame_h, posX >> 10, posY >> 10, w, h, 0, L.u.pr.win, tid);
  __syncthreads();
  for (int f = 0; f < 3; f++) {
    int kh[8], kv[8];
    filter_taps(w, h, f << 2, posX, posY, kh, kv);
    hpass(L.u.pr.win, w, h, kh, L.u.pr.mid[f], tid);
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const uint8_t *sp = src + fp.offset[0] + y0 * S + x0;
  for (int f = wave; f < 9; f += RW) {
    int kh[8], kv[8];
    filter_taps(w, h, f % 3, posX, posY, kh, kv);
    const int16_t *mid = L.u.pr.mid[f / 3];
    long long e = 0;
    for (int k = lane; k < vw * vh; k += 64) {
      const int r = k / vw, c = k - r * vw;
      int acc = 0;
#pragma unroll
      for (int t = 0; t < 8; t++) acc += kv[t] * mid[(r + t) * w + c];
      const int d = d_clip1(d_round2(acc, 11)) - (int)sp[r * S + c];
      e += d * d;
    }
    e = wave_sum64(e);
    if (lane == 0) L.fcost[f] = e * 16 + (long long)fp.lambda * ((f % 3 != 0) + (f / 3 != 0));
  }
  __syncthreads();
  long long best = -1;
  int bestF = 0;
  for (int f = 0; f < 9; f++)
    if (best < 0 || L.fcost[f] < best) {
      best = L.fcost[f];
      bestF = (f % 3) | ((f / 3) << 2);
    }
  __syncthreads();
  return bestF;
}

/* recon_block (k_recon.h:293-394) for L.blk. Inter blocks are fully coded
 * by the inter phase; the intra phase only marks their decoded area. */
NOINL void recon_block(const Mav1FrameParams &fp, const uint8_t *__restrict__ src, uint8_t *rec, bool tiled,
                       const uint8_t *__restrict__ ref, const uint8_t *__restrict__ gold, int32_t *__restrict__ coefs,
                       const int32_t *__restrict__ mats, bool useBd, bool markOnly) {
  RLds &L = gL;
  Mav1BlockInfo &b = L.blk;
  const int tid = threadIdx.x;
  const int bs = b.bsize, comp = b.is_inter && b.ref_frame1 > INTRA_FRAME;
  const uint8_t *r0 = (b.is_inter && b.ref_frame != LAST_FRAME) ? gold : ref;
  int maxLW = 0, maxLH = 0;
  if (!markOnly) {
    if (tid == 0) {
      b.tx_depth = 0;
      b.interp_filter &= 0x30;
    }
    __syncthreads();
    if (b.is_inter && fp.switchable_interp && !comp) {
      const int f = dual_filter_search(fp, r0, src);
      if (tid == 0) b.interp_filter = f;
      __syncthreads();
    }
  }
  for (int plane = 0; plane < 3; plane++) {
    const int sub = plane > 0, S = fp.stride[plane];
    const int baseX = (b.mi_col >> sub) * 4, baseY = (b.mi_row >> sub) * 4;
    const int w = BLOCK_WIDTH(bs) >> sub, h = BLOCK_HEIGHT(bs) >> sub;
    if (!markOnly) {
      if (comp) {
        CompLds &P = L.u.cp;
        predict_inter_plane(r0 + fp.offset[plane], S, fp.frame_w, fp.frame_h, baseX, baseY, w, h, b.mv_row, b.mv_col, sub,
                            b.interp_filter, nullptr, w, P.p0, 7, P.win, P.mid);
        /* the second prediction in bands of 16 rows, combined with the first */
        for (int yy = 0; yy < h; yy += 16) {
          const int hh = d_min(16, h - yy);
          predict_inter_plane(gold + fp.offset[plane], S, fp.frame_w, fp.frame_h, baseX, baseY + yy, w, hh, b.mv1_row,
                              b.mv1_col, sub, b.interp_filter, nullptr, w, P.p1, 7, P.win, P.mid);
          uint8_t *d = rec + fp.offset[plane] + (baseY + yy) * S + baseX;
          for (int k = tid; k < w * hh; k += RT) {
            const int i = k / w, j = k - i * w;
            d[i * S + j] = (uint8_t)d_clip1(d_round2(P.p0[(yy + i) * w + j] + P.p1[i * w + j], 5));
          }
          __syncthreads();
        }
      } else if (b.is_inter) {
        predict_inter_plane(r0 + fp.offset[plane], S, fp.frame_w, fp.frame_h, baseX, baseY, w, h, b.mv_row, b.mv_col,
                            sub, b.interp_filter, rec + fp.offset[plane] + baseY * S + baseX, S, nullptr, 11, L.u.pr.win,
                            L.u.pr.mid[0]);
      }
    }
    /* transform size: 4x4 when lossless, the chroma size, or the luma size
     * of the best tx_depth (RD search of intra luma only, never markOnly) */
    const int baseTx = fp.lossless ? TX_4X4 : plane > 0 ? uv_tx_size(bs) : Max_Tx_Size_Rect[bs];
    const int maxDepth =
        (plane == 0 && !fp.lossless && !b.is_inter && fp.tx_mode_select) ? d_min(MAX_TX_DEPTH, Max_Tx_Depth[bs]) : 0;
    int TS = S;
    uint8_t *px = (tiled ? tile_ptr(0, 0, b.mi_row, b.mi_col, TS) : rec + fp.offset[0]) + baseY * TS + baseX;
    if (maxDepth > 0) {
      for (int k = tid; k < w * h; k += RT) L.save[k] = px[(k / w) * TS + (k % w)];
      for (int k = tid; k < BDD * BDD; k += RT) L.bdSave[k] = L.bd[k];
      __syncthreads();
    }
    long long best = -1;
    int bestD = 0;
    for (int d = 0; d <= maxDepth + 1; d++) {
      const int depth = d <= maxDepth ? d : bestD;
      if (d == maxDepth + 1 && bestD == maxDepth) break;
      if (d > 0) {
        for (int k = tid; k < w * h; k += RT) px[(k / w) * TS + (k % w)] = L.save[k];
        for (int k = tid; k < BDD * BDD; k += RT) L.bd[k] = L.bdSave[k];
        __syncthreads();
      }
      int t = baseTx;
      for (int k = 0; k < depth; k++) t = Split_Tx_Size[t];
      const long long c = code_plane(fp, src, rec, tiled, coefs, mats, plane, t, maxLW, maxLH, useBd, markOnly) +
                          (long long)fp.lambda * 8 * depth;
      if (d <= maxDepth && (best < 0 || c < best)) {
        best = c;
        bestD = depth;
      }
      if (d == maxDepth + 1) break;
    }
    if (plane == 0 && tid == 0 && !markOnly) b.tx_depth = bestD;
    __syncthreads();
  }
}

/* ---- small intra blocks ----
 * Intra blocks of 8x8 .. 16x16 without a tx_depth search code one transform
 * block per plane. The intra phase pipelines them: while the workgroup codes
 * the luma of block i (prediction on wave 0, the transform types one per
 * wave, the chosen one on wave 0), waves SI_WU and SI_WV code the U and V
 * planes of block i-1. Each plane's BlockDecoded flags are set in decode
 * order by its own stream, and CfL reads the luma of its block, which the
 * previous step finished. */
#define SI_WU 6
#define SI_WV 7

DEVI bool si_eligible(const Mav1FrameParams &fp, const Mav1BlockInfo &b) {
  const int bs = b.bsize;
  return !b.is_inter && !fp.lossless && !fp.tx_mode_select &&
         (bs == BLOCK_8X8 || bs == BLOCK_8X16 || bs == BLOCK_16X8 || bs == BLOCK_16X16);
}

/* The BlockDecoded marks of one plane of block b, transform block by
 * transform block as code_plane sets them (one thread). */
DEVI void si_mark_g(const Mav1FrameParams &fp, int bs, int miRow, int miCol, int plane) {
  const int sub = plane > 0;
  const int txSz = fp.lossless ? TX_4X4 : plane ? uv_tx_size(bs) : Max_Tx_Size_Rect[bs];
  const int psz = SUBSAMPLED_SIZE(bs, sub, sub), n4w = Num_4x4_Blocks_Wide[psz], n4h = Num_4x4_Blocks_High[psz];
  const int baseX = (miCol >> sub) * 4, baseY = (miRow >> sub) * 4;
  const int stepX = Tx_Width[txSz] >> 2, stepY = Tx_Height[txSz] >> 2;
  for (int y = 0; y < n4h; y += stepY)
    for (int x = 0; x < n4w; x += stepX) {
      const int sx = baseX + 4 * x, sy = baseY + 4 * y;
      if (sx >= ((fp.mi_cols * 4) >> sub) || sy >= ((fp.mi_rows * 4) >> sub)) continue;
      const int sbr = ((sy << sub) >> 2) & 15, sbc = ((sx << sub) >> 2) & 15;
      for (int i = 0; i < stepY; i++)
        for (int j = 0; j < stepX; j++) gL.bd[BDI(plane, (sbr >> sub) + i, (sbc >> sub) + j)] = 1;
    }
}
DEVI void si_mark(const Mav1FrameParams &fp, const Mav1BlockInfo &b, int plane) {
  si_mark_g(fp, b.bsize, b.mi_row, b.mi_col, plane);
}

/* Geometry of plane `plane` of small block b. */
struct SiGeo {
  int sub, txSz, lw, lh, sx, sy, S;
  uint8_t *px;
  const uint8_t *spx;
};
DEVI SiGeo si_geo(const Mav1BlockInfo &b, int plane) {
  /* one block per wave: every field is the same in each lane */
  SiGeo g;
  plane = d_uni(plane);
  const int bs = d_uni(b.bsize), mr = d_uni(b.mi_row), mc = d_uni(b.mi_col);
  g.sub = plane > 0;
  g.txSz = d_uni(plane ? uv_tx_size(bs) : Max_Tx_Size_Rect[bs]);
  g.lw = d_uni(Tx_Width_Log2[g.txSz]);
  g.lh = d_uni(Tx_Height_Log2[g.txSz]);
  g.sx = (mc >> g.sub) * 4;
  g.sy = (mr >> g.sub) * 4;
  g.px = tile_ptr(0, plane, mr, mc, g.S);
  g.spx = tile_ptr(1, plane, mr, mc, g.S);
  return g;
}

/* Prediction of one plane of small block b by one wave (code_plane's
 * intra_tx_predict call); for CfL also the alphas' SSEs (cfl_tx), and for U
 * the chosen alpha applied. */
DEVI void si_predict(const Mav1FrameParams &fp, Mav1BlockInfo &b, int plane, SiWave &W, int har = -1, int hbl = -1) {
  const TxCtx c = wave_ctx();
  const int lane = c.tid;
  const SiGeo g = si_geo(b, plane);
  plane = d_uni(plane);
  const int sub = g.sub, stepX = d_uni(Tx_Width[g.txSz]) >> 2, stepY = d_uni(Tx_Height[g.txSz]) >> 2;
  const int sbr = ((g.sy << sub) >> 2) & 15, sbc = ((g.sx << sub) >> 2) & 15;
  const int mr = d_uni(b.mi_row), mc = d_uni(b.mi_col);
  const int availU = mr > 0, availL = mc > tile_start(fp, mc);
  const int uvMode = d_uni(b.uv_mode);
  const int isCfl = plane > 0 && uvMode == UV_CFL_PRED;
  intra_tx_predict(c, g.px, g.S, g.sx, g.sy, g.lw, g.lh, d_uni(plane == 0 ? b.y_mode : (isCfl ? DC_PRED : uvMode)),
                   d_uni(plane == 0 ? b.angle_y : b.angle_uv), availL, availU,
                   d_uni(har >= 0 ? har : gL.bd[BDI(plane, (sbr >> sub) - 1, (sbc >> sub) + stepX)]),
                   d_uni(hbl >= 0 ? hbl : gL.bd[BDI(plane, (sbr >> sub) + stepY, (sbc >> sub) - 1)]),
                   ((fp.mi_cols * 4) >> sub) - 1, ((fp.mi_rows * 4) >> sub) - 1, fp.enable_intra_edge_filter,
                   d_uni((b.interp_filter >> (plane == 0 ? 4 : 5)) & 1), W.eA, W.eL, W.ecp);
  if (!isCfl) return;
  /* CfL (7.11.5) over the DC prediction: the luma average of each sample,
   * then one alpha per lane */
  const int w = 1 << g.lw, h = 1 << g.lh, n = w * h;
  int LS;
  const uint8_t *luma = tile_ptr(0, 0, mr, mc, LS);
  const int lumaTx = d_uni(Max_Tx_Size_Rect[d_uni(b.bsize)]);
  const int maxLW = mc * 4 + d_uni(Tx_Width[lumaTx]), maxLH = mr * 4 + d_uni(Tx_Height[lumaTx]);
  const int dcv = g.px[g.sy * g.S + g.sx];
  int sum = 0;
  for (int k = lane; k < n; k += 64) {
    const int i = k >> g.lw, j = k & (w - 1);
    const int ly = d_min((g.sy + i) << 1, maxLH - 2), lx = d_min((g.sx + j) << 1, maxLW - 2);
    const uint8_t *l = luma + ly * LS + lx;
    const int v = ((int)l[0] + l[1] + l[LS] + l[LS + 1]) << 1;
    W.cfl[k] = v;
    sum += v;
  }
  const int avg = d_round2(wave_sum_dpp(sum), g.lw + g.lh);
  tx_sync(c);
  if (lane < 33) {
    const int a = lane - 16;
    long long e = 0;
    for (int k = 0; k < n; k++) {
      const int i = k >> g.lw, j = k & (w - 1);
      const int d = d_clip1(dcv + d_round2signed(a * (W.cfl[k] - avg), 6)) - g.spx[(g.sy + i) * g.S + g.sx + j];
      e += d * d;
    }
    W.aSse[lane] = e;
  }
  if (lane == 0) {
    W.cflAvg = avg;
    W.cflDc = dcv;
  }
  tx_sync(c);
}

/* CfL: the alpha of smallest SSE (first in -16..16, 0 excluded when asked),
 * applied to the DC prediction. */
DEVI int si_cfl_apply(const Mav1BlockInfo &b, int plane, SiWave &W, int excludeZero) {
  const TxCtx c = wave_ctx();
  const int lane = c.tid;
  const SiGeo g = si_geo(b, plane);
  const int w = 1 << g.lw, n = w << g.lh;
  long long v = (lane < 33 && !(excludeZero && lane == 16)) ? W.aSse[lane] : LLONG_MAX;
  int idx = lane;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const long long ov = __shfl_xor(v, m);
    const int oi = __shfl_xor(idx, m);
    if (ov < v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  const int alpha = d_uni(idx) - 16, avg = d_uni(W.cflAvg), dcv = d_uni(W.cflDc);
  tx_sync(c);
  for (int k = lane; k < n; k += 64) {
    const int i = k >> g.lw, j = k & (w - 1);
    g.px[(g.sy + i) * g.S + g.sx + j] = (uint8_t)d_clip1(dcv + d_round2signed(alpha * (W.cfl[k] - avg), 6));
  }
  tx_sync(c);
  return alpha;
}

/* The chroma plane of small block b by one wave: its single transform type
 * (code_plane), levels, reconstruction, flags. */
DEVI void si_code_chroma(const Mav1FrameParams &fp, Mav1BlockInfo &b, int plane, int32_t *__restrict__ coefs,
                         const int32_t *__restrict__ mats, SmallTx &T, bool mark = true) {
  const TxCtx c = wave_ctx();
  const SiGeo g = si_geo(b, plane);
  const int set = d_uni(get_tx_set(g.txSz, 0, fp.reduced_tx_set));
  int t = d_uni(Mode_To_Txfm[d_uni(b.uv_mode)]);
  if (d_uni(Tx_Size_Sqr_Up[g.txSz]) > TX_32X32 || !d_uni(in_set(set, t, 0))) t = DCT_DCT;
  long long cost;
  const int eob = tx_eval<true>(c, fp, g.spx, g.px, g.S, g.sx, g.sy, g.txSz, plane, t, true, T.resid, SRS, T.tmp, T.coef,
                          T.q, T.dq, SQS, mats, &cost);
  const int type = eob > 0 ? t : DCT_DCT;
  const int tw = d_min(32, 1 << g.lw), th = d_min(32, 1 << g.lh);
  int32_t *co = coefs + d_uni(b.coef_off[plane]);
  for (int k = c.tid; k < tw * th; k += 64) co[k] = T.q[k];
  if (eob > 0)
    tx_recon_small(c, T.q, T.resid, SRS, g.txSz, type, fp.dc_q[plane], fp.ac_q[plane], g.px, g.S, g.sx, g.sy);
  if (c.tid == 0) {
    b.eob[plane] = eob > 0;
    if (mark) si_mark(fp, b, plane);
  }
}

/* One pipeline step: luma of the block in slot ls (lk: 1 code, 2 mark
 * only, 0 none) and chroma of the block in slot cs (ck likewise). */
DEVI void si_step(const Mav1FrameParams &fp, int ls, int lk, int cs, int ck, int32_t *__restrict__ coefs,
                  const int32_t *__restrict__ mats) {
  RLds &L = gL;
  const int tid = threadIdx.x, wave = tid >> 6;
  /* P1: predictions (luma on wave 0, U and V on their waves) */
  const int cplane = wave == SI_WU ? 1 : wave == SI_WV ? 2 : 0;
  if ((lk == 1 && wave == 0) || (ck == 1 && cplane)) {
    Mav1BlockInfo &b = L.slot[cplane ? cs : ls];
    if (tid == 0) {
      b.tx_depth = 0;
      b.interp_filter &= 0x30;
    }
    si_predict(fp, b, cplane, L.si[cplane]);
    if (cplane == 1 && b.uv_mode == UV_CFL_PRED) {
      const int a = si_cfl_apply(b, 1, L.si[1], 0);
      if ((tid & 63) == 0) b.cfl_alpha_u = a;
    }
  }
  __syncthreads();
  /* P2: the luma transform types, one per wave */
  int bestType = DCT_DCT, bestEob = 0, bestWave = 0;
  bool reuse = false; /* the chosen candidate's wave holds its inverse transform */
  const SiGeo gy = lk == 1 ? si_geo(L.slot[ls], 0) : SiGeo{};
  if (lk == 1) {
    const Mav1BlockInfo &b = L.slot[ls];
    const int set = get_tx_set(gy.txSz, 0, fp.reduced_tx_set);
    int cands[16], nc = 0;
    if (!fp.tx_search || set == TX_SET_DCTONLY) cands[nc++] = DCT_DCT;
    else
      for (int t = 0; t < 16; t++)
        if (in_set(set, t, 0)) cands[nc++] = t;
    (void)b;
    const bool direct = nc == 1;
    reuse = !direct && nc <= RW;
    const int tw = d_min(32, 1 << gy.lw), th = d_min(32, 1 << gy.lh);
    const TxCtx c = wave_ctx();
    long long bestCost = -1;
    for (int r0 = 0; r0 < nc; r0 += RW) {
      const int ci = r0 + wave;
      if (ci < nc) {
        SmallTx &W = L.u.sm[wave];
        long long cost;
        const int eob = tx_eval<true>(c, fp, gy.spx, gy.px, gy.S, gy.sx, gy.sy, gy.txSz, 0, cands[ci], direct, W.resid, SRS,
                                W.tmp, W.coef, W.q, W.dq, SQS, mats, &cost);
        if (c.tid == 0) {
          L.candCost[ci] = cost;
          L.candEob[ci] = eob;
        }
      }
      __syncthreads();
      int pick = -1;
      for (int k = r0; k < d_min(nc, r0 + RW); k++)
        if (bestCost < 0 || L.candCost[k] < bestCost) {
          bestCost = L.candCost[k];
          pick = k;
        }
      if (pick >= 0) {
        bestEob = L.candEob[pick];
        bestType = bestEob > 0 ? cands[pick] : DCT_DCT;
        bestWave = pick - r0;
        const int32_t *from = L.u.sm[pick - r0].q;
        for (int k = tid; k < tw * th; k += RT) L.qbest[k] = from[k];
      }
      __syncthreads();
    }
  }
  /* P3: the chosen luma type on wave 0; chroma coding; marks */
  if (lk == 1 && wave == 0) {
    Mav1BlockInfo &b = L.slot[ls];
    const TxCtx c = wave_ctx();
    SmallTx &W = L.u.sm[0];
    const int tw = d_min(32, 1 << gy.lw), th = d_min(32, 1 << gy.lh);
    int32_t *co = coefs + b.coef_off[0];
    for (int k = c.tid; k < tw * th; k += 64) co[k] = L.qbest[k];
    /* the chroma waves code into their own candidate buffers meanwhile */
    if (bestEob > 0 && reuse && (ck != 1 || bestWave < SI_WU)) {
      tx_add(c, gy.px, gy.S, gy.sx, gy.sy, gy.lw, gy.lh, bestType, L.u.sm[bestWave].resid, SRS);
    } else if (bestEob > 0) {
      tx_dequant(c, L.qbest, W.dq, SQS, gy.txSz, fp.dc_q[0], fp.ac_q[0]);
      tx_inverse<true>(c, W.dq, SQS, W.resid, SRS, gy.txSz, bestType);
      tx_add(c, gy.px, gy.S, gy.sx, gy.sy, gy.lw, gy.lh, bestType, W.resid, SRS);
    }
    if (c.tid == 0) {
      b.tx_types_lo = (int32_t)(((uint32_t)b.tx_types_lo & ~15u) | (uint32_t)bestType);
      b.tx_type = bestType;
      b.eob[0] = bestEob > 0;
      si_mark(fp, b, 0);
    }
  }
  if (lk == 2 && tid == 0) si_mark(fp, L.slot[ls], 0);
  if (ck == 1 && cplane) {
    Mav1BlockInfo &b = L.slot[cs];
    if (cplane == 2 && b.uv_mode == UV_CFL_PRED) {
      const int a = si_cfl_apply(b, 2, L.si[2], b.cfl_alpha_u == 0);
      if ((tid & 63) == 0) b.cfl_alpha_v = a;
    }
    si_code_chroma(fp, b, cplane, coefs, mats, L.u.sm[wave]);
    /* then the chroma flags of the inter blocks that follow it */
    if ((tid & 63) == 0)
      for (int k = 0; k < L.nmq; k++) si_mark_g(fp, L.mq[k][0], L.mq[k][1], L.mq[k][2], cplane);
  }
  __syncthreads();
}

/* The workgroup's constant tables in LDS: the short forward matrices, the
 * small transforms' scans and the quantiser reciprocals. */
DEVI void stage_smats(const Mav1FrameParams &fp, const int32_t *__restrict__ mats) {
  for (int q = threadIdx.x; q < TX_SMATS; q += RT) {
    const int kind = q / 336, r = q - kind * 336;
    const int n = r < 16 ? 2 : r < 80 ? 3 : 4, idx = r - (n == 2 ? 0 : n == 3 ? 16 : 80);
    gL.smats[q] = mats[(kind * 5 + n - 2) * 4096 + idx];
  }
  const int sizes[9] = {TX_4X4, TX_8X8, TX_16X16, TX_4X8, TX_8X4, TX_8X16, TX_16X8, TX_4X16, TX_16X4};
  for (int q = threadIdx.x; q < 3 * SCAN_SMALL; q += RT) {
    const int kind = q / SCAN_SMALL, r = q - kind * SCAN_SMALL;
    int k = 8;
    while (small_scan_off(sizes[k]) > r) k--;
    const int off = small_scan_off(sizes[k]), v = scan_of(sizes[k], kind)[r - off];
    gL.scans[q] = v;
    gL.iscans[kind * SCAN_SMALL + off + v] = r - off;
  }
  if (threadIdx.x < 6) {
    const int p = threadIdx.x >> 1;
    gL.qrcp[p][threadIdx.x & 1] = tx_qrcp((threadIdx.x & 1) ? fp.ac_q[p] : fp.dc_q[p]);
  }
}

/* The inter blocks of the frame as a work list: work[0] = count, work[1] =
 * the next entry to take, work[2 + i] = block index. */
__global__ void k_recon_list_reset(int *work) {
  if (threadIdx.x < 2) work[threadIdx.x] = 0;
}
DEVI void recon_list_body(const Mav1FrameParams &fp, const Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo *__restrict__ sbs, int *work, const int blk_x, const int blk_y, const int blk_z) {
  const int sb = blk_x, k = threadIdx.x;
  if (k >= sbs[sb].num_blocks) return;
  const int bi = sbs[sb].first_block + k;
  if (blocks[bi].is_inter) work[2 + atomicAdd(&work[0], 1)] = bi;
}
__global__ __launch_bounds__(64) void k_recon_list(const Mav1FrameParams fp, const Mav1BlockInfo *__restrict__ blocks,
                                                   const Mav1SbInfo *__restrict__ sbs, int *work) {
  recon_list_body(fp, blocks, sbs, work, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* Inter phase: persistent workgroups take the inter blocks off the list
 * one at a time until it is empty. */
__global__ __launch_bounds__(RT) void k_recon_inter(const Mav1FrameParams fp, Mav1BlockInfo *__restrict__ blocks,
                                                    const uint8_t *__restrict__ src, uint8_t *rec,
                                                    const uint8_t *__restrict__ ref, const uint8_t *__restrict__ gold,
                                                    int32_t *__restrict__ coefs, const int32_t *__restrict__ mats,
                                                    int *work) {
  RLds &L = gL;
  const int tid = threadIdx.x;
  stage_tabs(tid, RT);
  stage_smats(fp, mats);
  const int count = work[0];
  for (;;) {
    if (tid == 0) L.ival[2] = atomicAdd(&work[1], 1);
    __syncthreads();
    const int idx = L.ival[2];
    __syncthreads();
    if (idx >= count) break;
    Mav1BlockInfo *g = blocks + work[2 + idx];
    if (tid < 27) ((int32_t *)&L.blk)[tid] = ((const int32_t *)g)[tid];
    __syncthreads();
    recon_block(fp, src, rec, false, ref, gold, coefs, mats, false, false);
    if (tid < 27) ((int32_t *)g)[tid] = ((const int32_t *)&L.blk)[tid];
    __syncthreads();
  }
}

/* ---- intra phase: one dataflow launch ----
 * k_intra_sched lists the superblocks holding intra blocks in order of
 * T = column + 2 row, then row. Every superblock whose reconstruction another
 * one reads (left, above-left, above, above-right) has a smaller T, so the
 * list is a topological order of the dependencies. k_recon_intra's
 * workgroups take the list's superblocks one at a time (an atomic counter);
 * before coding one they wait for its neighbours in the same tile column
 * that hold intra blocks, and they publish each finished superblock with an
 * agent-scope release. A workgroup only waits for superblocks taken earlier
 * by running workgroups, so every wait ends; it is still bounded, and a
 * wait that runs out sets iwork[IW_ERR] instead of hanging.
 * iwork: [0] count, [1] next, [2] error, [4 ..] the list, then per
 * superblock "holds intra blocks", then per superblock "done". */
#define IW_COUNT 0
#define IW_NEXT 1
#define IW_ERR 2
#define IW_HDR 4
#define SCHED_T 512
#define SCHED_KEYS 20480 /* (sb_cols + 2 sb_rows) sb_rows, checked on the host */

DEVI void intra_sched_body(const Mav1FrameParams &fp, const Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo *__restrict__ sbs, int *__restrict__ iwork, const int blk_x, const int blk_y, const int blk_z) {
  __shared__ uint16_t dense[SCHED_KEYS];
  __shared__ int cnt[SCHED_T + 1];
  const int n = fp.sb_cols * fp.sb_rows, tid = threadIdx.x;
  const int K = (fp.sb_cols + 2 * fp.sb_rows) * fp.sb_rows;
  int *list = iwork + IW_HDR, *has = list + n, *done = has + n;
  for (int k = tid; k < K; k += SCHED_T) dense[k] = 0;
  __syncthreads();
  for (int sb = tid; sb < n; sb += SCHED_T) {
    int h = 0;
    for (int i = 0; i < sbs[sb].num_blocks; i++) h |= !blocks[sbs[sb].first_block + i].is_inter;
    has[sb] = h;
    done[sb] = 0;
    const int r = sb / fp.sb_cols, c = sb - r * fp.sb_cols;
    if (h) dense[(c + 2 * r) * fp.sb_rows + r] = (uint16_t)(sb + 1);
  }
  __syncthreads();
  /* compaction in key order: per thread a contiguous range of keys */
  const int C = (K + SCHED_T - 1) / SCHED_T, k0 = tid * C, k1 = d_min(K, k0 + C);
  int m = 0;
  for (int k = k0; k < k1; k++) m += dense[k] != 0;
  cnt[tid + 1] = m;
  __syncthreads();
  if (tid == 0) {
    cnt[0] = 0;
    for (int i = 1; i <= SCHED_T; i++) cnt[i] += cnt[i - 1];
    iwork[IW_COUNT] = cnt[SCHED_T];
    iwork[IW_NEXT] = 0;
    iwork[IW_ERR] = 0;
  }
  __syncthreads();
  int o = cnt[tid];
  for (int k = k0; k < k1; k++)
    if (dense[k]) list[o++] = dense[k] - 1;
}
__global__ __launch_bounds__(SCHED_T) void k_intra_sched(const Mav1FrameParams fp, const Mav1BlockInfo *__restrict__ blocks,
                                                        const Mav1SbInfo *__restrict__ sbs, int *__restrict__ iwork) {
  intra_sched_body(fp, blocks, sbs, iwork, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* Waits (thread 0) until *flag is set; false when maxSpin polls ran out. */
DEVI bool wait_flag(const int *flag, int maxSpin) {
  for (int it = 0; it < maxSpin; it++) {
    if (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) return true;
    __builtin_amdgcn_s_sleep(4);
  }
  return false;
}

/* ---- the level schedule of a superblock ----
 * When every intra block of a superblock takes the small-block path (8x8 ..
 * 16x16, one transform block per plane, no tx_depth search) and no block is
 * smaller than 8x8, its intra blocks are coded by dependency level instead
 * of strictly one after another. A block's level is one more than the
 * highest level of the earlier (decode order) intra blocks of the
 * superblock that its prediction may read: those covering the column left
 * of it down to twice its height, the row above it right to twice its
 * width, or the corner; blocks of one level are therefore independent. The
 * BlockDecoded flags its prediction reads (above-right, below-left, per
 * plane) are those of decode order: the superblock's initial flags, or set
 * where an earlier block of the superblock covers the cell. Covering whole
 * blocks equals marking their transform blocks inside the frame, and a cell
 * outside the frame never changes a prediction (the edge fetch clamps to
 * the frame). When every intra block has a single transform type the
 * blocks run as a dataflow over these dependencies (lvl_dataflow below).
 * Otherwise each step takes the luma of up to four blocks of one level and
 * U and V of the previous step's blocks (CfL reads the luma finished one
 * step earlier), one task per wave: the predictions, behind a barrier the
 * luma transform types in rounds of one candidate per wave, then the chosen
 * types' reconstruction beside the chroma coding (lvl_step). Results equal
 * the decode-order pipeline's. The level pass also records per block the
 * mask of the earlier intra blocks it reads (dmask). */
/* luma blocks coded side by side in one step */
#define LVL_MAXW 4
DEVI void lvl_schedule(const Mav1FrameParams &fp, int n) {
  RLds &L = gL;
  const int tid = threadIdx.x;
  /* the BlockDecoded bits each intra block's prediction reads */
  for (int q = tid; q < n * 6; q += RT) {
    const int i = q / 6, p = (q >> 1) % 3, f = q & 1, sub = p > 0;
    const Mav1BlockInfo &b = L.sbb[i];
    if (b.is_inter) continue;
    const int txSz = p ? uv_tx_size(b.bsize) : Max_Tx_Size_Rect[b.bsize];
    const int stepX = Tx_Width[txSz] >> 2, stepY = Tx_Height[txSz] >> 2;
    const int by = (b.mi_row & 15) >> sub, bx = (b.mi_col & 15) >> sub;
    const int y = f ? by + stepY : by - 1, x = f ? bx - 1 : bx + stepX;
    int v = L.bd[BDI(p, y, x)];
    for (int j = 0; j < i && !v && y >= 0 && x >= 0; j++) {
      const Mav1BlockInfo &o = L.sbb[j];
      const int psz = SUBSAMPLED_SIZE(o.bsize, sub, sub);
      const int r0 = (o.mi_row & 15) >> sub, c0 = (o.mi_col & 15) >> sub;
      v = y >= r0 && y < r0 + Num_4x4_Blocks_High[psz] && x >= c0 && x < c0 + Num_4x4_Blocks_Wide[psz];
    }
    if (v) atomicOr(&L.pav[i], 1 << (2 * p + f));
  }
  /* levels, in decode order on wave 0 (lane j: block j as a dependency) */
  if (tid < 64) {
    const int j = tid;
    int myLv = 0, rj = 0, cj = 0, hj = 0, wj = 0, intraJ = 0;
    if (j < n) {
      const Mav1BlockInfo &o = L.sbb[j];
      rj = o.mi_row & 15;
      cj = o.mi_col & 15;
      hj = Num_4x4_Blocks_High[o.bsize];
      wj = Num_4x4_Blocks_Wide[o.bsize];
      intraJ = !o.is_inter;
    }
    for (int i = 0; i < n; i++) {
      const Mav1BlockInfo &b = L.sbb[i];
      if (d_uni(b.is_inter)) continue;
      const int r = d_uni(b.mi_row) & 15, c = d_uni(b.mi_col) & 15, bs = d_uni(b.bsize);
      const int h = d_uni(Num_4x4_Blocks_High[bs]), w = d_uni(Num_4x4_Blocks_Wide[bs]);
      bool dep = false;
      if (j < i && intraJ) {
        const bool left = c - 1 >= cj && c - 1 < cj + wj && r - 1 < rj + hj && r + 2 * h > rj;
        const bool above = r - 1 >= rj && r - 1 < rj + hj && c - 1 < cj + wj && c + 2 * w > cj;
        dep = left || above;
      }
      const unsigned long long dm = __ballot(dep);
      const int m = wave_max_dpp(dep ? myLv : 0);
      if (j == i) {
        myLv = m + 1;
        L.dmask[i] = dm;
      }
    }
    if (j < n) L.plv[j] = myLv;
  }
  __syncthreads();
  /* the schedule: intra blocks by level, then decode order */
  if (tid < n && L.plv[tid] > 0) {
    const int lv = L.plv[tid];
    int pos = 0;
    for (int k = 0; k < n; k++) {
      const int l = L.plv[k];
      pos += l > 0 && (l < lv || (l == lv && k < tid));
    }
    L.pord[pos] = tid;
  }
  if (tid == 64) {
    int cnt = 0;
    for (int k = 0; k < n; k++) cnt += L.plv[k] > 0;
    L.pn = cnt;
  }
  __syncthreads();
}

/* Number of luma transform-type candidates of a small intra block, and the
 * one at index ci (code_plane's candidate list). */
DEVI int lvl_ncands(const Mav1FrameParams &fp, int txSz) {
  const int set = get_tx_set(txSz, 0, fp.reduced_tx_set);
  if (!fp.tx_search || set == TX_SET_DCTONLY) return 1;
  int n = 0;
  for (int t = 0; t < 16; t++) n += in_set(set, t, 0) ? 1 : 0;
  return n;
}
DEVI int lvl_cand(const Mav1FrameParams &fp, int txSz, int ci) {
  const int set = get_tx_set(txSz, 0, fp.reduced_tx_set);
  if (!fp.tx_search || set == TX_SET_DCTONLY) return DCT_DCT;
  for (int t = 0; t < 16; t++)
    if (in_set(set, t, 0) && ci-- == 0) return t;
  return DCT_DCT;
}

/* The chosen luma type of small block b on one wave: levels out,
 * reconstruction into the tile, the block's transform fields. */
DEVI void lvl_luma_out(const Mav1FrameParams &fp, Mav1BlockInfo &b, const int32_t *q, int type, int eob,
                       int32_t *__restrict__ coefs, SmallTx &W) {
  const TxCtx c = wave_ctx();
  const SiGeo g = si_geo(b, 0);
  const int tw = d_min(32, 1 << g.lw), th = d_min(32, 1 << g.lh);
  type = d_uni(type);
  eob = d_uni(eob);
  int32_t *co = coefs + d_uni(b.coef_off[0]);
  for (int k = c.tid; k < tw * th; k += 64) co[k] = q[k];
  if (eob > 0) tx_recon_small(c, q, W.resid, SRS, g.txSz, type, fp.dc_q[0], fp.ac_q[0], g.px, g.S, g.sx, g.sy);
  if (c.tid == 0) {
    b.tx_types_lo = (int32_t)(((uint32_t)b.tx_types_lo & ~15u) | (uint32_t)type);
    b.tx_type = type;
    b.eob[0] = eob > 0;
  }
}

/* A chroma plane of the previous step's block k (task index 2k + plane - 1)
 * on one wave: V's CfL alpha (U's is known), then its coding. */
DEVI void lvl_chroma(const Mav1FrameParams &fp, int prev, int k, int32_t *__restrict__ coefs,
                     const int32_t *__restrict__ mats) {
  RLds &L = gL;
  k = d_uni(k);
  const int plane = 1 + (k & 1), wave = d_uni(threadIdx.x >> 6);
  Mav1BlockInfo &b = L.sbb[d_uni(L.pord[prev + (k >> 1)])];
  if (plane == 2 && d_uni(b.uv_mode) == UV_CFL_PRED) {
    /* U's alpha (another wave's task of the same step) decides whether V
     * may take alpha 0 */
    const int i = d_uni(L.pord[prev + (k >> 1)]);
    int spin = 0;
    while (!d_uni(__hip_atomic_load(&L.ufl[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) && spin < (1 << 24)) {
      __builtin_amdgcn_s_sleep(1);
      spin++;
    }
    if (spin == (1 << 24) && (threadIdx.x & 63) == 0) L.lerr = 1;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    const int a = si_cfl_apply(b, 2, L.psw[4 + k], d_uni(b.cfl_alpha_u) == 0);
    if ((threadIdx.x & 63) == 0) b.cfl_alpha_v = a;
  }
  si_code_chroma(fp, b, plane, coefs, mats, L.u.sm[wave], false);
}

/* The prediction of task t of a step (luma of block pord[cur + t] for
 * t < ncur, else U / V of block pord[prev + (t - ncur) / 2]) on one wave;
 * U's CfL alpha is chosen at once and published in ufl. */
DEVI void lvl_predict(const Mav1FrameParams &fp, int t, int cur, int ncur, int prev) {
  RLds &L = gL;
  const int lane = threadIdx.x & 63;
  if (t < ncur) {
    const int i = d_uni(L.pord[cur + t]);
    Mav1BlockInfo &b = L.sbb[i];
    if (lane == 0) {
      b.tx_depth = 0;
      b.interp_filter &= 0x30;
    }
    const int av = d_uni(L.pav[i]);
    si_predict(fp, b, 0, L.psw[t], av & 1, (av >> 1) & 1);
  } else {
    const int k = t - ncur, plane = 1 + (k & 1), i = d_uni(L.pord[prev + (k >> 1)]);
    Mav1BlockInfo &b = L.sbb[i];
    const int av = d_uni(L.pav[i]);
    si_predict(fp, b, plane, L.psw[4 + k], (av >> (2 * plane)) & 1, (av >> (2 * plane + 1)) & 1);
    if (plane == 1 && d_uni(b.uv_mode) == UV_CFL_PRED) {
      const int a = si_cfl_apply(b, 1, L.psw[4 + k], 0);
      if (lane == 0) {
        b.cfl_alpha_u = a;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __hip_atomic_store(&L.ufl[i], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
}

/* One step of the level schedule: luma of blocks pord[cur .. cur + ncur)
 * (one level), chroma of blocks pord[prev .. prev + nprev). */
DEVI void lvl_step(const Mav1FrameParams &fp, int32_t *__restrict__ coefs, const int32_t *__restrict__ mats, int cur,
                   int ncur, int prev, int nprev) {
  RLds &L = gL;
  const int tid = threadIdx.x, wave = d_uni(tid >> 6), lane = tid & 63;
  cur = d_uni(cur);
  ncur = d_uni(ncur);
  prev = d_uni(prev);
  nprev = d_uni(nprev);
  const int ntask = ncur + 2 * nprev;
  /* luma candidates: offsets per block */
  int off[5], total = 0;
  bool single = true;
  for (int k = 0; k < 4; k++) {
    off[k] = total;
    if (k < ncur) {
      const int nc = d_uni(lvl_ncands(fp, d_uni(Max_Tx_Size_Rect[d_uni(L.sbb[d_uni(L.pord[cur + k])].bsize)])));
      total += nc;
      single = single && nc == 1;
    }
  }
  off[4] = total;
  /* The tasks: the luma of this step's blocks (t < ncur) and U / V of the
   * previous step's blocks, one per wave. Their predictions touch disjoint
   * data: a luma prediction reads luma of lower levels and writes its own
   * block; a chroma prediction reads chroma of lower levels and its own
   * block's reconstructed luma (CfL), and writes its own chroma block; each
   * task has its own SiWave. The coding of a task reads only its own
   * prediction, except V's CfL, which waits for U's alpha (ufl). */
  if (single) {
    /* one type per block: each wave predicts and codes its tasks end to
     * end, with no barrier between prediction and coding */
    const TxCtx c = wave_ctx();
    for (int t = wave; t < ntask; t += RW) {
      lvl_predict(fp, t, cur, ncur, prev);
      if (t < ncur) {
        Mav1BlockInfo &b = L.sbb[d_uni(L.pord[cur + t])];
        const SiGeo g = si_geo(b, 0);
        SmallTx &W = L.u.sm[wave];
        long long cost;
        const int eob = tx_eval<true>(c, fp, g.spx, g.px, g.S, g.sx, g.sy, g.txSz, 0, DCT_DCT, true, W.resid, SRS,
                                      W.tmp, W.coef, W.q, W.dq, SQS, mats, &cost);
        lvl_luma_out(fp, b, W.q, DCT_DCT, eob, coefs, W);
      } else {
        lvl_chroma(fp, prev, t - ncur, coefs, mats);
      }
    }
    lds_barrier();
    return;
  }
  /* several types: the predictions first, behind a barrier the rounds */
  for (int t = wave; t < ntask; t += RW) lvl_predict(fp, t, cur, ncur, prev);
  lds_barrier();
  /* several types: rounds of one candidate per wave, the best kept per
   * block (strictly smaller cost wins, candidates in list order) */
  if (tid < ncur) L.pcost[tid] = -1;
  for (int r0 = 0; r0 < total; r0 += RW) {
    const int t = r0 + wave;
    if (t < total) {
      int k = 0;
      while (t >= off[k + 1]) k++;
      Mav1BlockInfo &b = L.sbb[d_uni(L.pord[cur + k])];
      const SiGeo g = si_geo(b, 0);
      SmallTx &W = L.u.sm[wave];
      long long cost;
      const int eob = tx_eval<true>(wave_ctx(), fp, g.spx, g.px, g.S, g.sx, g.sy, g.txSz, 0,
                                    lvl_cand(fp, g.txSz, t - off[k]), false, W.resid, SRS, W.tmp, W.coef, W.q, W.dq,
                                    SQS, mats, &cost);
      if (lane == 0) {
        L.rcost[wave] = cost;
        L.reob[wave] = eob;
      }
    }
    lds_barrier();
    if (tid < ncur) {
      const int k = tid;
      int win = -1;
      for (int u = d_max(off[k], r0); u < d_min(off[k + 1], r0 + RW); u++)
        if (L.pcost[k] < 0 || L.rcost[u - r0] < L.pcost[k]) {
          L.pcost[k] = L.rcost[u - r0];
          L.peob[k] = L.reob[u - r0];
          const int ty = lvl_cand(fp, Max_Tx_Size_Rect[L.sbb[L.pord[cur + k]].bsize], u - off[k]);
          L.ptype[k] = L.reob[u - r0] > 0 ? ty : DCT_DCT;
          win = u - r0;
        }
      L.pwin[k] = win;
    }
    lds_barrier();
    for (int k = 0; k < ncur; k++) {
      const int win = d_uni(L.pwin[k]);
      if (win < 0) continue;
      const Mav1BlockInfo &b = L.sbb[d_uni(L.pord[cur + k])];
      const int txSz = d_uni(Max_Tx_Size_Rect[d_uni(b.bsize)]);
      const int n = d_min(32, Tx_Width[txSz]) * d_min(32, Tx_Height[txSz]);
      for (int e = tid; e < n; e += RT) L.pq[k][e] = L.u.sm[win].q[e];
    }
    lds_barrier();
  }
  /* P3: the chosen luma types' reconstruction beside the chroma coding */
  for (int t = wave; t < ntask; t += RW) {
    if (t < ncur)
      lvl_luma_out(fp, L.sbb[d_uni(L.pord[cur + t])], L.pq[t], d_uni(L.ptype[t]), d_uni(L.peob[t]), coefs,
                   L.u.sm[wave]);
    else
      lvl_chroma(fp, prev, t - ncur, coefs, mats);
  }
  lds_barrier();
}

/* ---- the level schedule as a dataflow ----
 * When every intra block of the superblock has a single transform type,
 * the steps give way to per-block dependencies: luma of block i waits for
 * the luma of the blocks in dmask[i] (those its prediction reads); U (V)
 * of block i waits for its own luma and the U (V) of the blocks in
 * dmask[i] (the chroma prediction reads the same neighbours at half
 * resolution); V's CfL waits for U's alpha. Waves 0 .. LVL_LW - 1 take
 * the luma tasks in schedule order (task u on wave u % LVL_LW), the other
 * waves the chroma tasks (U then V per block, task ci on wave
 * LVL_LW + ci % (RW - LVL_LW)): chroma has two tasks per block and they
 * take as long as a luma task, so chroma gets the larger share (measured on
 * the 1080p key frame: 2 luma waves 13.4 ms, 3: 13.75 ms, 4: 14.2 ms).
 * Every task waits only
 * for tasks earlier in its kind's order, and luma waits for no chroma, so
 * the earliest unfinished task can always run: every wait ends. Waits are
 * still bounded (a wait that runs out sets lerr). Results equal the steps'. */
#define LVL_SPIN (1 << 20)
#define LVL_LW 2 /* waves for the luma tasks; the others take U and V */
DEVI unsigned long long lds_load64(const unsigned long long *p) {
  const unsigned long long v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return ((unsigned long long)(unsigned)d_uni((int)(v >> 32)) << 32) | (unsigned)d_uni((int)(unsigned)v);
}
/* Waits until (*m & need) == need and bit `own` of *own (if own) is set. */
DEVI void lvl_wait(const unsigned long long *m, unsigned long long need, const unsigned long long *own, int i) {
  for (int it = 0; it < LVL_SPIN; it++) {
    if ((lds_load64(m) & need) == need && (!own || ((lds_load64(own) >> i) & 1))) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      return;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  if ((threadIdx.x & 63) == 0) gL.lerr = 1;
}
DEVI void lvl_publish(unsigned long long *m, int i) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_or(m, 1ull << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
DEVI void lvl_dataflow(const Mav1FrameParams &fp, int32_t *__restrict__ coefs, const int32_t *__restrict__ mats) {
  RLds &L = gL;
  const int wave = d_uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int pn = d_uni(L.pn);
  const TxCtx c = wave_ctx();
  SmallTx &W = L.u.sm[wave];
  SiWave &P = L.psw[wave];
  if (wave < LVL_LW) {
    for (int u = wave; u < pn; u += LVL_LW) {
      const int i = d_uni(L.pord[u]);
      Mav1BlockInfo &b = L.sbb[i];
      lvl_wait(&L.ldm, L.dmask[i], nullptr, 0);
      if (lane == 0) {
        b.tx_depth = 0;
        b.interp_filter &= 0x30;
      }
      const int av = d_uni(L.pav[i]);
      si_predict(fp, b, 0, P, av & 1, (av >> 1) & 1);
      const SiGeo g = si_geo(b, 0);
      long long cost;
      const int eob = tx_eval<true>(c, fp, g.spx, g.px, g.S, g.sx, g.sy, g.txSz, 0, DCT_DCT, true, W.resid, SRS, W.tmp,
                                    W.coef, W.q, W.dq, SQS, mats, &cost);
      /* the reconstruction first: the blocks waiting for this one read only
       * its pixels; then the levels out and the block's transform fields
       * (lvl_luma_out's work, in that order) */
      if (eob > 0) tx_recon_small(c, W.q, W.resid, SRS, g.txSz, DCT_DCT, fp.dc_q[0], fp.ac_q[0], g.px, g.S, g.sx, g.sy);
      lvl_publish(&L.ldm, i);
      const int tw = d_min(32, 1 << g.lw), th = d_min(32, 1 << g.lh);
      int32_t *co = coefs + d_uni(b.coef_off[0]);
      for (int k = c.tid; k < tw * th; k += 64) co[k] = W.q[k];
      if (lane == 0) {
        b.tx_types_lo = (int32_t)(((uint32_t)b.tx_types_lo & ~15u) | (uint32_t)DCT_DCT);
        b.tx_type = DCT_DCT;
        b.eob[0] = eob > 0;
      }
    }
  } else {
    for (int ci = wave - LVL_LW; ci < 2 * pn; ci += RW - LVL_LW) {
      const int u = ci >> 1, plane = 1 + (ci & 1), i = d_uni(L.pord[u]);
      Mav1BlockInfo &b = L.sbb[i];
      unsigned long long *dn = plane == 1 ? &L.udm : &L.vdm;
      lvl_wait(dn, L.dmask[i], &L.ldm, i);
      const int av = d_uni(L.pav[i]);
      si_predict(fp, b, plane, P, (av >> (2 * plane)) & 1, (av >> (2 * plane + 1)) & 1);
      if (d_uni(b.uv_mode) == UV_CFL_PRED) {
        if (plane == 1) {
          const int a = si_cfl_apply(b, 1, P, 0);
          if (lane == 0) {
            b.cfl_alpha_u = a;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __hip_atomic_store(&L.ufl[i], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else {
          int spin = 0;
          while (!d_uni(__hip_atomic_load(&L.ufl[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) &&
                 spin < LVL_SPIN) {
            __builtin_amdgcn_s_sleep(1);
            spin++;
          }
          if (spin == LVL_SPIN && lane == 0) L.lerr = 1;
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
          const int a = si_cfl_apply(b, 2, P, d_uni(b.cfl_alpha_u) == 0);
          if (lane == 0) b.cfl_alpha_v = a;
        }
      }
      si_code_chroma(fp, b, plane, coefs, mats, W, false);
      lvl_publish(dn, i);
    }
  }
  __syncthreads();
}

/* The intra blocks of superblock (r, c) (recon_sb, k_recon.h:396-440) in
 * two parts. intra_sb_prep: what does not read the neighbours'
 * reconstruction (BlockDecoded flags, source tile, block records, the level
 * schedule), so it can run while the neighbours are still being coded. */
DEVI void intra_sb_prep(const Mav1FrameParams &fp, const Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo &S,
                        const uint8_t *__restrict__ src, int r, int c) {
  RLds &L = gL;
  const int tid = threadIdx.x;
  const int c0 = 0, cl = c;
  /* clear_block_decoded_flags (k_recon.h:402-414) */
  const int mr = r * 16, mc = (c0 + cl) * 16;
  const int tileEnd = d_min(fp.mi_cols, tile_start(fp, mc) + fp.tile_w_sb * 16);
  for (int q = tid; q < 3 * BDD * BDD; q += RT) {
    const int p = q / (BDD * BDD), e = q - p * BDD * BDD, y = e / BDD - 1, x = e % BDD - 1, sub = p > 0;
    const int w4 = (tileEnd - mc) >> sub, h4 = (fp.mi_rows - mr) >> sub;
    uint8_t v = (y < 0 && x < w4) ? 1 : (x < 0 && y < h4) ? 1 : 0;
    if (y == (16 >> sub) && x == -1) v = 0;
    if (y <= (16 >> sub) && x <= (16 >> sub)) L.bd[q] = v;
  }
  /* the superblock's source tile, as dwords */
  for (int p = 0; p < 3; p++) {
    const int sub = p > 0, TS = sub ? TS_C : TS_Y, n = 64 >> sub, wpr = TS / 4;
    const int oy = ((r * 64) >> sub) - 1, ox = (((c0 + cl) * 64) >> sub) - 4, stride = fp.stride[p];
    uint32_t *ts = (uint32_t *)(L.tile[1] + (p == 0 ? 0 : TILE_Y + (p - 1) * TILE_C));
    for (int q = tid; q < (n + 1) * wpr; q += RT) {
      const int ty = q / wpr, tx = (q - ty * wpr) * 4, Y = oy + ty, X = ox + tx;
      if (ty == 0 || tx < 4 || tx >= 4 + n || Y < 0 || X < 0 || X >= stride) continue;
      ts[q] = *(const uint32_t *)(src + fp.offset[p] + Y * stride + X);
    }
  }
  /* the block records; the level schedule applies when every intra block
   * takes the small-block path and no block is smaller than 8x8 */
  const int nb = S.num_blocks;
  for (int q = tid; q < nb * 27; q += RT) ((int32_t *)L.sbb)[q] = ((const int32_t *)(blocks + S.first_block))[q];
  for (int q = tid; q < 64; q += RT) {
    L.pav[q] = 0;
    L.ufl[q] = 0;
  }
  if (tid == 0) {
    L.pok = 1;
    L.lerr = 0;
    L.psingle = 1;
    L.ldm = L.udm = L.vdm = 0;
  }
  __syncthreads();
  if (tid < nb) {
    const Mav1BlockInfo &b = L.sbb[tid];
    const int bs = b.bsize;
    if (bs < BLOCK_8X8 || bs == BLOCK_4X16 || bs == BLOCK_16X4 || (!b.is_inter && !si_eligible(fp, b))) L.pok = 0;
    if (!b.is_inter && lvl_ncands(fp, Max_Tx_Size_Rect[bs]) != 1) L.psingle = 0;
  }
  __syncthreads();
  if (L.pok) lvl_schedule(fp, nb);
}

/* intra_sb: after intra_sb_prep and once the neighbouring superblocks are
 * done, the reconstruction tile (with the row above and the columns left),
 * the blocks, the tiles back. Returns non-zero when a wait inside ran out
 * (the results are then not to be used). */
DEVI int intra_sb(const Mav1FrameParams &fp, Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo &S,
                  const uint8_t *__restrict__ src, uint8_t *rec, const uint8_t *__restrict__ ref,
                  const uint8_t *__restrict__ gold, int32_t *__restrict__ coefs, const int32_t *__restrict__ mats,
                  int r, int c) {
  RLds &L = gL;
  const int tid = threadIdx.x;
  const int c0 = 0, cl = c;
  for (int p = 0; p < 3; p++) {
    const int sub = p > 0, TS = sub ? TS_C : TS_Y, n = 64 >> sub, wpr = TS / 4;
    const int oy = ((r * 64) >> sub) - 1, ox = (((c0 + cl) * 64) >> sub) - 4, stride = fp.stride[p];
    uint32_t *tr = (uint32_t *)(L.tile[0] + (p == 0 ? 0 : TILE_Y + (p - 1) * TILE_C));
    for (int q = tid; q < (n + 1) * wpr; q += RT) {
      const int ty = q / wpr, tx = (q - ty * wpr) * 4, Y = oy + ty, X = ox + tx;
      if ((ty > 0 && tx >= 4 + n) || Y < 0 || X < 0 || X >= stride) continue;
      tr[q] = *(const uint32_t *)(rec + fp.offset[p] + Y * stride + X);
    }
  }
  const int nb = S.num_blocks;
  __syncthreads();
  if (L.pok && L.psingle) {
    lvl_dataflow(fp, coefs, mats);
    for (int q = tid; q < nb * 27; q += RT)
      if (!L.sbb[q / 27].is_inter) ((int32_t *)(blocks + S.first_block))[q] = ((const int32_t *)L.sbb)[q];
    __syncthreads();
  } else if (L.pok) {
    int prev = 0, nprev = 0, cur = 0;
    const int pn = d_uni(L.pn);
    for (;;) {
      int ncur = 0;
      if (cur < pn) {
        const int lv = d_uni(L.plv[d_uni(L.pord[cur])]);
        while (ncur < LVL_MAXW && cur + ncur < pn && d_uni(L.plv[d_uni(L.pord[cur + ncur])]) == lv) ncur++;
      }
      if (!ncur && !nprev) break;
      lvl_step(fp, coefs, mats, cur, ncur, prev, nprev);
      prev = cur;
      nprev = ncur;
      cur += ncur;
    }
    for (int q = tid; q < nb * 27; q += RT)
      if (!L.sbb[q / 27].is_inter) ((int32_t *)(blocks + S.first_block))[q] = ((const int32_t *)L.sbb)[q];
    __syncthreads();
  } else {
  /* blocks in decode order: small intra blocks through the two-stage
   * pipeline (chroma one step behind luma), inter blocks' flags through
   * the same streams, any other intra block by the whole workgroup once the
   * pipeline has drained */
  /* inter blocks take no step: their luma flags are set at once, their
   * chroma flags queued behind the pending chroma (or set at once when
   * none is pending) */
  int pend = -1, pendIdx = 0;
  if (tid == 0) L.nmq = 0;
  for (int i = 0; i <= S.num_blocks; i++) {
    const int sl = pend == 0 ? 1 : 0;
    if (i < S.num_blocks && tid < 27)
      ((int32_t *)&L.slot[sl])[tid] = ((const int32_t *)(blocks + S.first_block + i))[tid];
    __syncthreads();
    const int kind = i == S.num_blocks ? 0 : L.slot[sl].is_inter ? 2 : si_eligible(fp, L.slot[sl]) ? 1 : 3;
    if (kind == 2) {
      if (tid == 0) {
        const Mav1BlockInfo &b = L.slot[sl];
        si_mark(fp, b, 0);
        if (pend >= 0) {
          L.mq[L.nmq][0] = b.bsize;
          L.mq[L.nmq][1] = b.mi_row;
          L.mq[L.nmq][2] = b.mi_col;
          L.nmq++;
        } else {
          si_mark(fp, b, 1);
          si_mark(fp, b, 2);
        }
      }
      __syncthreads();
      continue;
    }
    if (kind == 1) {
      si_step(fp, sl, 1, pend < 0 ? 0 : pend, pend < 0 ? 0 : 1, coefs, mats);
    } else if (pend >= 0) {
      si_step(fp, 0, 0, pend, 1, coefs, mats);
    }
    if (pend >= 0 && tid < 27)
      ((int32_t *)(blocks + S.first_block + pendIdx))[tid] = ((const int32_t *)&L.slot[pend])[tid];
    if (tid == 0) L.nmq = 0;
    pend = -1;
    if (kind == 1) {
      pend = sl;
      pendIdx = i;
    } else if (kind == 3) {
      if (tid < 27) ((int32_t *)&L.blk)[tid] = ((const int32_t *)&L.slot[sl])[tid];
      __syncthreads();
      recon_block(fp, src, rec, true, ref, gold, coefs, mats, true, false);
      if (tid < 27) ((int32_t *)(blocks + S.first_block + i))[tid] = ((const int32_t *)&L.blk)[tid];
    }
    __syncthreads();
  }
  }
  /* the superblock back into the frame */
  for (int p = 0; p < 3; p++) {
    const int sub = p > 0, TS = sub ? TS_C : TS_Y, n = 64 >> sub, wpr = n / 4;
    const int y0 = (r * 64) >> sub, x0 = ((c0 + cl) * 64) >> sub, stride = fp.stride[p];
    const uint32_t *tr = (const uint32_t *)(L.tile[0] + (p == 0 ? 0 : TILE_Y + (p - 1) * TILE_C));
    for (int q = tid; q < n * wpr; q += RT) {
      const int ty = q / wpr, tx = (q - ty * wpr) * 4;
      *(uint32_t *)(rec + fp.offset[p] + (y0 + ty) * stride + x0 + tx) = tr[(ty + 1) * (TS / 4) + 1 + tx / 4];
    }
  }
  return L.lerr;
}

__global__ __launch_bounds__(RT) void k_recon_intra(const Mav1FrameParams fp, Mav1BlockInfo *__restrict__ blocks,
                                                    const Mav1SbInfo *__restrict__ sbs, const uint8_t *__restrict__ src,
                                                    uint8_t *rec, const uint8_t *__restrict__ ref,
                                                    const uint8_t *__restrict__ gold, int32_t *__restrict__ coefs,
                                                    const int32_t *__restrict__ mats, int *__restrict__ iwork,
                                                    int maxSpin) {
  RLds &L = gL;
  const int tid = threadIdx.x, n = fp.sb_cols * fp.sb_rows;
  const int *list = iwork + IW_HDR, *has = list + n;
  int *done = iwork + IW_HDR + 2 * n;
  /* the first superblock before anything else: workgroups without one leave
   * at once */
  if (tid == 0) L.ival[2] = atomicAdd(&iwork[IW_NEXT], 1);
  __syncthreads();
  if (L.ival[2] >= iwork[IW_COUNT]) return;
  stage_tabs(tid, RT);
  stage_smats(fp, mats);
  for (;;) {
    const int idx = L.ival[2];
    __syncthreads();
    if (idx >= iwork[IW_COUNT]) break;
    const int sb = list[idx], r = sb / fp.sb_cols, c = sb - r * fp.sb_cols;
    /* everything that does not read the neighbours' reconstruction first,
     * beside the wait for them */
    intra_sb_prep(fp, blocks, sbs[sb], src, r, c);
    if (tid == 0) {
      /* the neighbours whose reconstruction this superblock reads */
      const int ts = (c / fp.tile_w_sb) * fp.tile_w_sb, te = d_min(fp.sb_cols, ts + fp.tile_w_sb);
      const int dep[4] = {c > ts ? sb - 1 : -1, r > 0 && c > ts ? sb - fp.sb_cols - 1 : -1, r > 0 ? sb - fp.sb_cols : -1,
                          r > 0 && c + 1 < te ? sb - fp.sb_cols + 1 : -1};
      for (int k = 0; k < 4; k++)
        if (dep[k] >= 0 && has[dep[k]] && !wait_flag(&done[dep[k]], maxSpin)) atomicOr(&iwork[IW_ERR], 1);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int err = intra_sb(fp, blocks, sbs[sb], src, rec, ref, gold, coefs, mats, r, c);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
      if (err) atomicOr(&iwork[IW_ERR], 1);
      __hip_atomic_store(&done[sb], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      L.ival[2] = atomicAdd(&iwork[IW_NEXT], 1);
    }
    __syncthreads();
  }
}

/* Edge filter types (src/encoder.c:319-346): smooth flags per mi, then per
 * block the above / left neighbours' flags into interp_filter bits 4 / 5. */
DEVI void edge_flags_body(const Mav1FrameParams &fp, const Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo *__restrict__ sbs, uint8_t *ys, uint8_t *uvs, const int blk_x, const int blk_y, const int blk_z) {
  const int sb = blk_x, lane = threadIdx.x;
  stage_tabs(lane, 64);
  __syncthreads();
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
__global__ __launch_bounds__(64) void k_edge_flags(const Mav1FrameParams fp, const Mav1BlockInfo *__restrict__ blocks,
                                                   const Mav1SbInfo *__restrict__ sbs, uint8_t *ys, uint8_t *uvs) {
  edge_flags_body(fp, blocks, sbs, ys, uvs, blockIdx.x, blockIdx.y, blockIdx.z);
}
DEVI void edge_types_body(const Mav1FrameParams &fp, Mav1BlockInfo *__restrict__ blocks, const Mav1SbInfo *__restrict__ sbs, const uint8_t *ys, const uint8_t *uvs, const int blk_x, const int blk_y, const int blk_z) {
  const int sb = blk_x, lane = threadIdx.x;
  if (lane >= sbs[sb].num_blocks) return;
  Mav1BlockInfo *b = &blocks[sbs[sb].first_block + lane];
  const int cols = fp.mi_cols, rows = fp.mi_rows, r = b->mi_row, c = b->mi_col;
  const int au = r > 0, al = c > tile_start(fp, c);
  const int fy = (au && ys[(r - 1) * cols + c]) || (al && ys[r * cols + c - 1]);
  const int fuv = (au && uvs[(r - 1) * cols + d_min(c + 1, cols - 1)]) || (al && uvs[d_min(r + 1, rows - 1) * cols + c - 1]);
  b->interp_filter = (b->interp_filter & 15) | fy << 4 | fuv << 5;
}
__global__ __launch_bounds__(64) void k_edge_types(const Mav1FrameParams fp, Mav1BlockInfo *__restrict__ blocks,
                                                   const Mav1SbInfo *__restrict__ sbs, const uint8_t *ys,
                                                   const uint8_t *uvs) {
  edge_types_body(fp, blocks, sbs, ys, uvs, blockIdx.x, blockIdx.y, blockIdx.z);
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
                              int32_t *coefs, const int32_t *mats, int *work, int workgroups) {
  const int nsb = fp.sb_cols * fp.sb_rows;
  hipLaunchKernelGGL(k_recon_list_reset, dim3(1), dim3(64), 0, st, work);
  hipLaunchKernelGGL(k_recon_list, dim3(nsb), dim3(64), 0, st, fp, (const Mav1BlockInfo *)blocks, sbs, work);
  hipLaunchKernelGGL(k_recon_inter, dim3(workgroups), dim3(RT), 0, st, fp, blocks, src, rec, ref, gold, coefs, mats,
                     work);
  return hipGetLastError();
}


/* list, has, done and (batched form) the pending-dependency counts */
size_t intra_work_ints(const Mav1FrameParams &fp) { return IW_HDR + 4 * (size_t)fp.sb_cols * fp.sb_rows; }
/* Workgroups for the intra phase of one frame: twice the superblocks the
 * dependency order lets run at once (per tile column, one superblock in
 * every second column of a diagonal c + 2 r), at least 16, at most the CU
 * count. More only hold CUs while they wait, which the other streams'
 * kernels could use. */
int intra_workgroups(const Mav1FrameParams &fp, int num_cu) {
  const int tiles = (fp.sb_cols + fp.tile_w_sb - 1) / fp.tile_w_sb;
  const int per = ((fp.tile_w_sb < 2 * fp.sb_rows ? fp.tile_w_sb : 2 * fp.sb_rows) + 1) / 2;
  const int wg = 2 * tiles * per < 16 ? 16 : 2 * tiles * per;
  return wg < num_cu ? wg : num_cu;
}

int intra_sched_fits(const Mav1FrameParams &fp) {
  return (fp.sb_cols + 2 * fp.sb_rows) * fp.sb_rows <= SCHED_KEYS && fp.sb_cols * fp.sb_rows < 65535;
}

hipError_t launch_intra_sched(hipStream_t st, const Mav1FrameParams &fp, const Mav1BlockInfo *blocks,
                              const Mav1SbInfo *sbs, int *iwork) {
  hipLaunchKernelGGL(k_intra_sched, dim3(1), dim3(SCHED_T), 0, st, fp, blocks, sbs, iwork);
  return hipGetLastError();
}

hipError_t launch_recon_intra(hipStream_t st, const Mav1FrameParams &fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                              const uint8_t *src, uint8_t *rec, const uint8_t *ref, const uint8_t *gold,
                              int32_t *coefs, const int32_t *mats, int *iwork, int workgroups) {
  hipLaunchKernelGGL(k_recon_intra, dim3(workgroups), dim3(RT), 0, st, fp, blocks, sbs, src, rec, ref, gold, coefs,
                     mats, iwork, 1 << 20);
  return hipGetLastError();
}

/* ---- batched launches (batch.h): the same kernels over several frames ---- */

__global__ __launch_bounds__(64) void k_edge_flags_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  edge_flags_body(F.fp, gptr(F.blocks), gptr(F.sbs), gptr(F.ys), gptr(F.uvs), b.x, b.y, b.z);
}
__global__ __launch_bounds__(64) void k_edge_types_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  edge_types_body(F.fp, gptr(F.blocks), gptr(F.sbs), gptr(F.ys), gptr(F.uvs), b.x, b.y, b.z);
}
__global__ void k_recon_list_reset_b(const BatchFrame *__restrict__ T) {
  if (threadIdx.x < 2) T[blockIdx.x].rwork[threadIdx.x] = 0;
}
__global__ __launch_bounds__(64) void k_recon_list_b(const BatchFrame *__restrict__ T, const BatchGrid g) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  recon_list_body(F.fp, gptr(F.blocks), gptr(F.sbs), gptr(F.rwork), b.x, b.y, b.z);
}

/* Inter phase over several frames: every workgroup takes the inter blocks
 * of frame 0 off its list until the list is empty, then frame 1's, and so
 * on; the frame-dependent LDS tables are staged when a workgroup takes its
 * first block of a frame. */
__global__ __launch_bounds__(RT) void k_recon_inter_b(const BatchFrame *__restrict__ T, int nfr) {
  RLds &L = gL;
  const int tid = threadIdx.x;
  stage_tabs(tid, RT);
  for (int f = 0; f < nfr; f++) {
    const BatchFrame &F = T[f];
    int *work = gptr(F.rwork);
    const int count = work[0];
    bool staged = false;
    for (;;) {
      if (tid == 0) L.ival[2] = atomicAdd(&work[1], 1);
      __syncthreads();
      const int idx = L.ival[2];
      __syncthreads();
      if (idx >= count) break;
      if (!staged) {
        stage_smats(F.fp, gptr(F.mats));
        __syncthreads();
        staged = true;
      }
      Mav1BlockInfo *gb = gptr(F.blocks) + work[2 + idx];
      if (tid < 27) ((int32_t *)&L.blk)[tid] = ((const int32_t *)gb)[tid];
      __syncthreads();
      recon_block(F.fp, gptr(F.src), gptr(F.rec), false, gptr(F.ref), gptr(F.gold), gptr(F.coefs), gptr(F.mats), false, false);
      if (tid < 27) ((int32_t *)gb)[tid] = ((const int32_t *)&L.blk)[tid];
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(SCHED_T) void k_intra_sched_b(const BatchFrame *__restrict__ T) {
  const BatchFrame &F = T[blockIdx.x];
  intra_sched_body(F.fp, gptr(F.blocks), gptr(F.sbs), gptr(F.iwork), 0, 0, 0);
}

/* Intra phase over several frames: a dataflow queue. A superblock holding
 * intra blocks enters the batch's ready queue when the last of its
 * dependencies (the neighbours in its tile column whose reconstruction it
 * reads and that hold intra blocks) is done, so a workgroup only ever takes
 * a superblock it can code at once, of whichever frame; superblocks without
 * such dependencies are queued first. q: [QH_HEAD] next entry to take,
 * [QH_TAIL] next entry to fill, [QH_TOTAL] superblocks of the round, then
 * the entries ((frame << 16) | superblock, -1 until filled). Every entry is
 * filled once all superblocks before it in dependency order are done, so a
 * workgroup's wait for its entry ends; the wait is still bounded, and one
 * that runs out sets every frame's error word. */
#define QH_HEAD 0
#define QH_TAIL 1
#define QH_TOTAL 2
#define QH_ITEMS 4

/* The superblocks whose reconstruction superblock sb reads (-1: none), as
 * the single-frame intra phase waits for them. */
DEVI void intra_deps(const Mav1FrameParams &fp, int sb, int dep[4]) {
  const int r = sb / fp.sb_cols, c = sb - r * fp.sb_cols;
  const int ts = (c / fp.tile_w_sb) * fp.tile_w_sb, te = d_min(fp.sb_cols, ts + fp.tile_w_sb);
  dep[0] = c > ts ? sb - 1 : -1;
  dep[1] = r > 0 && c > ts ? sb - fp.sb_cols - 1 : -1;
  dep[2] = r > 0 ? sb - fp.sb_cols : -1;
  dep[3] = r > 0 && c + 1 < te ? sb - fp.sb_cols + 1 : -1;
}

DEVI void queue_push(int *q, int v) {
  const int slot = atomicAdd(&q[QH_TAIL], 1);
  __hip_atomic_store(&q[QH_ITEMS + slot], v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_intra_queue_reset_b(int *__restrict__ q, int cap) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) q[QH_ITEMS + i] = -1;
  if (blockIdx.x == 0 && threadIdx.x < QH_ITEMS) q[threadIdx.x] = 0;
}

/* Per frame (one workgroup each, after the scheduling pass): the pending
 * dependency count of every superblock holding intra blocks; those with none
 * are queued. */
__global__ __launch_bounds__(256) void k_intra_ready_b(const BatchFrame *__restrict__ T, int *__restrict__ q) {
  const int f = blockIdx.x;
  const BatchFrame &F = T[f];
  const Mav1FrameParams &fp = F.fp;
  const int n = fp.sb_cols * fp.sb_rows;
  const int *has = gptr(F.iwork) + IW_HDR + n;
  int *cnt = gptr(F.iwork) + IW_HDR + 3 * n;
  int mine = 0;
  for (int sb = threadIdx.x; sb < n; sb += blockDim.x) {
    if (!has[sb]) continue;
    int dep[4], k = 0;
    intra_deps(fp, sb, dep);
    for (int j = 0; j < 4; j++) k += dep[j] >= 0 && has[dep[j]];
    cnt[sb] = k;
    mine++;
  }
  __syncthreads(); /* every count stored before any superblock is queued */
  for (int sb = threadIdx.x; sb < n; sb += blockDim.x)
    if (has[sb] && cnt[sb] == 0) queue_push(q, f << 16 | sb);
  if (mine) atomicAdd(&q[QH_TOTAL], mine);
}

__global__ __launch_bounds__(RT) void k_recon_intra_b(const BatchFrame *__restrict__ T, int nfr, int *__restrict__ q,
                                                      int maxSpin) {
  RLds &L = gL;
  const int tid = threadIdx.x;
  int staged = -1;
  for (;;) {
    if (tid == 0) {
      int v = -1;
      const int idx = atomicAdd(&q[QH_HEAD], 1);
      if (idx < q[QH_TOTAL]) {
        for (int it = 0; it < maxSpin && v < 0; it++) {
          v = __hip_atomic_load(&q[QH_ITEMS + idx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          if (v < 0) __builtin_amdgcn_s_sleep(4);
        }
        if (v < 0)
          for (int f = 0; f < nfr; f++) atomicOr(&T[f].iwork[IW_ERR], 1);
      }
      L.ival[2] = v;
    }
    __syncthreads();
    const int v = L.ival[2];
    __syncthreads();
    if (v < 0) break;
    const int f = v >> 16, sb = v & 0xffff;
    const BatchFrame &F = T[f];
    const Mav1FrameParams &fp = F.fp;
    if (staged < 0) stage_tabs(tid, RT);
    if (staged != f) {
      stage_smats(fp, gptr(F.mats));
      staged = f;
    }
    __syncthreads();
    const int r = sb / fp.sb_cols, c = sb - r * fp.sb_cols;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    intra_sb_prep(fp, gptr(F.blocks), gptr(F.sbs)[sb], gptr(F.src), r, c);
    const int err = intra_sb(fp, gptr(F.blocks), gptr(F.sbs)[sb], gptr(F.src), gptr(F.rec), gptr(F.ref), gptr(F.gold),
                             gptr(F.coefs), gptr(F.mats), r, c);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
      if (err) atomicOr(&gptr(F.iwork)[IW_ERR], 1);
      /* the superblocks that read this one: queued when this was their last
       * pending dependency */
      const int n = fp.sb_cols * fp.sb_rows;
      const int *has = gptr(F.iwork) + IW_HDR + n;
      int *cnt = gptr(F.iwork) + IW_HDR + 3 * n;
      const int cand[4] = {c + 1 < fp.sb_cols ? sb + 1 : -1, r + 1 < fp.sb_rows && c + 1 < fp.sb_cols ? sb + fp.sb_cols + 1 : -1,
                           r + 1 < fp.sb_rows ? sb + fp.sb_cols : -1, r + 1 < fp.sb_rows && c > 0 ? sb + fp.sb_cols - 1 : -1};
      for (int k = 0; k < 4; k++) {
        const int d = cand[k];
        if (d < 0 || !has[d]) continue;
        int dep[4];
        intra_deps(fp, d, dep);
        if (dep[0] != sb && dep[1] != sb && dep[2] != sb && dep[3] != sb) continue;
        if (__hip_atomic_fetch_add(&cnt[d], -1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 1) queue_push(q, f << 16 | d);
      }
    }
    __syncthreads();
  }
}

hipError_t launch_edge_types_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) batch_grid_add(g, H[f].fp.sb_cols * H[f].fp.sb_rows, 1, 1);
  hipLaunchKernelGGL(k_edge_flags_b, dim3(g.start[n]), dim3(64), 0, st, T, g);
  hipLaunchKernelGGL(k_edge_types_b, dim3(g.start[n]), dim3(64), 0, st, T, g);
  return hipGetLastError();
}

hipError_t launch_recon_inter_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int workgroups) {
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) batch_grid_add(g, H[f].fp.sb_cols * H[f].fp.sb_rows, 1, 1);
  hipLaunchKernelGGL(k_recon_list_reset_b, dim3(n), dim3(64), 0, st, T);
  hipLaunchKernelGGL(k_recon_list_b, dim3(g.start[n]), dim3(64), 0, st, T, g);
  hipLaunchKernelGGL(k_recon_inter_b, dim3(workgroups), dim3(RT), 0, st, T, n);
  return hipGetLastError();
}

hipError_t launch_intra_sched_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int *q, int qcap) {
  int need = 0;
  for (int f = 0; f < n; f++) need += H[f].fp.sb_cols * H[f].fp.sb_rows;
  if (need > qcap || n > (1 << 14)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_intra_sched_b, dim3(n), dim3(SCHED_T), 0, st, T);
  hipLaunchKernelGGL(k_intra_queue_reset_b, dim3(std::min((qcap + 255) / 256, 1024)), dim3(256), 0, st, q, qcap);
  hipLaunchKernelGGL(k_intra_ready_b, dim3(n), dim3(256), 0, st, T, q);
  return hipGetLastError();
}

hipError_t launch_recon_intra_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int *q,
                                int workgroups) {
  (void)H;
  hipLaunchKernelGGL(k_recon_intra_b, dim3(workgroups), dim3(RT), 0, st, T, n, q, 1 << 20);
  return hipGetLastError();
}
