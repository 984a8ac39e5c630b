/* ORACLE — test infrastructure only. Motion search restated from
 * src/kernels/k_me.h and the CPU driver src/backend_cpu.c:23-56. */
#include "ocommon.h"

#define TURBO(fp, bit) (((fp)->turbo & (bit)) != 0)

long oracle_qp_entry_bytes(const Mav1FrameParams *fp) {
  long W = fp->frame_w, H = fp->frame_h;
  return 16 * W * H + ((W + 3) >> 2) * ((H + 3) >> 2); /* k_me.h:105-107 */
}

/* k_me.h:289-313: plane fy * 4 + fx holds the EIGHTTAP prediction at
 * (y + fy / 4, x + fx / 4) of every luma sample, computed like the inter
 * predictor (horizontal Round2 3, vertical Round2 11, clip), with frame-edge
 * clamping of the 8 x 8 support. Planes are frame_w x frame_h, unpadded. */
void oracle_qpel_build(const Mav1FrameParams *fp, const uint8_t *ref, uint8_t *planes16) {
  const uint8_t *luma = ref + fp->offset[0];
  int W = fp->frame_w, H = fp->frame_h, S = fp->stride[0];
  size_t n = (size_t)W * H;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      int hz[4][8];
      for (int t = 0; t < 8; t++) {
        const uint8_t *row = luma + o_clip(0, H - 1, y + t - 3) * S;
        for (int fx = 0; fx < 4; fx++) {
          const int16_t *k = SUBPEL_FILTER(EIGHTTAP, 4 * fx);
          int acc = 0;
          for (int u = 0; u < 8; u++) acc += k[u] * row[o_clip(0, W - 1, x + u - 3)];
          hz[fx][t] = o_round2(acc, 3);
        }
      }
      for (int fy = 0; fy < 4; fy++) {
        const int16_t *k = SUBPEL_FILTER(EIGHTTAP, 4 * fy);
        for (int fx = 0; fx < 4; fx++) {
          int acc = 0;
          for (int t = 0; t < 8; t++) acc += k[t] * hz[fx][t];
          planes16[(size_t)(fy * 4 + fx) * n + (size_t)y * W + x] = (uint8_t)o_clip1(o_round2(acc, 11));
        }
      }
    }
}

/* k_me.h:112-118: rounded mean of each 4x4 luma block (edge-clamped). */
void oracle_dec_build(const Mav1FrameParams *fp, const uint8_t *frame, uint8_t *dec) {
  const uint8_t *luma = frame + fp->offset[0];
  int W = fp->frame_w, H = fp->frame_h, dw = (W + 3) >> 2, dh = (H + 3) >> 2;
  for (int y = 0; y < dh; y++)
    for (int x = 0; x < dw; x++) {
      int s = 0;
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
          s += luma[o_clip(0, H - 1, 4 * y + i) * fp->stride[0] + o_clip(0, W - 1, 4 * x + j)];
      dec[(size_t)y * dw + x] = (uint8_t)((s + 8) >> 4);
    }
}

/* Search state of one node (k_me.h:66-76, 141-178). */
typedef struct {
  int x, y, w;
  int cy, cx;      /* full-pel grid centre */
  int range, step; /* grid radius and spacing */
  int nstage;      /* full-pel stages (level 0: 2 or 3; 0 with T_ME_L3SUB at level 3) */
  int pr, pc;      /* parent vector, 1/8 pel */
  int copy;
} Node;

typedef struct {
  const Mav1FrameParams *fp;
  const uint8_t *src, *ref, *planes, *srcDec, *refDec;
} Refs;

static int grid_side(const Node *m) { return 2 * (m->range / m->step) + 1; }

/* Candidate k of the current stage: 0 = grid centre, then the grid in
 * raster order (k_me.h:199-209). */
static void cand(const Node *m, int k, int *dy, int *dx) {
  if (k == 0) { *dy = m->cy; *dx = m->cx; return; }
  int side = grid_side(m);
  *dy = m->cy - m->range + ((k - 1) / side) * m->step;
  *dx = m->cx - m->range + ((k - 1) % side) * m->step;
}

/* Full-resolution SAD with reference coordinates clamped to the frame;
 * rs > 1 visits every rs-th row and scales back (k_me.h:39-64). */
static int sad_full(const Refs *R, const Node *m, int dy, int dx, int rs) {
  const Mav1FrameParams *fp = R->fp;
  const uint8_t *s = R->src + fp->offset[0], *r = R->ref + fp->offset[0];
  int S = fp->stride[0], sum = 0;
  for (int i = 0; i < m->w; i += rs) {
    const uint8_t *rr = r + o_clip(0, fp->frame_h - 1, m->y + i + dy) * S;
    for (int j = 0; j < m->w; j++)
      sum += o_abs((int)s[(m->y + i) * S + m->x + j] - (int)rr[o_clip(0, fp->frame_w - 1, m->x + j + dx)]);
  }
  return sum * rs;
}

/* SAD on the 4x4-decimated planes at an offset that is a multiple of 4,
 * scaled by 16 (k_me.h:121-130). */
static int sad_dec(const Refs *R, const Node *m, int dy, int dx) {
  const Mav1FrameParams *fp = R->fp;
  int W = (fp->frame_w + 3) >> 2, H = (fp->frame_h + 3) >> 2, n = m->w >> 2;
  int x0 = m->x >> 2, y0 = m->y >> 2, oy = dy >> 2, ox = dx >> 2, sum = 0;
  for (int i = 0; i < n; i++) {
    int sy = o_min(y0 + i, H - 1), ry = o_clip(0, H - 1, y0 + i + oy);
    for (int j = 0; j < n; j++)
      sum += o_abs((int)R->srcDec[sy * W + o_min(x0 + j, W - 1)] - (int)R->refDec[ry * W + o_clip(0, W - 1, x0 + j + ox)]);
  }
  return sum * 16;
}

/* k_me.h:211-220: distortion + ac_q * (approximate vector bits) / 64. */
static int fullpel_cost(const Refs *R, const Node *m, int dy, int dx) {
  const Mav1FrameParams *fp = R->fp;
  int bits = 2 + 2 * o_floorlog2((unsigned)(o_abs(dy) * 8 + 1)) + 2 * o_floorlog2((unsigned)(o_abs(dx) * 8 + 1));
  int rate = fp->ac_q[0] * bits / 64;
  if (TURBO(fp, MAV1_T_ME_DEC) && m->step >= 4) return sad_dec(R, m, dy, dx) + rate;
  int rs = (TURBO(fp, MAV1_T_ME_RSUB) && m->step == 1) ? 2 : 1;
  return sad_full(R, m, dy, dx, rs) + rate;
}

/* Sub-pel distortion at an even 1/8-pel vector (k_me.h:247-258, 317-340):
 * from the quarter-pel planes (T_ME_HPEL; T_ME_USUB: every other 4x4 unit
 * of the raster order for nodes >= 32 wide, x2), else from the normative
 * EIGHTTAP prediction of the block. */
static int subpel_cost(const Refs *R, const Node *m, int mvr, int mvc) {
  const Mav1FrameParams *fp = R->fp;
  const uint8_t *s = R->src + fp->offset[0];
  int S = fp->stride[0];
  if (TURBO(fp, MAV1_T_ME_HPEL)) {
    int W = fp->frame_w, H = fp->frame_h, nr = m->w >> 2, nu = nr * nr;
    int st = (TURBO(fp, MAV1_T_ME_USUB) && m->w >= 32) ? 2 : 1;
    const uint8_t *pl = R->planes + (size_t)(((mvr >> 1) & 3) * 4 + ((mvc >> 1) & 3)) * W * H;
    int total = 0;
    for (int u = 0; u < nu; u += st) {
      int ux = m->x + (u % nr) * 4, uy = m->y + (u / nr) * 4, iy = uy + (mvr >> 3), ix = ux + (mvc >> 3);
      int32_t d[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
          d[4 * i + j] = (int)s[(uy + i) * S + ux + j] -
                         (int)pl[(size_t)o_clip(0, H - 1, iy + i) * W + o_clip(0, W - 1, ix + j)];
      total += oracle_satd4x4(d);
    }
    return total * st;
  }
  uint8_t pred[64 * 64];
  oracle_predict_inter(R->ref + fp->offset[0], S, fp->frame_w, fp->frame_h, pred, m->w, m->x, m->y, m->w, m->w, mvr,
                       mvc, 0, 0, EIGHTTAP);
  int total = 0;
  for (int by = 0; by < m->w; by += 4)
    for (int bx = 0; bx < m->w; bx += 4) {
      int32_t d[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) d[4 * i + j] = (int)s[(m->y + by + i) * S + m->x + bx + j] - (int)pred[(by + i) * m->w + bx + j];
      total += oracle_satd4x4(d);
    }
  return total;
}

/* Sub-pel neighbour order (k_me.h:234-243): 3x3 ring in raster order
 * without the centre; T_ME_CROSS keeps the 4 axial ones. */
static void ring(const Mav1FrameParams *fp, int j, int step, int *dr, int *dc) {
  static const int ry[8] = {-1, -1, -1, 0, 0, 1, 1, 1}, rx[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
  static const int cross[4] = {1, 3, 4, 6};
  int k = TURBO(fp, MAV1_T_ME_CROSS) ? cross[j] : j;
  *dr = ry[k] * step;
  *dc = rx[k] * step;
}

/* One node, as me_node_serial (k_me.h:343-379). */
static void search_node(const Refs *R, int sb, int level, int i, int isGold, int32_t *field) {
  const Mav1FrameParams *fp = R->fp;
  int n = 1 << level;
  Node m;
  memset(&m, 0, sizeof m);
  m.w = 64 >> level;
  m.x = (sb % fp->sb_cols) * 64 + (i % n) * m.w;
  m.y = (sb / fp->sb_cols) * 64 + (i / n) * m.w;
  int32_t *out = field + (sb * MAV1_NUM_NODES + o_node_base(level) + i) * 2;
  if (m.x >= fp->frame_w || m.y >= fp->frame_h) {
    out[0] = out[1] = 0;
    return;
  }
  m.range = fp->me_range;
  m.step = 1;
  m.nstage = 1;
  if (level > 0) {
    int parent = ((i / n) >> 1) * (n >> 1) + ((i % n) >> 1);
    const int32_t *pv = field + (sb * MAV1_NUM_NODES + o_node_base(level - 1) + parent) * 2;
    m.pr = pv[0];
    m.pc = pv[1];
    m.cy = m.pr >> 3;
    m.cx = m.pc >> 3;
    m.range = o_min(m.range, TURBO(fp, MAV1_T_ME_CR1) ? 1 : TURBO(fp, MAV1_T_ME_CR2) ? 2 : 4);
    int leafNo8 = TURBO(fp, MAV1_T_LEAF_NO8) && fp->leaf && !fp->frame_is_intra;
    m.copy = level == 3 && (TURBO(fp, MAV1_T_ME_L3COPY) || leafNo8);
  } else {
    m.step = TURBO(fp, MAV1_T_ME_HME3) ? 8 : 4;
    m.range = (m.range / m.step) * m.step;
    m.nstage = TURBO(fp, MAV1_T_ME_HME3) ? 3 : 2;
  }
  if (level == 3 && TURBO(fp, MAV1_T_ME_L3SUB)) m.nstage = 0;
  /* leaf frames: the second reference's 8x8 nodes copy their parent (k_me.h:94-96, 347) */
  if (isGold && level == 3 && TURBO(fp, MAV1_T_ME_G3COPY) && fp->leaf && !fp->frame_is_intra) m.copy = 1;
  if (m.copy) {
    out[0] = m.pr;
    out[1] = m.pc;
    return;
  }
  int by = 0, bx = 0;
  for (int st = 0; st < m.nstage; st++) {
    if (st > 0) { /* k_me.h:182-194 */
      m.cy = by;
      m.cx = bx;
      if (st == 1 && TURBO(fp, MAV1_T_ME_HME3)) {
        m.range = 4;
        m.step = 4;
      } else {
        m.range = TURBO(fp, MAV1_T_ME_REF2) ? 2 : 3;
        m.step = 1;
      }
    }
    int side = grid_side(&m), count = 1 + side * side, best = 0;
    for (int k = 0; k < count; k++) {
      int dy, dx;
      cand(&m, k, &dy, &dx);
      int c = fullpel_cost(R, &m, dy, dx);
      if (k == 0 || c < best) { /* first minimum in candidate order */
        best = c;
        by = dy;
        bx = dx;
      }
    }
  }
  if (level > 0 && TURBO(fp, MAV1_T_ME_KEEP) && by == m.cy && bx == m.cx) {
    out[0] = m.pr;
    out[1] = m.pc;
    return;
  }
  /* k_me.h:262-281: half-pel then quarter-pel ring steps (T_ME_HALF01:
   * half-pel only for nodes >= 32 wide); first minimum, centre first. */
  int mvr = m.nstage ? by * 8 : m.pr, mvc = m.nstage ? bx * 8 : m.pc;
  int bestS = subpel_cost(R, &m, mvr, mvc);
  int minStep = (m.w >= 32 && TURBO(fp, MAV1_T_ME_HALF01)) ? 4 : 2;
  int nring = TURBO(fp, MAV1_T_ME_CROSS) ? 4 : 8;
  for (int step = 4; step >= minStep; step >>= 1) {
    int cr = mvr, cc = mvc;
    for (int j = 0; j < nring; j++) {
      int dr, dc;
      ring(fp, j, step, &dr, &dc);
      int c = subpel_cost(R, &m, cr + dr, cc + dc);
      if (c < bestS) {
        bestS = c;
        mvr = cr + dr;
        mvc = cc + dc;
      }
    }
  }
  out[0] = mvr;
  out[1] = mvc;
}

/* backend_cpu.c:23-56: plane build (T_ME_HPEL), then levels 0..3, LAST and
 * (fp->golden) the second reference. */
void oracle_motion_search(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref, const uint8_t *gold,
                          int32_t *mvfield, int32_t *mvgold) {
  int nref = fp->golden ? 2 : 1, nsb = fp->sb_cols * fp->sb_rows;
  long eb = oracle_qp_entry_bytes(fp);
  size_t n16 = (size_t)16 * fp->frame_w * fp->frame_h;
  uint8_t *buf = (uint8_t *)calloc((size_t)(nref * eb + eb), 1);
  uint8_t *srcDec = buf + nref * eb;
  if (TURBO(fp, MAV1_T_ME_HPEL)) {
    for (int r = 0; r < nref; r++) {
      oracle_qpel_build(fp, r ? gold : ref, buf + r * eb);
      if (TURBO(fp, MAV1_T_ME_DEC)) oracle_dec_build(fp, r ? gold : ref, buf + r * eb + n16);
    }
    if (TURBO(fp, MAV1_T_ME_DEC)) oracle_dec_build(fp, src, srcDec);
  }
  for (int level = 0; level < 4; level++)
    for (int r = 0; r < nref; r++) {
      Refs R = {fp, src, r ? gold : ref, buf + r * eb, srcDec, buf + r * eb + n16};
      /* superblocks are independent within a level */
#pragma omp parallel for schedule(dynamic, 4)
      for (int sb = 0; sb < nsb; sb++)
        for (int i = 0; i < (1 << (2 * level)); i++) search_node(&R, sb, level, i, r, r ? mvgold : mvfield);
    }
  free(buf);
}
