/* Host-side encoder configuration and per-frame parameter setup: the
 * driver arithmetic whose results the analysis kernels consume through
 * Mav1FrameParams. Restated from src/encoder.c (line references per
 * function) and tools/mav1enc.c; plain C, no device. */
#include <string.h>

#include "../../include/mav1_hip.h"
#include "av1_defs.h"
#include "av1_tables.h"

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

void mav1hip_config_default(Mav1HipConfig *cfg, int32_t width, int32_t height, int32_t turbo) {
  memset(cfg, 0, sizeof *cfg);
  /* mav1_default_config (src/encoder.c:51-66) */
  cfg->width = width;
  cfg->height = height;
  cfg->qindex = 100;
  cfg->tx_search = 1;
  cfg->tx_depth = 1;
  cfg->restoration = 1;
  cfg->switchable_interp = 1;
  cfg->intra_edge_filter = 1;
  cfg->loop_filter = 1;
  cfg->cdef = 1;
  cfg->me_range = 32;
  /* CLI presets (tools/mav1enc.c:98-100) */
  cfg->turbo = turbo;
  cfg->fast = turbo;
  cfg->tile_cols_log2 = turbo ? 2 : 0;
  if (cfg->fast) cfg->tx_search = cfg->tx_depth = cfg->restoration = cfg->switchable_interp = 0;
  cfg->knob_turbo_mask = cfg->knob_dt = cfg->knob_dt_bias = cfg->knob_cdef_preset = cfg->knob_leafopt = -1;
}

void mav1hip_config_normalize(Mav1HipConfig *cfg) {
  /* src/encoder.c:75-92 */
  if (cfg->turbo) {
    cfg->fast = 1;
    if (cfg->altref < 2) cfg->altref = 2;
    cfg->leaf = 1;
  }
  if (cfg->altref == 1) cfg->altref = 0;
  if (cfg->altref > 16) cfg->altref = 16;
  if (cfg->altref) cfg->multi_ref = cfg->nonref_odd = 0;
  if (!cfg->altref) cfg->leaf = 0;
  if (cfg->qindex < 0) cfg->qindex = 0;
  if (cfg->qindex > 255) cfg->qindex = 255;
}

size_t mav1hip_geometry(const Mav1HipConfig *cfg, int32_t *mi_cols, int32_t *mi_rows, int32_t *sb_cols,
                        int32_t *sb_rows, int32_t *tile_w_sb, int32_t stride[3], int32_t offset[3],
                        int32_t plane_h[3]) {
  int w = cfg->width, h = cfg->height;
  if (w < 16 || h < 16 || w > 65536 || h > 65536) return 0;
  /* src/encoder.c:93-120 */
  int mc = 2 * ((w + 7) >> 3), mr = 2 * ((h + 7) >> 3);
  int sc = (mc + 15) >> 4, sr = (mr + 15) >> 4;
  int k = 0;
  while (((MAX_TILE_WIDTH >> 6) << k) < sc) k++;
  int lo = k, maxCols = imin(sc, MAX_TILE_COLS);
  for (k = 0; (1 << k) < maxCols; k++) {
  }
  int l = imin(imax(cfg->tile_cols_log2, lo), k);
  if (mi_cols) *mi_cols = mc;
  if (mi_rows) *mi_rows = mr;
  if (sb_cols) *sb_cols = sc;
  if (sb_rows) *sb_rows = sr;
  if (tile_w_sb) *tile_w_sb = (sc + (1 << l) - 1) >> l;
  int st[3] = {sc * 64, sc * 32, sc * 32}, ph[3] = {sr * 64, sr * 32, sr * 32};
  int of[3] = {0, st[0] * ph[0], st[0] * ph[0] + st[1] * ph[1]};
  for (int p = 0; p < 3; p++) {
    if (stride) stride[p] = st[p];
    if (offset) offset[p] = of[p];
    if (plane_h) plane_h[p] = ph[p];
  }
  return (size_t)of[2] + (size_t)st[2] * ph[2];
}

int32_t mav1hip_frame_qindex(const Mav1HipConfig *cfg, const Mav1HipFrame *fr) {
  /* frame_qindex (src/encoder.c:585-596) with the default offsets of the
   * MAV1_Q* knobs */
  int q = cfg->qindex;
  if (q == 0) return 0;
  int isKey = fr->is_key;
  int off = isKey ? -24 : (fr->order_hint % 4 == 0 ? -8 : 4);
  if (!isKey && fr->nonref) off = 12;
  if (!isKey && cfg->altref) off = fr->frame_kind == 1 ? -8 : 4;
  if (!isKey && cfg->leaf && fr->frame_kind == 3) off = 0;
  if (!isKey && cfg->leaf && fr->nonref) off = 8;
  q += off;
  return q < 1 ? 1 : q > 255 ? 255 : q;
}

int32_t mav1hip_setup_params(const Mav1HipConfig *cfg, const Mav1HipFrame *fr, int32_t q, Mav1FrameParams *fp) {
  int32_t mc, mr, sc, sr, tw, st[3], of[3], ph[3];
  memset(fp, 0, sizeof *fp);
  if (!mav1hip_geometry(cfg, &mc, &mr, &sc, &sr, &tw, st, of, ph)) return -1;
  int isKey = fr->is_key;
  /* the leaf-frame savings of this frame (src/encoder.c:1007-1008) */
  int leafoptCfg = cfg->knob_leafopt >= 0 ? cfg->knob_leafopt : 0x3F;
  int leafopt = (cfg->altref && fr->nonref) ? leafoptCfg : 0;
  /* setup_params (src/encoder.c:407-510) */
  fp->mi_cols = mc;
  fp->mi_rows = mr;
  fp->sb_cols = sc;
  fp->sb_rows = sr;
  for (int p = 0; p < 3; p++) {
    fp->plane_w[p] = st[p];
    fp->plane_h[p] = ph[p];
    fp->stride[p] = st[p];
    fp->offset[p] = of[p];
  }
  fp->base_q_idx = q;
  for (int p = 0; p < 3; p++) {
    fp->dc_q[p] = DC_Q(q);
    fp->ac_q[p] = AC_Q(q);
  }
  fp->frame_is_intra = isKey;
  fp->reduced_tx_set = 0;
  fp->enable_intra_edge_filter = cfg->intra_edge_filter;
  fp->tx_search = cfg->tx_search;
  fp->lossless = q == 0;
  fp->tx_mode_select = q != 0 && cfg->tx_depth;
  fp->switchable_interp = !isKey && cfg->switchable_interp;
  fp->golden = !isKey && cfg->multi_ref && fr->gold_differs;
  fp->ref2_frame = GOLDEN_FRAME;
  if (cfg->altref) {
    fp->golden = !isKey && fr->frame_kind >= 2;
    fp->ref2_frame = ALTREF_FRAME;
    fp->compound = fr->reference_select;
  }
  fp->fast = cfg->fast;
  fp->turbo = cfg->turbo ? (cfg->knob_turbo_mask >= 0 ? cfg->knob_turbo_mask : MAV1_TURBO_DEFAULT) : 0;
  fp->leaf = cfg->leaf ? (leafopt & 1) : fr->nonref;
  if (leafopt & 16) fp->tx_search = 0;
  {
    int dt = cfg->knob_dt >= 0 ? cfg->knob_dt : sr >= 16 ? 320 : 352;
    int bias = cfg->knob_dt_bias >= 0 ? cfg->knob_dt_bias : 0;
    fp->depth_thr = cfg->turbo ? dt * (q + bias) / (128 + bias) : 0;
  }
  fp->tile_w_sb = tw;
  long acq = fp->ac_q[0];
  fp->lambda = (int32_t)(acq * acq * 256 / 430);
  fp->frame_w = cfg->width;
  fp->frame_h = cfg->height;
  fp->me_range = cfg->me_range;
  if ((leafopt & 32) && fp->me_range > 16) fp->me_range = 16;
  /* deblocking level from q (libaom's LPF_PICK_FROM_Q rule for 8-bit) */
  int lvl = 0;
  if (cfg->loop_filter && q > 0 && !(cfg->leaf ? (leafopt & 2) : fr->nonref)) {
    lvl = (int)((acq * 20723 + 1015158 + (1 << 17)) >> 18);
    if (!isKey) lvl -= 4;
    lvl = lvl < 0 ? 0 : (lvl > MAX_LOOP_FILTER ? MAX_LOOP_FILTER : lvl);
  }
  for (int i = 0; i < 4; i++) fp->lf_level[i] = lvl;
  fp->lf_sharpness = 0;
  /* CDEF strength presets: (y_pri, y_sec, uv_pri, uv_sec) */
  static const int presets[4][4] = {{0, 0, 0, 0}, {2, 1, 1, 0}, {4, 2, 2, 1}, {8, 4, 4, 2}};
  fp->cdef_damping = 3 + (q >> 6);
  fp->cdef_bits = 2;
  if (leafopt & 4) {
    fp->cdef_damping = 3;
    fp->cdef_bits = 0;
  }
  int pi[4] = {0, 1, 2, 3};
  if ((fp->turbo & MAV1_T_CDEF2) && fp->cdef_bits == 2) {
    fp->cdef_bits = 1;
    pi[1] = cfg->knob_cdef_preset >= 0 ? cfg->knob_cdef_preset : 2;
  }
  for (int i = 0; i < 4; i++) {
    fp->cdef_y_pri[i] = presets[pi[i]][0];
    fp->cdef_y_sec[i] = presets[pi[i]][1];
    fp->cdef_uv_pri[i] = presets[pi[i]][2];
    fp->cdef_uv_sec[i] = presets[pi[i]][3];
  }
  /* loop-restoration units of 64 x 64 (count_units_in_frame) */
  int off = 0;
  for (int p = 0; p < 3; p++) {
    int sub = p > 0, size = RESTORATION_TILESIZE_MAX >> 2;
    int pw = (cfg->width + sub) >> sub, phh = (cfg->height + sub) >> sub;
    fp->lr_type[p] = RESTORE_NONE;
    fp->lr_unit_size[p] = size;
    fp->lr_unit_rows[p] = imax((phh + (size >> 1)) / size, 1);
    fp->lr_unit_cols[p] = imax((pw + (size >> 1)) / size, 1);
    fp->lr_unit_off[p] = off;
    off += fp->lr_unit_rows[p] * fp->lr_unit_cols[p];
  }
  /* presets the filter stage evaluates (src/encoder.c:618-631) */
  int presetsOn = cfg->cdef ? 4 : 0;
  if (presetsOn) presetsOn = 1 << fp->cdef_bits;
  if (leafopt & 4) presetsOn = 0;
  return presetsOn;
}

int32_t mav1hip_filter_plan(const Mav1HipConfig *cfg, const Mav1HipFrame *fr, const Mav1FrameParams *fp,
                            int32_t npresets, Mav1HipFilterPlan *plan) {
  if (!cfg || !fr || !fp || !plan) return MAV1HIP_EINVAL;
  /* run_frame_core's filter decisions (src/encoder.c:618-640) */
  int leafoptCfg = cfg->knob_leafopt >= 0 ? cfg->knob_leafopt : 0x3F;
  int leafopt = (cfg->altref && fr->nonref) ? leafoptCfg : 0;
  int cdefOff = (leafopt & 4) != 0;
  int lfOn = fp->lf_level[0] || fp->lf_level[1] || fp->lf_level[2] || fp->lf_level[3];
  int filt = lfOn || (npresets && !cdefOff) || !cfg->turbo;
  plan->filter_info = filt;
  plan->deblock = filt && lfOn;
  plan->cdef_sets = filt && !cdefOff ? npresets : 0;
  plan->cdef_select = cfg->cdef && fp->base_q_idx > 0 && cfg->qindex > 0 && !cdefOff && npresets > 0;
  plan->restoration = cfg->restoration && cfg->qindex > 0 && !(leafopt & 8);
  return MAV1HIP_OK;
}
