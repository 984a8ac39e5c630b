/* backend_hip.c: the HIP backend of the mav1 encoder, for the encoder's own
 * source tree (src/backend_hip.c beside backend_cpu.c / backend_metal.m).
 *
 * Fills the Backend vtable of src/backend.h:10-24 with the C-ABI of
 * libmav1hip (include/mav1_hip.h): every stage reads and writes the
 * encoder's host buffers exactly as backend_cpu.c does, so the rest of the
 * encoder (driver, entropy coding, OBU writer) is unchanged. The turbo
 * frame_core keeps the intermediate buffers (motion fields, analysis
 * tables, blocks) in HBM between the stages and copies only the stage
 * results the CPU side reads. Two optional driver hooks run the CDEF choice
 * and the restoration search on the GPU as well (see hip_cdef_select /
 * hip_lr_select).
 *
 * Build (from the encoder's src/):
 *   gcc -O2 -std=gnu11 -c backend_hip.c -I<repo>/include
 *   ... link with -L<repo>/mav1-encoder_amd -lmav1hip -Wl,-rpath,<repo>/mav1-encoder_amd
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "backend.h"
#include "enc_internal.h"
#include "mav1_hip.h"

_Static_assert(sizeof(FrameParams) == sizeof(Mav1FrameParams), "FrameParams layout");
_Static_assert(sizeof(BlockInfo) == sizeof(Mav1BlockInfo), "BlockInfo layout");
_Static_assert(sizeof(SbInfo) == sizeof(Mav1SbInfo), "SbInfo layout");

typedef struct {
  Backend base;
  mav1hip_ctx *ctx;
} HipBackend;

static mav1hip_ctx *ctx_of(Backend *b) { return ((HipBackend *)b)->ctx; }
static const Mav1FrameParams *params_of(Mav1Encoder *e) { return (const Mav1FrameParams *)ENC_PARAMS(e); }

/* GPU errors are logged and the stage returns, as the Metal backend does
 * (src/backend_metal.m:75). */
static void check(Backend *b, int rc, const char *stage) {
  if (rc != MAV1HIP_OK) fprintf(stderr, "mav1 hip: %s: %s\n", stage, mav1hip_error(ctx_of(b)));
}

static void hip_motion_search(Backend *b, Mav1Encoder *e) {
  check(b,
        mav1hip_motion_search(ctx_of(b), params_of(e), e->src.base, e->ref.base, e->gold.base,
                              (int32_t *)e->mvfield.base, (int32_t *)e->mvgold.base),
        "motion_search");
}

static void hip_analyze(Backend *b, Mav1Encoder *e) {
  check(b,
        mav1hip_analyze(ctx_of(b), params_of(e), e->src.base, e->ref.base, e->gold.base,
                        (const int32_t *)e->mvfield.base, (const int32_t *)e->mvgold.base,
                        (Mav1BlockInfo *)e->blocks.base, (Mav1SbInfo *)e->sbs.base, (int64_t *)e->acosts.base,
                        (Mav1BlockInfo *)e->ainfos.base),
        "analyze");
}

static void hip_recon(Backend *b, Mav1Encoder *e) {
  check(b,
        mav1hip_recon(ctx_of(b), params_of(e), (Mav1BlockInfo *)e->blocks.base, (const Mav1SbInfo *)e->sbs.base,
                      e->src.base, e->rec.base, e->ref.base, e->gold.base, (int32_t *)e->coefs.base),
        "recon");
}

static void hip_filter(Backend *b, Mav1Encoder *e) {
  check(b,
        mav1hip_filter(ctx_of(b), params_of(e), e->lfmask.base, e->rec.base, e->num_cdef_presets, e->cdefbuf.base),
        "filter");
}

static void hip_restore(Backend *b, Mav1Encoder *e) {
  check(b,
        mav1hip_restore(ctx_of(b), params_of(e), e->rec.base, e->dbk.base, (const Mav1LrUnit *)e->lrunits.base,
                        e->lrout.base),
        "restore");
}

/* Motion search, analysis, edge filter types and reconstruction with the
 * intermediate buffers resident on the device. */
static void hip_frame_core(Backend *b, Mav1Encoder *e) {
  mav1hip_ctx *c = ctx_of(b);
  const Mav1FrameParams *fp = params_of(e);
  const size_t fb = e->frame_bytes;
  int rc = mav1hip_set_params(c, fp);
  if (!rc) rc = mav1hip_upload(c, MAV1HIP_BUF_SRC, e->src.base, fb);
  if (!rc && !fp->frame_is_intra) rc = mav1hip_upload(c, MAV1HIP_BUF_REF, e->ref.base, fb);
  if (!rc && !fp->frame_is_intra && fp->golden) rc = mav1hip_upload(c, MAV1HIP_BUF_GOLD, e->gold.base, fb);
  if (!rc && !fp->frame_is_intra) rc = mav1hip_launch_motion_search(c);
  if (!rc) rc = mav1hip_launch_analyze(c);
  if (!rc && e->cfg.intra_edge_filter) rc = mav1hip_launch_edge_types(c);
  if (!rc) rc = mav1hip_launch_recon(c);
  if (!rc)
    rc = mav1hip_download(c, MAV1HIP_BUF_BLOCKS, e->blocks.base, mav1hip_buffer_bytes(c, MAV1HIP_BUF_BLOCKS));
  if (!rc) rc = mav1hip_download(c, MAV1HIP_BUF_SBS, e->sbs.base, mav1hip_buffer_bytes(c, MAV1HIP_BUF_SBS));
  if (!rc) rc = mav1hip_download(c, MAV1HIP_BUF_COEFS, e->coefs.base, mav1hip_buffer_bytes(c, MAV1HIP_BUF_COEFS));
  if (!rc) rc = mav1hip_download(c, MAV1HIP_BUF_REC, e->rec.base, fb);
  /* both streams drained, then the intra phase's error word: a dependency
   * wait that ran out leaves the reconstruction wrong */
  if (!rc) rc = mav1hip_sync(c);
  check(b, rc, "frame_core");
}

/* Optional driver hooks: the driver's cdef_select (src/encoder.c:530-563)
 * and lr_select (src/lr_search.c:235) on the GPU, same results. */
void hip_cdef_select(Backend *b, Mav1Encoder *e) {
  check(b,
        mav1hip_cdef_select(ctx_of(b), params_of(e), e->src.base, e->cdefbuf.base, e->num_cdef_presets, e->rec.base,
                            (int8_t *)e->cdef_choice),
        "cdef_select");
}

void hip_lr_select(Backend *b, Mav1Encoder *e) {
  FrameParams *fp = ENC_PARAMS(e);
  int32_t types[3];
  int rc = mav1hip_lr_select(ctx_of(b), (const Mav1FrameParams *)fp, e->src.base, e->rec.base, e->dbk.base,
                             e->cfg.lr_test, (Mav1LrUnit *)e->lrunits.base, types, NULL);
  check(b, rc, "lr_select");
  if (rc == MAV1HIP_OK)
    for (int p = 0; p < 3; p++) fp->lr_type[p] = e->hdr.lr_type[p] = types[p];
}

static void hip_destroy(Backend *b) {
  mav1hip_destroy(ctx_of(b));
  free(b);
}

Backend *backend_hip_create(struct Mav1Encoder *e, char *err, size_t errlen) {
  Mav1HipConfig cfg;
  mav1hip_config_default(&cfg, e->w, e->h, e->cfg.turbo);
  cfg.qindex = e->cfg.qindex;
  cfg.keyint = e->cfg.keyint;
  cfg.tx_search = e->cfg.tx_search;
  cfg.tx_depth = e->cfg.tx_depth;
  cfg.loop_filter = e->cfg.loop_filter;
  cfg.cdef = e->cfg.cdef;
  cfg.restoration = e->cfg.restoration;
  cfg.tile_cols_log2 = e->cfg.tile_cols_log2;
  cfg.intra_edge_filter = e->cfg.intra_edge_filter;
  cfg.nonref_odd = e->cfg.nonref_odd;
  cfg.fast = e->cfg.fast;
  cfg.multi_ref = e->cfg.multi_ref;
  cfg.altref = e->cfg.altref;
  cfg.leaf = e->cfg.leaf;
  cfg.switchable_interp = e->cfg.switchable_interp;
  cfg.me_range = e->cfg.me_range;
  cfg.turbo = e->cfg.turbo;
  mav1hip_ctx *c = mav1hip_create(&cfg, err, errlen);
  if (!c) return NULL;
  HipBackend *h = calloc(1, sizeof(*h));
  h->ctx = c;
  h->base.name = "hip";
  h->base.motion_search = hip_motion_search;
  h->base.analyze = hip_analyze;
  h->base.recon = hip_recon;
  h->base.filter = hip_filter;
  h->base.restore = hip_restore;
  h->base.destroy = hip_destroy;
  h->base.frame_core = hip_frame_core;
  return &h->base;
}
