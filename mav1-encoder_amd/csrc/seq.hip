/* Sequence driver (include/mav1_hip.h, mav1hip_seq_*): the encoder's frame
 * driver for the backend path, on HIP streams and events.
 *
 * What it restates from the reference:
 *  - the display-order input queue and the coding order of ALTREF
 *    mini-GOPs: alt_drain / pyr_code (src/encoder.c:1229-1293);
 *  - each frame's state (frame kind, order hint, reference slots,
 *    refresh flags, compound): alt_code (:1004-1030), then frame_qindex and
 *    setup_params;
 *  - the stages of run_frame_core (:598-647) and their filter decisions;
 *  - engines: the ALTREF chain on one engine and the other frames on a
 *    second (:1046-1052), slot copies into the engine that reads a slot
 *    (:1088-1118), slot writes after the core (:1128-1140).
 *
 * How it runs on the GPU:
 *  - an engine is a context (its own streams and buffers) plus kPics
 *    device pictures; each of the 8 reference slots maps, per engine, to
 *    one of its pictures and the generation of the slot it holds. A frame
 *    that refreshes a slot stores its final picture into a picture no slot
 *    of its engine maps to; an engine that reads a slot coded elsewhere
 *    copies it once on its own stream, after an event that completes the
 *    picture on the engine that wrote it. That engine waits for the copy
 *    only when it next writes the picture (a per-picture read event), so it
 *    never waits for the reading engine's earlier frames;
 *  - the driver uses four streams per device, created once per process
 *    (the GPU runs four hardware queues per process, and a stream gets its
 *    queue when it is created): engine 0's main and side streams, engine
 *    1's; uploads and downloads run on engine 1's side stream, which the
 *    critical ALTREF chain never waits on;
 *  - a frame of the second engine is enqueued after the next ALTREF's
 *    motion search, analysis and edge types, so it runs beside that
 *    ALTREF's reconstruction (whose intra phase leaves most CUs idle);
 *  - every frame's outputs are copied on its engine's stream into a device
 *    staging slot (one launch), then to pinned host memory on the transfer
 *    stream, so the engines never wait for PCIe; host-sent frames go up on
 *    the transfer stream into a per-engine ring of source frames.
 * All streams have the default priority. */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <mutex>
#include <vector>

#include "../../include/mav1_hip.h"
#include "internal.h"

namespace {

constexpr int kEngines = 2;
constexpr int kMaxDevices = 64;
constexpr int kPics = 10; /* pictures per engine: 8 slots, the frame's own store, one spare */
constexpr int kSrcRing = 4;

struct Input {
  int disp;
  int stage;          /* host staging slot, -1 for a device frame */
  const uint8_t *dev; /* the caller's device frame */
};

/* A frame in coding order with everything alt_code decides. */
struct Frame {
  Input in;
  int is_key, kind, last_slot, arf_slot, refresh, hidden, nonref;
  int order_hint, last_hint, arf_hint;
  int engine;
  int out;        /* output slot */
  long long coded; /* coding index */
  Mav1HipFrame fr;
  Mav1FrameParams fp;
  Mav1HipFilterPlan plan;
  int npresets;
  uint8_t *src;   /* device source frame the stages read */
  int src_ring;   /* ring entry of the engine (-1: the caller's device frame) */
};

struct OutSlot {
  uint8_t *d = nullptr; /* device staging */
  uint8_t *h = nullptr; /* pinned host mirror */
  hipEvent_t staged = nullptr, landed = nullptr;
  int busy = 0;
  Mav1HipCodedFrame info;
  int32_t hdr7[7];
};

struct Engine {
  mav1hip_ctx *ctx = nullptr;
  hipStream_t st = nullptr;
  uint8_t *pics = nullptr;
  int map[8];
  unsigned have[8];
  uint8_t *ring = nullptr; /* kSrcRing source frames */
  hipEvent_t ring_free[kSrcRing] = {};
  int ring_used[kSrcRing] = {};
  int ring_next = 0;
  hipEvent_t ev = nullptr; /* scratch */
  hipEvent_t pic_read[kPics] = {}; /* the last copy of the picture to the other engine */
  int pic_pend[kPics] = {};
};

}  // namespace

struct mav1hip_seq {
  Mav1HipConfig cfg;
  int lr_test = 0, dump = 0;
  char err[512] = {0};
  size_t frame_bytes = 0;
  int num_sbs = 0, lr_units = 0;
  size_t grid_bytes = 0;
  int32_t stride[3], offset[3], plane_h[3];
  Engine eng[kEngines];
  hipStream_t tx = nullptr; /* transfers: the second engine's side stream */
  /* host staging of sent frames */
  std::vector<uint8_t *> stage;
  std::vector<hipEvent_t> stage_done;
  std::vector<int> stage_busy;
  std::deque<Input> inq;
  int in_disp = 0, ended = 0;
  int key_disp = 0, anchor = 0;
  unsigned gen = 0;
  unsigned slot_gen[8] = {};
  int slot_owner[8];
  int slot_hint[8];
  /* outputs */
  std::vector<OutSlot> outs;
  std::deque<int> order; /* output slots in coding order, not yet received */
  int held = -1;         /* slot handed to the caller by the last receive */
  long long coded = 0;
  std::vector<Frame> pending; /* second-engine frames waiting for the next ALTREF's early stages */
  mav1hip_batch *batch = nullptr; /* coded in batches with other sequences (frames queue in `jobs`) */
  std::deque<Frame> jobs;
  size_t o_blocks, o_sbs, o_coefs, o_choice, o_units, o_lrtype, o_err, o_pic, o_grid, o_end;
};

namespace {
constexpr int kTabs = 8; /* batch tables in flight */

/* Frames of several sequences coded together: the same position of their
 * coding order, the same stages. */
struct Round {
  int engine = 0;
  int tab = -1;
  std::vector<mav1hip_seq *> seqs;
  std::vector<Frame> frames;
};
}  // namespace

struct mav1hip_batch {
  std::vector<mav1hip_seq *> seqs;
  int lr_test = 0, dump = 0;
  char err[512] = {0};
  int num_cu = 0;
  /* the frame tables: kTabs x kMaxBatch entries in device memory and their
   * pinned host mirror; a table is reused once the round that used it has
   * finished on the GPU */
  BatchFrame *d_tab = nullptr, *h_tab = nullptr;
  /* per engine (whose rounds run beside each other's), the intra phase's
   * ready queue: 4 header ints and one entry per superblock of a round (at
   * most every stream's superblocks) */
  int *d_iq[2] = {};
  int iq_cap = 0;
  hipEvent_t tab_free[kTabs] = {};
  int tab_used[kTabs] = {};
  int tab_next = 0;
  /* per engine: analysis beside motion search */
  hipEvent_t me_done[2][4] = {}, ana_done[2] = {};
  std::vector<Round> pending; /* second-engine rounds waiting for the next ALTREF round's early stages */
  /* launch shapes (mav1hip_batch_set_knob): persistent workgroups of the
   * intra phase in rounds of intra frames and of inter frames, and of the
   * inter phase; 0 = the default */
  int intra_wg_key = 0, intra_wg_inter = 0, inter_wg = 0;
  /* the learned pruner, per engine (mav1hip_batch_set_pruner): its network,
   * the per-superblock deepest levels, and an event after the table upload
   * that its inference on the side stream waits for */
  mav1hip_pruner *pruner[2] = {};
  int8_t *d_prune[2] = {};
  hipEvent_t tab_up[2] = {};
  float prune_tau = 0.f;
};

namespace {

int seq_fail(mav1hip_seq *s, int code, const char *fmt, ...) {
  if (s) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(s->err, sizeof s->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

#define SCHK(s, call)                                                                                       \
  do {                                                                                                      \
    hipError_t e_ = (call);                                                                                 \
    if (e_ != hipSuccess)                                                                                   \
      return seq_fail(s, MAV1HIP_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

/* A context call's status, with the context's message. */
int cchk(mav1hip_seq *s, int eng, int rc) {
  if (rc == MAV1HIP_OK) return rc;
  return seq_fail(s, rc, "engine %d: %s", eng, mav1hip_error(s->eng[eng].ctx));
}
#define CCHK(s, eng, call)              \
  do {                                  \
    int rc_ = cchk(s, eng, (call));     \
    if (rc_) return rc_;                \
  } while (0)

size_t al(size_t n) { return (n + 255) & ~(size_t)255; }

uint8_t *pic(mav1hip_seq *s, int e, int idx) { return s->eng[e].pics + (size_t)idx * s->frame_bytes; }

/* A picture of engine e that no slot maps to and that is none of `avoid`. */
int free_pic(mav1hip_seq *s, int e, const int *avoid, int navoid) {
  for (int i = 0; i < kPics; i++) {
    bool used = false;
    for (int k = 0; k < 8 && !used; k++) used = s->eng[e].map[k] == i;
    for (int j = 0; j < navoid && !used; j++) used = avoid[j] == i;
    if (!used) return i;
  }
  return -1;
}

/* The picture of engine e holding slot k's current version, copied there
 * from the engine that wrote it when e holds an older one. */
int slot_pic(mav1hip_seq *s, int e, int k, const int *avoid, int navoid, int *idx) {
  Engine &E = s->eng[e];
  const unsigned g = s->slot_gen[k];
  if (E.map[k] >= 0 && E.have[k] == g) {
    *idx = E.map[k];
    return MAV1HIP_OK;
  }
  /* another slot of this engine may hold the same picture (a key frame
   * refreshes every slot) */
  for (int j = 0; j < 8; j++)
    if (E.map[j] >= 0 && E.have[j] == g && s->slot_gen[j] == g) {
      E.map[k] = E.map[j];
      E.have[k] = g;
      *idx = E.map[k];
      return MAV1HIP_OK;
    }
  const int o = s->slot_owner[k];
  if (o < 0 || o == e || s->eng[o].map[k] < 0 || s->eng[o].have[k] != g)
    return seq_fail(s, MAV1HIP_EINVAL, "reference slot %d has no picture", k);
  E.map[k] = -1;
  const int b = free_pic(s, e, avoid, navoid);
  if (b < 0) return seq_fail(s, MAV1HIP_EINVAL, "engine %d: no free picture", e);
  /* the copy on this engine's stream, after everything the owner enqueued
   * so far (the picture is complete); the owner waits for it only before it
   * next writes the picture */
  Engine &O = s->eng[o];
  if (E.pic_pend[b]) {
    SCHK(s, hipStreamWaitEvent(E.st, E.pic_read[b], 0));
    E.pic_pend[b] = 0;
  }
  const int from = O.map[k];
  SCHK(s, hipEventRecord(O.ev, O.st));
  SCHK(s, hipStreamWaitEvent(E.st, O.ev, 0));
  SCHK(s, hipMemcpyAsync(pic(s, e, b), pic(s, o, from), s->frame_bytes, hipMemcpyDeviceToDevice, E.st));
  SCHK(s, hipEventRecord(O.pic_read[from], E.st));
  O.pic_pend[from] = 1;
  E.map[k] = b;
  E.have[k] = g;
  *idx = b;
  return MAV1HIP_OK;
}

int free_outs(const mav1hip_seq *s) {
  int n = 0;
  for (auto &o : s->outs) n += !o.busy;
  return n;
}

/* alt_code's frame state (src/encoder.c:1004-1030) and its parameters. */
int make_frame(mav1hip_seq *s, const Input &in, int isKey, int kind, int lastSlot, int arfSlot, int refresh,
               int hidden, Frame *f) {
  memset(f, 0, sizeof *f);
  f->in = in;
  if (isKey) s->key_disp = in.disp;
  f->is_key = isKey;
  f->kind = isKey ? 0 : kind;
  f->refresh = isKey ? 0xFF : refresh;
  f->hidden = hidden;
  f->nonref = s->cfg.leaf && !isKey && refresh == 0;
  f->order_hint = in.disp - s->key_disp;
  f->last_slot = isKey ? -1 : lastSlot;
  f->arf_slot = isKey ? -1 : arfSlot;
  f->last_hint = isKey ? -1 : s->slot_hint[lastSlot];
  f->arf_hint = !isKey && arfSlot >= 0 ? s->slot_hint[arfSlot] : -1;
  f->fr.is_key = isKey;
  f->fr.frame_kind = f->kind;
  f->fr.nonref = f->nonref;
  f->fr.order_hint = f->order_hint;
  f->fr.reference_select = !isKey && kind >= 2 && arfSlot >= 0;
  f->fr.gold_differs = 0;
  /* the reference chain on engine 0; pyramid middle frames and leaf frames
   * on engine 1 (src/encoder.c:1046-1052) */
  const int chain = f->refresh && !(f->kind == 3);
  f->engine = chain ? 0 : 1;
  const int q = mav1hip_frame_qindex(&s->cfg, &f->fr);
  f->npresets = mav1hip_setup_params(&s->cfg, &f->fr, q, &f->fp);
  if (f->npresets < 0) return seq_fail(s, MAV1HIP_EINVAL, "setup_params failed");
  mav1hip_filter_plan(&s->cfg, &f->fr, &f->fp, f->npresets, &f->plan);
  /* the slots' order hints change when the frame is handed over, as alt_code
   * updates them at once */
  for (int k = 0; k < 8; k++)
    if ((f->refresh >> k) & 1) s->slot_hint[k] = f->order_hint;
  /* output slot */
  f->out = -1;
  for (int i = 0; i < (int)s->outs.size(); i++)
    if (!s->outs[i].busy) {
      f->out = i;
      break;
    }
  if (f->out < 0) return seq_fail(s, MAV1HIP_EBUSY, "no free output slot");
  s->outs[f->out].busy = 1;
  s->order.push_back(f->out);
  f->coded = s->coded++;
  return MAV1HIP_OK;
}

/* The source frame on the engine: the caller's device frame, or the staged
 * host frame uploaded into the engine's ring on the upload stream. */
int bind_source(mav1hip_seq *s, Frame &f) {
  Engine &E = s->eng[f.engine];
  if (f.in.stage < 0) {
    f.src = const_cast<uint8_t *>(f.in.dev);
    f.src_ring = -1;
    return MAV1HIP_OK;
  }
  const int r = E.ring_next;
  E.ring_next = (E.ring_next + 1) % kSrcRing;
  if (E.ring_used[r]) SCHK(s, hipStreamWaitEvent(s->tx, E.ring_free[r], 0));
  uint8_t *dst = E.ring + (size_t)r * s->frame_bytes;
  SCHK(s, hipMemcpyAsync(dst, s->stage[f.in.stage], s->frame_bytes, hipMemcpyHostToDevice, s->tx));
  SCHK(s, hipEventRecord(s->stage_done[f.in.stage], s->tx));
  s->stage_busy[f.in.stage] = 2; /* reusable once stage_done has completed */
  SCHK(s, hipStreamWaitEvent(E.st, s->stage_done[f.in.stage], 0));
  E.ring_used[r] = 1;
  f.src = dst;
  f.src_ring = r;
  return MAV1HIP_OK;
}

/* The frame's source, references and parameters bound to its engine. */
int bind_frame(mav1hip_seq *s, Frame &f) {
  const int e = f.engine;
  Engine &E = s->eng[e];
  int rc;
  if ((rc = bind_source(s, f))) return rc;
  uint8_t *ref = nullptr, *gold = nullptr;
  int avoid[2] = {-1, -1};
  if (!f.is_key) {
    int li, ai = -1;
    if ((rc = slot_pic(s, e, f.last_slot, avoid, 0, &li))) return rc;
    avoid[0] = li;
    ref = pic(s, e, li);
    if (f.arf_slot >= 0 && f.fp.golden) {
      if ((rc = slot_pic(s, e, f.arf_slot, avoid, 1, &ai))) return rc;
      gold = pic(s, e, ai);
    }
  }
  CCHK(s, e, ctx_bind_ptrs(E.ctx, f.src, ref, gold));
  CCHK(s, e, mav1hip_set_params(E.ctx, &f.fp));
  return MAV1HIP_OK;
}

/* Source, references and parameters, then motion search, analysis and edge
 * types. */
int early_stages(mav1hip_seq *s, Frame &f) {
  const int e = f.engine;
  Engine &E = s->eng[e];
  int rc;
  if ((rc = bind_frame(s, f))) return rc;
  if (!f.is_key) CCHK(s, e, mav1hip_launch_motion_search(E.ctx));
  CCHK(s, e, mav1hip_launch_analyze(E.ctx));
  if (s->cfg.intra_edge_filter) CCHK(s, e, mav1hip_launch_edge_types(E.ctx));
  return MAV1HIP_OK;
}

/* The frame's outputs into the staging slot in one launch: each segment a
 * device buffer of the engine copied into the slot (and the final picture
 * also into the picture that a refreshed slot will map to), 16 bytes per
 * thread where source and destinations are 16-byte aligned. */
constexpr int kMaxSegs = 10;
struct StageSeg {
  const uint8_t *src;
  uint8_t *dst, *dst2;
  unsigned long long bytes;
};
struct StageList {
  StageSeg seg[kMaxSegs];
  unsigned long long start[kMaxSegs + 1]; /* prefix sums of the sizes rounded up to 16 */
  int n;
};

__device__ __forceinline__ bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

__global__ __launch_bounds__(256) void k_stage(const StageList L) {
  const unsigned long long total = L.start[L.n], step = (unsigned long long)gridDim.x * 256 * 16;
  for (unsigned long long o = ((unsigned long long)blockIdx.x * 256 + threadIdx.x) * 16; o < total; o += step) {
    int s = 0;
    while (o >= L.start[s + 1]) s++;
    const StageSeg &g = L.seg[s];
    const unsigned long long r = o - L.start[s];
    if (r >= g.bytes) continue;
    if (r + 16 <= g.bytes && al16(g.src + r) && al16(g.dst + r) && (!g.dst2 || al16(g.dst2 + r))) {
      const uint4 v = *(const uint4 *)(g.src + r);
      *(uint4 *)(g.dst + r) = v;
      if (g.dst2) *(uint4 *)(g.dst2 + r) = v;
    } else {
      const unsigned long long e = r + 16 < g.bytes ? r + 16 : g.bytes;
      for (unsigned long long q = r; q < e; q++) {
        g.dst[q] = g.src[q];
        if (g.dst2) g.dst2[q] = g.src[q];
      }
    }
  }
}

int stage_add(mav1hip_seq *s, StageList &L, const void *src, uint8_t *dst, uint8_t *dst2, size_t bytes) {
  if (!src || L.n >= kMaxSegs) return seq_fail(s, MAV1HIP_EINVAL, "staging: bad segment");
  L.seg[L.n] = StageSeg{(const uint8_t *)src, dst, dst2, bytes};
  L.start[L.n + 1] = L.start[L.n] + ((bytes + 15) & ~(size_t)15);
  L.n++;
  return MAV1HIP_OK;
}
int stage_buf(mav1hip_seq *s, StageList &L, Engine &E, int buf, uint8_t *dst, uint8_t *dst2, size_t bytes) {
  size_t cap = 0;
  void *src = ctx_buffer(E.ctx, buf, &cap);
  if (!src || bytes > cap) return seq_fail(s, MAV1HIP_EINVAL, "staging buffer %d", buf);
  return stage_add(s, L, src, dst, dst2, bytes);
}

int finish_frame(mav1hip_seq *s, Frame &f);

/* Reconstruction and the in-loop filters (run_frame_core,
 * src/encoder.c:611-641), the outputs staged and sent down, the refreshed
 * slots stored. */
int late_stages(mav1hip_seq *s, Frame &f) {
  const int e = f.engine;
  Engine &E = s->eng[e];
  mav1hip_ctx *c = E.ctx;
  const Mav1HipFilterPlan &pl = f.plan;
  CCHK(s, e, mav1hip_launch_recon(c));
  if (pl.filter_info) {
    CCHK(s, e, mav1hip_launch_filter_info(c));
    if (pl.deblock) CCHK(s, e, mav1hip_launch_deblock(c));
    if (pl.cdef_sets) CCHK(s, e, mav1hip_launch_cdef(c, pl.cdef_sets));
  }
  /* leafopt bit 4: CDEF skipped, every superblock codes set 0 */
  const int leafopt = (s->cfg.altref && f.nonref) ? (s->cfg.knob_leafopt >= 0 ? s->cfg.knob_leafopt : 0x3F) : 0;
  const int cdefOff = (leafopt & 4) != 0;
  if (pl.restoration) CCHK(s, e, mav1hip_launch_save_deblocked(c));
  if (pl.cdef_select)
    CCHK(s, e, mav1hip_launch_cdef_select(c, pl.cdef_sets));
  else if (cdefOff)
    CCHK(s, e, ctx_clear_choice(c));
  if (pl.restoration) {
    CCHK(s, e, mav1hip_launch_lr_select(c, s->lr_test));
    CCHK(s, e, ctx_rec_to_lrout(c));
    CCHK(s, e, mav1hip_launch_restore(c));
  }
  if (s->dump) CCHK(s, e, mav1hip_launch_decision_grid(c));
  return finish_frame(s, f);
}

/* The frame's outputs into its staging slot (the final picture also into
 * the picture its refreshed slots will map to), then down on the transfer
 * stream; the refreshed slots then map to that picture. */
int finish_frame(mav1hip_seq *s, Frame &f) {
  const int e = f.engine;
  Engine &E = s->eng[e];
  mav1hip_ctx *c = E.ctx;
  const Mav1HipFilterPlan &pl = f.plan;
  const int fin = pl.restoration ? MAV1HIP_BUF_LROUT : MAV1HIP_BUF_REC;
  OutSlot &o = s->outs[f.out];
  int rc;
  int b = -1;
  if (f.refresh) {
    b = free_pic(s, e, nullptr, 0);
    if (b < 0) return seq_fail(s, MAV1HIP_EINVAL, "engine %d: no free picture", e);
    if (E.pic_pend[b]) { /* a copy to the other engine still reads it */
      SCHK(s, hipStreamWaitEvent(E.st, E.pic_read[b], 0));
      E.pic_pend[b] = 0;
    }
  }
  StageList SL;
  SL.n = 0;
  SL.start[0] = 0;
  const size_t nb = sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)s->num_sbs;
  if ((rc = stage_buf(s, SL, E, MAV1HIP_BUF_BLOCKS, o.d + s->o_blocks, nullptr, nb))) return rc;
  if ((rc = stage_buf(s, SL, E, MAV1HIP_BUF_SBS, o.d + s->o_sbs, nullptr, sizeof(Mav1SbInfo) * (size_t)s->num_sbs)))
    return rc;
  if ((rc = stage_buf(s, SL, E, MAV1HIP_BUF_COEFS, o.d + s->o_coefs, nullptr,
                      sizeof(int32_t) * MAV1_COEFS_PER_SB * (size_t)s->num_sbs)))
    return rc;
  if ((rc = stage_buf(s, SL, E, MAV1HIP_BUF_CDEF_CHOICE, o.d + s->o_choice, nullptr, (size_t)s->num_sbs))) return rc;
  if (pl.restoration &&
      (rc = stage_buf(s, SL, E, MAV1HIP_BUF_LRUNITS, o.d + s->o_units, nullptr, sizeof(Mav1LrUnit) * (size_t)s->lr_units)))
    return rc;
  if ((rc = stage_buf(s, SL, E, fin, o.d + s->o_pic, b >= 0 ? pic(s, e, b) : nullptr, s->frame_bytes))) return rc;
  if (s->dump && (rc = stage_buf(s, SL, E, MAV1HIP_BUF_DECGRID, o.d + s->o_grid, nullptr, s->grid_bytes))) return rc;
  /* the restoration types chosen on the device, and the intra phase's error word */
  if ((rc = stage_add(s, SL, ctx_lrtype(c), o.d + s->o_lrtype, nullptr, sizeof(int) * 3))) return rc;
  if ((rc = stage_add(s, SL, ctx_intra_error(c), o.d + s->o_err, nullptr, sizeof(int)))) return rc;
  {
    const unsigned long long blocks16 = (SL.start[SL.n] / 16 + 255) / 256;
    const int grid = (int)(blocks16 < 1024 ? blocks16 : 1024);
    hipLaunchKernelGGL(k_stage, dim3(grid), dim3(256), 0, E.st, SL);
    SCHK(s, hipGetLastError());
  }
  SCHK(s, hipEventRecord(o.staged, E.st));
  SCHK(s, hipStreamWaitEvent(s->tx, o.staged, 0));
  SCHK(s, hipMemcpyAsync(o.h, o.d, s->o_end, hipMemcpyDeviceToHost, s->tx));
  SCHK(s, hipEventRecord(o.landed, s->tx));
  /* the source ring entry is free once this frame's stages have read it */
  if (f.src_ring >= 0) SCHK(s, hipEventRecord(E.ring_free[f.src_ring], E.st));
  /* refreshed slots now map to that picture */
  if (f.refresh) {
    s->gen++;
    for (int k = 0; k < 8; k++)
      if ((f.refresh >> k) & 1) {
        s->slot_gen[k] = s->gen;
        s->slot_owner[k] = e;
        E.map[k] = b;
        E.have[k] = s->gen;
      }
  }
  /* what receive reports */
  Mav1HipCodedFrame &info = o.info;
  memset(&info, 0, sizeof info);
  info.coded_index = (int32_t)f.coded;
  info.disp = f.in.disp;
  info.is_key = f.is_key;
  info.frame_kind = f.kind;
  info.nonref = f.nonref;
  info.hidden = f.hidden;
  info.order_hint = f.order_hint;
  info.refresh_frame_flags = f.refresh;
  info.last_slot = f.last_slot;
  info.arf_slot = f.arf_slot;
  info.last_hint = f.last_hint;
  info.arf_hint = f.arf_hint;
  info.engine = e;
  info.fp = f.fp;
  info.plan = f.plan;
  o.hdr7[0] = f.order_hint;
  o.hdr7[1] = f.last_hint;
  o.hdr7[2] = f.arf_hint;
  o.hdr7[3] = f.is_key ? 0 : 1;
  o.hdr7[4] = f.kind;
  o.hdr7[5] = f.nonref;
  o.hdr7[6] = 0;
  return MAV1HIP_OK;
}

int run_full(mav1hip_seq *s, Frame &f) {
  int rc = early_stages(s, f);
  return rc ? rc : late_stages(s, f);
}

int flush_pending(mav1hip_seq *s) {
  for (auto &f : s->pending) {
    int rc = run_full(s, f);
    if (rc) return rc;
  }
  s->pending.clear();
  return MAV1HIP_OK;
}

/* A frame handed over in coding order: second-engine frames wait for the
 * next ALTREF's early stages; anything else runs now. */
int code_frame(mav1hip_seq *s, const Input &in, int isKey, int kind, int lastSlot, int arfSlot, int refresh,
               int hidden) {
  Frame f;
  int rc = make_frame(s, in, isKey, kind, lastSlot, arfSlot, refresh, hidden, &f);
  if (rc) return rc;
  if (s->batch) { /* the batch codes it together with the other sequences' frames */
    s->jobs.push_back(f);
    return MAV1HIP_OK;
  }
  if (f.engine != 0) {
    s->pending.push_back(f);
    return MAV1HIP_OK;
  }
  const bool arf = !f.is_key && f.kind == 1;
  if (!arf && (rc = flush_pending(s))) return rc;
  if ((rc = early_stages(s, f))) return rc;
  if (arf && (rc = flush_pending(s))) return rc;
  return late_stages(s, f);
}

/* pyr_code (src/encoder.c:1229-1250): the frames strictly between a (slot
 * as) and b (slot bs). */
int pyr_code(mav1hip_seq *s, int a, int as, int b, int bs, int depth) {
  const int d0 = s->in_disp - (int)s->inq.size();
  int rc;
  if (b - a - 1 >= 3 && depth < 4) {
    const int m = (a + b) / 2, ms = 4 + depth;
    if ((rc = code_frame(s, s->inq[m - d0], 0, 3, as, bs, 1 << ms, 1))) return rc;
    if ((rc = pyr_code(s, a, as, m, ms, depth + 1))) return rc;
    return pyr_code(s, m, ms, b, bs, depth + 1);
  }
  for (int d = a + 1; d < b; d++)
    if ((rc = code_frame(s, s->inq[d - d0], 0, 2, as, bs, 0, 0))) return rc;
  return MAV1HIP_OK;
}

/* Frames the next mini-GOP of alt_drain would code (0: it cannot be formed
 * yet). */
int next_gop_len(const mav1hip_seq *s, int flush) {
  if (s->inq.empty()) return 0;
  const int N = s->cfg.altref;
  const int d0 = s->in_disp - (int)s->inq.size();
  const int isKey = d0 == 0 || (s->cfg.keyint > 0 && d0 % s->cfg.keyint == 0);
  int L = isKey ? 1 : N;
  if (!isKey && s->cfg.keyint > 0) L = std::min(L, (d0 / s->cfg.keyint + 1) * s->cfg.keyint - d0);
  if ((int)s->inq.size() < L) {
    if (!flush) return 0;
    L = (int)s->inq.size();
  }
  return L;
}

/* alt_drain (src/encoder.c:1252-1293). */
int drain(mav1hip_seq *s, int flush) {
  int L;
  while ((L = next_gop_len(s, flush)) > 0) {
    if (free_outs(s) < L) return MAV1HIP_EBUSY;
    const int d0 = s->in_disp - (int)s->inq.size();
    const int isKey = d0 == 0 || (s->cfg.keyint > 0 && d0 % s->cfg.keyint == 0);
    int rc;
    if (isKey || L == 1) {
      if ((rc = code_frame(s, s->inq[0], isKey, 0, s->anchor, -1, 0x01, 0))) return rc;
      s->anchor = 0;
    } else if (s->cfg.leaf) {
      const int arfSlot = s->anchor == 2 ? 3 : 2;
      if ((rc = code_frame(s, s->inq[L - 1], 0, 1, s->anchor, -1, 1 << arfSlot, 1))) return rc;
      if ((rc = pyr_code(s, d0 - 1, s->anchor, d0 + L - 1, arfSlot, 0))) return rc;
      s->anchor = arfSlot;
    } else {
      const int arfSlot = s->anchor == 2 ? 3 : 2;
      if ((rc = code_frame(s, s->inq[L - 1], 0, 1, s->anchor, -1, 1 << arfSlot, 1))) return rc;
      for (int i = 0; i < L - 1; i++)
        if ((rc = code_frame(s, s->inq[i], 0, 2, i == 0 ? s->anchor : 0, arfSlot, 0x01, 0))) return rc;
      s->anchor = arfSlot;
    }
    for (int i = 0; i < L; i++) s->inq.pop_front();
  }
  if (flush && s->inq.empty()) return flush_pending(s);
  return MAV1HIP_OK;
}

/* A host staging slot for a sent frame (waits for its previous upload). */
int take_stage(mav1hip_seq *s, int *idx) {
  for (int i = 0; i < (int)s->stage.size(); i++) {
    if (s->stage_busy[i] == 1) continue;
    if (s->stage_busy[i] == 2) {
      SCHK(s, hipEventSynchronize(s->stage_done[i]));
      s->stage_busy[i] = 0;
    }
    s->stage_busy[i] = 1;
    *idx = i;
    return MAV1HIP_OK;
  }
  return seq_fail(s, MAV1HIP_EBUSY, "no free input staging slot");
}

void destroy(mav1hip_seq *s) {
  if (!s) return;
  for (auto &E : s->eng)
    if (E.st) (void)hipStreamSynchronize(E.st);
  if (s->tx) (void)hipStreamSynchronize(s->tx);
  for (auto &E : s->eng) {
    if (E.ctx) mav1hip_destroy(E.ctx);
    (void)hipFree(E.pics);
    (void)hipFree(E.ring);
    for (auto &ev : E.ring_free)
      if (ev) (void)hipEventDestroy(ev);
    if (E.ev) (void)hipEventDestroy(E.ev);
    for (auto &ev : E.pic_read)
      if (ev) (void)hipEventDestroy(ev);
  }
  for (auto &o : s->outs) {
    (void)hipFree(o.d);
    if (o.h) (void)hipHostFree(o.h);
    if (o.staged) (void)hipEventDestroy(o.staged);
    if (o.landed) (void)hipEventDestroy(o.landed);
  }
  for (auto *p : s->stage)
    if (p) (void)hipHostFree(p);
  for (auto ev : s->stage_done)
    if (ev) (void)hipEventDestroy(ev);
  delete s;
}

/* The driver's four streams per device (engine 0 main and side, engine 1
 * main and side), created once per process and shared by every sequence on
 * the device: the GPU runs four hardware queues per process and assigns a
 * stream its queue when the stream is created, so streams created later
 * could share a queue with the ALTREF engine's. Sequences on one device
 * therefore run one after another on these streams. */
std::mutex g_streams_mu;
hipStream_t g_streams[kMaxDevices][2 * kEngines];

int device_streams(mav1hip_seq *s, hipStream_t **out) {
  int dev = 0;
  SCHK(s, hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return seq_fail(s, MAV1HIP_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_streams_mu);
  for (int i = 0; i < 2 * kEngines; i++)
    if (!g_streams[dev][i]) SCHK(s, hipStreamCreate(&g_streams[dev][i]));
  *out = g_streams[dev];
  return MAV1HIP_OK;
}

int setup(mav1hip_seq *s) {
  hipStream_t *st = nullptr;
  int rc = device_streams(s, &st);
  if (rc) return rc;
  for (int e = 0; e < kEngines; e++) {
    Engine &E = s->eng[e];
    char err[512] = {0};
    E.ctx = ctx_create_on_streams(&s->cfg, st[2 * e], st[2 * e + 1], err, sizeof err);
    if (!E.ctx) return seq_fail(s, MAV1HIP_EDEVICE, "engine %d: %s", e, err);
    E.st = ctx_stream(E.ctx);
    SCHK(s, hipMalloc((void **)&E.pics, s->frame_bytes * kPics));
    SCHK(s, hipMemset(E.pics, 0, s->frame_bytes * kPics));
    SCHK(s, hipMalloc((void **)&E.ring, s->frame_bytes * kSrcRing));
    for (auto &ev : E.ring_free) SCHK(s, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    SCHK(s, hipEventCreateWithFlags(&E.ev, hipEventDisableTiming));
    for (auto &ev : E.pic_read) SCHK(s, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    for (int k = 0; k < 8; k++) {
      E.map[k] = -1;
      E.have[k] = 0;
    }
  }
  s->tx = ctx_side_stream(s->eng[1].ctx);
  for (int k = 0; k < 8; k++) {
    s->slot_owner[k] = -1;
    s->slot_hint[k] = 0;
  }
  /* output staging: blocks | sbs | coefs | choice | units | lrtype | err | picture | grid */
  size_t o = 0;
  s->o_blocks = o;
  o += al(sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)s->num_sbs);
  s->o_sbs = o;
  o += al(sizeof(Mav1SbInfo) * (size_t)s->num_sbs);
  s->o_coefs = o;
  o += al(sizeof(int32_t) * MAV1_COEFS_PER_SB * (size_t)s->num_sbs);
  s->o_choice = o;
  o += al((size_t)s->num_sbs);
  s->o_units = o;
  o += al(sizeof(Mav1LrUnit) * (size_t)s->lr_units);
  s->o_lrtype = o;
  o += al(sizeof(int) * 3);
  s->o_err = o;
  o += al(sizeof(int));
  s->o_pic = o;
  o += al(s->frame_bytes);
  /* the decision record's 44-byte header goes in front of the cells */
  s->o_grid = o + 64;
  o += al(64 + s->grid_bytes);
  s->o_end = o;
  const int nout = s->cfg.altref + 4;
  s->outs.resize(nout);
  for (auto &os : s->outs) {
    SCHK(s, hipMalloc((void **)&os.d, s->o_end));
    SCHK(s, hipMemset(os.d, 0, s->o_end));
    SCHK(s, hipHostMalloc((void **)&os.h, s->o_end, hipHostMallocDefault));
    SCHK(s, hipEventCreateWithFlags(&os.staged, hipEventDisableTiming));
    SCHK(s, hipEventCreateWithFlags(&os.landed, hipEventDisableTiming));
  }
  /* input staging: the display-order queue plus the second engine's pending frames */
  const int nstage = 2 * s->cfg.altref + 4;
  s->stage.assign(nstage, nullptr);
  s->stage_done.assign(nstage, nullptr);
  s->stage_busy.assign(nstage, 0);
  for (int i = 0; i < nstage; i++) {
    SCHK(s, hipHostMalloc((void **)&s->stage[i], s->frame_bytes, hipHostMallocDefault));
    SCHK(s, hipEventCreateWithFlags(&s->stage_done[i], hipEventDisableTiming));
  }
  return MAV1HIP_OK;
}

/* ---- batches of sequences ----
 * Each sequence keeps its own driver state, engines, slot pictures and
 * outputs; its frames queue as jobs. A round takes the next job of every
 * sequence (the same coding position when the streams are fed in step) and
 * runs each stage as one launch over all of them (batch.h), on the streams
 * of the jobs' engine, which every sequence on the device shares. Rounds of
 * the second engine wait for the next ALTREF round's early stages, as single
 * frames do. */

int bfail(mav1hip_batch *b, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b->err, sizeof b->err, fmt, ap);
  va_end(ap);
  return code;
}

#define BCHK(b, call)                                                                                       \
  do {                                                                                                      \
    hipError_t e_ = (call);                                                                                 \
    if (e_ != hipSuccess)                                                                                   \
      return bfail(b, MAV1HIP_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

/* A sequence call's status, with the sequence's message. */
int from_seq(mav1hip_batch *b, int i, int rc) {
  if (rc != MAV1HIP_OK && rc != 1) snprintf(b->err, sizeof b->err, "stream %d: %s", i, b->seqs[i]->err);
  return rc;
}

/* What decides a frame's stages: frames batched together agree on it. */
struct Sig {
  int v[16];
};
Sig sig_of(const mav1hip_seq *s, const Frame &f) {
  Sig g;
  memset(&g, 0, sizeof g);
  const int leafopt = (s->cfg.altref && f.nonref) ? (s->cfg.knob_leafopt >= 0 ? s->cfg.knob_leafopt : 0x3F) : 0;
  const int v[] = {f.engine,           f.is_key,          f.kind,
                   f.fp.golden,        f.fp.turbo,        f.plan.filter_info,
                   f.plan.deblock,     f.plan.cdef_sets,  f.plan.cdef_select,
                   f.plan.restoration, (leafopt & 4) != 0, s->cfg.intra_edge_filter,
                   f.fp.frame_is_intra};
  memcpy(g.v, v, sizeof v);
  return g;
}
bool same_sig(const Sig &a, const Sig &b) { return memcmp(&a, &b, sizeof a) == 0; }

bool seq_finished(const mav1hip_seq *s) { return s->ended && s->inq.empty() && s->jobs.empty(); }

int take_table(mav1hip_batch *b, int *t) {
  const int i = b->tab_next;
  b->tab_next = (i + 1) % kTabs;
  if (b->tab_used[i]) {
    BCHK(b, hipEventSynchronize(b->tab_free[i]));
    b->tab_used[i] = 0;
  }
  *t = i;
  return MAV1HIP_OK;
}

#define LCHK(b, what, call)                                                                         \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return bfail(b, MAV1HIP_EDEVICE, "batched %s: %s", what, hipGetErrorString(e_)); \
  } while (0)

/* Sources, references and parameters of every frame, the table, then motion
 * search, analysis and edge types as batched launches. */
int round_early(mav1hip_batch *b, Round &R) {
  const int n = (int)R.frames.size(), e = R.engine;
  for (int k = 0; k < n; k++) {
    int rc = bind_frame(R.seqs[k], R.frames[k]);
    if (rc) return bfail(b, rc, "%s", R.seqs[k]->err);
  }
  int rc = take_table(b, &R.tab);
  if (rc) return rc;
  BatchFrame *T = b->d_tab + (size_t)R.tab * kMaxBatch, *H = b->h_tab + (size_t)R.tab * kMaxBatch;
  const bool prune = b->pruner[e] && b->prune_tau > 0.f;
  for (int k = 0, off = 0; k < n; k++) {
    if (ctx_batch_frame(R.seqs[k]->eng[e].ctx, &H[k], R.frames[k].plan.restoration))
      return bfail(b, MAV1HIP_EINVAL, "batch: frame %d has no parameters", k);
    H[k].prune = prune ? b->d_prune[e] + off : nullptr;
    off += H[k].fp.sb_cols * H[k].fp.sb_rows;
  }
  mav1hip_seq *s0 = R.seqs[0];
  hipStream_t st = s0->eng[e].st, side = ctx_side_stream(s0->eng[e].ctx);
  BCHK(b, hipMemcpyAsync(T, H, sizeof(BatchFrame) * n, hipMemcpyHostToDevice, st));
  const Frame &f0 = R.frames[0];
  const bool me = !f0.is_key;
  if (prune) {
    /* the network's levels before the analysis: on the side stream beside
     * motion search (after the table and the frames it reads are in place),
     * or in line on a key frame round */
    hipStream_t pst = me ? side : st;
    if (me) {
      BCHK(b, hipEventRecord(b->tab_up[e], st));
      BCHK(b, hipStreamWaitEvent(side, b->tab_up[e], 0));
    }
    const int prc = pruner_infer_frames(b->pruner[e], pst, T, H, n, b->prune_tau);
    if (prc) return bfail(b, prc, "pruner: %s", mav1hip_pruner_error(b->pruner[e]));
  }
  if (me) {
    const int turbo = f0.fp.turbo, nref = f0.fp.golden ? 2 : 1;
    const bool hp = (turbo & MAV1_T_ME_HPEL) != 0, dec = (turbo & MAV1_T_ME_DEC) != 0;
    if (!hp && dec) return bfail(b, MAV1HIP_EUNSUP, "motion search: turbo bit 32 without bit 16");
    if (hp) {
      for (int r = 0; r < nref; r++) {
        LCHK(b, "qpel_build", launch_qpel_build_b(st, T, H, n, r));
        if (dec) LCHK(b, "dec_build", launch_dec_build_b(st, T, H, n, r));
      }
      if (dec) LCHK(b, "dec_build", launch_dec_build_b(st, T, H, n, 2));
    }
    for (int level = 0; level < 4; level++) {
      LCHK(b, "me_level", launch_me_level_b(st, level, T, H, n, nref));
      BCHK(b, hipEventRecord(b->me_done[e][level], st));
    }
  }
  /* analysis level k after motion-search level k, beside level k + 1 */
  hipStream_t ast = me ? side : st;
  for (int level = 0; level < 4; level++) {
    if (me) BCHK(b, hipStreamWaitEvent(side, b->me_done[e][level], 0));
    LCHK(b, "analyze", launch_analyze_b(ast, level, T, H, n));
  }
  LCHK(b, "analyze_table", launch_analyze_table_b(ast, T, H, n));
  if (me) {
    BCHK(b, hipEventRecord(b->ana_done[e], side));
    BCHK(b, hipStreamWaitEvent(st, b->ana_done[e], 0));
  }
  if (s0->cfg.intra_edge_filter) LCHK(b, "edge_types", launch_edge_types_b(st, T, H, n));
  return MAV1HIP_OK;
}

/* Reconstruction and the in-loop filters as batched launches, then every
 * frame's outputs staged and sent down. */
int round_late(mav1hip_batch *b, Round &R) {
  const int n = (int)R.frames.size(), e = R.engine;
  mav1hip_seq *s0 = R.seqs[0];
  hipStream_t st = s0->eng[e].st;
  const BatchFrame *T = b->d_tab + (size_t)R.tab * kMaxBatch, *H = b->h_tab + (size_t)R.tab * kMaxBatch;
  /* the scheduling pass and the ready queue in line before the inter phase
   * (they read only the analysis' records): beside it they would wait for
   * CUs, since the persistent inter phase holds every CU */
  LCHK(b, "intra_sched", launch_intra_sched_b(st, T, H, n, b->d_iq[e], b->iq_cap));
  const int interWg = b->inter_wg > 0 ? b->inter_wg : 2 * b->num_cu;
  LCHK(b, "recon_inter", launch_recon_inter_b(st, T, H, n, interWg));
  /* an inter frame's intra superblocks form a few short chains: a quarter
   * of the CUs serve them and leave the rest to the other engine's round
   * and the transfers; intra frames use every CU */
  const bool intraFrame = R.frames[0].fp.frame_is_intra != 0;
  int intraWg = intraFrame ? b->intra_wg_key : b->intra_wg_inter;
  if (intraWg <= 0) intraWg = intraFrame ? b->num_cu : std::max(1, b->num_cu / 4);
  LCHK(b, "recon_intra", launch_recon_intra_b(st, T, H, n, b->d_iq[e], intraWg));
  const Frame &f0 = R.frames[0];
  const Mav1HipFilterPlan &pl = f0.plan;
  if (pl.filter_info) {
    LCHK(b, "filter_info", launch_filter_info_b(st, T, H, n));
    if (pl.deblock) LCHK(b, "deblock", launch_deblock_b(st, T, H, n));
    if (pl.cdef_sets) LCHK(b, "cdef", launch_cdef_b(st, T, H, n, pl.cdef_sets));
  }
  const int leafopt = (s0->cfg.altref && f0.nonref) ? (s0->cfg.knob_leafopt >= 0 ? s0->cfg.knob_leafopt : 0x3F) : 0;
  const int cdefOff = (leafopt & 4) != 0;
  if (pl.restoration) LCHK(b, "deblocked copy", launch_frame_op_b(st, T, H, n, 0));
  if (pl.cdef_select)
    LCHK(b, "cdef_select", launch_cdef_select_b(st, T, H, n, pl.cdef_sets));
  else if (cdefOff)
    LCHK(b, "cdef choice clear", launch_frame_op_b(st, T, H, n, 2));
  if (pl.restoration) {
    /* the fit, then chain launches with the refinements between them, as
     * the single-frame restoration search */
    const int LR_ROUNDS = 3;
    LCHK(b, "lr_fit", launch_lr_fit_b(st, T, H, n));
    for (int r = 0; r < LR_ROUNDS; r++) {
      if (r > 0) LCHK(b, "lr_path", launch_lr_path_b(st, T, H, n));
      LCHK(b, "lr_chain", launch_lr_chain_b(st, T, H, n, b->lr_test != 0, r < LR_ROUNDS - 1));
    }
    LCHK(b, "restoration copy", launch_frame_op_b(st, T, H, n, 1));
    LCHK(b, "lr_apply", launch_lr_apply_b(st, T, H, n));
  }
  if (b->dump) LCHK(b, "decision_grid", launch_decision_grid_b(st, T, H, n));
  BCHK(b, hipEventRecord(b->tab_free[R.tab], st));
  b->tab_used[R.tab] = 1;
  for (int k = 0; k < n; k++) {
    int rc = finish_frame(R.seqs[k], R.frames[k]);
    if (rc) return bfail(b, rc, "%s", R.seqs[k]->err);
  }
  return MAV1HIP_OK;
}

int flush_rounds(mav1hip_batch *b) {
  for (auto &R : b->pending) {
    int rc = round_early(b, R);
    if (!rc) rc = round_late(b, R);
    if (rc) return rc;
  }
  b->pending.clear();
  return MAV1HIP_OK;
}

/* code_frame for a round */
int exec_round(mav1hip_batch *b, Round &R) {
  if (R.engine != 0) {
    b->pending.push_back(R);
    return MAV1HIP_OK;
  }
  const Frame &f0 = R.frames[0];
  const bool arf = !f0.is_key && f0.kind == 1;
  int rc;
  if (!arf && (rc = flush_rounds(b))) return rc;
  if ((rc = round_early(b, R))) return rc;
  if (arf && (rc = flush_rounds(b))) return rc;
  return round_late(b, R);
}

/* Runs the rounds that can be formed: every sequence that can still code
 * a frame has a job (all: also rounds of the sequences that have one). */
int pump(mav1hip_batch *b, int all) {
  for (;;) {
    int first = -1;
    for (int i = 0; i < (int)b->seqs.size() && first < 0; i++)
      if (!b->seqs[i]->jobs.empty()) first = i;
    if (first < 0) break;
    if (!all)
      for (auto *s : b->seqs)
        if (s->jobs.empty() && !seq_finished(s)) return MAV1HIP_OK;
    Round R;
    R.engine = b->seqs[first]->jobs.front().engine;
    const Sig g0 = sig_of(b->seqs[first], b->seqs[first]->jobs.front());
    for (auto *s : b->seqs)
      if (!s->jobs.empty() && same_sig(sig_of(s, s->jobs.front()), g0) && (int)R.frames.size() < kMaxBatch) {
        R.seqs.push_back(s);
        R.frames.push_back(s->jobs.front());
        s->jobs.pop_front();
      }
    int rc = exec_round(b, R);
    if (rc) return rc;
  }
  bool done = true;
  for (auto *s : b->seqs) done = done && seq_finished(s);
  return done ? flush_rounds(b) : MAV1HIP_OK;
}

void batch_destroy(mav1hip_batch *b) {
  if (!b) return;
  for (auto *s : b->seqs) destroy(s);
  for (auto &ev : b->tab_free)
    if (ev) (void)hipEventDestroy(ev);
  for (int e = 0; e < 2; e++) {
    for (auto &ev : b->me_done[e])
      if (ev) (void)hipEventDestroy(ev);
    if (b->ana_done[e]) (void)hipEventDestroy(b->ana_done[e]);
  }
  (void)hipFree(b->d_tab);
  for (auto *q : b->d_iq) (void)hipFree(q);
  for (int e = 0; e < 2; e++) {
    if (b->pruner[e]) mav1hip_pruner_destroy(b->pruner[e]);
    (void)hipFree(b->d_prune[e]);
    if (b->tab_up[e]) (void)hipEventDestroy(b->tab_up[e]);
  }
  if (b->h_tab) (void)hipHostFree(b->h_tab);
  delete b;
}

int batch_setup(mav1hip_batch *b) {
  BCHK(b, hipMalloc((void **)&b->d_tab, sizeof(BatchFrame) * kTabs * kMaxBatch));
  BCHK(b, hipHostMalloc((void **)&b->h_tab, sizeof(BatchFrame) * kTabs * kMaxBatch, hipHostMallocDefault));
  for (auto &ev : b->tab_free) BCHK(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  for (int e = 0; e < 2; e++) {
    for (auto &ev : b->me_done[e]) BCHK(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    BCHK(b, hipEventCreateWithFlags(&b->ana_done[e], hipEventDisableTiming));
  }
  b->num_cu = ctx_num_cu(b->seqs[0]->eng[0].ctx);
  for (auto *s : b->seqs) b->iq_cap += s->num_sbs;
  for (auto &q : b->d_iq) BCHK(b, hipMalloc((void **)&q, sizeof(int) * (4 + (size_t)b->iq_cap)));
  return MAV1HIP_OK;
}

}  // namespace

extern "C" {

mav1hip_seq *mav1hip_seq_create(const Mav1HipConfig *cfg, int32_t lr_test, int32_t dump_decisions, char *err,
                                size_t errlen) {
  if (!cfg) {
    if (err && errlen) snprintf(err, errlen, "no configuration");
    return nullptr;
  }
  mav1hip_seq *s = new mav1hip_seq();
  s->cfg = *cfg;
  mav1hip_config_normalize(&s->cfg);
  if (s->cfg.altref < 2) {
    if (err && errlen) snprintf(err, errlen, "the sequence driver runs ALTREF configurations (altref >= 2)");
    delete s;
    return nullptr;
  }
  s->lr_test = lr_test;
  s->dump = dump_decisions != 0;
  int32_t mc, mr, sc, sr, tw;
  s->frame_bytes = mav1hip_geometry(&s->cfg, &mc, &mr, &sc, &sr, &tw, s->stride, s->offset, s->plane_h);
  if (!s->frame_bytes) {
    if (err && errlen) snprintf(err, errlen, "unsupported frame size %dx%d", cfg->width, cfg->height);
    delete s;
    return nullptr;
  }
  s->num_sbs = sc * sr;
  s->grid_bytes = (size_t)((mr + 1) >> 1) * ((mc + 1) >> 1) * 8;
  {
    Mav1HipFrame fr = {};
    fr.is_key = 1;
    Mav1FrameParams fp;
    mav1hip_setup_params(&s->cfg, &fr, s->cfg.qindex, &fp);
    s->lr_units = fp.lr_unit_off[2] + fp.lr_unit_rows[2] * fp.lr_unit_cols[2];
  }
  if (setup(s) != MAV1HIP_OK) {
    if (err && errlen) snprintf(err, errlen, "%s", s->err);
    destroy(s);
    return nullptr;
  }
  return s;
}

void mav1hip_seq_destroy(mav1hip_seq *s) { destroy(s); }

const char *mav1hip_seq_error(const mav1hip_seq *s) { return s ? s->err : "no sequence"; }

mav1hip_ctx *mav1hip_seq_engine(mav1hip_seq *s, int32_t i) {
  return s && i >= 0 && i < kEngines ? s->eng[i].ctx : nullptr;
}

static int send_input(mav1hip_seq *s, const Input &in) {
  s->inq.push_back(in);
  s->in_disp++;
  return drain(s, 0);
}

int mav1hip_seq_send_frame(mav1hip_seq *s, const uint8_t *const planes[3], const int32_t strides[3]) {
  if (!s) return MAV1HIP_EINVAL;
  if (s->ended) return seq_fail(s, MAV1HIP_EINVAL, "send_frame: the stream has ended");
  if (!planes) {
    if (free_outs(s) < (int)s->inq.size()) return seq_fail(s, MAV1HIP_EBUSY, "end of stream: receive first");
    s->ended = 1;
    return drain(s, 1);
  }
  if (!strides || !planes[0] || !planes[1] || !planes[2]) return seq_fail(s, MAV1HIP_EINVAL, "send_frame: missing plane");
  /* the frames a drain could code after this one need their output slots */
  if (free_outs(s) < (int)s->inq.size() + 1) return seq_fail(s, MAV1HIP_EBUSY, "send_frame: receive first");
  int st;
  int rc = take_stage(s, &st);
  if (rc) return rc;
  /* load_source (src/encoder.c:297-310): edges replicated to the padded size */
  uint8_t *dst0 = s->stage[st];
  for (int p = 0; p < 3; p++) {
    const int pw = p ? (s->cfg.width + 1) >> 1 : s->cfg.width;
    const int ph = p ? (s->cfg.height + 1) >> 1 : s->cfg.height;
    const int aw = s->stride[p], ah = s->plane_h[p];
    uint8_t *dst = dst0 + s->offset[p];
    for (int y = 0; y < ah; y++) {
      const uint8_t *src = planes[p] + (size_t)(y < ph ? y : ph - 1) * strides[p];
      uint8_t *d = dst + (size_t)y * aw;
      memcpy(d, src, pw);
      memset(d + pw, src[pw - 1], aw - pw);
    }
  }
  return send_input(s, Input{s->in_disp, st, nullptr});
}

int mav1hip_seq_send_device_frame(mav1hip_seq *s, const uint8_t *frame) {
  if (!s || !frame) return MAV1HIP_EINVAL;
  if (s->ended) return seq_fail(s, MAV1HIP_EINVAL, "send_device_frame: the stream has ended");
  if (free_outs(s) < (int)s->inq.size() + 1) return seq_fail(s, MAV1HIP_EBUSY, "send_device_frame: receive first");
  return send_input(s, Input{s->in_disp, -1, frame});
}

int mav1hip_seq_receive(mav1hip_seq *s, Mav1HipCodedFrame *out, int32_t wait) {
  if (!s || !out) return MAV1HIP_EINVAL;
  if (s->held >= 0) {
    s->outs[s->held].busy = 0;
    s->held = -1;
  }
  /* frames still waiting for the next ALTREF run when the caller waits for them */
  if (s->order.empty()) return 1;
  const int i = s->order.front();
  OutSlot &o = s->outs[i];
  if (wait && !s->pending.empty()) {
    bool queued = false;
    for (auto &f : s->pending) queued |= f.out == i;
    if (queued) {
      int rc = flush_pending(s);
      if (rc) return rc;
    }
  }
  for (auto &f : s->pending)
    if (f.out == i) return 1; /* not enqueued yet */
  if (wait) {
    SCHK(s, hipEventSynchronize(o.landed));
  } else {
    hipError_t q = hipEventQuery(o.landed);
    if (q == hipErrorNotReady) return 1;
    if (q != hipSuccess) return seq_fail(s, MAV1HIP_EDEVICE, "receive: %s", hipGetErrorString(q));
  }
  s->order.pop_front();
  s->held = i;
  int ierr = 0;
  memcpy(&ierr, o.h + s->o_err, sizeof ierr);
  if (ierr) return seq_fail(s, MAV1HIP_EDEVICE, "frame %d: an intra-phase dependency wait ran out", o.info.disp);
  *out = o.info;
  if (out->plan.restoration) memcpy(out->fp.lr_type, o.h + s->o_lrtype, sizeof(int) * 3);
  out->blocks = (const Mav1BlockInfo *)(o.h + s->o_blocks);
  out->sbs = (const Mav1SbInfo *)(o.h + s->o_sbs);
  out->coefs = (const int32_t *)(o.h + s->o_coefs);
  out->cdef_choice = (const int8_t *)(o.h + s->o_choice);
  out->lr_units = out->plan.restoration ? (const Mav1LrUnit *)(o.h + s->o_units) : nullptr;
  out->recon = o.h + s->o_pic;
  if (s->dump) {
    /* the record's header (src/encoder.c:659-704) in front of the cells */
    uint8_t *g = o.h + s->o_grid - 44;
    const int32_t hdr[11] = {0x4345444D, o.hdr7[0], o.hdr7[1], o.hdr7[2], o.hdr7[3], o.hdr7[4], o.hdr7[5],
                             out->fp.base_q_idx, (out->fp.mi_rows + 1) >> 1, (out->fp.mi_cols + 1) >> 1, o.hdr7[6]};
    memcpy(g, hdr, sizeof hdr);
    out->decision = g;
    out->decision_bytes = 44 + s->grid_bytes;
  }
  return 0;
}

mav1hip_batch *mav1hip_batch_create(const Mav1HipConfig *cfgs, int32_t n, int32_t lr_test, int32_t dump_decisions,
                                    char *err, size_t errlen) {
  if (!cfgs || n < 1 || n > kMaxBatch) {
    if (err && errlen) snprintf(err, errlen, "a batch holds 1 to %d streams", kMaxBatch);
    return nullptr;
  }
  mav1hip_batch *b = new mav1hip_batch();
  b->lr_test = lr_test;
  b->dump = dump_decisions != 0;
  for (int i = 0; i < n; i++) {
    char e2[512] = {0};
    mav1hip_seq *s = mav1hip_seq_create(&cfgs[i], lr_test, dump_decisions, e2, sizeof e2);
    if (!s) {
      if (err && errlen) snprintf(err, errlen, "stream %d: %s", i, e2);
      batch_destroy(b);
      return nullptr;
    }
    s->batch = b;
    b->seqs.push_back(s);
  }
  if (batch_setup(b) != MAV1HIP_OK) {
    if (err && errlen) snprintf(err, errlen, "%s", b->err);
    batch_destroy(b);
    return nullptr;
  }
  return b;
}

void mav1hip_batch_destroy(mav1hip_batch *b) { batch_destroy(b); }

const char *mav1hip_batch_error(const mav1hip_batch *b) { return b ? b->err : "no batch"; }

int mav1hip_batch_send_frame(mav1hip_batch *b, int32_t stream, const uint8_t *const planes[3],
                             const int32_t strides[3]) {
  if (!b || stream < 0 || stream >= (int)b->seqs.size()) return MAV1HIP_EINVAL;
  int rc = from_seq(b, stream, mav1hip_seq_send_frame(b->seqs[stream], planes, strides));
  if (rc) return rc;
  return pump(b, 0);
}

int mav1hip_batch_send_device_frame(mav1hip_batch *b, int32_t stream, const uint8_t *frame) {
  if (!b || stream < 0 || stream >= (int)b->seqs.size()) return MAV1HIP_EINVAL;
  int rc = from_seq(b, stream, mav1hip_seq_send_device_frame(b->seqs[stream], frame));
  if (rc) return rc;
  return pump(b, 0);
}

int mav1hip_batch_receive(mav1hip_batch *b, int32_t stream, Mav1HipCodedFrame *out, int32_t wait) {
  if (!b || !out || stream < 0 || stream >= (int)b->seqs.size()) return MAV1HIP_EINVAL;
  mav1hip_seq *s = b->seqs[stream];
  if (!s->order.empty()) {
    /* the next frame may still be a job or in a waiting round: with wait,
     * run what can run now, in rounds of the sequences that have a job */
    const int i0 = s->order.front();
    bool queued = false;
    for (auto &f : s->jobs) queued |= f.out == i0;
    for (auto &R : b->pending)
      for (size_t k = 0; k < R.seqs.size(); k++) queued |= R.seqs[k] == s && R.frames[k].out == i0;
    if (queued) {
      if (!wait) return 1;
      int rc = pump(b, 1);
      if (!rc) rc = flush_rounds(b);
      if (rc) return rc;
    }
  }
  return from_seq(b, stream, mav1hip_seq_receive(s, out, wait));
}

int mav1hip_batch_set_knob(mav1hip_batch *b, const char *name, int32_t value) {
  if (!b || !name || value < 0 || value > 65536) return MAV1HIP_EINVAL;
  if (!strcmp(name, "intra_wg_key"))
    b->intra_wg_key = value;
  else if (!strcmp(name, "intra_wg_inter"))
    b->intra_wg_inter = value;
  else if (!strcmp(name, "inter_wg"))
    b->inter_wg = value;
  else
    return bfail(b, MAV1HIP_EINVAL, "unknown knob %s", name);
  return MAV1HIP_OK;
}

int mav1hip_batch_set_pruner(mav1hip_batch *b, const float *params, float tau) {
  if (!b || tau >= 1.f) return MAV1HIP_EINVAL;
  if (!params || tau <= 0.f) {
    b->prune_tau = 0.f;
    return MAV1HIP_OK;
  }
  for (int e = 0; e < 2; e++) {
    if (!b->pruner[e]) {
      char err[256] = {0};
      b->pruner[e] = mav1hip_pruner_create(b->iq_cap, err, sizeof err);
      if (!b->pruner[e]) return bfail(b, MAV1HIP_EDEVICE, "pruner: %s", err);
      BCHK(b, hipMalloc((void **)&b->d_prune[e], (size_t)b->iq_cap));
      BCHK(b, hipEventCreateWithFlags(&b->tab_up[e], hipEventDisableTiming));
    }
    const int rc = mav1hip_pruner_set_params(b->pruner[e], params);
    if (rc) return bfail(b, rc, "pruner: %s", mav1hip_pruner_error(b->pruner[e]));
  }
  b->prune_tau = tau;
  return MAV1HIP_OK;
}

mav1hip_seq *mav1hip_batch_stream(mav1hip_batch *b, int32_t stream) {
  return b && stream >= 0 && stream < (int)b->seqs.size() ? b->seqs[stream] : nullptr;
}

}  // extern "C"
