/* C-ABI of the HIP analysis backend (include/mav1_hip.h): context, device
 * buffers, stage launches on the context's stream, host-buffer stage calls
 * and per-kernel event timing. */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/mav1_hip.h"
#include "internal.h"

namespace {

struct TimedLaunch {
  const char *name;
  hipEvent_t start, stop;
};

}  // namespace

struct mav1hip_ctx {
  Mav1HipConfig cfg;
  int mi_cols, mi_rows, sb_cols, sb_rows, num_sbs, tile_w_sb;
  int stride[3], offset[3], plane_h[3];
  size_t frame_bytes;
  int device;
  int own_streams; /* the streams were created for the context (else the caller keeps them) */
  hipStream_t stream;
  /* Analysis runs on `side` when motion search of the same frame was the
   * last work enqueued: analysis level k waits for me_done[k] only, so it
   * overlaps motion-search level k + 1; `stream` waits for ana_done before
   * it goes on. me_pending: me_done holds this frame's levels and nothing
   * was enqueued on `stream` since (every other launch, copy or parameter
   * change clears it). */
  hipStream_t side;
  hipEvent_t me_done[4], ana_done;
  int me_pending;
  /* cross-context ordering of mav1hip_pool_import */
  hipEvent_t xfer;
  /* the intra phase's scheduling pass on `side` beside the inter phase */
  hipEvent_t rec_go, sched_done;

  char err[512];
  char devname[256];
  Mav1FrameParams fp;
  int fp_set;
  /* own device buffers */
  uint8_t *d_src, *d_ref, *d_gold;
  int32_t *d_mvfield, *d_mvgold;
  uint8_t *d_planes;
  long long *d_acosts;
  Mav1BlockInfo *d_ainfos, *d_blocks;
  Mav1SbInfo *d_sbs;
  uint8_t *d_rec;
  int32_t *d_coefs, *d_mats;
  uint8_t *d_ys, *d_uvs;
  uint8_t *d_lfmask, *d_cdefbuf, *d_dbk, *d_lrout;
  int *d_dirvar;
  Mav1LrUnit *d_lrunits;
  int8_t *d_choice;
  void *d_lrstats;
  void *d_lrchain;
  int *d_lrtype, *d_lrreruns;
  uint8_t *d_decgrid;
  int *d_rwork;
  int *d_iwork; /* intra phase schedule and completion flags */
  int num_cu;
  size_t decgrid_bytes;
  int lr_units;
  int lr_gen; /* generation of the last restoration fit launch */
  long entry_bytes;
  size_t planes_bytes;
  /* frames the stages read (own buffers or pool frames) */
  uint8_t *b_src, *b_ref, *b_gold;
  uint8_t *d_pool;
  int pool_n;
  /* timing */
  int timing;
  std::string timing_only; /* empty: every kernel */
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<int, double>> stats;
};

static int fail(mav1hip_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIPCHK(c, call)                                                                                   \
  do {                                                                                                    \
    hipError_t e_ = (call);                                                                               \
    if (e_ != hipSuccess) return fail(c, MAV1HIP_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                                \
  } while (0)

static hipEvent_t take_event(mav1hip_ctx *c) {
  if (!c->free_events.empty()) {
    hipEvent_t e = c->free_events.back();
    c->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

/* Records the launch of `name` on stream st between two events when timing
 * is on. */
template <class F>
static int timed_on(mav1hip_ctx *c, hipStream_t st, const char *name, F &&launch) {
  hipEvent_t a = nullptr, b = nullptr;
  if (st == c->stream) c->me_pending = 0; /* analysis may overlap only what directly precedes it: motion search */
  const bool on = c->timing && (c->timing_only.empty() || c->timing_only == name);
  if (on) {
    a = take_event(c);
    b = take_event(c);
    if (a && b) HIPCHK(c, hipEventRecord(a, st));
  }
  hipError_t e = launch();
  if (e != hipSuccess) return fail(c, MAV1HIP_EDEVICE, "launch of %s: %s", name, hipGetErrorString(e));
  if (on && a && b) {
    HIPCHK(c, hipEventRecord(b, st));
    c->pending.push_back({name, a, b});
  }
  return MAV1HIP_OK;
}
template <class F>
static int timed(mav1hip_ctx *c, const char *name, F &&launch) {
  return timed_on(c, c->stream, name, launch);
}

static int collect_timing(mav1hip_ctx *c) {
  if (c->pending.empty()) return MAV1HIP_OK;
  HIPCHK(c, hipStreamSynchronize(c->side));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &t : c->pending) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, t.start, t.stop));
    auto &s = c->stats[t.name];
    s.first += 1;
    s.second += ms;
    c->free_events.push_back(t.start);
    c->free_events.push_back(t.stop);
  }
  c->pending.clear();
  return MAV1HIP_OK;
}

static int dmalloc(mav1hip_ctx *c, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 256);
  if (e != hipSuccess) return fail(c, MAV1HIP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return MAV1HIP_OK;
}

extern "C" {

static mav1hip_ctx *create_ctx(const Mav1HipConfig *cfg, hipStream_t extMain, hipStream_t extSide, char *err,
                               size_t errlen);

mav1hip_ctx *mav1hip_create(const Mav1HipConfig *cfg, char *err, size_t errlen) {
  return create_ctx(cfg, nullptr, nullptr, err, errlen);
}

}  // extern "C"

/* A context on two streams the caller keeps (the sequence driver's
 * process-wide streams): they are neither created nor destroyed here. */
mav1hip_ctx *ctx_create_on_streams(const Mav1HipConfig *cfg, hipStream_t main, hipStream_t side, char *err,
                                   size_t errlen) {
  if (!main || !side) {
    if (err && errlen) snprintf(err, errlen, "no streams");
    return nullptr;
  }
  return create_ctx(cfg, main, side, err, errlen);
}

static mav1hip_ctx *create_ctx(const Mav1HipConfig *cfg, hipStream_t extMain, hipStream_t extSide, char *err,
                               size_t errlen) {
  if (!cfg) {
    if (err && errlen) snprintf(err, errlen, "no configuration");
    return nullptr;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) {
    if (err && errlen) snprintf(err, errlen, "no HIP device available (%s)", hipGetErrorString(e));
    return nullptr;
  }
  mav1hip_ctx *c = new mav1hip_ctx();
  c->cfg = *cfg;
  mav1hip_config_normalize(&c->cfg);
  c->frame_bytes = mav1hip_geometry(&c->cfg, &c->mi_cols, &c->mi_rows, &c->sb_cols, &c->sb_rows, &c->tile_w_sb,
                                    c->stride, c->offset, c->plane_h);
  if (!c->frame_bytes) {
    if (err && errlen) snprintf(err, errlen, "unsupported frame size %dx%d", cfg->width, cfg->height);
    delete c;
    return nullptr;
  }
  c->num_sbs = c->sb_cols * c->sb_rows;
  int rc = MAV1HIP_OK;
  c->own_streams = extMain == nullptr;
  if (!c->own_streams) {
    c->stream = extMain;
    c->side = extSide;
  }
  if (hipGetDevice(&c->device) != hipSuccess ||
      (c->own_streams && (hipStreamCreate(&c->stream) != hipSuccess || hipStreamCreate(&c->side) != hipSuccess))) {
    if (err && errlen) snprintf(err, errlen, "cannot create a HIP stream");
    if (c->own_streams && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return nullptr;
  }
  hipEvent_t *evs[8] = {&c->me_done[0], &c->me_done[1], &c->me_done[2], &c->me_done[3],
                        &c->ana_done,   &c->xfer,       &c->rec_go,     &c->sched_done};
  for (int i = 0; i < 8; i++) {
    if (hipEventCreateWithFlags(evs[i], hipEventDisableTiming) != hipSuccess) {
      if (err && errlen) snprintf(err, errlen, "cannot create a HIP event");
      for (int j = 0; j < i; j++) (void)hipEventDestroy(*evs[j]);
      if (c->own_streams) {
        (void)hipStreamDestroy(c->side);
        (void)hipStreamDestroy(c->stream);
      }
      delete c;
      return nullptr;
    }
  }

  hipDeviceProp_t prop;
  c->num_cu = 256;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
    snprintf(c->devname, sizeof c->devname, "%s (%s)", prop.name, prop.gcnArchName);
    if (prop.multiProcessorCount > 0) c->num_cu = prop.multiProcessorCount;
  }
  long W = c->cfg.width, H = c->cfg.height;
  /* entries start 256-byte aligned (the plane build stores 32-bit words) */
  c->entry_bytes = (16 * W * H + ((W + 3) >> 2) * ((H + 3) >> 2) + 255) & ~255L;
  c->planes_bytes = (size_t)c->entry_bytes * 2 + (size_t)((W + 3) >> 2) * ((H + 3) >> 2);
  size_t mvb = sizeof(int32_t) * 2 * MAV1_NUM_NODES * (size_t)c->num_sbs;
  rc |= dmalloc(c, (void **)&c->d_src, c->frame_bytes);
  rc |= dmalloc(c, (void **)&c->d_ref, c->frame_bytes);
  rc |= dmalloc(c, (void **)&c->d_gold, c->frame_bytes);
  rc |= dmalloc(c, (void **)&c->d_mvfield, mvb);
  rc |= dmalloc(c, (void **)&c->d_mvgold, mvb);
  rc |= dmalloc(c, (void **)&c->d_planes, c->planes_bytes);
  rc |= dmalloc(c, (void **)&c->d_acosts, sizeof(long long) * MAV1_A_ITEMS * (size_t)c->num_sbs);
  rc |= dmalloc(c, (void **)&c->d_ainfos, sizeof(Mav1BlockInfo) * MAV1_A_ITEMS * (size_t)c->num_sbs);
  rc |= dmalloc(c, (void **)&c->d_blocks, sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs);
  rc |= dmalloc(c, (void **)&c->d_sbs, sizeof(Mav1SbInfo) * (size_t)c->num_sbs);
  rc |= dmalloc(c, (void **)&c->d_rec, c->frame_bytes);
  rc |= dmalloc(c, (void **)&c->d_coefs, sizeof(int32_t) * MAV1_COEFS_PER_SB * (size_t)c->num_sbs);
  rc |= dmalloc(c, (void **)&c->d_mats, sizeof(int32_t) * 4096 * 15);
  rc |= dmalloc(c, (void **)&c->d_ys, (size_t)c->mi_rows * c->mi_cols);
  rc |= dmalloc(c, (void **)&c->d_uvs, (size_t)c->mi_rows * c->mi_cols);
  rc |= dmalloc(c, (void **)&c->d_lfmask, (size_t)c->mi_rows * c->mi_cols);
  rc |= dmalloc(c, (void **)&c->d_cdefbuf, c->frame_bytes * MAV1HIP_MAX_CDEF_SETS);
  rc |= dmalloc(c, (void **)&c->d_dbk, c->frame_bytes);
  rc |= dmalloc(c, (void **)&c->d_lrout, c->frame_bytes);
  rc |= dmalloc(c, (void **)&c->d_dirvar, sizeof(int) * (size_t)(c->mi_rows >> 1) * (c->mi_cols >> 1));
  rc |= dmalloc(c, (void **)&c->d_choice, (size_t)c->num_sbs);
  c->decgrid_bytes = (size_t)((c->mi_rows + 1) >> 1) * ((c->mi_cols + 1) >> 1) * 8;
  rc |= dmalloc(c, (void **)&c->d_rwork, sizeof(int) * (2 + (size_t)c->num_sbs * MAV1_MAX_BLOCKS_PER_SB));
  {
    Mav1FrameParams g = {};
    g.sb_cols = c->sb_cols;
    g.sb_rows = c->sb_rows;
    if (!intra_sched_fits(g)) {
      if (err && errlen) snprintf(err, errlen, "frame of %d x %d superblocks is too large for the intra scheduler",
                                  c->sb_cols, c->sb_rows);
      mav1hip_destroy(c);
      return nullptr;
    }
    rc |= dmalloc(c, (void **)&c->d_iwork, sizeof(int) * intra_work_ints(g));
  }
  rc |= dmalloc(c, (void **)&c->d_decgrid, c->decgrid_bytes);
  {
    /* restoration units of 64x64 per plane (count_units_in_frame) */
    c->lr_units = 0;
    for (int p = 0; p < 3; p++) {
      int sub = p > 0, pw = (cfg->width + sub) >> sub, ph = (cfg->height + sub) >> sub;
      int rows = (ph + 32) / 64, cols = (pw + 32) / 64;
      c->lr_units += (rows < 1 ? 1 : rows) * (cols < 1 ? 1 : cols);
    }
    rc |= dmalloc(c, (void **)&c->d_lrunits, sizeof(Mav1LrUnit) * (size_t)c->lr_units);
    rc |= dmalloc(c, &c->d_lrstats, lr_stat_bytes() * (size_t)c->lr_units);
    /* zeroed once: the fit's generation numbers start at 1 */
    if (rc == MAV1HIP_OK && hipMemset(c->d_lrstats, 0, lr_stat_bytes() * (size_t)c->lr_units) != hipSuccess)
      rc = fail(c, MAV1HIP_EDEVICE, "hipMemset of the restoration statistics failed");
    rc |= dmalloc(c, &c->d_lrchain, lr_chain_bytes());
    rc |= dmalloc(c, (void **)&c->d_lrtype, sizeof(int) * 3);
    rc |= dmalloc(c, (void **)&c->d_lrreruns, sizeof(int) * 3);
  }
  if (rc != MAV1HIP_OK) {
    if (err && errlen) snprintf(err, errlen, "%s", c->err);
    mav1hip_destroy(c);
    return nullptr;
  }
  /* deterministic contents before the first frame (the reference's buffers
   * are zero-filled pages) */
  (void)hipMemsetAsync(c->d_planes, 0, c->planes_bytes, c->stream);
  (void)hipMemsetAsync(c->d_mvfield, 0, mvb, c->stream);
  if (c->d_iwork) (void)hipMemsetAsync(c->d_iwork, 0, sizeof(int) * 4, c->stream);
  (void)hipMemsetAsync(c->d_mvgold, 0, mvb, c->stream);
  (void)hipMemsetAsync(c->d_acosts, 0, sizeof(long long) * MAV1_A_ITEMS * (size_t)c->num_sbs, c->stream);
  (void)hipMemsetAsync(c->d_ainfos, 0, sizeof(Mav1BlockInfo) * MAV1_A_ITEMS * (size_t)c->num_sbs, c->stream);
  (void)hipMemsetAsync(c->d_blocks, 0, sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs, c->stream);
  (void)hipMemsetAsync(c->d_sbs, 0, sizeof(Mav1SbInfo) * (size_t)c->num_sbs, c->stream);
  (void)hipMemsetAsync(c->d_rec, 0, c->frame_bytes, c->stream);
  (void)hipMemsetAsync(c->d_coefs, 0, sizeof(int32_t) * MAV1_COEFS_PER_SB * (size_t)c->num_sbs, c->stream);
  (void)hipMemsetAsync(c->d_ys, 0, (size_t)c->mi_rows * c->mi_cols, c->stream);
  (void)hipMemsetAsync(c->d_uvs, 0, (size_t)c->mi_rows * c->mi_cols, c->stream);
  (void)hipMemsetAsync(c->d_lfmask, 0, (size_t)c->mi_rows * c->mi_cols, c->stream);
  (void)hipMemsetAsync(c->d_cdefbuf, 0, c->frame_bytes * MAV1HIP_MAX_CDEF_SETS, c->stream);
  (void)hipMemsetAsync(c->d_dbk, 0, c->frame_bytes, c->stream);
  (void)hipMemsetAsync(c->d_lrout, 0, c->frame_bytes, c->stream);
  (void)hipMemsetAsync(c->d_choice, 0, (size_t)c->num_sbs, c->stream);
  (void)hipMemsetAsync(c->d_lrunits, 0, sizeof(Mav1LrUnit) * (size_t)c->lr_units, c->stream);
  (void)hipMemsetAsync(c->d_lrtype, 0, sizeof(int) * 3, c->stream);
  {
    static int32_t mats[4096 * 15];
    static int have = 0;
    if (!have) {
      mav1hip_forward_matrices(mats);
      have = 1;
    }
    (void)hipMemcpyAsync(c->d_mats, mats, sizeof mats, hipMemcpyHostToDevice, c->stream);
  }
  (void)hipStreamSynchronize(c->stream);
  c->b_src = c->d_src;
  c->b_ref = c->d_ref;
  c->b_gold = c->d_gold;
  return c;
}

extern "C" {

void mav1hip_destroy(mav1hip_ctx *c) {
  if (!c) return;
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &t : c->pending) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (auto e : c->free_events) (void)hipEventDestroy(e);
  (void)hipFree(c->d_src);
  (void)hipFree(c->d_ref);
  (void)hipFree(c->d_gold);
  (void)hipFree(c->d_mvfield);
  (void)hipFree(c->d_mvgold);
  (void)hipFree(c->d_planes);
  (void)hipFree(c->d_acosts);
  (void)hipFree(c->d_ainfos);
  (void)hipFree(c->d_blocks);
  (void)hipFree(c->d_sbs);
  (void)hipFree(c->d_rec);
  (void)hipFree(c->d_coefs);
  (void)hipFree(c->d_mats);
  (void)hipFree(c->d_ys);
  (void)hipFree(c->d_uvs);
  (void)hipFree(c->d_lfmask);
  (void)hipFree(c->d_cdefbuf);
  (void)hipFree(c->d_dbk);
  (void)hipFree(c->d_lrout);
  (void)hipFree(c->d_dirvar);
  (void)hipFree(c->d_choice);
  (void)hipFree(c->d_decgrid);
  (void)hipFree(c->d_rwork);
  (void)hipFree(c->d_iwork);
  (void)hipFree(c->d_lrunits);
  (void)hipFree(c->d_lrstats);
  (void)hipFree(c->d_lrchain);
  (void)hipFree(c->d_lrtype);
  (void)hipFree(c->d_lrreruns);
  (void)hipFree(c->d_pool);
  for (int i = 0; i < 4; i++) (void)hipEventDestroy(c->me_done[i]);
  (void)hipEventDestroy(c->ana_done);
  (void)hipEventDestroy(c->xfer);
  (void)hipEventDestroy(c->rec_go);
  (void)hipEventDestroy(c->sched_done);
  if (c->own_streams && c->side) (void)hipStreamDestroy(c->side);
  if (c->own_streams && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *mav1hip_error(const mav1hip_ctx *c) { return c ? c->err : "no context"; }
const char *mav1hip_device_name(const mav1hip_ctx *c) { return c ? c->devname : ""; }

static void *buf_ptr(mav1hip_ctx *c, int buf, size_t *bytes) {
  size_t mvb = sizeof(int32_t) * 2 * MAV1_NUM_NODES * (size_t)c->num_sbs;
  switch (buf) {
    case MAV1HIP_BUF_SRC: *bytes = c->frame_bytes; return c->b_src;
    case MAV1HIP_BUF_REF: *bytes = c->frame_bytes; return c->b_ref;
    case MAV1HIP_BUF_GOLD: *bytes = c->frame_bytes; return c->b_gold;
    case MAV1HIP_BUF_MVFIELD: *bytes = mvb; return c->d_mvfield;
    case MAV1HIP_BUF_MVGOLD: *bytes = mvb; return c->d_mvgold;
    case MAV1HIP_BUF_PLANES: *bytes = c->planes_bytes; return c->d_planes;
    case MAV1HIP_BUF_ACOSTS: *bytes = sizeof(long long) * MAV1_A_ITEMS * (size_t)c->num_sbs; return c->d_acosts;
    case MAV1HIP_BUF_AINFOS: *bytes = sizeof(Mav1BlockInfo) * MAV1_A_ITEMS * (size_t)c->num_sbs; return c->d_ainfos;
    case MAV1HIP_BUF_BLOCKS:
      *bytes = sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs;
      return c->d_blocks;
    case MAV1HIP_BUF_SBS: *bytes = sizeof(Mav1SbInfo) * (size_t)c->num_sbs; return c->d_sbs;
    case MAV1HIP_BUF_REC: *bytes = c->frame_bytes; return c->d_rec;
    case MAV1HIP_BUF_COEFS: *bytes = sizeof(int32_t) * MAV1_COEFS_PER_SB * (size_t)c->num_sbs; return c->d_coefs;
    case MAV1HIP_BUF_LFMASK: *bytes = (size_t)c->mi_rows * c->mi_cols; return c->d_lfmask;
    case MAV1HIP_BUF_CDEFBUF: *bytes = c->frame_bytes * MAV1HIP_MAX_CDEF_SETS; return c->d_cdefbuf;
    case MAV1HIP_BUF_DBK: *bytes = c->frame_bytes; return c->d_dbk;
    case MAV1HIP_BUF_LROUT: *bytes = c->frame_bytes; return c->d_lrout;
    case MAV1HIP_BUF_LRUNITS: *bytes = sizeof(Mav1LrUnit) * (size_t)c->lr_units; return c->d_lrunits;
    case MAV1HIP_BUF_CDEF_CHOICE: *bytes = (size_t)c->num_sbs; return c->d_choice;
    case MAV1HIP_BUF_DECGRID: *bytes = c->decgrid_bytes; return c->d_decgrid;
    default: *bytes = 0; return nullptr;
  }
}

size_t mav1hip_buffer_bytes(const mav1hip_ctx *c, int32_t buf) {
  size_t b = 0;
  if (c) buf_ptr(const_cast<mav1hip_ctx *>(c), buf, &b);
  return b;
}
size_t mav1hip_planes_entry_bytes(const mav1hip_ctx *c) { return c ? (size_t)c->entry_bytes : 0; }

int mav1hip_upload(mav1hip_ctx *c, int32_t buf, const void *host, size_t bytes) {
  if (!c || !host) return MAV1HIP_EINVAL;
  c->me_pending = 0;
  size_t cap;
  void *d = buf_ptr(c, buf, &cap);
  if (!d || bytes > cap) return fail(c, MAV1HIP_EINVAL, "upload: buffer %d holds %zu bytes, %zu given", buf, cap, bytes);
  HIPCHK(c, hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, c->stream));
  return MAV1HIP_OK;
}

int mav1hip_download(mav1hip_ctx *c, int32_t buf, void *host, size_t bytes) {
  if (!c || !host) return MAV1HIP_EINVAL;
  size_t cap;
  void *d = buf_ptr(c, buf, &cap);
  if (!d || bytes > cap) return fail(c, MAV1HIP_EINVAL, "download: buffer %d holds %zu bytes, %zu asked", buf, cap, bytes);
  HIPCHK(c, hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

/* The parameters must describe the context's geometry: the kernels index
 * the device buffers with them. */
static int check_params(mav1hip_ctx *c, const Mav1FrameParams *fp) {
  if (fp->frame_w != c->cfg.width || fp->frame_h != c->cfg.height || fp->sb_cols != c->sb_cols ||
      fp->sb_rows != c->sb_rows || fp->mi_cols != c->mi_cols || fp->mi_rows != c->mi_rows)
    return fail(c, MAV1HIP_EINVAL, "frame parameters (%dx%d) do not match the context (%dx%d)", fp->frame_w,
                fp->frame_h, c->cfg.width, c->cfg.height);
  for (int p = 0; p < 3; p++)
    if (fp->stride[p] != c->stride[p] || fp->offset[p] != c->offset[p])
      return fail(c, MAV1HIP_EINVAL, "frame parameters: plane %d layout does not match the context", p);
  if (fp->me_range < 0 || fp->me_range > 256) return fail(c, MAV1HIP_EINVAL, "me_range %d out of range", fp->me_range);
  return MAV1HIP_OK;
}

int mav1hip_set_params(mav1hip_ctx *c, const Mav1FrameParams *fp) {
  if (!c || !fp) return MAV1HIP_EINVAL;
  int rc = check_params(c, fp);
  if (rc) return rc;
  c->fp = *fp;
  c->fp_set = 1;
  c->me_pending = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_lrtype, fp->lr_type, sizeof(int) * 3, hipMemcpyHostToDevice, c->stream));
  return MAV1HIP_OK;
}

int mav1hip_launch_motion_search(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "motion search: no frame parameters set");
  const Mav1FrameParams &fp = c->fp;
  const bool hp = (fp.turbo & MAV1_T_ME_HPEL) != 0;
  if (!hp && (fp.turbo & MAV1_T_ME_DEC))
    return fail(c, MAV1HIP_EUNSUP,
                "motion search: decimated-plane stages (turbo bit 32) without the quarter-pel planes (bit 16) read "
                "planes no stage builds");
  int nref = fp.golden ? 2 : 1;
  uint8_t *srcDec = c->d_planes + 2 * c->entry_bytes;
  size_t n16 = (size_t)16 * fp.frame_w * fp.frame_h;
  int rc;
  if (hp) {
    for (int r = 0; r < nref; r++) {
      const uint8_t *refp = r ? c->b_gold : c->b_ref;
      uint8_t *ent = c->d_planes + r * c->entry_bytes;
      if ((rc = timed(c, "qpel_build", [&] { return launch_qpel_build(c->stream, fp, refp, ent); }))) return rc;
      if (fp.turbo & MAV1_T_ME_DEC)
        if ((rc = timed(c, "dec_build", [&] { return launch_dec_build(c->stream, fp, refp, ent + n16); }))) return rc;
    }
    if (fp.turbo & MAV1_T_ME_DEC)
      if ((rc = timed(c, "dec_build", [&] { return launch_dec_build(c->stream, fp, c->b_src, srcDec); }))) return rc;
  }
  static const char *names[4] = {"me_level0", "me_level1", "me_level2", "me_level3"};
  c->me_pending = 0;
  for (int level = 0; level < 4; level++) {
    if ((rc = timed(c, names[level], [&] {
           return launch_me_level(c->stream, level, fp, c->b_src, c->b_ref, c->b_gold, c->d_planes, c->entry_bytes,
                                  srcDec, c->d_mvfield, c->d_mvgold, nref);
         })))
      return rc;
    HIPCHK(c, hipEventRecord(c->me_done[level], c->stream));
  }
  c->me_pending = 1;
  return MAV1HIP_OK;
}

int mav1hip_launch_analyze(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "analyze: no frame parameters set");
  const Mav1FrameParams &fp = c->fp;
  static const char *names[4] = {"analyze_level0", "analyze_level1", "analyze_level2", "analyze_level3"};
  int rc;
  /* Level k reads the motion field of level k only (and the costs of the
   * levels above it): after this frame's motion search it runs on the side
   * stream behind me_done[k], beside motion-search level k + 1. Level 0's
   * wait also orders it after everything enqueued on the main stream
   * before this frame's motion search. */
  const bool overlap = c->me_pending != 0;
  c->me_pending = 0;
  hipStream_t st = overlap ? c->side : c->stream;
  for (int level = 0; level < 4; level++) {
    if (overlap) HIPCHK(c, hipStreamWaitEvent(c->side, c->me_done[level], 0));
    if ((rc = timed_on(c, st, names[level], [&] {
           return launch_analyze(st, level, fp, c->b_src, c->b_ref, c->b_gold, c->d_mvfield, c->d_mvgold, c->d_acosts,
                                 c->d_ainfos);
         })))
      return rc;
  }
  if ((rc = timed_on(c, st, "analyze_table",
                     [&] { return launch_analyze_table(st, fp, c->d_acosts, c->d_ainfos, c->d_blocks, c->d_sbs); })))
    return rc;
  if (overlap) {
    HIPCHK(c, hipEventRecord(c->ana_done, c->side));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ana_done, 0));
  }
  return MAV1HIP_OK;
}

int mav1hip_analyze(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref,
                    const uint8_t *gold, const int32_t *mvfield, const int32_t *mvgold, Mav1BlockInfo *blocks,
                    Mav1SbInfo *sbs, int64_t *acosts, Mav1BlockInfo *ainfos) {
  if (!c || !fp || !src || !blocks || !sbs) return fail(c, MAV1HIP_EINVAL, "analyze: missing buffer");
  if (!fp->frame_is_intra && (!ref || !mvfield)) return fail(c, MAV1HIP_EINVAL, "analyze: inter frame without LAST");
  if (!fp->frame_is_intra && fp->golden && (!gold || !mvgold))
    return fail(c, MAV1HIP_EINVAL, "analyze: second reference missing");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  c->b_src = c->d_src;
  c->b_ref = c->d_ref;
  c->b_gold = c->d_gold;
  size_t mvb = sizeof(int32_t) * 2 * MAV1_NUM_NODES * (size_t)c->num_sbs;
  HIPCHK(c, hipMemcpyAsync(c->d_src, src, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if (!fp->frame_is_intra) {
    HIPCHK(c, hipMemcpyAsync(c->d_ref, ref, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mvfield, mvfield, mvb, hipMemcpyHostToDevice, c->stream));
    if (fp->golden) {
      HIPCHK(c, hipMemcpyAsync(c->d_gold, gold, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->d_mvgold, mvgold, mvb, hipMemcpyHostToDevice, c->stream));
    }
  }
  if ((rc = mav1hip_launch_analyze(c))) return rc;
  size_t nb = sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs;
  HIPCHK(c, hipMemcpyAsync(blocks, c->d_blocks, nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(sbs, c->d_sbs, sizeof(Mav1SbInfo) * (size_t)c->num_sbs, hipMemcpyDeviceToHost, c->stream));
  if (acosts)
    HIPCHK(c, hipMemcpyAsync(acosts, c->d_acosts, sizeof(long long) * MAV1_A_ITEMS * (size_t)c->num_sbs,
                             hipMemcpyDeviceToHost, c->stream));
  if (ainfos)
    HIPCHK(c, hipMemcpyAsync(ainfos, c->d_ainfos, sizeof(Mav1BlockInfo) * MAV1_A_ITEMS * (size_t)c->num_sbs,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_launch_edge_types(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "edge types: no frame parameters set");
  return timed(c, "edge_types",
               [&] { return launch_edge_types(c->stream, c->fp, c->d_blocks, c->d_sbs, c->d_ys, c->d_uvs); });
}

int mav1hip_launch_recon(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "recon: no frame parameters set");
  const Mav1FrameParams &fp = c->fp;
  /* The scheduling pass reads only is_inter of the block records, which the
   * inter phase writes back unchanged: it runs on the side stream beside the
   * inter phase, after everything enqueued so far (the previous frame's
   * intra phase included, whose work space it resets). */
  HIPCHK(c, hipEventRecord(c->rec_go, c->stream));
  HIPCHK(c, hipStreamWaitEvent(c->side, c->rec_go, 0));
  int rc = timed_on(c, c->side, "intra_sched",
                    [&] { return launch_intra_sched(c->side, fp, c->d_blocks, c->d_sbs, c->d_iwork); });
  if (rc) return rc;
  HIPCHK(c, hipEventRecord(c->sched_done, c->side));
  rc = timed(c, "recon_inter", [&] {
    return launch_recon_inter(c->stream, fp, c->d_blocks, c->d_sbs, c->b_src, c->d_rec, c->b_ref, c->b_gold,
                              c->d_coefs, c->d_mats, c->d_rwork, 2 * c->num_cu);
  });
  if (rc) return rc;
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->sched_done, 0));
  return timed(c, "recon_intra", [&] {
    return launch_recon_intra(c->stream, fp, c->d_blocks, c->d_sbs, c->b_src, c->d_rec, c->b_ref, c->b_gold,
                              c->d_coefs, c->d_mats, c->d_iwork, intra_workgroups(fp, c->num_cu));
  });
}

int mav1hip_launch_filter_info(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "filter info: no frame parameters set");
  return timed(c, "filter_info", [&] { return launch_filter_info(c->stream, c->fp, c->d_blocks, c->d_sbs, c->d_lfmask); });
}

int mav1hip_launch_deblock(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "deblock: no frame parameters set");
  return timed(c, "deblock", [&] { return launch_deblock(c->stream, c->fp, c->d_rec, c->d_lfmask); });
}

int mav1hip_launch_cdef(mav1hip_ctx *c, int32_t nsets) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "cdef: no frame parameters set");
  if (nsets < 0 || nsets > MAV1HIP_MAX_CDEF_SETS) return fail(c, MAV1HIP_EINVAL, "cdef: %d sets (at most %d)", nsets,
                                                           MAV1HIP_MAX_CDEF_SETS);
  if (!nsets) return MAV1HIP_OK;
  return timed(c, "cdef", [&] {
    return launch_cdef(c->stream, c->fp, c->d_rec, c->d_cdefbuf, (long long)c->frame_bytes, c->d_lfmask, c->d_dirvar,
                       nsets);
  });
}

int mav1hip_launch_save_deblocked(mav1hip_ctx *c) {
  if (!c) return MAV1HIP_EINVAL;
  c->me_pending = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_dbk, c->d_rec, c->frame_bytes, hipMemcpyDeviceToDevice, c->stream));
  return MAV1HIP_OK;
}

int mav1hip_launch_cdef_select(mav1hip_ctx *c, int32_t nsets) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "cdef select: no frame parameters set");
  if (nsets < 1 || nsets > MAV1HIP_MAX_CDEF_SETS) return fail(c, MAV1HIP_EINVAL, "cdef select: %d sets", nsets);
  return timed(c, "cdef_select", [&] {
    return launch_cdef_select(c->stream, c->fp, c->b_src, c->d_cdefbuf, (long long)c->frame_bytes, nsets, c->d_rec,
                              c->d_choice);
  });
}

static int check_lr_params(mav1hip_ctx *c, const Mav1FrameParams &fp) {
  int n = 0;
  for (int p = 0; p < 3; p++) {
    if (fp.lr_unit_size[p] != 64 || fp.lr_unit_rows[p] < 1 || fp.lr_unit_cols[p] < 1 || fp.lr_unit_off[p] != n)
      return fail(c, MAV1HIP_EINVAL, "loop restoration: unit layout of plane %d is not the 64x64 one", p);
    n += fp.lr_unit_rows[p] * fp.lr_unit_cols[p];
  }
  if (n != c->lr_units) return fail(c, MAV1HIP_EINVAL, "loop restoration: %d units, the context holds %d", n, c->lr_units);
  return MAV1HIP_OK;
}

int mav1hip_launch_restore(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "restore: no frame parameters set");
  int rc = check_lr_params(c, c->fp);
  if (rc) return rc;
  return timed(c, "lr_apply", [&] {
    return launch_lr_apply(c->stream, c->fp, c->d_rec, c->d_dbk, c->d_lrout, c->d_lrunits, c->d_lrtype);
  });
}

int mav1hip_launch_lr_select(mav1hip_ctx *c, int32_t lr_test) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "lr select: no frame parameters set");
  int rc = check_lr_params(c, c->fp);
  if (rc) return rc;
  c->lr_gen = c->lr_gen % 0x7ffffffe + 1;
  if ((rc = timed(c, "lr_fit", [&] {
         return launch_lr_fit(c->stream, c->fp, c->b_src, c->d_rec, c->d_dbk, c->d_lrstats, c->d_lrchain, c->lr_units,
                              c->lr_gen);
       })))
    return rc;
  /* the fit launch refines under the default reference; then up to
   * LR_ROUNDS - 1 chain launches that stop at a reference change, each
   * followed by the refinement of the remaining units under the new
   * reference; the last chain launch refines any further mismatching unit
   * itself */
  const int LR_ROUNDS = 3;
  for (int r = 0; r < LR_ROUNDS; r++) {
    if (r > 0 && (rc = timed(c, "lr_path", [&] {
                    return launch_lr_path(c->stream, c->fp, c->b_src, c->d_rec, c->d_dbk, c->d_lrstats, c->d_lrchain,
                                          c->lr_units);
                  })))
      return rc;
    if ((rc = timed(c, "lr_chain", [&] {
           return launch_lr_chain(c->stream, c->fp, c->b_src, c->d_rec, c->d_dbk, c->d_lrstats, c->d_lrunits,
                                  c->d_lrtype, lr_test != 0, c->d_lrreruns, c->d_lrchain, r < LR_ROUNDS - 1);
         })))
      return rc;
  }
  return MAV1HIP_OK;
}

int mav1hip_lr_select(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdef,
                      const uint8_t *dbk, int32_t lr_test, Mav1LrUnit *units, int32_t lr_type[3], int32_t reruns[3]) {
  if (!c || !fp || !src || !cdef || !dbk || !units || !lr_type) return fail(c, MAV1HIP_EINVAL, "lr select: missing buffer");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  if ((rc = check_lr_params(c, *fp))) return rc;
  c->b_src = c->d_src;
  HIPCHK(c, hipMemcpyAsync(c->d_src, src, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_rec, cdef, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_dbk, dbk, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_lr_select(c, lr_test))) return rc;
  HIPCHK(c, hipMemcpyAsync(units, c->d_lrunits, sizeof(Mav1LrUnit) * (size_t)c->lr_units, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(lr_type, c->d_lrtype, sizeof(int) * 3, hipMemcpyDeviceToHost, c->stream));
  if (reruns) HIPCHK(c, hipMemcpyAsync(reruns, c->d_lrreruns, sizeof(int) * 3, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_filter_info(mav1hip_ctx *c, const Mav1FrameParams *fp, const Mav1BlockInfo *blocks,
                        const Mav1SbInfo *sbs, uint8_t *info) {
  if (!c || !fp || !blocks || !sbs || !info) return fail(c, MAV1HIP_EINVAL, "filter info: missing buffer");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  size_t nb = sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs;
  HIPCHK(c, hipMemcpyAsync(c->d_blocks, blocks, nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_sbs, sbs, sizeof(Mav1SbInfo) * (size_t)c->num_sbs, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_filter_info(c))) return rc;
  HIPCHK(c, hipMemcpyAsync(info, c->d_lfmask, (size_t)c->mi_rows * c->mi_cols, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_filter(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *info, uint8_t *rec, int32_t nsets,
                   uint8_t *cdefbuf) {
  if (!c || !fp || !info || !rec || (nsets && !cdefbuf)) return fail(c, MAV1HIP_EINVAL, "filter: missing buffer");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_lfmask, info, (size_t)c->mi_rows * c->mi_cols, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_rec, rec, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_deblock(c))) return rc;
  if ((rc = mav1hip_launch_cdef(c, nsets))) return rc;
  HIPCHK(c, hipMemcpyAsync(rec, c->d_rec, c->frame_bytes, hipMemcpyDeviceToHost, c->stream));
  if (nsets)
    HIPCHK(c, hipMemcpyAsync(cdefbuf, c->d_cdefbuf, c->frame_bytes * nsets, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_cdef_select(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdefbuf,
                        int32_t nsets, uint8_t *rec, int8_t *choice) {
  if (!c || !fp || !src || !cdefbuf || !rec || !choice) return fail(c, MAV1HIP_EINVAL, "cdef select: missing buffer");
  if (nsets < 1 || nsets > MAV1HIP_MAX_CDEF_SETS) return fail(c, MAV1HIP_EINVAL, "cdef select: %d sets", nsets);
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  c->b_src = c->d_src;
  HIPCHK(c, hipMemcpyAsync(c->d_src, src, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_cdefbuf, cdefbuf, c->frame_bytes * nsets, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_rec, rec, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_cdef_select(c, nsets))) return rc;
  HIPCHK(c, hipMemcpyAsync(rec, c->d_rec, c->frame_bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(choice, c->d_choice, (size_t)c->num_sbs, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_restore(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *cdef, const uint8_t *dbk,
                    const Mav1LrUnit *units, uint8_t *out) {
  if (!c || !fp || !cdef || !dbk || !units || !out) return fail(c, MAV1HIP_EINVAL, "restore: missing buffer");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  if ((rc = check_lr_params(c, *fp))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_rec, cdef, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_dbk, dbk, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_lrunits, units, sizeof(Mav1LrUnit) * (size_t)c->lr_units, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_lrout, out, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_restore(c))) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->d_lrout, c->frame_bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_launch_decision_grid(mav1hip_ctx *c) {
  if (!c || !c->fp_set) return fail(c, MAV1HIP_EINVAL, "decision grid: no frame parameters set");
  return timed(c, "decision_grid",
               [&] { return launch_decision_grid(c->stream, c->fp, c->d_blocks, c->d_sbs, c->d_decgrid); });
}

size_t mav1hip_decision_record_bytes(const Mav1FrameParams *fp) {
  if (!fp) return 0;
  return 44 + (size_t)((fp->mi_rows + 1) >> 1) * ((fp->mi_cols + 1) >> 1) * 8;
}

int mav1hip_decision_record(mav1hip_ctx *c, const Mav1FrameParams *fp, const Mav1BlockInfo *blocks,
                            const Mav1SbInfo *sbs, const int32_t hdr7[7], uint8_t *out) {
  if (!c || !fp || !blocks || !sbs || !hdr7 || !out) return fail(c, MAV1HIP_EINVAL, "decision record: missing buffer");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  size_t nb = sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs;
  HIPCHK(c, hipMemcpyAsync(c->d_blocks, blocks, nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_sbs, sbs, sizeof(Mav1SbInfo) * (size_t)c->num_sbs, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_decision_grid(c))) return rc;
  const int32_t hdr[11] = {0x4345444D, hdr7[0], hdr7[1], hdr7[2], hdr7[3], hdr7[4], hdr7[5], fp->base_q_idx,
                           (c->mi_rows + 1) >> 1, (c->mi_cols + 1) >> 1, hdr7[6]};
  memcpy(out, hdr, sizeof hdr);
  HIPCHK(c, hipMemcpyAsync(out + sizeof hdr, c->d_decgrid, c->decgrid_bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_edge_types(mav1hip_ctx *c, const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs) {
  if (!c || !fp || !blocks || !sbs) return fail(c, MAV1HIP_EINVAL, "edge types: missing buffer");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  size_t nb = sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs;
  HIPCHK(c, hipMemcpyAsync(c->d_blocks, blocks, nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_sbs, sbs, sizeof(Mav1SbInfo) * (size_t)c->num_sbs, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_edge_types(c))) return rc;
  HIPCHK(c, hipMemcpyAsync(blocks, c->d_blocks, nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

/* After a synchronisation: the intra phase's error word (a dependency wait
 * that ran out, which would leave the reconstruction wrong). */
static int check_intra(mav1hip_ctx *c) {
  int e = 0;
  HIPCHK(c, hipMemcpy(&e, c->d_iwork + 2, sizeof e, hipMemcpyDeviceToHost));
  if (e) return fail(c, MAV1HIP_EDEVICE, "intra phase: a superblock dependency wait ran out");
  return MAV1HIP_OK;
}

int mav1hip_recon(mav1hip_ctx *c, const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                  const uint8_t *src, uint8_t *rec, const uint8_t *ref, const uint8_t *gold, int32_t *coefs) {
  if (!c || !fp || !blocks || !sbs || !src || !rec || !coefs) return fail(c, MAV1HIP_EINVAL, "recon: missing buffer");
  if (!fp->frame_is_intra && !ref) return fail(c, MAV1HIP_EINVAL, "recon: inter frame without LAST");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  c->b_src = c->d_src;
  c->b_ref = c->d_ref;
  c->b_gold = c->d_gold;
  size_t nb = sizeof(Mav1BlockInfo) * MAV1_MAX_BLOCKS_PER_SB * (size_t)c->num_sbs;
  size_t ncoef = sizeof(int32_t) * MAV1_COEFS_PER_SB * (size_t)c->num_sbs;
  HIPCHK(c, hipMemcpyAsync(c->d_blocks, blocks, nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_sbs, sbs, sizeof(Mav1SbInfo) * (size_t)c->num_sbs, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_src, src, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_rec, rec, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_coefs, coefs, ncoef, hipMemcpyHostToDevice, c->stream));
  if (ref) HIPCHK(c, hipMemcpyAsync(c->d_ref, ref, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if (gold) HIPCHK(c, hipMemcpyAsync(c->d_gold, gold, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_recon(c))) return rc;
  HIPCHK(c, hipMemcpyAsync(blocks, c->d_blocks, nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(rec, c->d_rec, c->frame_bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(coefs, c->d_coefs, ncoef, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_intra(c);
}

int mav1hip_sync(mav1hip_ctx *c) {
  if (!c) return MAV1HIP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->side));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_intra(c);
}

int mav1hip_motion_search(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref,
                          const uint8_t *gold, int32_t *mvfield, int32_t *mvgold) {
  if (!c || !fp || !src || !ref || !mvfield) return fail(c, MAV1HIP_EINVAL, "motion search: missing buffer");
  if (fp->golden && (!gold || !mvgold)) return fail(c, MAV1HIP_EINVAL, "motion search: second reference missing");
  int rc;
  if ((rc = mav1hip_set_params(c, fp))) return rc;
  c->b_src = c->d_src;
  c->b_ref = c->d_ref;
  c->b_gold = c->d_gold;
  HIPCHK(c, hipMemcpyAsync(c->d_src, src, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_ref, ref, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if (fp->golden) HIPCHK(c, hipMemcpyAsync(c->d_gold, gold, c->frame_bytes, hipMemcpyHostToDevice, c->stream));
  if ((rc = mav1hip_launch_motion_search(c))) return rc;
  size_t mvb = sizeof(int32_t) * 2 * MAV1_NUM_NODES * (size_t)c->num_sbs;
  HIPCHK(c, hipMemcpyAsync(mvfield, c->d_mvfield, mvb, hipMemcpyDeviceToHost, c->stream));
  if (fp->golden) HIPCHK(c, hipMemcpyAsync(mvgold, c->d_mvgold, mvb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_pool_alloc(mav1hip_ctx *c, int32_t n) {
  if (!c || n < 0) return MAV1HIP_EINVAL;
  if (c->d_pool) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->d_pool);
    c->d_pool = nullptr;
    c->pool_n = 0;
    c->b_src = c->d_src;
    c->b_ref = c->d_ref;
    c->b_gold = c->d_gold;
  }
  if (n == 0) return MAV1HIP_OK;
  int rc = dmalloc(c, (void **)&c->d_pool, c->frame_bytes * (size_t)n);
  if (rc) return rc;
  c->pool_n = n;
  return MAV1HIP_OK;
}

int mav1hip_pool_upload(mav1hip_ctx *c, int32_t idx, const uint8_t *frame) {
  if (!c || !frame || idx < 0 || idx >= c->pool_n) return fail(c, MAV1HIP_EINVAL, "pool_upload: bad index %d", idx);
  c->me_pending = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_pool + (size_t)idx * c->frame_bytes, frame, c->frame_bytes, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAV1HIP_OK;
}

int mav1hip_bind_frames(mav1hip_ctx *c, int32_t s, int32_t r, int32_t g) {
  if (!c) return MAV1HIP_EINVAL;
  if (s >= c->pool_n || r >= c->pool_n || g >= c->pool_n) return fail(c, MAV1HIP_EINVAL, "bind_frames: bad index");
  c->b_src = s < 0 ? c->d_src : c->d_pool + (size_t)s * c->frame_bytes;
  c->b_ref = r < 0 ? c->d_ref : c->d_pool + (size_t)r * c->frame_bytes;
  c->b_gold = g < 0 ? c->d_gold : c->d_pool + (size_t)g * c->frame_bytes;
  return MAV1HIP_OK;
}

int mav1hip_pool_store(mav1hip_ctx *c, int32_t idx, int32_t buf) {
  if (!c || idx < 0 || idx >= c->pool_n) return fail(c, MAV1HIP_EINVAL, "pool_store: bad index %d", idx);
  c->me_pending = 0;
  if (buf != MAV1HIP_BUF_REC && buf != MAV1HIP_BUF_LROUT) return fail(c, MAV1HIP_EINVAL, "pool_store: buffer %d", buf);
  uint8_t *dst = c->d_pool + (size_t)idx * c->frame_bytes;
  if (dst == c->b_src || dst == c->b_ref || dst == c->b_gold)
    return fail(c, MAV1HIP_EINVAL, "pool_store: frame %d is bound to a stage input", idx);
  HIPCHK(c, hipMemcpyAsync(dst, buf == MAV1HIP_BUF_REC ? c->d_rec : c->d_lrout, c->frame_bytes,
                           hipMemcpyDeviceToDevice, c->stream));
  return MAV1HIP_OK;
}

int mav1hip_pool_import(mav1hip_ctx *dst, int32_t dst_idx, mav1hip_ctx *src, int32_t src_idx) {
  if (!dst || !src || dst == src) return fail(dst, MAV1HIP_EINVAL, "pool_import: needs two contexts");
  if (dst_idx < 0 || dst_idx >= dst->pool_n) return fail(dst, MAV1HIP_EINVAL, "pool_import: bad index %d", dst_idx);
  if (src_idx < 0 || src_idx >= src->pool_n)
    return fail(dst, MAV1HIP_EINVAL, "pool_import: bad source index %d", src_idx);
  if (dst->device != src->device || dst->frame_bytes != src->frame_bytes)
    return fail(dst, MAV1HIP_EINVAL, "pool_import: contexts differ in device or frame geometry");
  uint8_t *to = dst->d_pool + (size_t)dst_idx * dst->frame_bytes;
  if (to == dst->b_src || to == dst->b_ref || to == dst->b_gold)
    return fail(dst, MAV1HIP_EINVAL, "pool_import: frame %d is bound to a stage input", dst_idx);
  dst->me_pending = src->me_pending = 0;
  /* after everything enqueued on src so far (the picture is complete), and
   * before src's later work (which may overwrite the source frame) */
  HIPCHK(dst, hipEventRecord(src->xfer, src->stream));
  HIPCHK(dst, hipStreamWaitEvent(dst->stream, src->xfer, 0));
  HIPCHK(dst, hipMemcpyAsync(to, src->d_pool + (size_t)src_idx * src->frame_bytes, dst->frame_bytes,
                             hipMemcpyDeviceToDevice, dst->stream));
  HIPCHK(dst, hipEventRecord(dst->xfer, dst->stream));
  HIPCHK(dst, hipStreamWaitEvent(src->stream, dst->xfer, 0));
  return MAV1HIP_OK;
}

int mav1hip_timing_enable(mav1hip_ctx *c, int32_t on) {
  if (!c) return MAV1HIP_EINVAL;
  c->timing = on != 0;
  return MAV1HIP_OK;
}

int mav1hip_timing_select(mav1hip_ctx *c, const char *kernel) {
  if (!c) return MAV1HIP_EINVAL;
  c->timing_only = kernel ? kernel : "";
  return MAV1HIP_OK;
}

int mav1hip_timing_reset(mav1hip_ctx *c) {
  if (!c) return MAV1HIP_EINVAL;
  int rc = collect_timing(c);
  c->stats.clear();
  return rc;
}

int mav1hip_timing_get(mav1hip_ctx *c, const char *kernel, int32_t *launches, double *total_ms) {
  if (!c || !kernel) return MAV1HIP_EINVAL;
  int rc = collect_timing(c);
  if (rc) return rc;
  auto it = c->stats.find(kernel);
  if (launches) *launches = it == c->stats.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.second;
  return MAV1HIP_OK;
}

}  // extern "C"

/* Library-internal access for the sequence driver (seq.hip). */
hipStream_t ctx_stream(mav1hip_ctx *c) { return c->stream; }
hipStream_t ctx_side_stream(mav1hip_ctx *c) { return c->side; }
int ctx_device(const mav1hip_ctx *c) { return c->device; }
size_t ctx_frame_bytes(const mav1hip_ctx *c) { return c->frame_bytes; }
void *ctx_buffer(mav1hip_ctx *c, int buf, size_t *bytes) { return buf_ptr(c, buf, bytes); }
const int *ctx_intra_error(const mav1hip_ctx *c) { return c->d_iwork + 2; }
const int *ctx_lrtype(const mav1hip_ctx *c) { return c->d_lrtype; }
int ctx_bind_ptrs(mav1hip_ctx *c, uint8_t *src, uint8_t *ref, uint8_t *gold) {
  if (!c || !src) return MAV1HIP_EINVAL;
  c->b_src = src;
  c->b_ref = ref ? ref : c->d_ref;
  c->b_gold = gold ? gold : c->d_gold;
  return MAV1HIP_OK;
}
/* Copies a device frame buffer into dst on the context's stream. */
int ctx_store(mav1hip_ctx *c, uint8_t *dst, int buf) {
  if (!c || !dst || (buf != MAV1HIP_BUF_REC && buf != MAV1HIP_BUF_LROUT))
    return fail(c, MAV1HIP_EINVAL, "store: buffer %d", buf);
  if (dst == c->b_src || dst == c->b_ref || dst == c->b_gold)
    return fail(c, MAV1HIP_EINVAL, "store: the frame is bound to a stage input");
  c->me_pending = 0;
  HIPCHK(c, hipMemcpyAsync(dst, buf == MAV1HIP_BUF_REC ? c->d_rec : c->d_lrout, c->frame_bytes,
                           hipMemcpyDeviceToDevice, c->stream));
  return MAV1HIP_OK;
}
/* REC copied into LROUT on the context's stream (the driver's copy before
 * the restoration stage, src/encoder.c:636: LROUT outside the restored
 * area keeps the reconstruction). */
int ctx_rec_to_lrout(mav1hip_ctx *c) {
  if (!c) return MAV1HIP_EINVAL;
  c->me_pending = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_lrout, c->d_rec, c->frame_bytes, hipMemcpyDeviceToDevice, c->stream));
  return MAV1HIP_OK;
}
/* CDEF choices set to 0 (the driver's memset when the frame skips CDEF,
 * src/encoder.c:624-627). */
int ctx_clear_choice(mav1hip_ctx *c) {
  if (!c) return MAV1HIP_EINVAL;
  c->me_pending = 0;
  HIPCHK(c, hipMemsetAsync(c->d_choice, 0, (size_t)c->num_sbs, c->stream));
  return MAV1HIP_OK;
}
int ctx_batch_frame(mav1hip_ctx *c, BatchFrame *bf, int new_lr_gen) {
  if (!c || !bf || !c->fp_set) return MAV1HIP_EINVAL;
  memset(bf, 0, sizeof *bf);
  bf->fp = c->fp;
  bf->src = c->b_src;
  bf->ref = c->b_ref;
  bf->gold = c->b_gold;
  bf->planes = c->d_planes;
  bf->entry_bytes = c->entry_bytes;
  bf->mvf = c->d_mvfield;
  bf->mvg = c->d_mvgold;
  bf->acosts = c->d_acosts;
  bf->ainfos = c->d_ainfos;
  bf->blocks = c->d_blocks;
  bf->sbs = c->d_sbs;
  bf->rec = c->d_rec;
  bf->coefs = c->d_coefs;
  bf->mats = c->d_mats;
  bf->ys = c->d_ys;
  bf->uvs = c->d_uvs;
  bf->rwork = c->d_rwork;
  bf->iwork = c->d_iwork;
  bf->lfmask = c->d_lfmask;
  bf->cdefbuf = c->d_cdefbuf;
  bf->dbk = c->d_dbk;
  bf->lrout = c->d_lrout;
  bf->decgrid = c->d_decgrid;
  bf->dirvar = c->d_dirvar;
  bf->choice = c->d_choice;
  bf->lrunits = c->d_lrunits;
  bf->lrstats = c->d_lrstats;
  bf->lrchain = c->d_lrchain;
  bf->lrtype = c->d_lrtype;
  bf->lrreruns = c->d_lrreruns;
  bf->frame_bytes = (long long)c->frame_bytes;
  bf->lr_units = c->lr_units;
  if (new_lr_gen) c->lr_gen = c->lr_gen % 0x7ffffffe + 1;
  bf->lr_gen = c->lr_gen;
  c->me_pending = 0;
  return MAV1HIP_OK;
}
int ctx_num_cu(const mav1hip_ctx *c) { return c ? c->num_cu : 0; }
