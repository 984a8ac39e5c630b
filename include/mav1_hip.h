/* mav1 HIP analysis backend: C ABI.
 *
 * libmav1hip.so runs the mav1 encoder's analysis path (everything behind
 * the encoder's `Backend` vtable, src/backend.h:10-24) on one MI355X GPU
 * (gfx950). It takes plain host pointers and sizes; the records are the
 * layout-identical structs of mav1_types.h. A mav1 build binds it through
 * host/backend_hip.c (a `Backend` whose slots call these functions).
 *
 * Two ways to drive a context:
 *  - stage calls with host buffers (mav1hip_motion_search, ...): inputs are
 *    copied to the device, the stage runs, outputs are copied back and the
 *    call returns when they are in the host buffers; the semantics of the
 *    matching Backend slot;
 *  - the device-resident form (mav1hip_upload / mav1hip_launch_* /
 *    mav1hip_download / mav1hip_sync): stages are enqueued on the context's
 *    HIP stream against device buffers and nothing crosses PCIe until asked.
 *
 * Return values: 0 on success, a negative MAV1HIP_E* code on failure, with a
 * message in mav1hip_error(ctx). A context serves one caller at a time; use
 * one context per concurrent caller (the encoder's frame engines each own a
 * Backend, src/encoder.c:196-208).
 */
#ifndef MAV1_HIP_H
#define MAV1_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "mav1_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MAV1HIP_OK 0
#define MAV1HIP_EINVAL (-1)    /* bad argument or size */
#define MAV1HIP_EDEVICE (-2)   /* HIP runtime / device error */
#define MAV1HIP_ENOMEM (-3)    /* device allocation failed */
#define MAV1HIP_EUNSUP (-4)    /* configuration this build does not run on the GPU */

/* ------------------------------------------------------------------------
 * Encoder configuration and per-frame parameters (host only; no device).
 * ------------------------------------------------------------------------ */

/* The encoder configuration (Mav1Config, src/mav1.h:20-42) plus the testing
 * knobs the reference reads from the environment; a knob of -1 takes the
 * reference's default. */
typedef struct {
  int32_t width, height;
  int32_t qindex;
  int32_t keyint;
  int32_t tx_search, tx_depth, loop_filter, cdef, restoration;
  int32_t tile_cols_log2, intra_edge_filter;
  int32_t nonref_odd, fast, multi_ref, altref, leaf, switchable_interp;
  int32_t me_range;
  int32_t turbo;
  int32_t knob_turbo_mask;  /* MAV1_TURBO (default 0x47CBF1) */
  int32_t knob_dt;          /* MAV1_DT (default 320 when sb_rows >= 16, else 352) */
  int32_t knob_dt_bias;     /* MAV1_DTB (default 0) */
  int32_t knob_cdef_preset; /* MAV1_CP (default 2) */
  int32_t knob_leafopt;     /* MAV1_LEAFOPT (default 0x3F) */
} Mav1HipConfig;

/* Frame-level state the parameter setup reads (src/encoder.c:1004-1030). */
typedef struct {
  int32_t is_key;
  int32_t frame_kind;       /* 0 regular, 1 hidden ALTREF, 2 inside a mini-GOP, 3 pyramid middle */
  int32_t nonref;           /* not kept as a reference (leaf frame) */
  int32_t order_hint;
  int32_t reference_select; /* compound LAST + ALTREF allowed */
  int32_t gold_differs;     /* the GOLDEN slot holds another picture than LAST */
} Mav1HipFrame;

/* mav1_default_config (src/encoder.c:51-66) followed by the CLI's speed
 * presets (tools/mav1enc.c:98-100): turbo implies fast and 4 tile columns,
 * fast turns off tx search, tx depth, restoration and switchable filters. */
void mav1hip_config_default(Mav1HipConfig *cfg, int32_t width, int32_t height, int32_t turbo);
/* The adjustments mav1_create applies to a configuration
 * (src/encoder.c:75-92): turbo => fast, altref >= 2, leaf; clamps. */
void mav1hip_config_normalize(Mav1HipConfig *cfg);
/* Frame geometry (src/encoder.c:93-120): fills mi/sb counts, tile width,
 * strides, offsets; returns the padded frame size in bytes (0: bad size). */
size_t mav1hip_geometry(const Mav1HipConfig *cfg, int32_t *mi_cols, int32_t *mi_rows, int32_t *sb_cols,
                        int32_t *sb_rows, int32_t *tile_w_sb, int32_t stride[3], int32_t offset[3],
                        int32_t plane_h[3]);
/* Frame quantizer cascade (src/encoder.c:585-596). */
int32_t mav1hip_frame_qindex(const Mav1HipConfig *cfg, const Mav1HipFrame *fr);
/* setup_params (src/encoder.c:407-510) for a frame coded at base_q_idx.
 * Returns the number of CDEF presets the filter stage evaluates. */
int32_t mav1hip_setup_params(const Mav1HipConfig *cfg, const Mav1HipFrame *fr, int32_t base_q_idx,
                             Mav1FrameParams *fp);

/* Which in-loop filter stages a frame runs (run_frame_core,
 * src/encoder.c:618-640): filter info and the filter stage unless the frame
 * has nothing to filter (turbo leaf frames), CDEF sets of the filter stage,
 * the per-superblock CDEF choice, loop restoration. */
typedef struct {
  int32_t filter_info;  /* build the per-mi filter info */
  int32_t deblock;      /* a deblocking level is non-zero */
  int32_t cdef_sets;    /* CDEF parameter sets the filter stage produces (0: none) */
  int32_t cdef_select;  /* choose a set per superblock into the reconstruction */
  int32_t restoration;  /* loop-restoration search and apply */
} Mav1HipFilterPlan;
int32_t mav1hip_filter_plan(const Mav1HipConfig *cfg, const Mav1HipFrame *fr, const Mav1FrameParams *fp,
                            int32_t npresets, Mav1HipFilterPlan *plan);

/* The forward-transform matrices in Q20 (fwd_build_matrices,
 * src/fwd_matrices.c:57-65): 15 slots of 4096 int32, slot
 * (kind * 5 + log2(n) - 2) for kind 0 DCT, 1 ADST, 2 identity. Every context
 * computes them at creation. */
void mav1hip_forward_matrices(int32_t *mats);

/* ------------------------------------------------------------------------
 * Context
 * ------------------------------------------------------------------------ */
typedef struct mav1hip_ctx mav1hip_ctx;

/* Creates a context with device buffers for frames of cfg's size (the
 * convention of backend_metal_create, src/backend.h:37: NULL and a message
 * in err when there is no usable device). */
mav1hip_ctx *mav1hip_create(const Mav1HipConfig *cfg, char *err, size_t errlen);
/* Backend.destroy (src/backend.h:20). */
void mav1hip_destroy(mav1hip_ctx *c);
const char *mav1hip_error(const mav1hip_ctx *c);
/* Name of the device the context runs on. */
const char *mav1hip_device_name(const mav1hip_ctx *c);

/* ------------------------------------------------------------------------
 * Stage calls with host buffers (Backend slots, src/backend.h:10-24)
 * ------------------------------------------------------------------------ */

/* Backend.motion_search (src/backend_cpu.c:23-56): quarter-pel and
 * decimated planes, then the 4-level search of every 64/32/16/8 node.
 * src, ref, gold: padded frames (frame_bytes); gold is read when
 * fp->golden. mvfield / mvgold: num_sbs * 85 * 2 int32 (mvgold written
 * when fp->golden, may be NULL otherwise). */
int mav1hip_motion_search(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref,
                          const uint8_t *gold, int32_t *mvfield, int32_t *mvgold);

/* Backend.analyze (src/backend_cpu.c:58-74): pass 1 evaluates every
 * (superblock, node, slot) table item, pass 2 picks each superblock's
 * partition tree and writes its blocks in decode order. mvfield / mvgold:
 * the motion search's fields (inter frames). blocks: num_sbs * 64 records,
 * sbs: num_sbs. acosts (num_sbs * 1105 int64, -1 = not evaluated) and
 * ainfos (num_sbs * 1105 records, defined where the cost is >= 0) receive
 * the pass-1 table when not NULL. */
int mav1hip_analyze(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *ref,
                    const uint8_t *gold, const int32_t *mvfield, const int32_t *mvgold, Mav1BlockInfo *blocks,
                    Mav1SbInfo *sbs, int64_t *acosts, Mav1BlockInfo *ainfos);

/* Intra edge filter types of the blocks (the driver's edge_types_rows,
 * src/encoder.c:319-346): interp_filter bits 4 / 5 from the smooth modes of
 * the above and left neighbours; blocks updated in place. */
int mav1hip_edge_types(mav1hip_ctx *c, const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs);

/* Backend.recon (src/backend_cpu.c:153-156, per superblock recon_sb of
 * src/kernels/k_recon.h): prediction, transform type / depth choice,
 * forward transform, quantisation and reconstruction of every block in
 * decode order. blocks are completed in place (tx_type, eob, cfl alphas,
 * tx_depth, tx_types, interp filter); rec (frame_bytes) receives the
 * reconstruction inside the mi area (its other bytes are kept); coefs
 * (num_sbs * 6144 int32) the quantised levels at each block's coef_off. */
int mav1hip_recon(mav1hip_ctx *c, const Mav1FrameParams *fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                  const uint8_t *src, uint8_t *rec, const uint8_t *ref, const uint8_t *gold, int32_t *coefs);

/* The driver's build_filter_info (src/encoder.c:513-526): per mi
 * (mi_rows * mi_cols bytes) bsize | skip << 5 | tx_depth << 6 of the coded
 * blocks. */
int mav1hip_filter_info(mav1hip_ctx *c, const Mav1FrameParams *fp, const Mav1BlockInfo *blocks,
                        const Mav1SbInfo *sbs, uint8_t *info);

/* Backend.filter (src/backend_cpu.c:158-177): deblocking of rec in place
 * (lf_line, src/kernels/k_lf.h), then CDEF of the deblocked frame with
 * parameter sets 0 .. nsets-1 into cdefbuf + k * frame_bytes
 * (cdef_block_all, src/kernels/k_cdef.h:124-150). nsets <= 8; cdefbuf may
 * be NULL when nsets is 0. */
int mav1hip_filter(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *info, uint8_t *rec, int32_t nsets,
                   uint8_t *cdefbuf);

/* The driver's cdef_select (src/encoder.c:530-563): per superblock the set
 * with the smallest SSE against src over the visible area (first minimum)
 * into choice[sb], copied from cdefbuf into rec over the mi-aligned area. */
int mav1hip_cdef_select(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdefbuf,
                        int32_t nsets, uint8_t *rec, int8_t *choice);

/* Backend.restore (src/backend_cpu.c:179-188, lr_block of
 * src/kernels/k_lr.h): loop restoration of the CDEF output cdef, with the
 * deblocked frame dbk above and below each stripe, into out; units holds
 * the per-unit parameters (fp->lr_unit_off indexing), planes whose lr_type
 * is RESTORE_NONE are copied. */
int mav1hip_restore(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *cdef, const uint8_t *dbk,
                    const Mav1LrUnit *units, uint8_t *out);

/* The driver's lr_select (src/lr_search.c:235-359): per restoration unit
 * the parameters of smallest 16 * SSE + lambda * bits among NONE, the
 * fitted and refined Wiener filter and the 16 self-guided sets, given the
 * source, the CDEF output cdef and the deblocked frame dbk. units receives
 * one record per unit (fp->lr_unit_off indexing), lr_type[3] the frame's
 * restoration type per plane (the reference writes fp->lr_type). lr_test
 * forces Wiener on even units and the self-guided set of smallest SSE on
 * odd ones (Mav1Config.lr_test). reruns[3] (may be NULL) receives the
 * number of units per plane whose tap refinement was recomputed with the
 * true references (a diagnostic of the speculative search). */
int mav1hip_lr_select(mav1hip_ctx *c, const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdef,
                      const uint8_t *dbk, int32_t lr_test, Mav1LrUnit *units, int32_t lr_type[3], int32_t reruns[3]);

/* No counterpart in the reference: mav1hip_lr_select for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It needs no context and
 * calls no allocator: src, cdef and dbk (padded frames), units (one record
 * per unit, fp->lr_unit_off indexing) and lr_type (int32[3]) are DEVICE
 * pointers the caller owns; the search's scratch is a static device array
 * sized for MAV1HIP_STATIC_LR_UNITS restoration units (about 1.2 MB per
 * device, held until the process ends), taken and cleared on the first call
 * on each device. The kernels are enqueued on `stream` (a hipStream_t; NULL
 * is the null stream) and the call returns without waiting. More units
 * than the scratch holds return MAV1HIP_EINVAL. Every call on one device
 * shares that device's scratch: calls that overlap on one device race.
 * err (may be NULL) receives a message on failure. */
#define MAV1HIP_STATIC_LR_UNITS 3072
int mav1hip_lr_select_static(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdef, const uint8_t *dbk,
                             int32_t lr_test, Mav1LrUnit *units, int32_t *lr_type, void *stream, char *err,
                             size_t errlen);

/* The per-frame decision record of the driver's dump_decisions
 * (src/encoder.c:659-704, MAV1_DUMP_DEC): 11 int32 (magic 'MDEC',
 * order_hint, last_hint, arf_hint, frame_type 0 key / 1 inter, frame_kind,
 * nonref, base_q_idx, rows8, cols8, tile_bytes) then rows8 * cols8 cells of
 * 8 bytes (bsize, mode 0 intra / 1 single / 2 compound, ref_frame, 0,
 * int16 mv_row, int16 mv_col) of the last block covering each 8x8 cell.
 * hdr7 = {order_hint, last_hint, arf_hint, frame_type, frame_kind, nonref,
 * tile_bytes} (driver state); the cells come from the coded blocks. out
 * holds mav1hip_decision_record_bytes(fp) bytes; the caller appends it to
 * its dump file. */
size_t mav1hip_decision_record_bytes(const Mav1FrameParams *fp);
int mav1hip_decision_record(mav1hip_ctx *c, const Mav1FrameParams *fp, const Mav1BlockInfo *blocks,
                            const Mav1SbInfo *sbs, const int32_t hdr7[7], uint8_t *out);

/* ------------------------------------------------------------------------
 * Device-resident form
 * ------------------------------------------------------------------------ */
enum {
  MAV1HIP_BUF_SRC = 0,   /* padded source frame */
  MAV1HIP_BUF_REF = 1,   /* LAST reference picture */
  MAV1HIP_BUF_GOLD = 2,  /* second reference picture */
  MAV1HIP_BUF_MVFIELD = 3,
  MAV1HIP_BUF_MVGOLD = 4,
  MAV1HIP_BUF_PLANES = 5, /* motion-search planes: per reference 16 quarter-pel planes + decimated
                            plane (mav1hip_planes_entry_bytes each), LAST then second reference,
                            then the decimated source */
  MAV1HIP_BUF_ACOSTS = 6,  /* pass-1 analysis costs, int64 per item */
  MAV1HIP_BUF_AINFOS = 7,  /* pass-1 analysis decisions */
  MAV1HIP_BUF_BLOCKS = 8,  /* coded blocks, 64 records per superblock */
  MAV1HIP_BUF_SBS = 9,     /* superblock records */
  MAV1HIP_BUF_REC = 10,    /* reconstruction */
  MAV1HIP_BUF_COEFS = 11,  /* quantised levels, 6144 int32 per superblock */
  MAV1HIP_BUF_LFMASK = 12, /* per-mi filter info */
  MAV1HIP_BUF_CDEFBUF = 13, /* CDEF outputs, MAV1HIP_MAX_CDEF_SETS frames */
  MAV1HIP_BUF_DBK = 14,    /* deblocked frame kept for loop restoration */
  MAV1HIP_BUF_LROUT = 15,  /* loop-restoration output */
  MAV1HIP_BUF_LRUNITS = 16, /* Mav1LrUnit per restoration unit */
  MAV1HIP_BUF_CDEF_CHOICE = 17, /* CDEF set per superblock, int8 */
  MAV1HIP_BUF_DECGRID = 18 /* decision cells, 8 bytes per 8x8 */
};
#define MAV1HIP_MAX_CDEF_SETS 8
/* Size in bytes of a device buffer. */
size_t mav1hip_buffer_bytes(const mav1hip_ctx *c, int32_t buf);
size_t mav1hip_planes_entry_bytes(const mav1hip_ctx *c);
int mav1hip_upload(mav1hip_ctx *c, int32_t buf, const void *host, size_t bytes);
int mav1hip_download(mav1hip_ctx *c, int32_t buf, void *host, size_t bytes);
/* Parameters of the next launches (copied; the pointer is not kept). */
int mav1hip_set_params(mav1hip_ctx *c, const Mav1FrameParams *fp);
int mav1hip_launch_motion_search(mav1hip_ctx *c);
int mav1hip_launch_analyze(mav1hip_ctx *c);
int mav1hip_launch_edge_types(mav1hip_ctx *c);
int mav1hip_launch_recon(mav1hip_ctx *c);
/* Filter stages on the device buffers: BLOCKS -> LFMASK; REC deblocked in
 * place; REC -> CDEFBUF (nsets); REC copied to DBK; CDEFBUF -> REC and
 * CDEF_CHOICE; REC (CDEF output) + DBK + LRUNITS -> LROUT with the
 * restoration types of the last mav1hip_launch_lr_select, or of the frame
 * parameters when LRUNITS was uploaded (mav1hip_set_params copies them). */
int mav1hip_launch_filter_info(mav1hip_ctx *c);
int mav1hip_launch_deblock(mav1hip_ctx *c);
int mav1hip_launch_cdef(mav1hip_ctx *c, int32_t nsets);
int mav1hip_launch_save_deblocked(mav1hip_ctx *c);
int mav1hip_launch_cdef_select(mav1hip_ctx *c, int32_t nsets);
/* REC (CDEF output) + DBK + SRC -> LRUNITS and the per-plane restoration
 * types that mav1hip_launch_restore applies. */
int mav1hip_launch_lr_select(mav1hip_ctx *c, int32_t lr_test);
int mav1hip_launch_restore(mav1hip_ctx *c);
/* BLOCKS -> DECGRID (the cells of the decision record). */
int mav1hip_launch_decision_grid(mav1hip_ctx *c);
/* Waits for everything enqueued on the context's stream. */
int mav1hip_sync(mav1hip_ctx *c);

/* A pool of device-resident frames (the benchmark keeps its whole input
 * sequence in HBM): mav1hip_pool_alloc reserves n padded frames,
 * mav1hip_pool_upload fills one, mav1hip_bind_frames points the SRC / REF /
 * GOLD buffers at pool frames (-1: the context's own buffer). */
int mav1hip_pool_alloc(mav1hip_ctx *c, int32_t n);
int mav1hip_pool_upload(mav1hip_ctx *c, int32_t idx, const uint8_t *frame);
int mav1hip_bind_frames(mav1hip_ctx *c, int32_t src_idx, int32_t ref_idx, int32_t gold_idx);
/* Copies a device frame buffer (MAV1HIP_BUF_REC or MAV1HIP_BUF_LROUT) into
 * pool frame idx on the context's stream: a reconstructed picture becomes a
 * reference for the frames that follow. */
int mav1hip_pool_store(mav1hip_ctx *c, int32_t idx, int32_t buf);
/* Copies pool frame src_idx of context src into pool frame dst_idx of
 * context dst (same device and frame geometry), ordered after all work
 * enqueued on src so far and before the later work of both contexts.
 * Stands in for an engine's slot copy (encoder.c:1088-1118, the reference
 * frames copied into the engine's ref / gold buffers before its frame
 * runs): frame engines are separate contexts, each with its own stream, and
 * a reference picture passes from the engine that coded it to the engine
 * that reads it. Calls for one pair of contexts must not overlap. */
int mav1hip_pool_import(mav1hip_ctx *dst, int32_t dst_idx, mav1hip_ctx *src, int32_t src_idx);

/* ------------------------------------------------------------------------
 * Sequence driver: the encoder's frame driver on HIP streams and events
 * ------------------------------------------------------------------------
 * The queue interface of the encoder (mav1_send_frame /
 * mav1_receive_packet, src/mav1.h:55-64) for the backend path: frames go in
 * in display order, the driver forms the ALTREF mini-GOPs and the coding
 * order of src/encoder.c:1252-1293 (alt_drain, pyr_code: key frames every
 * keyint, hidden ALTREF frames, pyramid middle frames, leaf frames, a
 * regular frame when a mini-GOP cannot be completed), picks each frame's
 * parameters (alt_code, :1004-1030; setup_params) and runs every stage of
 * run_frame_core (:598-647) on the device. Reference pictures stay in HBM:
 * each of the 8 reference slots maps to a picture on the engine that coded
 * it, and an engine that reads a slot another engine refreshed copies it
 * once (the engine slot copy of :1088-1118, ordered by events). Two frame
 * engines (contexts with their own streams) code a stream as the
 * reference's engines do (:1150-1224): the key frames, regular frames and
 * the ALTREF chain on the first, the other frames on the second, each
 * running beside the next ALTREF's reconstruction. The outputs the CPU side
 * reads (the entropy coder's inputs and the final picture) are staged on
 * the device and copied to pinned host memory on a copy stream.
 *
 * Only ALTREF configurations (altref >= 2, which --turbo implies) are
 * driven; altref 0 returns NULL from mav1hip_seq_create. */
typedef struct mav1hip_seq mav1hip_seq;

/* One coded frame, in coding order (the order its OBU is written). The
 * pointers are pinned host buffers owned by the driver, valid until the next
 * mav1hip_seq_receive or mav1hip_seq_destroy. */
typedef struct {
  int32_t coded_index;   /* 0, 1, 2, ... in coding order */
  int32_t disp;          /* display index of the frame */
  int32_t is_key, frame_kind, nonref, hidden;
  int32_t order_hint;
  int32_t refresh_frame_flags;
  int32_t last_slot, arf_slot;  /* reference slots read as LAST / ALTREF (-1: none) */
  int32_t last_hint, arf_hint;  /* their order hints (-1: none) */
  int32_t engine;               /* the frame engine that coded it */
  Mav1FrameParams fp;           /* parameters, lr_type as the restoration search chose */
  Mav1HipFilterPlan plan;       /* the in-loop filter stages that ran */
  const Mav1BlockInfo *blocks;  /* num_sbs * 64 records (sbs[s] lists superblock s's) */
  const Mav1SbInfo *sbs;        /* num_sbs */
  const int32_t *coefs;         /* num_sbs * 6144 quantised levels */
  const int8_t *cdef_choice;    /* num_sbs (0 when the frame skips CDEF) */
  const Mav1LrUnit *lr_units;   /* one record per unit when restoration ran */
  const uint8_t *recon;         /* final picture (padded layout, frame_bytes) */
  const uint8_t *decision;      /* decision record (mav1hip_decision_record layout, tile_bytes 0) or NULL */
  size_t decision_bytes;
} Mav1HipCodedFrame;

/* Creates the driver with its two engines for frames of cfg's size.
 * lr_test: Mav1Config.lr_test. dump_decisions: produce the per-frame
 * decision record (MAV1_DUMP_DEC). NULL and a message in err on failure. */
mav1hip_seq *mav1hip_seq_create(const Mav1HipConfig *cfg, int32_t lr_test, int32_t dump_decisions, char *err,
                                size_t errlen);
void mav1hip_seq_destroy(mav1hip_seq *s);
const char *mav1hip_seq_error(const mav1hip_seq *s);
/* Sends one I420 frame (planes / strides as mav1_send_frame; the frame is
 * copied at once), or ends the stream with planes == NULL. Frames whose
 * mini-GOP is complete are enqueued on the device at once. Returns
 * MAV1HIP_EBUSY when every output slot holds a frame not yet received: the
 * frame was not taken; receive and send it again. */
#define MAV1HIP_EBUSY (-5)
int mav1hip_seq_send_frame(mav1hip_seq *s, const uint8_t *const planes[3], const int32_t strides[3]);
/* Sends a frame already resident in device memory in the padded layout
 * (frame_bytes, edges replicated as load_source does, src/encoder.c:297-310).
 * It must stay unchanged until the frame is received. */
int mav1hip_seq_send_device_frame(mav1hip_seq *s, const uint8_t *frame);
/* The next coded frame in coding order: 0 and *out filled, 1 when none is
 * ready (wait == 0) or none is left (the stream ended and every frame was
 * received), a negative code on failure. wait != 0 blocks until the next
 * frame's outputs are in host memory. */
int mav1hip_seq_receive(mav1hip_seq *s, Mav1HipCodedFrame *out, int32_t wait);
/* The context of engine i (0: key frames, regular frames and the ALTREF
 * chain; 1: the other frames), for the timing calls below; owned by the
 * driver. NULL for another index. */
mav1hip_ctx *mav1hip_seq_engine(mav1hip_seq *s, int32_t i);

/* ------------------------------------------------------------------------
 * Batched streams: several sequences coded together on one GPU
 * ------------------------------------------------------------------------
 * No counterpart in the reference, which batches only within a stream (its
 * engines, src/encoder.c:1150-1224): n independent streams, each with its
 * own configuration (sizes may differ) and its own driver state, engines,
 * reference pictures and outputs, as n mav1hip_seq would have. Their frames
 * at the same position of their coding order are coded as one round: every
 * stage is one launch over all of the round's frames (a table of per-frame
 * parameters and buffers in device memory; the persistent inter and intra
 * phases take work from every frame of the round). Each stream's coded
 * frames equal what its own mav1hip_seq codes from the same frames.
 * Rounds form when every stream that can still code has a frame ready:
 * feed the streams in step (frame t to every stream, then frame t + 1).
 * A receive that waits runs the frames that are ready without waiting for
 * the other streams. At most 32 streams. */
typedef struct mav1hip_batch mav1hip_batch;
mav1hip_batch *mav1hip_batch_create(const Mav1HipConfig *cfgs, int32_t n, int32_t lr_test, int32_t dump_decisions,
                                    char *err, size_t errlen);
void mav1hip_batch_destroy(mav1hip_batch *b);
const char *mav1hip_batch_error(const mav1hip_batch *b);
/* As mav1hip_seq_send_frame / _send_device_frame / _receive for stream
 * `stream` (0 .. n - 1). */
int mav1hip_batch_send_frame(mav1hip_batch *b, int32_t stream, const uint8_t *const planes[3],
                             const int32_t strides[3]);
int mav1hip_batch_send_device_frame(mav1hip_batch *b, int32_t stream, const uint8_t *frame);
int mav1hip_batch_receive(mav1hip_batch *b, int32_t stream, Mav1HipCodedFrame *out, int32_t wait);
/* Launch shapes of the batched form, chosen by mav1-encoder_amd/autotune.py
 * (0 = the default): "intra_wg_key" / "intra_wg_inter", the persistent
 * workgroups of the intra phase in rounds of intra frames (default: one per
 * CU) and of inter frames (default: a quarter of the CUs); "inter_wg", those
 * of the inter phase (default: two per CU). Results do not depend on them. */
int mav1hip_batch_set_knob(mav1hip_batch *b, const char *name, int32_t value);
/* Pruning of the batched analysis by the learned pruner (the network of
 * mav1hip_pruner_*, parameters as mav1hip_pruner_set_params takes them):
 * per superblock wholly inside its frame, the analysis evaluates no level
 * deeper than the smallest l whose cumulative partition probability reaches
 * 1 - tau (the exhaustive search where the network is not that confident).
 * params NULL or tau <= 0 turns it off (the default: outputs then equal the
 * sequence driver's). Changes the coded decisions when on. */
int mav1hip_batch_set_pruner(mav1hip_batch *b, const float *params, float tau);
/* The sequence of stream i (owned by the batch; for mav1hip_seq_engine). */
mav1hip_seq *mav1hip_batch_stream(mav1hip_batch *b, int32_t stream);

/* ------------------------------------------------------------------------
 * Learned partition / mode pruner (training and inference on the GPU)
 * ------------------------------------------------------------------------
 * No counterpart in the reference, whose analysis searches every partition
 * slot and inter candidate (src/kernels/k_analysis.h:256-311, 656-808). A
 * small CNN per 64x64 superblock: input the superblock's luma / 255 and its
 * residual against the co-located reference block / 255 (64 x 64 x 2
 * floats, channels innermost) and the frame's quantiser q_idx / 255;
 * three 3x3 stride-2 convolutions (32, 64, 64 channels, ReLU) on BF16
 * operands with FP32 accumulation on the matrix cores, mean pooling, and
 * two heads: MAV1HIP_PRUNER_PART_CLASSES partition logits and
 * MAV1HIP_PRUNER_MODE_CLASSES mode logits (labels: mav1-encoder_amd/
 * pruner_data.py). Parameters are one FP32 vector of
 * mav1hip_pruner_num_params() values (layout in csrc/pruner.hip). */
#define MAV1HIP_PRUNER_PART_CLASSES 4
#define MAV1HIP_PRUNER_MODE_CLASSES 8
typedef struct mav1hip_pruner mav1hip_pruner;
int32_t mav1hip_pruner_num_params(void);
/* Device buffers for batches of up to max_batch superblocks. */
mav1hip_pruner *mav1hip_pruner_create(int32_t max_batch, char *err, size_t errlen);
void mav1hip_pruner_destroy(mav1hip_pruner *p);
const char *mav1hip_pruner_error(const mav1hip_pruner *p);
/* Parameters in and out (host FP32); setting them resets the optimizer. */
int mav1hip_pruner_set_params(mav1hip_pruner *p, const float *params);
int mav1hip_pruner_get_params(mav1hip_pruner *p, float *params);
/* A batch: x batch x 64 x 64 x 2 floats, q batch floats, labels batch x 2
 * (partition class, mode class; -1 = no label; NULL = none). */
int mav1hip_pruner_load(mav1hip_pruner *p, const float *x, const float *q, const int32_t *labels, int32_t batch);
/* Forward pass of the loaded batch: logits (batch x 12: 4 partition, 8
 * mode) and the mean loss (sum of the two heads' softmax cross-entropies)
 * into host memory when not NULL. */
int mav1hip_pruner_forward(mav1hip_pruner *p, float *logits, float *loss);
/* Gradient of the mean loss for every parameter (after a forward pass);
 * into host memory when not NULL. */
int mav1hip_pruner_backward(mav1hip_pruner *p, float *grads);
/* One Adam step (step = 1, 2, ...) with the last gradients. */
int mav1hip_pruner_adam(mav1hip_pruner *p, float lr, float beta1, float beta2, float eps, int32_t step);
int mav1hip_pruner_sync(mav1hip_pruner *p);

/* ------------------------------------------------------------------------
 * Input side and quality measure (host only)
 * ------------------------------------------------------------------------ */
/* The y4m reader of the command-line front end (tools/mav1enc.c:67-77):
 * opens a YUV4MPEG2 file, reads W, H and F (frame rate num:den, default
 * 30:1) from its header and refuses colour spaces other than 4:2:0. NULL
 * and a message in err on failure. */
typedef struct mav1hip_y4m mav1hip_y4m;
mav1hip_y4m *mav1hip_y4m_open(const char *path, int32_t *width, int32_t *height, int32_t *fps_num, int32_t *fps_den,
                              char *err, size_t errlen);
/* Reads the next frame into I420 planes (rows of the visible width at the
 * given strides): 0 on success, 1 at the end of the file, MAV1HIP_EINVAL
 * on a truncated or malformed frame. */
int mav1hip_y4m_read(mav1hip_y4m *y, uint8_t *const planes[3], const int32_t strides[3]);
void mav1hip_y4m_close(mav1hip_y4m *y);
/* PSNR of n samples (tools/mav1enc.c:22-34): 10 log10(255^2 n / SSE) with
 * the SSE summed exactly in integers; 99 when the samples are equal. */
double mav1hip_psnr(const uint8_t *a, const uint8_t *b, size_t n);

/* Per-kernel timing with HIP events recorded on the context's stream
 * around every launch while enabled. mav1hip_timing_get returns the number
 * of launches of the named kernel and their total duration in ms since the
 * last reset (synchronises the stream). */
int mav1hip_timing_enable(mav1hip_ctx *c, int32_t on);
/* Restricts the events to the launches of one kernel (NULL: every kernel);
 * each timed launch adds two event records to its stream. */
int mav1hip_timing_select(mav1hip_ctx *c, const char *kernel);
int mav1hip_timing_reset(mav1hip_ctx *c);
int mav1hip_timing_get(mav1hip_ctx *c, const char *kernel, int32_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif
