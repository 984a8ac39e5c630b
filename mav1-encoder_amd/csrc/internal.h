/* Launchers of the gfx950 kernels, called by the C-ABI layer (api.hip).
 * Each enqueues its kernels on `st` and returns the launch status. */
#ifndef MAV1_INTERNAL_H
#define MAV1_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mav1_hip.h"
#include "../../include/mav1_types.h"
#include "batch.h"

/* Motion-search planes of one reference: 16 quarter-pel planes
 * (frame_w x frame_h each) followed by the 4x4-decimated plane. */
long me_entry_bytes(const Mav1FrameParams &fp);
size_t me_dec_bytes(const Mav1FrameParams &fp);

hipError_t launch_qpel_build(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *ref, uint8_t *planes16);
hipError_t launch_dec_build(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *frame, uint8_t *dec);
/* One level of the node search for nref references (LAST: ref0 / field0 /
 * entry 0; second reference: ref1 / field1 / entry 1). */
hipError_t launch_me_level(hipStream_t st, int level, const Mav1FrameParams &fp, const uint8_t *src,
                           const uint8_t *ref0, const uint8_t *ref1, const uint8_t *planes, long entryBytes,
                           const uint8_t *srcDec, int32_t *field0, int32_t *field1, int nref);

/* Analysis: pass 1 of one level (every slot of every node), then pass 2. */
int analysis_level_slots(const Mav1FrameParams &fp, int level);
hipError_t launch_analyze(hipStream_t st, int level, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *ref,
                          const uint8_t *gold, const int32_t *mvf, const int32_t *mvg, long long *costs,
                          Mav1BlockInfo *infos);
hipError_t launch_analyze_table(hipStream_t st, const Mav1FrameParams &fp, const long long *costs,
                                const Mav1BlockInfo *infos, Mav1BlockInfo *blocks, Mav1SbInfo *sbs);

/* Reconstruction: inter blocks (all at once), then the intra wavefront. */
hipError_t launch_edge_types(hipStream_t st, const Mav1FrameParams &fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                             uint8_t *ys, uint8_t *uvs);
/* work: 2 + num_sbs * 64 ints of device scratch (the inter block list);
 * workgroups: the persistent grid. */
hipError_t launch_recon_inter(hipStream_t st, const Mav1FrameParams &fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                              const uint8_t *src, uint8_t *rec, const uint8_t *ref, const uint8_t *gold,
                              int32_t *coefs, const int32_t *mats, int *work, int workgroups);
/* Intra phase work space (ints) and whether the frame fits its scheduler. */
size_t intra_work_ints(const Mav1FrameParams &fp);
int intra_sched_fits(const Mav1FrameParams &fp);
int intra_workgroups(const Mav1FrameParams &fp, int num_cu);
/* The intra phase: the scheduling pass, then `workgroups` persistent
 * workgroups over the superblocks holding intra blocks; iwork[2] becomes
 * non-zero if a dependency wait ran out. */
hipError_t launch_intra_sched(hipStream_t st, const Mav1FrameParams &fp, const Mav1BlockInfo *blocks,
                              const Mav1SbInfo *sbs, int *iwork);
hipError_t launch_recon_intra(hipStream_t st, const Mav1FrameParams &fp, Mav1BlockInfo *blocks, const Mav1SbInfo *sbs,
                              const uint8_t *src, uint8_t *rec, const uint8_t *ref, const uint8_t *gold,
                              int32_t *coefs, const int32_t *mats, int *iwork, int workgroups);

/* filters.hip */
hipError_t launch_filter_info(hipStream_t st, const Mav1FrameParams &fp, const Mav1BlockInfo *blocks,
                              const Mav1SbInfo *sbs, uint8_t *info);
hipError_t launch_deblock(hipStream_t st, const Mav1FrameParams &fp, uint8_t *rec, const uint8_t *info);
hipError_t launch_cdef(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *cur, uint8_t *out,
                       long long frameBytes, const uint8_t *info, int *dirvar, int n);
hipError_t launch_cdef_select(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdefbuf,
                              long long frameBytes, int n, uint8_t *rec, int8_t *choice);
hipError_t launch_lr_apply(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *cdef, const uint8_t *dbk,
                           uint8_t *out, const Mav1LrUnit *units, const int *lrType);

hipError_t launch_decision_grid(hipStream_t st, const Mav1FrameParams &fp, const Mav1BlockInfo *blocks,
                                const Mav1SbInfo *sbs, uint8_t *grid);

/* lrsearch.hip */
/* Upper bounds of lr_stat_bytes() and lr_chain_bytes() (the static scratch
 * of mav1hip_lr_select_static is laid out with them). */
#define LR_STAT_MAX_BYTES 384
#define LR_CHAIN_MAX_BYTES 192
size_t lr_stat_bytes();
size_t lr_chain_bytes();
/* The fit of every unit with its refinement under the default reference,
 * and the chain state reset. gen: a number that differs from the previous
 * launch's on the same statistics buffer (the buffer starts zeroed). */
hipError_t launch_lr_fit(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdef,
                         const uint8_t *dbk, void *stats, void *chain, int nunits, int gen);
/* Refinement of the units from each plane's chain restart point on. */
hipError_t launch_lr_path(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdef,
                          const uint8_t *dbk, void *stats, const void *chain, int nunits);
/* The unit chains; stop: end at the first reference mismatch instead of
 * refining the unit again. */
hipError_t launch_lr_chain(hipStream_t st, const Mav1FrameParams &fp, const uint8_t *src, const uint8_t *cdef,
                           const uint8_t *dbk, const void *stats, Mav1LrUnit *units, int *lrType, int lrTest,
                           int *reruns, void *chain, int stop);

/* api.hip: library-internal access to a context for the sequence driver */
/* A context on two streams the caller keeps (neither created nor destroyed
 * by the context). */
mav1hip_ctx *ctx_create_on_streams(const Mav1HipConfig *cfg, hipStream_t main, hipStream_t side, char *err,
                                   size_t errlen);
hipStream_t ctx_stream(mav1hip_ctx *c);
hipStream_t ctx_side_stream(mav1hip_ctx *c);
int ctx_device(const mav1hip_ctx *c);
size_t ctx_frame_bytes(const mav1hip_ctx *c);
void *ctx_buffer(mav1hip_ctx *c, int buf, size_t *bytes);
const int *ctx_intra_error(const mav1hip_ctx *c);
/* restoration types of the last restoration search (int[3], device) */
const int *ctx_lrtype(const mav1hip_ctx *c);
/* Points the stage inputs at device frames (ref / gold NULL: the context's
 * own buffers). */
int ctx_bind_ptrs(mav1hip_ctx *c, uint8_t *src, uint8_t *ref, uint8_t *gold);
int ctx_store(mav1hip_ctx *c, uint8_t *dst, int buf);
int ctx_rec_to_lrout(mav1hip_ctx *c);
int ctx_clear_choice(mav1hip_ctx *c);
/* The context's frame as a batch table entry: its parameters (set by
 * mav1hip_set_params), bound inputs and buffers. new_lr_gen: take the next
 * restoration-fit generation, as a restoration search of this context does. */
int ctx_batch_frame(mav1hip_ctx *c, BatchFrame *bf, int new_lr_gen);
/* workgroups of the persistent inter and intra phases */
int ctx_num_cu(const mav1hip_ctx *c);

#endif
