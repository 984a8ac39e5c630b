/* Batched launches: one launch of a stage over the frames of several
 * sequences (streams of possibly different sizes) coded at the same
 * position of their coding order.
 *
 * Each frame's parameters and device buffers are one BatchFrame in a table
 * in device memory, written once per batch of frames. A launch splits its
 * workgroups over the frames with a BatchGrid passed by value: frame f owns
 * the workgroups start[f] .. start[f + 1] - 1, which it reads as a grid of
 * gx[f] x gy[f] x (rest) workgroups, exactly the grid the single-frame
 * launch of that frame would use. The kernel bodies are the single-frame
 * ones, called with the frame's table entry and its own block index, so a
 * frame's results do not depend on the frames batched with it. */
#ifndef MAV1_BATCH_H
#define MAV1_BATCH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mav1_types.h"

constexpr int kMaxBatch = 32;

struct BatchFrame {
  Mav1FrameParams fp;
  /* frames the stages read */
  const uint8_t *src, *ref, *gold;
  /* motion search */
  uint8_t *planes;
  long long entry_bytes;
  int32_t *mvf, *mvg;
  /* analysis */
  long long *acosts;
  Mav1BlockInfo *ainfos, *blocks;
  Mav1SbInfo *sbs;
  /* reconstruction */
  uint8_t *rec;
  int32_t *coefs;
  const int32_t *mats;
  uint8_t *ys, *uvs;
  int *rwork, *iwork;
  /* filters and restoration */
  uint8_t *lfmask, *cdefbuf, *dbk, *lrout, *decgrid;
  int *dirvar;
  int8_t *choice;
  Mav1LrUnit *lrunits;
  void *lrstats, *lrchain;
  int *lrtype, *lrreruns;
  long long frame_bytes;
  int lr_units, lr_gen;
  /* the deepest analysis level per superblock the pruner allows (NULL: no
   * pruning) */
  int8_t *prune;
};

struct BatchGrid {
  int n;
  int start[kMaxBatch + 1];
  int gx[kMaxBatch], gy[kMaxBatch];
  int aux[kMaxBatch]; /* a per-frame launch argument (slots per node, nodes per reference, ...) */
};

/* Host: frame f's single-frame grid gx x gy x gz, with its argument. */
inline void batch_grid_add(BatchGrid &g, int gx, int gy, int gz, int aux = 0) {
  if (g.n == 0) g.start[0] = 0;
  g.gx[g.n] = gx;
  g.gy[g.n] = gy;
  g.aux[g.n] = aux;
  g.start[g.n + 1] = g.start[g.n] + gx * gy * gz;
  g.n++;
}
inline BatchGrid batch_grid() {
  BatchGrid g;
  g.n = 0;
  g.start[0] = 0;
  return g;
}

#ifdef __HIPCC__
struct BatchIdx {
  int f, x, y, z;
};
/* The frame that owns workgroup b and the workgroup's index in that frame's
 * grid (the frame count is small: a scan). */
__device__ __forceinline__ BatchIdx batch_idx(const BatchGrid &g, int b) {
  int f = 0;
  while (f + 1 < g.n && b >= g.start[f + 1]) f++;
  const int l = b - g.start[f], gx = g.gx[f], gy = g.gy[f];
  BatchIdx r;
  r.f = f;
  r.x = l % gx;
  r.y = (l / gx) % gy;
  r.z = l / (gx * gy);
  return r;
}
__device__ __forceinline__ BatchIdx batch_idx(const BatchGrid &g) { return batch_idx(g, blockIdx.x); }
/* A table pointer as a global-memory pointer (the table holds device
 * allocations), rebuilt from its integer value so the compiler keeps the
 * address space and emits global rather than flat instructions. */
template <class T>
__device__ __forceinline__ T *gptr(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}
#endif

/* Batched launchers: T is the table in device memory, H the same table on
 * the host (for the grids), n the frame count. Each enqueues on `st` what
 * the single-frame launcher of the same name enqueues, for every frame. */
hipError_t launch_qpel_build_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int which);
hipError_t launch_dec_build_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int which);
hipError_t launch_me_level_b(hipStream_t st, int level, const BatchFrame *T, const BatchFrame *H, int n, int nref);
hipError_t launch_analyze_b(hipStream_t st, int level, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_analyze_table_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_edge_types_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_recon_inter_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int workgroups);
/* The scheduling pass of every frame, then the batch's ready queue (q: 4 +
 * qcap ints, qcap >= the round's superblocks) reset and filled with the
 * superblocks that depend on none. */
hipError_t launch_intra_sched_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int *q, int qcap);
hipError_t launch_recon_intra_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int *q,
                                int workgroups);
hipError_t launch_filter_info_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_deblock_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_cdef_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int nsets);
hipError_t launch_cdef_select_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int nsets);
hipError_t launch_lr_apply_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_decision_grid_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_lr_fit_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_lr_path_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n);
hipError_t launch_lr_chain_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int lrTest, int stop);
/* Frame copies and clears of the driver: which 0 = reconstruction into the
 * deblocked copy, 1 = reconstruction into the restoration output, 2 = the
 * CDEF choices cleared. */
hipError_t launch_frame_op_b(hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, int which);

/* The pruner on the superblocks of the round's frames (pruner.hip): every
 * superblock's sample built from the frame's source and LAST reference in
 * HBM, the network's forward pass on `st`, and into each frame's prune[]
 * the deepest analysis level: the smallest level l whose cumulative
 * partition probability reaches 1 - tau (3, no pruning, when none does or
 * the superblock is not wholly inside the frame). */
struct mav1hip_pruner;
int pruner_infer_frames(mav1hip_pruner *p, hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, float tau);

#endif
