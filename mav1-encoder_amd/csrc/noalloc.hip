/* The loop-restoration search without allocation (mav1hip_lr_select_static):
 * the same kernels as mav1hip_launch_lr_select, on device buffers the
 * caller owns and a stream the caller names, with the search's scratch
 * (per-unit fit statistics, the per-plane chain state and the rerun counts)
 * in one file-scope device array instead of context buffers.
 *
 * The array's device address is taken once per device with
 * hipGetSymbolAddress (the host-side name of a __device__ array is not a
 * device address) and cleared once then; every later call only launches
 * kernels. All calls on one device share the array: calls that overlap on
 * one device race. */
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdio.h>

#include "../../include/mav1_hip.h"
#include "internal.h"

namespace {

constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
/* statistics of MAV1HIP_STATIC_LR_UNITS units, then the chain state, then
 * the rerun counts */
constexpr size_t kStatsBytes = align_up((size_t)MAV1HIP_STATIC_LR_UNITS * LR_STAT_MAX_BYTES);
constexpr size_t kChainBytes = align_up(LR_CHAIN_MAX_BYTES);
constexpr size_t kScratchBytes = kStatsBytes + kChainBytes + kAlign;

constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct DeviceScratch {
  unsigned char *ptr = nullptr; /* scratch's address on this device */
  int gen = 0;                  /* generation of the last fit launch on it */
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

int fail(char *err, size_t errlen, int code, const char *msg, hipError_t e = hipSuccess) {
  if (err && errlen) {
    if (e != hipSuccess)
      snprintf(err, errlen, "%s: %s", msg, hipGetErrorString(e));
    else
      snprintf(err, errlen, "%s", msg);
  }
  return code;
}

/* The current device's scratch address, taken and cleared on its first use. */
int device_scratch(DeviceScratch **out, char *err, size_t errlen) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(err, errlen, MAV1HIP_EDEVICE, "lr_select_static: hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) return fail(err, errlen, MAV1HIP_EDEVICE, "lr_select_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceScratch &d = g_dev[dev];
  if (!d.ptr) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(err, errlen, MAV1HIP_EDEVICE, "lr_select_static: hipGetSymbolAddress", e);
    /* the fit's generation numbers start from zeroed statistics */
    e = hipMemset(p, 0, kScratchBytes);
    if (e != hipSuccess) return fail(err, errlen, MAV1HIP_EDEVICE, "lr_select_static: clearing the scratch", e);
    d.ptr = (unsigned char *)p;
  }
  *out = &d;
  return MAV1HIP_OK;
}

}  // namespace

int mav1hip_lr_select_static(const Mav1FrameParams *fp, const uint8_t *src, const uint8_t *cdef, const uint8_t *dbk,
                             int32_t lr_test, Mav1LrUnit *units, int32_t *lr_type, void *stream, char *err,
                             size_t errlen) {
  if (!fp || !src || !cdef || !dbk || !units || !lr_type)
    return fail(err, errlen, MAV1HIP_EINVAL, "lr_select_static: missing buffer");
  int n = 0;
  for (int p = 0; p < 3; p++) {
    if (fp->lr_unit_size[p] != 64 || fp->lr_unit_rows[p] < 1 || fp->lr_unit_cols[p] < 1 || fp->lr_unit_off[p] != n)
      return fail(err, errlen, MAV1HIP_EINVAL, "lr_select_static: the unit layout is not the 64x64 one");
    n += fp->lr_unit_rows[p] * fp->lr_unit_cols[p];
  }
  if (n > MAV1HIP_STATIC_LR_UNITS) {
    char msg[160];
    snprintf(msg, sizeof msg, "lr_select_static: %d restoration units, the static scratch holds %d", n,
             MAV1HIP_STATIC_LR_UNITS);
    return fail(err, errlen, MAV1HIP_EINVAL, msg);
  }
  if (lr_stat_bytes() > LR_STAT_MAX_BYTES || lr_chain_bytes() > LR_CHAIN_MAX_BYTES)
    return fail(err, errlen, MAV1HIP_EDEVICE, "lr_select_static: scratch layout out of date");
  DeviceScratch *d = nullptr;
  int rc = device_scratch(&d, err, errlen);
  if (rc) return rc;
  unsigned char *stats = d->ptr, *chain = d->ptr + kStatsBytes;
  int *reruns = (int *)(d->ptr + kStatsBytes + kChainBytes);
  int gen;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    d->gen = d->gen % 0x7ffffffe + 1;
    gen = d->gen;
  }
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = launch_lr_fit(st, *fp, src, cdef, dbk, stats, chain, n, gen);
  /* the rounds of mav1hip_launch_lr_select: chain launches that stop at a
   * reference change, each followed by the refinement of the remaining
   * units; the last refines any further mismatching unit itself */
  const int LR_ROUNDS = 3;
  for (int r = 0; r < LR_ROUNDS && e == hipSuccess; r++) {
    if (r > 0) e = launch_lr_path(st, *fp, src, cdef, dbk, stats, chain, n);
    if (e == hipSuccess)
      e = launch_lr_chain(st, *fp, src, cdef, dbk, stats, units, lr_type, lr_test != 0, reruns, chain,
                          r < LR_ROUNDS - 1);
  }
  if (e != hipSuccess) return fail(err, errlen, MAV1HIP_EDEVICE, "lr_select_static: launch", e);
  return MAV1HIP_OK;
}
