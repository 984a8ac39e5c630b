/* Learned partition / mode pruner: a small CNN trained and run on the GPU
 * with hand-written BF16 MFMA kernels (FP32 accumulation).
 *
 * No counterpart in the reference, whose analysis searches every partition
 * slot (src/kernels/k_analysis.h:656-679, the DP of :697-808) and inter
 * candidate (:256-311) exhaustively. The network predicts, per 64x64
 * superblock, a 4-way partition class and an 8-way inter-mode class from
 * the superblock's luma and its residual against the co-located reference
 * block, and the frame's quantiser.
 *
 * Network (activations NHWC, channels innermost):
 *   x0  64x64x4   luma / 255, residual / 255 (channels 2, 3 zero)
 *   y1  32x32x32  ReLU(conv3x3 stride 2 pad 1 (x0) + b1)
 *   y2  16x16x64  ReLU(conv3x3 s2 p1 (y1) + b2)
 *   y3  8x8x64    ReLU(conv3x3 s2 p1 (y2) + b3)
 *   h   64        mean of y3 over the 8x8 positions + q * wq + bq
 *   partition logits h . Wp + bp (4), mode logits h . Wm + bm (8)
 *   loss: mean over the batch of the two softmax cross-entropies.
 * Conv operands (inputs, weights, the stored activations and the gradients
 * of the pre-activations) are BF16, rounded to nearest even; the products
 * accumulate in FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16); the
 * pooling, heads, loss and the optimizer run in FP32.
 *
 * Each convolution is an implicit GEMM, one wave per 16x16 output tile:
 *   forward      out[pixel][co]  = sum_k im2col(in)[pixel][k] W[co][k]
 *   data grad    din[pixel][ci]  = sum_k gather(dpre)[pixel][k] Wt[ci][k]
 *   weight grad  dW[co][k]       = sum_pixel dpre[pixel][co] im2col(in)[pixel][k]
 * with k = tap * C + c (tap = 3 ky + kx); the weight gradient splits the
 * pixel sum over slices and adds the slices in a fixed order, so every
 * result is deterministic. */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "../../include/mav1_hip.h"
#include "batch.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

/* parameter layout (FP32, flat): conv weights [co][tap][ci] */
constexpr int C0 = 2, C1 = 32, C2 = 64, C3 = 64, NP = 4, NM = 8, NF = 64;
constexpr int P_W1 = 0, P_B1 = P_W1 + C1 * 9 * C0, P_W2 = P_B1 + C1, P_B2 = P_W2 + C2 * 9 * C1, P_W3 = P_B2 + C2,
              P_B3 = P_W3 + C3 * 9 * C2, P_WQ = P_B3 + C3, P_BQ = P_WQ + NF, P_WP = P_BQ + NF, P_BP = P_WP + NF * NP,
              P_WM = P_BP + NP, P_BM = P_WM + NF * NM, P_N = P_BM + NM;
constexpr int CI0 = 4;              /* the input's channels in memory (2 used) */
constexpr int K1 = 48;              /* 9 * CI0 = 36 padded to a multiple of 16 */
constexpr int K2 = 9 * C1, K3 = 9 * C2;
constexpr int WG_SLICES = 16;       /* pixel slices of the weight gradient */

__device__ __forceinline__ uint16_t f2bf(float f) {
  /* round to nearest even on the upper 16 bits (finite inputs) */
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f4 mfma16(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

/* BF16 copies of the conv weights: W [co][Kp] with k = tap * Ci + ci (the
 * input's channels in memory), zero where there is no parameter; Wt
 * [ci][tap * Co + co] for the data gradients of layers 2 and 3. */
__global__ void k_pack(const float *__restrict__ P, uint16_t *w1, uint16_t *w2, uint16_t *w3, uint16_t *wt2,
                       uint16_t *wt3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < C1 * K1) {
    const int co = i / K1, k = i % K1, tap = k / CI0, ci = k % CI0;
    w1[i] = tap < 9 && ci < C0 ? f2bf(P[P_W1 + (co * 9 + tap) * C0 + ci]) : 0;
  }
  if (i < C2 * K2) {
    const int co = i / K2, k = i % K2, tap = k / C1, ci = k % C1;
    const uint16_t v = f2bf(P[P_W2 + (co * 9 + tap) * C1 + ci]);
    w2[i] = v;
    wt2[ci * (9 * C2) + tap * C2 + co] = v;
  }
  if (i < C3 * K3) {
    const int co = i / K3, k = i % K3, tap = k / C2, ci = k % C2;
    const uint16_t v = f2bf(P[P_W3 + (co * 9 + tap) * C2 + ci]);
    w3[i] = v;
    wt3[ci * (9 * C3) + tap * C3 + co] = v;
  }
}

/* x [b][64][64][2] float -> x0 [b][64][64][4] bf16 */
__global__ void k_input(const float *__restrict__ x, uint16_t *__restrict__ x0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x0[4 * i + 0] = f2bf(x[2 * i + 0]);
  x0[4 * i + 1] = f2bf(x[2 * i + 1]);
  x0[4 * i + 2] = 0;
  x0[4 * i + 3] = 0;
}

/* Forward convolution, 3x3 stride 2 pad 1, bias and ReLU fused; out BF16.
 * Block: 4 waves, each 16 output pixels x all NT * 16 output channels (the
 * pixels' im2col fragment gathered once per k step for every channel tile);
 * grid M / 64. Ci % 4 == 0 (the 4 consecutive k of a lane share one tap). */
template <int NT>
__global__ __launch_bounds__(256) void k_conv_fwd(const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias, uint16_t *__restrict__ out, int B,
                                                  int H, int W, int Ci, int Kp) {
  constexpr int Co = NT * 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int Ho = H / 2, Wo = W / 2, M = B * Ho * Wo;
  const int m0 = (blockIdx.x * 4 + wave) * 16;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  const int m = m0 + r;
  const int b = m / (Ho * Wo), rem = m - b * (Ho * Wo), oy = rem / Wo, ox = rem - oy * Wo;
  const int kreal = 9 * Ci;
  f4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kp; k0 += 16) {
    const int k = k0 + kq;
    s4 a = {0, 0, 0, 0};
    if (m < M && k < kreal) {
      const int tap = k / Ci, ci = k - tap * Ci;
      const int iy = 2 * oy - 1 + tap / 3, ix = 2 * ox - 1 + tap % 3;
      if (iy >= 0 && iy < H && ix >= 0 && ix < W) a = *(const s4 *)(in + (((size_t)b * H + iy) * W + ix) * Ci + ci);
    }
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = mfma16(a, *(const s4 *)(w + (size_t)(t * 16 + r) * Kp + k), acc[t]);
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int co = t * 16 + r;
    const float bb = bias[co];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int om = m0 + (lane >> 4) * 4 + j;
      if (om < M) {
        const float v = acc[t][j] + bb;
        out[(size_t)om * Co + co] = f2bf(v > 0.f ? v : 0.f);
      }
    }
  }
}

/* Data gradient of a 3x3 stride 2 pad 1 convolution, times the ReLU mask of
 * its input (the activation below): dbelow[pixel][ci] = (sum over co, taps
 * of dpre[out pixel][co] Wt[ci][tap * Co + co]) * (ybelow > 0), BF16.
 * M = B H W input pixels, N = Ci, K = 9 Co. Co % 4 == 0, Ci % 16 == 0. */
__global__ __launch_bounds__(256) void k_conv_dgrad(const uint16_t *__restrict__ dpre, const uint16_t *__restrict__ wt,
                                                    const uint16_t *__restrict__ ybelow, uint16_t *__restrict__ dbelow,
                                                    int B, int H, int W, int Ci, int Co) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int Ho = H / 2, Wo = W / 2, M = B * H * W, K = 9 * Co;
  const int m0 = (blockIdx.x * 4 + wave) * 16, n0 = blockIdx.y * 16;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  const int m = m0 + r;
  const int b = m / (H * W), rem = m - b * (H * W), iy = rem / W, ix = rem - iy * W;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int k = k0 + kq;
    const int tap = k / Co, co = k - tap * Co;
    s4 a = {0, 0, 0, 0};
    /* the output pixel whose tap reads (iy, ix): iy = 2 oy - 1 + ky */
    const int ty = iy + 1 - tap / 3, tx = ix + 1 - tap % 3;
    if (m < M && ty >= 0 && tx >= 0 && !(ty & 1) && !(tx & 1) && (ty >> 1) < Ho && (tx >> 1) < Wo)
      a = *(const s4 *)(dpre + (((size_t)b * Ho + (ty >> 1)) * Wo + (tx >> 1)) * Co + co);
    const s4 bv = *(const s4 *)(wt + (size_t)(n0 + r) * K + k);
    acc = mfma16(a, bv, acc);
  }
  const int ci = n0 + r;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int om = m0 + (lane >> 4) * 4 + j;
    if (om < M) {
      const size_t o = (size_t)om * Ci + ci;
      dbelow[o] = bf2f(ybelow[o]) > 0.f ? f2bf(acc[j]) : 0;
    }
  }
}

/* Weight gradient, one pixel slice (blockIdx.z of WG_SLICES) per block:
 * part[slice][co][k] = sum over the slice's pixels of dpre[pixel][co] *
 * im2col(in)[pixel][k], k < 9 Ci. One wave per 16 x 16 tile, 4 waves per
 * block on consecutive k tiles; grid (Co / 16, ceil(9 Ci / 64), slices). */
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t *__restrict__ dpre, const uint16_t *__restrict__ in,
                                                    float *__restrict__ part, int B, int H, int W, int Ci, int Co) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int Ho = H / 2, Wo = W / 2, P = B * Ho * Wo, KR = 9 * Ci;
  const int m0 = blockIdx.x * 16, n0 = (blockIdx.y * 4 + wave) * 16;
  if (n0 >= KR) return; /* whole wave */
  const int r = lane & 15, kq = (lane >> 4) * 4;
  const int per = (P / 16 + WG_SLICES - 1) / WG_SLICES * 16;
  const int p0 = blockIdx.z * per, p1 = p0 + per < P ? p0 + per : P;
  const int co = m0 + r, kk = n0 + r;
  const int tap = kk / Ci, ci = kk - tap * Ci, ky = tap / 3, kx = tap % 3;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int p = p0; p < p1; p += 16) {
    s4 a, bv;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int px = p + kq + j;
      a[j] = (short)dpre[(size_t)px * Co + co];
      const int b = px / (Ho * Wo), rem = px - b * (Ho * Wo), oy = rem / Wo, ox = rem - oy * Wo;
      const int iy = 2 * oy - 1 + ky, ix = 2 * ox - 1 + kx;
      bv[j] = (short)(kk < KR && iy >= 0 && iy < H && ix >= 0 && ix < W ? in[(((size_t)b * H + iy) * W + ix) * Ci + ci] : 0);
    }
    acc = mfma16(a, bv, acc);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int orow = m0 + (lane >> 4) * 4 + j;
    if (kk < KR) part[((size_t)blockIdx.z * Co + orow) * KR + kk] = acc[j];
  }
}

/* The slices added in order into the FP32 gradient at the layer's parameter
 * offset ([co][tap][ci_param], ci_param < Cparam of the Ci channels in
 * memory); bias gradient: the column sums of dpre. */
__global__ void k_wgrad_sum(const float *__restrict__ part, float *__restrict__ G, int off, int Co, int Ci, int Cparam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, KR = 9 * Ci;
  if (i >= Co * KR) return;
  const int co = i / KR, k = i - co * KR, tap = k / Ci, ci = k - tap * Ci;
  if (ci >= Cparam) return;
  float s = 0.f;
  for (int z = 0; z < WG_SLICES; z++) s += part[((size_t)z * Co + co) * KR + k];
  G[off + (co * 9 + tap) * Cparam + ci] = s;
}
__global__ void k_bias_grad(const uint16_t *__restrict__ dpre, float *__restrict__ G, int off, int M, int Co) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (int m = threadIdx.x; m < M; m += 256) s += bf2f(dpre[(size_t)m * Co + co]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) G[off + co] = red[0];
}

/* Heads, one 64-thread block per sample: features, logits, the sample's
 * loss (softmax cross-entropy of both heads, labels < 0 count nothing). */
__global__ __launch_bounds__(64) void k_head_fwd(const uint16_t *__restrict__ y3, const float *__restrict__ q,
                                                 const float *__restrict__ P, const int *__restrict__ labels,
                                                 float *__restrict__ h, float *__restrict__ logits,
                                                 float *__restrict__ loss) {
  __shared__ float hs[NF];
  __shared__ float lg[NP + NM];
  const int b = blockIdx.x, c = threadIdx.x;
  float s = 0.f;
  for (int pos = 0; pos < 64; pos++) s += bf2f(y3[((size_t)b * 64 + pos) * C3 + c]);
  const float hv = s / 64.f + q[b] * P[P_WQ + c] + P[P_BQ + c];
  hs[c] = hv;
  h[b * NF + c] = hv;
  __syncthreads();
  if (c < NP + NM) {
    float z;
    if (c < NP) {
      z = P[P_BP + c];
      for (int f = 0; f < NF; f++) z += hs[f] * P[P_WP + f * NP + c];
    } else {
      z = P[P_BM + c - NP];
      for (int f = 0; f < NF; f++) z += hs[f] * P[P_WM + f * NM + c - NP];
    }
    lg[c] = z;
    logits[b * (NP + NM) + c] = z;
  }
  __syncthreads();
  if (c == 0) {
    float L = 0.f;
    for (int head = 0; head < 2; head++) {
      const int base = head ? NP : 0, n = head ? NM : NP, y = labels[2 * b + head];
      if (y < 0) continue;
      float mx = lg[base];
      for (int j = 1; j < n; j++) mx = fmaxf(mx, lg[base + j]);
      float se = 0.f;
      for (int j = 0; j < n; j++) se += expf(lg[base + j] - mx);
      L += logf(se) + mx - lg[base + y];
    }
    loss[b] = L;
  }
}

__global__ void k_loss_sum(const float *__restrict__ per, float *__restrict__ out, int B) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int b = 0; b < B; b++) s += per[b];
    out[0] = s / B;
  }
}

/* Head backward per sample: d logits = (softmax - onehot) / B, dh, and the
 * pre-activation gradient of layer 3: dh / 64 where y3 > 0, BF16. */
__global__ __launch_bounds__(64) void k_head_bwd(const float *__restrict__ logits, const int *__restrict__ labels,
                                                 const float *__restrict__ P, const uint16_t *__restrict__ y3,
                                                 float *__restrict__ dlog, float *__restrict__ dh,
                                                 uint16_t *__restrict__ dpre3, int B) {
  __shared__ float dl[NP + NM];
  const int b = blockIdx.x, c = threadIdx.x;
  if (c < 2) {
    const int base = c ? NP : 0, n = c ? NM : NP, y = labels[2 * b + c];
    const float *lg = logits + b * (NP + NM) + base;
    float mx = lg[0];
    for (int j = 1; j < n; j++) mx = fmaxf(mx, lg[j]);
    float se = 0.f;
    for (int j = 0; j < n; j++) se += expf(lg[j] - mx);
    for (int j = 0; j < n; j++) {
      const float g = y < 0 ? 0.f : (expf(lg[j] - mx) / se - (j == y ? 1.f : 0.f)) / B;
      dl[base + j] = g;
      dlog[b * (NP + NM) + base + j] = g;
    }
  }
  __syncthreads();
  float g = 0.f;
  for (int j = 0; j < NP; j++) g += dl[j] * P[P_WP + c * NP + j];
  for (int j = 0; j < NM; j++) g += dl[NP + j] * P[P_WM + c * NM + j];
  dh[b * NF + c] = g;
  for (int pos = 0; pos < 64; pos++) {
    const size_t o = ((size_t)b * 64 + pos) * C3 + c;
    dpre3[o] = bf2f(y3[o]) > 0.f ? f2bf(g / 64.f) : 0;
  }
}

/* Head parameter gradients, one thread per parameter, summed over the batch. */
__global__ void k_head_grads(const float *__restrict__ h, const float *__restrict__ q, const float *__restrict__ dlog,
                             const float *__restrict__ dh, float *__restrict__ G, int B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.f;
  if (i < NF * NP) {
    const int f = i / NP, j = i % NP;
    for (int b = 0; b < B; b++) s += h[b * NF + f] * dlog[b * (NP + NM) + j];
    G[P_WP + i] = s;
  } else if (i < NF * NP + NP) {
    const int j = i - NF * NP;
    for (int b = 0; b < B; b++) s += dlog[b * (NP + NM) + j];
    G[P_BP + j] = s;
  } else if (i < NF * NP + NP + NF * NM) {
    const int t = i - NF * NP - NP, f = t / NM, j = t % NM;
    for (int b = 0; b < B; b++) s += h[b * NF + f] * dlog[b * (NP + NM) + NP + j];
    G[P_WM + t] = s;
  } else if (i < NF * NP + NP + NF * NM + NM) {
    const int j = i - NF * NP - NP - NF * NM;
    for (int b = 0; b < B; b++) s += dlog[b * (NP + NM) + NP + j];
    G[P_BM + j] = s;
  } else if (i < NF * NP + NP + NF * NM + NM + NF) {
    const int f = i - (NF * NP + NP + NF * NM + NM);
    for (int b = 0; b < B; b++) s += dh[b * NF + f] * q[b];
    G[P_WQ + f] = s;
  } else if (i < NF * NP + NP + NF * NM + NM + 2 * NF) {
    const int f = i - (NF * NP + NP + NF * NM + NM + NF);
    for (int b = 0; b < B; b++) s += dh[b * NF + f];
    G[P_BQ + f] = s;
  }
}

__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m1, float *__restrict__ m2,
                       float lr, float b1, float b2, float eps, float c1, float c2, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float a = b1 * m1[i] + (1.f - b1) * g, v = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = v;
  P[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

/* A sample per superblock of the round's frames, straight from HBM: luma /
 * 255 and (inter frames) luma minus the LAST reference's co-located luma /
 * 255, as the training samples are built; q = base_q_idx / 255. */
__global__ __launch_bounds__(256) void k_frame_input(const BatchFrame *__restrict__ T, const BatchGrid g,
                                                     uint16_t *__restrict__ x0, float *__restrict__ q) {
  const BatchIdx b = batch_idx(g);
  const BatchFrame &F = T[b.f];
  const Mav1FrameParams &fp = F.fp;
  const int sr = b.x / fp.sb_cols, sc = b.x - sr * fp.sb_cols, S = fp.stride[0];
  const uint8_t *Y = F.src + fp.offset[0];
  const uint8_t *R = fp.frame_is_intra ? nullptr : F.ref + fp.offset[0];
  for (int pix = threadIdx.x; pix < 4096; pix += 256) {
    const int o = (sr * 64 + (pix >> 6)) * S + sc * 64 + (pix & 63);
    const int l = Y[o];
    uint16_t *d = x0 + ((size_t)blockIdx.x * 4096 + pix) * 4;
    d[0] = f2bf((float)l / 255.f);
    d[1] = R ? f2bf((float)(l - (int)R[o]) / 255.f) : 0;
    d[2] = 0;
    d[3] = 0;
  }
  if (threadIdx.x == 0) q[blockIdx.x] = (float)fp.base_q_idx / 255.f;
}

__global__ void k_prune_levels(const BatchFrame *__restrict__ T, const BatchGrid g, const float *__restrict__ logits,
                               float tau) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.start[g.n]) return;
  const BatchIdx b = batch_idx(g, s);
  const BatchFrame &F = T[b.f];
  const Mav1FrameParams &fp = F.fp;
  const int sr = b.x / fp.sb_cols, sc = b.x - sr * fp.sb_cols;
  int lvl = 3;
  if ((sr + 1) * 16 <= fp.mi_rows && (sc + 1) * 16 <= fp.mi_cols) {
    const float *lg = logits + (size_t)s * (NP + NM);
    float mx = lg[0];
    for (int j = 1; j < NP; j++) mx = fmaxf(mx, lg[j]);
    float e[NP], se = 0.f;
    for (int j = 0; j < NP; j++) se += (e[j] = expf(lg[j] - mx));
    float cdf = 0.f;
    for (int l = 0; l < 3; l++) {
      cdf += e[l] / se;
      if (cdf >= 1.f - tau) {
        lvl = l;
        break;
      }
    }
  }
  F.prune[b.x] = (int8_t)lvl;
}

}  // namespace

struct mav1hip_pruner {
  int max_batch = 0, batch = 0;
  char err[256] = {0};
  hipStream_t st = nullptr;
  float *P = nullptr, *G = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *wt2 = nullptr, *wt3 = nullptr;
  float *x = nullptr, *q = nullptr;
  int *labels = nullptr;
  uint16_t *x0 = nullptr, *y1 = nullptr, *y2 = nullptr, *y3 = nullptr;
  uint16_t *d1 = nullptr, *d2 = nullptr, *d3 = nullptr;
  float *h = nullptr, *logits = nullptr, *dlog = nullptr, *dh = nullptr, *lossb = nullptr, *loss = nullptr;
  float *part = nullptr;
  int packed = 0, forwarded = 0;
};

namespace {

int pfail(mav1hip_pruner *p, int code, const char *fmt, ...) {
  if (p) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(p->err, sizeof p->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

#define PCHK(p, call)                                                                                      \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) return pfail(p, MAV1HIP_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                       __FILE__, __LINE__);                                                \
  } while (0)

template <class T>
int palloc(mav1hip_pruner *p, T **ptr, size_t n) {
  PCHK(p, hipMalloc((void **)ptr, n * sizeof(T)));
  PCHK(p, hipMemset(*ptr, 0, n * sizeof(T)));
  return MAV1HIP_OK;
}

int pack(mav1hip_pruner *p) {
  if (p->packed) return MAV1HIP_OK;
  const int n = C3 * K3;
  hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, p->st, p->P, p->w1, p->w2, p->w3, p->wt2, p->wt3);
  PCHK(p, hipGetLastError());
  p->packed = 1;
  return MAV1HIP_OK;
}

/* the three convolutions' shapes: input H (= W), Ci in memory, Co, Kp */
struct Layer {
  int H, Ci, Co, Kp;
};
constexpr Layer L1 = {64, CI0, C1, K1}, L2 = {32, C1, C2, K2}, L3 = {16, C2, C3, K3};

int conv_fwd(mav1hip_pruner *p, const Layer &L, const uint16_t *in, const uint16_t *w, const float *bias, uint16_t *out) {
  const int M = p->batch * (L.H / 2) * (L.H / 2);
  const dim3 grid((M + 63) / 64);
  if (L.Co == 32)
    hipLaunchKernelGGL(k_conv_fwd<2>, grid, dim3(256), 0, p->st, in, w, bias, out, p->batch, L.H, L.H, L.Ci, L.Kp);
  else if (L.Co == 64)
    hipLaunchKernelGGL(k_conv_fwd<4>, grid, dim3(256), 0, p->st, in, w, bias, out, p->batch, L.H, L.H, L.Ci, L.Kp);
  else
    return pfail(p, MAV1HIP_EINVAL, "conv: %d output channels", L.Co);
  PCHK(p, hipGetLastError());
  return MAV1HIP_OK;
}

int conv_wgrad(mav1hip_pruner *p, const Layer &L, const uint16_t *dpre, const uint16_t *in, int off, int cparam,
               int boff) {
  const int KR = 9 * L.Ci;
  hipLaunchKernelGGL(k_conv_wgrad, dim3(L.Co / 16, (KR + 63) / 64, WG_SLICES), dim3(256), 0, p->st, dpre, in, p->part,
                     p->batch, L.H, L.H, L.Ci, L.Co);
  hipLaunchKernelGGL(k_wgrad_sum, dim3((L.Co * KR + 255) / 256), dim3(256), 0, p->st, p->part, p->G, off, L.Co, L.Ci,
                     cparam);
  hipLaunchKernelGGL(k_bias_grad, dim3(L.Co), dim3(256), 0, p->st, dpre, p->G, boff, p->batch * (L.H / 2) * (L.H / 2),
                     L.Co);
  PCHK(p, hipGetLastError());
  return MAV1HIP_OK;
}

int conv_dgrad(mav1hip_pruner *p, const Layer &L, const uint16_t *dpre, const uint16_t *wt, const uint16_t *ybelow,
               uint16_t *dbelow) {
  const int M = p->batch * L.H * L.H;
  hipLaunchKernelGGL(k_conv_dgrad, dim3((M + 63) / 64, L.Ci / 16), dim3(256), 0, p->st, dpre, wt, ybelow, dbelow,
                     p->batch, L.H, L.H, L.Ci, L.Co);
  PCHK(p, hipGetLastError());
  return MAV1HIP_OK;
}

void destroy_pruner(mav1hip_pruner *p) {
  if (!p) return;
  if (p->st) (void)hipStreamSynchronize(p->st);
  void *bufs[] = {p->P,  p->G,  p->m1, p->m2, p->w1, p->w2, p->w3,     p->wt2,  p->wt3, p->x,     p->q,   p->labels,
                  p->x0, p->y1, p->y2, p->y3, p->d1, p->d2, p->d3,     p->h,    p->logits, p->dlog, p->dh,  p->lossb,
                  p->loss, p->part};
  for (void *b : bufs) (void)hipFree(b);
  if (p->st) (void)hipStreamDestroy(p->st);
  delete p;
}

}  // namespace

int pruner_infer_frames(mav1hip_pruner *p, hipStream_t st, const BatchFrame *T, const BatchFrame *H, int n, float tau) {
  if (!p || n < 1) return MAV1HIP_EINVAL;
  BatchGrid g = batch_grid();
  for (int f = 0; f < n; f++) batch_grid_add(g, H[f].fp.sb_cols * H[f].fp.sb_rows, 1, 1);
  const int total = g.start[n];
  if (total > p->max_batch) return pfail(p, MAV1HIP_EINVAL, "pruner: %d superblocks, room for %d", total, p->max_batch);
  hipStream_t own = p->st;
  p->st = st;
  p->batch = total;
  int rc = pack(p);
  if (!rc) {
    hipLaunchKernelGGL(k_frame_input, dim3(total), dim3(256), 0, st, T, g, p->x0, p->q);
    rc = hipGetLastError() == hipSuccess ? MAV1HIP_OK : pfail(p, MAV1HIP_EDEVICE, "pruner input launch failed");
  }
  if (!rc) rc = conv_fwd(p, L1, p->x0, p->w1, p->P + P_B1, p->y1);
  if (!rc) rc = conv_fwd(p, L2, p->y1, p->w2, p->P + P_B2, p->y2);
  if (!rc) rc = conv_fwd(p, L3, p->y2, p->w3, p->P + P_B3, p->y3);
  if (!rc) {
    hipLaunchKernelGGL(k_head_fwd, dim3(total), dim3(64), 0, st, p->y3, p->q, p->P, p->labels, p->h, p->logits,
                       p->lossb);
    hipLaunchKernelGGL(k_prune_levels, dim3((total + 255) / 256), dim3(256), 0, st, T, g, p->logits, tau);
    rc = hipGetLastError() == hipSuccess ? MAV1HIP_OK : pfail(p, MAV1HIP_EDEVICE, "pruner launch failed");
  }
  p->st = own;
  return rc;
}

extern "C" {

int32_t mav1hip_pruner_num_params(void) { return P_N; }

mav1hip_pruner *mav1hip_pruner_create(int32_t max_batch, char *err, size_t errlen) {
  if (max_batch < 1 || max_batch > 4096) {
    if (err && errlen) snprintf(err, errlen, "max_batch must be 1 .. 4096");
    return nullptr;
  }
  mav1hip_pruner *p = new mav1hip_pruner();
  p->max_batch = max_batch;
  const size_t B = (size_t)max_batch;
  int rc = hipStreamCreate(&p->st) == hipSuccess ? MAV1HIP_OK : pfail(p, MAV1HIP_EDEVICE, "hipStreamCreate failed");
  if (!rc) rc = palloc(p, &p->P, P_N);
  if (!rc) rc = palloc(p, &p->G, P_N);
  if (!rc) rc = palloc(p, &p->m1, P_N);
  if (!rc) rc = palloc(p, &p->m2, P_N);
  if (!rc) rc = palloc(p, &p->w1, (size_t)C1 * K1);
  if (!rc) rc = palloc(p, &p->w2, (size_t)C2 * K2);
  if (!rc) rc = palloc(p, &p->w3, (size_t)C3 * K3);
  if (!rc) rc = palloc(p, &p->wt2, (size_t)C1 * 9 * C2);
  if (!rc) rc = palloc(p, &p->wt3, (size_t)C2 * 9 * C3);
  if (!rc) rc = palloc(p, &p->x, B * 64 * 64 * C0);
  if (!rc) rc = palloc(p, &p->q, B);
  if (!rc) rc = palloc(p, &p->labels, 2 * B);
  if (!rc) rc = palloc(p, &p->x0, B * 64 * 64 * CI0);
  if (!rc) rc = palloc(p, &p->y1, B * 32 * 32 * C1);
  if (!rc) rc = palloc(p, &p->y2, B * 16 * 16 * C2);
  if (!rc) rc = palloc(p, &p->y3, B * 8 * 8 * C3);
  if (!rc) rc = palloc(p, &p->d1, B * 32 * 32 * C1);
  if (!rc) rc = palloc(p, &p->d2, B * 16 * 16 * C2);
  if (!rc) rc = palloc(p, &p->d3, B * 8 * 8 * C3);
  if (!rc) rc = palloc(p, &p->h, B * NF);
  if (!rc) rc = palloc(p, &p->logits, B * (NP + NM));
  if (!rc) rc = palloc(p, &p->dlog, B * (NP + NM));
  if (!rc) rc = palloc(p, &p->dh, B * NF);
  if (!rc) rc = palloc(p, &p->lossb, B);
  if (!rc) rc = palloc(p, &p->loss, 1);
  if (!rc) rc = palloc(p, &p->part, (size_t)WG_SLICES * C3 * K3);
  if (rc) {
    if (err && errlen) snprintf(err, errlen, "%s", p->err);
    destroy_pruner(p);
    return nullptr;
  }
  return p;
}

void mav1hip_pruner_destroy(mav1hip_pruner *p) { destroy_pruner(p); }

const char *mav1hip_pruner_error(const mav1hip_pruner *p) { return p ? p->err : "no pruner"; }

int mav1hip_pruner_set_params(mav1hip_pruner *p, const float *params) {
  if (!p || !params) return MAV1HIP_EINVAL;
  PCHK(p, hipMemcpyAsync(p->P, params, sizeof(float) * P_N, hipMemcpyHostToDevice, p->st));
  PCHK(p, hipMemsetAsync(p->m1, 0, sizeof(float) * P_N, p->st));
  PCHK(p, hipMemsetAsync(p->m2, 0, sizeof(float) * P_N, p->st));
  PCHK(p, hipStreamSynchronize(p->st));
  p->packed = 0;
  return MAV1HIP_OK;
}

int mav1hip_pruner_get_params(mav1hip_pruner *p, float *params) {
  if (!p || !params) return MAV1HIP_EINVAL;
  PCHK(p, hipMemcpyAsync(params, p->P, sizeof(float) * P_N, hipMemcpyDeviceToHost, p->st));
  PCHK(p, hipStreamSynchronize(p->st));
  return MAV1HIP_OK;
}

int mav1hip_pruner_load(mav1hip_pruner *p, const float *x, const float *q, const int32_t *labels, int32_t batch) {
  if (!p || !x || !q || batch < 1 || batch > p->max_batch) return pfail(p, MAV1HIP_EINVAL, "load: bad batch");
  if (labels)
    for (int b = 0; b < batch; b++)
      if (labels[2 * b] < -1 || labels[2 * b] >= NP || labels[2 * b + 1] < -1 || labels[2 * b + 1] >= NM)
        return pfail(p, MAV1HIP_EINVAL, "load: sample %d has labels %d, %d (partition 0..3, mode 0..7, -1 none)", b,
                     labels[2 * b], labels[2 * b + 1]);
  p->batch = batch;
  PCHK(p, hipMemcpyAsync(p->x, x, sizeof(float) * (size_t)batch * 64 * 64 * C0, hipMemcpyHostToDevice, p->st));
  PCHK(p, hipMemcpyAsync(p->q, q, sizeof(float) * (size_t)batch, hipMemcpyHostToDevice, p->st));
  if (labels)
    PCHK(p, hipMemcpyAsync(p->labels, labels, sizeof(int32_t) * 2 * (size_t)batch, hipMemcpyHostToDevice, p->st));
  else
    PCHK(p, hipMemsetAsync(p->labels, 0xFF, sizeof(int32_t) * 2 * (size_t)batch, p->st));
  const int n = batch * 64 * 64;
  hipLaunchKernelGGL(k_input, dim3((n + 255) / 256), dim3(256), 0, p->st, p->x, p->x0, n);
  PCHK(p, hipGetLastError());
  p->forwarded = 0;
  return MAV1HIP_OK;
}

int mav1hip_pruner_forward(mav1hip_pruner *p, float *logits, float *loss) {
  if (!p || p->batch < 1) return pfail(p, MAV1HIP_EINVAL, "forward: no batch loaded");
  int rc;
  if ((rc = pack(p))) return rc;
  if ((rc = conv_fwd(p, L1, p->x0, p->w1, p->P + P_B1, p->y1))) return rc;
  if ((rc = conv_fwd(p, L2, p->y1, p->w2, p->P + P_B2, p->y2))) return rc;
  if ((rc = conv_fwd(p, L3, p->y2, p->w3, p->P + P_B3, p->y3))) return rc;
  hipLaunchKernelGGL(k_head_fwd, dim3(p->batch), dim3(64), 0, p->st, p->y3, p->q, p->P, p->labels, p->h, p->logits,
                     p->lossb);
  hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(64), 0, p->st, p->lossb, p->loss, p->batch);
  PCHK(p, hipGetLastError());
  if (logits)
    PCHK(p, hipMemcpyAsync(logits, p->logits, sizeof(float) * (NP + NM) * (size_t)p->batch, hipMemcpyDeviceToHost,
                           p->st));
  if (loss) PCHK(p, hipMemcpyAsync(loss, p->loss, sizeof(float), hipMemcpyDeviceToHost, p->st));
  if (logits || loss) PCHK(p, hipStreamSynchronize(p->st));
  p->forwarded = 1;
  return MAV1HIP_OK;
}

int mav1hip_pruner_backward(mav1hip_pruner *p, float *grads) {
  if (!p || !p->forwarded) return pfail(p, MAV1HIP_EINVAL, "backward: run forward first");
  hipLaunchKernelGGL(k_head_bwd, dim3(p->batch), dim3(64), 0, p->st, p->logits, p->labels, p->P, p->y3, p->dlog, p->dh,
                     p->d3, p->batch);
  hipLaunchKernelGGL(k_head_grads, dim3((NF * NP + NP + NF * NM + NM + 2 * NF + 255) / 256), dim3(256), 0, p->st, p->h,
                     p->q, p->dlog, p->dh, p->G, p->batch);
  PCHK(p, hipGetLastError());
  int rc;
  if ((rc = conv_wgrad(p, L3, p->d3, p->y2, P_W3, C2, P_B3))) return rc;
  if ((rc = conv_dgrad(p, L3, p->d3, p->wt3, p->y2, p->d2))) return rc;
  if ((rc = conv_wgrad(p, L2, p->d2, p->y1, P_W2, C1, P_B2))) return rc;
  if ((rc = conv_dgrad(p, L2, p->d2, p->wt2, p->y1, p->d1))) return rc;
  if ((rc = conv_wgrad(p, L1, p->d1, p->x0, P_W1, C0, P_B1))) return rc;
  if (grads) {
    PCHK(p, hipMemcpyAsync(grads, p->G, sizeof(float) * P_N, hipMemcpyDeviceToHost, p->st));
    PCHK(p, hipStreamSynchronize(p->st));
  }
  return MAV1HIP_OK;
}

int mav1hip_pruner_adam(mav1hip_pruner *p, float lr, float beta1, float beta2, float eps, int32_t step) {
  if (!p || step < 1) return MAV1HIP_EINVAL;
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((P_N + 255) / 256), dim3(256), 0, p->st, p->P, p->G, p->m1, p->m2, lr, beta1, beta2,
                     eps, c1, c2, P_N);
  PCHK(p, hipGetLastError());
  p->packed = 0;
  return MAV1HIP_OK;
}

int mav1hip_pruner_sync(mav1hip_pruner *p) {
  if (!p) return MAV1HIP_EINVAL;
  PCHK(p, hipStreamSynchronize(p->st));
  return MAV1HIP_OK;
}

}  // extern "C"
