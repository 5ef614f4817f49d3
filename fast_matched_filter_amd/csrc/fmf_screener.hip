// Learned screener: features, 1D convolutional network (forward, backward,
// Adam) and candidate gating for the exact matched filter.
//
// The network maps, per template, a decimated feature sequence (8 channels at
// one value per 8 lags) to the decimated peak |cc| (max |cc| over each run of
// 8 lags): 4 layers of 1D convolution, kernel 9 ('same' padding), 32 hidden
// channels, ReLU between layers.  Weights are stored in FP16; every kernel
// accumulates in FP32.  These kernels are plain vector-ALU code: one
// screening pass over a day costs ~1e10 MACs per template, 1/400 of the
// exact correlation, and training runs on short crops.
//
// Layouts: activations (B, C, X) float32, row-major; weights (Cout, Cin, 9)
// float16 with a float32 copy for compute; bias (Cout) float32.
#include <hip/hip_fp16.h>

#include <algorithm>

#include "fmf_internal.h"

namespace fmf {

namespace {

constexpr int kK = 9;       // kernel taps
constexpr int kPad = 4;     // 'same' padding
constexpr int kTile = 256;  // positions per block

// ---- features --------------------------------------------------------------

// env[ch][q] = mean of d^2 over samples [8q, 8q+8) of channel ch.
__global__ void envelope_kernel(const float *__restrict__ data, int n_data,
                                int n_dec, float *__restrict__ env) {
  const int ch = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_dec) return;
  const float *d = data + (size_t)ch * n_data + 8 * (size_t)q;
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const float v = (8 * q + i < n_data) ? d[i] : 0.0f;
    s = __builtin_fmaf(v, v, s);
  }
  env[(size_t)ch * n_dec + q] = s * 0.125f;
}

// out[ch][q] = mean of env[ch][q .. q + width) (clipped at the end).
__global__ void box_mean_kernel(const float *__restrict__ env, int n_dec,
                                int width, float *__restrict__ out) {
  const int ch = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_dec) return;
  const float *e = env + (size_t)ch * n_dec;
  const int hi = q + width < n_dec ? q + width : n_dec;
  float s = 0.0f;
  for (int i = q; i < hi; i++) s += e[i];
  out[(size_t)ch * n_dec + q] = s / (float)(hi - q);
}

// Per-template features on the decimated lag grid j (lags 8j..8j+7), built
// from the moveout-aligned envelopes of the active channels (w != 0):
//  f0: weighted mean over channels of the window energy (mean envelope over
//      the template duration starting at j + mo/8), relative to the channel's
//      mean energy  (energy arriving with the template's moveout pattern)
//  f1: maximum over channels of the same ratio
//  f2: weighted mean of the envelope at the window start
//  f3: weighted mean of the window energy ratio over the first half window
//  f4: weighted mean of the window energy ratio over the second half window
//  f5: fraction of active channels whose window ratio exceeds 1.05
//  f6: template length / 1000
//  f7: constant 1
// win[ch][q] is the prefix-free box mean of env over `nw` decimated samples
// and chmean[ch] the channel's mean envelope (both prepared by the host-side
// driver from env with cumulative sums).
__global__ void features_kernel(const float *__restrict__ win,
                                const float *__restrict__ winh,
                                const float *__restrict__ env,
                                const float *__restrict__ chinv,
                                const int *__restrict__ moveouts,
                                const float *__restrict__ weights,
                                int n_channels, int n_dec, int n_out, int nwh,
                                float len_feature, float *__restrict__ feat) {
  const int t = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_out) return;
  float f0 = 0, f1 = 0, f2 = 0, f3 = 0, f4 = 0, f5 = 0, wsum = 0;
  int used = 0;
  for (int ch = 0; ch < n_channels; ch++) {
    const float w = weights[t * n_channels + ch];
    if (w == 0.0f) continue;
    int q = j + moveouts[t * n_channels + ch] / 8;
    q = q < 0 ? 0 : (q >= n_dec ? n_dec - 1 : q);
    const size_t base = (size_t)ch * n_dec;
    const float r = win[base + q] * chinv[ch];
    const int q2 = q + nwh < n_dec ? q + nwh : n_dec - 1;
    f0 += w * r;
    f1 = r > f1 ? r : f1;
    f2 += w * env[base + q] * chinv[ch];
    f3 += w * winh[base + q] * chinv[ch];
    f4 += w * winh[base + q2] * chinv[ch];
    f5 += r > 1.05f ? 1.0f : 0.0f;
    wsum += w;
    used++;
  }
  const float iw = wsum > 0 ? 1.0f / wsum : 0.0f;
  float *o = feat + (size_t)t * 8 * n_out + j;
  o[0 * (size_t)n_out] = f0 * iw;
  o[1 * (size_t)n_out] = f1;
  o[2 * (size_t)n_out] = f2 * iw;
  o[3 * (size_t)n_out] = f3 * iw;
  o[4 * (size_t)n_out] = f4 * iw;
  o[5 * (size_t)n_out] = used ? f5 / used : 0.0f;
  o[6 * (size_t)n_out] = len_feature;
  o[7 * (size_t)n_out] = 1.0f;
}

// ---- convolution layers ----------------------------------------------------

// y[b][co][x] = bias[co] + sum_ci sum_k w[co][ci][k] * a[b][ci][x + k - 4]
// where a = relu(in) when relu_in, else in.  One block per (b, 256 positions);
// the weights and the input tile (+-4 halo) of every input channel sit in
// LDS, each thread keeps COUT accumulators for one position.
template <int COUT>
__global__ __launch_bounds__(kTile) void conv_fwd_kernel(
    const float *__restrict__ in, int cin, int nx, int relu_in,
    const float *__restrict__ w, const float *__restrict__ bias,
    float *__restrict__ out) {
  // LDS: weights as [ci][k][co] (the COUT weights of one (ci, k) are
  // contiguous, read as broadcast float4s), then the input tile with halo.
  extern __shared__ float4 smem4[];
  float *wl = reinterpret_cast<float *>(smem4);
  float *tile = wl + cin * kK * COUT;
  const int b = blockIdx.y;
  const int x0 = blockIdx.x * kTile;
  const int tid = threadIdx.x;
  const int tw = kTile + 2 * kPad;
  for (int i = tid; i < cin * kK * COUT; i += kTile) {
    const int co = i % COUT, r = i / COUT;  // r = ci * kK + k
    wl[i] = w[co * cin * kK + r];
  }
  for (int i = tid; i < cin * tw; i += kTile) {
    const int ci = i / tw, xx = x0 - kPad + (i - ci * tw);
    float v = (xx >= 0 && xx < nx) ? in[((size_t)b * cin + ci) * nx + xx] : 0.0f;
    if (relu_in) v = v > 0.0f ? v : 0.0f;
    tile[i] = v;
  }
  __syncthreads();
  const int x = x0 + tid;
  if (x >= nx) return;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; co++) acc[co] = bias[co];
  for (int ci = 0; ci < cin; ci++) {
#pragma unroll
    for (int k = 0; k < kK; k++) {
      const float a = tile[ci * tw + tid + k];
      const float *wk = wl + (ci * kK + k) * COUT;
      if constexpr (COUT % 4 == 0) {
#pragma unroll
        for (int c4 = 0; c4 < COUT / 4; c4++) {
          const float4 wv = reinterpret_cast<const float4 *>(wk)[c4];
          acc[4 * c4 + 0] = __builtin_fmaf(wv.x, a, acc[4 * c4 + 0]);
          acc[4 * c4 + 1] = __builtin_fmaf(wv.y, a, acc[4 * c4 + 1]);
          acc[4 * c4 + 2] = __builtin_fmaf(wv.z, a, acc[4 * c4 + 2]);
          acc[4 * c4 + 3] = __builtin_fmaf(wv.w, a, acc[4 * c4 + 3]);
        }
      } else {
#pragma unroll
        for (int co = 0; co < COUT; co++) acc[co] = __builtin_fmaf(wk[co], a, acc[co]);
      }
    }
  }
#pragma unroll
  for (int co = 0; co < COUT; co++) out[((size_t)b * COUT + co) * nx + x] = acc[co];
}

// ---- matrix-core forward (inference) ----------------------------------------
//
// The same convolution as an implicit GEMM on v_mfma_f32_32x32x16_f16:
//   out[co][x] = bias[co] + sum_{kidx < 9*CIN} A[co][kidx] * B[kidx][x],
//   kidx = k * CIN + ci,  A[co][kidx] = w[co][ci][k],  B[kidx][x] = a[ci][x + k - 4],
// with activations and weights in FP16 and FP32 accumulation (inference: the
// activations are rounded to FP16 as they are staged; training keeps the
// FP32 kernel above).  Ordering K as (tap, channel) makes the 8 values of
// a lane's B fragment 8 consecutive channels of one input position: the input
// tile sits in LDS as [position][channel] (rows padded to CIN + 8 halves for
// CIN = 32 so the 16-byte reads of 8 consecutive lanes hit distinct banks).
// A fragments are packed once per call (conv_pack_kernel) in fragment order
// and held in registers.  Block: 256 threads, 256 positions of one batch row;
// each wave owns 2 tiles of 32 positions, all 32 output channels.

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16v __attribute__((ext_vector_type(16)));

template <int CIN>
struct MfmaConv {
  static constexpr int kChunks = (CIN * kK + 15) / 16;  // K steps of 16
  static constexpr int kRow = CIN == 32 ? 40 : CIN;     // LDS row (halves)
  static constexpr int kRows = kTile + 16;              // positions x0-4 .. x0+267
};

// wf[(kc * 32 + co) * 16 + j] = w[co][ci][k] for kidx = 16 kc + j = k * CIN + ci
// (0 past 9 * CIN).
__global__ void conv_pack_kernel(const float *__restrict__ w, int cin, int chunks,
                                 _Float16 *__restrict__ wf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= chunks * 32 * 16) return;
  const int j = i & 15, co = (i >> 4) & 31, kc = i >> 9;
  const int kidx = 16 * kc + j;
  float v = 0.0f;
  if (kidx < cin * kK) {
    const int k = kidx / cin, ci = kidx - k * cin;
    v = w[(co * cin + ci) * kK + k];
  }
  wf[i] = (_Float16)v;
}

template <int CIN>
__global__ __launch_bounds__(kTile) void conv_fwd_mfma_kernel(
    const float *__restrict__ in, int nx, int relu_in,
    const _Float16 *__restrict__ wf, const float *__restrict__ bias,
    float *__restrict__ out) {
  using C = MfmaConv<CIN>;
  __shared__ __align__(16) _Float16 tile[C::kRows * C::kRow];
  const int b = blockIdx.y;
  const int x0 = blockIdx.x * kTile;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;
  // A fragments of every K step
  f16x8 afr[C::kChunks];
#pragma unroll
  for (int kc = 0; kc < C::kChunks; kc++)
    afr[kc] = *reinterpret_cast<const f16x8 *>(wf + (kc * 32 + n) * 16 + 8 * h);
  // input tile: row r = position x0 - 4 + r, FP16 (ReLU of the layer input)
  const float *src = in + (size_t)b * CIN * nx;
  for (int i = tid; i < CIN * C::kRows; i += kTile) {
    const int ci = i / C::kRows, r = i - ci * C::kRows;
    const int xx = x0 - kPad + r;
    float v = (xx >= 0 && xx < nx) ? src[(size_t)ci * nx + xx] : 0.0f;
    if (relu_in) v = v > 0.0f ? v : 0.0f;
    tile[r * C::kRow + ci] = (_Float16)v;
  }
  __syncthreads();
  f32x16v acc[2];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[p][j] = 0.0f;
#pragma unroll
  for (int kc = 0; kc < C::kChunks; kc++) {
    const int kidx = 16 * kc + 8 * h;
    const int k = kidx / CIN, ci0 = kidx - k * CIN;  // 8 channels of tap k
#pragma unroll
    for (int p = 0; p < 2; p++) {
      const int row = wave * 64 + 32 * p + n + k;
      const f16x8 bfr = *reinterpret_cast<const f16x8 *>(tile + row * C::kRow + ci0);
      acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_f16(afr[kc], bfr, acc[p], 0, 0, 0);
    }
  }
  float *dst = out + (size_t)b * 32 * nx;
#pragma unroll
  for (int p = 0; p < 2; p++) {
    const int x = x0 + wave * 64 + 32 * p + n;
    if (x < nx) {
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int co = 8 * (j >> 2) + 4 * h + (j & 3);
        dst[(size_t)co * nx + x] = acc[p][j] + bias[co];
      }
    }
  }
}

// ---- coarse correlation features -------------------------------------------

// out[r][i] = mean of in[r][D*i .. D*i + D) (samples past n_in count as 0),
// summed in sample order.  When D and the row length are multiples of 4 and
// the input is 16-byte aligned, a thread's D samples are D / 4 16-byte loads
// (consecutive threads read consecutive 4 D bytes).
__global__ void decimate_kernel(const float *__restrict__ in, long long n_in,
                                int D, long long n_out, float *__restrict__ out) {
  const long long r = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  const float *src = in + r * n_in;
  float s = 0.0f;
  if ((D & 3) == 0 && (n_in & 3) == 0 && ((uintptr_t)in & 15) == 0 && (i + 1) * D <= n_in) {
    const float4 *v = reinterpret_cast<const float4 *>(src + i * D);
    for (int k = 0; k < D / 4; k++) {
      const float4 x = v[k];
      s += x.x;
      s += x.y;
      s += x.z;
      s += x.w;
    }
  } else {
    for (int k = 0; k < D; k++) {
      const long long g = i * D + k;
      s += g < n_in ? src[g] : 0.0f;
    }
  }
  out[r * n_out + i] = s / (float)D;
}

// log(1 + x) for x >= 0 by Goldberg's correction (D. Goldberg, "What every
// computer scientist should know about floating-point arithmetic", 1991,
// Theorem 4): with u = fl(1 + x), log1p(x) = log(u) * x / (u - 1), exact to a
// few ulp given an accurate log(u); log(u) from the hardware base-2
// logarithm.  About 8 instructions against the ~120 of the library's
// compensated log1pf, and well inside the FP16 rounding the network applies
// to its inputs.
__device__ __forceinline__ float log1p_nonneg(float x) {
  const float u = 1.0f + x;
  if (u == 1.0f) return x;
  if (u == __builtin_inff()) return u;
  return __logf(u) * (x * __builtin_amdgcn_rcpf(u - 1.0f));
}

// Features of screener position j (exact lags 8j .. 8j+7) from the coarse
// cc (T, n_coarse) of the decimated scan (lag q covers exact lags D*q ..),
// in units of the coarse cc's MAD, log-compressed:
//   f0 max |cc| over the position's coarse lags, f1 max(cc, 0), f2 max(-cc, 0),
//   f3 max |cc| and f4 mean |cc| and f5 rms cc over +-3 positions,
//   f6 log(MAD) + 5 (the template's noise level), f7 = 1.
// OutT float (training) or _Float16 (inference: the fused network reads its
// input as FP16, so storing FP16 halves the bytes and changes no prediction).
template <typename OutT>
__global__ void coarse_features_kernel(const float *__restrict__ cc, int n_coarse,
                                       const float *__restrict__ mad, int D, int n_out,
                                       OutT *__restrict__ feat) {
  const int t = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_out) return;
  const float *row = cc + (size_t)t * n_coarse;
  const float inv = 1.0f / fmaxf(mad[t], 1e-20f);
  const int q0 = (8 * j) / D, q1 = (8 * j + 7) / D;
  const int span = (8 + D - 1) / D;  // coarse lags per position
  const int w0 = q0 - 3 * span, w1 = q1 + 3 * span;
  float mx = 0.0f, pos = 0.0f, neg = 0.0f, wmx = 0.0f, wsum = 0.0f, wsq = 0.0f;
  int wn = 0;
  for (int q = w0; q <= w1; q++) {
    const float v = (q >= 0 && q < n_coarse) ? row[q] : 0.0f;
    const float a = fabsf(v);
    if (q >= q0 && q <= q1) {
      mx = fmaxf(mx, a);
      pos = fmaxf(pos, v);
      neg = fmaxf(neg, -v);
    }
    wmx = fmaxf(wmx, a);
    wsum += a;
    wsq += v * v;
    wn++;
  }
  OutT *f = feat + (size_t)t * 8 * n_out + j;
  f[0] = (OutT)log1p_nonneg(mx * inv);
  f[(size_t)1 * n_out] = (OutT)log1p_nonneg(pos * inv);
  f[(size_t)2 * n_out] = (OutT)log1p_nonneg(neg * inv);
  f[(size_t)3 * n_out] = (OutT)log1p_nonneg(wmx * inv);
  f[(size_t)4 * n_out] = (OutT)log1p_nonneg(wsum / wn * inv);
  f[(size_t)5 * n_out] = (OutT)log1p_nonneg(sqrtf(wsq / wn) * inv);
  f[(size_t)6 * n_out] = (OutT)(logf(fmaxf(mad[t], 1e-20f)) + 5.0f);
  f[(size_t)7 * n_out] = (OutT)1.0f;
}

// ---- fused network (inference) ----------------------------------------------
//
// The four layers in one persistent kernel.  A block loads the packed FP16
// weight fragments of the three hidden layers into registers once and then
// walks tiles of 224 positions: it reads the 8 input features of a tile and
// its halo (16 positions before, 32 after; prefetched into registers while
// the previous tile computes), keeps the hidden activations in LDS (FP16,
// rows of 32 channels padded to 40 halves) and writes only the prediction,
// instead of writing and re-reading 32 FP32 channels per layer.  Hidden
// layers run as in conv_fwd_mfma_kernel (8 tiles of 32 positions per layer,
// one per wave); each computes the positions the next layer's window needs
// (halo 12, 8, 4) and writes 0 at positions outside [0, nx), which is the
// next layer's 'same' zero padding.
//
// The 32 -> 1 output layer runs on the matrix cores too, straight from the
// last hidden layer's accumulators: element j of a lane's accumulator is
// channel 8 (j / 4) + 4 h + j % 4 of row n, so accumulator elements 8c .. 8c+7
// are, as they stand, the B fragment (K = 16 channels in the order
// (j / 4, h, j % 4)) of K step c.  With A[m][K] = w4[channel(K)][m] (m < 9,
// zero rows past) one pair of MFMAs gives u_m[row] = sum_ci h3[row][ci] w4[ci][m]
// for the 9 taps, and the prediction is b4 + sum_m u_m[x + m], one shift-sum
// through LDS.

constexpr int kNetOut = 224;                 // predictions per tile
constexpr int kNetRows = 256;                // rows per hidden layer (8 x 32)
constexpr int kNetRowH = 40;                 // halves per activation row
constexpr int kNetInRows = kNetRows + 16;    // input rows (taps up to 9)
constexpr int kNetBufRows = kNetRows + 8;    // activation rows (taps up to 8)
constexpr int kNetTaps = 16;                 // rows of u (taps 0..8 used)

struct NetArgs {
  const _Float16 *wf1, *wf2, *wf3;           // packed A fragments (5, 18, 18 chunks)
  const float *b1, *b2, *b3, *w4, *b4;
};

template <int KC>
__device__ __forceinline__ void load_frags(const _Float16 *__restrict__ wf, f16x8 (&afr)[KC]) {
  const int lane = threadIdx.x & 63;
  const int n = lane & 31, h = lane >> 5;
#pragma unroll
  for (int kc = 0; kc < KC; kc++)
    afr[kc] = *reinterpret_cast<const f16x8 *>(wf + (kc * 32 + n) * 16 + 8 * h);
}

// ReLU of 4 accumulator values in FP16 (convert, then a packed max: the
// conversion is monotonic and keeps 0).
__device__ __forceinline__ f16x4 relu_f16x4(float a, float b, float c, float d) {
  const f16x4 v = {(_Float16)a, (_Float16)b, (_Float16)c, (_Float16)d};
  const f16x4 z = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
  return __builtin_elementwise_max(v, z);
}

// One hidden layer's products: rows r in [0, 256), inputs at row r + k of
// `src` (row stride src_row halves), acc[p] = bias + products for the rows
// 32 (P wave + p) + n of this lane (P tiles of 32 rows per wave).  The B
// fragments of K step kc + 2 are read while the MFMAs of step kc run; the
// scheduling barriers keep that order.
template <int CIN, int P, int KC>
__device__ __forceinline__ void net_products(const _Float16 *src, int src_row,
                                             const f16x8 (&afr)[KC], const float *bias_lds,
                                             f32x16v (&acc)[P]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 31, h = lane >> 5;
  auto bread = [&](int kc, int p) {
    const int kidx = 16 * kc + 8 * h;
    const int k = kidx / CIN, ci0 = kidx - k * CIN;
    const int row = 32 * (P * wave + p) + n + k;
    return *reinterpret_cast<const f16x8 *>(src + row * src_row + ci0);
  };
  // accumulators start at the bias: element j of a lane is channel
  // 8 (j / 4) + 4 h + j % 4, four consecutive channels per group of 4
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const float4 b = *reinterpret_cast<const float4 *>(bias_lds + 8 * g + 4 * h);
#pragma unroll
    for (int p = 0; p < P; p++) {
      acc[p][4 * g] = b.x;
      acc[p][4 * g + 1] = b.y;
      acc[p][4 * g + 2] = b.z;
      acc[p][4 * g + 3] = b.w;
    }
  }
  // ring of R K steps: the reads of step kc + R - 1 are issued before the
  // MFMAs of step kc (two steps ahead for one tile per wave, one step ahead
  // for two, whose reads are twice as many)
  constexpr int R = P == 1 ? 3 : 2;
  f16x8 bq[R][P];
#pragma unroll
  for (int q = 0; q < R - 1 && q < KC; q++)
#pragma unroll
    for (int p = 0; p < P; p++) bq[q][p] = bread(q, p);
#pragma unroll
  for (int kc = 0; kc < KC; kc++) {
    if (kc + R - 1 < KC) {
#pragma unroll
      for (int p = 0; p < P; p++) bq[(kc + R - 1) % R][p] = bread(kc + R - 1, p);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int p = 0; p < P; p++)
      acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_f16(afr[kc], bq[kc % R][p], acc[p], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ReLU(acc) as FP16 to the rows of `dst`, four channels per 8-byte store.
// MASK (first and last tiles of a row only): rows at positions outside
// [0, nx) are written as 0.
template <bool MASK, int P>
__device__ __forceinline__ void net_store(const f32x16v (&acc)[P], _Float16 *dst, int pos0,
                                          int nx) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 31, h = lane >> 5;
#pragma unroll
  for (int p = 0; p < P; p++) {
    const int r = 32 * (P * wave + p) + n;
    const bool keep = !MASK || (pos0 + r >= 0 && pos0 + r < nx);
#pragma unroll
    for (int g = 0; g < 4; g++) {
      f16x4 o = relu_f16x4(acc[p][4 * g], acc[p][4 * g + 1], acc[p][4 * g + 2],
                           acc[p][4 * g + 3]);
      if (!keep) o = f16x4{(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
      *reinterpret_cast<f16x4 *>(dst + r * kNetRowH + 8 * g + 4 * h) = o;
    }
  }
}

// Output layer products from the last hidden layer's accumulators (see the
// section comment): u[m][r] for the taps m of this lane.  MASK as in
// net_store.
template <bool MASK, int P>
__device__ __forceinline__ void net_out(const f32x16v (&acc)[P], const f16x8 (&a4)[2],
                                        float *u, int pos0, int nx) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 31, h = lane >> 5;
#pragma unroll
  for (int p = 0; p < P; p++) {
    const int r = 32 * (P * wave + p) + n;
    const bool keep = !MASK || (pos0 + r >= 0 && pos0 + r < nx);
    f16x8 bo[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const f16x4 lo = relu_f16x4(acc[p][8 * c], acc[p][8 * c + 1], acc[p][8 * c + 2],
                                  acc[p][8 * c + 3]);
      const f16x4 hi = relu_f16x4(acc[p][8 * c + 4], acc[p][8 * c + 5], acc[p][8 * c + 6],
                                  acc[p][8 * c + 7]);
      bo[c] = f16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      if (!keep) bo[c] = f16x8{};
    }
    f32x16v d = {};
    d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a4[0], bo[0], d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a4[1], bo[1], d, 0, 0, 0);
    // d[j] = u_m[r], m = 8 (j / 4) + 4 h + j % 4: taps 4h .. 4h+3 and 8 + 4h
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const int m = 8 * (j >> 2) + 4 * h + (j & 3);
      u[m * kNetRows + r] = d[j];
    }
  }
}

// Two blocks per CU of 4 waves, two tiles of 32 rows per wave and layer (256
// registers, no spills): each SIMD holds one wave of each block, and the
// blocks' barriers, stores and LDS waits fall at unrelated times, so one
// block's matrix-core work fills the other's gaps (4.7 ms per day against
// 5.0-5.4 ms with one block of 8 waves of one tile each, whose waves all
// reach every barrier together; 5.8 ms with one block of 4 waves).
constexpr int kNetThreads = 256;
constexpr int kNetBlocksPerCu = 2;

// InT: the input features as float (converted to FP16 when staged) or
// already as _Float16.
template <typename InT>
__global__ __launch_bounds__(kNetThreads, kNetBlocksPerCu) void screen_net_kernel(
    const InT *__restrict__ feat, int nb, int nx, int n_xt, NetArgs a,
    float *__restrict__ pred) {
  constexpr int NT = kNetThreads, P = kNetRows / 32 / (NT / 64);
  static_assert(NT / 64 * P * 32 == kNetRows, "rows per layer");
  static_assert(NT >= kNetOut && 2 * NT >= kNetInRows, "one output and two input rows per thread");
  __shared__ __align__(16) _Float16 in_t[kNetInRows * 8];
  __shared__ __align__(16) _Float16 buf_a[kNetBufRows * kNetRowH];
  __shared__ __align__(16) _Float16 buf_b[kNetBufRows * kNetRowH];
  __shared__ __align__(16) float bias[3 * 32];
  // u[m][row] of the output layer, over the first rows of buf_a (free while
  // the last hidden layer reads buf_b; rows 256.. of buf_a stay 0)
  static_assert(kNetTaps * kNetRows * 4 <= kNetRows * kNetRowH * 2, "u fits in buf_a");
  float *u = reinterpret_cast<float *>(buf_a);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int n = lane & 31, h = lane >> 5;
  if (tid < 96) bias[tid] = tid < 32 ? a.b1[tid] : (tid < 64 ? a.b2[tid - 32] : a.b3[tid - 64]);
  for (int i = tid; i < 8 * kNetRowH; i += NT) {  // rows past the computed ones
    buf_a[kNetRows * kNetRowH + i] = (_Float16)0.0f;
    buf_b[kNetRows * kNetRowH + i] = (_Float16)0.0f;
  }
  f16x8 a1[MfmaConv<8>::kChunks], a2[MfmaConv<32>::kChunks], a3[MfmaConv<32>::kChunks];
  load_frags(a.wf1, a1);
  load_frags(a.wf2, a2);
  load_frags(a.wf3, a3);
  // output layer A fragments: A[m = n][K = 8 h + i] of K step c is
  // w4[ch][n], ch = 8 (2 c + i / 4) + 4 h + i % 4 (w4 is (32, 9))
  f16x8 a4[2];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int ch = 8 * (2 * c + i / 4) + 4 * h + i % 4;
      a4[c][i] = (_Float16)(n < kK ? a.w4[ch * kK + n] : 0.0f);
    }
  const float b4 = a.b4[0];
  const long long total = (long long)nb * n_xt;
  // input of a tile, rows r at positions x0 - 16 + r: thread t prefetches
  // the 8 features of row t; rows NT .. 271 are loaded when staged
  constexpr int kInB = (int)sizeof(InT);
  auto load_feat = [&](__amdgpu_buffer_rsrc_t rs, int off) -> InT {
    if constexpr (kInB == 4) return buffer_load_f32(rs, off);
    else return buffer_load_f16(rs, off);
  };
  InT pre[8];
  auto load_input = [&](long long it) {
    const int b = (int)(it / n_xt), x0 = (int)(it % n_xt) * kNetOut;
    const auto rs = buffer_rsrc(feat + (size_t)b * 8 * nx, (unsigned)(8 * nx) * kInB);
    const int pos = x0 - 16 + tid;
    const bool ok = pos >= 0 && pos < nx && tid < kNetInRows;
#pragma unroll
    for (int ci = 0; ci < 8; ci++) pre[ci] = load_feat(rs, ok ? (ci * nx + pos) * kInB : kOob);
  };
  long long it = blockIdx.x;
  if (it < total) load_input(it);
  for (; it < total; it += gridDim.x) {
    const int b = (int)(it / n_xt), x0 = (int)(it % n_xt) * kNetOut;
    __syncthreads();  // the previous tile's output sum has read u (buf_a)
    if (tid < kNetInRows) {
      f16x8 row;
#pragma unroll
      for (int ci = 0; ci < 8; ci++) row[ci] = (_Float16)pre[ci];
      *reinterpret_cast<f16x8 *>(in_t + tid * 8) = row;
    }
    if (tid + NT < kNetInRows) {
      const auto rs = buffer_rsrc(feat + (size_t)b * 8 * nx, (unsigned)(8 * nx) * kInB);
      const int pos = x0 - 16 + tid + NT;
      const bool ok = pos >= 0 && pos < nx;
      f16x8 row;
#pragma unroll
      for (int ci = 0; ci < 8; ci++)
        row[ci] = (_Float16)load_feat(rs, ok ? (ci * nx + pos) * kInB : kOob);
      *reinterpret_cast<f16x8 *>(in_t + (tid + NT) * 8) = row;
    }
    __syncthreads();
    if (it + gridDim.x < total) load_input(it + gridDim.x);
    // every layer's rows lie inside [0, nx) except in the first and last
    // tiles of a row, which take the masked variants
    const bool interior = x0 - 12 >= 0 && x0 - 4 + kNetRows <= nx;
    f32x16v acc[P];
    net_products<8, P>(in_t, 8, a1, bias, acc);
    if (interior) net_store<false, P>(acc, buf_a, x0 - 12, nx);
    else net_store<true, P>(acc, buf_a, x0 - 12, nx);
    __syncthreads();
    net_products<32, P>(buf_a, kNetRowH, a2, bias + 32, acc);
    if (interior) net_store<false, P>(acc, buf_b, x0 - 8, nx);
    else net_store<true, P>(acc, buf_b, x0 - 8, nx);
    __syncthreads();
    net_products<32, P>(buf_b, kNetRowH, a3, bias + 64, acc);
    if (interior) net_out<false, P>(acc, a4, u, x0 - 4, nx);
    else net_out<true, P>(acc, a4, u, x0 - 4, nx);
    __syncthreads();
    if (tid < kNetOut && x0 + tid < nx) {
      float s = b4;
#pragma unroll
      for (int k = 0; k < kK; k++) s += u[k * kNetRows + tid + k];
      pred[(size_t)b * nx + x0 + tid] = s;
    }
  }
}

// Input gradient: ga[b][ci][x] = sum_co sum_k w[co][ci][k] * gz[b][co][x - k + 4],
// then multiplied by the ReLU mask of the layer input pre-activation (zin > 0)
// when mask_relu.  One block per (b, 256 positions), gz tile in LDS.
__global__ __launch_bounds__(kTile) void conv_bwd_data_kernel(
    const float *__restrict__ gz, int cout, int nx,
    const float *__restrict__ w, int cin, const float *__restrict__ zin,
    int mask_relu, float *__restrict__ ga) {
  extern __shared__ float tile[];  // cout * (kTile + 8)
  const int b = blockIdx.y;
  const int x0 = blockIdx.x * kTile;
  const int tid = threadIdx.x;
  const int tw = kTile + 2 * kPad;
  for (int i = tid; i < cout * tw; i += kTile) {
    const int co = i / tw, xx = x0 - kPad + (i - co * tw);
    tile[i] = (xx >= 0 && xx < nx) ? gz[((size_t)b * cout + co) * nx + xx] : 0.0f;
  }
  __syncthreads();
  const int x = x0 + tid;
  if (x >= nx) return;
  for (int ci = 0; ci < cin; ci++) {
    float s = 0.0f;
    for (int co = 0; co < cout; co++) {
#pragma unroll
      for (int k = 0; k < kK; k++)  // gz index x - k + 4 -> tile offset tid + 8 - k
        s = __builtin_fmaf(w[(co * cin + ci) * kK + k], tile[co * tw + tid + 2 * kPad - k], s);
    }
    const size_t idx = ((size_t)b * cin + ci) * nx + x;
    if (mask_relu && !(zin[idx] > 0.0f)) s = 0.0f;
    ga[idx] = s;
  }
}

// Weight and bias gradients: gw[co][ci][k] = sum_b sum_x gz[b][co][x] * a[b][ci][x + k - 4]
// (a = relu(ain) when relu_in), gb[co] = sum_b sum_x gz[b][co][x].  One block
// per (co, ci); block-level tree reduction in FP32.
__global__ __launch_bounds__(kTile) void conv_bwd_weight_kernel(
    const float *__restrict__ gz, int cout, const float *__restrict__ ain,
    int cin, int relu_in, int nb, int nx, float *__restrict__ gw,
    float *__restrict__ gb) {
  __shared__ float red[kK + 1][kTile / 64];
  const int co = blockIdx.x / cin, ci = blockIdx.x - co * cin;
  float s[kK + 1];
#pragma unroll
  for (int k = 0; k <= kK; k++) s[k] = 0.0f;
  const long long n = (long long)nb * nx;
  for (long long i = threadIdx.x; i < n; i += kTile) {
    const int b = (int)(i / nx), x = (int)(i - (long long)b * nx);
    const float g = gz[((size_t)b * cout + co) * nx + x];
    const float *arow = ain + ((size_t)b * cin + ci) * nx;
#pragma unroll
    for (int k = 0; k < kK; k++) {
      const int xx = x + k - kPad;
      float a = (xx >= 0 && xx < nx) ? arow[xx] : 0.0f;
      if (relu_in) a = a > 0.0f ? a : 0.0f;
      s[k] = __builtin_fmaf(g, a, s[k]);
    }
    s[kK] += g;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k <= kK; k++) {
    float v = s[k];
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane == 0) red[k][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x <= kK) {
    float v = 0.0f;
    for (int wv = 0; wv < kTile / 64; wv++) v += red[threadIdx.x][wv];
    if (threadIdx.x < kK)
      gw[(co * cin + ci) * kK + threadIdx.x] = v;
    else if (ci == 0)
      gb[co] = v;
  }
}

// Weighted squared error of the prediction against the target:
// loss = sum_i m_i (p_i - y_i)^2 / sum_i m_i with m_i = 1 + boost * y_i
// (peaks weigh more); writes dloss/dp and per-block partial losses.
__global__ __launch_bounds__(kTile) void loss_kernel(
    const float *__restrict__ pred, const float *__restrict__ target,
    long long n, float boost, float inv_norm, float *__restrict__ grad,
    float *__restrict__ partial) {
  __shared__ float red[kTile / 64];
  float l = 0.0f;
  for (long long i = (long long)blockIdx.x * kTile + threadIdx.x; i < n;
       i += (long long)gridDim.x * kTile) {
    const float y = target[i], d = pred[i] - y, m = 1.0f + boost * y;
    l += m * d * d;
    grad[i] = 2.0f * m * d * inv_norm;
  }
  for (int d = 32; d >= 1; d >>= 1) l += __shfl_xor(l, d, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float v = 0.0f;
    for (int wv = 0; wv < kTile / 64; wv++) v += red[wv];
    partial[blockIdx.x] = v * inv_norm;
  }
}

// Adam on float32 master weights; the FP16 weights and their float32 copy
// used by the forward kernels are refreshed from the updated master.
__global__ void adam_kernel(float *__restrict__ master, const float *__restrict__ grad,
                            float *__restrict__ m, float *__restrict__ v, int n,
                            float lr, float b1, float b2, float eps, float c1,
                            float c2, __half *__restrict__ w16,
                            float *__restrict__ w32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float p = master[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  master[i] = p;
  const __half h = __float2half(p);
  w16[i] = h;
  w32[i] = __half2float(h);
}

// Candidate gate: tile (t, k) of `tile_out` outputs is kept when any
// decimated prediction inside it (lags of the tile widened by `margin`
// decimated samples) exceeds thr[t].  One wave per tile: the lanes read
// consecutive predictions (coalesced) and the wave votes.
constexpr int kGateThreads = 256;

__global__ __launch_bounds__(kGateThreads) void gate_kernel(
    const float *__restrict__ pred, int n_out, const float *__restrict__ thr, int n_tiles,
    int tile_dec, int margin, int n_templates, int *__restrict__ keep) {
  const long long idx =
      ((long long)blockIdx.x * kGateThreads + threadIdx.x) >> 6;  // whole waves
  const int lane = threadIdx.x & 63;
  if (idx >= (long long)n_templates * n_tiles) return;
  const int t = (int)(idx / n_tiles), k = (int)(idx - (long long)t * n_tiles);
  int lo = k * tile_dec - margin, hi = (k + 1) * tile_dec + margin;
  lo = lo < 0 ? 0 : lo;
  hi = hi > n_out ? n_out : hi;
  const float *p = pred + (size_t)t * n_out;
  const float th = thr[t];
  int hit = 0;
  for (int j = lo + lane; j < hi; j += 64) hit |= p[j] > th;
  const unsigned long long any = __ballot(hit);
  if (lane == 0) keep[idx] = any != 0ull;
}

// target[t][j] = max |cc[t][8j .. 8j+7]| over the valid outputs.
__global__ void peak_target_kernel(const float *__restrict__ cc, int n_corr,
                                   int n_out, int n_templates,
                                   float *__restrict__ target) {
  const int t = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_out) return;
  const float *c = cc + (size_t)t * n_corr + 8 * (size_t)j;
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (8 * j + i < n_corr) m = fmaxf(m, fabsf(c[i]));
  target[(size_t)t * n_out + j] = m;
}

}  // namespace

int launch_envelope(const float *data, int n_channels, int n_data, int n_dec,
                    float *env, hipStream_t s) {
  dim3 grid((n_dec + 255) / 256, n_channels);
  hipLaunchKernelGGL(envelope_kernel, grid, dim3(256), 0, s, data, n_data, n_dec, env);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_box_mean(const float *env, int n_channels, int n_dec, int width,
                    float *out, hipStream_t s) {
  dim3 grid((n_dec + 255) / 256, n_channels);
  hipLaunchKernelGGL(box_mean_kernel, grid, dim3(256), 0, s, env, n_dec,
                     width < 1 ? 1 : width, out);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_features(const float *win, const float *winh, const float *env,
                    const float *chinv, const int *moveouts,
                    const float *weights, int n_templates, int n_channels,
                    int n_dec, int n_out, int nwh, float len_feature,
                    float *feat, hipStream_t s) {
  dim3 grid((n_out + 255) / 256, n_templates);
  hipLaunchKernelGGL(features_kernel, grid, dim3(256), 0, s, win, winh, env,
                     chinv, moveouts, weights, n_channels, n_dec, n_out, nwh,
                     len_feature, feat);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_conv_fwd(const float *in, int nb, int cin, int nx, int relu_in,
                    const float *w, const float *bias, int cout, float *out,
                    hipStream_t s) {
  dim3 grid((nx + kTile - 1) / kTile, nb);
  const size_t lds = ((size_t)cin * kK * cout + (size_t)cin * (kTile + 2 * kPad)) * sizeof(float);
  if (lds > 160 * 1024) return set_error(kInvalidArgument, "conv layer too wide for LDS");
  if (cout == 32) {
    if (lds > 64 * 1024) {
      const int rc = ensure_lds_limit(reinterpret_cast<const void *>(&conv_fwd_kernel<32>),
                                      160 * 1024);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(conv_fwd_kernel<32>, grid, dim3(kTile), lds, s, in, cin, nx, relu_in, w, bias, out);
  } else if (cout == 1) {
    hipLaunchKernelGGL(conv_fwd_kernel<1>, grid, dim3(kTile), lds, s, in, cin, nx, relu_in, w, bias, out);
  } else {
    return set_error(kInvalidArgument, "conv layers have 32 or 1 output channels");
  }
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_conv_fwd_fast(const float *in, int nb, int cin, int nx, int relu_in,
                         const float *w, const float *bias, int cout, float *out,
                         void *wf_scratch, hipStream_t s) {
  if (cout != 32 || (cin != 8 && cin != 32))
    return set_error(kInvalidArgument, "matrix-core conv: cout 32 and cin 8 or 32 only");
  if (nb <= 0 || nx <= 0) return kOk;
  if (nb > 65535) return set_error(kInvalidArgument, "matrix-core conv: batch > 65535");
  const int chunks = cin == 32 ? MfmaConv<32>::kChunks : MfmaConv<8>::kChunks;
  _Float16 *wf = static_cast<_Float16 *>(wf_scratch);
  hipLaunchKernelGGL(conv_pack_kernel, dim3((chunks * 512 + 255) / 256), dim3(256), 0, s,
                     w, cin, chunks, wf);
  const dim3 grid((nx + kTile - 1) / kTile, nb);
  if (cin == 32)
    hipLaunchKernelGGL(conv_fwd_mfma_kernel<32>, grid, dim3(kTile), 0, s, in, nx, relu_in,
                       wf, bias, out);
  else
    hipLaunchKernelGGL(conv_fwd_mfma_kernel<8>, grid, dim3(kTile), 0, s, in, nx, relu_in,
                       wf, bias, out);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

size_t conv_fwd_fast_scratch_bytes(int cin) {
  const int chunks = cin == 32 ? MfmaConv<32>::kChunks : MfmaConv<8>::kChunks;
  return (size_t)chunks * 512 * sizeof(_Float16);
}

size_t screen_net_scratch_bytes() {
  return (size_t)(MfmaConv<8>::kChunks + 2 * MfmaConv<32>::kChunks) * 512 * sizeof(_Float16);
}

template <typename InT>
int launch_screen_net(const InT *feat, int nb, int nx, const float *w1, const float *b1,
                      const float *w2, const float *b2, const float *w3, const float *b3,
                      const float *w4, const float *b4, void *scratch, float *pred,
                      hipStream_t s) {
  if (nb <= 0 || nx <= 0) return kOk;
  if (nx >= (1 << 26))
    return set_error(kInvalidArgument, "screen_net: at most 2^26 - 1 positions per row");
  _Float16 *wf1 = static_cast<_Float16 *>(scratch);
  _Float16 *wf2 = wf1 + MfmaConv<8>::kChunks * 512;
  _Float16 *wf3 = wf2 + MfmaConv<32>::kChunks * 512;
  hipLaunchKernelGGL(conv_pack_kernel, dim3((MfmaConv<8>::kChunks * 512 + 255) / 256), dim3(256),
                     0, s, w1, 8, MfmaConv<8>::kChunks, wf1);
  hipLaunchKernelGGL(conv_pack_kernel, dim3((MfmaConv<32>::kChunks * 512 + 255) / 256),
                     dim3(256), 0, s, w2, 32, MfmaConv<32>::kChunks, wf2);
  hipLaunchKernelGGL(conv_pack_kernel, dim3((MfmaConv<32>::kChunks * 512 + 255) / 256),
                     dim3(256), 0, s, w3, 32, MfmaConv<32>::kChunks, wf3);
  NetArgs a{wf1, wf2, wf3, b1, b2, b3, w4, b4};
  const int n_xt = (nx + kNetOut - 1) / kNetOut;
  const long long tiles = (long long)nb * n_xt;
  // persistent: two blocks of 4 waves per CU (256 CUs; the weight fragments
  // of the three hidden layers take most of a wave's registers), each
  // walking tiles
  const unsigned blocks = (unsigned)std::min<long long>(tiles, 256 * kNetBlocksPerCu);
  hipLaunchKernelGGL(screen_net_kernel<InT>, dim3(blocks), dim3(kNetThreads), 0, s, feat, nb, nx,
                     n_xt, a, pred);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

template int launch_screen_net<float>(const float *, int, int, const float *, const float *,
                                      const float *, const float *, const float *,
                                      const float *, const float *, const float *, void *,
                                      float *, hipStream_t);
template int launch_screen_net<_Float16>(const _Float16 *, int, int, const float *,
                                         const float *, const float *, const float *,
                                         const float *, const float *, const float *,
                                         const float *, void *, float *, hipStream_t);

int launch_decimate(const float *in, long long n_rows, long long n_in, int D,
                    long long n_out, float *out, hipStream_t s) {
  if (n_rows <= 0 || n_out <= 0) return kOk;
  if (D < 1 || n_rows > 65535)
    return set_error(kInvalidArgument, "decimate: D >= 1 and at most 65535 rows");
  hipLaunchKernelGGL(decimate_kernel, dim3((unsigned)((n_out + 255) / 256), (unsigned)n_rows),
                     dim3(256), 0, s, in, n_in, D, n_out, out);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

template <typename OutT>
int launch_coarse_features(const float *cc, int n_templates, int n_coarse, const float *mad,
                           int D, int n_out, OutT *feat, hipStream_t s) {
  if (n_templates <= 0 || n_out <= 0) return kOk;
  if (D < 1 || D > 16 || n_templates > 65535)
    return set_error(kInvalidArgument, "coarse features: 1 <= D <= 16, at most 65535 templates");
  hipLaunchKernelGGL(coarse_features_kernel<OutT>, dim3((n_out + 255) / 256, n_templates),
                     dim3(256), 0, s, cc, n_coarse, mad, D, n_out, feat);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

template int launch_coarse_features<float>(const float *, int, int, const float *, int, int,
                                           float *, hipStream_t);
template int launch_coarse_features<_Float16>(const float *, int, int, const float *, int, int,
                                              _Float16 *, hipStream_t);

int launch_conv_bwd(const float *gz, const float *ain, const float *zin,
                    int nb, int cin, int nx, int relu_in, const float *w,
                    int cout, float *ga, float *gw, float *gb, hipStream_t s) {
  const size_t lds = (size_t)cout * (kTile + 2 * kPad) * sizeof(float);
  if (ga) {
    dim3 grid((nx + kTile - 1) / kTile, nb);
    hipLaunchKernelGGL(conv_bwd_data_kernel, grid, dim3(kTile), lds, s, gz,
                       cout, nx, w, cin, zin, relu_in, ga);
    FMF_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(conv_bwd_weight_kernel, dim3(cout * cin), dim3(kTile), 0,
                     s, gz, cout, ain, cin, relu_in, nb, nx, gw, gb);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_loss(const float *pred, const float *target, long long n,
                float boost, float inv_norm, float *grad, float *partial,
                int blocks, hipStream_t s) {
  hipLaunchKernelGGL(loss_kernel, dim3(blocks), dim3(kTile), 0, s, pred,
                     target, n, boost, inv_norm, grad, partial);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_adam(float *master, const float *grad, float *m, float *v, int n,
                float lr, float b1, float b2, float eps, int t, void *w16,
                float *w32, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                     master, grad, m, v, n, lr, b1, b2, eps, c1, c2,
                     static_cast<__half *>(w16), w32);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_gate(const float *pred, int n_out, const float *thr, int n_tiles,
                int tile_dec, int margin, int n_templates, int *keep,
                hipStream_t s) {
  // one wave per tile, kGateThreads / 64 tiles per block
  constexpr long long kTilesPerBlock = kGateThreads / 64;
  const long long n = (long long)n_templates * n_tiles;
  const long long blocks = (n + kTilesPerBlock - 1) / kTilesPerBlock;
  if (blocks > 0x7fffffffLL) return set_error(kInvalidArgument, "too many tiles (%lld)", n);
  hipLaunchKernelGGL(gate_kernel, dim3((unsigned)blocks), dim3(kGateThreads), 0, s, pred,
                     n_out, thr, n_tiles, tile_dec, margin, n_templates, keep);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_peak_target(const float *cc, int n_corr, int n_out,
                       int n_templates, float *target, hipStream_t s) {
  dim3 grid((n_out + 255) / 256, n_templates);
  hipLaunchKernelGGL(peak_target_kernel, grid, dim3(256), 0, s, cc, n_corr,
                     n_out, n_templates, target);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace fmf
