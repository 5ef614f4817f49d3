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
// decimated samples) exceeds thr[t].
__global__ void gate_kernel(const float *__restrict__ pred, int n_out,
                            const float *__restrict__ thr, int n_tiles,
                            int tile_dec, int margin, int n_templates,
                            int *__restrict__ keep) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_templates * n_tiles) return;
  const int t = idx / n_tiles, k = idx - t * n_tiles;
  int lo = k * tile_dec - margin, hi = (k + 1) * tile_dec + margin;
  lo = lo < 0 ? 0 : lo;
  hi = hi > n_out ? n_out : hi;
  const float *p = pred + (size_t)t * n_out;
  const float th = thr[t];
  int hit = 0;
  for (int j = lo; j < hi && !hit; j++) hit = p[j] > th;
  keep[idx] = hit;
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
    static bool raised = false;
    if (!raised && lds > 64 * 1024) {
      FMF_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&conv_fwd_kernel<32>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      raised = true;
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
  const int n = n_templates * n_tiles;
  hipLaunchKernelGGL(gate_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pred,
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
