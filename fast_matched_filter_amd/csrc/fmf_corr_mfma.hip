// bf16x3 split-precision MFMA matched-filter kernel (the 'hip_mfma' arch).
//
// Same contract as the FP32 kernel (network_corr/corrc of
// fast_matched_filter/src/matched_filter.c:416-455, :611-627), with the
// numerator computed on the matrix cores.  Correlating one template channel
// against a run of 256 consecutive outputs is written as a Toeplitz-block
// product so that the moveout only shifts the data operand:
//
//   out[o0 + i + 16n] = sum_m sum_{r<32} A_m[i][r] * B_m[r][n]
//   A_m[i][r] = T[32m + r - step*i]                 (16x32, template Toeplitz)
//   B_m[r][n] = d[mo + o0*step + 16*step*n + 32m + r] (32x16, plain data rows)
//
// one v_mfma_f32_16x16x32_bf16 per (m, split product).  Every FP32 operand is
// split into bf16 hi + lo (x = hi + lo + O(2^-16 x)) and three products
// (lo*hi, hi*lo, hi*hi) are accumulated in FP32, which keeps the numerator at
// FP32-level accuracy (the dropped lo*lo term is O(2^-16) relative).
//
// Work decomposition (MI355X): a 256-thread block owns one template and
// 4 * G * 256 consecutive outputs and walks the template's channels in the
// reference's order.  Per channel it stages the template (hi/lo, zero-padded)
// and the moveout-shifted data window (hi/lo) in LDS, splitting on the fly.
// Each wave owns G output tiles: for every 32-tap step m it builds the A
// fragment once (unaligned LDS gather, two dword funnel shifts per pair) and
// feeds it to 3*G MFMAs whose B fragments are 16-byte LDS reads.  The
// normalisation by the window energy, the threshold and the weighted network
// sum are fused into the per-channel epilogue; cc sums stay in registers and
// are written once.  Blocks are remapped so each XCD walks a contiguous range
// of output tiles with all templates of a tile back to back (data window hot
// in that XCD's L2).
#include "fmf_internal.h"

namespace fmf {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kTileOut = 256;  // one 16x16 MFMA output tile

template <int STEP>
struct MCfg {
  static constexpr int kG = 8 / STEP;               // tiles per wave
  static constexpr int kWaveOut = kG * kTileOut;
  static constexpr int kBlockOut = kWaves * kWaveOut;
  static constexpr int kPadF = 16 * STEP;           // template front zero pad
};

__host__ __device__ inline int round8(int v) { return (v + 7) & ~7; }

template <int STEP>
__host__ __device__ inline int tpl_len(int n_m) {
  return round8(32 * n_m + MCfg<STEP>::kPadF + 8);
}

template <int STEP>
__host__ __device__ inline int span_len(int n_m) {
  return round8((MCfg<STEP>::kBlockOut - 16) * STEP + 32 * n_m);
}

template <int STEP>
size_t mfma_lds_bytes(int n_m) {
  return (size_t)(2 * tpl_len<STEP>(n_m) + 2 * span_len<STEP>(n_m)) * sizeof(uint16_t);
}

__device__ __forceinline__ void split2(float x0, float x1, uint32_t &hi,
                                       uint32_t &lo) {
  const __bf16 h0 = (__bf16)x0, h1 = (__bf16)x1;
  const __bf16 l0 = (__bf16)(x0 - (float)h0), l1 = (__bf16)(x1 - (float)h1);
  hi = (uint32_t)__builtin_bit_cast(uint16_t, h0) |
       ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
  lo = (uint32_t)__builtin_bit_cast(uint16_t, l0) |
       ((uint32_t)__builtin_bit_cast(uint16_t, l1) << 16);
}

// 8 consecutive bf16 starting at element x (any parity) of an LDS array.
__device__ __forceinline__ bf16x8 gather8(const uint32_t *a32, int x) {
  const int d = x >> 1;
  const uint32_t sh = (uint32_t)(x & 1) * 16u;
  uint32_t w[5];
#pragma unroll
  for (int i = 0; i < 5; i++) w[i] = a32[d + i];
  uint32_t o[4];
#pragma unroll
  for (int i = 0; i < 4; i++) o[i] = __builtin_amdgcn_alignbit(w[i + 1], w[i], sh);
  return __builtin_bit_cast(bf16x8, o);
}

__device__ __forceinline__ float normalize_f(float num, float sst, float e,
                                             int mode) {
  const float den = sst * e;
  if (mode == kThrBeforeSqrt)
    return den > kStabilityThreshold ? num / sqrtf(den) : 0.0f;
  const float dq = sqrtf(den);
  return dq > kStabilityThreshold ? num / dq : 0.0f;
}

template <int STEP>
__global__ __launch_bounds__(kThreads) void corr_mfma_kernel(
    const float *__restrict__ templates, const float *__restrict__ sst,
    const int *__restrict__ moveouts, const float *__restrict__ weights,
    const float *__restrict__ data, const float *__restrict__ energy,
    const int *__restrict__ out_range, float *__restrict__ out,
    long long n_templates, long long n_channels, long long n_tp,
    long long n_pad, long long n_data, long long n_win, long long n_corr,
    int n_m, int mode, int summed, long long n_tiles, long long per_xcd) {
  using Cfg = MCfg<STEP>;
  constexpr int G = Cfg::kG;
  extern __shared__ uint4 lds_raw[];
  const int n_tpl = tpl_len<STEP>(n_m);
  const int span = span_len<STEP>(n_m);
  uint16_t *tpl_hi = reinterpret_cast<uint16_t *>(lds_raw);
  uint16_t *tpl_lo = tpl_hi + n_tpl;
  uint16_t *dat_hi = tpl_lo + n_tpl;
  uint16_t *dat_lo = dat_hi + span;

  const long long b = blockIdx.x;
  const long long w = (b & 7) * per_xcd + (b >> 3);
  if (w >= n_tiles * n_templates) return;
  const long long tile = w / n_templates;
  const long long t = w - tile * n_templates;
  const long long o_blk = tile * Cfg::kBlockOut;
  const long long olo = out_range[2 * t], ohi = out_range[2 * t + 1];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int c = lane & 15, g = lane >> 4;
  const long long o_lane = o_blk + (long long)wave * Cfg::kWaveOut + 16 * c + 4 * g;

  float sum[G][4];
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int j = 0; j < 4; j++) sum[q][j] = 0.0f;

  const bool active = (o_blk < ohi) && (o_blk + Cfg::kBlockOut > olo);
  if (active) {
    for (long long ch = 0; ch < n_channels; ch++) {
      const float wch = weights[t * n_channels + ch];
      if (wch == 0.0f) continue;  // matched_filter.c:436
      const int mo = moveouts[t * n_channels + ch];
      const long long base = o_blk * STEP + mo;
      const float *tp = templates + (t * n_channels + ch) * n_pad;
      const float *dch = data + ch * n_data;

      __syncthreads();  // previous channel's fragments are consumed
      for (int i2 = tid; i2 < n_tpl / 2; i2 += kThreads) {
        const int x = 2 * i2 - Cfg::kPadF;
        const float v0 = (x >= 0 && x < n_tp) ? tp[x] : 0.0f;
        const float v1 = (x + 1 >= 0 && x + 1 < n_tp) ? tp[x + 1] : 0.0f;
        uint32_t hi, lo;
        split2(v0, v1, hi, lo);
        reinterpret_cast<uint32_t *>(tpl_hi)[i2] = hi;
        reinterpret_cast<uint32_t *>(tpl_lo)[i2] = lo;
      }
      for (int i2 = tid; i2 < span / 2; i2 += kThreads) {
        const long long gi = base + 2 * i2;
        const float v0 = (gi >= 0 && gi < n_data) ? dch[gi] : 0.0f;
        const float v1 = (gi + 1 >= 0 && gi + 1 < n_data) ? dch[gi + 1] : 0.0f;
        uint32_t hi, lo;
        split2(v0, v1, hi, lo);
        reinterpret_cast<uint32_t *>(dat_hi)[i2] = hi;
        reinterpret_cast<uint32_t *>(dat_lo)[i2] = lo;
      }
      __syncthreads();

      f32x4 acc[G];
#pragma unroll
      for (int q = 0; q < G; q++) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

      const uint32_t *th32 = reinterpret_cast<const uint32_t *>(tpl_hi);
      const uint32_t *tl32 = reinterpret_cast<const uint32_t *>(tpl_lo);
      const int xa = 8 * g - STEP * c + Cfg::kPadF;
      const int yb = wave * Cfg::kWaveOut * STEP + 16 * STEP * c + 8 * g;
      for (int m = 0; m < n_m; m++) {
        const bf16x8 a_hi = gather8(th32, 32 * m + xa);
        const bf16x8 a_lo = gather8(tl32, 32 * m + xa);
#pragma unroll
        for (int q = 0; q < G; q++) {
          const int y = yb + q * kTileOut * STEP + 32 * m;
          const bf16x8 b_hi = *reinterpret_cast<const bf16x8 *>(dat_hi + y);
          const bf16x8 b_lo = *reinterpret_cast<const bf16x8 *>(dat_lo + y);
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo, b_hi, acc[q], 0, 0, 0);
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, b_lo, acc[q], 0, 0, 0);
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, b_hi, acc[q], 0, 0, 0);
        }
      }

      // Epilogue: normalise, weight, accumulate (or store per channel).
      const float s = sst[t * n_channels + ch];
#pragma unroll
      for (int q = 0; q < G; q++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const long long o = o_lane + q * kTileOut + j;
          if (o >= olo && o < ohi) {
            const long long p = o * STEP + mo;
            const float cc = normalize_f(acc[q][j], s, energy[ch * n_win + p], mode);
            if (summed)
              sum[q][j] = __builtin_fmaf(cc, wch, sum[q][j]);
            else
              out[(t * n_corr + o) * n_channels + ch] = cc;
          }
        }
      }
    }
  }
  if (summed) {
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long long o = o_lane + q * kTileOut + j;
        if (o < n_corr) out[t * n_corr + o] = (o >= olo && o < ohi) ? sum[q][j] : 0.0f;
      }
  }
}

constexpr size_t kDefaultLds = 64 * 1024;
constexpr size_t kMaxLds = 160 * 1024;

template <int STEP>
int launch_step(const CorrArgs &a, hipStream_t s) {
  using Cfg = MCfg<STEP>;
  const int n_m = (int)((a.n_tp - 1 + 15 * STEP) / 32 + 1);
  const size_t lds = mfma_lds_bytes<STEP>(n_m);
  if (lds > kMaxLds) return kUnsupported;
  if (lds > kDefaultLds) {
    static bool raised = false;
    if (!raised) {
      FMF_HIP_CHECK(hipFuncSetAttribute(
          reinterpret_cast<const void *>(&corr_mfma_kernel<STEP>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds));
      raised = true;
    }
  }
  const long long n_tiles = (a.n_corr + Cfg::kBlockOut - 1) / Cfg::kBlockOut;
  const long long work = n_tiles * a.n_templates;
  if (work == 0) return kOk;
  const long long per_xcd = (work + 7) / 8;
  const long long blocks = per_xcd * 8;
  if (blocks > 0x7fffffffLL)
    return set_error(kInvalidArgument, "grid too large (%lld blocks)", blocks);
  hipLaunchKernelGGL((corr_mfma_kernel<STEP>), dim3((unsigned)blocks),
                     dim3(kThreads), lds, s, a.templates, a.sst, a.moveouts,
                     a.weights, a.data, a.energy, a.out_range, a.out,
                     a.n_templates, a.n_channels, a.n_tp, a.n_pad, a.n_data,
                     a.n_win, a.n_corr, n_m, a.threshold_mode, a.summed,
                     n_tiles, per_xcd);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace

// Returns kUnsupported (without launching) for steps or template lengths the
// MFMA tiling does not cover; the caller then runs the FP32 kernel.
int launch_corr_mfma(const CorrArgs &a, hipStream_t s) {
  if (a.step == 1) return launch_step<1>(a, s);
  if (a.step == 2) return launch_step<2>(a, s);
  return kUnsupported;
}

}  // namespace fmf
