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

template <int STEP, int G_>
struct MCfg {
  static constexpr int kG = G_;                     // tiles per wave
  static constexpr int kWaveOut = kG * kTileOut;
  static constexpr int kBlockOut = kWaves * kWaveOut;
  static constexpr int kPadF = 16 * STEP;           // template front zero pad
  static constexpr int kMaxM = 40;                  // 32-tap steps (N <= 1280 - 15*STEP)
  static constexpr int kSpanMax = ((kBlockOut - 16) * STEP + 32 * kMaxM + 7) & ~7;
  static constexpr int kTplMax = (32 * kMaxM + kPadF + 8 + 7) & ~7;
  static constexpr int kDataPairs = (kSpanMax / 2 + kThreads - 1) / kThreads;
  static constexpr int kTplPairs = (kTplMax / 2 + kThreads - 1) / kThreads;
};

__host__ __device__ inline int round8(int v) { return (v + 7) & ~7; }

template <int STEP, int G>
__host__ __device__ inline int tpl_len(int n_m) {
  return round8(32 * n_m + MCfg<STEP, G>::kPadF + 8);
}

template <int STEP, int G>
__host__ __device__ inline int span_len(int n_m) {
  return round8((MCfg<STEP, G>::kBlockOut - 16) * STEP + 32 * n_m);
}

template <int STEP, int G>
size_t mfma_lds_bytes(int n_m) {
  return (size_t)(2 * tpl_len<STEP, G>(n_m) + 2 * span_len<STEP, G>(n_m)) * sizeof(uint16_t);
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

// Waves per SIMD the register allocation is sized for: eight tiles per wave
// keep ~350 VGPRs live (one wave per SIMD), fewer tiles allow two or more.
template <int G>
constexpr int min_waves() { return G >= 8 ? 1 : 2; }

template <int STEP, int G_>
__global__ __launch_bounds__(kThreads, min_waves<G_>()) void corr_mfma_kernel(
    const float *__restrict__ templates, const float *__restrict__ sst,
    const int *__restrict__ moveouts, const float *__restrict__ weights,
    const float *__restrict__ data, const float *__restrict__ energy,
    const int *__restrict__ out_range, float *__restrict__ out,
    long long n_templates, long long n_channels, long long n_tp,
    long long n_pad, long long n_data, long long n_win, long long n_corr,
    int n_m, int mode, int summed, long long n_tiles, long long per_xcd) {
  using Cfg = MCfg<STEP, G_>;
  constexpr int G = Cfg::kG;
  extern __shared__ uint4 lds_raw[];
  const int n_tpl = tpl_len<STEP, G>(n_m);
  const int span = span_len<STEP, G>(n_m);
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
  const float *w_t = weights + t * n_channels;
  // Next channel with a non-zero weight after `ch` (matched_filter.c:436).
  auto next_active = [&](long long ch) {
    for (ch = ch + 1; ch < n_channels; ch++)
      if (w_t[ch] != 0.0f) break;
    return ch;
  };
  // Staging is software-pipelined: the next channel's template and data
  // window are loaded into registers while the current channel runs on the
  // matrix cores, then split and written to LDS between two barriers.
  float pd[2 * Cfg::kDataPairs];
  float pt[2 * Cfg::kTplPairs];
  auto issue = [&](long long ch) {
    const long long base = o_blk * STEP + moveouts[t * n_channels + ch];
    const float *dch = data + ch * n_data;
    const float *tp = templates + (t * n_channels + ch) * n_pad;
#pragma unroll
    for (int k = 0; k < Cfg::kDataPairs; k++) {
      const int i2 = tid + k * kThreads;
      const long long gi = base + 2 * i2;
      const bool in = i2 < span / 2;
      pd[2 * k] = (in && gi >= 0 && gi < n_data) ? dch[gi] : 0.0f;
      pd[2 * k + 1] = (in && gi + 1 >= 0 && gi + 1 < n_data) ? dch[gi + 1] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < Cfg::kTplPairs; k++) {
      const int x = 2 * (tid + k * kThreads) - Cfg::kPadF;
      pt[2 * k] = (x >= 0 && x < n_tp) ? tp[x] : 0.0f;
      pt[2 * k + 1] = (x + 1 >= 0 && x + 1 < n_tp) ? tp[x + 1] : 0.0f;
    }
  };

  long long ch = active ? next_active(-1) : n_channels;
  if (ch < n_channels) issue(ch);
  while (ch < n_channels) {
    const float wch = w_t[ch];
    const int mo = moveouts[t * n_channels + ch];
    __syncthreads();  // previous channel's fragments are consumed
#pragma unroll
    for (int k = 0; k < Cfg::kTplPairs; k++) {
      const int i2 = tid + k * kThreads;
      if (i2 < n_tpl / 2) {
        uint32_t hi, lo;
        split2(pt[2 * k], pt[2 * k + 1], hi, lo);
        reinterpret_cast<uint32_t *>(tpl_hi)[i2] = hi;
        reinterpret_cast<uint32_t *>(tpl_lo)[i2] = lo;
      }
    }
#pragma unroll
    for (int k = 0; k < Cfg::kDataPairs; k++) {
      const int i2 = tid + k * kThreads;
      if (i2 < span / 2) {
        uint32_t hi, lo;
        split2(pd[2 * k], pd[2 * k + 1], hi, lo);
        reinterpret_cast<uint32_t *>(dat_hi)[i2] = hi;
        reinterpret_cast<uint32_t *>(dat_lo)[i2] = lo;
      }
    }
    __syncthreads();
    const long long nxt = next_active(ch);
    if (nxt < n_channels) issue(nxt);

    // Window energies of this lane's outputs, loaded ahead of the MFMA loop.
    float e[G][4];
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long long o = o_lane + q * kTileOut + j;
        e[q][j] = (o >= olo && o < ohi) ? energy[ch * n_win + o * STEP + mo] : 0.0f;
      }

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
          const float cc = normalize_f(acc[q][j], s, e[q][j], mode);
          if (summed)
            sum[q][j] = __builtin_fmaf(cc, wch, sum[q][j]);
          else
            out[(t * n_corr + o) * n_channels + ch] = cc;
        }
      }
    }
    ch = nxt;
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

template <int STEP, int G>
int launch_step(const CorrArgs &a, hipStream_t s) {
  using Cfg = MCfg<STEP, G>;
  const int n_m = (int)((a.n_tp - 1 + 15 * STEP) / 32 + 1);
  if (n_m > Cfg::kMaxM) return kUnsupported;
  const size_t lds = mfma_lds_bytes<STEP, G>(n_m);
  if (lds > kMaxLds) return kUnsupported;
  if (lds > kDefaultLds) {
    static bool raised = false;
    if (!raised) {
      FMF_HIP_CHECK(hipFuncSetAttribute(
          reinterpret_cast<const void *>(&corr_mfma_kernel<STEP, G>),
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
  hipLaunchKernelGGL((corr_mfma_kernel<STEP, G>), dim3((unsigned)blocks),
                     dim3(kThreads), lds, s, a.templates, a.sst, a.moveouts,
                     a.weights, a.data, a.energy, a.out_range, a.out,
                     a.n_templates, a.n_channels, a.n_tp, a.n_pad, a.n_data,
                     a.n_win, a.n_corr, n_m, a.threshold_mode, a.summed,
                     n_tiles, per_xcd);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

template <int STEP>
int launch_tiles(const CorrArgs &a, int tiles, hipStream_t s) {
  switch (tiles) {
    case 2: return launch_step<STEP, 2>(a, s);
    case 8: return launch_step<STEP, 8>(a, s);
    default: return launch_step<STEP, 4>(a, s);
  }
}

}  // namespace

// Returns kUnsupported (without launching) for steps or template lengths the
// MFMA tiling does not cover; the caller then runs the FP32 kernel.
int launch_corr_mfma(const CorrArgs &a, int tiles_per_wave, hipStream_t s) {
  if (a.step == 1) return launch_tiles<1>(a, tiles_per_wave, s);
  if (a.step == 2) return launch_tiles<2>(a, tiles_per_wave, s);
  return kUnsupported;
}

}  // namespace fmf
