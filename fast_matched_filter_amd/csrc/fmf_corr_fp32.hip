// Plain FP32 vector-ALU matched-filter kernel (the 'hip' arch).
//
// Computes, for every template t and output index o (lag i = o*step),
//   cc[t][o] = sum_{ch: w != 0} w[t][ch] * num / sqrt(sst[t][ch] * energy[ch][i+mo])
//   num      = sum_k T[t][ch][k] * d[ch][i + mo[t][ch] + k]
// the reference's network_corr/corrc pair
// (fast_matched_filter/src/matched_filter.c:416-455, :611-627) and its
// per-channel variant (:496-531) when summed == 0.
//
// Work decomposition (MI355X): one 256-thread block owns one template and
// 4096 consecutive outputs and walks the channels in the reference's order.
// Per channel the block stages the moveout-shifted data window
// (4095*step + N samples) in LDS with coalesced loads; each thread keeps 16
// consecutive outputs in registers and slides a 31-sample (step 1) window
// over them, so one ds_read_b128 feeds 32 FMAs.  Template taps are uniform
// across the block and are read with scalar loads.  The LDS image is padded
// by 4 floats per 16*step samples so that the 16-byte reads of 8 adjacent
// lanes land in disjoint banks.  Block indices are remapped so that each of
// the 8 XCDs walks a contiguous range of output tiles, all templates of a
// tile back to back, keeping the shared data window hot in that XCD's L2.
//
// Exact-order mode reproduces the compiled reference's float rounding
// sequence: products rounded then added in k order for k < 4*floor(N/4),
// fused multiply-adds for the last N mod 4 taps, FMA for the weighted sum,
// and the cc division in FP64 (gcc -O3 -ftree-vectorize's in-order
// vectorised reduction of matched_filter.c:618-621, :450, :625).
#include "fmf_internal.h"

namespace fmf {

namespace {

constexpr int kNT = 256;        // threads per block
constexpr int kR = 16;          // consecutive outputs per thread
constexpr int kKB = 16;         // taps per inner block
constexpr int kTO = kNT * kR;   // outputs per block
constexpr int kSlack = 64;      // staged samples beyond the last window

template <int STEP>
struct Lds {
  static constexpr int kChunk = kR * STEP;  // one thread's window origin stride
  static constexpr int kWin = ((kR - 1) * STEP + kKB + 3) / 4 * 4;
  __device__ __forceinline__ static int phys(int i) {
    return i + (i / kChunk) * 4;
  }
  // Same mapping in float4 units (i4 = i / 4 for i a multiple of 4).
  __device__ __forceinline__ static int phys4(int i4) {
    return i4 + i4 / (kChunk / 4);
  }
  static size_t bytes(long long n_pad) {
    long long span = (long long)(kTO - 1) * STEP + n_pad + kSlack;
    return (size_t)(span + (span / kChunk + 1) * 4) * sizeof(float);
  }
};

// Default-order epilogue: hardware reciprocal square root (about 1 ulp), as
// the reference GPU kernel (matched_filter.cu:136-138).
__device__ __forceinline__ float normalize_fast(float num, float sst, float e,
                                                int mode) {
  const float den = sst * e;
  const float thr = mode == kThrBeforeSqrt
                        ? kStabilityThreshold
                        : kStabilityThreshold * kStabilityThreshold;
  return den > thr ? num * __builtin_amdgcn_rsqf(den) : 0.0f;
}

// Exact-order epilogue: the reference CPU build's FP64 division.
__device__ __forceinline__ float normalize_cc(float num, float sst, float e,
                                              int mode) {
  const float den = sst * e;
  if (mode == kThrBeforeSqrt) {
    // matched_filter.c:622-625
    if (den > kStabilityThreshold)
      return (float)((double)num / sqrt((double)den));
    return 0.0f;
  }
  // matched_filter.c:650-655
  const float dq = (float)sqrt((double)den);
  if (dq > kStabilityThreshold) return num / dq;
  return 0.0f;
}

template <int STEP, bool EXACT>
__device__ __forceinline__ void accumulate_block(
    float (&acc)[kR], const float (&tv)[kKB], const float *win, int kb,
    int n_nf) {
  if (!EXACT) {
#pragma unroll
    for (int k = 0; k < kKB; k++)
#pragma unroll
      for (int r = 0; r < kR; r++)
        acc[r] = __builtin_fmaf(tv[k], win[r * STEP + k], acc[r]);
  } else {
#pragma clang fp contract(off)
    if (kb + kKB <= n_nf) {
#pragma unroll
      for (int k = 0; k < kKB; k++)
#pragma unroll
        for (int r = 0; r < kR; r++) {
          const float prod = tv[k] * win[r * STEP + k];
          acc[r] = acc[r] + prod;
        }
    } else {
#pragma unroll
      for (int k = 0; k < kKB; k++) {
        if (kb + k < n_nf) {
#pragma unroll
          for (int r = 0; r < kR; r++) {
            const float prod = tv[k] * win[r * STEP + k];
            acc[r] = acc[r] + prod;
          }
        } else {
#pragma unroll
          for (int r = 0; r < kR; r++)
            acc[r] = __builtin_fmaf(tv[k], win[r * STEP + k], acc[r]);
        }
      }
    }
  }
}

// STEP = 1 or 2: LDS-staged register-window path.  STEP = 0: any step,
// direct global loads (L1/L2 cached), for steps the LDS image cannot hold.
template <int STEP, bool EXACT>
__global__ __launch_bounds__(kNT) void corr_fp32_kernel(
    const float *__restrict__ templates, const float *__restrict__ sst,
    const int *__restrict__ moveouts, const float *__restrict__ weights,
    const float *__restrict__ data, const float *__restrict__ energy,
    const int *__restrict__ out_range, float *__restrict__ out,
    long long n_templates, long long n_channels, long long n_tp,
    long long n_pad, long long n_data, long long n_win, long long n_corr,
    int step_rt, int mode, int summed, long long n_tiles, long long per_xcd) {
  extern __shared__ float4 lds_v4[];
  float *lds = reinterpret_cast<float *>(lds_v4);
  const long long b = blockIdx.x;
  const long long w = (b & 7) * per_xcd + (b >> 3);
  if (w >= n_tiles * n_templates) return;
  const long long tile = w / n_templates;
  const long long t = w - tile * n_templates;
  const int step = STEP ? STEP : step_rt;
  const long long o0 = tile * kTO;
  const long long olo = out_range[2 * t], ohi = out_range[2 * t + 1];
  const int tid = threadIdx.x;
  const long long my_o = o0 + (long long)tid * kR;
  const int n_nf = (int)(n_tp / 4) * 4;  // exact mode: unfused taps

  float sum[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) sum[r] = 0.0f;

  const bool active = (o0 < ohi) && (o0 + kTO > olo);
  if (active) {
    for (long long ch = 0; ch < n_channels; ch++) {
      const float wch = weights[t * n_channels + ch];
      if (wch == 0.0f) continue;  // matched_filter.c:436
      const int mo = moveouts[t * n_channels + ch];
      const long long base = o0 * step + mo;
      const float *dch = data + ch * n_data;
      const float *tp = templates + (t * n_channels + ch) * n_pad;
      float acc[kR];
#pragma unroll
      for (int r = 0; r < kR; r++) acc[r] = 0.0f;

      if constexpr (STEP > 0) {
        const int span = (kTO - 1) * STEP + (int)n_pad + kSlack;
        __syncthreads();
        for (int i = tid; i < span; i += kNT) {
          const long long g = base + i;
          lds[Lds<STEP>::phys(i)] = (g >= 0 && g < n_data) ? dch[g] : 0.0f;
        }
        __syncthreads();
        const int lb = tid * kR * STEP;
        for (int kb = 0; kb < n_pad; kb += kKB) {
          float win[Lds<STEP>::kWin];
#pragma unroll
          for (int q = 0; q < Lds<STEP>::kWin / 4; q++) {
            const float4 v = lds_v4[Lds<STEP>::phys4((lb + kb) / 4 + q)];
            win[4 * q + 0] = v.x;
            win[4 * q + 1] = v.y;
            win[4 * q + 2] = v.z;
            win[4 * q + 3] = v.w;
          }
          float tv[kKB];
#pragma unroll
          for (int k = 0; k < kKB; k++) tv[k] = tp[kb + k];
          accumulate_block<STEP, EXACT>(acc, tv, win, kb, n_nf);
        }
      } else {
        const long long lb = base + (long long)tid * kR * step;
        for (int kb = 0; kb < n_pad; kb += kKB) {
          float tv[kKB];
#pragma unroll
          for (int k = 0; k < kKB; k++) tv[k] = tp[kb + k];
#pragma unroll
          for (int k = 0; k < kKB; k++) {
#pragma unroll
            for (int r = 0; r < kR; r++) {
              const long long g = lb + (long long)r * step + kb + k;
              const float x = (g >= 0 && g < n_data) ? dch[g] : 0.0f;
              if (EXACT && kb + k < n_nf) {
#pragma clang fp contract(off)
                const float prod = tv[k] * x;
                acc[r] = acc[r] + prod;
              } else {
                acc[r] = __builtin_fmaf(tv[k], x, acc[r]);
              }
            }
          }
        }
      }

      // Epilogue: normalise with the window energy, weight, accumulate.
      const float s = sst[t * n_channels + ch];
#pragma unroll
      for (int r = 0; r < kR; r++) {
        const long long o = my_o + r;
        if (o >= olo && o < ohi) {
          const long long p = o * step + mo;
          const float e = energy[ch * n_win + p];
          const float cc = EXACT ? normalize_cc(acc[r], s, e, mode)
                                 : normalize_fast(acc[r], s, e, mode);
          if (summed)
            sum[r] = __builtin_fmaf(cc, wch, sum[r]);  // matched_filter.c:450
          else
            out[(t * n_corr + o) * n_channels + ch] = cc;
        }
      }
    }
  }
  if (summed) {
#pragma unroll
    for (int r = 0; r < kR; r++) {
      const long long o = my_o + r;
      if (o < n_corr)
        out[t * n_corr + o] = (o >= olo && o < ohi) ? sum[r] : 0.0f;
    }
  }
}

template <int STEP, bool EXACT>
int launch_impl(const CorrArgs &a, hipStream_t s) {
  const long long n_tiles = (a.n_corr + kTO - 1) / kTO;
  const long long work = n_tiles * a.n_templates;
  if (work == 0) return kOk;
  const long long per_xcd = (work + 7) / 8;
  const long long blocks = per_xcd * 8;
  if (blocks > 0x7fffffffLL)
    return set_error(kInvalidArgument, "grid too large (%lld blocks)", blocks);
  size_t shmem = 0;
  if constexpr (STEP > 0) shmem = Lds<STEP>::bytes(a.n_pad);
  hipLaunchKernelGGL((corr_fp32_kernel<STEP, EXACT>), dim3((unsigned)blocks),
                     dim3(kNT), shmem, s, a.templates, a.sst, a.moveouts,
                     a.weights, a.data, a.energy, a.out_range, a.out,
                     a.n_templates, a.n_channels, a.n_tp, a.n_pad, a.n_data,
                     a.n_win, a.n_corr, a.step, a.threshold_mode, a.summed,
                     n_tiles, per_xcd);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace

// Largest LDS image the staged path is allowed to request.
constexpr size_t kMaxStagedLds = 64 * 1024;

int launch_corr_fp32(const CorrArgs &a, int exact_order, hipStream_t s) {
  if (a.n_pad % kKB != 0)
    return set_error(kInvalidArgument, "n_pad must be a multiple of %d", kKB);
  if (a.step == 1 && Lds<1>::bytes(a.n_pad) <= kMaxStagedLds)
    return exact_order ? launch_impl<1, true>(a, s) : launch_impl<1, false>(a, s);
  if (a.step == 2 && Lds<2>::bytes(a.n_pad) <= kMaxStagedLds)
    return exact_order ? launch_impl<2, true>(a, s) : launch_impl<2, false>(a, s);
  return exact_order ? launch_impl<0, true>(a, s) : launch_impl<0, false>(a, s);
}

}  // namespace fmf
