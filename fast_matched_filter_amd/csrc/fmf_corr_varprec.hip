// Variable-length, Fisher-combined matched filter (the 'variable_precise'
// path of the reference, fast_matched_filter/src/matched_filter.c:321-412,
// :569-606, :629-657).
//
// Every (template, channel) has its own length n[t][ch] <= Nmax (templates are
// stored with row stride Nmax).  Per channel the correlation follows
// corrc_precise: optional float window mean (normalize = 1), float sums of
// T*(d - m) and (d - m)^2 in tap order, den = sqrt(sst * ssd) tested after the
// square root.  Channels are combined through the Fisher transform
//   z = sum_ch atanh(cc - 10 eps) * cn[t][ch],  cn = w / sqrt(1 / (n - 3))
// and every one of the n_corr outputs becomes
//   (tanh(z / sum_ch cn) + used * eps * 10) * wsum[t]
// where wsum[t] is the running sum of the weights of templates 0..t (the
// reference never resets it) — outputs outside the lag range have z = 0.
// The per-template scalars are computed on the host in the reference's float
// operation order (fmf_api.hip).
//
// Work decomposition: one 256-thread block per (template, 2048 outputs); per
// channel the block stages the moveout-shifted window in LDS and each thread
// evaluates 8 consecutive outputs, two passes over the taps when the mean is
// removed.  This path is a precision path: it carries 3 flops per tap more
// than the short-normalised kernels and is not tuned further.
#include <float.h>

#include "fmf_internal.h"

namespace fmf {

namespace {

constexpr int kNT = 256;
constexpr int kR = 8;
constexpr int kTO = kNT * kR;  // outputs per block

template <bool LDS>
__global__ __launch_bounds__(kNT) void corr_varprec_kernel(
    const float *__restrict__ templates, const float *__restrict__ sst,
    const int *__restrict__ moveouts, const float *__restrict__ weights,
    const int *__restrict__ lengths, const float *__restrict__ cn,
    const float *__restrict__ tscal, const int *__restrict__ trange,
    const float *__restrict__ data, float *__restrict__ out, int n_templates,
    int n_channels, int n_max, int n_data, int n_corr, int step,
    int normalize, int n_tiles) {
  extern __shared__ float lds[];
  const int b = blockIdx.x;
  if (b >= n_tiles * n_templates) return;
  const int tile = b / n_templates;
  const int t = b - tile * n_templates;
  const int tid = threadIdx.x;
  const int o0 = tile * kTO;
  const int my_o = o0 + tid * kR;
  // valid output indices: lags i = o * step in [start, stop) of this
  // template (matched_filter.c:384-385)
  const int start = trange[2 * t], stop = trange[2 * t + 1];
  const int olo = start / step;
  int ohi = stop > start ? (stop + step - 1) / step : olo;
  ohi = ohi < n_corr ? ohi : n_corr;
  float z[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) z[r] = 0.0f;
  const bool active = o0 < ohi && o0 + kTO > olo;
  const int tsc = t * n_channels;
  if (active) {
    for (int ch = 0; ch < n_channels; ch++) {
      const float wch = weights[tsc + ch];
      if (wch == 0.0f) continue;  // matched_filter.c:588
      const int n = lengths[tsc + ch];
      const int mo = moveouts[tsc + ch];
      const int base = o0 * step + mo;
      const float *dch = data + (size_t)ch * n_data;
      const float *tp = templates + (size_t)(tsc + ch) * n_max;
      const int span = (kTO - 1) * step + n;
      if (LDS) {
        __syncthreads();
        for (int i = tid; i < span; i += kNT) {
          const int g = base + i;
          lds[i] = (g >= 0 && g < n_data) ? dch[g] : 0.0f;
        }
        __syncthreads();
      }
      auto at = [&](int i) -> float {
        if (LDS) return lds[i];
        const int g = base + i;
        return (g >= 0 && g < n_data) ? dch[g] : 0.0f;
      };
      const int lb = tid * kR * step;
      float mean[kR];
#pragma unroll
      for (int r = 0; r < kR; r++) mean[r] = 0.0f;
      if (normalize > 0) {
        for (int k = 0; k < n; k++)
#pragma unroll
          for (int r = 0; r < kR; r++) mean[r] += at(lb + r * step + k);
#pragma unroll
        for (int r = 0; r < kR; r++) mean[r] /= (float)n;
      }
      float num[kR], ssd[kR];
#pragma unroll
      for (int r = 0; r < kR; r++) num[r] = ssd[r] = 0.0f;
      for (int k = 0; k < n; k++) {
        const float tk = tp[k];
#pragma unroll
        for (int r = 0; r < kR; r++) {
          const float v = at(lb + r * step + k) - mean[r];
          num[r] = __builtin_fmaf(tk, v, num[r]);
          ssd[r] = __builtin_fmaf(v, v, ssd[r]);
        }
      }
      const float s = sst[tsc + ch];
      const float c = cn[tsc + ch];
#pragma unroll
      for (int r = 0; r < kR; r++) {
        const int o = my_o + r;
        if (o >= olo && o < ohi) {
          const float den = (float)sqrt((double)(s * ssd[r]));
          const float cc = den > kStabilityThreshold ? num[r] / den : 0.0f;
          z[r] += atanhf(cc - 10 * FLT_EPSILON) * c;  // matched_filter.c:601
        }
      }
    }
  }
  // matched_filter.c:405-408, over every output index
  const float zsum = tscal[3 * t + 0];
  const float eps_used = tscal[3 * t + 1];
  const float wsum = tscal[3 * t + 2];
#pragma unroll
  for (int r = 0; r < kR; r++) {
    const int o = my_o + r;
    if (o < n_corr)
      out[(size_t)t * n_corr + o] = (tanhf(z[r] / zsum) + eps_used) * wsum;
  }
}

}  // namespace

int launch_corr_varprec(const VarprecArgs &a, hipStream_t s) {
  const long long n_tiles = (a.n_corr + kTO - 1) / kTO;
  const long long blocks = n_tiles * a.n_templates;
  if (blocks == 0) return kOk;
  if (blocks > 0x7fffffffLL)
    return set_error(kInvalidArgument, "grid too large (%lld blocks)", blocks);
  const size_t span = (size_t)(kTO - 1) * a.step + a.n_max;
  const size_t lds = span * sizeof(float);
  if (lds <= 64 * 1024) {
    hipLaunchKernelGGL((corr_varprec_kernel<true>), dim3((unsigned)blocks),
                       dim3(kNT), lds, s, a.templates, a.sst, a.moveouts,
                       a.weights, a.lengths, a.cn, a.tscal, a.trange, a.data,
                       a.out, a.n_templates, a.n_channels, a.n_max, a.n_data,
                       a.n_corr, a.step, a.normalize, (int)n_tiles);
  } else {
    hipLaunchKernelGGL((corr_varprec_kernel<false>), dim3((unsigned)blocks),
                       dim3(kNT), 0, s, a.templates, a.sst, a.moveouts,
                       a.weights, a.lengths, a.cn, a.tscal, a.trange, a.data,
                       a.out, a.n_templates, a.n_channels, a.n_max, a.n_data,
                       a.n_corr, a.step, a.normalize, (int)n_tiles);
  }
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace fmf
