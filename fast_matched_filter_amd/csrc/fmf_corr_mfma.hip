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

// cc = num / sqrt(sst * e) with the hardware reciprocal square root (about
// 1 ulp), as the reference GPU kernel does with rsqrtf
// (matched_filter.cu:136-138); the threshold is tested before the square
// root ('cpu'/'gpu') or on its value ('precise': sqrt(den) > thr <=> den > thr^2).
__device__ __forceinline__ float normalize_f(float num, float sst, float e,
                                             int mode) {
  const float den = sst * e;
  const float thr = mode == kThrBeforeSqrt
                        ? kStabilityThreshold
                        : kStabilityThreshold * kStabilityThreshold;
  return den > thr ? num * __builtin_amdgcn_rsqf(den) : 0.0f;
}

// W_ = waves per SIMD the register allocation is sized for (512 / W_ VGPRs
// each): eight tiles per wave keep ~360 VGPRs live (one wave), four tiles fit
// two waves without spilling and three with a few spills outside the MFMA
// loop.  The autotuner picks the pair.
template <int STEP, int G_, int W_>
__global__ __launch_bounds__(kThreads, W_) void corr_mfma_kernel(
    const float *__restrict__ templates, const float *__restrict__ sst,
    const int *__restrict__ moveouts, const float *__restrict__ weights,
    const float *__restrict__ data, const float *__restrict__ energy,
    const int *__restrict__ out_range, float *__restrict__ out,
    int n_templates, int n_channels, int n_tp_all, int n_pad, int n_data,
    int n_corr, const int *__restrict__ t_len,
    const long long *__restrict__ t_eoff, int mode, int summed, int n_tiles,
    int per_xcd, const int *__restrict__ items, int n_listed) {
  using Cfg = MCfg<STEP, G_>;
  constexpr int G = Cfg::kG;
  extern __shared__ uint4 lds_raw[];
  // Sizes and indices are 32-bit (the host checks L, n_corr < 2^31); only
  // element offsets of whole rows are formed in 64 bits.
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int c = lane & 15, g = lane >> 4;

  // Work items (output tile, template), tile-major, split into 8 contiguous
  // ranges, one per XCD (block b runs on XCD b mod 8).  The blocks of an XCD
  // stride through its range, so a grid smaller than the item count is a
  // persistent launch and a grid of 8 * per_xcd blocks is one item each.
  // With an item list (screened calls) the ranges run over the list.
  const int n_items = items ? n_listed : n_tiles * n_templates;
  const int xcd = blockIdx.x & 7;
  const int blocks_per_xcd = gridDim.x >> 3;
  for (int k = blockIdx.x >> 3; k < per_xcd; k += blocks_per_xcd) {
  const int wi = xcd * per_xcd + k;
  if (wi >= n_items) break;
  const int w = items ? items[wi] : wi;
  // listed items outside [0, n_tiles * n_templates) are skipped
  if ((unsigned)w >= (unsigned)(n_tiles * n_templates)) continue;
  const int tile = w / n_templates;
  const int t = w - tile * n_templates;
  // Per-template length (batched calls) and its window-energy table.
  int n_tp = t_len ? t_len[t] : n_tp_all;
  n_tp = n_tp < 1 ? 1 : (n_tp > n_tp_all ? n_tp_all : n_tp);
  const int n_m = (n_tp - 1 + 15 * STEP) / 32 + 1;
  const int n_win = n_data - n_tp + 1;
  const long long eoff = t_eoff ? t_eoff[t] : 0;
  const float *energy_t = energy + (eoff > 0 ? eoff : 0);
  const int n_tpl = tpl_len<STEP, G>(n_m);
  const int span = span_len<STEP, G>(n_m);
  uint16_t *tpl_hi = reinterpret_cast<uint16_t *>(lds_raw);
  uint16_t *tpl_lo = tpl_hi + n_tpl;
  uint16_t *dat_hi = tpl_lo + n_tpl;
  uint16_t *dat_lo = dat_hi + span;
  const int o_blk = tile * Cfg::kBlockOut;
  const int olo = out_range[2 * t], ohi = out_range[2 * t + 1];
  const int o_lane = o_blk + wave * Cfg::kWaveOut + 16 * c + 4 * g;
  const int tsc = t * n_channels;

  float sum[G][4];
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int j = 0; j < 4; j++) sum[q][j] = 0.0f;

  // eoff < 0: a batched template whose length has no energy table (the
  // caller's precondition is violated) — its row is written as zeros.
  const bool active = (o_blk < ohi) && (o_blk + Cfg::kBlockOut > olo) && eoff >= 0;
  const float *w_t = weights + tsc;
  // Next channel with a non-zero weight after `ch` (matched_filter.c:436).
  ActiveChannels act(w_t, n_channels);
  auto next_active = [&](int ch) { return act.next(ch); };
  // Staging is software-pipelined: the next channel's template and data
  // window are loaded into registers while the current channel runs on the
  // matrix cores, then split and written to LDS between two barriers.
  float pd[2 * Cfg::kDataPairs];
  float pt[2 * Cfg::kTplPairs];
  auto issue = [&](int ch) {
    const int base = o_blk * STEP + moveouts[tsc + ch];
    const float *dch = data + (size_t)ch * n_data;
    const float *tp = templates + (size_t)(tsc + ch) * n_pad;
#pragma unroll
    for (int k2 = 0; k2 < Cfg::kDataPairs; k2++) {
      const int i2 = tid + k2 * kThreads;
      const int gi = base + 2 * i2;
      const bool in = i2 < span / 2;
      pd[2 * k2] = (in && gi >= 0 && gi < n_data) ? dch[gi] : 0.0f;
      pd[2 * k2 + 1] = (in && gi + 1 >= 0 && gi + 1 < n_data) ? dch[gi + 1] : 0.0f;
    }
#pragma unroll
    for (int k2 = 0; k2 < Cfg::kTplPairs; k2++) {
      const int x = 2 * (tid + k2 * kThreads) - Cfg::kPadF;
      pt[2 * k2] = (x >= 0 && x < n_tp) ? tp[x] : 0.0f;
      pt[2 * k2 + 1] = (x + 1 >= 0 && x + 1 < n_tp) ? tp[x + 1] : 0.0f;
    }
  };

  int ch = active ? next_active(-1) : n_channels;
  if (ch < n_channels) issue(ch);
  while (ch < n_channels) {
    const float wch = w_t[ch];
    const int mo = moveouts[tsc + ch];
    __syncthreads();  // previous channel's fragments are consumed
#pragma unroll
    for (int k2 = 0; k2 < Cfg::kTplPairs; k2++) {
      const int i2 = tid + k2 * kThreads;
      if (i2 < n_tpl / 2) {
        uint32_t hi, lo;
        split2(pt[2 * k2], pt[2 * k2 + 1], hi, lo);
        reinterpret_cast<uint32_t *>(tpl_hi)[i2] = hi;
        reinterpret_cast<uint32_t *>(tpl_lo)[i2] = lo;
      }
    }
#pragma unroll
    for (int k2 = 0; k2 < Cfg::kDataPairs; k2++) {
      const int i2 = tid + k2 * kThreads;
      if (i2 < span / 2) {
        uint32_t hi, lo;
        split2(pd[2 * k2], pd[2 * k2 + 1], hi, lo);
        reinterpret_cast<uint32_t *>(dat_hi)[i2] = hi;
        reinterpret_cast<uint32_t *>(dat_lo)[i2] = lo;
      }
    }
    __syncthreads();
    const int nxt = next_active(ch);
    if (nxt < n_channels) issue(nxt);

    // Window energies of this lane's outputs, loaded ahead of the MFMA loop.
    const float *e_ch = energy_t + (size_t)ch * n_win + mo;
    float e[G][4];
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int o = o_lane + q * kTileOut + j;
        e[q][j] = (o >= olo && o < ohi) ? e_ch[o * STEP] : 0.0f;
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
    const float s = sst[tsc + ch];
#pragma unroll
    for (int q = 0; q < G; q++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int o = o_lane + q * kTileOut + j;
        if (o >= olo && o < ohi) {
          const float cc = normalize_f(acc[q][j], s, e[q][j], mode);
          if (summed)
            sum[q][j] = __builtin_fmaf(cc, wch, sum[q][j]);
          else
            out[((size_t)t * n_corr + o) * n_channels + ch] = cc;
        }
      }
    }
    ch = nxt;
  }
  if (summed) {
    float *row = out + (size_t)t * n_corr;
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int o = o_lane + q * kTileOut + j;
        if (o < n_corr) row[o] = (o >= olo && o < ohi) ? sum[q][j] : 0.0f;
      }
  }
  }  // work items
}

constexpr size_t kDefaultLds = 64 * 1024;
constexpr size_t kMaxLds = 160 * 1024;

template <int STEP, int G, int W>
int launch_step(const CorrArgs &a, hipStream_t s) {
  using Cfg = MCfg<STEP, G>;
  const long long n_tp_max = a.n_tp;
  const int n_m = (int)((n_tp_max - 1 + 15 * STEP) / 32 + 1);
  if (n_m > Cfg::kMaxM) return kUnsupported;
  const size_t lds = mfma_lds_bytes<STEP, G>(n_m);
  if (lds > kMaxLds) return kUnsupported;
  if (lds > kDefaultLds) {
    const int rc = ensure_lds_limit(
        reinterpret_cast<const void *>(&corr_mfma_kernel<STEP, G, W>), (int)kMaxLds);
    if (rc) return rc;
  }
  const long long n_tiles = (a.n_corr + Cfg::kBlockOut - 1) / Cfg::kBlockOut;
  const long long work = a.items ? a.n_items : n_tiles * a.n_templates;
  if (work == 0) return kOk;
  if (work > 0x7fffffffLL)
    return set_error(kInvalidArgument, "too many work items (%lld)", work);
  const long long per_xcd = (work + 7) / 8;
  long long blocks = per_xcd * 8;
  if (a.persistent_blocks > 0 && a.persistent_blocks < blocks)
    blocks = (a.persistent_blocks + 7) / 8 * 8;
  hipLaunchKernelGGL((corr_mfma_kernel<STEP, G, W>), dim3((unsigned)blocks),
                     dim3(kThreads), lds, s, a.templates, a.sst, a.moveouts,
                     a.weights, a.data, a.energy, a.out_range, a.out,
                     (int)a.n_templates, (int)a.n_channels, (int)a.n_tp,
                     (int)a.n_pad, (int)a.n_data, (int)a.n_corr, a.t_len,
                     a.t_eoff, a.threshold_mode, a.summed, (int)n_tiles,
                     (int)per_xcd, a.items, (int)a.n_items);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

// Tile shapes: code = 10 * tiles_per_wave + waves_per_simd.
template <int STEP>
int launch_tiles(const CorrArgs &a, int code, hipStream_t s) {
  switch (code) {
    case 24: return launch_step<STEP, 2, 4>(a, s);
    case 43: return launch_step<STEP, 4, 3>(a, s);
    case 81: return launch_step<STEP, 8, 1>(a, s);
    default: return launch_step<STEP, 4, 2>(a, s);
  }
}

}  // namespace

bool mfma16_covers(long long step, long long n_tp, long long n_data, long long n_corr,
                   long long n_channels, int summed) {
  (void)n_data;
  (void)n_corr;
  (void)n_channels;
  (void)summed;
  if (step != 1 && step != 2) return false;
  return n_tp >= 1 && (n_tp - 1 + 15 * step) / 32 + 1 <= MCfg<1, 4>::kMaxM;
}

// Returns kUnsupported (without launching) for steps or template lengths the
// MFMA tiling does not cover.
bool mfma16_shape(int shape) {
  return shape == 24 || shape == 42 || shape == 43 || shape == 81;
}

int mfma_block_outputs(int shape) {
  if (mfma32_shape(shape)) return mfma32_block_outputs(shape);
  return kWaves * (shape / 10) * kTileOut;
}

int launch_corr_mfma(const CorrArgs &a, int shape, hipStream_t s) {
  if (a.screen || mfma32_shape(shape)) return launch_corr_mfma32(a, shape, s);
  if (a.step == 1) return launch_tiles<1>(a, shape, s);
  if (a.step == 2) return launch_tiles<2>(a, shape, s);
  return kUnsupported;
}

}  // namespace fmf
