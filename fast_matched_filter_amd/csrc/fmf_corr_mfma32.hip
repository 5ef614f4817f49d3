// bf16x3 split-precision matched filter on 32x32x16 MFMA tiles.
//
// Same contract and numerics as the 16x16x32 kernel in fmf_corr_mfma.hip
// (network_corr/corrc of fast_matched_filter/src/matched_filter.c:416-455,
// :611-627): FP32 numerator from three bf16 products, FP32 window energies,
// the reference's stability threshold, weighted network sum over channels in
// the reference's order.  The Toeplitz-block form uses the larger MFMA:
//
//   out[o0 + i + 32n] = sum_m sum_{r<16} A_m[i][r] * B_m[r][n]
//   A_m[i][r] = T[16m + r - step*i]                   (32x16, template Toeplitz)
//   B_m[r][n] = d[mo + o0*step + 32*step*n + 16m + r]  (16x32, plain data rows)
//
// one v_mfma_f32_32x32x16_bf16 per (m, split product): 16384 MACs per
// instruction for the same 2 KiB of operands a 16x16x32 MFMA takes for 8192,
// so every byte read from LDS and every address instruction serves twice the
// arithmetic.
//
// Layout in LDS (bf16 planes, hi and lo of every FP32 value):
//   * the template, zero padded in front by 32*step taps, in two copies for
//     step 1 (the second shifted by one element) so that every lane's 8
//     consecutive taps start on a dword: A fragments are two ds_read2_b32 per
//     plane, with no funnel shifts;
//   * the moveout-shifted data window of the block, with 8 pad elements after
//     every 32*step elements.  Lane n of a B fragment reads 16 bytes at
//     element 32*step*n + 16m + 8h; the pad turns the row stride of 16*step
//     dwords (every lane on the same four banks) into 16*step + 4 dwords, and
//     the 16 B reads of 16 consecutive lanes then fall on distinct banks.
//
// Global memory is read and written through buffer resources whose record
// count is the row length: reads outside a template row, a data channel or an
// energy table return 0 and stores past the end of a cc row are dropped by the
// hardware, so staging and epilogue have neither clamps nor branches.
// Energies of a lane's 16*G outputs are loaded before the MFMA loop of each
// channel, the next channel's template and data window while it runs, and the
// LDS fragments of step m+1 while the MFMAs of step m execute.
//
// Screening variant (NP = 1, no reference counterpart): one bf16 product per
// MAC (hi planes only, about three significant digits) and an energy table
// that already holds 1/sqrt(e) (0 where e = 0), so the epilogue is
// cc = num * r_e * (1/sqrt(sst)) with no threshold test: a third of the
// MFMAs, half the staging and a quarter of the epilogue of the bf16x3 kernel,
// for passes whose output only ranks windows (the screener's coarse scan).
// Its energies are prefetched with coalesced loads (consecutive lanes on
// consecutive outputs) and staged in LDS, where the epilogue reads a lane's
// 4 runs of 4 outputs per tile as 16-byte reads, instead of the 16 scalar
// loads per tile whose 64 lanes touch 32 cache lines each; it runs on shape
// 414 (one tile per wave, 4 waves per SIMD).
#include "fmf_internal.h"

namespace fmf {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kTile = 1024;   // outputs of one 32x32 MFMA tile
constexpr int kMaxN16 = 84;   // 16-tap steps: N <= 16 * 84 - 31 * step

template <int STEP, int W, int G>
struct Cfg32 {
  static constexpr int kThreads = 64 * W;
  static constexpr int kBlockOut = W * G * kTile;
  static constexpr int kPad = 32 * STEP;          // template front zero pad
  static constexpr int kCopies = STEP == 1 ? 2 : 1;
  static constexpr int kChunkLog = STEP == 1 ? 5 : 6;  // 32 * STEP elements
  static constexpr int kTileStride = kTile * STEP + 8 * 32;  // padded tile
  static constexpr int kLaneStride = 32 * STEP + 8;          // padded row
  static constexpr int kStepLog = STEP == 1 ? 1 : 2;         // log2(2 * STEP)
  static constexpr int kSpanMax = (kBlockOut - 32) * STEP + 16 * (kMaxN16 + 1);
  static constexpr int kTplMax = 16 * (kMaxN16 + 1) + kPad + 8 + 64;
  static constexpr int kDataPairs = (kSpanMax / 2 + kThreads - 1) / kThreads;
  // dwords between the physical positions of consecutive staging pairs of a
  // thread (2 * kThreads elements, whole chunks of 32 * STEP plus their pads)
  static constexpr int kPairDwords = kThreads + 4 * (2 * kThreads / (32 * STEP));
  // pairs inside every window: the shortest span (n16 = 2) holds
  // (kBlockOut - 32) * STEP + 48 elements
  static constexpr int kSafePairs = ((kBlockOut - 32) * STEP + 48) / (2 * kThreads);
  static constexpr int kTplPairs = (kTplMax / 2 + kThreads - 1) / kThreads;
};

__host__ __device__ inline int round8(int v) { return (v + 7) & ~7; }

// 16-tap steps covering taps [-31 * step, n_tp) of every row of A, rounded
// up to a multiple of 2 (the MFMA loop runs four steps per iteration and a
// final pair when the count is 2 mod 4).
__host__ __device__ inline int n16_of(int n_tp, int step) {
  return ((n_tp + 31 * step + 15) / 16 + 1) & ~1;
}

// Template plane length: taps of n16 + 1 steps (the loop reads one step past
// the last) behind the front pad.  For step 1 the second copy starts 2 * lt
// elements (lt dwords) after the first; lt = 16 (mod 64) puts the dwords the
// odd lanes read from it on other banks than the even lanes' dwords of the
// first copy (with 32 or 64 banks).
template <int STEP, int W, int G>
__host__ __device__ inline int tpl_len32(int n16) {
  const int v = round8(16 * (n16 + 1) + Cfg32<STEP, W, G>::kPad + 8);
  return STEP == 1 ? v + ((16 - v) & 63) : v;
}

template <int STEP, int W, int G>
__host__ __device__ inline int span32(int n16) {
  return round8((Cfg32<STEP, W, G>::kBlockOut - 32) * STEP + 16 * (n16 + 1));
}

// Physical length of a padded data plane holding `span` elements.
template <int STEP, int W, int G>
__host__ __device__ inline int span32_phys(int span) {
  constexpr int chunk = 32 * STEP;
  return round8(span + 8 * ((span + chunk - 1) / chunk));
}

// Energies of a block's outputs staged in LDS (NP = 1): 4 pad floats after
// every 32, so that the 16-byte reads of the epilogue (lanes 32 floats apart)
// fall on distinct banks.
__host__ __device__ inline int epad(int p) { return p + 4 * (p >> 5); }

// Template planes (hi and lo, per copy), NP == 3 ? 2 : 1 data planes and, for
// NP = 1, the block's energies.
template <int STEP, int W, int G, int NP>
size_t mfma32_lds_bytes(int n16) {
  using C = Cfg32<STEP, W, G>;
  const int lt = tpl_len32<STEP, W, G>(n16);
  const int sp = span32_phys<STEP, W, G>(span32<STEP, W, G>(n16));
  const size_t e = NP == 1 ? (size_t)epad(C::kBlockOut) * sizeof(float) : 0;
  return (size_t)(2 * C::kCopies * lt + (NP == 3 ? 2 : 1) * sp) * sizeof(uint16_t) + e;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// hi = bf16(x) (round to nearest even), lo = bf16(x - hi), two values packed
// per dword: two packed conversions, two shifts/masks back to FP32, two
// subtractions.
__device__ __forceinline__ void split2(float x0, float x1, uint32_t &hi,
                                       uint32_t &lo) {
  const f32x2 x = {x0, x1};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
  const f32x2 h = {__builtin_bit_cast(float, hi << 16),
                   __builtin_bit_cast(float, hi & 0xffff0000u)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - h, bf16x2));
}

// hi = bf16(x) of two values packed per dword (the NP = 1 staging).
__device__ __forceinline__ uint32_t pack_bf16x2(float x0, float x1) {
  const f32x2 x = {x0, x1};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
}

// 8 bf16 starting at a dword-aligned LDS element.
__device__ __forceinline__ bf16x8 load4dw(const uint32_t *p) {
  uint32_t o[4];
#pragma unroll
  for (int i = 0; i < 4; i++) o[i] = p[i];
  return __builtin_bit_cast(bf16x8, o);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *base, unsigned bytes) {
  return buffer_rsrc(base, bytes);
}

__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, int off) {
  return buffer_load_f32(r, off);
}

__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, int off, float v) {
  buffer_store_f32(r, off, v);
}

// The value of v, opaque to the optimiser: per-element offsets derived from
// it are recomputed where they are used instead of being hoisted out of the
// channel loop, one register each.
__device__ __forceinline__ int opaque(int v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

// cc = num / sqrt(sst * e) with the hardware reciprocal square root, the
// threshold tested before the square root ('cpu'/'gpu') or on its value
// ('precise': sqrt(den) > thr <=> den > thr^2).
__device__ __forceinline__ float normalize32(float num, float sst, float e,
                                             int mode) {
  const float den = sst * e;
  const float thr = mode == kThrBeforeSqrt
                        ? kStabilityThreshold
                        : kStabilityThreshold * kStabilityThreshold;
  return den > thr ? num * __builtin_amdgcn_rsqf(den) : 0.0f;
}

// A and B fragments (hi and lo planes; hi only for NP = 1) of one 16-tap
// step for G tiles.  read<U>() reads step m + U for plane pointers at step m
// (m a multiple of 4): every address is the pointer plus an immediate offset.
template <int STEP, int W, int G, int NP>
struct Frag {
  bf16x8 ahi, alo, bhi[G], blo[G];
  template <int U>
  __device__ __forceinline__ void read(const uint32_t *a_hi, const uint32_t *a_lo,
                                       const uint16_t *b_hi, const uint16_t *b_lo) {
    using C = Cfg32<STEP, W, G>;
    ahi = load4dw(a_hi + 8 * U);
    if (NP == 3) alo = load4dw(a_lo + 8 * U);
    constexpr int bu = 16 * U + 8 * (U >> C::kStepLog);
#pragma unroll
    for (int q = 0; q < G; q++) {
      bhi[q] = *reinterpret_cast<const bf16x8 *>(b_hi + bu + q * C::kTileStride);
      if (NP == 3) blo[q] = *reinterpret_cast<const bf16x8 *>(b_lo + bu + q * C::kTileStride);
    }
  }
  __device__ __forceinline__ void mfma(f32x16 *acc) const {
    if (NP == 3) {
#pragma unroll
      for (int q = 0; q < G; q++)
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi[q], acc[q], 0, 0, 0);
#pragma unroll
      for (int q = 0; q < G; q++)
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo[q], acc[q], 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < G; q++)
      acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi[q], acc[q], 0, 0, 0);
  }
};

// W waves per block, G tiles of 1024 outputs per wave, launch bounds sized
// for WPS waves per SIMD, NP bf16 products per MAC (3: bf16x3; 1: screening).
template <int STEP, int W, int G, int WPS, int NP>
__global__ __launch_bounds__(64 * W, WPS) void corr_mfma32_kernel(
    const float *__restrict__ templates, const float *__restrict__ sst,
    const int *__restrict__ moveouts, const float *__restrict__ weights,
    const float *__restrict__ data, const float *__restrict__ energy,
    const int *__restrict__ out_range, float *__restrict__ out,
    int n_templates, int n_channels, int n_tp_all, int n_pad, int n_data,
    int n_corr, const int *__restrict__ t_len,
    const long long *__restrict__ t_eoff, int mode, int summed, int n_tiles,
    int per_xcd, const int *__restrict__ items, int n_listed, int rsq_inline) {
  using C = Cfg32<STEP, W, G>;
  constexpr int kThreads = C::kThreads;
  extern __shared__ uint4 lds_raw[];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int c = lane & 31, h = lane >> 5;

  // Work items (output tile, template), tile-major, in 8 contiguous ranges,
  // one per XCD (block b runs on XCD b mod 8); the blocks of an XCD stride
  // through its range (persistent launch when the grid is smaller).
  const int n_items = items ? n_listed : n_tiles * n_templates;
  const int xcd = blockIdx.x & 7;
  const int blocks_per_xcd = gridDim.x >> 3;
  for (int k = blockIdx.x >> 3; k < per_xcd; k += blocks_per_xcd) {
  const int wi = xcd * per_xcd + k;
  if (wi >= n_items) break;
  const int w = items ? items[wi] : wi;
  const int tile = w / n_templates;
  const int t = w - tile * n_templates;
  int n_tp = t_len ? t_len[t] : n_tp_all;
  n_tp = n_tp < 1 ? 1 : (n_tp > n_tp_all ? n_tp_all : n_tp);
  const int n16 = n16_of(n_tp, STEP);
  const int n_win = n_data - n_tp + 1;
  const long long eoff = t_eoff ? t_eoff[t] : 0;
  const float *energy_t = energy + (eoff > 0 ? eoff : 0);
  const int lt = tpl_len32<STEP, W, G>(n16);
  const int span = span32<STEP, W, G>(n16);
  const int sph = span32_phys<STEP, W, G>(span);
  uint16_t *tpl = reinterpret_cast<uint16_t *>(lds_raw);
  uint16_t *dat_hi = tpl + 2 * C::kCopies * lt;
  uint16_t *dat_lo = dat_hi + sph;
  float *ebuf = reinterpret_cast<float *>(dat_hi + sph);  // NP = 1 (16-byte aligned)
  const int o_blk = tile * C::kBlockOut;
  const int olo = out_range[2 * t], ohi = out_range[2 * t + 1];
  // First output of this lane in tile q: + q * 1024, then + 8 * (j / 4) +
  // (j % 4) for accumulator element j.
  const int o_lane = o_blk + wave * G * kTile + 32 * c + 4 * h;
  const int tsc = t * n_channels;

  float sum[G][16];
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int j = 0; j < 16; j++) sum[q][j] = 0.0f;

  // eoff < 0: a batched template whose length has no energy table; its row
  // is written as zeros.
  const bool active = (o_blk < ohi) && (o_blk + C::kBlockOut > olo) && eoff >= 0;
  const float *w_t = weights + tsc;
  auto next_active = [&](int ch) {  // matched_filter.c:436
    for (ch = ch + 1; ch < n_channels; ch++)
      if (w_t[ch] != 0.0f) break;
    return ch;
  };

  // Prefetch registers of the next channel: data pairs and template taps
  // (three per pair index for step 1: the shifted copy shares them).
  constexpr int kTplVals = STEP == 1 ? 3 : 2;
  float pd[2 * C::kDataPairs];
  float pt[kTplVals * C::kTplPairs];
  constexpr int kEPer = C::kBlockOut / kThreads;  // energies per thread (NP = 1)
  float pe[NP == 1 ? kEPer : 1];
  // Offsets stay non-negative: a window that starts before the data (a
  // negative moveout in the first block) takes the per-element path, where
  // elements before the data get the out-of-range offset kOob.
  auto issue = [&](int ch) {
    const int tid = opaque(threadIdx.x);
    const int base = o_blk * STEP + moveouts[tsc + ch];
    const auto rd = rsrc(data + (size_t)ch * n_data, 4u * (unsigned)n_data);
    const auto rt = rsrc(templates + (size_t)(tsc + ch) * n_pad, 4u * (unsigned)n_tp);
    if (base >= 0) {
      const int od = 4 * (base + 2 * tid);
#pragma unroll
      for (int k2 = 0; k2 < C::kDataPairs; k2++) {
        pd[2 * k2] = bload(rd, od + 8 * k2 * kThreads);
        pd[2 * k2 + 1] = bload(rd, od + 8 * k2 * kThreads + 4);
      }
    } else {
#pragma unroll
      for (int k2 = 0; k2 < C::kDataPairs; k2++) {
        const int gi = base + 2 * (tid + k2 * kThreads);
        pd[2 * k2] = bload(rd, gi >= 0 ? 4 * gi : kOob);
        pd[2 * k2 + 1] = bload(rd, gi + 1 >= 0 ? 4 * (gi + 1) : kOob);
      }
    }
    // Taps x = 2 * pair - pad: only the first pairs of the front pad are
    // negative (pad <= 64 < 2 * threads).
#pragma unroll
    for (int k2 = 0; k2 < C::kTplPairs; k2++) {
      const int x = 2 * (tid + k2 * kThreads) - C::kPad;
      if (k2 == 0) {
        if (STEP == 1) pt[0] = bload(rt, x - 1 >= 0 ? 4 * (x - 1) : kOob);
        pt[STEP == 1 ? 1 : 0] = bload(rt, x >= 0 ? 4 * x : kOob);
        pt[STEP == 1 ? 2 : 1] = bload(rt, x + 1 >= 0 ? 4 * (x + 1) : kOob);
      } else if (STEP == 1) {
        pt[3 * k2] = bload(rt, 4 * (x - 1));
        pt[3 * k2 + 1] = bload(rt, 4 * x);
        pt[3 * k2 + 2] = bload(rt, 4 * (x + 1));
      } else {
        pt[2 * k2] = bload(rt, 4 * x);
        pt[2 * k2 + 1] = bload(rt, 4 * (x + 1));
      }
    }
    if (NP == 1) {
      // Energies of the block's outputs, consecutive lanes on consecutive
      // outputs (two cache lines per load at step 1); staged in LDS and read
      // in the accumulator layout by the epilogue.
      const auto re = rsrc(energy_t + (size_t)ch * n_win, 4u * (unsigned)n_win);
      if (base >= 0) {
        const int oe = 4 * (base + tid * STEP);
#pragma unroll
        for (int v = 0; v < kEPer; v++) pe[v] = bload(re, oe + 4 * v * kThreads * STEP);
      } else {
#pragma unroll
        for (int v = 0; v < kEPer; v++) {
          const int x = base + (tid + v * kThreads) * STEP;
          pe[v] = bload(re, x >= 0 ? 4 * x : kOob);
        }
      }
    }
  };

  // Lane constants of the fragment reads (elements of the bf16 planes).
  // A: taps x0 = 16m + 8h - step*c + pad; copy p = x0 & 1 starts one element
  // later, so its element x0 + p is even.
  const int x0 = 8 * h - STEP * c + C::kPad;
  const int p = STEP == 1 ? (x0 & 1) : 0;
  const uint32_t *a_hi =
      reinterpret_cast<const uint32_t *>(tpl + 2 * p * lt + x0 + p);
  const uint32_t *a_lo = a_hi + lt / 2;
  // B: element e = tile*1024*step + 32*step*c + 16m + 8h of the window at
  // physical e + 8 * (e >> log2(32*step)); the m-dependent part of the pad is
  // 8 * (m >> log2(2*step)).
  const int b_lane = wave * G * C::kTileStride + c * C::kLaneStride + 8 * h;
  const uint16_t *b_hi = dat_hi + b_lane;
  const uint16_t *b_lo = dat_lo + b_lane;

  int ch = active ? next_active(-1) : n_channels;
  if (ch < n_channels) issue(ch);
  while (ch < n_channels) {
    const float wch = w_t[ch];
    const int mo = moveouts[tsc + ch];
    __syncthreads();  // previous channel's fragments are consumed
    {
    const int tid = opaque(threadIdx.x);
#pragma unroll
    for (int k2 = 0; k2 < C::kTplPairs; k2++) {
      const int y = 2 * (tid + k2 * kThreads);
      if (y < lt && NP == 1) {
        if (STEP == 1) {
          reinterpret_cast<uint32_t *>(tpl)[y / 2] = pack_bf16x2(pt[3 * k2 + 1], pt[3 * k2 + 2]);
          reinterpret_cast<uint32_t *>(tpl + 2 * lt)[y / 2] =
              pack_bf16x2(pt[3 * k2], pt[3 * k2 + 1]);
        } else {
          reinterpret_cast<uint32_t *>(tpl)[y / 2] = pack_bf16x2(pt[2 * k2], pt[2 * k2 + 1]);
        }
      } else if (y < lt) {
        uint32_t hi, lo;
        if (STEP == 1) {
          split2(pt[3 * k2 + 1], pt[3 * k2 + 2], hi, lo);
          reinterpret_cast<uint32_t *>(tpl)[y / 2] = hi;
          reinterpret_cast<uint32_t *>(tpl + lt)[y / 2] = lo;
          split2(pt[3 * k2], pt[3 * k2 + 1], hi, lo);
          reinterpret_cast<uint32_t *>(tpl + 2 * lt)[y / 2] = hi;
          reinterpret_cast<uint32_t *>(tpl + 3 * lt)[y / 2] = lo;
        } else {
          split2(pt[2 * k2], pt[2 * k2 + 1], hi, lo);
          reinterpret_cast<uint32_t *>(tpl)[y / 2] = hi;
          reinterpret_cast<uint32_t *>(tpl + lt)[y / 2] = lo;
        }
      }
    }
    // Pair k2 starts 2 * k2 * threads elements, a whole number of padded
    // chunks, after pair 0: its physical dword is pair 0's plus a constant.
    // The first kSafePairs pairs lie inside the shortest window (n16 = 2).
    const int e0 = 2 * tid;
    const int pw0 = (e0 + 8 * (e0 >> C::kChunkLog)) >> 1;
#pragma unroll
    for (int k2 = 0; k2 < C::kDataPairs; k2++) {
      const int e = e0 + 2 * k2 * kThreads;
      const int pw = pw0 + k2 * C::kPairDwords;
      if (k2 >= C::kSafePairs && e >= span) continue;
      if (NP == 1) {
        reinterpret_cast<uint32_t *>(dat_hi)[pw] = pack_bf16x2(pd[2 * k2], pd[2 * k2 + 1]);
      } else {
        uint32_t hi, lo;
        split2(pd[2 * k2], pd[2 * k2 + 1], hi, lo);
        reinterpret_cast<uint32_t *>(dat_hi)[pw] = hi;
        reinterpret_cast<uint32_t *>(dat_lo)[pw] = lo;
      }
    }
    if (NP == 1) {
      // epad(tid + v * threads) = epad(tid) + v * (threads + threads / 8)
      const int eb = epad(tid);
      if (rsq_inline) {  // the table holds energies: 1/sqrt here
#pragma unroll
        for (int v = 0; v < kEPer; v++)
          ebuf[eb + v * (kThreads + kThreads / 8)] =
              pe[v] > 0.0f ? __builtin_amdgcn_rsqf(pe[v]) : 0.0f;
      } else {
#pragma unroll
        for (int v = 0; v < kEPer; v++) ebuf[eb + v * (kThreads + kThreads / 8)] = pe[v];
      }
    }
    }
    __syncthreads();
    const int nxt = next_active(ch);
    if (nxt < n_channels) issue(nxt);

    // Window energies of this lane's outputs.  Outputs outside [olo, ohi)
    // read whatever lies there, or 0, and are masked when stored; windows
    // before the data (o * step + mo < 0, only below olo) read kOob.
    const auto re = rsrc(energy_t + (size_t)ch * n_win, 4u * (unsigned)n_win);
    float e[G][16];
    const int ol = opaque(o_lane);
    if (NP == 1) {
      // read from LDS by the epilogue
    } else if (o_blk * STEP + mo >= 0) {
      const int oe = 4 * (ol * STEP + mo);
#pragma unroll
      for (int q = 0; q < G; q++)
#pragma unroll
        for (int j = 0; j < 16; j++)
          e[q][j] = bload(re, oe + 4 * STEP * (q * kTile + 8 * (j >> 2) + (j & 3)));
    } else {
#pragma unroll
      for (int q = 0; q < G; q++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const int x = (ol + q * kTile + 8 * (j >> 2) + (j & 3)) * STEP + mo;
          e[q][j] = bload(re, x >= 0 ? 4 * x : kOob);
        }
    }

    f32x16 acc[G];
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 16; j++) acc[q][j] = 0.0f;

    // Two fragment sets in turn: those of step m + 1 are read while the
    // MFMAs of step m run.  With n16 a multiple of 4 the last read of the
    // loop is one step past the template (inside the planes, which hold
    // n16 + 1 steps; never used); otherwise a final pair of steps follows.
    // The scheduling barriers keep this order: left to itself the compiler
    // moves every read next to its MFMA and waits for it there.
    {
      const uint32_t *ah = a_hi, *al = a_lo;
      const uint16_t *bh = b_hi, *bl = b_lo;
      constexpr int kB4 = 64 + 8 * (4 >> C::kStepLog);  // 4 steps of B
      Frag<STEP, W, G, NP> f0, f1;
      f0.template read<0>(ah, al, bh, bl);
      int m = 0;
      for (; m + 4 <= n16; m += 4) {
        f1.template read<1>(ah, al, bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f0.mfma(acc);
        __builtin_amdgcn_sched_barrier(0);
        f0.template read<2>(ah, al, bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f1.mfma(acc);
        __builtin_amdgcn_sched_barrier(0);
        f1.template read<3>(ah, al, bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f0.mfma(acc);
        __builtin_amdgcn_sched_barrier(0);
        f0.template read<4>(ah, al, bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f1.mfma(acc);
        __builtin_amdgcn_sched_barrier(0);
        ah += 32;
        al += 32;
        bh += kB4;
        bl += kB4;
      }
      if (m < n16) {  // steps n16 - 2 (already in f0) and n16 - 1
        f1.template read<1>(ah, al, bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f0.mfma(acc);
        __builtin_amdgcn_sched_barrier(0);
        f1.mfma(acc);
        __builtin_amdgcn_sched_barrier(0);
      }
    }

    // Epilogue: normalise, weight, accumulate (or store per channel).  NP = 1:
    // e holds 1/sqrt(energy) and the template's factor is 1/sqrt(sst).
    const float s = sst[tsc + ch];
    const float rs = NP == 1 ? (s > 0.0f ? __builtin_amdgcn_rsqf(s) : 0.0f) : 0.0f;
    const float wrs = wch * rs;
    if (NP == 1) {
      // the lane's outputs: 4 runs of 4 consecutive ones per tile, each run
      // one aligned 16-byte read (o_lane - o_blk is a multiple of 4); the
      // runs of a tile share one padded chunk: epad(lo + 1024q + 8g) =
      // epad(lo) + 1152q + 8g
      const int eb = epad(ol - o_blk);
#pragma unroll
      for (int q = 0; q < G; q++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
          const float4 v =
              *reinterpret_cast<const float4 *>(ebuf + eb + q * (kTile + kTile / 8) + 8 * g);
          e[q][4 * g] = v.x;
          e[q][4 * g + 1] = v.y;
          e[q][4 * g + 2] = v.z;
          e[q][4 * g + 3] = v.w;
        }
    }
    const auto rc = rsrc(out + (size_t)t * n_corr * n_channels,
                         summed ? 0u : 4u * (unsigned)n_corr * (unsigned)n_channels);
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 16; j++) {
        if (NP == 1 && summed) {
          sum[q][j] = __builtin_fmaf(acc[q][j] * e[q][j], wrs, sum[q][j]);
          continue;
        }
        const float cc =
            NP == 1 ? acc[q][j] * e[q][j] * rs : normalize32(acc[q][j], s, e[q][j], mode);
        if (summed) {
          sum[q][j] = __builtin_fmaf(cc, wch, sum[q][j]);
        } else {
          // Per-channel output (t, o, ch); offsets within the template's
          // block of n_corr * n_channels values (host: < 2^29 values).
          const int o = opaque(o_lane) + q * kTile + 8 * (j >> 2) + (j & 3);
          bstore(rc, (o >= olo && o < ohi) ? 4 * (o * n_channels + ch) : kOob, cc);
        }
      }
    ch = nxt;
  }
  if (summed) {
    // Stores past n_corr fall outside the row's buffer and are dropped.
    const auto ro = rsrc(out + (size_t)t * n_corr, 4u * (unsigned)n_corr);
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int o = o_lane + q * kTile + 8 * (j >> 2) + (j & 3);
        bstore(ro, 4 * o, (o >= olo && o < ohi) ? sum[q][j] : 0.0f);
      }
  }
  }  // work items
}

constexpr size_t kDefaultLds = 64 * 1024;
constexpr size_t kMaxLds = 160 * 1024;

template <int STEP, int W, int G, int WPS, int NP>
int launch32(const CorrArgs &a, hipStream_t s) {
  using C = Cfg32<STEP, W, G>;
  const int n16 = n16_of((int)a.n_tp, STEP);
  if (a.n_tp > 0x7fffffffLL || n16 > kMaxN16) return kUnsupported;
  // Buffer offsets are 32-bit byte offsets within one channel or row.
  if (a.n_data >= (1LL << 29) || a.n_corr >= (1LL << 29)) return kUnsupported;
  if (!a.summed && a.n_corr * a.n_channels >= (1LL << 29)) return kUnsupported;
  const size_t lds = mfma32_lds_bytes<STEP, W, G, NP>(n16);
  if (lds > kMaxLds) return kUnsupported;
  if (lds > kDefaultLds) {
    static bool raised = false;
    if (!raised) {
      FMF_HIP_CHECK(hipFuncSetAttribute(
          reinterpret_cast<const void *>(&corr_mfma32_kernel<STEP, W, G, WPS, NP>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds));
      raised = true;
    }
  }
  const long long n_tiles = (a.n_corr + C::kBlockOut - 1) / C::kBlockOut;
  const long long work = a.items ? a.n_items : n_tiles * a.n_templates;
  if (work == 0) return kOk;
  if (work > 0x7fffffffLL)
    return set_error(kInvalidArgument, "too many work items (%lld)", work);
  const long long per_xcd = (work + 7) / 8;
  long long blocks = per_xcd * 8;
  if (a.persistent_blocks > 0 && a.persistent_blocks < blocks)
    blocks = (a.persistent_blocks + 7) / 8 * 8;
  hipLaunchKernelGGL((corr_mfma32_kernel<STEP, W, G, WPS, NP>), dim3((unsigned)blocks),
                     dim3(C::kThreads), lds, s, a.templates, a.sst, a.moveouts,
                     a.weights, a.data, a.energy, a.out_range, a.out,
                     (int)a.n_templates, (int)a.n_channels, (int)a.n_tp,
                     (int)a.n_pad, (int)a.n_data, (int)a.n_corr, a.t_len,
                     a.t_eoff, a.threshold_mode, a.summed, (int)n_tiles,
                     (int)per_xcd, a.items, (int)a.n_items, a.rsq_inline);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

// Tile shapes of this kernel: code = 100 * waves_per_block + 10 *
// tiles_per_wave + waves_per_simd.
template <int STEP>
int launch32_code(const CorrArgs &a, int code, hipStream_t s) {
  // The screening variant has its own shape: one tile per wave and 4 waves
  // per SIMD (120 registers).  Its channel iterations are short (10 16-tap
  // steps for a 100-tap coarse template), so occupancy, not the tile size,
  // hides their barriers and loads: 5.2 ms on the coarse day against 7.3 ms
  // with the bf16x3 default shape 422 (shapes 413, 424 and 814 run within
  // 5 % of it).  Code 212 asks for items of 2048 lags (two waves per block,
  // 3 waves per SIMD: 4 spill).
  if (a.screen) {
    if (code == 212) return launch32<STEP, 2, 1, 3, 1>(a, s);
    return launch32<STEP, 4, 1, 4, 1>(a, s);
  }
  switch (code) {
    case 412: return launch32<STEP, 4, 1, 2, 3>(a, s);
    case 212: return launch32<STEP, 2, 1, 2, 3>(a, s);
    case 413: return launch32<STEP, 4, 1, 3, 3>(a, s);
    case 421: return launch32<STEP, 4, 2, 1, 3>(a, s);
    case 242: return launch32<STEP, 2, 4, 2, 3>(a, s);
    case 241: return launch32<STEP, 2, 4, 1, 3>(a, s);
    default: return launch32<STEP, 4, 2, 2, 3>(a, s);  // 422
  }
}

}  // namespace

bool mfma32_shape(int code) {
  switch (code) {
    case 212: case 412: case 413: case 421: case 422: case 241: case 242: return true;
    default: return false;
  }
}

int mfma32_block_outputs(int code) {
  return (code / 100) * ((code / 10) % 10) * kTile;
}

int launch_corr_mfma32(const CorrArgs &a, int code, hipStream_t s) {
  if (a.step == 1) return launch32_code<1>(a, code, s);
  if (a.step == 2) return launch32_code<2>(a, code, s);
  return kUnsupported;
}

}  // namespace fmf
