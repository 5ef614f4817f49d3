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
// Operands are split into bf16 hi and lo once per call, not per work item:
//   * the data as four planes per channel (hi and lo of pairs of consecutive
//     samples, once starting at an even and once at an odd sample), so that a
//     window starting at any sample is a run of whole dwords of one of them;
//   * every template channel as an image of its LDS layout (below).
// A block then stages a channel with 16-byte copies (global -> registers ->
// LDS) and no arithmetic; only a window that starts before the data (a
// negative moveout in the first block of a row) is split on the fly from the
// FP32 samples.
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
// count is the row length: reads outside a template image, a data plane or
// an energy table return 0 and stores past the end of a cc row are dropped
// by the hardware, so staging and epilogue have neither clamps nor branches.
// The energy table holds r = 1/sqrt(e) (0 where e = 0), so the epilogue is
// cc = num * r * (1/sqrt(sst)) with the threshold sst * e > thr tested as
// r < sqrt(sst / thr): no square root per output.
//
// Screening variant (NP = 1, no reference counterpart): one bf16 product per
// MAC (hi planes only, about three significant digits) and no threshold test:
// a third of the MFMAs, half the staging and a quarter of the epilogue of the
// bf16x3 kernel, for passes whose output only ranks windows (the screener's
// coarse scan).  Its energies are prefetched with coalesced loads
// (consecutive lanes on consecutive outputs) and staged in LDS, where the
// epilogue reads a lane's 4 runs of 4 outputs per tile as 16-byte reads; it
// runs on shape 414 (one tile per wave, 4 waves per SIMD).
#include <mutex>
#include <set>
#include <type_traits>
#include <utility>

#include "fmf_internal.h"

namespace fmf {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 1024;   // outputs of one 32x32 MFMA tile
constexpr int kMaxN16 = 84;   // 16-tap steps: N <= 16 * 84 - 31 * step

template <int STEP>
struct Pad32 {
  static constexpr int kPad = 32 * STEP;                // template front zero pad
  static constexpr int kCopies = STEP == 1 ? 2 : 1;
  static constexpr int kChunkLog = STEP == 1 ? 5 : 6;   // 32 * STEP elements
};

// IL = 1 (interleaved tiles, step 1): the G tiles of a wave are 32 outputs
// apart and their B columns 32 * G outputs apart, so that tile q's data
// fragment at 16-tap step m is tile 0's fragment at step m + 2q: a wave reads
// each data fragment once and uses it for every tile.
template <int STEP, int W, int G, int IL = 0>
struct Cfg32 : Pad32<STEP> {
  static_assert(!IL || STEP == 1, "interleaved tiles are built for step 1");
  static constexpr int kThreads = 64 * W;
  static constexpr int kBlockOut = W * G * kTile;
  // data elements between consecutive B columns (a pad of 8 follows each)
  static constexpr int kChunk = IL ? 32 * G : 32 * STEP;
  static constexpr int kChunkLog = kChunk == 32 ? 5 : (kChunk == 64 ? 6 : 7);
  static_assert((1 << kChunkLog) == kChunk, "B column spacing is a power of 2");
  static constexpr int kLaneStride = kChunk + 8;             // padded column
  // padded elements of one wave's G * 1024 outputs; of one tile (IL = 0)
  static constexpr int kWaveStride = G * kTile * STEP + 8 * (G * kTile * STEP / kChunk);
  static constexpr int kTileStride = kTile * STEP + 8 * (kTile * STEP / kChunk);
  // outputs between tiles of a wave and between B columns of a tile
  static constexpr int kTileOff = IL ? 32 : kTile;
  static constexpr int kColOff = IL ? 32 * G : 32;
  static constexpr int kStepLog = kChunkLog - 4;             // log2(16-tap steps per chunk)
  // B fragments a step reads beyond the last template step: tile G - 1 runs
  // 2 * (G - 1) steps ahead of tile 0 (IL)
  static constexpr int kAhead = IL ? 2 * (G - 1) : 0;
  static constexpr int kSpanMax = (kBlockOut - kColOff) * STEP + 16 * (kMaxN16 + 1 + kAhead);
  static constexpr int kTplMax = 16 * (kMaxN16 + 1) + 32 * STEP + 8 + 64;
  // 8-element groups of the data window per thread
  static constexpr int kGroups = (kSpanMax / 8 + kThreads - 1) / kThreads;
  // groups inside every window: the shortest span (n16 = 2) holds
  // (kBlockOut - kColOff) * STEP + 16 * (3 + kAhead) elements
  static constexpr int kSafeGroups =
      ((kBlockOut - kColOff) * STEP + 16 * (3 + kAhead)) / (8 * kThreads);
  // 8-element chunks of the template image per thread
  static constexpr int kTplChunks =
      (2 * Pad32<STEP>::kCopies * kTplMax / 8 + kThreads - 1) / kThreads;
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
template <int STEP>
__host__ __device__ inline int tpl_len32(int n16) {
  const int v = round8(16 * (n16 + 1) + Pad32<STEP>::kPad + 8);
  return STEP == 1 ? v + ((16 - v) & 63) : v;
}

template <int STEP, int W, int G, int IL = 0>
__host__ __device__ inline int span32(int n16) {
  using C = Cfg32<STEP, W, G, IL>;
  return round8((C::kBlockOut - C::kColOff) * STEP + 16 * (n16 + 1 + C::kAhead));
}

// Physical length of a padded data plane holding `span` elements, 8 pad
// elements after every `chunk`.
template <int CHUNK>
__host__ __device__ inline int span32_phys(int span) {
  return round8(span + 8 * ((span + CHUNK - 1) / CHUNK));
}

// Energies of a block's outputs staged in LDS (NP = 1): 4 pad floats after
// every 32, so that the 16-byte reads of the epilogue (lanes 32 floats apart)
// fall on distinct banks.
__host__ __device__ inline int epad(int p) { return p + 4 * (p >> 5); }

// Template planes (hi and lo, per copy), NP == 3 ? 2 : 1 data planes and, for
// NP = 1, the block's energies.
template <int STEP, int W, int G, int NP, int IL = 0>
size_t mfma32_lds_bytes(int n16) {
  using C = Cfg32<STEP, W, G, IL>;
  const int lt = tpl_len32<STEP>(n16);
  const int sp = span32_phys<C::kChunk>(span32<STEP, W, G, IL>(n16));
  const size_t e = NP == 1 ? (size_t)epad(C::kBlockOut) * sizeof(float) : 0;
  return (size_t)(2 * C::kCopies * lt + (NP == 3 ? 2 : 1) * sp) * sizeof(uint16_t) + e;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// hi = bf16(x) (round to nearest even), lo = bf16(x - hi), two values packed
// per dword (first value in the low half).
__device__ __forceinline__ void split2(float x0, float x1, uint32_t &hi,
                                       uint32_t &lo) {
  const f32x2 x = {x0, x1};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
  const f32x2 h = {__builtin_bit_cast(float, hi << 16),
                   __builtin_bit_cast(float, hi & 0xffff0000u)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - h, bf16x2));
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

__device__ __forceinline__ u32x4 bload4(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
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

// A and B fragments (hi and lo planes; hi only for NP = 1) of one 16-tap
// step for G tiles.  read<U>() reads step m + U for plane pointers at step m
// (m a multiple of 4): every address is the pointer plus an immediate offset.
template <int STEP, int W, int G, int NP, int IL>
struct Frag {
  bf16x8 ahi, alo, bhi[G], blo[G];
  template <int U>
  __device__ __forceinline__ void read(const uint32_t *a_hi, const uint32_t *a_lo,
                                       const uint16_t *b_hi, const uint16_t *b_lo) {
    using C = Cfg32<STEP, W, G, IL>;
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

// Fragments of the interleaved-tile loop (IL = 1, step 1, G = 2): A of step
// m in slot m & 1, data fragment F_s (B columns 64 outputs apart) in slot
// s & 3.  Step m runs tile 0 on F_m and tile 1 on F_{m+2}; while it runs,
// A_{m+1} and F_{m+3} are read.  Pointers are at a step m0 that is a multiple
// of 4: every address is the pointer plus an immediate offset.
struct RingIL {
  bf16x8 ahi[2], alo[2], bhi[4], blo[4];
  template <int U>
  __device__ __forceinline__ void read_a(const uint32_t *a_hi, const uint32_t *a_lo) {
    ahi[U & 1] = load4dw(a_hi + 8 * U);
    alo[U & 1] = load4dw(a_lo + 8 * U);
  }
  template <int U>
  __device__ __forceinline__ void read_f(const uint16_t *b_hi, const uint16_t *b_lo) {
    constexpr int bu = 16 * U + 8 * (U >> 2);  // a pad of 8 after every 4 steps
    bhi[U & 3] = *reinterpret_cast<const bf16x8 *>(b_hi + bu);
    blo[U & 3] = *reinterpret_cast<const bf16x8 *>(b_lo + bu);
  }
  // the first step of a channel: the first product of each tile starts from
  // zero (an inline-constant accumulator input instead of zeroed registers)
  template <int U>
  __device__ __forceinline__ void mfma_first(f32x16 *acc) const {
    constexpr int a = U & 1, f0 = U & 3, f1 = (U + 2) & 3;
    const f32x16 z = {};
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[a], bhi[f0], z, 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[a], bhi[f1], z, 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], blo[f0], acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], blo[f1], acc[1], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], bhi[f0], acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], bhi[f1], acc[1], 0, 0, 0);
  }
  template <int U>
  __device__ __forceinline__ void mfma(f32x16 *acc) const {
    constexpr int a = U & 1, f0 = U & 3, f1 = (U + 2) & 3;
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[a], bhi[f0], acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[a], bhi[f1], acc[1], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], blo[f0], acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], blo[f1], acc[1], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], bhi[f0], acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[a], bhi[f1], acc[1], 0, 0, 0);
  }
};

// W waves per block, G tiles of 1024 outputs per wave, launch bounds sized
// for WPS waves per SIMD, NP bf16 products per MAC (3: bf16x3; 1: screening),
// IL interleaved tiles (Cfg32).
template <int STEP, int W, int G, int WPS, int NP, int IL>
__global__ __launch_bounds__(64 * W, WPS) void corr_mfma32_kernel(
    const uint16_t *__restrict__ images, int img_stride, const float *__restrict__ sst,
    const int *__restrict__ moveouts, const float *__restrict__ weights,
    const float *__restrict__ data, const uint32_t *__restrict__ planes, int plane_pairs,
    const float *__restrict__ energy, const int *__restrict__ out_range,
    float *__restrict__ out, int n_templates, int n_channels, int n_tp_all,
    int n_data, int n_corr, const int *__restrict__ t_len,
    const long long *__restrict__ t_eoff, const int *__restrict__ t_order, int mode,
    int summed, int n_tiles, int per_xcd, const int *__restrict__ items, int n_listed) {
  using C = Cfg32<STEP, W, G, IL>;
  static_assert(!IL || (NP == 3 && G == 2), "interleaved tiles: bf16x3, two tiles per wave");
  constexpr int kThreads = C::kThreads;
  extern __shared__ uint4 lds_raw[];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int c = lane & 31, h = lane >> 5;

  // Work items (output tile, template slot), tile-major, in 8 contiguous
  // ranges, one per XCD (block b runs on XCD b mod 8); the blocks of an XCD
  // stride through its range (persistent launch when the grid is smaller).
  // A batched call's slots follow its schedule (t_order).  Listed items
  // outside [0, n_tiles * n_templates) are skipped.
  const int n_all = n_tiles * n_templates;
  const int n_items = items ? n_listed : n_all;
  const int xcd = blockIdx.x & 7;
  const int blocks_per_xcd = gridDim.x >> 3;
  for (int k = blockIdx.x >> 3; k < per_xcd; k += blocks_per_xcd) {
  const int wi = xcd * per_xcd + k;
  if (wi >= n_items) break;
  const int w = items ? items[wi] : wi;
  if ((unsigned)w >= (unsigned)n_all) continue;
  const int tile = w / n_templates;
  const int slot = w - tile * n_templates;
  const int t = t_order ? t_order[slot] : slot;
  int n_tp = t_len ? t_len[t] : n_tp_all;
  n_tp = n_tp < 1 ? 1 : (n_tp > n_tp_all ? n_tp_all : n_tp);
  const int n16 = n16_of(n_tp, STEP);
  const int n_win = n_data - n_tp + 1;
  const long long eoff = t_eoff ? t_eoff[t] : 0;
  const float *energy_t = energy + (eoff > 0 ? eoff : 0);
  const int lt = tpl_len32<STEP>(n16);
  const int span = span32<STEP, W, G, IL>(n16);
  const int sph = span32_phys<C::kChunk>(span);
  uint16_t *tpl = reinterpret_cast<uint16_t *>(lds_raw);
  uint16_t *dat_hi = tpl + 2 * C::kCopies * lt;
  uint16_t *dat_lo = dat_hi + sph;
  float *ebuf = reinterpret_cast<float *>(dat_hi + sph);  // NP = 1 (16-byte aligned)
  const int o_blk = tile * C::kBlockOut;
  const int olo = out_range[2 * t], ohi = out_range[2 * t + 1];
  // First output of this lane in tile q: + q * kTileOff, then + 8 * (j / 4)
  // + (j % 4) for accumulator element j.
  const int o_lane = o_blk + wave * G * kTile + C::kColOff * c + 4 * h;
  // per channel, the largest 1/sqrt(e) of its windows, behind the table
  const unsigned *r_top = reinterpret_cast<const unsigned *>(energy_t + (size_t)n_channels * n_win);
  const int tsc = t * n_channels;
  const int img_bytes = 4 * C::kCopies * lt;  // the image of one template channel

  float sum[G][16];
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int j = 0; j < 16; j++) sum[q][j] = 0.0f;

  // eoff < 0: a batched template whose length has no energy table; its row
  // is written as zeros.
  const bool active = (o_blk < ohi) && (o_blk + C::kBlockOut > olo) && eoff >= 0;
  const float *w_t = weights + tsc;
  ActiveChannels act(w_t, n_channels);
  auto next_active = [&](int ch) { return act.next(ch); };

  // Prefetch registers of the next channel: 8-element groups of the data
  // window (hi and lo) and 8-element chunks of the template image.
  u32x4 vh[C::kGroups], vl[NP == 3 ? C::kGroups : 1];
  u32x4 vt[C::kTplChunks];
  constexpr int kEPer = C::kBlockOut / kThreads;  // energies per thread (NP = 1)
  float pe[NP == 1 ? kEPer : 1];
  auto issue = [&](int ch, int mo_ch) {
    const int tid = opaque(threadIdx.x);
    const int base = o_blk * STEP + mo_ch;
    const auto ri = rsrc(images + (size_t)(tsc + ch) * img_stride, (unsigned)img_bytes);
#pragma unroll
    for (int k2 = 0; k2 < C::kTplChunks; k2++)
      vt[k2] = bload4(ri, 16 * (tid + k2 * kThreads));
    if (base >= 0) {
      // group g of the window: elements [8g, 8g + 8) = pairs p0 + 4g .. + 3 of
      // the plane whose pairs start at the window's parity
      const int a = base & 1;
      const unsigned pb = 4u * (unsigned)plane_pairs;
      const auto rh = rsrc(planes + ((size_t)ch * 4 + a) * plane_pairs, pb);
      const int od = 4 * (base >> 1) + 16 * tid;
#pragma unroll
      for (int g = 0; g < C::kGroups; g++) vh[g] = bload4(rh, od + 16 * g * kThreads);
      if (NP == 3) {
        const auto rl = rsrc(planes + ((size_t)ch * 4 + 2 + a) * plane_pairs, pb);
#pragma unroll
        for (int g = 0; g < C::kGroups; g++) vl[g] = bload4(rl, od + 16 * g * kThreads);
      }
    } else {
      // The window starts before the data: split the FP32 samples here;
      // samples before the data get the out-of-range offset kOob (zero).
      const auto rd = rsrc(data + (size_t)ch * n_data, 4u * (unsigned)n_data);
#pragma unroll
      for (int g = 0; g < C::kGroups; g++) {
        uint32_t hv[4], lv[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int gi = base + 8 * (tid + g * kThreads) + 2 * j;
          const float x0 = bload(rd, gi >= 0 ? 4 * gi : kOob);
          const float x1 = bload(rd, gi + 1 >= 0 ? 4 * (gi + 1) : kOob);
          split2(x0, x1, hv[j], lv[j]);
        }
        vh[g] = u32x4{hv[0], hv[1], hv[2], hv[3]};
        if (NP == 3) vl[g] = u32x4{lv[0], lv[1], lv[2], lv[3]};
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
  // B: element e = (wave*G + tile)*1024*step + chunk*c + 16m + 8h of the
  // window (IL: wave*G*1024 + 64c + 16m + 8h, tiles via the ring) at physical
  // e + 8 * (e >> log2(chunk)); the m-dependent part of the pad is
  // 8 * (m >> log2(chunk / 16)).
  const int b_lane = wave * C::kWaveStride + c * C::kLaneStride + 8 * h;
  const uint16_t *b_hi = dat_hi + b_lane;
  const uint16_t *b_lo = dat_lo + b_lane;

  int ch = active ? next_active(-1) : n_channels;
  if (ch < n_channels) issue(ch, moveouts[tsc + ch]);
  int nxt = ch < n_channels ? next_active(ch) : n_channels;
  while (ch < n_channels) {
    // Scalars of this channel and the next channel's moveout, loaded before
    // the staging barriers so that their latency passes with the LDS writes
    // (pinned in SGPRs after the second barrier; the epilogue and the next
    // channel's loads would otherwise wait for them).
    float wch = w_t[ch];
    int mo = moveouts[tsc + ch];
    float s = sst[tsc + ch];
    unsigned rt = r_top[ch];
    int mo_nxt = moveouts[tsc + (nxt < n_channels ? nxt : n_channels - 1)];
    __syncthreads();  // previous channel's fragments are consumed
    {
    const int tid = opaque(threadIdx.x);
    // The template image is the LDS layout of the planes: a flat copy.
#pragma unroll
    for (int k2 = 0; k2 < C::kTplChunks; k2++) {
      const int y = tid + k2 * kThreads;
      if (16 * y < img_bytes) reinterpret_cast<u32x4 *>(tpl)[y] = vt[k2];
    }
    // Group g starts 8 * threads * g elements, whole padded chunks, after
    // group 0: its physical element is group 0's plus a constant.  The first
    // kSafeGroups groups lie inside the shortest window (n16 = 2).
    const int e0 = 8 * tid;
    const int ph0 = e0 + 8 * (e0 >> C::kChunkLog);
    constexpr int kGroupPhys = 8 * kThreads + 8 * (8 * kThreads / C::kChunk);
#pragma unroll
    for (int g = 0; g < C::kGroups; g++) {
      if (g >= C::kSafeGroups && e0 + 8 * g * kThreads >= span) continue;
      const int ph = ph0 + g * kGroupPhys;
      *reinterpret_cast<u32x4 *>(dat_hi + ph) = vh[g];
      if (NP == 3) *reinterpret_cast<u32x4 *>(dat_lo + ph) = vl[g];
    }
    if (NP == 1) {
      // epad(tid + v * threads) = epad(tid) + v * (threads + threads / 8)
      const int eb = epad(tid);
#pragma unroll
      for (int v = 0; v < kEPer; v++) ebuf[eb + v * (kThreads + kThreads / 8)] = pe[v];
    }
    }
    __syncthreads();
    __asm__ volatile("" : "+s"(wch), "+s"(mo), "+s"(s), "+s"(rt), "+s"(mo_nxt));
    // With at least 8 steps, the next channel's loads and this channel's
    // energy loads issue between the MFMA groups of the first 8 steps
    // instead of all at once after the barrier (the waves of a block would
    // otherwise queue some 80 loads on the CU's address unit together, and
    // each wave's MFMAs would wait behind them).  bf16x3 shapes of up to two
    // tiles per wave whose window takes at most 5 load groups per thread.
    constexpr bool kSpreadable = NP == 3 && G <= 2 && C::kGroups <= 5;
    const bool spread = kSpreadable && n16 >= 8;
    if (nxt < n_channels && !spread) issue(nxt, mo_nxt);

    // 1/sqrt of the window energies of this lane's outputs in tile q.
    // Outputs outside [olo, ohi) read whatever lies there, or 0, and are
    // masked when stored; windows before the data (o * step + mo < 0, only
    // below olo) read kOob.
    const auto re = rsrc(energy_t + (size_t)ch * n_win, 4u * (unsigned)n_win);
    float e[G][16];
    const int ol = opaque(o_lane);
    auto load_e = [&](int q) {
      if (STEP == 1 && o_blk + mo >= 0) {
        // a lane's outputs are runs of 4 consecutive ones: one 16-byte load each
        const int oe = 4 * (ol + mo);
#pragma unroll
        for (int g = 0; g < 4; g++) {
          const u32x4 v = bload4(re, oe + 4 * (q * C::kTileOff + 8 * g));
#pragma unroll
          for (int i = 0; i < 4; i++) e[q][4 * g + i] = __uint_as_float(v[i]);
        }
      } else if (STEP == 2 && G == 1 && o_blk * 2 + mo >= 0) {
        // a run of 4 consecutive outputs reads every second of 8 energies (the
        // wider loads' registers would spill with two tiles per wave)
        const int oe = 4 * (2 * ol + mo);
#pragma unroll
        for (int g = 0; g < 4; g++) {
          const u32x4 v0 = bload4(re, oe + 8 * (q * kTile + 8 * g));
          const u32x4 v1 = bload4(re, oe + 8 * (q * kTile + 8 * g) + 16);
          e[q][4 * g] = __uint_as_float(v0[0]);
          e[q][4 * g + 1] = __uint_as_float(v0[2]);
          e[q][4 * g + 2] = __uint_as_float(v1[0]);
          e[q][4 * g + 3] = __uint_as_float(v1[2]);
        }
      } else if (o_blk * STEP + mo >= 0) {
        const int oe = 4 * (ol * STEP + mo);
#pragma unroll
        for (int j = 0; j < 16; j++)
          e[q][j] = bload(re, oe + 4 * STEP * (q * C::kTileOff + 8 * (j >> 2) + (j & 3)));
      } else {
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const int x = (ol + q * C::kTileOff + 8 * (j >> 2) + (j & 3)) * STEP + mo;
          e[q][j] = bload(re, x >= 0 ? 4 * x : kOob);
        }
      }
    };
    if (NP == 3 && !spread) {
#pragma unroll
      for (int q = 0; q < G; q++) load_e(q);
    }
    // NP = 1: the epilogue reads the energies from LDS

    // Parts of the spread loads. 0: the template image (and a window before
    // the data, whole); 1-5: data load groups; 6, 7: energies of tiles 0, 1.
    const bool pf = nxt < n_channels;
    const int base_n = o_blk * STEP + mo_nxt;
    auto part = [&](auto pc) {
      constexpr int P = decltype(pc)::value;
      if constexpr (kSpreadable) {
        const int tq = opaque(threadIdx.x);
        if constexpr (P == 0) {
          if (pf) {
            const auto ri = rsrc(images + (size_t)(tsc + nxt) * img_stride, (unsigned)img_bytes);
#pragma unroll
            for (int k2 = 0; k2 < C::kTplChunks; k2++)
              vt[k2] = bload4(ri, 16 * (tq + k2 * kThreads));
            if (base_n < 0) issue(nxt, mo_nxt);
          }
        } else if constexpr (P <= 5) {
          if constexpr (P - 1 < C::kGroups) {
            if (pf && base_n >= 0) {
              const unsigned pb = 4u * (unsigned)plane_pairs;
              const int a_n = base_n & 1;
              const int od = 4 * (base_n >> 1) + 16 * tq;
              const auto rh = rsrc(planes + ((size_t)nxt * 4 + a_n) * plane_pairs, pb);
              const auto rl = rsrc(planes + ((size_t)nxt * 4 + 2 + a_n) * plane_pairs, pb);
              vh[P - 1] = bload4(rh, od + 16 * (P - 1) * kThreads);
              vl[P - 1] = bload4(rl, od + 16 * (P - 1) * kThreads);
            }
          }
        } else if constexpr (P - 6 < G) {
          load_e(P - 6);
        }
      }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    using I6 = std::integral_constant<int, 6>;
    using I7 = std::integral_constant<int, 7>;

    f32x16 acc[G];
    if (!spread || !IL) {
#pragma unroll
      for (int q = 0; q < G; q++)
#pragma unroll
        for (int j = 0; j < 16; j++) acc[q][j] = 0.0f;
    }

    // Two fragment sets in turn: those of step m + 1 are read while the
    // MFMAs of step m run.  With n16 a multiple of 4 the last read of the
    // loop is one step past the template (inside the planes, which hold
    // n16 + 1 steps; never used); otherwise a final pair of steps follows.
    // The scheduling barriers keep this order: left to itself the compiler
    // moves every read next to its MFMA and waits for it there.
    if constexpr (IL) {
      const uint32_t *ah = a_hi, *al = a_lo;
      const uint16_t *bh = b_hi, *bl = b_lo;
      RingIL f;
      f.read_a<0>(ah, al);
      f.read_f<0>(bh, bl);
      f.read_f<1>(bh, bl);
      f.read_f<2>(bh, bl);
      int m = 0;
      if (spread) {
        auto group = [&](auto first, auto p0, auto p1, auto p2, auto p3) {
          f.read_a<1>(ah, al);
          f.read_f<3>(bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          if constexpr (decltype(first)::value)
            f.mfma_first<0>(acc);
          else
            f.mfma<0>(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p0);
          __builtin_amdgcn_sched_barrier(0);
          f.read_a<2>(ah, al);
          f.read_f<4>(bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          f.mfma<1>(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p1);
          __builtin_amdgcn_sched_barrier(0);
          f.read_a<3>(ah, al);
          f.read_f<5>(bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          f.mfma<2>(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p2);
          __builtin_amdgcn_sched_barrier(0);
          f.read_a<4>(ah, al);
          f.read_f<6>(bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          f.mfma<3>(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p3);
          __builtin_amdgcn_sched_barrier(0);
          ah += 32;
          al += 32;
          bh += 72;
          bl += 72;
        };
        group(std::true_type{}, I0{}, I1{}, I2{}, I3{});
        group(std::false_type{}, I4{}, I5{}, I6{}, I7{});
        m = 8;
      }
      // When the taps need an odd number of 16-tap steps, the last of the
      // n16 steps (rounded up to even for the loop) holds template taps past
      // n_tp only, which are zero, and its MFMAs are skipped: a last group of
      // four that holds it is left out of the loop and run below without its
      // fourth step's MFMAs, and in the two-step tail a uniform branch skips
      // the second (the fragment reads stay as in the loop; no test inside
      // the loop, which cost the lengths without such a step 0.5 %).  The
      // spread groups (steps 0-7) run every step.
      const bool zero_last = ((n_tp + 31 * STEP + 15) / 16) & 1;
      const int n_loop = zero_last && (n16 & 3) == 0 ? n16 - 4 : n16;
      for (; m + 4 <= n_loop; m += 4) {
        f.read_a<1>(ah, al);
        f.read_f<3>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<0>(acc);
        __builtin_amdgcn_sched_barrier(0);
        f.read_a<2>(ah, al);
        f.read_f<4>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<1>(acc);
        __builtin_amdgcn_sched_barrier(0);
        f.read_a<3>(ah, al);
        f.read_f<5>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<2>(acc);
        __builtin_amdgcn_sched_barrier(0);
        f.read_a<4>(ah, al);
        f.read_f<6>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<3>(acc);
        __builtin_amdgcn_sched_barrier(0);
        ah += 32;
        al += 32;
        bh += 72;
        bl += 72;
      }
      if (m + 4 <= n16) {  // the last four steps, the fourth all zero
        f.read_a<1>(ah, al);
        f.read_f<3>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<0>(acc);
        __builtin_amdgcn_sched_barrier(0);
        f.read_a<2>(ah, al);
        f.read_f<4>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<1>(acc);
        __builtin_amdgcn_sched_barrier(0);
        f.read_a<3>(ah, al);
        f.read_f<5>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<2>(acc);
        __builtin_amdgcn_sched_barrier(0);
        m += 4;
      }
      if (m < n16) {  // steps n16 - 2 and n16 - 1
        f.read_a<1>(ah, al);
        f.read_f<3>(bh, bl);
        __builtin_amdgcn_sched_barrier(0);
        f.mfma<0>(acc);
        __builtin_amdgcn_sched_barrier(0);
        if (!zero_last) f.mfma<1>(acc);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      const uint32_t *ah = a_hi, *al = a_lo;
      const uint16_t *bh = b_hi, *bl = b_lo;
      constexpr int kB4 = 64 + 8 * (4 >> C::kStepLog);  // 4 steps of B
      Frag<STEP, W, G, NP, IL> f0, f1;
      f0.template read<0>(ah, al, bh, bl);
      int m = 0;
      if (spread) {
        auto group = [&](auto p0, auto p1, auto p2, auto p3) {
          f1.template read<1>(ah, al, bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          f0.mfma(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p0);
          __builtin_amdgcn_sched_barrier(0);
          f0.template read<2>(ah, al, bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          f1.mfma(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p1);
          __builtin_amdgcn_sched_barrier(0);
          f1.template read<3>(ah, al, bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          f0.mfma(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p2);
          __builtin_amdgcn_sched_barrier(0);
          f0.template read<4>(ah, al, bh, bl);
          __builtin_amdgcn_sched_barrier(0);
          f1.mfma(acc);
          __builtin_amdgcn_sched_barrier(0);
          part(p3);
          __builtin_amdgcn_sched_barrier(0);
          ah += 32;
          al += 32;
          bh += kB4;
          bl += kB4;
        };
        group(I0{}, I1{}, I2{}, I3{});
        group(I4{}, I5{}, I6{}, I7{});
        m = 8;
      }
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

    // Epilogue: normalise, weight, accumulate (or store per channel).  e holds
    // r = 1/sqrt(energy) (0 for an all-zero window, whose numerator is 0 too);
    // the template's factor is 1/sqrt(sst).  bf16x3: the reference's test
    // sst * e > thr (thr = 1e-6, or 1e-12 for the 'precise' rule
    // sqrt(sst * e) > 1e-6) is r < sqrt(sst / thr).
    const float rs = s > 0.0f ? __builtin_amdgcn_rsqf(s) : 0.0f;
    const float thr = mode == kThrBeforeSqrt ? kStabilityThreshold
                                             : kStabilityThreshold * kStabilityThreshold;
    const float rmax = __builtin_sqrtf(s / thr);
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
    // Per-channel output (t, o, ch) through a resource based at the block's
    // first output: offsets stay below kBlockOut * n_channels values.
    const long long rem = ((long long)n_corr - o_blk) * n_channels * 4;
    const auto rc = rsrc(out + ((size_t)t * n_corr + o_blk) * n_channels,
                         summed ? 0u : (unsigned)(rem < 0x7fffffffLL ? rem : 0x7fffffffLL));
    // The threshold test can fail only where r reaches rmax: when the
    // channel's largest r (behind its energy table; NaN compares false) stays
    // below, every window passes and the test is skipped (same values).  The
    // one-wave, four-tile blocks at step 2 keep the tested form only (a
    // second epilogue would make them spill registers).
    constexpr bool kTwoForms = !(STEP == 2 && G == 4);
    const bool tested = NP == 3 && (!kTwoForms || !(__uint_as_float(rt) < rmax));
    auto epilogue = [&](auto test) {
#pragma unroll
      for (int q = 0; q < G; q++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const float x = acc[q][j] * e[q][j];
          const bool pass = !decltype(test)::value || e[q][j] < rmax;
          if (summed) {
            sum[q][j] = __builtin_fmaf(pass ? x : 0.0f, wrs, sum[q][j]);
            continue;
          }
          const float cc = pass ? x * rs : 0.0f;
          const int o = opaque(o_lane) + q * C::kTileOff + 8 * (j >> 2) + (j & 3);
          bstore(rc, (o >= olo && o < ohi) ? 4 * ((o - o_blk) * n_channels + ch) : kOob, cc);
        }
    };
    if (tested)
      epilogue(std::true_type{});
    else
      epilogue(std::false_type{});
    ch = nxt;
    nxt = ch < n_channels ? next_active(ch) : n_channels;
  }
  if (summed) {
    // Stores past n_corr fall outside the row's buffer and are dropped.
    const auto ro = rsrc(out + (size_t)t * n_corr, 4u * (unsigned)n_corr);
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int o = o_lane + q * C::kTileOff + 8 * (j >> 2) + (j & 3);
        bstore(ro, 4 * o, (o >= olo && o < ohi) ? sum[q][j] : 0.0f);
      }
  }
  }  // work items
}

// ---- operand preparation (once per call) --------------------------------

// Four planes per channel of `pairs` dwords each: plane 2b + a holds, in pair
// p, the bf16 hi (b = 0) or lo (b = 1) parts of samples 2p + a and 2p + a + 1
// (first sample in the low half); samples at or past L are 0.
__global__ __launch_bounds__(256) void split_planes_kernel(
    const float *__restrict__ data, long long n_data, long long pairs,
    uint32_t *__restrict__ planes) {
  const long long ch = blockIdx.y;
  const float *d = data + ch * n_data;
  uint32_t *pl = planes + ch * 4 * pairs;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long pp = (long long)blockIdx.x * blockDim.x + threadIdx.x; pp < pairs;
       pp += stride) {
    const long long i = 2 * pp;
    const float x0 = i < n_data ? d[i] : 0.0f;
    const float x1 = i + 1 < n_data ? d[i + 1] : 0.0f;
    const float x2 = i + 2 < n_data ? d[i + 2] : 0.0f;
    uint32_t hi, lo;
    split2(x0, x1, hi, lo);
    pl[pp] = hi;
    pl[2 * pairs + pp] = lo;
    split2(x1, x2, hi, lo);
    pl[pairs + pp] = hi;
    pl[3 * pairs + pp] = lo;
  }
}

// LDS image of every template channel (rows of `stride` bf16): for each copy
// (two for step 1, the second shifted by one tap) the hi plane then the lo
// plane, each lt = tpl_len32(n16) elements, element y of copy p holding tap
// y - 32*step - p of the prepared template (0 outside [0, n_tp)).
template <int STEP>
__global__ __launch_bounds__(256) void template_images_kernel(
    const float *__restrict__ tp, long long n_pad, int n_tp_all,
    const int *__restrict__ t_len, long long rows_per_template, long long stride,
    uint16_t *__restrict__ img) {
  const long long row = blockIdx.x;
  int n_tp = t_len ? t_len[row / rows_per_template] : n_tp_all;
  n_tp = n_tp < 1 ? 1 : (n_tp > n_tp_all ? n_tp_all : n_tp);
  const int lt = tpl_len32<STEP>(n16_of(n_tp, STEP));
  const float *src = tp + row * n_pad;
  uint16_t *dst = img + row * stride;
  const int n = 2 * Pad32<STEP>::kCopies * lt;
  for (int y = threadIdx.x; y < n; y += blockDim.x) {
    const int part = y / lt, yy = y - part * lt;   // part = 2 * copy + plane
    const int x = yy - Pad32<STEP>::kPad - (part >> 1);
    const float v = (x >= 0 && x < n_tp) ? src[x] : 0.0f;
    uint32_t hi, lo;
    split2(v, 0.0f, hi, lo);
    dst[y] = (uint16_t)((part & 1) ? lo : hi);
  }
}

constexpr size_t kDefaultLds = 64 * 1024;
constexpr size_t kMaxLds = 160 * 1024;

template <int STEP, int W, int G, int WPS, int NP, int IL = 0>
int launch32(const CorrArgs &a, hipStream_t s) {
  using C = Cfg32<STEP, W, G, IL>;
  const int n16 = n16_of((int)a.n_tp, STEP);
  if (!mfma32_covers(a.step, a.n_tp, a.n_data, a.n_corr, a.n_channels)) return kUnsupported;
  if (!a.planes || !a.images)
    return set_error(kInvalidArgument, "32x32 MFMA launch without split operands");
  const size_t lds = mfma32_lds_bytes<STEP, W, G, NP, IL>(n16);
  if (lds > kMaxLds) return kUnsupported;
  if (lds > kDefaultLds) {
    int rc = ensure_lds_limit(
        reinterpret_cast<const void *>(&corr_mfma32_kernel<STEP, W, G, WPS, NP, IL>),
        (int)kMaxLds);
    if (rc) return rc;
  }
  const long long n_tiles = (a.n_corr + C::kBlockOut - 1) / C::kBlockOut;
  if (n_tiles * a.n_templates > 0x7fffffffLL)
    return set_error(kInvalidArgument, "too many work items (%lld)", n_tiles * a.n_templates);
  const long long work = a.items ? a.n_items : n_tiles * a.n_templates;
  if (work == 0) return kOk;
  if (work > 0x7fffffffLL)
    return set_error(kInvalidArgument, "too many work items (%lld)", work);
  const long long per_xcd = (work + 7) / 8;
  long long blocks = per_xcd * 8;
  if (a.persistent_blocks > 0 && a.persistent_blocks < blocks)
    blocks = (a.persistent_blocks + 7) / 8 * 8;
  hipLaunchKernelGGL((corr_mfma32_kernel<STEP, W, G, WPS, NP, IL>), dim3((unsigned)blocks),
                     dim3(C::kThreads), lds, s, a.images, (int)a.img_stride, a.sst,
                     a.moveouts, a.weights, a.data, a.planes, (int)a.plane_pairs,
                     a.energy, a.out_range, a.out, (int)a.n_templates,
                     (int)a.n_channels, (int)a.n_tp, (int)a.n_data, (int)a.n_corr,
                     a.t_len, a.t_eoff, a.t_order, a.threshold_mode, a.summed, (int)n_tiles,
                     (int)per_xcd, a.items, (int)a.n_items);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

// Tile shapes of this kernel: code = 100 * waves_per_block + 10 *
// tiles_per_wave + waves_per_simd; 1422: 422 with interleaved tiles (at step
// 2 it runs as 422).
template <int STEP>
int launch32_code(const CorrArgs &a, int code, hipStream_t s) {
  // The screening variant has its own shape: one tile per wave and 4 waves
  // per SIMD.  Its channel iterations are short (10 16-tap steps for a
  // 100-tap coarse template), so occupancy, not the tile size, hides their
  // barriers and loads.  Code 212 asks for items of 2048 lags (two waves per
  // block, 3 waves per SIMD).
  if (a.screen) {
    if (code == 212) return launch32<STEP, 2, 1, 3, 1>(a, s);
    return launch32<STEP, 4, 1, 4, 1>(a, s);
  }
  switch (code) {
    case 1422:
      if constexpr (STEP == 1) return launch32<STEP, 4, 2, 2, 3, 1>(a, s);
      return launch32<STEP, 4, 2, 2, 3>(a, s);
    case 121: return launch32<STEP, 1, 2, 1, 3>(a, s);
    case 141: return launch32<STEP, 1, 4, 1, 3>(a, s);
    case 412: return launch32<STEP, 4, 1, 2, 3>(a, s);
    case 212: return launch32<STEP, 2, 1, 2, 3>(a, s);
    case 413: return launch32<STEP, 4, 1, 3, 3>(a, s);
    case 421: return launch32<STEP, 4, 2, 1, 3>(a, s);
    case 241: return launch32<STEP, 2, 4, 1, 3>(a, s);
    default: return launch32<STEP, 4, 2, 2, 3>(a, s);  // 422
  }
}

}  // namespace

// Raise a kernel's dynamic LDS limit to `bytes` once per (kernel, device);
// thread-safe.
int ensure_lds_limit(const void *fn, int bytes) {
  static std::mutex mu;
  static std::set<std::pair<const void *, int>> done;
  int dev = 0;
  FMF_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if (done.count({fn, dev})) return kOk;
  FMF_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    bytes));
  done.insert({fn, dev});
  return kOk;
}

bool mfma32_shape(int code) {
  switch (code) {
    case 121: case 141:
    case 212: case 412: case 413: case 421: case 422: case 241: case 1422: return true;
    default: return false;
  }
}

int mfma32_block_outputs(int code) {
  code %= 1000;
  return (code / 100) * ((code / 10) % 10) * kTile;
}

// Steps and sizes the 32x32 kernels take: step 1 or 2, N <= 16 * 84 - 31 *
// step - 15, 32-bit byte offsets into data rows, energy tables and cc rows,
// and per-block per-channel output offsets (8192 outputs x channels) below
// 2^31 bytes.
bool mfma32_covers(long long step, long long n_tp, long long n_data, long long n_corr,
                   long long n_channels) {
  if (step != 1 && step != 2) return false;
  if (n_tp < 1 || n_tp > 0x7fffffffLL || n16_of((int)n_tp, (int)step) > kMaxN16) return false;
  if (n_data >= (1LL << 29) || n_corr >= (1LL << 29)) return false;
  if (n_channels < 1 || n_channels >= (1LL << 16)) return false;
  return true;
}

long long mfma32_plane_pairs(long long n_data) {
  // pairs holding samples [0, L] and, behind them, zeros for windows that
  // run past the data (at most 8192 outputs * step 2 + the longest
  // template's steps beyond a block's last valid output)
  const long long need = (n_data + 2) / 2 + 8192 + 64 * 16;
  return (need + 63) / 64 * 64;
}

long long mfma32_image_stride(long long n_tp_max, int step) {
  const int n16 = n16_of((int)n_tp_max, step);
  const int lt = step == 1 ? tpl_len32<1>(n16) : tpl_len32<2>(n16);
  return 2LL * (step == 1 ? Pad32<1>::kCopies : Pad32<2>::kCopies) * lt;
}

int launch_split_planes(const float *data, long long n_channels, long long n_data,
                        long long pairs, uint32_t *planes, hipStream_t s) {
  if (n_channels <= 0) return kOk;
  const long long bx = std::min<long long>((pairs + 255) / 256, 4096);
  hipLaunchKernelGGL(split_planes_kernel, dim3((unsigned)bx, (unsigned)n_channels), dim3(256),
                     0, s, data, n_data, pairs, planes);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_template_images(const float *prepared, long long n_rows, long long n_pad,
                           long long n_tp, int step, const int *t_len,
                           long long rows_per_template, long long stride, uint16_t *img,
                           hipStream_t s) {
  if (n_rows <= 0) return kOk;
  // every row's image must fit its stride (the kernel writes whole images)
  if (step == 1 || step == 2) {
    const int n16 = n16_of((int)n_tp, step);
    const long long need = step == 1 ? 2LL * Pad32<1>::kCopies * tpl_len32<1>(n16)
                                     : 2LL * Pad32<2>::kCopies * tpl_len32<2>(n16);
    if (stride < need)
      return set_error(kInvalidArgument, "template image stride %lld < %lld", stride, need);
  }
  if (step == 1)
    hipLaunchKernelGGL(template_images_kernel<1>, dim3((unsigned)n_rows), dim3(256), 0, s,
                       prepared, n_pad, (int)n_tp, t_len,
                       rows_per_template > 0 ? rows_per_template : 1, stride, img);
  else if (step == 2)
    hipLaunchKernelGGL(template_images_kernel<2>, dim3((unsigned)n_rows), dim3(256), 0, s,
                       prepared, n_pad, (int)n_tp, t_len,
                       rows_per_template > 0 ? rows_per_template : 1, stride, img);
  else
    return set_error(kUnsupported, "template images are built for step 1 or 2");
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_corr_mfma32(const CorrArgs &a, int code, hipStream_t s) {
  if (a.step == 1) return launch32_code<1>(a, code, s);
  if (a.step == 2) return launch32_code<2>(a, code, s);
  return kUnsupported;
}

}  // namespace fmf
