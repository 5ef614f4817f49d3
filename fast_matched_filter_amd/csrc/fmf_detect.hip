// Detection post-processing of cc rows on the GPU: per-template median and
// MAD, threshold, and peak picking with a minimum separation.  The reference
// stops at cc_sum (fast_matched_filter.py:318-362); these kernels let a
// caller keep the (T, n_corr) rows on the device and copy back only the
// detections.
//
// Definitions (row t, valid lags [lo, hi) of that template):
//   median  = numpy.median(x) in float32 (the mean of the two middle values
//             for an even count: (a + b) * 0.5 in float32),
//   mad     = numpy.median(|x - median|) (differences in float32),
//   thr     = median + n_mad * mad (float32, no fused multiply-add),
//   lag j is a detection when x[j] > thr and x[j] is the first maximum of
//   x over [max(lo, j - s), min(hi - 1, j + s)]: x[j] > x[i] for the lags i
//   before j and x[j] >= x[i] for those after it.
//
// Medians are exact: a radix select on order-preserving 32-bit keys in three
// passes of 11, 11 and 10 bits (histograms in LDS, merged with global
// atomics); the last pass also yields the upper middle value of an even
// count.  Rows of at least 2^18 valid lags first take a sampled bracket
// (Floyd-Rivest style): 16384 keys from 4096 chunks of 4 lags spread over
// the row, sorted in LDS, give keys a and b 256 sample ranks either side of
// the wanted rank; one pass counts the keys below a, copies those in [a, b]
// (about 3 % of the row) into a candidate buffer and finds the smallest
// key above b; the radix select then runs on the candidates.  A row whose
// rank falls outside its bracket, or whose bracket overflows the buffer
// (n/8 keys), runs the radix select over the whole row instead: the result
// is exact either way, two full reads per median and MAD instead of six.
// The window test reads maxima of 64-lag segments for whole segments inside
// the window and single lags at its ends, so its cost does not grow with s
// beyond s / 64 reads, and only segments whose maximum is above the
// threshold are read at all.  Detections are compacted in (template, lag)
// order: one pass counts them per 4096-lag chunk, a scan turns the counts
// into offsets and a second pass, over the chunks that hold detections,
// writes them.
#include "fmf_internal.h"

namespace fmf {

namespace {

constexpr int kThreads = 256;
constexpr int kBins = 2048;
constexpr int kSeg = 64;           // lags per segment maximum
constexpr int kLagsPerChunk = 4096;  // lags per detection-count chunk
constexpr int kUnroll = 8;         // row loads in flight per thread

struct SelectState {
  unsigned prefix;   // key bits fixed so far
  unsigned mask;     // which bits are fixed
  long long k;       // rank still sought among keys matching prefix
  long long n;       // valid lags of the row
  unsigned min_gt;   // smallest key above the last pass's prefix group
  unsigned key_hi;   // key of rank k + 1 (the upper middle value)
};

__device__ __forceinline__ unsigned fkey(float x) {
  const unsigned b = __float_as_uint(x);
  return b ^ ((b >> 31) ? 0xffffffffu : 0x80000000u);
}

__device__ __forceinline__ float kval(unsigned k) {
  return __uint_as_float((k >> 31) ? (k ^ 0x80000000u) : ~k);
}

__device__ __forceinline__ void row_range(const int *ranges, int t, long long n_corr,
                                          long long &lo, long long &hi) {
  if (ranges) {
    lo = ranges[2 * t];
    hi = ranges[2 * t + 1];
    lo = lo < 0 ? 0 : (lo > n_corr ? n_corr : lo);
    hi = hi < lo ? lo : (hi > n_corr ? n_corr : hi);
  } else {
    lo = 0;
    hi = n_corr;
  }
}

// Element value of a select: the cc value, or |cc - median| for the MAD.
__device__ __forceinline__ float sel_value(float x, int abs_dev, float med) {
  return abs_dev ? fabsf(__fsub_rn(x, med)) : x;
}

__global__ __launch_bounds__(kThreads) void init_select_kernel(
    const int *ranges, long long n_corr, int n_templates, SelectState *st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_templates) return;
  long long lo, hi;
  row_range(ranges, t, n_corr, lo, hi);
  SelectState s;
  s.prefix = 0;
  s.mask = 0;
  s.n = hi - lo;
  s.k = s.n > 0 ? (s.n - 1) / 2 : 0;  // lower middle rank
  s.min_gt = 0xffffffffu;
  s.key_hi = 0;
  st[t] = s;
}

// Sampled bracket of one row (see the file comment).  The three fields that
// every block of the counting pass updates atomically sit on cache lines of
// their own (128 bytes apart), so that their atomics do not queue behind one
// another.
struct Bracket {
  unsigned a, b;             // candidate keys: a <= key <= b
  int fallback;              // 1: the radix select over the whole row
  unsigned pad0[29];
  unsigned min_above;        // smallest key above b
  unsigned pad1[31];
  unsigned count;            // keys in [a, b] (past the buffer's capacity: overflow)
  unsigned pad2[31];
  unsigned long long below;  // keys below a
  unsigned pad3[30];
};
static_assert(sizeof(Bracket) == 512, "four 128-byte lines per row");

// Histogram of the digit (key >> shift) & (2^bits - 1) over the keys that
// match the row's prefix.  In the last pass (shift 0) it also finds the
// smallest key above the prefix group, the candidate for rank k + 1 when the
// group holds no larger key.  br: only the rows whose bracket fell back
// (nullptr: every row).  grid = (blocks per row, T).
__global__ __launch_bounds__(kThreads) void hist_kernel(
    const float *__restrict__ cc, long long n_corr, const int *ranges,
    int abs_dev, const float *__restrict__ med, SelectState *__restrict__ st,
    int shift, int bits, unsigned *__restrict__ hist, const Bracket *__restrict__ br) {
  __shared__ unsigned h[kBins];
  const int t = blockIdx.y;
  if (br && !br[t].fallback) return;  // the whole block
  for (int i = threadIdx.x; i < kBins; i += kThreads) h[i] = 0;
  __syncthreads();
  long long lo, hi;
  row_range(ranges, t, n_corr, lo, hi);
  const unsigned prefix = st[t].prefix, mask = st[t].mask;
  const unsigned group_max = prefix | ~mask;  // largest key matching the prefix
  const unsigned dmask = (1u << bits) - 1u;
  const float m = abs_dev ? med[t] : 0.0f;
  const float *row = cc + (size_t)t * n_corr;
  unsigned mn = 0xffffffffu;
  // kUnroll loads in flight per thread before any is used
  const long long stride = (long long)gridDim.x * kThreads * kUnroll;
  for (long long i0 = lo + (long long)blockIdx.x * kThreads * kUnroll + threadIdx.x; i0 < hi;
       i0 += stride) {
    float v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const long long i = i0 + u * kThreads;
      v[u] = row[i < hi ? i : lo];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      if (i0 + u * kThreads >= hi) break;
      const unsigned k = fkey(sel_value(v[u], abs_dev, m));
      if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & dmask], 1u);
      else if (k > group_max) mn = k < mn ? k : mn;
    }
  }
  if (shift == 0) {
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned o = __shfl_down(mn, off);
      mn = o < mn ? o : mn;
    }
    if ((threadIdx.x & 63) == 0 && mn != 0xffffffffu) atomicMin(&st[t].min_gt, mn);
  }
  __syncthreads();
  unsigned *g = hist + (size_t)t * kBins;
  for (int i = threadIdx.x; i <= (int)dmask; i += kThreads)
    if (h[i]) atomicAdd(&g[i], h[i]);
}

// One block per row: the bin holding rank k fixes the next digit; the
// histogram is cleared for the next pass.
__global__ __launch_bounds__(kThreads) void select_kernel(
    SelectState *st, int shift, int bits, unsigned *hist) {
  __shared__ unsigned long long part[kThreads];
  // last pass: the bin of rank k when rank k + 1 lies past it, the prefix
  // before this pass, and the next non-empty bin (searched by every thread)
  __shared__ int found_b, next_b;
  __shared__ unsigned old_prefix;
  const int t = blockIdx.x;
  // read before the barriers below: one thread rewrites it after them
  const long long k = st[t].k;
  if (threadIdx.x == 0) {
    found_b = -1;
    next_b = 0x7fffffff;
  }
  unsigned *g = hist + (size_t)t * kBins;
  const int nb = 1 << bits;
  const int per = (nb + kThreads - 1) / kThreads;
  const int b0 = threadIdx.x * per;
  unsigned long long s = 0;
  for (int b = b0; b < b0 + per && b < nb; b++) s += g[b];
  part[threadIdx.x] = s;
  __syncthreads();
  // Inclusive scan of the per-thread sums (Hillis-Steele over 256 entries).
  for (int off = 1; off < kThreads; off <<= 1) {
    const unsigned long long v = threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  const unsigned long long before = threadIdx.x ? part[threadIdx.x - 1] : 0ull;
  if ((long long)before <= k && k < (long long)part[threadIdx.x]) {
    unsigned long long c = before;
    for (int b = b0; b < b0 + per && b < nb; b++) {
      if (k < (long long)(c + g[b])) {
        const unsigned base = st[t].prefix;
        const unsigned key = base | ((unsigned)b << shift);
        st[t].k = k - (long long)c;
        st[t].prefix = key;
        st[t].mask |= (unsigned)(nb - 1) << shift;
        if (shift == 0) {
          // rank k + 1: the same key if the bin holds more, else the next
          // non-empty bin (found below by all threads), else the smallest
          // key above the group
          if (k + 1 < (long long)(c + g[b])) {
            st[t].key_hi = key;
          } else {
            found_b = b;
            old_prefix = base;
          }
        }
        break;
      }
      c += g[b];
    }
  }
  __syncthreads();
  if (shift == 0 && found_b >= 0) {  // the same in every thread (LDS)
    const int from = b0 > found_b + 1 ? b0 : found_b + 1;
    for (int b = from; b < b0 + per && b < nb; b++)
      if (g[b]) {
        atomicMin(&next_b, b);
        break;
      }
    __syncthreads();
    if (threadIdx.x == 0)
      st[t].key_hi = next_b != 0x7fffffff ? (old_prefix | (unsigned)next_b) : st[t].min_gt;
    __syncthreads();
  }
  for (int b = threadIdx.x; b < kBins; b += kThreads) g[b] = 0;
}

// Median of the row from the select state: the lower middle value, or the
// mean of the two middle values for an even count.  Empty rows give 0.
__global__ __launch_bounds__(kThreads) void median_kernel(
    const SelectState *st, int n_templates, float *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_templates) return;
  const SelectState s = st[t];
  if (s.n <= 0) {
    out[t] = 0.0f;
    return;
  }
  const float a = kval(s.prefix);
  if (s.n & 1) {
    out[t] = a;
    return;
  }
  out[t] = __fmul_rn(__fadd_rn(a, kval(s.key_hi)), 0.5f);
}

// ---- sampled bracket -------------------------------------------------------

constexpr int kSample = 16384;                // sampled keys per row
constexpr int kChunkLen = 4;                  // contiguous lags per sample chunk
constexpr int kChunks = kSample / kChunkLen;  // chunks spread evenly over the row
constexpr int kSampleThreads = 1024;
// sample ranks either side of the target: 4 standard deviations of the
// target's sample rank for independent lags (sqrt(16384 / 4) = 64); where
// neighbouring lags are correlated the chunks count as fewer samples, and a
// row whose rank falls outside its bracket takes the exact whole-row select
constexpr int kBracketHalf = 256;
constexpr long long kBracketMinRow = 16LL * kSample;
constexpr int kWaveBuf = 2 * 64 * kUnroll;    // candidate keys buffered per wave

// One block per row: the stratified sample's keys sorted in LDS (bitonic),
// then a and b at the target rank's sample position -+ kBracketHalf (the
// key range's ends where that passes the sample's ends).  Rows shorter than
// kBracketMinRow fall back.
__global__ __launch_bounds__(kSampleThreads) void bracket_sample_kernel(
    const float *__restrict__ cc, long long n_corr, const int *ranges, int abs_dev,
    const float *__restrict__ med, Bracket *__restrict__ br) {
  __shared__ unsigned key[kSample];
  const int t = blockIdx.x;
  long long lo, hi;
  row_range(ranges, t, n_corr, lo, hi);
  const long long n = hi - lo;
  if (n < kBracketMinRow) {
    if (threadIdx.x == 0) {
      br[t].fallback = 1;
      br[t].min_above = 0xffffffffu;
      br[t].count = 0u;
      br[t].below = 0ull;
    }
    return;
  }
  const float m = abs_dev ? med[t] : 0.0f;
  const float *row = cc + (size_t)t * n_corr;
  for (int s = threadIdx.x; s < kSample; s += kSampleThreads) {
    const int c = s / kChunkLen, r = s - c * kChunkLen;
    const long long start = lo + (long long)c * (n - kChunkLen) / (kChunks - 1);
    key[s] = fkey(sel_value(row[start + r], abs_dev, m));
  }
  __syncthreads();
  for (int k = 2; k <= kSample; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < kSample / 2; i += kSampleThreads) {
        const int p = 2 * i - (i & (j - 1));  // i with a zero bit inserted at j
        const int q = p + j;
        const unsigned x = key[p], y = key[q];
        if ((x > y) == ((p & k) == 0)) {
          key[p] = y;
          key[q] = x;
        }
      }
      __syncthreads();
    }
  if (threadIdx.x == 0) {
    const long long k = (n - 1) / 2;
    const long long sr = (long long)((double)k / (double)n * kSample);
    const long long ra = sr - kBracketHalf, rb = sr + kBracketHalf;
    br[t].a = ra <= 0 ? 0u : key[ra];
    br[t].b = rb >= kSample - 1 ? 0xffffffffu : key[rb];
    br[t].fallback = 0;
    br[t].min_above = 0xffffffffu;
    br[t].count = 0u;
    br[t].below = 0ull;
  }
}

// One pass over the rows that have a bracket: keys below a are counted,
// keys in [a, b] go to the row's candidates (per-wave LDS buffers, flushed
// with one atomic per wave; the loop is wave-uniform so every flush runs
// with the whole wave), the smallest key above b is kept.  ABS: the keys of
// |x - median| (non-negative, so the key is the bits with the top bit set).
// BUF: rows below 2^29 lags, read through a buffer resource with the
// unrolled loads' offsets in scalar registers.  Only a wave's last
// iteration of a row tests which lags lie before hi.  The tests use
// wrap-around differences: k in [a, b] iff k - a <= b - a, and k > b iff
// k - (b + 1) < ~b, whose minimum over the lags is the smallest key above b
// less b + 1 (no key is above b = 0xffffffff, and then ~b = 0).
template <int ABS, int BUF>
__global__ __launch_bounds__(kThreads) void bracket_count_kernel(
    const float *__restrict__ cc, long long n_corr, const int *ranges,
    const float *__restrict__ med, Bracket *__restrict__ br, unsigned *__restrict__ cand,
    long long cap) {
  __shared__ unsigned wbuf[kThreads / 64][kWaveBuf];
  const int t = blockIdx.y;
  if (br[t].fallback) return;  // the whole block
  long long lo, hi;
  row_range(ranges, t, n_corr, lo, hi);
  const unsigned a = br[t].a, b = br[t].b;
  const unsigned span = b - a, b1 = b + 1u, nb_ = ~b;
  const float m = ABS ? med[t] : 0.0f;
  const float *row = cc + (size_t)t * n_corr;
  const auto rr = buffer_rsrc(row, BUF ? 4u * (unsigned)n_corr : 0u);
  unsigned *ck = cand + (size_t)t * cap;
  const int lane = threadIdx.x & 63;
  unsigned *wb = wbuf[threadIdx.x >> 6];
  unsigned wn = 0;  // keys in this wave's buffer (the same in every lane)
  auto flush = [&]() {
    __builtin_amdgcn_wave_barrier();
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&br[t].count, wn);
    base = __shfl(base, 0);
    for (unsigned j = lane; j < wn; j += 64)
      if ((long long)base + j < cap) ck[base + j] = wb[j];
    __builtin_amdgcn_wave_barrier();
    wn = 0;
  };
  auto key = [&](float x) {
    if (ABS) return __float_as_uint(fabsf(__fsub_rn(x, m))) | 0x80000000u;
    const unsigned u = __float_as_uint(x);
    return u ^ ((unsigned)((int)u >> 31) | 0x80000000u);
  };
  unsigned long long n_below = 0;  // wave-uniform
  unsigned mo = 0xffffffffu;       // min of k - (b + 1) over this lane's keys
  auto count = [&](const float (&v)[kUnroll], long long i0, bool tail) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const bool valid = !tail || i0 + u * kThreads < hi;
      const unsigned k = key(v[u]);
      const bool in = valid && k - a <= span;
      n_below += (unsigned)__popcll(__ballot(valid && k < a));
      if (valid) mo = min(mo, k - b1);
      const unsigned long long bal = __ballot(in);
      const unsigned below = __builtin_amdgcn_mbcnt_hi(
          (unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
      if (in) wb[wn + below] = k;
      wn += (unsigned)__popcll(bal);
    }
  };
  const long long stride = (long long)gridDim.x * kThreads * kUnroll;
  for (long long w0 = lo + (long long)blockIdx.x * kThreads * kUnroll + (threadIdx.x & ~63);
       w0 < hi; w0 += stride) {
    const long long i0 = w0 + lane;
    const bool tail = w0 + (kUnroll - 1) * kThreads + 64 > hi;  // wave-uniform
    float v[kUnroll];
    if (BUF) {
      // 4 * i < 2^31; past the row's end (only in the tail) the load reads 0
#pragma unroll
      for (int u = 0; u < kUnroll; u++)
        v[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                             rr, (int)(4 * i0), 4 * u * kThreads, 0));
    } else {
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        const long long i = i0 + u * kThreads;
        v[u] = row[i < hi ? i : lo];
      }
    }
    if (tail) count(v, i0, true);
    else count(v, i0, false);
    if (wn > kWaveBuf / 2) flush();  // at most 64 * kUnroll more per iteration
  }
  unsigned mn = mo < nb_ ? mo + b1 : 0xffffffffu;
  // the block's remainder: one atomic per field for the whole block (the
  // fields of a row share a cache line, so per-wave atomics would queue)
  __shared__ unsigned w_n[kThreads / 64], w_min[kThreads / 64], w_base[kThreads / 64];
  __shared__ unsigned long long w_below[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = __shfl_down(mn, off);
    mn = o < mn ? o : mn;
  }
  const int wave = threadIdx.x >> 6;
  if (lane == 0) {
    w_n[wave] = wn;
    w_min[wave] = mn;
    w_below[wave] = n_below;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned total = 0, m_all = 0xffffffffu;
    unsigned long long b_all = 0;
    for (int q = 0; q < kThreads / 64; q++) {
      w_base[q] = total;
      total += w_n[q];
      m_all = w_min[q] < m_all ? w_min[q] : m_all;
      b_all += w_below[q];
    }
    const unsigned base = total ? atomicAdd(&br[t].count, total) : 0u;
    for (int q = 0; q < kThreads / 64; q++) w_base[q] += base;
    if (b_all) atomicAdd(&br[t].below, b_all);
    if (m_all != 0xffffffffu) atomicMin(&br[t].min_above, m_all);
  }
  __syncthreads();
  const unsigned base = w_base[wave];
  for (unsigned j = lane; j < wn; j += 64)
    if ((long long)base + j < cap) ck[base + j] = wb[j];
}

// Per row: the select state of the lower middle rank, over the candidates
// when the rank lies inside the bracket and the buffer held them all
// (rank k - below, the smallest key above b as the start of the rank k + 1
// search), else over the whole row (fallback).
__global__ __launch_bounds__(kThreads) void bracket_decide_kernel(
    const int *ranges, long long n_corr, int n_templates, long long cap,
    Bracket *__restrict__ br, SelectState *__restrict__ st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_templates) return;
  long long lo, hi;
  row_range(ranges, t, n_corr, lo, hi);
  SelectState s;
  s.prefix = 0;
  s.mask = 0;
  s.n = hi - lo;
  s.k = s.n > 0 ? (s.n - 1) / 2 : 0;
  s.min_gt = 0xffffffffu;
  s.key_hi = 0;
  if (!br[t].fallback) {
    const long long below = (long long)br[t].below, count = (long long)br[t].count;
    if (count <= cap && below <= s.k && s.k < below + count) {
      s.k -= below;
      s.min_gt = br[t].min_above;
    } else {
      br[t].fallback = 1;
    }
  }
  st[t] = s;
}

// hist_kernel over a row's candidate keys (rows with a bracket only).
__global__ __launch_bounds__(kThreads) void hist_cand_kernel(
    const Bracket *__restrict__ br, const unsigned *__restrict__ cand, long long cap,
    SelectState *__restrict__ st, int shift, int bits, unsigned *__restrict__ hist) {
  __shared__ unsigned h[kBins];
  const int t = blockIdx.y;
  if (br[t].fallback) return;  // the whole block
  for (int i = threadIdx.x; i < kBins; i += kThreads) h[i] = 0;
  __syncthreads();
  const long long count = br[t].count;  // <= cap: the decision checked it
  const unsigned prefix = st[t].prefix, mask = st[t].mask;
  const unsigned group_max = prefix | ~mask;
  const unsigned dmask = (1u << bits) - 1u;
  const unsigned *ck = cand + (size_t)t * cap;
  const int lane = threadIdx.x & 63;
  unsigned mn = 0xffffffffu;
  // kUnroll loads in flight per thread; the loop is wave-uniform (ballots).
  // Candidates lie in a narrow key range, so many lanes of a wave share a
  // bin: the first counted lane's bin takes one atomic for all the lanes in
  // it, the other lanes one each.
  const long long stride = (long long)gridDim.x * kThreads * kUnroll;
  for (long long w0 = (long long)blockIdx.x * kThreads * kUnroll + (threadIdx.x & ~63);
       w0 < count; w0 += stride) {
    const long long i0 = w0 + lane;
    unsigned kk[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const long long i = i0 + u * kThreads;
      kk[u] = ck[i < count ? i : 0];  // count >= 1 inside the loop
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const unsigned k = kk[u];
      const bool valid = i0 + u * kThreads < count;
      const bool in = valid && (k & mask) == prefix;
      const unsigned bin = (k >> shift) & dmask;
      const unsigned long long inb = __ballot(in);
      if (inb) {
        const int lead = (int)__builtin_ctzll(inb);
        const unsigned lbin = __shfl(bin, lead);
        const bool same = in && bin == lbin;
        const unsigned long long sb = __ballot(same);
        if (lane == lead) atomicAdd(&h[lbin], (unsigned)__popcll(sb));
        if (in && !same) atomicAdd(&h[bin], 1u);
      }
      if (valid && !in && k > group_max) mn = k < mn ? k : mn;
    }
  }
  if (shift == 0) {
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned o = __shfl_down(mn, off);
      mn = o < mn ? o : mn;
    }
    if ((threadIdx.x & 63) == 0 && mn != 0xffffffffu) atomicMin(&st[t].min_gt, mn);
  }
  __syncthreads();
  unsigned *g = hist + (size_t)t * kBins;
  for (int i = threadIdx.x; i <= (int)dmask; i += kThreads)
    if (h[i]) atomicAdd(&g[i], h[i]);
}

__global__ __launch_bounds__(kThreads) void threshold_kernel(
    const float *med, const float *mad, float n_mad, int n_templates, float *thr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_templates) return;
  thr[t] = __fadd_rn(med[t], __fmul_rn(n_mad, mad[t]));
}

// Maximum of each 64-lag segment over the row's valid lags (-inf if none).
// A block covers kSegPerBlock segments; each wave loads its 16 segments'
// lags before reducing them.
constexpr int kSegPerBlock = 64;

// 16 lanes per segment, 4 consecutive lags each: a wave covers 4 segments
// per step (1 KB contiguous) and reduces each inside its 16-lane group in 4
// shuffles (against 6 per segment with one segment per wave).
__global__ __launch_bounds__(kThreads) void segmax_kernel(
    const float *__restrict__ cc, long long n_corr, const int *ranges,
    long long n_seg, float *__restrict__ segmax) {
  constexpr int kPerWave = kSegPerBlock / (kThreads / 64);  // segments per wave
  constexpr int kSteps = kPerWave / 4;
  const int t = blockIdx.y;
  long long lo, hi;
  row_range(ranges, t, n_corr, lo, hi);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane >> 4, l16 = lane & 15;
  const long long seg0 = (long long)blockIdx.x * kSegPerBlock + wave * kPerWave;
  const float *row = cc + (size_t)t * n_corr;
  float v[kSteps][4];
  if (n_corr < (1LL << 29)) {
    // a lane's 4 lags in one 16-byte load through a buffer resource over the
    // row (lags past the row read 0 — the bounds are checked per dword, so a
    // load across the end of the row returns its lags before the end; every
    // lag outside [lo, hi) becomes -inf)
    const auto rr = buffer_rsrc(row, 4u * (unsigned)n_corr);
#pragma unroll
    for (int u = 0; u < kSteps; u++) {
      const long long i0 = (seg0 + 4 * u + sub) * kSeg + 4 * l16;
      typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
      // cache policy nt (aux bit 1 on gfx950): a streaming read, the rows
      // read again only where a segment passes the threshold; detection pass
      // of the batched output 12.36 -> 11.65 ms on one box, while the same
      // policy on the counting pass's loads made that pass 10 % slower
      const u32x4_t q = __builtin_amdgcn_raw_buffer_load_b128(rr, (int)(4 * i0), 0, 2);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const long long i = i0 + e;
        v[u][e] = (i >= lo && i < hi) ? __uint_as_float(q[e]) : -__builtin_inff();
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < kSteps; u++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const long long i = (seg0 + 4 * u + sub) * kSeg + 4 * l16 + e;
        const bool in = i >= lo && i < hi;
        const float x = row[in ? i : 0];  // lag 0 always exists (n_corr >= 1)
        v[u][e] = in ? x : -__builtin_inff();
      }
  }
#pragma unroll
  for (int u = 0; u < kSteps; u++) {
    float x = fmaxf(fmaxf(v[u][0], v[u][1]), fmaxf(v[u][2], v[u][3]));
    for (int off = 8; off > 0; off >>= 1) {
      const float o = __shfl_xor(x, off, 16);
      x = o > x ? o : x;
    }
    const long long seg = seg0 + 4 * u + sub;
    if (l16 == 0 && seg < n_seg) segmax[(size_t)t * n_seg + seg] = x;
  }
}

// Is lag j (value v) the first maximum of [a, b]?
__device__ bool is_first_max(const float *row, const float *sm, long long a,
                             long long j, long long b, float v) {
  // lags before j: strictly smaller
  long long i = a;
  while (i < j) {
    if ((i & (kSeg - 1)) == 0 && i + kSeg <= j) {
      // a whole segment inside [a, j): its maximum decides
      if (!(sm[i / kSeg] < v)) return false;
      i += kSeg;
    } else {
      if (!(row[i] < v)) return false;
      i++;
    }
  }
  // lags after j: not larger
  i = j + 1;
  while (i <= b) {
    if ((i & (kSeg - 1)) == 0 && i + kSeg - 1 <= b) {
      if (sm[i / kSeg] > v) return false;
      i += kSeg;
    } else {
      if (row[i] > v) return false;
      i++;
    }
  }
  return true;
}

// write = 0: count the detections of every 4096-lag chunk of the rows;
// write = 1: write them at offsets[t][chunk] in lag order (entries past
// `capacity` are not written; chunks without detections are skipped).  One
// wave owns a chunk at a time, lane s holding the maximum of the chunk's
// segment s: a lag can only be a detection inside a segment whose maximum is
// above the threshold, so the wave reads the lags of those segments alone
// (`<=` skips a segment, so a NaN maximum or threshold keeps it), one segment
// per wave load, in segment order — the order of the output.  Each wave
// takes the chunks blockIdx.x * 4 + wave + k * gridDim.x * 4 of row
// blockIdx.y and loads the next chunk's maxima before it tests the current
// one.
__global__ __launch_bounds__(kThreads) void detect_kernel(
    const float *__restrict__ cc, long long n_corr, const int *ranges,
    const float *__restrict__ thr, const float *__restrict__ segmax,
    long long n_seg, long long sep, int n_blk, int write,
    unsigned *__restrict__ counts, const unsigned long long *__restrict__ offsets,
    int *__restrict__ det_t, int *__restrict__ det_lag, float *__restrict__ det_val,
    long long capacity) {
  constexpr int kSegPerChunk = kLagsPerChunk / kSeg;
  static_assert(kSegPerChunk == 64, "one lane per segment of a chunk");
  constexpr int kWaves = kThreads / 64;
  const int t = blockIdx.y;
  long long lo, hi;
  row_range(ranges, t, n_corr, lo, hi);
  const float *row = cc + (size_t)t * n_corr;
  const float *sm = segmax + (size_t)t * n_seg;
  const float th = thr[t];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned long long below = (1ull << lane) - 1ull;
  const long long stride = (long long)gridDim.x * kWaves;
  auto seg_max = [&](long long c) {
    const long long seg = c * kSegPerChunk + lane;
    const bool in = c < n_blk && seg < n_seg;
    const float v = sm[in ? seg : 0];  // segment 0 always exists
    return in && !(v <= th);
  };
  long long c = (long long)blockIdx.x * kWaves + wave;
  bool live_next = seg_max(c);
  for (; c < n_blk; c += stride) {
    unsigned long long live = __ballot(live_next);
    live_next = seg_max(c + stride);
    const size_t cidx = (size_t)t * n_blk + c;
    unsigned long long base = 0;
    if (write) {
      if (live == 0 || counts[cidx] == 0) continue;
      base = offsets[cidx];
    }
    unsigned cnt = 0;
    while (live) {
      const int s = (int)__builtin_ctzll(live);
      live &= live - 1ull;
      const long long j = c * kLagsPerChunk + (long long)s * kSeg + lane;
      bool d = false;
      float v = 0.0f;
      if (j >= lo && j < hi) {
        v = row[j];
        if (v > th) {
          const long long a = j - sep < lo ? lo : j - sep;
          const long long b = j + sep > hi - 1 ? hi - 1 : j + sep;
          d = is_first_max(row, sm, a, j, b, v);
        }
      }
      const unsigned long long bal = __ballot(d);
      if (write && d) {
        const unsigned long long pos = base + cnt + (unsigned)__popcll(bal & below);
        if ((long long)pos < capacity) {
          det_t[pos] = t;
          det_lag[pos] = (int)j;
          det_val[pos] = v;
        }
      }
      cnt += (unsigned)__popcll(bal);
    }
    if (!write && lane == 0) counts[cidx] = cnt;
  }
}

// Exclusive scan of n counts into offsets in three coalesced steps: sums of
// chunks of 4096 counts, a one-block scan of the chunk sums (total to
// *n_out), and a rescan of every chunk from its offset.
constexpr int kScanChunk = 4096;

__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long v,
                                                              unsigned long long *sh,
                                                              unsigned long long &total) {
  // 256 threads: wave inclusive scan, then the wave totals
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long o = __shfl_up(incl, off);
    if (lane >= off) incl += o;
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  unsigned long long base = 0;
  total = 0;
  for (int w = 0; w < kThreads / 64; w++) {
    if (w < wave) base += sh[w];
    total += sh[w];
  }
  __syncthreads();
  return base + incl - v;
}

__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const unsigned *counts,
                                                               long long n,
                                                               unsigned long long *chunk_sum) {
  __shared__ unsigned long long sh[kThreads / 64];
  const long long c0 = (long long)blockIdx.x * kScanChunk;
  unsigned long long v = 0;
  for (int i = threadIdx.x; i < kScanChunk; i += kThreads)
    if (c0 + i < n) v += counts[c0 + i];
  unsigned long long total;
  block_excl_scan(v, sh, total);
  if (threadIdx.x == 0) chunk_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void scan_top_kernel(unsigned long long *chunk_sum,
                                                            long long n_chunks,
                                                            long long *n_out) {
  __shared__ unsigned long long sh[kThreads / 64];
  unsigned long long carry = 0;
  for (long long c0 = 0; c0 < n_chunks; c0 += kThreads) {
    const long long i = c0 + threadIdx.x;
    const unsigned long long v = i < n_chunks ? chunk_sum[i] : 0ull;
    unsigned long long total;
    const unsigned long long ex = block_excl_scan(v, sh, total);
    if (i < n_chunks) chunk_sum[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *n_out = (long long)carry;
}

__global__ __launch_bounds__(kThreads) void scan_down_kernel(const unsigned *counts,
                                                             long long n,
                                                             const unsigned long long *chunk_off,
                                                             unsigned long long *offsets) {
  __shared__ unsigned long long sh[kThreads / 64];
  constexpr int kPer = kScanChunk / kThreads;
  const long long c0 = (long long)blockIdx.x * kScanChunk + threadIdx.x * kPer;
  unsigned v[kPer];
  unsigned long long mine = 0;
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    v[q] = c0 + q < n ? counts[c0 + q] : 0u;
    mine += v[q];
  }
  unsigned long long total;
  unsigned long long run = chunk_off[blockIdx.x] + block_excl_scan(mine, sh, total);
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    if (c0 + q < n) offsets[c0 + q] = run;
    run += v[q];
  }
}

__global__ void fill_u32_kernel(unsigned *p, long long n, unsigned v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct DetectWs {
  SelectState *st;
  unsigned *hist;
  Bracket *br;       // (T) sampled brackets
  unsigned *cand;    // (T, cand_cap) candidate keys
  long long cand_cap;
  float *segmax;
  unsigned *counts;
  unsigned long long *offsets;
  float *thr;
  unsigned long long *chunk;
  size_t bytes;
};

DetectWs carve(void *base, long long T, long long n_corr) {
  DetectWs w{};
  const long long n_seg = (n_corr + kSeg - 1) / kSeg;
  const long long n_blk = (n_corr + kLagsPerChunk - 1) / kLagsPerChunk;
  char *p = static_cast<char *>(base);
  size_t off = 0;
  auto take = [&](size_t b) { char *r = p ? p + off : nullptr; off += align256(b); return r; };
  w.st = reinterpret_cast<SelectState *>(take(sizeof(SelectState) * T));
  w.hist = reinterpret_cast<unsigned *>(take(sizeof(unsigned) * kBins * T));
  w.br = reinterpret_cast<Bracket *>(take(sizeof(Bracket) * T));
  // candidate buffers only for rows long enough for a bracket
  w.cand_cap = n_corr >= kBracketMinRow ? (n_corr + 7) / 8 : 0;
  w.cand = reinterpret_cast<unsigned *>(take(sizeof(unsigned) * w.cand_cap * T));
  w.segmax = reinterpret_cast<float *>(take(sizeof(float) * n_seg * T));
  w.counts = reinterpret_cast<unsigned *>(take(sizeof(unsigned) * n_blk * T));
  w.offsets = reinterpret_cast<unsigned long long *>(take(sizeof(unsigned long long) * n_blk * T));
  w.thr = reinterpret_cast<float *>(take(sizeof(float) * T));
  w.chunk = reinterpret_cast<unsigned long long *>(
      take(sizeof(unsigned long long) * ((n_blk * T + kScanChunk - 1) / kScanChunk)));
  w.bytes = off;
  return w;
}

unsigned row_blocks(long long n_corr) {
  long long b = (n_corr + 16383) / 16384;
  return (unsigned)(b < 1 ? 1 : (b > 512 ? 512 : b));
}

// Exact median of every row (abs_dev = 0) or of |row - med| (abs_dev = 1).
int row_median(const float *cc, long long T, long long n_corr, const int *ranges,
               int abs_dev, const float *med, float *out, const DetectWs &w,
               hipStream_t s) {
  const unsigned tb = (unsigned)((T + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(init_select_kernel, dim3(tb), dim3(kThreads), 0, s, ranges,
                     n_corr, (int)T, w.st);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(256), dim3(256), 0, s, w.hist,
                     (long long)kBins * T, 0u);
  const int shifts[3] = {21, 10, 0}, widths[3] = {11, 11, 10};
  const dim3 grid(row_blocks(n_corr), (unsigned)T);
  for (int p = 0; p < 3; p++) {
    hipLaunchKernelGGL(hist_kernel, grid, dim3(kThreads), 0, s, cc, n_corr, ranges,
                       abs_dev, med, w.st, shifts[p], widths[p], w.hist,
                       static_cast<const Bracket *>(nullptr));
    hipLaunchKernelGGL(select_kernel, dim3((unsigned)T), dim3(kThreads), 0, s, w.st,
                       shifts[p], widths[p], w.hist);
  }
  hipLaunchKernelGGL(median_kernel, dim3(tb), dim3(kThreads), 0, s, w.st, (int)T, out);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

// The same median through the sampled bracket: sample and sort, one counting
// pass, the radix select over the candidates of the rows whose bracket holds
// their rank, and over the whole row for the others (their blocks return at
// once in the other launches, and a select over an empty histogram leaves a
// row's state as it is).
int row_median_bracket(const float *cc, long long T, long long n_corr, const int *ranges,
                       int abs_dev, const float *med, float *out, const DetectWs &w,
                       hipStream_t s) {
  const unsigned tb = (unsigned)((T + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(bracket_sample_kernel, dim3((unsigned)T), dim3(kSampleThreads), 0, s, cc,
                     n_corr, ranges, abs_dev, med, w.br);
  const dim3 grid(row_blocks(n_corr), (unsigned)T);
  // at most 128 blocks per row: each block then runs several iterations, and
  // its end (block reduction, three atomics) is paid less often
  const dim3 bgrid(row_blocks(n_corr) < 128 ? row_blocks(n_corr) : 128, (unsigned)T);
  // buffer loads while every offset, the last wave's overrun included, stays
  // below 2^31 bytes
  const bool buf = n_corr <= (1LL << 29) - 4096;
  auto count = abs_dev ? (buf ? bracket_count_kernel<1, 1> : bracket_count_kernel<1, 0>)
                       : (buf ? bracket_count_kernel<0, 1> : bracket_count_kernel<0, 0>);
  hipLaunchKernelGGL(count, bgrid, dim3(kThreads), 0, s, cc, n_corr, ranges, med, w.br,
                     w.cand, w.cand_cap);
  hipLaunchKernelGGL(bracket_decide_kernel, dim3(tb), dim3(kThreads), 0, s, ranges, n_corr,
                     (int)T, w.cand_cap, w.br, w.st);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(256), dim3(256), 0, s, w.hist,
                     (long long)kBins * T, 0u);
  const int shifts[3] = {21, 10, 0}, widths[3] = {11, 11, 10};
  const dim3 cgrid(row_blocks(w.cand_cap), (unsigned)T);
  for (int p = 0; p < 3; p++) {
    hipLaunchKernelGGL(hist_cand_kernel, cgrid, dim3(kThreads), 0, s, w.br, w.cand, w.cand_cap,
                       w.st, shifts[p], widths[p], w.hist);
    hipLaunchKernelGGL(select_kernel, dim3((unsigned)T), dim3(kThreads), 0, s, w.st,
                       shifts[p], widths[p], w.hist);
  }
  // whole-row passes for the rows that fell back (rare: the other rows'
  // blocks return at once, so a smaller grid per row keeps those launches
  // cheap when no row fell back)
  const dim3 fgrid(row_blocks(n_corr) < 64 ? row_blocks(n_corr) : 64, (unsigned)T);
  for (int p = 0; p < 3; p++) {
    hipLaunchKernelGGL(hist_kernel, fgrid, dim3(kThreads), 0, s, cc, n_corr, ranges, abs_dev,
                       med, w.st, shifts[p], widths[p], w.hist,
                       static_cast<const Bracket *>(w.br));
    hipLaunchKernelGGL(select_kernel, dim3((unsigned)T), dim3(kThreads), 0, s, w.st,
                       shifts[p], widths[p], w.hist);
  }
  hipLaunchKernelGGL(median_kernel, dim3(tb), dim3(kThreads), 0, s, w.st, (int)T, out);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace

size_t detect_workspace_bytes(long long T, long long n_corr) {
  return carve(nullptr, T, n_corr).bytes;
}

int launch_median_mad(const float *cc, long long T, long long n_corr,
                      const int *ranges, float *med, float *mad, void *ws,
                      size_t ws_bytes, hipStream_t s) {
  const DetectWs w = carve(ws, T, n_corr);
  if (ws_bytes < w.bytes)
    return set_error(kInvalidArgument, "detection workspace too small (%zu < %zu bytes)",
                     ws_bytes, w.bytes);
  auto sel = n_corr >= kBracketMinRow ? row_median_bracket : row_median;
  int rc = sel(cc, T, n_corr, ranges, 0, nullptr, med, w, s);
  if (rc) return rc;
  return sel(cc, T, n_corr, ranges, 1, med, mad, w, s);
}

int launch_detect(const float *cc, long long T, long long n_corr, const int *ranges,
                  const float *med, const float *mad, float n_mad, long long sep,
                  int *det_t, int *det_lag, float *det_val, long long capacity,
                  long long *n_det, void *ws, size_t ws_bytes, hipStream_t s) {
  const DetectWs w = carve(ws, T, n_corr);
  if (ws_bytes < w.bytes)
    return set_error(kInvalidArgument, "detection workspace too small (%zu < %zu bytes)",
                     ws_bytes, w.bytes);
  const long long n_seg = (n_corr + kSeg - 1) / kSeg;
  const long long n_blk = (n_corr + kLagsPerChunk - 1) / kLagsPerChunk;
  const unsigned tb = (unsigned)((T + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(threshold_kernel, dim3(tb), dim3(kThreads), 0, s, med, mad, n_mad,
                     (int)T, w.thr);
  hipLaunchKernelGGL(segmax_kernel,
                     dim3((unsigned)((n_seg + kSegPerBlock - 1) / kSegPerBlock), (unsigned)T),
                     dim3(kThreads), 0, s, cc, n_corr, ranges, n_seg, w.segmax);
  // about 8 chunks per wave (4 waves per block) in each row
  const dim3 grid((unsigned)((n_blk + 31) / 32), (unsigned)T);
  hipLaunchKernelGGL(detect_kernel, grid, dim3(kThreads), 0, s, cc, n_corr, ranges,
                     w.thr, w.segmax, n_seg, sep, (int)n_blk, 0, w.counts, w.offsets,
                     det_t, det_lag, det_val, capacity);
  const long long n_cnt = n_blk * T;
  const long long n_chunks = (n_cnt + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)n_chunks), dim3(kThreads), 0, s,
                     w.counts, n_cnt, w.chunk);
  hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kThreads), 0, s, w.chunk, n_chunks, n_det);
  hipLaunchKernelGGL(scan_down_kernel, dim3((unsigned)n_chunks), dim3(kThreads), 0, s,
                     w.counts, n_cnt, w.chunk, w.offsets);
  hipLaunchKernelGGL(detect_kernel, grid, dim3(kThreads), 0, s, cc, n_corr, ranges,
                     w.thr, w.segmax, n_seg, sep, (int)n_blk, 1, w.counts, w.offsets,
                     det_t, det_lag, det_val, capacity);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace fmf
