// Window energies (sliding sums of squares) and small preparation kernels.
//
// The reference normalises every correlation by the energy of the data
// window, taken as a difference of an FP64 running sum of squares per
// channel (fast_matched_filter/src/matched_filter.c:660-705, used at :622).
// Here each block owns a tile of window positions of one channel and builds
// FP64 prefix sums local to that tile, so the window sums carry the rounding
// of a few thousand terms instead of the whole trace.  A parallel integer
// count of non-zero samples makes windows that lie entirely in a data gap
// exactly 0, as the reference's constant running sum does.
#include <algorithm>

#include "fmf_internal.h"

namespace fmf {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr int kTile = kThreads * kPerThread;  // window positions per block

// Exclusive block-wide scan of one double per thread (wave64 shuffles + LDS).
template <typename V>
__device__ V block_exclusive_scan(V v, V *wave_tot, V &total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  V inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    V o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  V off = 0;
  V tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; w++) {
    if (w < wave) off += wave_tot[w];
    tot += wave_tot[w];
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

// Block reduction of a double / int over the block.
template <typename V>
__device__ V block_sum(V v, V *wave_tot) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (lane == 0) wave_tot[wave] = v;
  __syncthreads();
  V tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; w++) tot += wave_tot[w];
  __syncthreads();
  return tot;
}

// energy[ch][p] = sum_{k<N} d[p+k]^2              (full = 0)
//              = sum d^2 - (sum d)^2 / N           (full = 1, window mean removed)
// for p in [0, n_win), n_win = L - N + 1.  Window [p, p+N) of a tile starting
// at p0 is S + B[j] - A[j] with j = p - p0, A the exclusive prefix over
// [p0, p0+kTile), B the one over [p0+N, p0+N+kTile) and S the sum over
// [p0, p0+N).
__global__ __launch_bounds__(kThreads) void window_energy_kernel(
    const float *__restrict__ data, long long n_data, long long n_tp,
    long long n_win, int full, float *__restrict__ energy) {
  __shared__ double sh_d[kThreads / 64];
  __shared__ int sh_i[kThreads / 64];
  const long long ch = blockIdx.y;
  const long long p0 = (long long)blockIdx.x * kTile;
  const float *d = data + ch * n_data;

  // S over [p0, p0 + N): strided block reduction.
  double s2 = 0.0, s1 = 0.0;
  int sc = 0;
  for (long long i = threadIdx.x; i < n_tp; i += kThreads) {
    long long g = p0 + i;
    float v = g < n_data ? d[g] : 0.0f;
    s2 += (double)v * (double)v;
    s1 += (double)v;
    sc += v != 0.0f;
  }
  const double S2 = block_sum(s2, sh_d);
  const double S1 = full ? block_sum(s1, sh_d) : 0.0;
  const int SC = block_sum(sc, sh_i);

  // Local values of the two segments.
  float va[kPerThread], vb[kPerThread];
  const long long j0 = (long long)threadIdx.x * kPerThread;
#pragma unroll
  for (int e = 0; e < kPerThread; e++) {
    long long ga = p0 + j0 + e;
    long long gb = p0 + n_tp + j0 + e;
    va[e] = ga < n_data ? d[ga] : 0.0f;
    vb[e] = gb < n_data ? d[gb] : 0.0f;
  }
  double ta2 = 0, tb2 = 0, ta1 = 0, tb1 = 0;
  int tac = 0, tbc = 0;
#pragma unroll
  for (int e = 0; e < kPerThread; e++) {
    ta2 += (double)va[e] * (double)va[e];
    tb2 += (double)vb[e] * (double)vb[e];
    ta1 += (double)va[e];
    tb1 += (double)vb[e];
    tac += va[e] != 0.0f;
    tbc += vb[e] != 0.0f;
  }
  double tot_d;
  int tot_i;
  double a2 = block_exclusive_scan(ta2, sh_d, tot_d);
  double b2 = block_exclusive_scan(tb2, sh_d, tot_d);
  double a1 = 0, b1 = 0;
  if (full) {
    a1 = block_exclusive_scan(ta1, sh_d, tot_d);
    b1 = block_exclusive_scan(tb1, sh_d, tot_d);
  }
  int ac = block_exclusive_scan(tac, sh_i, tot_i);
  int bc = block_exclusive_scan(tbc, sh_i, tot_i);

  const double inv_n = 1.0 / (double)n_tp;
#pragma unroll
  for (int e = 0; e < kPerThread; e++) {
    const long long p = p0 + j0 + e;
    if (p < n_win) {
      const int cnt = SC + bc - ac;
      double w2 = (S2 + b2) - a2;
      float out;
      if (cnt == 0) {
        out = 0.0f;
      } else if (full) {
        double w1 = (S1 + b1) - a1;
        double v = w2 - w1 * w1 * inv_n;
        out = (float)(v > 0.0 ? v : 0.0);
      } else {
        out = (float)(w2 > 0.0 ? w2 : 0.0);
      }
      energy[ch * n_win + p] = out;
    }
    const double xa = (double)va[e], xb = (double)vb[e];
    a2 += xa * xa;
    b2 += xb * xb;
    a1 += xa;
    b1 += xb;
    ac += va[e] != 0.0f;
    bc += vb[e] != 0.0f;
  }
}

// Same energies, one pass over the union of a tile's windows: a block owns
// `tile` window positions [p0, p0 + tile) and stages the samples
// [p0, p0 + tile + N) once (reads 1 + N / tile per output instead of
// 2 + N / 1024).  Every thread takes a contiguous run of the union, the
// block scans the runs' FP64 sums of squares (and sums, for 'full'), the
// exclusive prefixes of every position go to LDS, and window p is the
// difference of the prefixes at j + N and j (j = p - p0), written by
// consecutive lanes.  A window of zeros lies between two equal prefixes (an
// FP64 sum does not change when 0 is added), so its energy is exactly 0, as
// with the reference's running sum; no separate count is needed.
//
// One launch serves up to kTabsMax template lengths over the same data: the
// union is staged for the longest, the prefixes are built once and every
// table takes its windows from them, so the data is read once for all of
// them.  The prefix at a position depends on the run length and the tile
// only (runs past the union add 0), and both are fixed for every length up
// to kFixedRunMaxN: a table is the same bits whichever lengths share its
// launch.
constexpr int kUThreads = 256;
constexpr int kUMaxPer = 32;  // union positions per thread (union <= 8192)
constexpr int kTabsMax = 8;   // tables per launch
constexpr int kUnionTile = 4096;
// odd (a wave's run reads on distinct LDS banks), covering the union of a
// 4096-window tile for templates of up to 1280 samples (longer ones take a
// launch of their own in every call, with the run length of their union)
constexpr int kFixedRun = 21;
constexpr int kFixedRunMaxN = kUThreads * kFixedRun - kUnionTile;

struct EnergyTabs {
  int n;         // tables
  int n_tp_max;  // longest template: the union is tile + n_tp_max samples
  int per;       // union positions per thread
  int n_tp[kTabsMax];
  int n_blk[kTabsMax];  // blocks of the table: ceil(n_win / tile)
  long long n_win[kTabsMax];
  float *energy[kTabsMax];
  unsigned *blk_top[kTabsMax];  // RSQ: per-block maxima, (channel, block); nullptr: none
};

// RSQ = 1 writes 1/sqrt(energy) instead (0 for a zero energy), the table the
// 32x32 MFMA kernels read, and each block's largest such value.  MULTI = 0:
// one table (tabs.n = 1), compiled without the loop over tables (the loop
// cost a single table about 5 %).
template <int FULL, int RSQ, int MULTI>
__global__ __launch_bounds__(kUThreads) void window_energy_union_kernel(
    const float *__restrict__ data, long long n_data, int tile, const EnergyTabs tabs) {
  extern __shared__ __attribute__((aligned(16))) double u_lds[];
  __shared__ double sh_d[kUThreads / 64];
  const int U = tile + tabs.n_tp_max;
  double *P2 = u_lds;                                  // U + 1
  double *P1 = P2 + (U + 1);                           // U + 1 ('full')
  float *D = reinterpret_cast<float *>(u_lds);         // staging, then P2
  const long long ch = blockIdx.y;
  const int p0 = blockIdx.x * tile;
  // samples past the data read as 0 through the buffer resource; all loads
  // of a thread are in flight before the first LDS store
  const auto rd = buffer_rsrc(data + ch * n_data, 4u * (unsigned)n_data);
  float v[kUMaxPer];
  // 4 consecutive samples per thread and 16-byte load: positions 4 * (tid +
  // 256 q) .. + 3 of the union; groups wholly past the union take an
  // out-of-range offset (no memory read), samples past the data read 0 (the
  // bounds are checked per dword: a load across the end of the channel
  // returns its samples before the end)
  typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int q = 0; q < kUMaxPer / 4; q++) {
    const int u = 4 * (threadIdx.x + q * kUThreads);
    const u32x4_t x = __builtin_amdgcn_raw_buffer_load_b128(
        rd, u < U ? 4 * (p0 + u) : (int)0x80000000u, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) v[4 * q + i] = __uint_as_float(x[i]);
  }
#pragma unroll
  for (int q = 0; q < kUMaxPer / 4; q++) {
    const int u = 4 * (threadIdx.x + q * kUThreads);
    if (u + 3 < U) {
      *reinterpret_cast<float4 *>(D + u) = float4{v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++)
        if (u + i < U) D[u + i] = v[4 * q + i];
    }
  }
  __syncthreads();
  const int per = tabs.per;
  const int u0 = threadIdx.x * per;
  double s2 = 0.0, s1 = 0.0;
#pragma unroll
  for (int e = 0; e < kUMaxPer; e++) {
    v[e] = (e < per && u0 + e < U) ? D[u0 + e] : 0.0f;
    s2 += (double)v[e] * (double)v[e];
    s1 += (double)v[e];
  }
  double t2, t1;
  double a2 = block_exclusive_scan(s2, sh_d, t2);
  double a1 = FULL ? block_exclusive_scan(s1, sh_d, t1) : 0.0;
  __syncthreads();  // every staged sample is in registers before P2 overwrites D
#pragma unroll
  for (int e = 0; e < kUMaxPer; e++) {
    const int u = u0 + e;
    if (e < per && u < U) {
      P2[u] = a2;
      if (FULL) P1[u] = a1;
      a2 += (double)v[e] * (double)v[e];
      a1 += (double)v[e];
    }
  }
  if (threadIdx.x == 0) {
    P2[U] = t2;
    if (FULL) P1[U] = t1;
  }
  __syncthreads();
  // the block's largest r per table (RSQ), one word per (channel, block);
  // reduced per channel by rmax_reduce_kernel (atomics from every XCD onto a
  // few cache lines doubled this kernel's time)
  constexpr int kTabs = MULTI ? kTabsMax : 1;
  __shared__ unsigned sh_top[kTabs][kUThreads / 64];
  const int n_tabs = MULTI ? tabs.n : 1;
  for (int k = 0; k < n_tabs; k++) {  // block-uniform
    const long long n_win = tabs.n_win[k];
    if (p0 >= n_win) continue;
    const int n_tp = tabs.n_tp[k];
    float *en = tabs.energy[k] + ch * n_win;
    const double inv_n = 1.0 / (double)n_tp;
    unsigned top = 0;  // RSQ: largest 1/sqrt(e) of this thread's windows, as bits
    // four windows per thread at a time, their prefix reads issued together
    for (int j0 = threadIdx.x; j0 < tile; j0 += 4 * kUThreads) {
      double w2[4], w1[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int j = j0 + i * kUThreads;
        const int jj = j < tile ? j : 0;
        w2[i] = P2[jj + n_tp] - P2[jj];
        w1[i] = FULL ? P1[jj + n_tp] - P1[jj] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int j = j0 + i * kUThreads;
        const long long p = p0 + j;
        if (j >= tile || p >= n_win) continue;
        float out;
        if (FULL) {
          const double var = w2[i] - w1[i] * w1[i] * inv_n;
          out = (float)(var > 0.0 ? var : 0.0);
        } else {
          out = (float)(w2[i] > 0.0 ? w2[i] : 0.0);
        }
        if (RSQ) {
          out = out > 0.0f ? __builtin_amdgcn_rsqf(out) : 0.0f;
          // r >= 0, so its bits order as its values; a NaN's bits exceed
          // every number's and stick
          top = max(top, __float_as_uint(out));
        }
        en[p] = out;
      }
    }
    if (RSQ) {
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) top = max(top, (unsigned)__shfl_xor((int)top, d));
      if ((threadIdx.x & 63) == 0) sh_top[k][threadIdx.x >> 6] = top;
    }
  }
  if (RSQ) {
    __syncthreads();
    if (threadIdx.x < n_tabs) {
      const int k = threadIdx.x;
      if (tabs.blk_top[k] && blockIdx.x < tabs.n_blk[k]) {
        unsigned t = sh_top[k][0];
        for (int w = 1; w < kUThreads / 64; w++) t = max(t, sh_top[k][w]);
        tabs.blk_top[k][ch * tabs.n_blk[k] + blockIdx.x] = t;
      }
    }
  }
}

// Per channel, the largest of its blocks' maxima.
__global__ __launch_bounds__(256) void rmax_reduce_kernel(const unsigned *__restrict__ blk_top,
                                                          int n_blk,
                                                          unsigned *__restrict__ rmax_bits) {
  __shared__ unsigned sh[4];
  const unsigned *b = blk_top + (size_t)blockIdx.x * n_blk;
  unsigned top = 0;
  for (int k = threadIdx.x; k < n_blk; k += blockDim.x) top = max(top, b[k]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) top = max(top, (unsigned)__shfl_xor((int)top, d));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = top;
  __syncthreads();
  if (threadIdx.x == 0) rmax_bits[blockIdx.x] = max(max(sh[0], sh[1]), max(sh[2], sh[3]));
}

// Valid output index range of each template: matched_filter.c:53-79 (lags
// i in [start, stop) with step, output index i / step).  min/max moveout
// start at 0 and include weight-zero channels.
__global__ void out_range_kernel(const int *__restrict__ moveouts,
                                 long long n_templates, long long n_channels,
                                 int step, long long n_tp_all, long long n_data,
                                 long long n_corr, const int *__restrict__ t_len,
                                 int *__restrict__ out_range) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_templates) return;
  long long n_tp = t_len ? (long long)t_len[t] : n_tp_all;
  n_tp = n_tp < 1 ? 1 : n_tp;
  int lo = 0, hi = 0;
  for (long long ch = 0; ch < n_channels; ch++) {
    int m = moveouts[t * n_channels + ch];
    lo = m < lo ? m : lo;
    hi = m > hi ? m : hi;
  }
  long long start = lo < 0 ? ((long long)(-lo) + step - 1) / step * step : 0;
  long long stop = 1 + n_data - n_tp - hi;
  long long olo = start / step;
  long long ohi = stop > 0 ? (stop + step - 1) / step : 0;
  if (ohi > n_corr) ohi = n_corr;
  if (olo > ohi) olo = ohi;
  out_range[2 * t] = (int)olo;
  out_range[2 * t + 1] = (int)ohi;
}

// Copy templates into rows of n_pad floats, zero beyond n_tp.  demean = 1
// subtracts the row mean (FP64) first: sum_k T[k](d[k] - m) equals
// sum_k (T[k] - mean(T)) d[k] exactly, so the 'full' numerator needs no
// per-window mean inside the correlation loop.
__global__ __launch_bounds__(kThreads) void prepare_templates_kernel(
    const float *__restrict__ tp, long long n_stride, long long n_pad,
    int demean, const int *__restrict__ t_len, long long rows_per_template,
    float *__restrict__ out) {
  __shared__ double sh[kThreads / 64];
  const long long row = blockIdx.x;
  const float *src = tp + row * n_stride;
  // Batched calls: row length of its template, zero beyond it.
  long long n_tp = t_len ? (long long)t_len[row / rows_per_template] : n_stride;
  n_tp = n_tp < 1 ? 1 : (n_tp > n_stride ? n_stride : n_tp);
  double m = 0.0;
  if (demean) {
    double s = 0.0;
    for (long long k = threadIdx.x; k < n_tp; k += kThreads) s += src[k];
    m = block_sum(s, sh) / (double)n_tp;
  }
  for (long long k = threadIdx.x; k < n_pad; k += kThreads)
    out[row * n_pad + k] = k < n_tp ? (float)((double)src[k] - m) : 0.0f;
}

// Running FP64 sum of squares of one channel in the reference's exact
// operation order (matched_filter.c:680-705: Neumaier compensation, branch on
// the float-converted magnitudes, output = running sum + correction), so the
// prefix array is bit-identical to the reference's.  One wave per channel:
// the wave stages 1024 samples in LDS, lane 0 runs the sequential chain and
// the wave writes the results back coalesced.  Sequential by construction;
// used by the exact-order mode and the cumsum_square_data entry point.
constexpr int kSeqChunk = 1024;

__global__ __launch_bounds__(64) void neumaier_prefix_kernel(
    const float *__restrict__ data, long long n_data, double *__restrict__ P) {
  __shared__ float xs[kSeqChunk];
  __shared__ double ps[kSeqChunk];
  const long long ch = blockIdx.x;
  const float *d = data + ch * n_data;
  double *out = P + ch * (n_data + 1);
  if (threadIdx.x == 0) out[0] = 0.0;
  double acc = 0.0, comp = 0.0;
  for (long long c0 = 0; c0 < n_data; c0 += kSeqChunk) {
    const int n = (int)((n_data - c0) < kSeqChunk ? (n_data - c0) : kSeqChunk);
    for (int i = threadIdx.x; i < n; i += 64) xs[i] = d[c0 + i];
    __syncthreads();
    if (threadIdx.x == 0) {
      int i = 0;
      if (c0 == 0) {
        acc = (double)xs[0] * (double)xs[0];
        ps[0] = acc;
        i = 1;
      }
      for (; i < n; i++) {
        const double sq = (double)xs[i] * (double)xs[i];
        const double nxt = acc + sq;
        if (fabsf((float)acc) >= fabsf((float)sq))
          comp += (acc - nxt) + sq;
        else
          comp += (sq - nxt) + acc;
        acc = nxt;
        ps[i] = nxt + comp;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 64) out[1 + c0 + i] = ps[i];
    __syncthreads();
  }
}

// energy[ch][p] = (float)(P[p+N] - P[p]) from the reference-order prefix
// (matched_filter.c:622).
__global__ void energy_from_prefix_kernel(const double *__restrict__ P,
                                          long long n_data, long long n_tp,
                                          long long n_win,
                                          float *__restrict__ energy) {
  const long long ch = blockIdx.y;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_win) return;
  const double *q = P + ch * (n_data + 1);
  energy[ch * n_win + p] = (float)(q[p + n_tp] - q[p]);
}

// Energy-table offset of each template of a batched call: the table whose
// length equals the template's; -1 when none does.
__global__ void batch_offsets_kernel(const int *__restrict__ t_len,
                                     long long n_templates, BatchTables tab,
                                     long long *__restrict__ t_eoff) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_templates) return;
  long long off = -1;
  for (int d = 0; d < tab.n; d++)
    if (tab.len[d] == t_len[t]) off = tab.offset[d];
  t_eoff[t] = off;
}

// Schedule of a batched call's templates: grouped by length (the index of
// its energy table), then by decreasing number of active channels, so that
// the blocks resident on an XCD at one time — consecutive templates of one
// output tile — read the same energy table and walk their channels at a
// similar pace, and their data windows and energies are still in L2 when the
// next block asks for them.  A counting sort in one block: a histogram of
// the bucket keys in LDS, its exclusive scan, then every template takes the
// next slot of its bucket (the order inside a bucket is whichever thread
// gets there first; results do not depend on it: every row is computed the
// same way whenever it runs).  Buckets: (length index, active channels) when
// they fit in LDS, else the length index alone.
constexpr int kOrderThreads = 1024;
constexpr int kOrderBuckets = 16384;

__global__ __launch_bounds__(kOrderThreads) void batch_order_kernel(
    const int *__restrict__ t_len, const float *__restrict__ weights, long long n_templates,
    long long n_channels, BatchTables tab, int *__restrict__ order) {
  __shared__ unsigned hist[kOrderBuckets];
  __shared__ unsigned part[kOrderThreads];
  const bool by_active = (long long)tab.n * (n_channels + 1) <= kOrderBuckets;
  const int n_buckets = by_active ? tab.n * (int)(n_channels + 1) : tab.n;
  auto key = [&](long long t) {
    int d = 0;
    for (int j = 0; j < tab.n; j++)
      if (tab.len[j] == t_len[t]) d = j;
    if (!by_active) return d;
    int act = 0;
    for (long long ch = 0; ch < n_channels; ch++) act += weights[t * n_channels + ch] != 0.0f;
    return d * (int)(n_channels + 1) + (int)(n_channels - act);
  };
  for (int i = threadIdx.x; i < kOrderBuckets; i += kOrderThreads) hist[i] = 0;
  __syncthreads();
  for (long long t = threadIdx.x; t < n_templates; t += kOrderThreads)
    atomicAdd(&hist[key(t)], 1u);
  __syncthreads();
  // exclusive scan of hist[0, n_buckets): runs of 16 buckets per thread
  constexpr int kRun = kOrderBuckets / kOrderThreads;
  const int b0 = threadIdx.x * kRun;
  unsigned run = 0;
  for (int i = 0; i < kRun; i++) run += b0 + i < n_buckets ? hist[b0 + i] : 0u;
  part[threadIdx.x] = run;
  __syncthreads();
  for (int d = 1; d < kOrderThreads; d <<= 1) {
    const unsigned v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned off = part[threadIdx.x] - run;
  for (int i = 0; i < kRun; i++) {
    if (b0 + i >= n_buckets) break;
    const unsigned c = hist[b0 + i];
    hist[b0 + i] = off;
    off += c;
  }
  __syncthreads();
  for (long long t = threadIdx.x; t < n_templates; t += kOrderThreads)
    order[atomicAdd(&hist[key(t)], 1u)] = (int)t;
}

__global__ void fill_kernel(float *p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void rsqrt_table_kernel(float *e, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const float v = e[i];
    e[i] = v > 0.0f ? __builtin_amdgcn_rsqf(v) : 0.0f;
  }
}

}  // namespace

int launch_rsqrt_table(float *energy, long long n, hipStream_t s) {
  if (n <= 0) return kOk;
  const long long blocks = std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(rsqrt_table_kernel, dim3((unsigned)blocks), dim3(256), 0, s, energy, n);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

// Windows per block of the union kernel, or 0 when the two-segment kernel
// runs (templates of more than 7168 samples, or 32-bit buffer offsets out of
// reach).
long long union_tile(long long n_data, long long n_tp) {
  const long long tile = std::min<long long>(4096, (8192 - n_tp) / 256 * 256);
  return (tile >= 1024 && n_data < (1LL << 29)) ? tile : 0;
}

long long energy_tail_words(long long n_channels, long long n_data, long long n_tp) {
  const long long n_win = n_data - n_tp + 1;
  if (n_win <= 0 || n_channels <= 0) return 0;
  const long long tile = union_tile(n_data, n_tp);
  return n_channels * (1 + (tile ? (n_win + tile - 1) / tile : 0));
}

namespace {

// One launch of the union kernel over the tables tabs.energy[0 .. n) (lengths
// sharing tile and run length), then each table's per-channel maximum.
int launch_union(const float *data, long long n_channels, long long n_data, int full,
                 int rsq, long long tile, EnergyTabs &tabs, unsigned *const *rmax_bits,
                 hipStream_t s) {
  const long long U = tile + tabs.n_tp_max;
  const size_t lds = (size_t)(U + 1) * (full ? 16 : 8);
  const bool multi = tabs.n > 1;
#define FMF_UNION_FN(F, R) \
  (multi ? reinterpret_cast<const void *>(&window_energy_union_kernel<F, R, 1>) \
         : reinterpret_cast<const void *>(&window_energy_union_kernel<F, R, 0>))
  const void *k = full ? (rsq ? FMF_UNION_FN(1, 1) : FMF_UNION_FN(1, 0))
                       : (rsq ? FMF_UNION_FN(0, 1) : FMF_UNION_FN(0, 0));
#undef FMF_UNION_FN
  if (lds > 64 * 1024) {
    // 160 KiB per block minus room for the kernel's static LDS
    const int rc = ensure_lds_limit(k, 159 * 1024);
    if (rc) return rc;
  }
  int n_blk = 0;
  for (int i = 0; i < tabs.n; i++) n_blk = std::max(n_blk, tabs.n_blk[i]);
  dim3 grid((unsigned)n_blk, (unsigned)n_channels);
#define FMF_UNION(F, R)                                                                  \
  do {                                                                                   \
    if (multi)                                                                           \
      hipLaunchKernelGGL((window_energy_union_kernel<F, R, 1>), grid, dim3(kUThreads), lds, \
                         s, data, n_data, (int)tile, tabs);                              \
    else                                                                                 \
      hipLaunchKernelGGL((window_energy_union_kernel<F, R, 0>), grid, dim3(kUThreads), lds, \
                         s, data, n_data, (int)tile, tabs);                              \
  } while (0)
  if (full && rsq) FMF_UNION(1, 1);
  else if (full) FMF_UNION(1, 0);
  else if (rsq) FMF_UNION(0, 1);
  else FMF_UNION(0, 0);
#undef FMF_UNION
  FMF_HIP_CHECK(hipGetLastError());
  for (int i = 0; i < tabs.n && rsq; i++) {
    hipLaunchKernelGGL(rmax_reduce_kernel, dim3((unsigned)n_channels), dim3(256), 0, s,
                       tabs.blk_top[i], tabs.n_blk[i], rmax_bits[i]);
    FMF_HIP_CHECK(hipGetLastError());
  }
  return kOk;
}

// A table that no union launch can hold (template longer than the union
// allows, or data of 2^29 samples and more): the two-segment kernel.
int launch_segments(const float *data, long long n_channels, long long n_data,
                    long long n_tp, int full, float *energy, unsigned *rmax_bits, int rsq,
                    hipStream_t s) {
  const long long n_win = n_data - n_tp + 1;
  dim3 grid((unsigned)((n_win + kTile - 1) / kTile), (unsigned)n_channels);
  hipLaunchKernelGGL(window_energy_kernel, grid, dim3(kThreads), 0, s, data,
                     n_data, n_tp, n_win, full, energy);
  FMF_HIP_CHECK(hipGetLastError());
  if (rsq) {
    // no maxima from this kernel: NaN sends every channel to the full test
    const int rc = launch_fill(reinterpret_cast<float *>(rmax_bits), (size_t)n_channels,
                               __builtin_nanf(""), s);
    if (rc) return rc;
    return launch_rsqrt_table(energy, n_channels * n_win, s);
  }
  return kOk;
}

}  // namespace

int launch_window_energies(const float *data, long long n_channels, long long n_data,
                           const long long *n_tp, float *const *energy, int n_tab, int full,
                           hipStream_t s, int rsq) {
  if (n_channels <= 0) return kOk;
  // rsq: behind each table, per channel, the largest 1/sqrt(e) of its windows
  // (as float bits; a NaN where a window's energy is NaN or not computed
  // here), then the per-block maxima it is reduced from
  EnergyTabs group{};
  unsigned *group_rmax[kTabsMax];
  auto flush = [&]() -> int {
    if (!group.n) return kOk;
    group.per = kFixedRun;
    const int rc = launch_union(data, n_channels, n_data, full, rsq, kUnionTile, group,
                                group_rmax, s);
    group = EnergyTabs{};
    return rc;
  };
  for (int i = 0; i < n_tab; i++) {
    const long long n_win = n_data - n_tp[i] + 1;
    if (n_win <= 0) continue;
    unsigned *rmax_bits =
        rsq ? reinterpret_cast<unsigned *>(energy[i] + n_channels * n_win) : nullptr;
    const long long tile = union_tile(n_data, n_tp[i]);
    const int n_blk = tile ? (int)((n_win + tile - 1) / tile) : 0;
    if (tile == kUnionTile && n_tp[i] <= kFixedRunMaxN) {
      // lengths of the fixed run share launches, kTabsMax at a time
      const int g = group.n++;
      group.n_tp_max = std::max(group.n_tp_max, (int)n_tp[i]);
      group.n_tp[g] = (int)n_tp[i];
      group.n_blk[g] = n_blk;
      group.n_win[g] = n_win;
      group.energy[g] = energy[i];
      group.blk_top[g] = rsq ? rmax_bits + n_channels : nullptr;
      group_rmax[g] = rmax_bits;
      if (group.n == kTabsMax) {
        const int rc = flush();
        if (rc) return rc;
      }
    } else if (tile) {
      // a longer template: a launch of its own, its run length from its union
      EnergyTabs one{};
      one.n = 1;
      one.n_tp_max = (int)n_tp[i];
      one.per = std::min((int)(((tile + n_tp[i] + kUThreads - 1) / kUThreads) | 1), kUMaxPer);
      one.n_tp[0] = (int)n_tp[i];
      one.n_blk[0] = n_blk;
      one.n_win[0] = n_win;
      one.energy[0] = energy[i];
      one.blk_top[0] = rsq ? rmax_bits + n_channels : nullptr;
      const int rc = launch_union(data, n_channels, n_data, full, rsq, tile, one, &rmax_bits, s);
      if (rc) return rc;
    } else {
      const int rc = launch_segments(data, n_channels, n_data, n_tp[i], full, energy[i],
                                     rmax_bits, rsq, s);
      if (rc) return rc;
    }
  }
  return flush();
}

int launch_window_energy(const float *data, long long n_channels,
                         long long n_data, long long n_tp, int full,
                         float *energy, hipStream_t s, int rsq) {
  return launch_window_energies(data, n_channels, n_data, &n_tp, &energy, 1, full, s, rsq);
}

int launch_out_ranges(const int *moveouts, long long n_templates,
                      long long n_channels, int step, long long n_tp,
                      long long n_data, long long n_corr, int *out_range,
                      hipStream_t s, const int *t_len) {
  if (n_templates <= 0) return kOk;
  hipLaunchKernelGGL(out_range_kernel, dim3((unsigned)((n_templates + 255) / 256)),
                     dim3(256), 0, s, moveouts, n_templates, n_channels, step,
                     n_tp, n_data, n_corr, t_len, out_range);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_prepare_templates(const float *templates, long long n_rows,
                             long long n_tp, long long n_pad, int demean,
                             float *out, hipStream_t s, const int *t_len,
                             long long rows_per_template) {
  if (n_rows <= 0) return kOk;
  hipLaunchKernelGGL(prepare_templates_kernel, dim3((unsigned)n_rows),
                     dim3(kThreads), 0, s, templates, n_tp, n_pad, demean,
                     t_len, rows_per_template > 0 ? rows_per_template : 1, out);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_cumsum_square(const float *data, long long n_channels,
                         long long n_data, double *P, hipStream_t s) {
  if (n_channels <= 0 || n_data <= 0) return kOk;
  hipLaunchKernelGGL(neumaier_prefix_kernel, dim3((unsigned)n_channels),
                     dim3(64), 0, s, data, n_data, P);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_energy_from_prefix(const double *P, long long n_channels,
                              long long n_data, long long n_tp, float *energy,
                              hipStream_t s) {
  const long long n_win = n_data - n_tp + 1;
  if (n_win <= 0 || n_channels <= 0) return kOk;
  dim3 grid((unsigned)((n_win + 255) / 256), (unsigned)n_channels);
  hipLaunchKernelGGL(energy_from_prefix_kernel, grid, dim3(256), 0, s, P,
                     n_data, n_tp, n_win, energy);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_batch_offsets(const int *t_len, long long n_templates,
                         const BatchTables &tab, long long *t_eoff,
                         hipStream_t s) {
  if (n_templates <= 0) return kOk;
  hipLaunchKernelGGL(batch_offsets_kernel, dim3((unsigned)((n_templates + 255) / 256)),
                     dim3(256), 0, s, t_len, n_templates, tab, t_eoff);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_batch_order(const int *t_len, const float *weights, long long n_templates,
                       long long n_channels, const BatchTables &tab, int *order, hipStream_t s) {
  if (n_templates <= 0) return kOk;
  if (n_templates > 0x7fffffffLL) return set_error(kInvalidArgument, "too many templates");
  hipLaunchKernelGGL(batch_order_kernel, dim3(1), dim3(kOrderThreads), 0, s, t_len, weights,
                     n_templates, n_channels, tab, order);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

int launch_fill(float *p, size_t n, float v, hipStream_t s) {
  if (n == 0) return kOk;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  FMF_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace fmf
