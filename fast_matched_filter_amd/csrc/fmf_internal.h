// Internal declarations shared by the HIP translation units of libfmf_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fmf {

// Status codes returned by every extern "C" entry point (0 = success).
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kOutOfMemory = 2,
  kDeviceError = 3,
  kUnsupported = 4,
};

// Records a message for fmf_last_error() and returns the code.
int set_error(int code, const char *fmt, ...);

// Raises a kernel's dynamic LDS limit to `bytes` on the current device, once
// per (kernel, device); safe to call from several threads.
int ensure_lds_limit(const void *fn, int bytes);

#define FMF_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::fmf::set_error(e_ == hipErrorOutOfMemory ? ::fmf::kOutOfMemory \
                                                        : ::fmf::kDeviceError,\
                              "%s failed: %s (%s:%d)", #expr,                 \
                              hipGetErrorString(e_), __FILE__, __LINE__);     \
  } while (0)

// Threshold rule of the reference paths:
//  kThrBeforeSqrt — cc = 0 unless ssT*ssd > 1e-6 (matched_filter.c:622-625,
//                   matched_filter.cu:132-138; archs 'cpu' and 'gpu')
//  kThrAfterSqrt  — cc = 0 unless sqrt(ssT*ssd) > 1e-6 (matched_filter.c:650-655;
//                   arch 'precise')
enum ThresholdMode : int { kThrBeforeSqrt = 0, kThrAfterSqrt = 1 };

constexpr float kStabilityThreshold = 0.000001f;

// Raw buffer resource over `bytes` bytes at `base` (gfx9 descriptor word 3:
// 32-bit float data, no swizzle).  Loads at byte offsets at or past `bytes`
// (negative offsets included: they wrap to large unsigned values) return 0
// and stores there are dropped, so guarded accesses need no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base,
                                                              unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                           (int)bytes, 0x00020000);
}

// Offset that is out of range of every resource built with buffer_rsrc
// (record counts stay below 2^31 bytes) and stays so when small constants
// are added to it.
constexpr int kOob = (int)0x80000000u;

__device__ __forceinline__ float buffer_load_f32(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

__device__ __forceinline__ _Float16 buffer_load_f16(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
}

__device__ __forceinline__ void buffer_store_f32(__amdgpu_buffer_rsrc_t r, int off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, 0);
}

// Channels of one template with a non-zero weight (matched_filter.c:436),
// found 64 at a time: every lane of a wave reads one weight and a ballot
// gives the window's mask, so skipping weight-0 channels costs one load per
// 64 channels instead of one dependent load per channel.  Every lane of the
// wave calls next() with the same argument; results are wave-uniform.
struct ActiveChannels {
  const float *w;
  int n, base;
  unsigned long long mask;
  __device__ __forceinline__ ActiveChannels(const float *w_t, int n_channels)
      : w(w_t), n(n_channels), base(-64), mask(0) {}
  // The first active channel after ch, or n_channels.
  __device__ __forceinline__ int next(int ch) {
    int c = ch + 1;
    while (c < n) {
      if (c < base || c >= base + 64) {
        base = c & ~63;
        const int i = base + (int)(threadIdx.x & 63);
        const float v = i < n ? w[i] : 0.0f;
        mask = __builtin_amdgcn_ballot_w64(v != 0.0f);
      }
      const unsigned long long m = mask >> (c - base);
      if (m) return c + __builtin_ctzll(m);
      c = base + 64;
    }
    return n;
  }
};

// Everything the correlation kernels read, all device pointers.
struct CorrArgs {
  const float *templates;   // (T, SC, n_pad) prepared templates, zero beyond N
  const float *sst;         // (T, SC) sum of squares of the caller's templates
  const int *moveouts;      // (T, SC)
  const float *weights;     // (T, SC)
  const float *data;        // (SC, L)
  const float *energy;      // (SC, n_win) window energies, n_win = L - N + 1
  const int *out_range;     // (T, 2) valid output index range [lo, hi)
  float *out;               // (T, n_corr) or (T, n_corr, SC)
  long long n_templates, n_channels, n_tp, n_pad, n_data, n_win, n_corr;
  int step;
  int threshold_mode;
  int summed;               // 1: weighted network sum, 0: per-channel cc
  // Batched calls (MFMA kernel): per-template lengths (n_tp is then the
  // maximum) and per-template element offsets of their window-energy table
  // inside `energy`; nullptr for one length per call.
  const int *t_len = nullptr;
  const long long *t_eoff = nullptr;
  // Batched calls (32x32 kernels): template of the i-th template slot of a
  // tile (a permutation of [0, T)); nullptr: slot i is template i.
  const int *t_order = nullptr;
  // > 0: persistent launch with about this many blocks.
  long long persistent_blocks = 0;
  // Optional list of work items (tile * n_templates + template, ascending):
  // only these are computed and written (screened calls).
  const int *items = nullptr;
  long long n_items = 0;
  // 1: screening variant of the MFMA kernel (one bf16 product per MAC, no
  // threshold).
  int screen = 0;
  // 32x32 MFMA kernels: operands split into bf16 hi/lo once per call
  // (launch_split_planes, launch_template_images); `energy` then holds
  // 1/sqrt(e), 0 where e = 0 (launch_rsqrt_table).
  const uint32_t *planes = nullptr;   // (SC, 4, plane_pairs)
  long long plane_pairs = 0;
  const uint16_t *images = nullptr;   // (T, SC, img_stride)
  long long img_stride = 0;
};

// Kernel launchers (stream-ordered, no synchronisation).
int launch_out_ranges(const int *moveouts, long long n_templates,
                      long long n_channels, int step, long long n_tp,
                      long long n_data, long long n_corr, int *out_range,
                      hipStream_t s, const int *t_len = nullptr);
// rsq = 1: the table holds 1/sqrt(energy), 0 for a zero energy.
// rsq: the table holds 1/sqrt(e) (0 where e = 0), followed by n_channels
// words, per channel the largest of those values as float bits (NaN when a
// window's energy is NaN, or when the long-template kernel computed it), and
// the per-block maxima they are reduced from: the table needs
// n_channels * n_win + energy_tail_words(...) floats.
int launch_window_energy(const float *data, long long n_channels,
                         long long n_data, long long n_tp, int full,
                         float *energy, hipStream_t s, int rsq = 0);
// The tables of n_tab template lengths over the same data (table i of
// length n_tp[i] at energy[i], laid out as above): lengths of up to 1280
// samples share launches, 8 tables per launch reading the data once, and
// each table holds the same bits as from launch_window_energy alone.
int launch_window_energies(const float *data, long long n_channels, long long n_data,
                           const long long *n_tp, float *const *energy, int n_tab, int full,
                           hipStream_t s, int rsq = 0);
long long energy_tail_words(long long n_channels, long long n_data, long long n_tp);
long long union_tile(long long n_data, long long n_tp);
// energy[i] = energy[i] > 0 ? 1/sqrt(energy[i]) : 0 (screening kernel).
int launch_rsqrt_table(float *energy, long long n, hipStream_t s);
int launch_prepare_templates(const float *templates, long long n_rows,
                             long long n_tp, long long n_pad, int demean,
                             float *out, hipStream_t s,
                             const int *t_len = nullptr,
                             long long rows_per_template = 1);

// Distinct template lengths of a batched call and the element offsets of
// their window-energy tables in the workspace (passed by value to kernels).
constexpr int kMaxBatchLengths = 16;
struct BatchTables {
  int n;
  int len[kMaxBatchLengths];
  long long offset[kMaxBatchLengths];
};
int launch_batch_offsets(const int *t_len, long long n_templates,
                         const BatchTables &tab, long long *t_eoff,
                         hipStream_t s);
// Permutation of a batched call's templates grouped by length and active
// channels (the order of their work items within an output tile).
int launch_batch_order(const int *t_len, const float *weights, long long n_templates,
                       long long n_channels, const BatchTables &tab, int *order, hipStream_t s);
int launch_corr_fp32(const CorrArgs &a, int exact_order, hipStream_t s);
// Returns kUnsupported without launching when the MFMA tiling does not cover
// the step or template length.
// shape = 10 * tiles_per_wave + waves_per_simd (24, 42, 43, 81) for the
// 16x16x32 kernel, 100 * waves_per_block + 10 * tiles_per_wave +
// waves_per_simd (121, 141, 212, 412, 413, 421, 422, 241) for the 32x32x16 kernel.
int launch_corr_mfma(const CorrArgs &a, int shape, hipStream_t s);
int launch_corr_mfma32(const CorrArgs &a, int shape, hipStream_t s);
bool mfma16_shape(int shape);
bool mfma32_shape(int shape);
// Sizes the 32x32 MFMA kernels take (step, template length, data length,
// outputs, channels); the 16x16 kernels take the same steps and
// N <= 1280 - 15 * step.
bool mfma32_covers(long long step, long long n_tp, long long n_data, long long n_corr,
                   long long n_channels);
bool mfma16_covers(long long step, long long n_tp, long long n_data, long long n_corr,
                   long long n_channels, int summed);
// Split operands of the 32x32 kernels: dwords per data plane (four planes
// per channel) and bf16 elements per template-image row.
long long mfma32_plane_pairs(long long n_data);
long long mfma32_image_stride(long long n_tp_max, int step);
int launch_split_planes(const float *data, long long n_channels, long long n_data,
                        long long pairs, uint32_t *planes, hipStream_t s);
int launch_template_images(const float *prepared, long long n_rows, long long n_pad,
                           long long n_tp, int step, const int *t_len,
                           long long rows_per_template, long long stride, uint16_t *img,
                           hipStream_t s);
// Detection post-processing (fmf_detect.hip).  ranges: (T, 2) valid lag
// range per template, or nullptr for [0, n_corr).
size_t detect_workspace_bytes(long long T, long long n_corr);
int launch_median_mad(const float *cc, long long T, long long n_corr,
                      const int *ranges, float *med, float *mad, void *ws,
                      size_t ws_bytes, hipStream_t s);
int launch_detect(const float *cc, long long T, long long n_corr, const int *ranges,
                  const float *med, const float *mad, float n_mad, long long sep,
                  int *det_t, int *det_lag, float *det_val, long long capacity,
                  long long *n_det, void *ws, size_t ws_bytes, hipStream_t s);
// Outputs per work item (output tile) of an MFMA tile shape.
int mfma_block_outputs(int shape);
int mfma32_block_outputs(int shape);
// Variable-length Fisher-combined path (arch 'variable_precise').
struct VarprecArgs {
  const float *templates;   // (T, SC, n_max)
  const float *sst;         // (T, SC)
  const int *moveouts;      // (T, SC)
  const float *weights;     // (T, SC)
  const int *lengths;       // (T, SC)
  const float *cn;          // (T, SC) Fisher weights w / sqrt(1 / (n - 3))
  const float *tscal;       // (T, 3) sum of cn, used*eps*10, running weight sum
  const int *trange;        // (T, 2) lag range [start, stop)
  const float *data;        // (SC, L)
  float *out;               // (T, n_corr)
  int n_templates, n_channels, n_max, n_data, n_corr, step, normalize;
};
int launch_corr_varprec(const VarprecArgs &a, hipStream_t s);
int launch_fill(float *p, size_t n, float v, hipStream_t s);

// Device -> pageable host copy after the work queued on `s` (fmf_host_io.hip):
// the destination's pages are faulted in by host threads while the GPU
// works, then the data streams through pinned staging buffers.  Returns when
// the copy is complete.
int copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s);

// Screener kernels (fmf_screener.hip).
int launch_envelope(const float *data, int n_channels, int n_data, int n_dec,
                    float *env, hipStream_t s);
int launch_box_mean(const float *env, int n_channels, int n_dec, int width,
                    float *out, hipStream_t s);
int launch_features(const float *win, const float *winh, const float *env,
                    const float *chinv, const int *moveouts,
                    const float *weights, int n_templates, int n_channels,
                    int n_dec, int n_out, int nwh, float len_feature,
                    float *feat, hipStream_t s);
int launch_conv_fwd(const float *in, int nb, int cin, int nx, int relu_in,
                    const float *w, const float *bias, int cout, float *out,
                    hipStream_t s);
int launch_conv_fwd_fast(const float *in, int nb, int cin, int nx, int relu_in,
                         const float *w, const float *bias, int cout, float *out,
                         void *wf_scratch, hipStream_t s);
size_t conv_fwd_fast_scratch_bytes(int cin);
size_t screen_net_scratch_bytes();
template <typename InT>
int launch_screen_net(const InT *feat, int nb, int nx, const float *w1, const float *b1,
                      const float *w2, const float *b2, const float *w3, const float *b3,
                      const float *w4, const float *b4, void *scratch, float *pred,
                      hipStream_t s);
int launch_decimate(const float *in, long long n_rows, long long n_in, int D,
                    long long n_out, float *out, hipStream_t s);
template <typename OutT>
int launch_coarse_features(const float *cc, int n_templates, int n_coarse, const float *mad,
                           int D, int n_out, OutT *feat, hipStream_t s);
int launch_conv_bwd(const float *gz, const float *ain, const float *zin,
                    int nb, int cin, int nx, int relu_in, const float *w,
                    int cout, float *ga, float *gw, float *gb, hipStream_t s);
int launch_loss(const float *pred, const float *target, long long n,
                float boost, float inv_norm, float *grad, float *partial,
                int blocks, hipStream_t s);
int launch_adam(float *master, const float *grad, float *m, float *v, int n,
                float lr, float b1, float b2, float eps, int t, void *w16,
                float *w32, hipStream_t s);
int launch_gate(const float *pred, int n_out, const float *thr, int n_tiles,
                int tile_dec, int margin, int n_templates, int *keep,
                hipStream_t s);
int launch_peak_target(const float *cc, int n_corr, int n_out,
                       int n_templates, float *target, hipStream_t s);
int launch_cumsum_square(const float *data, long long n_channels,
                         long long n_data, double *P, hipStream_t s);
int launch_energy_from_prefix(const double *P, long long n_channels,
                              long long n_data, long long n_tp, float *energy,
                              hipStream_t s);

}  // namespace fmf
