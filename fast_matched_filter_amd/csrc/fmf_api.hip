// C ABI of libfmf_hip: host driver and device-resident entry points.
//
// The reference GPU host driver (fast_matched_filter/src/matched_filter.cu:
// 171-399) replicates all inputs on every device, launches two kernels per
// template and per 1/20 of the lags with a device synchronisation and a
// device-to-host copy after each, and exit()s on errors.  This driver keeps
// one stream per call, uploads the inputs once, computes the window energies
// once for all templates, runs one launch per block of templates that fits
// the device's memory, and returns a status code.
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "fmf_internal.h"
#include "../../include/fmf_hip.h"

namespace fmf {

static thread_local char g_err[512] = "";
// Kernel that ran in the calling thread's last matched-filter call (-1: none).
static thread_local int g_last_kernel = -1;

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

namespace {

// Optional per-kernel timing: while enabled, device calls bracket their
// window-energy and correlation launches with HIP events on the call's
// stream; fmf_kernel_timer_read() waits for and sums them.
struct TimedLaunch {
  int which;
  hipEvent_t start, stop;
};
std::mutex g_timer_mu;
bool g_timer_on = false;
std::vector<TimedLaunch> g_timed;

struct ScopedTimer {
  bool on = false;
  TimedLaunch tl{};
  ScopedTimer(int which, hipStream_t s) {
    {
      std::lock_guard<std::mutex> lk(g_timer_mu);
      on = g_timer_on;
    }
    if (!on) return;
    tl.which = which;
    if (hipEventCreate(&tl.start) != hipSuccess || hipEventCreate(&tl.stop) != hipSuccess) {
      on = false;
      return;
    }
    (void)hipEventRecord(tl.start, s);
  }
  void done(hipStream_t s) {
    if (!on) return;
    (void)hipEventRecord(tl.stop, s);
    std::lock_guard<std::mutex> lk(g_timer_mu);
    g_timed.push_back(tl);
    on = false;
  }
};

constexpr int kTapBlock = 16;  // template rows are padded to a multiple of this

// Tile shape of the MFMA kernel per step (1, 2), set by the autotuner through
// fmf_set_mfma_tiles.  Defaults (32x32x16 tiles, 4 waves per block): step 1
// two interleaved tiles per wave (data fragments shared by both tiles) and 2
// waves per SIMD (1422); step 2, whose data windows are twice as long, one
// tile per wave and 3 waves per SIMD (413) — the fastest on every measured
// workload of each step.  A call reads its step's shape once, when it starts.
std::atomic<int> g_mfma_tiles[3] = {{1422}, {1422}, {413}};

int mfma_shape_of(int step) { return g_mfma_tiles[step >= 1 && step <= 2 ? step : 0].load(); }

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Sizes {
  long long T, S, C, SC, N, L, n_corr, n_win, n_pad;
  int step, normalize, summed, kernel, flags;
};

// The kernel family that runs a call: the FP32 vector kernel, the 16x16x32
// MFMA kernel (e energy table) or the 32x32x16 MFMA kernel (split operands,
// 1/sqrt(e) energy table).
enum RunKind : int { kRunFP32 = 0, kRunMfma16 = 1, kRunMfma32 = 2 };

// Whether calls of these sizes on an MFMA kernel carry split operands in the
// workspace: the 32x32 kernels cover the template length at step 1 (step 2
// covers fewer lengths).
bool split_operands(const Sizes &z) {
  return z.kernel != FMF_KERNEL_FP32 && mfma32_covers(1, z.N, z.L, 1, z.SC);
}

struct WsLayout {
  size_t energy, templates, range, prefix, planes, images, total;
  long long plane_pairs, img_stride;
};

WsLayout layout(const Sizes &z, long long t_chunk) {
  WsLayout w{};
  size_t off = 0;
  const long long n_win = z.n_win > 0 ? z.n_win : 0;
  w.energy = off;  // the table, then per channel its largest 1/sqrt(e)
  off = align_up(off + ((size_t)z.SC * n_win + energy_tail_words(z.SC, z.L, z.N)) *
                           sizeof(float), 256);
  w.templates = off;
  off = align_up(off + (size_t)t_chunk * z.SC * z.n_pad * sizeof(float), 256);
  w.range = off;
  off = align_up(off + (size_t)t_chunk * 2 * sizeof(int), 256);
  w.prefix = off;
  if (z.flags & FMF_FLAG_EXACT_ORDER)
    off = align_up(off + (size_t)z.SC * (z.L + 1) * sizeof(double), 256);
  w.planes = w.images = off;
  if (split_operands(z)) {
    w.plane_pairs = mfma32_plane_pairs(z.L);
    w.img_stride = std::max(mfma32_image_stride(z.N, 1), mfma32_image_stride(z.N, 2));
    off = align_up(off + (size_t)z.SC * 4 * w.plane_pairs * sizeof(uint32_t), 256);
    w.images = off;
    off = align_up(off + (size_t)t_chunk * z.SC * w.img_stride * sizeof(uint16_t), 256);
  }
  w.total = off;
  return w;
}

// Kernel family for a call on tile shape `shape`, or -1 with an error set.
// The bf16x3 kernel falls back to the FP32 kernel for sizes the MFMA tiling
// does not take, unless FMF_FLAG_STRICT_KERNEL is set or work items are
// listed; the screening kernel has no fallback.
int run_kind(const Sizes &z, int shape, bool listed) {
  if (z.kernel == FMF_KERNEL_FP32) return kRunFP32;
  if (z.kernel == FMF_KERNEL_BF16 || mfma32_shape(shape)) {
    if (mfma32_covers(z.step, z.N, z.L, z.n_corr, z.SC) && split_operands(z)) return kRunMfma32;
  } else if (mfma16_covers(z.step, z.N, z.L, z.n_corr, z.SC, z.summed)) {
    return kRunMfma16;
  }
  if (z.kernel == FMF_KERNEL_BF16)
    return set_error(kUnsupported, "the screening kernel needs step 1 or 2 and "
                                   "N <= 1344 - 31*step"), -1;
  const char *limit = mfma32_shape(shape) ? "N <= 1344 - 31*step" : "N <= 1280 - 15*step";
  if (listed)
    return set_error(kUnsupported, "item lists need step 1 or 2 and %s", limit), -1;
  if (z.flags & FMF_FLAG_STRICT_KERNEL)
    return set_error(kUnsupported,
                     "the bf16x3 MFMA kernel (tile shape %d) does not take this call (step %d, "
                     "N %lld, L %lld, %lld channels; it needs step 1 or 2, %s and L < 2^29) "
                     "and FMF_FLAG_STRICT_KERNEL forbids the FP32 kernel",
                     shape, z.step, z.N, z.L, z.SC, limit), -1;
  return kRunFP32;
}

// Element counts of every array a call touches (templates, data, outputs,
// workspace tables) stay below 2^50, so that no size or byte count computed
// from them overflows 64 bits: channels below 2^31, templates x channels
// below 2^40.
bool sizes_fit(size_t T, size_t S, size_t C, size_t N, size_t L, size_t n_corr,
               int summed) {
  const unsigned long long k50 = 1ULL << 50;
  unsigned long long sc, tsc, tscn, scl, out;
  if (__builtin_mul_overflow((unsigned long long)S, (unsigned long long)C, &sc) ||
      sc >= (1ULL << 31))
    return false;
  if (__builtin_mul_overflow((unsigned long long)T, sc, &tsc) || tsc >= (1ULL << 40))
    return false;
  if (__builtin_mul_overflow(tsc, (unsigned long long)N + kTapBlock, &tscn) || tscn >= k50)
    return false;
  if (__builtin_mul_overflow(sc, (unsigned long long)L + 1, &scl) || scl >= k50) return false;
  if (__builtin_mul_overflow((unsigned long long)T, (unsigned long long)n_corr, &out) ||
      __builtin_mul_overflow(out, summed ? 1ULL : sc, &out) || out >= k50)
    return false;
  return true;
}

// Fault injection for tests of the error paths (fmf_debug_inject): the nth
// following event of each kind in the calling thread fails; 0 = off.
thread_local int g_inject[3] = {0, 0, 0};

bool injected(int point) {
  if (g_inject[point] <= 0) return false;
  return --g_inject[point] == 0;
}

int create_stream(hipStream_t *s) {
  if (injected(FMF_INJECT_STREAM))
    return set_error(kDeviceError, "hipStreamCreateWithFlags failed: injected failure");
  FMF_HIP_CHECK(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
  return kOk;
}

int validate(size_t step, size_t N, size_t L, size_t T, size_t S, size_t C,
             size_t n_corr, int normalize, int sum_cc_mode, int kernel,
             int flags, Sizes &z) {
  g_last_kernel = -1;
  if (step < 1) return set_error(kInvalidArgument, "step must be >= 1");
  if (N < 1) return set_error(kInvalidArgument, "n_samples_template must be >= 1");
  if (T < 1 || S < 1 || C < 1)
    return set_error(kInvalidArgument, "n_templates, n_stations, n_components must be >= 1");
  if (normalize != 0 && normalize != 1)
    return set_error(kInvalidArgument, "normalize must be 0 (short) or 1 (full)");
  if (sum_cc_mode != 0 && sum_cc_mode != 1)
    return set_error(kInvalidArgument, "sum_cc_mode must be 0 or 1");
  if (kernel != FMF_KERNEL_FP32 && kernel != FMF_KERNEL_BF16X3 && kernel != FMF_KERNEL_BF16)
    return set_error(kInvalidArgument, "unknown kernel %d", kernel);
  if (flags & ~(FMF_FLAG_EXACT_ORDER | FMF_FLAG_THRESHOLD_AFTER_SQRT | FMF_FLAG_STRICT_KERNEL))
    return set_error(kInvalidArgument, "unknown flags 0x%x", flags);
  if ((flags & FMF_FLAG_EXACT_ORDER) && (kernel != FMF_KERNEL_FP32 || normalize != 0))
    return set_error(kUnsupported, "exact-order mode needs the FP32 kernel and short normalization");
  // The screening kernel rounds the data, not the demeaned data, to bf16: with
  // 'full' a DC offset would scale its error by sqrt(sum d^2 / sum (d - m)^2).
  if (kernel == FMF_KERNEL_BF16 && normalize != 0)
    return set_error(kUnsupported, "the screening kernel supports short normalization only");
  if (N > ((size_t)1 << 30))
    return set_error(kInvalidArgument, "n_samples_template must stay below 2^30");
  if (n_corr >= (size_t)0x7fffffff || L >= (size_t)0x7fffffff - 2 * N)
    return set_error(kInvalidArgument, "n_corr and n_samples_data must stay below 2^31");
  if (step > (size_t)0x40000000)
    return set_error(kInvalidArgument, "step too large");
  if (!sizes_fit(T, S, C, N, L, n_corr, sum_cc_mode))
    return set_error(kInvalidArgument,
                     "sizes too large: T=%zu S=%zu C=%zu N=%zu L=%zu n_corr=%zu give arrays "
                     "of 2^50 elements or more", T, S, C, N, L, n_corr);
  z.T = (long long)T;
  z.S = (long long)S;
  z.C = (long long)C;
  z.SC = z.S * z.C;
  z.N = (long long)N;
  z.L = (long long)L;
  z.n_corr = (long long)n_corr;
  z.n_win = z.L - z.N + 1;
  z.n_pad = (long long)align_up(N, kTapBlock);
  z.step = (int)step;
  z.normalize = normalize;
  z.summed = sum_cc_mode;
  z.kernel = kernel;
  z.flags = flags;
  return kOk;
}

struct DevIn {
  const float *templates, *sst, *data, *weights;
  const int *moveouts;
};

// Template-independent passes: the window energies of every channel (their
// reciprocal square roots for the 32x32 MFMA kernels) and, for those
// kernels, the data's split planes.
int compute_energy(const Sizes &z, const DevIn &in, char *ws, const WsLayout &w,
                   int run, hipStream_t s) {
  float *energy = reinterpret_cast<float *>(ws + w.energy);
  ScopedTimer timer(1, s);
  int rc;
  if (z.flags & FMF_FLAG_EXACT_ORDER) {
    double *P = reinterpret_cast<double *>(ws + w.prefix);
    rc = launch_cumsum_square(in.data, z.SC, z.L, P, s);
    if (!rc) rc = launch_energy_from_prefix(P, z.SC, z.L, z.N, energy, s);
  } else {
    rc = launch_window_energy(in.data, z.SC, z.L, z.N, z.normalize, energy, s,
                              run == kRunMfma32);
  }
  timer.done(s);
  if (!rc && run == kRunMfma32)
    rc = launch_split_planes(in.data, z.SC, z.L, w.plane_pairs,
                             reinterpret_cast<uint32_t *>(ws + w.planes), s);
  return rc;
}

// Correlate templates [t0, t0 + tc) into out (rows of that block only) on
// the kernel family `run`.  items/n_items (MFMA kernels only): compute only
// the listed work items.  shape: MFMA tile shape.
int correlate(const Sizes &z, const DevIn &in, long long t0, long long tc,
              float *out, char *ws, const WsLayout &w, int run, int shape,
              hipStream_t s, const int *items = nullptr, long long n_items = 0) {
  const long long SC = z.SC;
  float *tp = reinterpret_cast<float *>(ws + w.templates);
  int *range = reinterpret_cast<int *>(ws + w.range);
  int rc = launch_prepare_templates(in.templates + t0 * SC * z.N, tc * SC, z.N,
                                    z.n_pad, z.normalize, tp, s);
  if (rc) return rc;
  uint16_t *img = reinterpret_cast<uint16_t *>(ws + w.images);
  if (run == kRunMfma32 &&
      (rc = launch_template_images(tp, tc * SC, z.n_pad, z.N, z.step, nullptr, 1,
                                   w.img_stride, img, s)))
    return rc;
  rc = launch_out_ranges(in.moveouts + t0 * SC, tc, SC, z.step, z.N, z.L,
                         z.n_corr, range, s);
  if (rc) return rc;
  if (!z.summed && !items) {
    rc = launch_fill(out, (size_t)tc * z.n_corr * SC, 0.0f, s);
    if (rc) return rc;
  }
  if (z.n_win <= 0) {
    if (z.summed) return launch_fill(out, (size_t)tc * z.n_corr, 0.0f, s);
    return kOk;
  }
  CorrArgs a{};
  a.templates = tp;
  a.sst = in.sst + t0 * SC;
  a.moveouts = in.moveouts + t0 * SC;
  a.weights = in.weights + t0 * SC;
  a.data = in.data;
  a.energy = reinterpret_cast<const float *>(ws + w.energy);
  a.out_range = range;
  a.out = out;
  a.n_templates = tc;
  a.n_channels = SC;
  a.n_tp = z.N;
  a.n_pad = z.n_pad;
  a.n_data = z.L;
  a.n_win = z.n_win;
  a.n_corr = z.n_corr;
  a.step = z.step;
  a.threshold_mode = (z.flags & FMF_FLAG_THRESHOLD_AFTER_SQRT) ? kThrAfterSqrt : kThrBeforeSqrt;
  a.summed = z.summed;
  a.items = items;
  a.n_items = n_items;
  a.screen = z.kernel == FMF_KERNEL_BF16;
  a.planes = reinterpret_cast<const uint32_t *>(ws + w.planes);
  a.plane_pairs = w.plane_pairs;
  a.images = img;
  a.img_stride = w.img_stride;
  if (items && run == kRunFP32)
    return set_error(kUnsupported, "item lists run on the MFMA kernels");
  ScopedTimer timer(0, s);
  if (run == kRunFP32)
    rc = launch_corr_fp32(a, (z.flags & FMF_FLAG_EXACT_ORDER) ? 1 : 0, s);
  else
    rc = launch_corr_mfma(a, shape, s);
  if (rc == kUnsupported)  // the family was chosen by run_kind: not expected
    rc = set_error(kUnsupported, "the %s kernel refused a call run_kind gave it",
                   run == kRunFP32 ? "FP32" : "MFMA");
  timer.done(s);
  if (!rc)
    g_last_kernel = run == kRunFP32 ? FMF_KERNEL_FP32
                                    : (z.kernel == FMF_KERNEL_BF16 ? FMF_KERNEL_BF16
                                                                   : FMF_KERNEL_BF16X3);
  return rc;
}

// RAII holder for the device buffers of one host call.
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes, const char *what) {
    if (bytes == 0) bytes = 256;
    if (injected(FMF_INJECT_ALLOC))
      return set_error(kOutOfMemory, "hipMalloc of %zu bytes for %s failed: injected failure",
                       bytes, what);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return set_error(kOutOfMemory, "hipMalloc of %zu bytes for %s failed: %s",
                       bytes, what, hipGetErrorString(e));
    }
    return kOk;
  }
};

struct StreamHolder {
  hipStream_t s = nullptr;
  ~StreamHolder() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct EventHolder {
  hipEvent_t e = nullptr;
  ~EventHolder() {
    if (e) (void)hipEventDestroy(e);
  }
};

int host_call(const float *templates, const float *sst, const int *moveouts,
              const float *data, const float *weights, size_t step, size_t N,
              size_t L, size_t T, size_t S, size_t C, size_t n_corr,
              float *cc_out, int normalize, int sum_cc_mode, int kernel,
              int flags) {
  if (!templates || !sst || !moveouts || !data || !weights || !cc_out)
    return set_error(kInvalidArgument, "null pointer argument");
  Sizes z;
  int rc = validate(step, N, L, T, S, C, n_corr, normalize, sum_cc_mode,
                    kernel, flags, z);
  if (rc) return rc;
  const long long SC = z.SC;
  StreamHolder st;
  if ((rc = create_stream(&st.s))) return rc;
  hipStream_t s = st.s;

  size_t free_b = 0, total_b = 0;
  FMF_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));

  DevBuf d_tp, d_sst, d_mo, d_w, d_data, d_ws, d_out;
  const size_t n_tsc = (size_t)z.T * SC;
  if ((rc = d_data.alloc((size_t)SC * z.L * sizeof(float), "data"))) return rc;
  if ((rc = d_tp.alloc(n_tsc * z.N * sizeof(float), "templates"))) return rc;
  if ((rc = d_sst.alloc(n_tsc * sizeof(float), "sum_square_templates"))) return rc;
  if ((rc = d_mo.alloc(n_tsc * sizeof(int), "moveouts"))) return rc;
  if ((rc = d_w.alloc(n_tsc * sizeof(float), "weights"))) return rc;

  const int shape = mfma_shape_of(z.step);
  const int run = run_kind(z, shape, false);
  if (run < 0) return kUnsupported;

  // Templates per launch: two output blocks (one computing while the other
  // is copied to the host) plus prepared templates within about half of the
  // memory that is still free after the inputs; at least four blocks when
  // there are that many templates, so that all but the last block's copy
  // overlaps the correlation of the next.
  size_t used = (size_t)SC * z.L * 4 + n_tsc * (z.N + 3) * 4;
  size_t avail = free_b > used ? (free_b - used) / 2 : 0;
  const WsLayout w0 = layout(z, 0);
  const size_t fixed = w0.total;
  const size_t out_row = (size_t)z.n_corr * (z.summed ? 1 : SC);
  const size_t per_t = 2 * out_row * 4 + (size_t)SC * z.n_pad * 4 + 8 +
                       (size_t)SC * w0.img_stride * sizeof(uint16_t);
  long long tc = avail > fixed ? (long long)((avail - fixed) / per_t) : 0;
  if (tc > (z.T + 3) / 4) tc = (z.T + 3) / 4;
  if (tc < 1) tc = 1;
  const long long n_blocks = (z.T + tc - 1) / tc;
  const WsLayout w = layout(z, tc);
  if ((rc = d_ws.alloc(w.total, "workspace"))) return rc;
  DevBuf d_out2;
  if ((rc = d_out.alloc((size_t)tc * out_row * sizeof(float), "cc output"))) return rc;
  if (n_blocks > 1 && (rc = d_out2.alloc((size_t)tc * out_row * sizeof(float), "cc output")))
    return rc;

  FMF_HIP_CHECK(hipMemcpyAsync(d_data.p, data, (size_t)SC * z.L * 4, hipMemcpyHostToDevice, s));
  FMF_HIP_CHECK(hipMemcpyAsync(d_tp.p, templates, n_tsc * z.N * 4, hipMemcpyHostToDevice, s));
  FMF_HIP_CHECK(hipMemcpyAsync(d_sst.p, sst, n_tsc * 4, hipMemcpyHostToDevice, s));
  FMF_HIP_CHECK(hipMemcpyAsync(d_mo.p, moveouts, n_tsc * 4, hipMemcpyHostToDevice, s));
  FMF_HIP_CHECK(hipMemcpyAsync(d_w.p, weights, n_tsc * 4, hipMemcpyHostToDevice, s));

  DevIn in{static_cast<const float *>(d_tp.p), static_cast<const float *>(d_sst.p),
           static_cast<const float *>(d_data.p), static_cast<const float *>(d_w.p),
           static_cast<const int *>(d_mo.p)};
  char *ws = static_cast<char *>(d_ws.p);
  if ((rc = compute_energy(z, in, ws, w, run, s))) return rc;
  if (n_blocks == 1) {
    if ((rc = correlate(z, in, 0, z.T, static_cast<float *>(d_out.p), ws, w, run, shape, s)))
      return rc;
    return copy_to_host(cc_out, d_out.p, (size_t)z.T * out_row * sizeof(float), s);
  }
  // Block k is correlated into buffer k % 2 on stream s and copied out on
  // stream s2; the correlation of block k + 1 is queued before the host
  // runs the copy of block k (copy_to_host returns when it is done), so it
  // computes meanwhile.  Block k + 2 waits for the copy of block k to free
  // the buffer.  Both streams have the default priority.
  StreamHolder st2;
  if ((rc = create_stream(&st2.s))) return rc;
  EventHolder done[2], freed[2];
  for (int b = 0; b < 2; b++) {
    FMF_HIP_CHECK(hipEventCreateWithFlags(&done[b].e, hipEventDisableTiming));
    FMF_HIP_CHECK(hipEventCreateWithFlags(&freed[b].e, hipEventDisableTiming));
  }
  float *buf[2] = {static_cast<float *>(d_out.p), static_cast<float *>(d_out2.p)};
  auto launch_block = [&](long long k) -> int {
    const long long t0 = k * tc, n = std::min(tc, z.T - t0);
    if (k >= 2) FMF_HIP_CHECK(hipStreamWaitEvent(s, freed[k % 2].e, 0));
    int r = correlate(z, in, t0, n, buf[k % 2], ws, w, run, shape, s);
    if (r) return r;
    FMF_HIP_CHECK(hipEventRecord(done[k % 2].e, s));
    return kOk;
  };
  if ((rc = launch_block(0))) return rc;
  for (long long k = 0; k < n_blocks; k++) {
    if (k + 1 < n_blocks && (rc = launch_block(k + 1))) return rc;
    const long long t0 = k * tc, n = std::min(tc, z.T - t0);
    FMF_HIP_CHECK(hipStreamWaitEvent(st2.s, done[k % 2].e, 0));
    if ((rc = copy_to_host(cc_out + (size_t)t0 * out_row, buf[k % 2],
                           (size_t)n * out_row * sizeof(float), st2.s)))
      return rc;
    FMF_HIP_CHECK(hipEventRecord(freed[k % 2].e, st2.s));
  }
  FMF_HIP_CHECK(hipStreamSynchronize(st2.s));
  FMF_HIP_CHECK(hipStreamSynchronize(s));
  return kOk;
}

// ---- batched calls: one length per template -------------------------------

struct BatchLayout {
  BatchTables tab;
  size_t templates, range, t_eoff, order, planes, images, total;
  long long plane_pairs, img_stride;
};

// Energy tables (one per distinct length), prepared templates, ranges, table
// offsets and, when the 32x32 kernels take the longest template, the split
// operands (data planes and template images, sized for step 1 or 2).
int batch_layout(const int *distinct, int n_distinct, long long T, long long SC,
                 long long L, long long n_pad, BatchLayout &b, long long n_max = 0) {
  if (!distinct || n_distinct < 1 || n_distinct > kMaxBatchLengths)
    return set_error(kInvalidArgument, "1..%d distinct template lengths are supported",
                     kMaxBatchLengths);
  b.tab.n = n_distinct;
  size_t off = 0;  // bytes
  for (int d = 0; d < n_distinct; d++) {
    const long long n = distinct[d];
    if (n < 1 || n > n_pad)
      return set_error(kInvalidArgument, "template length %lld outside [1, max]", n);
    for (int e = 0; e < d; e++)
      if (distinct[e] == n) return set_error(kInvalidArgument, "lengths must be distinct");
    b.tab.len[d] = (int)n;
    b.tab.offset[d] = (long long)(off / sizeof(float));
    const long long n_win = L - n + 1 > 0 ? L - n + 1 : 0;
    // the table, then per channel its largest 1/sqrt(e)
    off = align_up(off + ((size_t)SC * n_win + energy_tail_words(SC, L, n)) * sizeof(float),
                   256);
  }
  b.templates = off;
  off = align_up(off + (size_t)T * SC * n_pad * sizeof(float), 256);
  b.range = off;
  off = align_up(off + (size_t)T * 2 * sizeof(int), 256);
  b.t_eoff = off;
  off = align_up(off + (size_t)T * sizeof(long long), 256);
  b.order = off;
  off = align_up(off + (size_t)T * sizeof(int), 256);
  b.planes = b.images = off;
  b.plane_pairs = b.img_stride = 0;
  if (n_max <= 0) n_max = n_pad;
  if (mfma32_covers(1, n_max, L, 1, SC)) {
    b.plane_pairs = mfma32_plane_pairs(L);
    b.img_stride = std::max(mfma32_image_stride(n_max, 1), mfma32_image_stride(n_max, 2));
    off = align_up(off + (size_t)SC * 4 * b.plane_pairs * sizeof(uint32_t), 256);
    b.images = off;
    off = align_up(off + (size_t)T * SC * b.img_stride * sizeof(uint16_t), 256);
  }
  b.total = off;
  return kOk;
}

int batched_device(const float *templates, const float *sst, const int *moveouts,
                   const float *data, const float *weights, const int *lengths,
                   const int *distinct, int n_distinct, size_t step,
                   size_t n_max, size_t L, size_t T, size_t S, size_t C,
                   size_t n_corr, float *out, int normalize, int flags,
                   long long persistent_blocks, void *workspace,
                   size_t ws_bytes, hipStream_t s) {
  if (!templates || !sst || !moveouts || !data || !weights || !lengths ||
      !out || !workspace)
    return set_error(kInvalidArgument, "null pointer argument");
  if (step != 1 && step != 2)
    return set_error(kUnsupported, "batched calls run the MFMA kernel: step 1 or 2");
  if (flags & FMF_FLAG_EXACT_ORDER)
    return set_error(kUnsupported, "exact-order mode is a single-length FP32 mode");
  Sizes z;
  int rc = validate(step, n_max, L, T, S, C, n_corr, normalize, 1,
                    FMF_KERNEL_BF16X3, flags, z);
  if (rc) return rc;
  BatchLayout b;
  if ((rc = batch_layout(distinct, n_distinct, z.T, z.SC, z.L, z.n_pad, b, z.N))) return rc;
  if (ws_bytes < b.total)
    return set_error(kInvalidArgument, "workspace too small: %zu < %zu bytes", ws_bytes, b.total);
  const int shape = mfma_shape_of(z.step);
  const bool k32 = mfma32_shape(shape);
  if (k32 ? !(mfma32_covers(z.step, z.N, z.L, z.n_corr, z.SC) && b.plane_pairs > 0)
          : !mfma16_covers(z.step, z.N, z.L, z.n_corr, z.SC, 1))
    return set_error(kUnsupported, "template length beyond the MFMA tiling (N <= %s)",
                     k32 ? "1344 - 31*step" : "1280 - 15*step");
  char *ws = static_cast<char *>(workspace);
  float *energy = reinterpret_cast<float *>(ws);
  {
    ScopedTimer timer(1, s);
    // every distinct length's table from one read of the data
    long long lens[kMaxBatchLengths];
    float *tabs[kMaxBatchLengths];
    for (int d = 0; d < b.tab.n; d++) {
      lens[d] = b.tab.len[d];
      tabs[d] = energy + b.tab.offset[d];
    }
    rc = launch_window_energies(data, z.SC, z.L, lens, tabs, b.tab.n, normalize, s, k32);
    if (rc) return rc;
    timer.done(s);
  }
  float *tp = reinterpret_cast<float *>(ws + b.templates);
  int *range = reinterpret_cast<int *>(ws + b.range);
  long long *t_eoff = reinterpret_cast<long long *>(ws + b.t_eoff);
  uint32_t *planes = reinterpret_cast<uint32_t *>(ws + b.planes);
  uint16_t *img = reinterpret_cast<uint16_t *>(ws + b.images);
  if ((rc = launch_batch_offsets(lengths, z.T, b.tab, t_eoff, s))) return rc;
  if ((rc = launch_prepare_templates(templates, z.T * z.SC, z.N, z.n_pad, normalize,
                                     tp, s, lengths, z.SC)))
    return rc;
  int *order = reinterpret_cast<int *>(ws + b.order);
  if (k32 && (rc = launch_batch_order(lengths, weights, z.T, z.SC, b.tab, order, s))) return rc;
  if (k32 && ((rc = launch_split_planes(data, z.SC, z.L, b.plane_pairs, planes, s)) ||
              (rc = launch_template_images(tp, z.T * z.SC, z.n_pad, z.N, z.step, lengths,
                                           z.SC, b.img_stride, img, s))))
    return rc;
  if ((rc = launch_out_ranges(moveouts, z.T, z.SC, z.step, z.N, z.L, z.n_corr,
                              range, s, lengths)))
    return rc;
  CorrArgs a{};
  a.templates = tp;
  a.sst = sst;
  a.moveouts = moveouts;
  a.weights = weights;
  a.data = data;
  a.energy = energy;
  a.out_range = range;
  a.out = out;
  a.n_templates = z.T;
  a.n_channels = z.SC;
  a.n_tp = z.N;
  a.n_pad = z.n_pad;
  a.n_data = z.L;
  a.n_win = z.L - z.N + 1;
  a.n_corr = z.n_corr;
  a.step = z.step;
  a.threshold_mode = (flags & FMF_FLAG_THRESHOLD_AFTER_SQRT) ? kThrAfterSqrt : kThrBeforeSqrt;
  a.summed = 1;
  a.t_len = lengths;
  a.t_eoff = t_eoff;
  a.persistent_blocks = persistent_blocks;
  a.planes = planes;
  a.plane_pairs = b.plane_pairs;
  a.images = img;
  a.img_stride = b.img_stride;
  a.t_order = k32 ? order : nullptr;
  ScopedTimer timer(0, s);
  rc = launch_corr_mfma(a, shape, s);
  timer.done(s);
  if (rc == kUnsupported)
    return set_error(kUnsupported, "template length beyond the MFMA tiling");
  if (!rc) g_last_kernel = FMF_KERNEL_BF16X3;
  return rc;
}

// Body of the device-resident call on a workspace of ws_bytes at ws.
int device_call(const float *templates, const float *sst, const int *moveouts,
                const float *data, const float *weights, size_t step, size_t N, size_t L,
                size_t T, size_t S, size_t C, size_t n_corr, float *cc_out, int normalize,
                int sum_cc_mode, int kernel, int flags, char *ws, size_t ws_bytes,
                hipStream_t s) {
  if (!templates || !sst || !moveouts || !data || !weights || !cc_out)
    return set_error(kInvalidArgument, "null pointer argument");
  Sizes z;
  int rc = validate(step, N, L, T, S, C, n_corr, normalize, sum_cc_mode, kernel, flags, z);
  if (rc) return rc;
  const WsLayout w = layout(z, z.T);
  if (ws_bytes < w.total)
    return set_error(kInvalidArgument, "workspace too small: %zu < %zu bytes", ws_bytes,
                     w.total);
  DevIn in{templates, sst, data, weights, moveouts};
  const int shape = mfma_shape_of(z.step);
  const int run = run_kind(z, shape, false);
  if (run < 0) return kUnsupported;
  if ((rc = compute_energy(z, in, ws, w, run, s))) return rc;
  return correlate(z, in, 0, z.T, cc_out, ws, w, run, shape, s);
}

// Workspace of fmf_matched_filter_device_noalloc: a fixed array in device
// memory, one per GPU the library runs on, held until the process ends.  In
// host code `scratch` names a host-side stand-in, not device memory: host
// code uses only the device address that hipGetSymbolAddress returns for the
// current device (kept per device below), never the array's name.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device with that device current.
int scratch_of_current_device(char **ptr) {
  int dev = 0;
  FMF_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(kUnsupported, "device %d: the scratch table holds %d devices", dev,
                     kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    FMF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(kDeviceError, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<char *>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace
}  // namespace fmf

using namespace fmf;

extern "C" {

const char *fmf_last_error(void) { return g_err; }

int fmf_last_kernel(void) { return g_last_kernel; }

int fmf_debug_inject(int point, int nth) {
  g_err[0] = 0;
  if ((point != FMF_INJECT_ALLOC && point != FMF_INJECT_STREAM) || nth < 0)
    return set_error(kInvalidArgument, "fmf_debug_inject: point %d, nth %d", point, nth);
  g_inject[point] = nth;
  return kOk;
}

int fmf_version(void) { return 10000; }

int fmf_copy_to_host(void *dst, const void *src, size_t bytes, void *stream) {
  g_err[0] = 0;
  if (bytes == 0) return kOk;
  if (!dst || !src) return set_error(kInvalidArgument, "null pointer argument");
  return copy_to_host(dst, src, bytes, static_cast<hipStream_t>(stream));
}

int matched_filter_variable_precise(float *templates,
                                    float *sum_square_templates, int *moveouts,
                                    float *data, float *weights, size_t step,
                                    int *n_samples_template,
                                    size_t n_samples_data, size_t n_templates,
                                    size_t n_stations, size_t n_components,
                                    size_t n_corr, float *cc_sum,
                                    int normalize) {
  g_err[0] = 0;
  if (!templates || !sum_square_templates || !moveouts || !data || !weights ||
      !n_samples_template || !cc_sum)
    return set_error(kInvalidArgument, "null pointer argument");
  if (step < 1 || n_templates < 1 || n_stations < 1 || n_components < 1)
    return set_error(kInvalidArgument, "sizes must be >= 1");
  if (normalize != 0 && normalize != 1)
    return set_error(kInvalidArgument, "normalize must be 0 or 1");
  if (n_samples_data >= 0x7fffffff || n_corr >= 0x7fffffff)
    return set_error(kInvalidArgument, "n_samples_data and n_corr must stay below 2^31");
  if (!sizes_fit(n_templates, n_stations, n_components, 1, n_samples_data, n_corr, 1))
    return set_error(kInvalidArgument, "sizes too large");
  const size_t SC = n_stations * n_components, TSC = n_templates * SC;
  // Host-side scalars in the reference's float operation order
  // (matched_filter.c:338-385).
  int n_max = 0;
  for (size_t i = 0; i < TSC; i++) {
    if (n_samples_template[i] < 1)
      return set_error(kInvalidArgument, "template lengths must be >= 1");
    if (n_samples_template[i] > n_max) n_max = n_samples_template[i];
  }
  if (!sizes_fit(n_templates, n_stations, n_components, (size_t)n_max, n_samples_data, n_corr,
                 1))
    return set_error(kInvalidArgument, "sizes too large");
  float *cn = (float *)malloc(TSC * sizeof(float));
  float *tscal = (float *)malloc(n_templates * 3 * sizeof(float));
  int *trange = (int *)malloc(n_templates * 2 * sizeof(int));
  if (!cn || !tscal || !trange) {
    free(cn); free(tscal); free(trange);
    return set_error(kOutOfMemory, "host allocation failed");
  }
  float wsum = 0.0f;  // never reset between templates (matched_filter.c:329,365)
  for (size_t t = 0; t < n_templates; t++) {
    const size_t net = t * SC;
    int lo = 0, hi = 0, used = 0;
    for (size_t ch = 0; ch < SC; ch++) {
      if (moveouts[net + ch] < lo) lo = moveouts[net + ch];
      if (moveouts[net + ch] > hi) hi = moveouts[net + ch];
      wsum += weights[net + ch];
      if (weights[net + ch] != 0.0f) used += 1;
    }
    float zsum = 0.0f;
    for (size_t ch = 0; ch < SC; ch++) {
      cn[net + ch] = 1. / sqrtf(1. / (n_samples_template[net + ch] - 3)) * weights[net + ch];
      zsum += cn[net + ch];
    }
    const long long start = (long long)(ceilf(abs(lo) / (float)step)) * (long long)step;
    long long stop = (long long)n_samples_data - n_max - hi - (long long)step;
    if (stop < start) stop = start;
    trange[2 * t] = (int)start;
    trange[2 * t + 1] = (int)stop;
    tscal[3 * t + 0] = zsum;
    tscal[3 * t + 1] = used * FLT_EPSILON * 10;
    tscal[3 * t + 2] = wsum;
  }
  StreamHolder st;
  DevBuf d_tp, d_sst, d_mo, d_w, d_len, d_cn, d_ts, d_tr, d_data, d_out;
  int rc = kOk;
  const size_t L = n_samples_data;
  rc = create_stream(&st.s);
  if (!rc) rc = d_tp.alloc(TSC * n_max * 4, "templates");
  if (!rc) rc = d_sst.alloc(TSC * 4, "sum_square_templates");
  if (!rc) rc = d_mo.alloc(TSC * 4, "moveouts");
  if (!rc) rc = d_w.alloc(TSC * 4, "weights");
  if (!rc) rc = d_len.alloc(TSC * 4, "lengths");
  if (!rc) rc = d_cn.alloc(TSC * 4, "fisher weights");
  if (!rc) rc = d_ts.alloc(n_templates * 12, "template scalars");
  if (!rc) rc = d_tr.alloc(n_templates * 8, "lag ranges");
  if (!rc) rc = d_data.alloc(SC * L * 4, "data");
  if (!rc) rc = d_out.alloc(n_templates * n_corr * 4, "cc output");
  struct { void *d; const void *h; size_t n; } up[] = {
      {d_tp.p, templates, TSC * n_max * 4}, {d_sst.p, sum_square_templates, TSC * 4},
      {d_mo.p, moveouts, TSC * 4}, {d_w.p, weights, TSC * 4},
      {d_len.p, n_samples_template, TSC * 4}, {d_cn.p, cn, TSC * 4},
      {d_ts.p, tscal, n_templates * 12}, {d_tr.p, trange, n_templates * 8},
      {d_data.p, data, SC * L * 4}};
  for (auto &u : up) {
    if (rc) break;
    if (hipMemcpyAsync(u.d, u.h, u.n, hipMemcpyHostToDevice, st.s) != hipSuccess)
      rc = set_error(kDeviceError, "host-to-device copy failed");
  }
  if (!rc) {
    VarprecArgs a{};
    a.templates = static_cast<const float *>(d_tp.p);
    a.sst = static_cast<const float *>(d_sst.p);
    a.moveouts = static_cast<const int *>(d_mo.p);
    a.weights = static_cast<const float *>(d_w.p);
    a.lengths = static_cast<const int *>(d_len.p);
    a.cn = static_cast<const float *>(d_cn.p);
    a.tscal = static_cast<const float *>(d_ts.p);
    a.trange = static_cast<const int *>(d_tr.p);
    a.data = static_cast<const float *>(d_data.p);
    a.out = static_cast<float *>(d_out.p);
    a.n_templates = (int)n_templates;
    a.n_channels = (int)SC;
    a.n_max = n_max;
    a.n_data = (int)L;
    a.n_corr = (int)n_corr;
    a.step = (int)step;
    a.normalize = normalize;
    rc = launch_corr_varprec(a, st.s);
  }
  if (!rc && hipMemcpyAsync(cc_sum, d_out.p, n_templates * n_corr * 4,
                            hipMemcpyDeviceToHost, st.s) != hipSuccess)
    rc = set_error(kDeviceError, "device-to-host copy failed");
  if (!rc && hipStreamSynchronize(st.s) != hipSuccess)
    rc = set_error(kDeviceError, "stream synchronisation failed");
  free(cn);
  free(tscal);
  free(trange);
  return rc;
}

int fmf_matched_filter_items_device(
    const float *templates, const float *sum_square_templates,
    const int *moveouts, const float *data, const float *weights, size_t step,
    size_t n_samples_template, size_t n_samples_data, size_t n_templates,
    size_t n_stations, size_t n_components, size_t n_corr, float *cc_out,
    int normalize, int flags, const int *items, long long n_items, int tile_shape,
    void *workspace, size_t workspace_bytes, void *stream) {
  g_err[0] = 0;
  if (!templates || !sum_square_templates || !moveouts || !data || !weights ||
      !cc_out || !workspace || (!items && n_items))
    return set_error(kInvalidArgument, "null pointer argument");
  if (n_items < 0) return set_error(kInvalidArgument, "n_items must be >= 0");
  if (tile_shape != 0 && !mfma16_shape(tile_shape) && !mfma32_shape(tile_shape))
    return set_error(kInvalidArgument, "unknown MFMA tile shape %d", tile_shape);
  Sizes z;
  int rc = validate(step, n_samples_template, n_samples_data, n_templates,
                    n_stations, n_components, n_corr, normalize, 1,
                    FMF_KERNEL_BF16X3, flags, z);
  if (rc) return rc;
  const WsLayout w = layout(z, z.T);
  if (workspace_bytes < w.total)
    return set_error(kInvalidArgument, "workspace too small: %zu < %zu bytes",
                     workspace_bytes, w.total);
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevIn in{templates, sum_square_templates, data, weights, moveouts};
  char *ws = static_cast<char *>(workspace);
  const int shape = tile_shape ? tile_shape : mfma_shape_of(z.step);
  const int run = run_kind(z, shape, true);
  if (run < 0) return kUnsupported;
  if ((rc = compute_energy(z, in, ws, w, run, s))) return rc;
  if (n_items == 0) return kOk;
  return correlate(z, in, 0, z.T, cc_out, ws, w, run, shape, s, items, n_items);
}

int fmf_matched_filter_items_mixed_device(
    const float *templates, const float *sum_square_templates,
    const int *moveouts, const float *data, const float *weights, size_t step,
    size_t n_samples_template, size_t n_samples_data, size_t n_templates,
    size_t n_stations, size_t n_components, size_t n_corr, float *cc_out,
    const int *items, long long n_items, const int *screen_items,
    long long n_screen_items, int tile_shape, void *workspace, size_t workspace_bytes,
    void *stream) {
  g_err[0] = 0;
  if (!templates || !sum_square_templates || !moveouts || !data || !weights ||
      !cc_out || !workspace || (!items && n_items) || (!screen_items && n_screen_items))
    return set_error(kInvalidArgument, "null pointer argument");
  if (n_items < 0 || n_screen_items < 0)
    return set_error(kInvalidArgument, "item counts must be >= 0");
  // shapes whose bf16x3 and screening blocks cover the same lags, so that
  // both lists number their tiles alike
  if (tile_shape != 212 && tile_shape != 412)
    return set_error(kInvalidArgument, "mixed item lists take tile shape 212 or 412, not %d",
                     tile_shape);
  Sizes z;
  int rc = validate(step, n_samples_template, n_samples_data, n_templates,
                    n_stations, n_components, n_corr, 0, 1, FMF_KERNEL_BF16X3, 0, z);
  if (rc) return rc;
  const WsLayout w = layout(z, z.T);
  if (workspace_bytes < w.total)
    return set_error(kInvalidArgument, "workspace too small: %zu < %zu bytes",
                     workspace_bytes, w.total);
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevIn in{templates, sum_square_templates, data, weights, moveouts};
  char *ws = static_cast<char *>(workspace);
  // Both kernels are 32x32 kernels: one 1/sqrt(e) table and one set of split
  // operands serve both lists.
  const int run = run_kind(z, tile_shape, true);
  if (run < 0) return kUnsupported;
  if ((rc = compute_energy(z, in, ws, w, run, s))) return rc;
  if (n_items && (rc = correlate(z, in, 0, z.T, cc_out, ws, w, run, tile_shape, s, items,
                                 n_items)))
    return rc;
  if (n_screen_items) {
    Sizes zs = z;
    zs.kernel = FMF_KERNEL_BF16;
    rc = correlate(zs, in, 0, z.T, cc_out, ws, w, run, tile_shape, s, screen_items,
                   n_screen_items);
  }
  return rc;
}

int fmf_mfma_item_outputs(int step) {
  if (step < 1 || step > 2) return 0;
  return mfma_block_outputs(mfma_shape_of(step));
}

int fmf_mfma_shape_outputs(int shape) {
  if (!mfma16_shape(shape) && !mfma32_shape(shape)) return 0;
  return mfma_block_outputs(shape);
}

// ---- detection post-processing --------------------------------------------

size_t fmf_detect_workspace_bytes(long long n_templates, long long n_corr) {
  if (n_templates <= 0 || n_corr <= 0) return 0;
  return detect_workspace_bytes(n_templates, n_corr);
}

int fmf_lag_ranges_device(const int *moveouts, long long n_templates,
                          long long n_channels, int step,
                          long long n_samples_template, long long n_samples_data,
                          long long n_corr, int *lag_ranges, void *stream) {
  g_err[0] = 0;
  if (n_templates <= 0 || n_channels <= 0 || step < 1 || n_samples_template < 1 ||
      n_corr < 0 || !moveouts || !lag_ranges)
    return set_error(kInvalidArgument, "fmf_lag_ranges_device: invalid arguments");
  return launch_out_ranges(moveouts, n_templates, n_channels, step, n_samples_template,
                           n_samples_data, n_corr, lag_ranges,
                           static_cast<hipStream_t>(stream));
}

namespace {
int check_detect_args(const float *cc, long long T, long long n_corr, void *ws) {
  if (!cc || T <= 0 || n_corr <= 0 || !ws)
    return set_error(kInvalidArgument, "detection: invalid arguments");
  if (T > 65535)
    return set_error(kInvalidArgument, "detection: at most 65535 templates per call (got %lld)", T);
  if (n_corr >= (1LL << 31))
    return set_error(kInvalidArgument, "detection: n_corr must be < 2^31");
  return kOk;
}
}  // namespace

int fmf_median_mad_device(const float *cc, long long n_templates, long long n_corr,
                          const int *lag_ranges, float *median, float *mad,
                          void *workspace, size_t workspace_bytes, void *stream) {
  g_err[0] = 0;
  int rc = check_detect_args(cc, n_templates, n_corr, workspace);
  if (rc) return rc;
  if (!median || !mad) return set_error(kInvalidArgument, "detection: null output");
  return launch_median_mad(cc, n_templates, n_corr, lag_ranges, median, mad, workspace,
                           workspace_bytes, static_cast<hipStream_t>(stream));
}

int fmf_detect_device(const float *cc, long long n_templates, long long n_corr,
                      const int *lag_ranges, const float *median, const float *mad,
                      float n_mad, long long min_separation, int *det_template,
                      int *det_lag, float *det_value, long long capacity,
                      long long *n_detections, void *workspace, size_t workspace_bytes,
                      void *stream) {
  g_err[0] = 0;
  int rc = check_detect_args(cc, n_templates, n_corr, workspace);
  if (rc) return rc;
  if (!median || !mad || !n_detections || capacity < 0 || min_separation < 0 ||
      (capacity > 0 && (!det_template || !det_lag || !det_value)))
    return set_error(kInvalidArgument, "detection: invalid outputs or separation");
  return launch_detect(cc, n_templates, n_corr, lag_ranges, median, mad, n_mad,
                       min_separation, det_template, det_lag, det_value, capacity,
                       n_detections, workspace, workspace_bytes,
                       static_cast<hipStream_t>(stream));
}

// ---- screener ----------------------------------------------------------

#define FMF_STREAM(p) static_cast<hipStream_t>(p)

int fmf_screen_envelope(const float *data, int n_channels, int n_data,
                        int n_dec, float *env, void *stream) {
  return launch_envelope(data, n_channels, n_data, n_dec, env, FMF_STREAM(stream));
}

int fmf_screen_box_mean(const float *env, int n_channels, int n_dec, int width,
                        float *out, void *stream) {
  return launch_box_mean(env, n_channels, n_dec, width, out, FMF_STREAM(stream));
}

int fmf_screen_features(const float *win, const float *winh, const float *env,
                        const float *chinv, const int *moveouts,
                        const float *weights, int n_templates, int n_channels,
                        int n_dec, int n_out, int nwh, float len_feature,
                        float *feat, void *stream) {
  return launch_features(win, winh, env, chinv, moveouts, weights, n_templates,
                         n_channels, n_dec, n_out, nwh, len_feature, feat,
                         FMF_STREAM(stream));
}

int fmf_conv1d_forward(const float *in, int nb, int cin, int nx, int relu_in,
                       const float *w, const float *bias, int cout, float *out,
                       void *stream) {
  return launch_conv_fwd(in, nb, cin, nx, relu_in, w, bias, cout, out, FMF_STREAM(stream));
}

size_t fmf_conv1d_forward_fast_scratch_bytes(int cin) {
  return (cin == 8 || cin == 32) ? conv_fwd_fast_scratch_bytes(cin) : 0;
}

int fmf_conv1d_forward_fast(const float *in, int nb, int cin, int nx, int relu_in,
                            const float *w, const float *bias, int cout, float *out,
                            void *scratch, size_t scratch_bytes, void *stream) {
  g_err[0] = 0;
  if (!scratch || scratch_bytes < fmf_conv1d_forward_fast_scratch_bytes(cin))
    return set_error(kInvalidArgument, "fmf_conv1d_forward_fast: scratch too small");
  return launch_conv_fwd_fast(in, nb, cin, nx, relu_in, w, bias, cout, out, scratch,
                              FMF_STREAM(stream));
}

size_t fmf_screen_net_scratch_bytes(void) { return screen_net_scratch_bytes(); }

int fmf_screen_net_forward(const float *feat, int nb, int nx, const float *w1,
                           const float *b1, const float *w2, const float *b2,
                           const float *w3, const float *b3, const float *w4,
                           const float *b4, void *scratch, size_t scratch_bytes,
                           float *pred, void *stream) {
  g_err[0] = 0;
  if (!scratch || scratch_bytes < screen_net_scratch_bytes())
    return set_error(kInvalidArgument, "fmf_screen_net_forward: scratch too small");
  return launch_screen_net(feat, nb, nx, w1, b1, w2, b2, w3, b3, w4, b4, scratch, pred,
                           FMF_STREAM(stream));
}

int fmf_screen_net_forward_f16(const void *feat, int nb, int nx, const float *w1,
                               const float *b1, const float *w2, const float *b2,
                               const float *w3, const float *b3, const float *w4,
                               const float *b4, void *scratch, size_t scratch_bytes,
                               float *pred, void *stream) {
  g_err[0] = 0;
  if (!scratch || scratch_bytes < screen_net_scratch_bytes())
    return set_error(kInvalidArgument, "fmf_screen_net_forward_f16: scratch too small");
  return launch_screen_net(static_cast<const _Float16 *>(feat), nb, nx, w1, b1, w2, b2, w3, b3,
                           w4, b4, scratch, pred, FMF_STREAM(stream));
}

int fmf_screen_decimate(const float *in, long long n_rows, long long n_in, int factor,
                        long long n_out, float *out, void *stream) {
  g_err[0] = 0;
  return launch_decimate(in, n_rows, n_in, factor, n_out, out, FMF_STREAM(stream));
}

int fmf_screen_coarse_features(const float *cc, int n_templates, int n_coarse,
                               const float *mad, int factor, int n_out, float *feat,
                               void *stream) {
  g_err[0] = 0;
  return launch_coarse_features(cc, n_templates, n_coarse, mad, factor, n_out, feat,
                                FMF_STREAM(stream));
}

int fmf_screen_coarse_features_f16(const float *cc, int n_templates, int n_coarse,
                                   const float *mad, int factor, int n_out, void *feat,
                                   void *stream) {
  g_err[0] = 0;
  return launch_coarse_features(cc, n_templates, n_coarse, mad, factor, n_out,
                                static_cast<_Float16 *>(feat), FMF_STREAM(stream));
}

int fmf_conv1d_backward(const float *gz, const float *ain, const float *zin,
                        int nb, int cin, int nx, int relu_in, const float *w,
                        int cout, float *ga, float *gw, float *gb,
                        void *stream) {
  return launch_conv_bwd(gz, ain, zin, nb, cin, nx, relu_in, w, cout, ga, gw,
                         gb, FMF_STREAM(stream));
}

int fmf_screen_loss(const float *pred, const float *target, long long n,
                    float boost, float inv_norm, float *grad, float *partial,
                    int blocks, void *stream) {
  return launch_loss(pred, target, n, boost, inv_norm, grad, partial, blocks,
                     FMF_STREAM(stream));
}

int fmf_adam_step(float *master, const float *grad, float *m, float *v, int n,
                  float lr, float b1, float b2, float eps, int t, void *w16,
                  float *w32, void *stream) {
  return launch_adam(master, grad, m, v, n, lr, b1, b2, eps, t, w16, w32,
                     FMF_STREAM(stream));
}

int fmf_screen_gate(const float *pred, int n_out, const float *thr,
                    int n_tiles, int tile_dec, int margin, int n_templates,
                    int *keep, void *stream) {
  return launch_gate(pred, n_out, thr, n_tiles, tile_dec, margin, n_templates,
                     keep, FMF_STREAM(stream));
}

int fmf_screen_peak_target(const float *cc, int n_corr, int n_out,
                           int n_templates, float *target, void *stream) {
  return launch_peak_target(cc, n_corr, n_out, n_templates, target, FMF_STREAM(stream));
}

size_t fmf_batched_workspace_bytes(const int *distinct_lengths, int n_distinct,
                                   size_t max_samples_template,
                                   size_t n_samples_data, size_t n_templates,
                                   size_t n_stations, size_t n_components) {
  BatchLayout b;
  Sizes z;
  if (validate(1, max_samples_template, n_samples_data, n_templates, n_stations, n_components,
               1, 0, 1, FMF_KERNEL_BF16X3, 0, z))
    return 0;
  const long long n_pad = (long long)align_up(max_samples_template, kTapBlock);
  if (batch_layout(distinct_lengths, n_distinct, (long long)n_templates,
                   (long long)(n_stations * n_components),
                   (long long)n_samples_data, n_pad, b, (long long)max_samples_template))
    return 0;
  return b.total;
}

int fmf_matched_filter_batched_device(
    const float *templates, const float *sum_square_templates,
    const int *moveouts, const float *data, const float *weights,
    const int *lengths, const int *distinct_lengths, int n_distinct,
    size_t step, size_t max_samples_template, size_t n_samples_data,
    size_t n_templates, size_t n_stations, size_t n_components, size_t n_corr,
    float *cc_out, int normalize, int flags, long long persistent_blocks,
    void *workspace, size_t workspace_bytes, void *stream) {
  g_err[0] = 0;
  return batched_device(templates, sum_square_templates, moveouts, data,
                        weights, lengths, distinct_lengths, n_distinct, step,
                        max_samples_template, n_samples_data, n_templates,
                        n_stations, n_components, n_corr, cc_out, normalize,
                        flags, persistent_blocks, workspace, workspace_bytes,
                        static_cast<hipStream_t>(stream));
}

int fmf_set_mfma_tiles(int step, int shape) {
  if (step < 1 || step > 2)
    return set_error(kInvalidArgument, "MFMA tiles are tuned for step 1 and 2");
  if (!mfma16_shape(shape) && !mfma32_shape(shape))
    return set_error(kInvalidArgument,
                     "shape must be 24, 42, 43 or 81 (16x16 tiles: 10 * tiles per wave + "
                     "waves per SIMD) or 121, 141, 212, 412, 413, 421, 422, 241 or 1422 "
                     "(32x32 tiles: 100 * waves per block + 10 * tiles per wave + "
                     "waves per SIMD; 1422: 422 with interleaved tiles)");
  g_mfma_tiles[step].store(shape);
  return kOk;
}

int fmf_get_mfma_tiles(int step) {
  return (step >= 1 && step <= 2) ? g_mfma_tiles[step].load() : 0;
}

int fmf_kernel_timer_enable(int on) {
  std::lock_guard<std::mutex> lk(g_timer_mu);
  g_timer_on = on != 0;
  return kOk;
}

int fmf_kernel_timer_read(int which, double *total_ms, int *launches) {
  g_err[0] = 0;
  std::vector<TimedLaunch> mine, keep;
  {
    std::lock_guard<std::mutex> lk(g_timer_mu);
    for (auto &t : g_timed) (t.which == which ? mine : keep).push_back(t);
    g_timed.swap(keep);
  }
  double sum = 0.0;
  int rc = kOk;
  for (auto &t : mine) {
    float ms = 0.0f;
    hipError_t e = hipEventSynchronize(t.stop);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t.start, t.stop);
    if (e != hipSuccess && rc == kOk)
      rc = set_error(kDeviceError, "event timing failed: %s", hipGetErrorString(e));
    sum += ms;
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  if (total_ms) *total_ms = sum;
  if (launches) *launches = (int)mine.size();
  return rc;
}

int fmf_matched_filter_host(const float *templates,
                            const float *sum_square_templates,
                            const int *moveouts, const float *data,
                            const float *weights, size_t step,
                            size_t n_samples_template, size_t n_samples_data,
                            size_t n_templates, size_t n_stations,
                            size_t n_components, size_t n_corr, float *cc_out,
                            int normalize, int sum_cc_mode, int kernel,
                            int flags) {
  g_err[0] = 0;
  return host_call(templates, sum_square_templates, moveouts, data, weights,
                   step, n_samples_template, n_samples_data, n_templates,
                   n_stations, n_components, n_corr, cc_out, normalize,
                   sum_cc_mode, kernel, flags);
}

int matched_filter(float *templates, float *sum_square_templates,
                   int *moveouts, float *data, float *weights, size_t step,
                   size_t n_samples_template, size_t n_samples_data,
                   size_t n_templates, size_t n_stations, size_t n_components,
                   size_t n_corr, float *cc_out, int normalize,
                   int sum_cc_mode) {
  return fmf_matched_filter_host(templates, sum_square_templates, moveouts,
                                 data, weights, step, n_samples_template,
                                 n_samples_data, n_templates, n_stations,
                                 n_components, n_corr, cc_out, normalize,
                                 sum_cc_mode, FMF_KERNEL_FP32, 0);
}

int matched_filter_precise(float *templates, float *sum_square_templates,
                           int *moveouts, float *data, float *weights,
                           size_t step, size_t n_samples_template,
                           size_t n_samples_data, size_t n_templates,
                           size_t n_stations, size_t n_components,
                           size_t n_corr, float *cc_sum, int normalize) {
  return fmf_matched_filter_host(templates, sum_square_templates, moveouts,
                                 data, weights, step, n_samples_template,
                                 n_samples_data, n_templates, n_stations,
                                 n_components, n_corr, cc_sum, normalize, 1,
                                 FMF_KERNEL_FP32, FMF_FLAG_THRESHOLD_AFTER_SQRT);
}

int matched_filter_no_sum(float *templates, float *sum_square_templates,
                          int *moveouts, float *data, float *weights,
                          size_t step, size_t n_samples_template,
                          size_t n_samples_data, size_t n_templates,
                          size_t n_stations, size_t n_components,
                          size_t n_corr, float *cc) {
  return fmf_matched_filter_host(templates, sum_square_templates, moveouts,
                                 data, weights, step, n_samples_template,
                                 n_samples_data, n_templates, n_stations,
                                 n_components, n_corr, cc, 0, 0,
                                 FMF_KERNEL_FP32, 0);
}

int matched_filter_precise_no_sum(float *templates,
                                  float *sum_square_templates, int *moveouts,
                                  float *data, float *weights, size_t step,
                                  size_t n_samples_template,
                                  size_t n_samples_data, size_t n_templates,
                                  size_t n_stations, size_t n_components,
                                  size_t n_corr, float *cc, int normalize) {
  return fmf_matched_filter_host(templates, sum_square_templates, moveouts,
                                 data, weights, step, n_samples_template,
                                 n_samples_data, n_templates, n_stations,
                                 n_components, n_corr, cc, normalize, 0,
                                 FMF_KERNEL_FP32, FMF_FLAG_THRESHOLD_AFTER_SQRT);
}

int cumsum_square_data(float *data, size_t n_samples_data, float *weights,
                       size_t n_stations, size_t n_components,
                       double *csum_square_data) {
  (void)weights;  // unused by the reference too (matched_filter.c:660-677)
  g_err[0] = 0;
  if (!data || !csum_square_data)
    return set_error(kInvalidArgument, "null pointer argument");
  if (n_samples_data < 1 || n_stations < 1 || n_components < 1)
    return set_error(kInvalidArgument, "sizes must be >= 1");
  if (n_samples_data >= 0x7fffffff ||
      !sizes_fit(1, n_stations, n_components, 1, n_samples_data, 1, 1))
    return set_error(kInvalidArgument, "sizes too large");
  const size_t SC = n_stations * n_components;
  StreamHolder st;
  int rc;
  if ((rc = create_stream(&st.s))) return rc;
  DevBuf d_data, d_P;
  if ((rc = d_data.alloc(SC * n_samples_data * 4, "data"))) return rc;
  if ((rc = d_P.alloc(SC * (n_samples_data + 1) * 8, "csum"))) return rc;
  FMF_HIP_CHECK(hipMemcpyAsync(d_data.p, data, SC * n_samples_data * 4,
                               hipMemcpyHostToDevice, st.s));
  if ((rc = launch_cumsum_square(static_cast<const float *>(d_data.p), (long long)SC,
                                 (long long)n_samples_data,
                                 static_cast<double *>(d_P.p), st.s)))
    return rc;
  FMF_HIP_CHECK(hipMemcpyAsync(csum_square_data, d_P.p, SC * (n_samples_data + 1) * 8,
                               hipMemcpyDeviceToHost, st.s));
  FMF_HIP_CHECK(hipStreamSynchronize(st.s));
  return kOk;
}

size_t fmf_workspace_bytes(size_t n_samples_template, size_t n_samples_data,
                           size_t n_templates, size_t n_stations,
                           size_t n_components, int normalize, int kernel,
                           int flags) {
  Sizes z;
  if (validate(1, n_samples_template, n_samples_data, n_templates, n_stations,
               n_components, 1, normalize, 1, kernel, flags, z))
    return 0;
  return layout(z, z.T).total;
}

int fmf_matched_filter_device(const float *templates,
                              const float *sum_square_templates,
                              const int *moveouts, const float *data,
                              const float *weights, size_t step,
                              size_t n_samples_template, size_t n_samples_data,
                              size_t n_templates, size_t n_stations,
                              size_t n_components, size_t n_corr,
                              float *cc_out, int normalize, int sum_cc_mode,
                              int kernel, int flags, void *workspace,
                              size_t workspace_bytes, void *stream) {
  g_err[0] = 0;
  if (!workspace) return set_error(kInvalidArgument, "null pointer argument");
  return device_call(templates, sum_square_templates, moveouts, data, weights, step,
                     n_samples_template, n_samples_data, n_templates, n_stations,
                     n_components, n_corr, cc_out, normalize, sum_cc_mode, kernel, flags,
                     static_cast<char *>(workspace), workspace_bytes,
                     static_cast<hipStream_t>(stream));
}

size_t fmf_noalloc_scratch_bytes(void) { return kScratchBytes; }

int fmf_matched_filter_device_noalloc(const float *templates,
                                      const float *sum_square_templates,
                                      const int *moveouts, const float *data,
                                      const float *weights, size_t step,
                                      size_t n_samples_template, size_t n_samples_data,
                                      size_t n_templates, size_t n_stations,
                                      size_t n_components, size_t n_corr, float *cc_out,
                                      int normalize, int sum_cc_mode, int kernel, int flags,
                                      void *stream) {
  g_err[0] = 0;
  char *ws = nullptr;
  int rc = scratch_of_current_device(&ws);
  if (rc) return rc;
  return device_call(templates, sum_square_templates, moveouts, data, weights, step,
                     n_samples_template, n_samples_data, n_templates, n_stations,
                     n_components, n_corr, cc_out, normalize, sum_cc_mode, kernel, flags, ws,
                     kScratchBytes, static_cast<hipStream_t>(stream));
}

}  // extern "C"
