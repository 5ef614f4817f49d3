/*
 * libfmf_hip — MI355X (gfx950) matched-filter engine, C ABI.
 *
 * Drop-in boundary for the reference's native libraries
 * (beridel/fast_matched_filter v1.5.0).  Every pointer of the reference-ABI
 * entry points is a host pointer to C-contiguous memory owned by the caller
 * (NumPy in the reference wrapper); arrays use the reference layouts:
 *   templates            float32 (T, S, C, N)
 *   sum_square_templates float32 (T, S, C)
 *   moveouts             int32   (T, S, C)
 *   weights              float32 (T, S, C)
 *   data                 float32 (S, C, L)
 *   cc_out               float32 (T, n_corr)        summed over the network
 *                        float32 (T, n_corr, S, C)  per channel
 * Lag i runs over [start, 1 + L - N - max_mo) in steps of `step` and lands at
 * output index i / step (start > 0 only for negative moveouts); every other
 * entry of cc_out in [0, n_corr) is written as 0.
 *
 * All entry points return 0 on success and a non-zero fmf_status on failure
 * (never exit()); fmf_last_error() describes the last failure of the calling
 * thread.  Host-pointer calls are synchronous and use streams of their own on
 * the current device.  Calls from several threads may run at the same time:
 * the library's process-wide state (the MFMA tile shape of each step, read
 * once when a call starts; the kernel timer; the one-time raise of a
 * kernel's LDS limit per device) is atomic or lock-protected, and errors and
 * fmf_last_kernel() are per thread.  Device-resident calls that share a
 * workspace or an output buffer must not overlap.
 */
#ifndef FMF_HIP_H
#define FMF_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fmf_status {
  FMF_OK = 0,
  FMF_INVALID_ARGUMENT = 1,
  FMF_OUT_OF_MEMORY = 2,
  FMF_DEVICE_ERROR = 3,
  FMF_UNSUPPORTED = 4
};

/* Correlation kernels. */
enum fmf_kernel {
  FMF_KERNEL_FP32 = 0,      /* FP32 vector-ALU kernel                          */
  FMF_KERNEL_BF16X3 = 1,    /* bf16x3 split-precision MFMA, FP32 accumulation  */
  /* Screening pass (no reference counterpart): one bf16 product per MAC
   * (about three significant digits, |dcc| <= 2^-8 per unit weight) and
   * cc = num / sqrt(sst * e) for e > 0, else 0, without the 1e-6
   * threshold.  Short normalization, step 1 or 2, N <= 1344 - 31*step, no
   * item lists. */
  FMF_KERNEL_BF16 = 2
};

/* Flags of the extended entry points. */
enum fmf_flags {
  /* Reproduce the reference CPU build's float rounding order (FP32 kernel
   * only): unfused products in k order, FMA network sum, FP64 division, and
   * the reference-order Neumaier running sum for the window energies. */
  FMF_FLAG_EXACT_ORDER = 1,
  /* Threshold after the square root (arch 'precise', matched_filter.c:650-655)
   * instead of before it (archs 'cpu'/'gpu', matched_filter.c:622-625). */
  FMF_FLAG_THRESHOLD_AFTER_SQRT = 2,
  /* With FMF_KERNEL_BF16X3: fail with FMF_UNSUPPORTED instead of running the
   * FP32 kernel when the MFMA tiling does not take the call (step other than
   * 1 or 2, L >= 2^29, or N > 1344 - 31*step with the 32x32 tile shapes, the
   * default; N > 1280 - 15*step with the 16x16 ones). */
  FMF_FLAG_STRICT_KERNEL = 4
};

/* ---- Reference-ABI entry points -------------------------------------- */

/* Replaces the reference GPU library's only symbol,
 * fast_matched_filter/src/matched_filter.cu:171-176 (declared at
 * matched_filter_GPU.h:1-4, redeclared at matched_filter_CPU.h:74-77),
 * bound at fast_matched_filter.py:52-55 and called at :314-316.
 * normalize: 0 'short', 1 'full'.  sum_cc_mode: 1 network sum, 0 per channel.
 * Runs the FP32 kernel with the 'cpu'/'gpu' threshold rule. */
int matched_filter(float *templates, float *sum_square_templates,
                   int *moveouts, float *data, float *weights, size_t step,
                   size_t n_samples_template, size_t n_samples_data,
                   size_t n_templates, size_t n_stations, size_t n_components,
                   size_t n_corr, float *cc_out, int normalize,
                   int sum_cc_mode);

/* Replaces matched_filter_precise, fast_matched_filter/src/matched_filter.c:98-168
 * (matched_filter_CPU.h:14-18; bound at fast_matched_filter.py:38, called at
 * :301-303): 'precise' threshold rule, short or full normalization. */
int matched_filter_precise(float *templates, float *sum_square_templates,
                           int *moveouts, float *data, float *weights,
                           size_t step, size_t n_samples_template,
                           size_t n_samples_data, size_t n_templates,
                           size_t n_stations, size_t n_components,
                           size_t n_corr, float *cc_sum, int normalize);

/* Replaces matched_filter_no_sum, matched_filter.c:170-247
 * (matched_filter_CPU.h:27-31; fast_matched_filter.py:39, :299-300). */
int matched_filter_no_sum(float *templates, float *sum_square_templates,
                          int *moveouts, float *data, float *weights,
                          size_t step, size_t n_samples_template,
                          size_t n_samples_data, size_t n_templates,
                          size_t n_stations, size_t n_components,
                          size_t n_corr, float *cc);

/* Replaces matched_filter_precise_no_sum, matched_filter.c:249-319
 * (matched_filter_CPU.h:41-45; fast_matched_filter.py:40, :304-306). */
int matched_filter_precise_no_sum(float *templates,
                                  float *sum_square_templates, int *moveouts,
                                  float *data, float *weights, size_t step,
                                  size_t n_samples_template,
                                  size_t n_samples_data, size_t n_templates,
                                  size_t n_stations, size_t n_components,
                                  size_t n_corr, float *cc, int normalize);

/* Replaces matched_filter_variable_precise, matched_filter.c:321-412 with
 * :569-606 (matched_filter_CPU.h:55-59; bound at fast_matched_filter.py:41,
 * called at :307-309): per-channel template lengths n_samples_template
 * (T, S, C) int32, templates (T, S, C, max length) zero padded; channels
 * combined with atanh/tanh (Fisher z) weighted by w / sqrt(1 / (n - 3)).
 * The reference's own conventions are kept: the running sum of weights used
 * as the output scale is not reset between templates, the lag loop stops at
 * L - max length - max moveout - step, and every one of the n_corr outputs is
 * mapped through tanh. */
int matched_filter_variable_precise(float *templates,
                                    float *sum_square_templates, int *moveouts,
                                    float *data, float *weights, size_t step,
                                    int *n_samples_template,
                                    size_t n_samples_data, size_t n_templates,
                                    size_t n_stations, size_t n_components,
                                    size_t n_corr, float *cc_sum,
                                    int normalize);

/* Replaces cumsum_square_data, matched_filter.c:660-677 (+ :680-705):
 * per-channel FP64 running sum of squares, csum (S*C, L+1), csum[ch][0] = 0,
 * bit-identical to the reference's Neumaier order.  `weights` is accepted and
 * ignored, as in the reference. */
int cumsum_square_data(float *data, size_t n_samples_data, float *weights,
                       size_t n_stations, size_t n_components,
                       double *csum_square_data);

/* ---- Extended entry points (no reference counterpart) ----------------- */

/* Host-pointer call with an explicit kernel and flags; the reference-ABI
 * entry points above are this call with fixed kernel/flags.  With
 * FMF_KERNEL_BF16X3 the FP32 kernel runs instead where the MFMA tiling does
 * not take the call, unless flags hold FMF_FLAG_STRICT_KERNEL;
 * fmf_last_kernel() tells which ran.  With more than one block of templates
 * (memory, or more than 4 templates) the copy of each block's output to the
 * host overlaps the correlation of the next. */
int fmf_matched_filter_host(const float *templates,
                            const float *sum_square_templates,
                            const int *moveouts, const float *data,
                            const float *weights, size_t step,
                            size_t n_samples_template, size_t n_samples_data,
                            size_t n_templates, size_t n_stations,
                            size_t n_components, size_t n_corr, float *cc_out,
                            int normalize, int sum_cc_mode, int kernel,
                            int flags);

/* Device-resident call: every array pointer is device memory on the current
 * device, `stream` is a hipStream_t (NULL = default stream), `workspace` is
 * device memory of at least fmf_workspace_bytes(...) bytes.  Stream-ordered;
 * returns without synchronising. */
size_t fmf_workspace_bytes(size_t n_samples_template, size_t n_samples_data,
                           size_t n_templates, size_t n_stations,
                           size_t n_components, int normalize, int kernel,
                           int flags);

int fmf_matched_filter_device(const float *templates,
                              const float *sum_square_templates,
                              const int *moveouts, const float *data,
                              const float *weights, size_t step,
                              size_t n_samples_template, size_t n_samples_data,
                              size_t n_templates, size_t n_stations,
                              size_t n_components, size_t n_corr,
                              float *cc_out, int normalize, int sum_cc_mode,
                              int kernel, int flags, void *workspace,
                              size_t workspace_bytes, void *stream);

/* Allocation-free variant of fmf_matched_filter_device (no counterpart in the
 * reference), for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * caller passes inputs and outputs (device pointers) and a stream; the
 * workspace is a fixed array of fmf_noalloc_scratch_bytes() (64 MiB) in
 * device memory that the library holds on each GPU it runs on until the
 * process ends.  Calls whose workspace (fmf_workspace_bytes) exceeds it fail
 * with FMF_INVALID_ARGUMENT.  The call allocates and frees nothing.  Every
 * call on one GPU shares that GPU's array, and every byte the kernels read
 * from it is written earlier in the same call (nothing needs to start at
 * zero); calls that overlap on one GPU — concurrent streams or threads —
 * race on it and must be serialised by the caller. */
size_t fmf_noalloc_scratch_bytes(void);
int fmf_matched_filter_device_noalloc(const float *templates,
                                      const float *sum_square_templates,
                                      const int *moveouts, const float *data,
                                      const float *weights, size_t step,
                                      size_t n_samples_template, size_t n_samples_data,
                                      size_t n_templates, size_t n_stations,
                                      size_t n_components, size_t n_corr, float *cc_out,
                                      int normalize, int sum_cc_mode, int kernel, int flags,
                                      void *stream);

/* Batched device call (no reference counterpart; the reference takes one
 * template length per call, matched_filter.c:24): template t uses its first
 * lengths[t] samples (templates are (T, S, C, max_samples_template), zero
 * padded), with its own lag range and window energies, on the bf16x3 kernel
 * (step 1 or 2), in one launch.  `lengths` is device memory (T int32);
 * `distinct_lengths` (host, n_distinct <= 16) lists every value it holds —
 * a template whose length is not listed gets a zero row.
 * persistent_blocks > 0 caps the grid (persistent launch); 0 = one block per
 * work item.  Output (T, n_corr) summed over the network; n_corr is usually
 * (L - min length) / step + 1.  Stream-ordered, no host synchronisation and
 * no host-to-device copies, so it can be captured in a HIP graph. */
size_t fmf_batched_workspace_bytes(const int *distinct_lengths, int n_distinct,
                                   size_t max_samples_template,
                                   size_t n_samples_data, size_t n_templates,
                                   size_t n_stations, size_t n_components);

int fmf_matched_filter_batched_device(
    const float *templates, const float *sum_square_templates,
    const int *moveouts, const float *data, const float *weights,
    const int *lengths, const int *distinct_lengths, int n_distinct,
    size_t step, size_t max_samples_template, size_t n_samples_data,
    size_t n_templates, size_t n_stations, size_t n_components, size_t n_corr,
    float *cc_out, int normalize, int flags, long long persistent_blocks,
    void *workspace, size_t workspace_bytes, void *stream);

/* Screened device call (no reference counterpart): computes the window
 * energies, then runs the bf16x3 kernel only on the listed work items.  An
 * item is one output tile of one template, numbered tile * n_templates +
 * template; `items` (device, ascending) lists n_items of them; entries
 * outside [0, n_tiles * n_templates), n_tiles = ceil(n_corr / lags per item),
 * are skipped.  The tile is
 * the block of the MFMA tile shape `tile_shape` (fmf_mfma_shape_outputs(shape)
 * lags; 0: the process-wide shape of the step, fmf_mfma_item_outputs(step)
 * lags).  Only those outputs are written (network sum, layout (T, n_corr));
 * the caller initialises the rest.  Workspace as
 * fmf_workspace_bytes(..., kernel = FMF_KERNEL_BF16X3, ...). */
int fmf_matched_filter_items_device(
    const float *templates, const float *sum_square_templates,
    const int *moveouts, const float *data, const float *weights, size_t step,
    size_t n_samples_template, size_t n_samples_data, size_t n_templates,
    size_t n_stations, size_t n_components, size_t n_corr, float *cc_out,
    int normalize, int flags, const int *items, long long n_items, int tile_shape,
    void *workspace, size_t workspace_bytes, void *stream);
/* Two item lists in one call (no reference counterpart): one energy pass,
 * then the bf16x3 kernel on `items` and the screening kernel
 * (FMF_KERNEL_BF16: one bf16 product per MAC, |dcc| <= 2^-8 per unit
 * weight, no 1e-6 threshold) on `screen_items`, both numbered as for
 * fmf_matched_filter_items_device on tile shape 212 (2048 lags) or 412
 * (4096), whose screening blocks cover the same lags.  The lists must not
 * share an item.  Short normalization, network sum, step 1 or 2.  For
 * callers that need exact cc on some tiles and only a statistic (a median
 * absolute deviation) of others. */
int fmf_matched_filter_items_mixed_device(
    const float *templates, const float *sum_square_templates,
    const int *moveouts, const float *data, const float *weights, size_t step,
    size_t n_samples_template, size_t n_samples_data, size_t n_templates,
    size_t n_stations, size_t n_components, size_t n_corr, float *cc_out,
    const int *items, long long n_items, const int *screen_items,
    long long n_screen_items, int tile_shape, void *workspace, size_t workspace_bytes,
    void *stream);
int fmf_mfma_item_outputs(int step);
/* Output lags of one block (one work item) of an MFMA tile shape; 0 if the
 * shape is unknown. */
int fmf_mfma_shape_outputs(int shape);

/* Learned screener building blocks (no reference counterpart; all device
 * pointers, stream-ordered).  Decimation is 8 lags per screener sample.
 *  fmf_screen_envelope:   env (S*C, n_dec) = mean of d^2 over 8 samples
 *  fmf_screen_box_mean:   out = mean of env over `width` decimated samples
 *  fmf_screen_features:   feat (T, 8, n_out) moveout-aligned features
 *  fmf_conv1d_forward:    out (nb, cout, nx) = conv9(relu?(in)) + bias
 *  fmf_conv1d_backward:   input (ga, may be NULL), weight and bias gradients
 *  fmf_screen_loss:       peak-weighted squared error and its gradient
 *  fmf_adam_step:         Adam on float32 master weights -> fp16 + fp32 copy
 *  fmf_screen_gate:       keep[t][tile] = any prediction in the tile > thr[t]
 *  fmf_screen_peak_target: target (T, n_out) = max |cc| over each 8 lags */
int fmf_screen_envelope(const float *data, int n_channels, int n_data,
                        int n_dec, float *env, void *stream);
int fmf_screen_box_mean(const float *env, int n_channels, int n_dec, int width,
                        float *out, void *stream);
int fmf_screen_features(const float *win, const float *winh, const float *env,
                        const float *chinv, const int *moveouts,
                        const float *weights, int n_templates, int n_channels,
                        int n_dec, int n_out, int nwh, float len_feature,
                        float *feat, void *stream);
int fmf_conv1d_forward(const float *in, int nb, int cin, int nx, int relu_in,
                       const float *w, const float *bias, int cout, float *out,
                       void *stream);
/* Screener, coarse-correlation front end and inference:
 *  fmf_screen_decimate:        out (rows, n_out) = mean of `factor` consecutive
 *                              samples of each row (past n_in: 0)
 *  fmf_screen_coarse_features: feat (T, 8, n_out) from the cc (T, n_coarse) of a
 *                              scan decimated by `factor` and its MAD (T);
 *                              the _f16 variant stores IEEE half values (the
 *                              network rounds its input to FP16 either way)
 *  fmf_conv1d_forward_fast:    fmf_conv1d_forward on the matrix cores (FP16
 *                              activations and weights, FP32 accumulation),
 *                              cout 32 and cin 8 or 32; scratch of
 *                              fmf_conv1d_forward_fast_scratch_bytes(cin) */
/* The screener network (conv 8->32, 32->32, 32->32, 32->1, kernel 9, ReLU
 * between layers) in one fused inference kernel: feat (nb, 8, nx) ->
 * pred (nb, nx); hidden activations FP16 in LDS, FP32 accumulation.
 * Weights (cout, cin, 9) and biases are float32 device arrays (values the
 * FP16 rounding of the trained weights); scratch of
 * fmf_screen_net_scratch_bytes() holds the packed FP16 weights.  The _f16
 * variant reads feat as IEEE half values (same predictions for features
 * that are the FP16 rounding of the float ones). */
size_t fmf_screen_net_scratch_bytes(void);
int fmf_screen_net_forward(const float *feat, int nb, int nx, const float *w1,
                           const float *b1, const float *w2, const float *b2,
                           const float *w3, const float *b3, const float *w4,
                           const float *b4, void *scratch, size_t scratch_bytes,
                           float *pred, void *stream);
int fmf_screen_net_forward_f16(const void *feat, int nb, int nx, const float *w1,
                               const float *b1, const float *w2, const float *b2,
                               const float *w3, const float *b3, const float *w4,
                               const float *b4, void *scratch, size_t scratch_bytes,
                               float *pred, void *stream);
int fmf_screen_decimate(const float *in, long long n_rows, long long n_in, int factor,
                        long long n_out, float *out, void *stream);
int fmf_screen_coarse_features(const float *cc, int n_templates, int n_coarse,
                               const float *mad, int factor, int n_out, float *feat,
                               void *stream);
int fmf_screen_coarse_features_f16(const float *cc, int n_templates, int n_coarse,
                                   const float *mad, int factor, int n_out, void *feat,
                                   void *stream);
size_t fmf_conv1d_forward_fast_scratch_bytes(int cin);
int fmf_conv1d_forward_fast(const float *in, int nb, int cin, int nx, int relu_in,
                            const float *w, const float *bias, int cout, float *out,
                            void *scratch, size_t scratch_bytes, void *stream);
int fmf_conv1d_backward(const float *gz, const float *ain, const float *zin,
                        int nb, int cin, int nx, int relu_in, const float *w,
                        int cout, float *ga, float *gw, float *gb,
                        void *stream);
int fmf_screen_loss(const float *pred, const float *target, long long n,
                    float boost, float inv_norm, float *grad, float *partial,
                    int blocks, void *stream);
int fmf_adam_step(float *master, const float *grad, float *m, float *v, int n,
                  float lr, float b1, float b2, float eps, int t, void *w16,
                  float *w32, void *stream);
int fmf_screen_gate(const float *pred, int n_out, const float *thr,
                    int n_tiles, int tile_dec, int margin, int n_templates,
                    int *keep, void *stream);
int fmf_screen_peak_target(const float *cc, int n_corr, int n_out,
                           int n_templates, float *target, void *stream);

/* Detection post-processing on device-resident cc rows (no reference
 * counterpart: the reference returns cc_sum, fast_matched_filter.py:318-362,
 * and leaves detection to the caller).  Row t of `cc` (T, n_corr) is taken
 * over its valid lags [lag_ranges[2t], lag_ranges[2t+1]) (NULL: all lags):
 *   fmf_median_mad_device: median[t] = numpy.median(x) and
 *     mad[t] = numpy.median(|x - median[t]|), exactly, in float32;
 *   fmf_detect_device: thr = median + n_mad * mad (float32, unfused); lag j
 *     is a detection when x[j] > thr and x[j] is the first maximum of x over
 *     [max(lo, j - s), min(hi - 1, j + s)], s = min_separation.  Detections
 *     are written in (template, lag) order to det_template / det_lag /
 *     det_value (first `capacity` of them); *n_detections (device) receives
 *     the total, which may exceed capacity.
 *   fmf_lag_ranges_device: the reference's valid lag range per template
 *     (matched_filter.c:53-79) for moveouts (T, S*C).
 * Stream-ordered, no allocation; workspace of fmf_detect_workspace_bytes. */
size_t fmf_detect_workspace_bytes(long long n_templates, long long n_corr);
int fmf_lag_ranges_device(const int *moveouts, long long n_templates,
                          long long n_channels, int step,
                          long long n_samples_template, long long n_samples_data,
                          long long n_corr, int *lag_ranges, void *stream);
int fmf_median_mad_device(const float *cc, long long n_templates, long long n_corr,
                          const int *lag_ranges, float *median, float *mad,
                          void *workspace, size_t workspace_bytes, void *stream);
int fmf_detect_device(const float *cc, long long n_templates, long long n_corr,
                      const int *lag_ranges, const float *median, const float *mad,
                      float n_mad, long long min_separation, int *det_template,
                      int *det_lag, float *det_value, long long capacity,
                      long long *n_detections, void *workspace,
                      size_t workspace_bytes, void *stream);

/* Tile shape of the bf16x3 kernel for `step` 1 or 2.  16x16x32 MFMA tiles:
 * 10 * (output tiles of 256 lags per wave) + (waves per SIMD the register
 * allocation targets): 24, 42, 43 or 81.  32x32x16 MFMA tiles:
 * 100 * (waves per block) + 10 * (output tiles of 1024 lags per wave) +
 * (waves per SIMD): 212, 412, 413 (default at step 2), 421, 422 or 241, and
 * one-wave blocks 121 or 141 (shapes whose registers would spill at two
 * waves per SIMD are not built); 1422 (default at step 1): 422 with the two
 * tiles of a wave 32 lags apart, so that both take their data operands from
 * the same LDS reads (same products in the same order as 422: same bits);
 * at step 2 it runs as 422.  Chosen per workload by the autotuner
 * (fast_matched_filter_amd/autotune.py); process-wide, read once by each call
 * when it starts. */
int fmf_set_mfma_tiles(int step, int shape);
int fmf_get_mfma_tiles(int step);

/* Kernel timing for measurement: while enabled, every call brackets its
 * correlation launch (which = 0) and its window-energy launch (which = 1)
 * with HIP events on the call's stream.  fmf_kernel_timer_read waits for the
 * recorded launches of one kind, returns their summed duration and count,
 * and forgets them. */
int fmf_kernel_timer_enable(int on);
int fmf_kernel_timer_read(int which, double *total_ms, int *launches);

/* Fault injection for tests of the error paths (no reference counterpart):
 * the nth following event of kind `point` in the calling thread fails, as a
 * real failure of that kind would — FMF_INJECT_ALLOC: a device allocation of
 * a host-pointer call (FMF_OUT_OF_MEMORY); FMF_INJECT_STREAM: the creation of
 * a call's stream (FMF_DEVICE_ERROR).  nth = 0 switches the point off.
 * Returns FMF_INVALID_ARGUMENT for an unknown point or a negative nth. */
enum fmf_inject_point { FMF_INJECT_ALLOC = 1, FMF_INJECT_STREAM = 2 };
int fmf_debug_inject(int point, int nth);

/* Message of the calling thread's last failure ("" if none). */
const char *fmf_last_error(void);

/* The correlation kernel (enum fmf_kernel) that the calling thread's last
 * matched-filter call ran, or -1 if that call failed before running one. */
int fmf_last_kernel(void);

/* Library version, major*10000 + minor*100 + patch. */
int fmf_version(void);

/* Copies `bytes` from device memory `src` to host memory `dst` after the work
 * queued on `stream` (0: the null stream) and returns when the copy is done.
 * No counterpart in the reference, whose wrapper receives the cc rows from
 * its own host driver; here the host-array paths use it for their outputs.
 * Built for pageable destinations the caller has just allocated: host
 * threads fault the destination's pages in while the GPU finishes the queued
 * work, then the data streams through two pinned staging buffers of 64 MiB
 * (kept by the library, one pair per concurrent call) while host threads
 * copy the previous chunk out.  Copies below 32 MiB are one hipMemcpy. */
int fmf_copy_to_host(void *dst, const void *src, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FMF_HIP_H */
