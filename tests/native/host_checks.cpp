// Host-side argument checks of libfmf_hip under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/native/Makefile builds the host driver,
// fmf_api.hip, with both; tests/test_sanitize.py runs this on the CPU).
//
// Every call below must be refused with the documented status before it
// touches a device: null pointers, zero and negative-looking sizes, sizes
// whose products overflow 64 bits, unknown kernels, flags and tile shapes,
// malformed length lists, and the fault-injection switch's stream failure.
// The sanitizers report any out-of-bounds access or undefined behaviour on
// the way; the process exits 0 only if every status is the expected one.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/fmf_hip.h"

static int failures = 0;

static void expect(int rc, int want, const char *what, const char *needle = nullptr) {
  const char *msg = fmf_last_error();
  if (rc != want || (needle && !strstr(msg, needle))) {
    fprintf(stderr, "FAIL %s: status %d (want %d), message \"%s\"\n", what, rc, want, msg);
    failures++;
  }
}

static void expect_size(size_t got, bool zero, const char *what) {
  if ((got == 0) != zero) {
    fprintf(stderr, "FAIL %s: %zu\n", what, got);
    failures++;
  }
}

int main() {
  const size_t big = (size_t)1 << 62, max = SIZE_MAX;
  float f[64] = {0};
  int i[64] = {0};
  float *F = f;
  int *I = i;

  // ---- reference-ABI entry points --------------------------------------
  expect(matched_filter(nullptr, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 1), FMF_INVALID_ARGUMENT,
         "null templates", "null");
  expect(matched_filter(F, F, I, F, F, 0, 4, 32, 1, 1, 1, 29, F, 0, 1), FMF_INVALID_ARGUMENT,
         "step 0", "step");
  expect(matched_filter(F, F, I, F, F, 1, 0, 32, 1, 1, 1, 29, F, 0, 1), FMF_INVALID_ARGUMENT,
         "N 0");
  expect(matched_filter(F, F, I, F, F, 1, 4, 32, 0, 1, 1, 29, F, 0, 1), FMF_INVALID_ARGUMENT,
         "T 0");
  expect(matched_filter(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 2, 1), FMF_INVALID_ARGUMENT,
         "normalize 2", "normalize");
  expect(matched_filter(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 7), FMF_INVALID_ARGUMENT,
         "sum_cc_mode 7");
  expect(matched_filter(F, F, I, F, F, 1, max, 32, 1, 1, 1, 29, F, 0, 1), FMF_INVALID_ARGUMENT,
         "N = SIZE_MAX");
  expect(matched_filter(F, F, I, F, F, 1, 4, max, 1, 1, 1, 29, F, 0, 1), FMF_INVALID_ARGUMENT,
         "L = SIZE_MAX");
  expect(matched_filter(F, F, I, F, F, 1, 4, 32, 1, 1, 1, max, F, 0, 1), FMF_INVALID_ARGUMENT,
         "n_corr = SIZE_MAX");
  expect(matched_filter(F, F, I, F, F, 1, 4, 1000, big, big, 1, 29, F, 0, 1),
         FMF_INVALID_ARGUMENT, "T * S overflows", "too large");
  expect(matched_filter(F, F, I, F, F, 1, 4, 1000, 1, max, max, 29, F, 0, 1),
         FMF_INVALID_ARGUMENT, "S * C overflows", "too large");
  expect(matched_filter(F, F, I, F, F, max, 4, 1000, 1, 1, 1, 29, F, 0, 1),
         FMF_INVALID_ARGUMENT, "step = SIZE_MAX");
  expect(matched_filter_precise(F, F, I, F, F, 1, 4, 32, (size_t)1 << 20, (size_t)1 << 21, 1,
                                29, F, 0),
         FMF_INVALID_ARGUMENT, "T * channels at 2^41", "too large");
  expect(matched_filter_no_sum(F, F, I, F, F, 1, 4, 100000, 1, 1 << 16, 1 << 15, 99997, F),
         FMF_INVALID_ARGUMENT, "per-channel output of 2^50 values", "too large");
  expect(matched_filter_precise_no_sum(F, F, I, F, F, 1, 4, 32, 1, 0, 1, 29, F, 0),
         FMF_INVALID_ARGUMENT, "S 0");
  expect(matched_filter_variable_precise(F, F, I, F, F, 1, nullptr, 32, 1, 1, 1, 29, F, 0),
         FMF_INVALID_ARGUMENT, "variable_precise null lengths", "null");
  expect(matched_filter_variable_precise(F, F, I, F, F, 1, I, 32, big, big, 1, 29, F, 0),
         FMF_INVALID_ARGUMENT, "variable_precise overflowing sizes", "too large");
  i[0] = 0;
  expect(matched_filter_variable_precise(F, F, I, F, F, 1, I, 32, 1, 1, 1, 29, F, 0),
         FMF_INVALID_ARGUMENT, "variable_precise length 0", "lengths");
  expect(cumsum_square_data(nullptr, 32, F, 1, 1, nullptr), FMF_INVALID_ARGUMENT,
         "cumsum null", "null");
  expect(cumsum_square_data(F, 32, F, big, big, (double *)F), FMF_INVALID_ARGUMENT,
         "cumsum overflowing channels", "too large");

  // ---- extended entry points --------------------------------------------
  expect(fmf_matched_filter_host(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 1, 9, 0),
         FMF_INVALID_ARGUMENT, "unknown kernel", "kernel");
  expect(fmf_matched_filter_host(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 1, 0, 64),
         FMF_INVALID_ARGUMENT, "unknown flag", "flags");
  expect(fmf_matched_filter_host(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 1, 1,
                                 FMF_FLAG_EXACT_ORDER),
         FMF_UNSUPPORTED, "exact order on bf16x3");
  expect(fmf_matched_filter_host(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 1, 1, 2, 0),
         FMF_UNSUPPORTED, "screening kernel with full", "short");
  expect(fmf_matched_filter_device(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 1, 0, 0,
                                   nullptr, 0, nullptr),
         FMF_INVALID_ARGUMENT, "device call without workspace", "null");
  expect(fmf_matched_filter_device(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 1, 0, 0, F, 16,
                                   nullptr),
         FMF_INVALID_ARGUMENT, "device call, workspace too small", "workspace");
  expect(fmf_matched_filter_device_noalloc(F, F, I, F, F, 1, 4, 32, big, big, 1, 29, F, 0, 1,
                                           0, 0, nullptr) == FMF_OK ? -1 : FMF_INVALID_ARGUMENT,
         FMF_INVALID_ARGUMENT, "noalloc call with overflowing sizes");
  expect_size(fmf_workspace_bytes(4, 32, 1, 1, 1, 0, 0, 0), false, "workspace bytes, small");
  expect_size(fmf_workspace_bytes(4, 32, big, big, 1, 0, 0, 0), true, "workspace bytes, T*S");
  expect_size(fmf_workspace_bytes(max, 32, 1, 1, 1, 0, 0, 0), true, "workspace bytes, N");
  expect_size(fmf_workspace_bytes(4, 32, 1, 1, 1, 0, 5, 0), true, "workspace bytes, kernel");
  expect_size(fmf_workspace_bytes(800, 8640000, (size_t)1 << 38, 4, 1, 0, 1, 0), true,
              "workspace bytes, 2^40 template channels");

  int lengths[20] = {400, 600, 800, 1200, 400};
  expect_size(fmf_batched_workspace_bytes(lengths, 4, 1200, 100000, 10, 4, 3), false,
              "batched workspace, valid");
  expect_size(fmf_batched_workspace_bytes(lengths, 5, 1200, 100000, 10, 4, 3), true,
              "batched workspace, duplicate length");
  expect_size(fmf_batched_workspace_bytes(lengths, 17, 1200, 100000, 10, 4, 3), true,
              "batched workspace, 17 lengths");
  expect_size(fmf_batched_workspace_bytes(lengths, 4, 800, 100000, 10, 4, 3), true,
              "batched workspace, length above the maximum");
  expect_size(fmf_batched_workspace_bytes(nullptr, 1, 1200, 100000, 10, 4, 3), true,
              "batched workspace, null lengths");
  expect_size(fmf_batched_workspace_bytes(lengths, 4, 1200, 100000, big, big, 3), true,
              "batched workspace, overflowing sizes");
  expect(fmf_matched_filter_batched_device(F, F, I, F, F, I, lengths, 4, 3, 1200, 100000, 10, 4,
                                           3, 99, F, 0, 0, 0, F, 1 << 20, nullptr),
         FMF_UNSUPPORTED, "batched step 3", "step");
  expect(fmf_matched_filter_batched_device(F, F, I, F, F, I, lengths, 4, 1, 1200, 100000, 10, 4,
                                           3, 99, F, 0, FMF_FLAG_EXACT_ORDER, 0, F, 1 << 20,
                                           nullptr),
         FMF_UNSUPPORTED, "batched exact order");
  expect(fmf_matched_filter_items_device(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 0, nullptr,
                                         3, 0, F, 1 << 20, nullptr),
         FMF_INVALID_ARGUMENT, "items without a list", "null");
  expect(fmf_matched_filter_items_device(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 0, I, -1,
                                         0, F, 1 << 20, nullptr),
         FMF_INVALID_ARGUMENT, "negative item count");
  expect(fmf_matched_filter_items_device(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 0, I, 1,
                                         999, F, 1 << 20, nullptr),
         FMF_INVALID_ARGUMENT, "unknown item tile shape", "shape");
  expect(fmf_matched_filter_items_mixed_device(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, I, 1, I,
                                               1, 422, F, 1 << 20, nullptr),
         FMF_INVALID_ARGUMENT, "mixed items on shape 422", "212 or 412");

  // ---- detection ---------------------------------------------------------
  expect(fmf_median_mad_device(F, 70000, 100, nullptr, F, F, F, 1 << 20, nullptr),
         FMF_INVALID_ARGUMENT, "median/mad, 70000 rows", "65535");
  expect(fmf_median_mad_device(F, 1, (long long)1 << 31, nullptr, F, F, F, 1 << 20, nullptr),
         FMF_INVALID_ARGUMENT, "median/mad, 2^31 lags");
  expect(fmf_detect_device(F, 1, 100, nullptr, F, F, 8.0f, -1, I, I, F, 4, (long long *)F, F,
                           1 << 20, nullptr),
         FMF_INVALID_ARGUMENT, "negative separation");
  expect(fmf_lag_ranges_device(I, 0, 3, 1, 4, 100, 97, I, nullptr), FMF_INVALID_ARGUMENT,
         "lag ranges, no templates");
  expect_size(fmf_detect_workspace_bytes(0, 100), true, "detect workspace, no rows");

  // ---- host copy ---------------------------------------------------------
  expect(fmf_copy_to_host(nullptr, F, 64, nullptr), FMF_INVALID_ARGUMENT, "copy, null dst");
  expect(fmf_copy_to_host(F, nullptr, 64, nullptr), FMF_INVALID_ARGUMENT, "copy, null src");
  expect(fmf_copy_to_host(nullptr, nullptr, 0, nullptr), FMF_OK, "copy of 0 bytes");

  // ---- process settings and the fault-injection switch -------------------
  expect(fmf_set_mfma_tiles(3, 422), FMF_INVALID_ARGUMENT, "tiles for step 3");
  expect(fmf_set_mfma_tiles(1, 7), FMF_INVALID_ARGUMENT, "unknown tile shape", "shape");
  expect(fmf_debug_inject(9, 1), FMF_INVALID_ARGUMENT, "unknown injection point");
  expect(fmf_debug_inject(FMF_INJECT_ALLOC, -1), FMF_INVALID_ARGUMENT, "negative nth");
  expect(fmf_debug_inject(FMF_INJECT_STREAM, 1), FMF_OK, "arm stream failure");
  expect(fmf_matched_filter_host(F, F, I, F, F, 1, 4, 32, 1, 1, 1, 29, F, 0, 1, 0, 0),
         FMF_DEVICE_ERROR, "injected stream failure", "injected");
  expect(fmf_debug_inject(FMF_INJECT_STREAM, 2), FMF_OK, "arm second stream failure");
  expect(fmf_debug_inject(FMF_INJECT_STREAM, 0), FMF_OK, "disarm");
  if (fmf_last_kernel() != -1) {
    fprintf(stderr, "FAIL last kernel after refused calls: %d\n", fmf_last_kernel());
    failures++;
  }

  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("sanitized host checks: all statuses as documented\n");
  return 0;
}
