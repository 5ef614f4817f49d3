/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of the reference matched-filter algorithm
 * (beridel/fast_matched_filter v1.5.0, fast_matched_filter/src/matched_filter.c),
 * written from its documented semantics.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker or the timed CPU baseline; the product path never calls it.
 *
 * Compiled with the reference's CPU flags (Makefile:36-39 of the reference:
 * -O3 -fopenmp -fPIC -ftree-vectorize -march=... -std=gnu99, no fast-math) so
 * that its float reduction order and contraction match the reference build.
 *
 * Every exported function keeps the reference's argument order and meaning.
 * Parity pins (tests/test_oracle.py): bit-identical outputs against the
 * compiled reference (oracle/_ref) on the golden vectors, FP64 closed forms,
 * brute force on tiny inputs and the self-detection invariant.
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>

/* matched_filter.c:17 — cc is left at 0 when the denominator is not above it */
#define ORC_THRESHOLD 0.000001f

/* Per-channel FP64 running sum of squares with Neumaier compensation.
 * Follows matched_filter.c:680-705 (neumaier_cumsum_squared), including the
 * branch test that compares the float-converted magnitudes (fabsf on doubles,
 * :693).  out[0..n) receives the inclusive prefix (the caller's P[1..n]). */
static void orc_neumaier_prefix_sq(const float *x, size_t n, double *out)
{
    double acc = (double)x[0] * (double)x[0];
    double comp = 0.0;
    out[0] = acc;
    for (size_t i = 1; i < n; i++) {
        double sq = (double)x[i] * (double)x[i];
        double nxt = acc + sq;
        if (fabsf((float)acc) >= fabsf((float)sq))
            comp += (acc - nxt) + sq;
        else
            comp += (sq - nxt) + acc;
        acc = nxt;
        out[i] = nxt + comp;
    }
}

/* Layout of matched_filter.c:660-677 (cumsum_square_data): channel ch owns
 * P[ch*(L+1) .. ch*(L+1)+L], P[ch*(L+1)] = 0.  OpenMP over channels. */
void orc_cumsum_square_data(const float *data, size_t n_samples_data,
                            size_t n_channels, double *P)
{
    memset(P, 0, (n_samples_data + 1) * n_channels * sizeof(double));
#pragma omp parallel for
    for (size_t ch = 0; ch < n_channels; ch++)
        orc_neumaier_prefix_sq(data + ch * n_samples_data, n_samples_data,
                               P + ch * (n_samples_data + 1) + 1);
}

/* matched_filter.c:611-627 (corrc): float numerator in k order, denominator
 * from the FP64 prefix difference converted to float, test before the sqrt,
 * double-precision sqrt/division. */
static float orc_cc_prefix(const float *tp, float ss_tp, const float *win,
                           const double *Pwin, size_t n)
{
    float num = 0.0f;
    for (size_t k = 0; k < n; k++)
        num += tp[k] * win[k];
    float den = ss_tp * (float)(Pwin[n] - Pwin[0]);
    float cc = 0.0f;
    if (den > ORC_THRESHOLD)
        cc = num / sqrt(den);
    return cc;
}

/* matched_filter.c:629-657 (corrc_precise): optional float mean removal
 * (normalize > 0), direct float sum of squares, threshold after the sqrt. */
static float orc_cc_direct(const float *tp, float ss_tp, const float *win,
                           size_t n, int normalize)
{
    float mean = 0.0f, num = 0.0f, ssd = 0.0f;
    if (normalize > 0) {
        for (size_t k = 0; k < n; k++)
            mean += win[k];
        mean /= n;
    }
    for (size_t k = 0; k < n; k++) {
        float v = win[k] - mean;
        num += tp[k] * v;
        ssd += v * v;
    }
    float den = sqrt(ss_tp * ssd);
    float cc = 0.0f;
    if (den > ORC_THRESHOLD)
        cc = num / den;
    return cc;
}

/* Lag range of one template: matched_filter.c:53-76.  min/max moveouts start
 * at 0 and run over every channel, weight-zero ones included. */
static void orc_lag_range(const int *mo, size_t n_ch, size_t step,
                          size_t n_tp, size_t n_data,
                          size_t *start, size_t *stop, int *ok)
{
    int lo = 0, hi = 0;
    for (size_t ch = 0; ch < n_ch; ch++) {
        if (mo[ch] < lo) lo = mo[ch];
        if (mo[ch] > hi) hi = mo[ch];
    }
    *start = (lo < 0) ? (size_t)ceilf(-lo / (float)step) * step : 0;
    /* The reference evaluates 1 + L - N - max_mo in size_t; a negative value
     * would wrap.  The restatement reports "no lags" in that case. */
    long long s = 1 + (long long)n_data - (long long)n_tp - (long long)hi;
    *ok = s > 0;
    *stop = s > 0 ? (size_t)s : 0;
}

/* Mode bits of the single driver below. */
enum { ORC_PREFIX = 0, ORC_DIRECT = 1 };

static void orc_driver(const float *templates, const float *ss_templates,
                       const int *moveouts, const float *data,
                       const float *weights, size_t step, size_t n_tp,
                       size_t n_data, size_t n_templates, size_t n_stations,
                       size_t n_components, size_t n_corr, float *out,
                       int method, int normalize, int summed)
{
    const size_t n_ch = n_stations * n_components;
    double *P = NULL;
    if (method == ORC_PREFIX) {
        /* matched_filter.c:41-43 */
        P = malloc((n_data + 1) * n_ch * sizeof(double));
        orc_cumsum_square_data(data, n_data, n_ch, P);
    }
    for (size_t t = 0; t < n_templates; t++) {
        const size_t net = t * n_ch;
        const float *tp_t = templates + net * n_tp;
        const float *ss_t = ss_templates + net;
        const int *mo_t = moveouts + net;
        const float *w_t = weights + net;
        size_t start, stop;
        int ok;
        orc_lag_range(mo_t, n_ch, step, n_tp, n_data, &start, &stop, &ok);
        if (!ok)
            continue;
        /* matched_filter.c:78-92: OpenMP over lags, output index i/step */
#pragma omp parallel for
        for (size_t i = start; i < stop; i += step) {
            const size_t o = i / step;
            float acc = 0.0f;
            /* matched_filter.c:427-451: station-major, skip weight 0 */
            for (size_t ch = 0; ch < n_ch; ch++) {
                if (w_t[ch] == 0.0f)
                    continue;
                const float *win = data + ch * n_data + i + mo_t[ch];
                float cc;
                if (method == ORC_PREFIX)
                    cc = orc_cc_prefix(tp_t + ch * n_tp, ss_t[ch], win,
                                       P + ch * (n_data + 1) + i + mo_t[ch],
                                       n_tp);
                else
                    cc = orc_cc_direct(tp_t + ch * n_tp, ss_t[ch], win, n_tp,
                                       normalize);
                if (summed)
                    acc += cc * w_t[ch];
                else /* matched_filter.c:199,230: (T, n_corr, S, C) */
                    out[(t * n_corr + o) * n_ch + ch] = cc;
            }
            if (summed)
                out[t * n_corr + o] = acc;
        }
    }
    free(P);
}

/* matched_filter.c:22-96 */
void orc_matched_filter(const float *templates, const float *ss_templates,
                        const int *moveouts, const float *data,
                        const float *weights, size_t step, size_t n_tp,
                        size_t n_data, size_t n_templates, size_t n_stations,
                        size_t n_components, size_t n_corr, float *cc_sum)
{
    orc_driver(templates, ss_templates, moveouts, data, weights, step, n_tp,
               n_data, n_templates, n_stations, n_components, n_corr, cc_sum,
               ORC_PREFIX, 0, 1);
}

/* matched_filter.c:98-168 */
void orc_matched_filter_precise(const float *templates,
                                const float *ss_templates, const int *moveouts,
                                const float *data, const float *weights,
                                size_t step, size_t n_tp, size_t n_data,
                                size_t n_templates, size_t n_stations,
                                size_t n_components, size_t n_corr,
                                float *cc_sum, int normalize)
{
    orc_driver(templates, ss_templates, moveouts, data, weights, step, n_tp,
               n_data, n_templates, n_stations, n_components, n_corr, cc_sum,
               ORC_DIRECT, normalize, 1);
}

/* matched_filter.c:170-247 */
void orc_matched_filter_no_sum(const float *templates,
                               const float *ss_templates, const int *moveouts,
                               const float *data, const float *weights,
                               size_t step, size_t n_tp, size_t n_data,
                               size_t n_templates, size_t n_stations,
                               size_t n_components, size_t n_corr, float *cc)
{
    orc_driver(templates, ss_templates, moveouts, data, weights, step, n_tp,
               n_data, n_templates, n_stations, n_components, n_corr, cc,
               ORC_PREFIX, 0, 0);
}

/* matched_filter.c:249-319 */
void orc_matched_filter_precise_no_sum(const float *templates,
                                       const float *ss_templates,
                                       const int *moveouts, const float *data,
                                       const float *weights, size_t step,
                                       size_t n_tp, size_t n_data,
                                       size_t n_templates, size_t n_stations,
                                       size_t n_components, size_t n_corr,
                                       float *cc, int normalize)
{
    orc_driver(templates, ss_templates, moveouts, data, weights, step, n_tp,
               n_data, n_templates, n_stations, n_components, n_corr, cc,
               ORC_DIRECT, normalize, 0);
}

/* matched_filter.c:321-412 and :569-606 (variable_precise).  Per-channel
 * template lengths; channels are combined through the Fisher transform
 * atanh(cc - 10 eps) * sqrt(n - 3) * w and the sum is mapped back with tanh.
 * Kept as in the reference: the running maximum possible cc (sum of weights)
 * is never reset between templates, the lag loop stops at
 * L - Nmax - max_mo - step (exclusive), the start uses |min_mo| and the final
 * tanh pass covers every one of the n_corr entries. */
void orc_matched_filter_variable_precise(
    const float *templates, const float *ss_templates, const int *moveouts,
    const float *data, const float *weights, size_t step, const int *n_tp,
    size_t n_data, size_t n_templates, size_t n_stations, size_t n_components,
    size_t n_corr, float *cc_sum, int normalize)
{
    const size_t n_ch = n_stations * n_components;
    int n_max = 0;
    for (size_t i = 0; i < n_templates * n_ch; i++)
        if (n_tp[i] > n_max)
            n_max = n_tp[i];
    float *zscale = malloc(n_ch * sizeof(float));
    float wsum = 0.0f; /* accumulates over templates, as in the reference */
    for (size_t t = 0; t < n_templates; t++) {
        const size_t net = t * n_ch;
        int lo = 0, hi = 0, used = 0;
        for (size_t ch = 0; ch < n_ch; ch++) {
            if (moveouts[net + ch] < lo) lo = moveouts[net + ch];
            if (moveouts[net + ch] > hi) hi = moveouts[net + ch];
            wsum += weights[net + ch];
            if (weights[net + ch] != 0.0f)
                used += 1;
        }
        float zsum = 0.0f;
        for (size_t ch = 0; ch < n_ch; ch++) {
            zscale[ch] = 1. / sqrtf(1. / (n_tp[net + ch] - 3)) * weights[net + ch];
            zsum += zscale[ch];
        }
        const float *tp_t = templates + net * (size_t)n_max;
        const float *ss_t = ss_templates + net;
        const int *mo_t = moveouts + net;
        const float *w_t = weights + net;
        const int *n_t = n_tp + net;
        size_t start = (int)(ceilf(abs(lo) / (float)step)) * step;
        long long stop_s = (long long)n_data - n_max - hi - (long long)step;
        size_t stop = stop_s > 0 ? (size_t)stop_s : 0;
#pragma omp parallel for
        for (size_t i = start; i < stop; i += step) {
            float acc = 0.0f;
            for (size_t ch = 0; ch < n_ch; ch++) {
                if (w_t[ch] == 0.0f)
                    continue;
                float cc = orc_cc_direct(tp_t + ch * (size_t)n_max, ss_t[ch],
                                         data + ch * n_data + i + mo_t[ch],
                                         (size_t)n_t[ch], normalize);
                acc += atanhf(cc - 10 * FLT_EPSILON) * zscale[ch];
            }
            cc_sum[t * n_corr + i / step] = acc;
        }
#pragma omp parallel for
        for (size_t i = 0; i < n_corr; i++)
            cc_sum[t * n_corr + i] =
                (tanhf(cc_sum[t * n_corr + i] / zsum) + used * FLT_EPSILON * 10) * wsum;
    }
    free(zscale);
}
