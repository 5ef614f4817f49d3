/*
 * ORACLE — test infrastructure only.
 *
 * Drives every function of the C restatement (fmf_oracle.c) on small inputs
 * under AddressSanitizer and UndefinedBehaviorSanitizer (built by
 * `make -C oracle sanitize`, run by tests/test_sanitize.py): every array is
 * its own heap allocation of exactly the size the reference layout gives it,
 * so a read or write past any of them is reported.  Cases: negative
 * moveouts, a weight-0 channel, steps 1 to 3, N not a multiple of 4, a zero
 * gap, a DC offset, per-channel lengths, and a template longer than the data
 * (no lags).  Besides the sanitizers' reports it checks what the reference's
 * definitions fix: finite outputs, per-channel |cc| <= 1, zeros outside the
 * lag range (matched_filter.c:53-79), the network sum equal to the weighted
 * sum of the per-channel output, and exact running sums of squares on
 * integer-valued data.  Exit status 0 when every check holds.
 *
 * The restatement keeps the reference's preconditions: variable_precise
 * needs the wrapper's layout (rows of the longest channel length, n_corr
 * from it); with a shorter longest length and the padded n_corr it writes
 * past the output, as the reference does (matched_filter.c:383-395).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_cumsum_square_data(const float *data, size_t n_samples_data, size_t n_channels,
                            double *P);
void orc_matched_filter(const float *, const float *, const int *, const float *,
                        const float *, size_t, size_t, size_t, size_t, size_t, size_t,
                        size_t, float *);
void orc_matched_filter_precise(const float *, const float *, const int *, const float *,
                                const float *, size_t, size_t, size_t, size_t, size_t,
                                size_t, size_t, float *, int);
void orc_matched_filter_no_sum(const float *, const float *, const int *, const float *,
                               const float *, size_t, size_t, size_t, size_t, size_t,
                               size_t, size_t, float *);
void orc_matched_filter_precise_no_sum(const float *, const float *, const int *,
                                       const float *, const float *, size_t, size_t, size_t,
                                       size_t, size_t, size_t, size_t, float *, int);
void orc_matched_filter_variable_precise(const float *, const float *, const int *,
                                         const float *, const float *, size_t, const int *,
                                         size_t, size_t, size_t, size_t, size_t, float *, int);

static int failures = 0;
#define CHECK(cond, ...)                                 \
    do {                                                 \
        if (!(cond)) {                                   \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                \
            fputc('\n', stderr);                         \
            failures++;                                  \
        }                                                \
    } while (0)

static unsigned long long rng_state = 88172645463325252ULL;
static double urand(void) /* xorshift64 in [0, 1) */
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) / 9007199254740992.0;
}
static float nrand(void)
{
    double u = urand() + 1e-12, v = urand();
    return (float)(sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v));
}

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

typedef struct {
    size_t T, S, C, N, L, step;
    int mo_lo, mo_hi, gap, zero_weight;
} Case;

static void run_case(const Case *k)
{
    const size_t SC = k->S * k->C, TSC = k->T * SC;
    const size_t n_corr = k->L >= k->N ? (k->L - k->N) / k->step + 1 : 0;
    float *tp = xmalloc(TSC * k->N * sizeof(float));
    float *sst = xmalloc(TSC * sizeof(float));
    int *mo = xmalloc(TSC * sizeof(int));
    float *w = xmalloc(TSC * sizeof(float));
    float *d = xmalloc(SC * k->L * sizeof(float));
    float *sum = xmalloc((n_corr ? k->T * n_corr : 1) * sizeof(float));
    float *per = xmalloc((n_corr ? k->T * n_corr * SC : 1) * sizeof(float));
    int *len = xmalloc(TSC * sizeof(int));
    for (size_t i = 0; i < SC * k->L; i++) d[i] = nrand() + (i / k->L == SC - 1 ? 2.0f : 0.0f);
    if (k->gap && k->L > 64)
        memset(d + k->L / 3, 0, (k->L / 4) * sizeof(float)); /* gap in channel 0 */
    for (size_t r = 0; r < TSC; r++) {
        double s = 0.0;
        for (size_t j = 0; j < k->N; j++) {
            tp[r * k->N + j] = nrand();
            s += (double)tp[r * k->N + j] * tp[r * k->N + j];
        }
        sst[r] = (float)s;
        mo[r] = k->mo_lo + (int)(urand() * (k->mo_hi - k->mo_lo + 1));
        w[r] = (float)(0.05 + 0.2 * urand());
        len[r] = (int)(k->N / 2 + urand() * (k->N - k->N / 2)) + 1;
        if (len[r] > (int)k->N) len[r] = (int)k->N;
        if (len[r] < 4) len[r] = 4 <= (int)k->N ? 4 : (int)k->N;
    }
    if (k->zero_weight) w[TSC - 1] = 0.0f;

    for (int method = 0; method < 2; method++) {
        for (int norm = 0; norm < 2; norm++) {
            if (method == 0 && norm == 1) continue; /* 'cpu' is short only */
            if (!n_corr) continue;
            for (size_t i = 0; i < k->T * n_corr; i++) sum[i] = 0.0f;
            for (size_t i = 0; i < k->T * n_corr * SC; i++) per[i] = 0.0f;
            if (method == 0) {
                orc_matched_filter(tp, sst, mo, d, w, k->step, k->N, k->L, k->T, k->S, k->C,
                                   n_corr, sum);
                orc_matched_filter_no_sum(tp, sst, mo, d, w, k->step, k->N, k->L, k->T, k->S,
                                          k->C, n_corr, per);
            } else {
                orc_matched_filter_precise(tp, sst, mo, d, w, k->step, k->N, k->L, k->T, k->S,
                                           k->C, n_corr, sum, norm);
                orc_matched_filter_precise_no_sum(tp, sst, mo, d, w, k->step, k->N, k->L, k->T,
                                                  k->S, k->C, n_corr, per, norm);
            }
            for (size_t t = 0; t < k->T; t++) {
                int lo = 0, hi = 0;
                for (size_t ch = 0; ch < SC; ch++) {
                    if (mo[t * SC + ch] < lo) lo = mo[t * SC + ch];
                    if (mo[t * SC + ch] > hi) hi = mo[t * SC + ch];
                }
                const long long start = lo < 0 ? (long long)ceil(-lo / (double)k->step) * (long long)k->step : 0;
                const long long stop = 1 + (long long)k->L - (long long)k->N - hi;
                for (size_t o = 0; o < n_corr; o++) {
                    const long long lag = (long long)(o * k->step);
                    const int valid = lag >= start && lag < stop;
                    double lin = 0.0;
                    for (size_t ch = 0; ch < SC; ch++) {
                        const float c = per[(t * n_corr + o) * SC + ch];
                        CHECK(isfinite(c) && fabsf(c) <= 1.0f + 1e-5f,
                              "per-channel cc %g (t %zu o %zu ch %zu)", c, t, o, ch);
                        if (!valid) CHECK(c == 0.0f, "per-channel value outside the lags");
                        if (w[t * SC + ch] != 0.0f) lin += (double)w[t * SC + ch] * c;
                        else CHECK(c == 0.0f, "weight-0 channel written");
                    }
                    const float s = sum[t * n_corr + o];
                    CHECK(isfinite(s), "network cc not finite");
                    if (!valid) CHECK(s == 0.0f, "network value outside the lags");
                    CHECK(fabs(s - lin) <= 1e-5, "sum %g vs weighted per-channel %g", s, lin);
                }
            }
        }
    }
    /* variable_precise: the wrapper's layout (fast_matched_filter.py:216-224,
     * 292-295) — rows of the longest channel length (one channel is N long),
     * zeros beyond each channel's own length, sums of squares over it, and
     * lengths of at least 4 (the Fisher weight uses sqrt(1 / (n - 3))) */
    if (n_corr && k->N >= 4) {
        len[0] = (int)k->N;
        for (size_t r = 0; r < TSC; r++) {
            double s = 0.0;
            for (size_t j = 0; j < k->N; j++) {
                if ((int)j >= len[r]) tp[r * k->N + j] = 0.0f;
                s += (double)tp[r * k->N + j] * tp[r * k->N + j];
            }
            sst[r] = (float)s;
        }
        for (int norm = 0; norm < 2; norm++) {
            orc_matched_filter_variable_precise(tp, sst, mo, d, w, k->step, len, k->L, k->T,
                                                k->S, k->C, n_corr, sum, norm);
            for (size_t i = 0; i < k->T * n_corr; i++)
                CHECK(!isnan(sum[i]), "variable_precise NaN at %zu", i);
        }
    }
    free(tp); free(sst); free(mo); free(w); free(d); free(sum); free(per); free(len);
}

static void run_cumsum(void)
{
    const size_t n_ch = 3, L = 20011;
    float *x = xmalloc(n_ch * L * sizeof(float));
    double *P = xmalloc(n_ch * (L + 1) * sizeof(double));
    for (size_t i = 0; i < n_ch * L; i++) x[i] = (float)((long)(urand() * 2000.0) - 1000);
    orc_cumsum_square_data(x, L, n_ch, P);
    for (size_t ch = 0; ch < n_ch; ch++) {
        long long acc = 0;
        CHECK(P[ch * (L + 1)] == 0.0, "P[0] not 0");
        for (size_t i = 0; i < L; i++) {
            const long long v = (long long)x[ch * L + i];
            acc += v * v;
            CHECK(P[ch * (L + 1) + i + 1] == (double)acc, "prefix %zu of channel %zu", i, ch);
        }
    }
    free(x);
    free(P);
}

int main(void)
{
    const Case cases[] = {
        {2, 2, 3, 37, 900, 1, 0, 40, 1, 1},
        {3, 3, 2, 16, 400, 3, -20, 50, 0, 1},
        {1, 1, 1, 1, 50, 2, 0, 0, 0, 0},
        {2, 2, 1, 64, 300, 1, -70, 5, 1, 0},
        {1, 2, 2, 80, 60, 1, 0, 3, 0, 0},     /* template longer than the data */
        {2, 1, 3, 24, 200, 2, -9, 120, 0, 1}, /* moveouts past most of the data */
    };
    for (size_t i = 0; i < sizeof(cases) / sizeof(cases[0]); i++) run_case(&cases[i]);
    run_cumsum();
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("sanitized oracle run: all checks hold\n");
    return 0;
}
