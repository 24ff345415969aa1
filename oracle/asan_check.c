/*
 * Memory check of the CPU oracle (test infrastructure): built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by `make -C oracle asan`, it calls every entry point of
 * oracle.h on small codes and on the largest (BG1 Z=384), with shortening, filler bits,
 * repetition (E > N) and every modulation, serial and on several threads, and checks the
 * few facts that hold for any correct decoder: a noiseless lifted codeword decodes to itself
 * in 0 iterations, an encoded block satisfies every check, a clean block of redundancy
 * version 0 decodes to its information bits under every rule. A sanitizer report
 * or a failed check ends the run with a non-zero status; "asan_check: ok" ends a clean one.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            fprintf(stderr, "asan_check: " __VA_ARGS__);  \
            fprintf(stderr, " (%s:%d)\n", __FILE__, __LINE__); \
            ++failures;                                   \
        }                                                 \
    } while (0)

static void* xmalloc(size_t n) {
    void* p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "asan_check: out of memory\n");
        exit(2);
    }
    return p;
}

/* encode, full LLRs and every decoder on one NR code (k_info, e, bg, rv, qm) */
static void nr_code(int32_t k_info, int32_t e, int bg, int rv, int32_t qm, int threads) {
    or_code* c = or_nr_create(k_info, e, bg, rv, qm);
    CHECK(c != NULL, "or_nr_create(%d, %d): %s", k_info, e, or_last_error());
    if (!c) return;
    const int32_t n = (int32_t)or_code_info(c, 0), ki = (int32_t)or_code_info(c, 3);
    const int32_t ee = (int32_t)or_code_info(c, 4);
    CHECK(ki == k_info && ee == e, "dimensions of (%d, %d)", k_info, e);
    uint8_t* info = xmalloc((size_t)ki);
    uint8_t* tx = xmalloc((size_t)ee);
    uint8_t* bits = xmalloc((size_t)ki);
    double* llr = xmalloc(sizeof(double) * (size_t)ee);
    double* full = xmalloc(sizeof(double) * (size_t)n);
    double* post = xmalloc(sizeof(double) * (size_t)n);
    or_rng r;
    or_rng_new(&r, 0x5eedu + (uint64_t)k_info);
    for (int32_t i = 0; i < ki; ++i) info[i] = or_rng_bit(&r);
    CHECK(or_nr_encode(c, info, tx) == 0, "or_nr_encode");
    for (int32_t i = 0; i < ee; ++i) llr[i] = tx[i] ? -4.0 : 4.0;
    CHECK(or_nr_full_llr(c, llr, full) == 0, "or_nr_full_llr");
    const or_dec rules[4] = {{OR_SPA, 0.0, 1, 20}, {OR_NMS, 0.75, 1, 20}, {OR_OMS, 0.5, 1, 20}, {OR_NMS, 0.8, 0, 20}};
    for (int k = 0; k < 4; ++k) {
        const int it = or_nr_decode_block(c, &rules[k], llr, bits);
        CHECK(it >= 0, "or_nr_decode_block rule %d", k);
        /* rv 0 sends the systematic bits first; a later rv with E <= N - k0 may send none */
        if (rv == 0) CHECK(memcmp(bits, info, (size_t)ki) == 0, "clean block, rule %d (%d, %d)", k, k_info, e);
        CHECK(or_decode_llr(c, &rules[k], full, post) >= 0, "or_decode_llr rule %d", k);
    }
    /* noisy BPSK through the QAM chain, many codewords on threads, every precision */
    const int count = 3;
    const int32_t bps = qm == 1 ? 2 : qm;
    uint8_t* infos = xmalloc((size_t)ki * count);
    double* llrs = xmalloc(sizeof(double) * (size_t)ee * count);
    or_trial_nr_qam_many(c, 1 << bps, 0.5, 7, "asan", 1, 0, count, infos, llrs, threads);
    double* fulls = xmalloc(sizeof(double) * (size_t)n * count);
    float* f32 = xmalloc(sizeof(float) * (size_t)n * count);
    for (int t = 0; t < count; ++t) {
        or_nr_full_llr(c, llrs + (size_t)t * ee, fulls + (size_t)t * n);
        for (int32_t i = 0; i < n; ++i) f32[(size_t)t * n + i] = (float)fulls[(size_t)t * n + i];
    }
    double* posts = xmalloc(sizeof(double) * (size_t)n * count);
    float* post32 = xmalloc(sizeof(float) * (size_t)n * count);
    int32_t its[3], lim[3] = {5, 0, 12};
    or_decode_many(c, &rules[1], fulls, count, posts, its, threads);
    or_decode_many_f32(c, &rules[1], f32, count, post32, its, threads);
    or_decode_many_f16(c, &rules[2], f32, count, lim, post32, its, threads);
    CHECK(its[1] == 0, "an iteration limit of 0 runs no iteration");
    or_decode_many_f16(c, &rules[1], f32, count, NULL, post32, its, 1);
    CHECK(or_decode_llr_f32(c, &rules[2], f32, post32) >= 0, "or_decode_llr_f32");
    CHECK(or_decode_llr_f16(c, &rules[1], f32, post32) >= 0, "or_decode_llr_f16");
    double* trace = xmalloc(sizeof(double) * (size_t)(10 + rules[1].max_iterations));
    CHECK(or_trace_f32_f64(c, &rules[1], fulls, trace) >= 0, "or_trace_f32_f64");
    CHECK(or_trace_f32_f64(c, &rules[0], fulls, trace) == -1, "the tracer refuses the sum-product rule");
    /* the device generator's trial in f64 */
    double* yre = xmalloc(sizeof(double) * (size_t)ee);
    double* yim = xmalloc(sizeof(double) * (size_t)ee);
    or_philox_trial(c, qm == 1 ? 1 : qm, 0.3, 9, 2, 5, info, yre, yim, llr);
    or_philox_trial(c, qm == 1 ? 1 : qm, 0.3, 9, 2, 6, info, NULL, NULL, llr);
    free(yre);
    free(yim);
    free(trace);
    free(post32);
    free(posts);
    free(f32);
    free(fulls);
    free(llrs);
    free(infos);
    free(post);
    free(full);
    free(llr);
    free(bits);
    free(tx);
    free(info);
    or_code_free(c);
}

/* lifted codes: systematic encoder, BPSK trials, learned decoder and its gradient */
static void lifted(int bg, int32_t z) {
    or_code* c = or_code_lifted(bg, z);
    CHECK(c != NULL, "or_code_lifted(%d, %d)", bg, z);
    if (!c) return;
    const int32_t n = (int32_t)or_code_info(c, 0), k = (int32_t)or_code_info(c, 2);
    const int rows = bg == 1 ? 46 : 42, T = 4;
    uint8_t* u = xmalloc((size_t)k);
    uint8_t* cw = xmalloc((size_t)n);
    double* llr = xmalloc(sizeof(double) * (size_t)n);
    double* post = xmalloc(sizeof(double) * (size_t)n);
    or_rng r;
    or_rng_new(&r, (uint64_t)z);
    for (int32_t i = 0; i < k; ++i) u[i] = or_rng_bit(&r);
    CHECK(or_encode(c, u, cw) == 0 && or_is_codeword(c, cw) == 1, "lifted encode (%d, %d)", bg, z);
    for (int32_t i = 0; i < n; ++i) llr[i] = cw[i] ? -3.0 : 3.0;
    const or_dec d = {OR_NMS, 0.75, 1, 10};
    CHECK(or_decode_llr(c, &d, llr, post) == 0, "a codeword decodes in 0 iterations");
    double* a = xmalloc(sizeof(double) * (size_t)(T * rows));
    double* b = xmalloc(sizeof(double) * (size_t)(T * rows));
    double* ga = xmalloc(sizeof(double) * (size_t)(T * rows));
    double* gb = xmalloc(sizeof(double) * (size_t)(T * rows));
    for (int i = 0; i < T * rows; ++i) {
        a[i] = 0.75;
        b[i] = 0.1;
    }
    or_trial_bpsk_lifted(c, 1.0, 3, "asan", 0, 4, cw, llr);
    CHECK(or_learned_decode(c, T, a, b, llr, post) >= 0, "or_learned_decode");
    const double loss = or_learned_loss_grad(c, T, a, b, llr, cw, ga, gb);
    CHECK(isfinite(loss) && loss > 0.0, "learned loss %g", loss);
    free(gb);
    free(ga);
    free(b);
    free(a);
    free(post);
    free(llr);
    free(cw);
    free(u);
    or_code_free(c);
}

static void link_and_statistics(void) {
    double re[64], im[64], metric[64], out[6], llr[6 * 8];
    uint8_t bits[6 * 8];
    or_rng r;
    or_rng_new(&r, 11);
    for (int m = 4; m <= 64; m *= 4) {
        or_qam_points(m, re, im);
        int bps = 0;
        while ((1 << bps) < m) ++bps;
        for (int i = 0; i < m; ++i) metric[i] = or_rng_uniform(&r);
        or_llr_from_metrics(metric, m, out);
        CHECK(fabs(or_label_halves_llr(metric, m, bps - 1) - out[0]) < 1e-12, "label halves, M = %d", m);
        for (int i = 0; i < bps * 8; ++i) bits[i] = or_rng_bit(&r);
        or_qam_awgn_llr(m, bits, bps * 8, 0.2, &r, llr);
        or_qam_llr(m, re, im, 8, 0.2, llr);
    }
    uint8_t c[16];
    double l16[16];
    for (int i = 0; i < 16; ++i) c[i] = (uint8_t)(i & 1);
    or_bpsk_awgn_llr(c, 16, 0.5, &r, l16);
    const uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6};
    uint32_t po[4];
    or_philox4x32_10(ctr, key, po);
    uint8_t info[100];
    or_philox_info_bits(1, 2, 3, 100, info);
    double g0, g1;
    or_philox_normals(1, 2, 3, 4, &g0, &g1);
    CHECK(isfinite(g0) && isfinite(g1), "philox normals");
    double lo, hi;
    or_wilson(0, 1000000, 0.95, &lo, &hi);
    CHECK(lo == 0.0 && hi > 0.0 && hi < 1e-5, "Wilson at zero errors");
    or_wilson(50, 1000, 0.95, &lo, &hi);
    CHECK(lo < 0.05 && hi > 0.05, "Wilson around 0.05");
    CHECK(fabs(or_erfc(0.0) - 1.0) < 1e-15 && fabs(or_erfcinv(1.0)) < 1e-12, "erfc");
    CHECK(fabs(or_qfuncinv(0.5)) < 1e-9 && fabs(or_z_score(0.95) - 1.959963984540054) < 1e-9, "z score");
    CHECK(or_db2lin(10.0) == 10.0 && or_bpsk_noise_variance(0.0, 0.5) == 1.0, "units");
    CHECK(or_ebn0_noise_variance(1.0, 0.0, 1.0, 1.0) == 1.0, "noise variance");
    uint64_t st = 1;
    (void)or_splitmix64(&st);
    (void)or_fnv1a64("asan", 4);
    or_stream(&r, 7, "asan", 1, 2);
    (void)or_rng_next(&r);
    (void)or_rng_gaussian(&r);
    or_rng_complex_gaussian(&r, &g0, &g1);
    for (uint32_t h = 0; h < 65536u; h += 257u) (void)or_f32_to_f16(or_f16_to_f32((uint16_t)h));
    /* a small code from rows, and the errors of bad sizes */
    const int32_t row_start[3] = {0, 3, 6}, cols[6] = {0, 1, 2, 2, 3, 4};
    or_code* t = or_code_from_rows(5, 2, row_start, cols);
    CHECK(t != NULL && or_code_info(t, 8) == 6, "or_code_from_rows");
    or_code_free(t);
    CHECK(or_nr_create(0, 100, 0, 0, 1) == NULL && or_last_error()[0] != '\0', "k_info 0 is refused");
    CHECK(or_nr_create(100, 301, 0, 0, 2) == NULL, "an E that Qm does not divide is refused");
    or_code* q3 = or_nr_create(100, 300, 0, 0, 3); /* any Qm that divides E (nr/ldpc.rs:103) */
    CHECK(q3 != NULL, "Qm 3 with E 300");
    if (q3) or_code_free(q3);
}

int main(void) {
    link_and_statistics();
    nr_code(100, 300, 0, 0, 1, 1);       /* the reference's doctest code */
    nr_code(500, 1200, 2, 2, 4, 2);      /* base graph 2, filler bits, rv 2 */
    nr_code(60, 1000, 0, 1, 2, 3);       /* E > N: repetition */
    nr_code(1000, 1500, 0, 3, 6, 2);
    nr_code(8448, 12672, 0, 2, 6, 4);    /* base graph 1, Z = 384 */
    nr_code(3840, 11520, 0, 0, 4, 2);
    lifted(2, 52);
    lifted(1, 384);
    lifted(1, 2);
    lifted(2, 15);
    if (failures) {
        fprintf(stderr, "asan_check: %d check(s) failed\n", failures);
        return 1;
    }
    printf("asan_check: ok\n");
    return 0;
}
