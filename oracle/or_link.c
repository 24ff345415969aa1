/*
 * Oracle (test infrastructure only): random streams, channel, LLRs and statistics,
 * restated from the reference in f64 with the platform libm (the libm Rust calls on Linux).
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <pthread.h>
#include "oracle.h"

/* crates/com-core/src/rng.rs:49-55 */
uint64_t or_splitmix64(uint64_t* state) {
    *state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* rng.rs:58-65 */
uint64_t or_fnv1a64(const char* bytes, size_t n) {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < n; ++i) {
        h ^= (uint8_t)bytes[i];
        h *= 0x00000100000001B3ULL;
    }
    return h;
}

/* rng.rs:75-79: four SplitMix64 outputs */
void or_rng_new(or_rng* r, uint64_t seed) {
    uint64_t st = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = or_splitmix64(&st);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* rng.rs:98-111 (xoshiro256**) */
uint64_t or_rng_next(or_rng* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* rng.rs:115-117 */
double or_rng_uniform(or_rng* r) {
    return (double)(or_rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

/* rng.rs:146-148 */
uint8_t or_rng_bit(or_rng* r) { return (uint8_t)(or_rng_next(r) >> 63); }

static const double OR_PI = 3.14159265358979323846;

/* rng.rs:152-157: Box-Muller, cosine branch */
double or_rng_gaussian(or_rng* r) {
    double u1 = or_rng_uniform(r);
    double u2 = or_rng_uniform(r);
    double rad = sqrt(-2.0 * log(1.0 - u1));
    return rad * cos(2.0 * OR_PI * u2);
}

/* rng.rs:161-167 */
void or_rng_complex_gaussian(or_rng* r, double* re, double* im) {
    double u1 = or_rng_uniform(r);
    double u2 = or_rng_uniform(r);
    double rad = sqrt(-2.0 * log(1.0 - u1)) * 0.70710678118654752440;
    double a = 2.0 * OR_PI * u2;
    double s = sin(a), c = cos(a);
    *re = rad * c;
    *im = rad * s;
}

/* rng.rs:270-278 */
void or_stream(or_rng* r, uint64_t base, const char* name, uint64_t cond, uint64_t trial) {
    uint64_t st = base ^ or_fnv1a64(name, strlen(name));
    uint64_t a = or_splitmix64(&st);
    st = a ^ cond;
    uint64_t b = or_splitmix64(&st);
    st = b ^ trial;
    uint64_t c = or_splitmix64(&st);
    or_rng_new(r, c);
}

/* crates/com-core/src/units.rs:38-40 */
double or_db2lin(double x_db) { return pow(10.0, x_db / 10.0); }

/* crates/com-fec/src/bounds.rs:58-60 */
double or_bpsk_noise_variance(double ebn0_db, double rate) {
    return 1.0 / (2.0 * rate * or_db2lin(ebn0_db));
}

/* crates/com-channel/src/awgn.rs:139-141 */
double or_ebn0_noise_variance(double p, double ebn0_db, double bits, double sps) {
    return p * sps / (bits * or_db2lin(ebn0_db));
}

/* validation/src/lib.rs:67-71 with bpsk (constellation.rs:149-151), bpsk_llr (demapper.rs:309-311) */
void or_bpsk_awgn_llr(const uint8_t* c, int64_t n, double noise_var, or_rng* r, double* llr) {
    double s = sqrt(noise_var);
    for (int64_t i = 0; i < n; ++i) {
        double x = 1.0 - 2.0 * (double)c[i];
        double y = x + s * or_rng_gaussian(r);
        llr[i] = 2.0 * y / noise_var;
    }
}

/* constellation.rs:168-205 (QAM natural order, serpentine rows, Gray labels) and
   normalized() (:396-398) with power = energy / M (com-math lib.rs:138-148) */
void or_qam_points(int m, double* re, double* im) {
    int l = (int)lround(sqrt((double)m));
    double* nat_re = (double*)malloc(sizeof(double) * m);
    double* nat_im = (double*)malloc(sizeof(double) * m);
    int p = 0;
    for (int row = 0; row < l; ++row) {
        double yv = 2.0 * (double)(l - 1 - row) - ((double)l - 1.0);
        for (int c = 0; c < l; ++c) {
            int cc = (row % 2 == 1) ? (l - 1 - c) : c;
            nat_re[p] = 2.0 * (double)cc - ((double)l - 1.0);
            nat_im[p] = yv;
            ++p;
        }
    }
    for (int i = 0; i < m; ++i) {
        int g = i ^ (i >> 1);
        re[g] = nat_re[i];
        im[g] = nat_im[i];
    }
    double e = 0.0;
    for (int i = 0; i < m; ++i) e += re[i] * re[i] + im[i] * im[i];
    double pw = e / (double)m;
    double s = 1.0 / sqrt(pw);
    for (int i = 0; i < m; ++i) {
        re[i] = re[i] * s;
        im[i] = im[i] * s;
    }
    free(nat_re);
    free(nat_im);
}

/* LabelHalves::new and LabelHalves::llr with Jacobian::MaxLog (demapper.rs:94-146): the
   largest metric of the labels whose bit `shift` is 0, minus the largest of those where it
   is 1 (f64::max, which ignores a NaN like fmax) */
double or_label_halves_llr(const double* metric, int m, int shift) {
    double m0 = -INFINITY, m1 = -INFINITY;
    for (int i = 0; i < m; ++i) {
        if (((i >> shift) & 1) == 0) m0 = fmax(m0, metric[i]);
        else m1 = fmax(m1, metric[i]);
    }
    return m0 - m1;
}

/* llr_from_metrics (demapper.rs:130-137), max-log: b = log2 m LLRs, the most significant
   bit first */
void or_llr_from_metrics(const double* metric, int m, double* out) {
    int b = 0;
    while ((1 << b) < m) ++b;
    for (int o = 0; o < b; ++o) out[o] = or_label_halves_llr(metric, m, b - 1 - o);
}

/* demapper.rs:270-297 (llr_kernel with the uniform prior -ln M added to every metric before
   the maxima) with llr_from_metrics in max-log */
void or_qam_llr(int m, const double* yre, const double* yim, int64_t nsym, double noise_var,
                double* llr) {
    double pre[256], pim[256], metric[256];
    int b = 0;
    while ((1 << b) < m) ++b;
    or_qam_points(m, pre, pim);
    double uniform = -log((double)m);
    for (int64_t n = 0; n < nsym; ++n) {
        for (int i = 0; i < m; ++i) {
            double dr = yre[n] - pre[i], di = yim[n] - pim[i];
            metric[i] = -(dr * dr + di * di) / noise_var + uniform;
        }
        or_llr_from_metrics(metric, m, llr + n * b);
    }
}

/* map_bits (constellation.rs:417-419, labels MSB first: bits.rs:68-70), add_noise (awgn.rs:94-99),
   then the max-log LLRs */
void or_qam_awgn_llr(int m, const uint8_t* bits, int64_t nbits, double noise_var, or_rng* r,
                     double* llr) {
    double pre[256], pim[256];
    int b = 0;
    while ((1 << b) < m) ++b;
    or_qam_points(m, pre, pim);
    int64_t nsym = nbits / b;
    double* yre = (double*)malloc(sizeof(double) * (size_t)nsym);
    double* yim = (double*)malloc(sizeof(double) * (size_t)nsym);
    double sd = sqrt(noise_var);
    for (int64_t s = 0; s < nsym; ++s) {
        int idx = 0;
        for (int k = 0; k < b; ++k) idx = (idx << 1) | (bits[s * b + k] & 1);
        double gr, gi;
        or_rng_complex_gaussian(r, &gr, &gi);
        yre[s] = pre[idx] + gr * sd;
        yim[s] = pim[idx] + gi * sd;
    }
    or_qam_llr(m, yre, yim, nsym, noise_var, llr);
    free(yre);
    free(yim);
}

/* crates/com-math/src/special.rs:180-255: Cody's rational Chebyshev erfc (CALERF) */
double or_erfc(double x) {
    static const double A[5] = {3.16112374387056560e00, 1.13864154151050156e02,
                                3.77485237685302021e02, 3.20937758913846947e03,
                                1.85777706184603153e-1};
    static const double B[4] = {2.36012909523441209e01, 2.44024637934444173e02,
                                1.28261652607737228e03, 2.84423683343917062e03};
    static const double C[9] = {5.64188496988670089e-1, 8.88314979438837594e00,
                                6.61191906371416295e01, 2.98635138197400131e02,
                                8.81952221241769090e02, 1.71204761263407058e03,
                                2.05107837782607147e03, 1.23033935479799725e03,
                                2.15311535474403846e-8};
    static const double D[8] = {1.57449261107098347e01, 1.17693950891312499e02,
                                5.37181101862009858e02, 1.62138957456669019e03,
                                3.29079923573345963e03, 4.36261909014324716e03,
                                3.43936767414372164e03, 1.23033935480374942e03};
    static const double P[6] = {3.05326634961232344e-1, 3.60344899949804439e-1,
                                1.25781726111229246e-1, 1.60837851487422766e-2,
                                6.58749161529837803e-4, 1.63153871373020978e-2};
    static const double Q[5] = {2.56852019228982242e00, 1.87295284992346725e00,
                                5.27905102951428412e-1, 6.05183413124413191e-2,
                                2.33520497626869185e-3};
    const double FRAC_1_SQRT_PI = 5.6418958354775628695e-1;
    double y = fabs(x);
    if (y <= 0.46875) {
        double ysq = y > 1.11e-16 ? y * y : 0.0;
        double num = A[4] * ysq, den = ysq;
        for (int i = 0; i < 3; ++i) {
            num = (num + A[i]) * ysq;
            den = (den + B[i]) * ysq;
        }
        return 1.0 - x * (num + A[3]) / (den + B[3]);
    }
    double r;
    if (y > 27.3) {
        r = 0.0;
    } else {
        double rr;
        if (y <= 4.0) {
            double num = C[8] * y, den = y;
            for (int i = 0; i < 7; ++i) {
                num = (num + C[i]) * y;
                den = (den + D[i]) * y;
            }
            rr = (num + C[7]) / (den + D[7]);
        } else {
            double ysq = 1.0 / (y * y);
            double num = P[5] * ysq, den = ysq;
            for (int i = 0; i < 4; ++i) {
                num = (num + P[i]) * ysq;
                den = (den + Q[i]) * ysq;
            }
            rr = (FRAC_1_SQRT_PI - ysq * (num + P[4]) / (den + Q[4])) / y;
        }
        double u = trunc(y * 16.0) / 16.0;
        r = exp(-u * u) * exp(-(y - u) * (y + u)) * rr;
    }
    return x < 0.0 ? 2.0 - r : r;
}

/* special.rs:265-273 with search.rs:69-80 (bisection, 200 steps on [-27, 27]) */
double or_erfcinv(double y) {
    if (y <= 0.0) return INFINITY;
    if (y >= 2.0) return -INFINITY;
    double lo = -27.0, hi = 27.0;
    for (int i = 0; i < 200; ++i) {
        double mid = 0.5 * (lo + hi);
        if (or_erfc(mid) > y) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* special.rs:291-293 */
double or_qfuncinv(double p) { return 1.41421356237309504880 * or_erfcinv(2.0 * p); }

/* crates/com-sim/src/stats.rs:43-45 */
double or_z_score(double confidence) {
    double c = confidence;
    double top = 1.0 - 1e-16;
    if (c < 0.0) c = 0.0;
    if (c > top) c = top;
    return or_qfuncinv((1.0 - c) / 2.0);
}

/* stats.rs:74-87 */
void or_wilson(uint64_t errors, uint64_t total, double confidence, double* lo, double* hi) {
    if (total == 0) {
        *lo = 0.0;
        *hi = 1.0;
        return;
    }
    double n = (double)total;
    double p = (double)errors / (double)total;
    double z = or_z_score(confidence);
    double z2 = z * z;
    double centre = (p + z2 / (2.0 * n)) / (1.0 + z2 / n);
    double half = z * sqrt(p * (1.0 - p) / n + z2 / (4.0 * n * n)) / (1.0 + z2 / n);
    double lower = errors == 0 ? 0.0 : fmax(centre - half, 0.0);
    *lo = lower;
    *hi = fmin(centre + half, 1.0);
}

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): the counter-based generator of the
   device trial generator (no reference counterpart), restated for its known-answer tests
   and to regenerate the device's information bits. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* The K' information bits of trial `trial` at grid point `point` of the device generator:
   block b of 128 bits is Philox(counter = (b, trial low, trial high, point*16 + 0), seed). */
void or_philox_info_bits(uint64_t seed, uint32_t point, uint64_t trial, int32_t k_info, uint8_t* bits) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int32_t words = (k_info + 31) / 32;
    for (int32_t blk = 0; blk * 4 < words; ++blk) {
        uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)trial, (uint32_t)(trial >> 32), point << 4};
        uint32_t out[4];
        or_philox4x32_10(ctr, key, out);
        for (int q = 0; q < 4; ++q)
            for (int b = 0; b < 32; ++b) {
                int32_t p = (blk * 4 + q) * 32 + b;
                if (p < k_info) bits[p] = (uint8_t)((out[q] >> b) & 1u);
            }
    }
}

/* The channel noise of the device trial generator (no reference counterpart), restated in
   f64: draw `index` of purpose 1 of the trial is Philox(counter = (index, trial low, trial
   high, point*16 + 1), seed); u1, u2 are its first two words as (x >> 8) + 1 over 2^24 (on
   (0, 1]), and Box-Muller gives g0 = sqrt(-2 ln u1) cos(2 pi u2), g1 = ... sin(2 pi u2). */
void or_philox_normals(uint64_t seed, uint32_t point, uint64_t trial, uint32_t index, double* g0,
                       double* g1) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {index, (uint32_t)trial, (uint32_t)(trial >> 32), (point << 4) | 1u};
    uint32_t out[4];
    or_philox4x32_10(ctr, key, out);
    double u1 = (double)((out[0] >> 8) + 1u) / 16777216.0;
    double u2 = (double)((out[1] >> 8) + 1u) / 16777216.0;
    double rad = sqrt(-2.0 * log(u1));
    *g0 = rad * cos(2.0 * OR_PI * u2);
    *g1 = rad * sin(2.0 * OR_PI * u2);
}

/* One trial of the device trial generator on an NR code, in f64: the Philox information bits,
   NrLdpc::encode_block (nr/ldpc.rs:208-214), then for bps == 1 BPSK y = (1 - 2c) + sd g0 with
   sd = sqrt(noise_var) and LLR 2y/noise_var (validation/src/lib.rs:67-71), or Gray QAM of 2^bps
   points with the labels taken MSB first from consecutive transmitted bits (bits.rs:68-70),
   y = p + sqrt(noise_var / 2) (g0 + j g1) (noise_var = N0 per symbol, awgn.rs:94-99) and the
   max-log LLRs of demapper.rs:270-297. yre, yim (may be NULL) get the received symbols. */
void or_philox_trial(const or_code* c, int bps, double noise_var, uint64_t seed, uint32_t point,
                     uint64_t trial, uint8_t* info, double* yre, double* yim, double* llr) {
    or_philox_info_bits(seed, point, trial, (int32_t)or_code_info(c, 3), info);
    int32_t e = (int32_t)or_code_info(c, 4);
    uint8_t* tx = (uint8_t*)malloc((size_t)e);
    or_nr_encode(c, info, tx);
    int32_t nsym = e / bps;
    if (bps == 1) {
        double sd = sqrt(noise_var);
        for (int32_t s = 0; s < nsym; ++s) {
            double g0, g1;
            or_philox_normals(seed, point, trial, (uint32_t)s, &g0, &g1);
            double y = (1.0 - 2.0 * (double)tx[s]) + sd * g0;
            if (yre) yre[s] = y;
            if (yim) yim[s] = 0.0;
            llr[s] = 2.0 * y / noise_var;
        }
    } else {
        int m = 1 << bps;
        double pre[256], pim[256];
        or_qam_points(m, pre, pim);
        double sd = sqrt(noise_var / 2.0);
        double* yr = (double*)malloc(sizeof(double) * (size_t)nsym);
        double* yi = (double*)malloc(sizeof(double) * (size_t)nsym);
        for (int32_t s = 0; s < nsym; ++s) {
            int label = 0;
            for (int l = 0; l < bps; ++l) label = (label << 1) | (tx[s * bps + l] & 1);
            double g0, g1;
            or_philox_normals(seed, point, trial, (uint32_t)s, &g0, &g1);
            yr[s] = pre[label] + sd * g0;
            yi[s] = pim[label] + sd * g1;
            if (yre) yre[s] = yr[s];
            if (yim) yim[s] = yi[s];
        }
        or_qam_llr(m, yr, yi, nsym, noise_var, llr);
        free(yr);
        free(yi);
    }
    free(tx);
}

/* or_trial_nr_qam for trials trial0 .. trial0 + count - 1 on `threads` POSIX threads: the
   information bits count x k_info, the LLRs count x e. */
typedef struct {
    const or_code* c;
    int m;
    double noise_var;
    uint64_t seed;
    const char* name;
    uint64_t cond, trial0;
    int64_t lo, hi;
    uint8_t* info;
    double* llr;
} qam_job;

static void* qam_worker(void* p) {
    qam_job* j = (qam_job*)p;
    int64_t k = or_code_info(j->c, 3), e = or_code_info(j->c, 4);
    for (int64_t t = j->lo; t < j->hi; ++t)
        or_trial_nr_qam(j->c, j->m, j->noise_var, j->seed, j->name, j->cond, j->trial0 + (uint64_t)t,
                        j->info + t * k, j->llr + t * e);
    return NULL;
}

void or_trial_nr_qam_many(const or_code* c, int m, double noise_var, uint64_t seed, const char* name,
                          uint64_t cond, uint64_t trial0, int64_t count, uint8_t* info, double* llr,
                          int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    qam_job jobs[256];
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (qam_job){c, m, noise_var, seed, name, cond, trial0, count * t / threads,
                            count * (t + 1) / threads, info, llr};
        if (threads == 1) qam_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, qam_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}
