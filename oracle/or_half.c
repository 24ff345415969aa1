/*
 * Oracle (test infrastructure only): the layered min-sum decoder of
 * crates/com-fec/src/ldpc.rs:468-491 (check node), :509-531 (layered propagate) and
 * :546-557 (syndrome stop), with every stored message in IEEE binary16 and the arithmetic
 * in FP32 — the device decoder's FP16 storage mode, restated edge by edge.
 *
 * Storage rules (the definition of the FP16 mode; the device kernel follows them):
 *   - a posterior is stored as half(clamp(x, -65504, 65504)) rounded to nearest even, and
 *     a zero is stored as +0.0 (so a stored posterior is never -0.0, and its sign bit is
 *     the reference's test x < 0.0);
 *   - a check-to-variable magnitude g(m) is stored as half(min(g, 65504)), and the message
 *     is that magnitude with the sign of the reference's product;
 *   - v2c = float(post) - float(c2v) and post' = v2c + float(c2v') are FP32 operations,
 *     and g(m) = (float)alpha * m (normalized) or fmaxf(m - (float)beta, 0) (offset).
 * Everything else is the reference's order: rows in lifted order, edges in table order,
 * the strict-< two-minimum scan, the syndrome over all rows after every iteration.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "or_internal.h"

/* IEEE 754 binary32 -> binary16, round to nearest, ties to even. */
uint16_t or_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);  /* NaN */
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520: infinity */
    const int exp = (int)(ax >> 23);
    if (exp >= 113) { /* |f| >= 2^-14: a normal half */
        const uint32_t mant = ax & 0x7fffffu;
        uint32_t h = ((uint32_t)(exp - 112) << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++; /* a carry moves into the exponent */
        return (uint16_t)(sign | h);
    }
    /* a subnormal half: round(|f| / 2^-24) = mant * 2^(exp - 126), mant with its leading 1 */
    const int shift = 126 - exp;
    if (exp == 0 || shift > 24) return sign; /* |f| <= 2^-25: rounds to zero */
    const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float or_f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        const float v = (float)m * 5.9604644775390625e-08f; /* m * 2^-24, exact */
        return sign ? -v : v;
    }
    if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else x = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    memcpy(&f, &x, 4);
    return f;
}

static uint16_t store_post(float x) {
    if (x > 65504.0f) x = 65504.0f;
    if (x < -65504.0f) x = -65504.0f;
    const uint16_t h = or_f32_to_f16(x);
    return (h & 0x7fffu) ? h : 0;
}

static float store_mag(float g) { return or_f16_to_f32(or_f32_to_f16(g > 65504.0f ? 65504.0f : g)); }

static int is_codeword_h(const or_code* c, const uint16_t* post) {
    for (int32_t r = 0; r < c->m; ++r) {
        unsigned a = 0;
        for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e) a ^= (post[c->cols[e]] >> 15) & 1u;
        if (a) return 0;
    }
    return 1;
}

int or_decode_llr_f16(const or_code* c, const or_dec* d, const float* llr, float* post_out) {
    const int32_t n = c->n, edges = c->row_start[c->m];
    uint16_t* post = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)n);
    float* c2v = (float*)calloc((size_t)edges, sizeof(float));
    int dm = 1;
    for (int32_t r = 0; r < c->m; ++r)
        if (c->row_start[r + 1] - c->row_start[r] > dm) dm = c->row_start[r + 1] - c->row_start[r];
    float* in = (float*)malloc(sizeof(float) * (size_t)dm);
    for (int32_t p = 0; p < n; ++p) post[p] = store_post(llr[p]);
    int result = 0;
    if (!is_codeword_h(c, post)) {
        result = d->max_iterations;
        const float alpha = (float)d->alpha;
        for (int it = 1; it <= d->max_iterations; ++it) {
            for (int32_t r = 0; r < c->m; ++r) {
                const int32_t s = c->row_start[r];
                const int deg = c->row_start[r + 1] - s;
                float sign = 1.0f, m1 = INFINITY, m2 = INFINITY;
                int i1 = 0;
                for (int i = 0; i < deg; ++i) {
                    in[i] = or_f16_to_f32(post[c->cols[s + i]]) - c2v[s + i];
                    sign *= in[i] < 0.0f ? -1.0f : 1.0f;
                }
                for (int i = 0; i < deg; ++i) {
                    const float a = fabsf(in[i]);
                    if (a < m1) {
                        m2 = m1;
                        m1 = a;
                        i1 = i;
                    } else if (a < m2) {
                        m2 = a;
                    }
                }
                float g1, g2;
                if (d->rule == OR_NMS) {
                    g1 = alpha * m1;
                    g2 = alpha * m2;
                } else {
                    g1 = fmaxf(m1 - alpha, 0.0f);
                    g2 = fmaxf(m2 - alpha, 0.0f);
                }
                g1 = store_mag(g1);
                g2 = store_mag(g2);
                for (int i = 0; i < deg; ++i) {
                    const float s_i = sign * (in[i] < 0.0f ? -1.0f : 1.0f);
                    c2v[s + i] = s_i * (i == i1 ? g2 : g1);
                    post[c->cols[s + i]] = store_post(in[i] + c2v[s + i]);
                }
            }
            if (is_codeword_h(c, post)) {
                result = it;
                break;
            }
        }
    }
    if (post_out)
        for (int32_t p = 0; p < n; ++p) post_out[p] = or_f16_to_f32(post[p]);
    free(post);
    free(c2v);
    free(in);
    return result;
}

typedef struct {
    const or_code* c;
    const or_dec* d;
    const float* llr;
    const int32_t* max_it;
    float* post;
    int32_t* its;
    int64_t begin, end;
} half_job;

static void* half_worker(void* arg) {
    half_job* j = (half_job*)arg;
    const int32_t n = j->c->n;
    for (int64_t w = j->begin; w < j->end; ++w) {
        or_dec d = *j->d;
        if (j->max_it) d.max_iterations = j->max_it[w];
        j->its[w] = or_decode_llr_f16(j->c, &d, j->llr + w * n, j->post + w * n);
    }
    return NULL;
}

void or_decode_many_f16(const or_code* c, const or_dec* d, const float* llr, int64_t count,
                        const int32_t* max_iterations, float* post, int32_t* iterations, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    half_job jobs[256];
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (half_job){c, d, llr, max_iterations, post, iterations, count * t / threads,
                             count * (t + 1) / threads};
        if (threads == 1) half_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, half_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}
