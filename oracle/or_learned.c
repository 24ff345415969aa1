/*
 * Oracle (test infrastructure only): the learned offset/scale min-sum, in f64.
 *
 * No reference counterpart (crates/com-fec has no trainable decoder, README.md:109-111).
 * It is the layered schedule of crates/com-fec/src/ldpc.rs:509-531 on a lifted NR code,
 * with the check node of ldpc.rs:468-491 whose magnitude map is, for iteration t and base
 * row i, g(m) = a[t][i] * max(m - b[t][i], 0): a = alpha, b = 0 is the normalized min-sum
 * and a = 1, b = beta the offset min-sum, bit for bit.
 *
 * Training loss: the mean over the T iterations and the N codeword bits of the binary
 * cross-entropy of the soft output sigmoid(-post) against the bit,
 *     BCE = softplus(-s post), s = 1 - 2c,
 * evaluated on the posteriors at the end of every iteration (no early stop). The backward
 * pass below is written by hand (reverse layers, reverse iterations); its pin is the
 * central finite difference of the loss (tests/test_oracle_learned.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#include "or_internal.h"

typedef struct {
    double m1, m2;
    int i1, i2;
    int neg;        /* parity of the negative inputs */
} row_min;

/* one scan with strict <, as ldpc.rs:471-480; i2 follows the second minimum */
static row_min scan(const double* in, int deg) {
    row_min s = {INFINITY, INFINITY, 0, 0, 0};
    for (int q = 0; q < deg; ++q) {
        double a = fabs(in[q]);
        s.neg ^= in[q] < 0.0;
        if (a < s.m1) {
            s.m2 = s.m1;
            s.i2 = s.i1;
            s.m1 = a;
            s.i1 = q;
        } else if (a < s.m2) {
            s.m2 = a;
            s.i2 = q;
        }
    }
    return s;
}

static double g_of(double m, double a, double b) { return a * fmax(m - b, 0.0); }

/* out_q = sign_q * g(q == i1 ? m2 : m1), sign_q the product of the other signs */
static void outputs(const double* in, int deg, const row_min* s, double a, double b, double* out) {
    double g1 = g_of(s->m1, a, b), g2 = g_of(s->m2, a, b);
    for (int q = 0; q < deg; ++q) {
        double mag = q == s->i1 ? g2 : g1;
        int neg = s->neg ^ (in[q] < 0.0);
        out[q] = neg ? -mag : mag;
    }
}

static int syndrome_ok(const or_code* c, const double* post) {
    for (int32_t r = 0; r < c->m; ++r) {
        int p = 0;
        for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e) p ^= post[c->cols[e]] < 0.0;
        if (p) return 0;
    }
    return 1;
}

/* Inference: layered iterations t = 0..T-1 with the syndrome stop of ldpc.rs:510-512, 546-548.
   Returns the iterations used (0 when the input is a codeword). */
int or_learned_decode(const or_code* c, int iterations, const double* alpha, const double* beta,
                      const double* llr, double* post) {
    int32_t edges = c->row_start[c->m];
    int rows_b = c->m / c->z;
    memcpy(post, llr, sizeof(double) * c->n);
    if (syndrome_ok(c, post)) return 0;
    double* c2v = (double*)calloc((size_t)edges, sizeof(double));
    double in[32], out[32];
    int used = iterations;
    for (int t = 0; t < iterations; ++t) {
        for (int32_t r = 0; r < c->m; ++r) {
            int i = r / c->z;
            double a = alpha[t * rows_b + i], b = beta[t * rows_b + i];
            int32_t s0 = c->row_start[r];
            int deg = c->row_start[r + 1] - s0;
            for (int q = 0; q < deg; ++q) in[q] = post[c->cols[s0 + q]] - c2v[s0 + q];
            row_min s = scan(in, deg);
            outputs(in, deg, &s, a, b, out);
            for (int q = 0; q < deg; ++q) {
                c2v[s0 + q] = out[q];
                post[c->cols[s0 + q]] = in[q] + out[q];
            }
        }
        if (syndrome_ok(c, post)) {
            used = t + 1;
            break;
        }
    }
    free(c2v);
    return used;
}

static double softplus(double x) { return x > 0.0 ? x + log1p(exp(-x)) : log1p(exp(x)); }
static double sigmoid(double x) { return x >= 0.0 ? 1.0 / (1.0 + exp(-x)) : exp(x) / (1.0 + exp(x)); }

double or_learned_loss_grad(const or_code* c, int iterations, const double* alpha, const double* beta,
                            const double* llr, const uint8_t* cw, double* g_alpha, double* g_beta) {
    const int32_t edges = c->row_start[c->m];
    const int32_t n = c->n;
    const int rows_b = c->m / c->z;
    const int T = iterations;
    const double scale = 1.0 / ((double)T * (double)n);
    /* forward, keeping every iteration's variable-to-check messages and posteriors */
    double* ckpt = (double*)malloc(sizeof(double) * (size_t)T * edges);
    double* posts = (double*)malloc(sizeof(double) * (size_t)T * n);
    double* post = (double*)malloc(sizeof(double) * n);
    double* c2v = (double*)calloc((size_t)edges, sizeof(double));
    double in[32], out[32];
    memcpy(post, llr, sizeof(double) * n);
    double loss = 0.0;
    for (int t = 0; t < T; ++t) {
        for (int32_t r = 0; r < c->m; ++r) {
            int i = r / c->z;
            double a = alpha[t * rows_b + i], b = beta[t * rows_b + i];
            int32_t s0 = c->row_start[r];
            int deg = c->row_start[r + 1] - s0;
            for (int q = 0; q < deg; ++q) {
                in[q] = post[c->cols[s0 + q]] - c2v[s0 + q];
                ckpt[(size_t)t * edges + s0 + q] = in[q];
            }
            row_min s = scan(in, deg);
            outputs(in, deg, &s, a, b, out);
            for (int q = 0; q < deg; ++q) {
                c2v[s0 + q] = out[q];
                post[c->cols[s0 + q]] = in[q] + out[q];
            }
        }
        memcpy(posts + (size_t)t * n, post, sizeof(double) * n);
        for (int32_t v = 0; v < n; ++v) loss += scale * softplus(-(1.0 - 2.0 * cw[v]) * post[v]);
    }
    /* backward */
    double* pbar = (double*)calloc((size_t)n, sizeof(double));
    double* cbar = (double*)calloc((size_t)edges, sizeof(double));
    for (int p = 0; p < T * rows_b; ++p) g_alpha[p] = g_beta[p] = 0.0;
    for (int t = T - 1; t >= 0; --t) {
        for (int32_t v = 0; v < n; ++v) {
            double s = 1.0 - 2.0 * cw[v];
            pbar[v] += scale * (-s) * sigmoid(-s * posts[(size_t)t * n + v]);
        }
        for (int32_t r = c->m - 1; r >= 0; --r) {
            int i = r / c->z;
            double a = alpha[t * rows_b + i], b = beta[t * rows_b + i];
            int32_t s0 = c->row_start[r];
            int deg = c->row_start[r + 1] - s0;
            for (int q = 0; q < deg; ++q) in[q] = ckpt[(size_t)t * edges + s0 + q];
            row_min s = scan(in, deg);
            double obar[32], ibar[32];
            double m1bar = 0.0, m2bar = 0.0;
            for (int q = 0; q < deg; ++q) {
                int v = c->cols[s0 + q];
                obar[q] = pbar[v] + (t < T - 1 ? cbar[s0 + q] : 0.0);
                double m = q == s.i1 ? s.m2 : s.m1;
                double sign = (s.neg ^ (in[q] < 0.0)) ? -1.0 : 1.0;
                double active = m > b ? 1.0 : 0.0;
                g_alpha[t * rows_b + i] += obar[q] * sign * fmax(m - b, 0.0);
                g_beta[t * rows_b + i] += obar[q] * sign * (-a) * active;
                if (q == s.i1) m2bar += obar[q] * sign * a * active;
                else m1bar += obar[q] * sign * a * active;
                ibar[q] = pbar[v];
            }
            /* d|x|/dx = sign(x), 0 at 0 */
            ibar[s.i1] += m1bar * (in[s.i1] > 0.0 ? 1.0 : in[s.i1] < 0.0 ? -1.0 : 0.0);
            ibar[s.i2] += m2bar * (in[s.i2] > 0.0 ? 1.0 : in[s.i2] < 0.0 ? -1.0 : 0.0);
            for (int q = 0; q < deg; ++q) {
                pbar[c->cols[s0 + q]] = ibar[q];
                cbar[s0 + q] = -ibar[q];
            }
        }
    }
    free(ckpt);
    free(posts);
    free(post);
    free(c2v);
    free(pbar);
    free(cbar);
    return loss;
}
