/*
 * Oracle (test infrastructure only): LDPC codes of 5G NR and the belief-propagation
 * decoder, restated in f64 from crates/com-fec/src/ldpc.rs and crates/com-fec/src/nr/ldpc.rs.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "../com-rs_amd/csrc/nr_base_graphs.inc"

#include "or_internal.h"

static __thread char g_err[256];
const char* or_last_error(void) { return g_err; }

static void set_err(const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); }

void or_code_free(or_code* c) {
    if (!c) return;
    free(c->row_start);
    free(c->cols);
    free(c->solve_start);
    free(c->solve);
    free(c->map);
    free(c);
}

or_code* or_code_from_rows(int32_t n, int32_t m, const int32_t* row_start, const int32_t* cols) {
    or_code* c = (or_code*)calloc(1, sizeof(or_code));
    c->n = n;
    c->m = m;
    c->row_start = (int32_t*)malloc(sizeof(int32_t) * (m + 1));
    memcpy(c->row_start, row_start, sizeof(int32_t) * (m + 1));
    c->cols = (int32_t*)malloc(sizeof(int32_t) * row_start[m]);
    memcpy(c->cols, cols, sizeof(int32_t) * row_start[m]);
    return c;
}

/* The rows of the lifted base graph: rows[i Z + r] gets j Z + (r + V mod Z) mod Z for every
   entry (i, j, V) in table order (nr/ldpc.rs:124-133). */
static void lift(or_code* c, int bg, int32_t z, int ils) {
    const nr_bg_entry* t = bg == 1 ? NR_BG1 : NR_BG2;
    int ne = bg == 1 ? NR_BG1_ENTRIES : NR_BG2_ENTRIES;
    int rows_n = bg == 1 ? 46 : 42;
    int cols_n = bg == 1 ? 68 : 52;
    int deg[46] = {0};
    for (int e = 0; e < ne; ++e) deg[t[e].i]++;
    c->m = rows_n * z;
    c->n = cols_n * z;
    c->row_start = (int32_t*)malloc(sizeof(int32_t) * (c->m + 1));
    c->row_start[0] = 0;
    for (int i = 0; i < rows_n; ++i)
        for (int32_t r = 0; r < z; ++r) {
            int32_t row = i * z + r;
            c->row_start[row + 1] = c->row_start[row] + deg[i];
        }
    c->cols = (int32_t*)malloc(sizeof(int32_t) * c->row_start[c->m]);
    int32_t* fill = (int32_t*)calloc(c->m, sizeof(int32_t));
    for (int e = 0; e < ne; ++e) {
        int32_t shift = (int32_t)(t[e].v[ils] % z);
        for (int32_t r = 0; r < z; ++r) {
            int32_t row = t[e].i * z + r;
            c->cols[c->row_start[row] + fill[row]++] = t[e].j * z + (r + shift) % z;
        }
    }
    free(fill);
}

/* The inverse of the core block B (core checks x core parity bits k..k+core) over GF(2),
   row j listing the checks whose syndromes sum to parity bit k + j (ldpc.rs:568-584). The
   inverse is unique, so a plain Gauss-Jordan on [B | I] gives the reference's rows. */
static int core_inverse(or_code* c) {
    int32_t core = c->core, k = c->k;
    int words = (2 * core + 63) / 64;
    uint64_t* a = (uint64_t*)calloc((size_t)core * words, sizeof(uint64_t));
    for (int32_t r = 0; r < core; ++r) {
        uint64_t* row = a + (size_t)r * words;
        for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e) {
            int32_t v = c->cols[e];
            if (v >= k) {
                int32_t col = v - k;
                row[col / 64] ^= 1ULL << (col % 64);
            }
        }
        int32_t col = core + r;
        row[col / 64] ^= 1ULL << (col % 64);
    }
    for (int32_t col = 0; col < core; ++col) {
        int32_t piv = -1;
        for (int32_t r = col; r < core; ++r)
            if ((a[(size_t)r * words + col / 64] >> (col % 64)) & 1) { piv = r; break; }
        if (piv < 0) { free(a); set_err("the core block is singular"); return -1; }
        if (piv != col)
            for (int w = 0; w < words; ++w) {
                uint64_t tmp = a[(size_t)piv * words + w];
                a[(size_t)piv * words + w] = a[(size_t)col * words + w];
                a[(size_t)col * words + w] = tmp;
            }
        for (int32_t r = 0; r < core; ++r)
            if (r != col && ((a[(size_t)r * words + col / 64] >> (col % 64)) & 1))
                for (int w = 0; w < words; ++w) a[(size_t)r * words + w] ^= a[(size_t)col * words + w];
    }
    /* row j of [I | B^-1]: the checks of parity bit j */
    c->solve_start = (int32_t*)malloc(sizeof(int32_t) * (core + 1));
    int64_t total = 0;
    for (int32_t j = 0; j < core; ++j)
        for (int32_t r = 0; r < core; ++r)
            total += (a[(size_t)j * words + (core + r) / 64] >> ((core + r) % 64)) & 1;
    c->solve = (int32_t*)malloc(sizeof(int32_t) * (total ? total : 1));
    int64_t p = 0;
    for (int32_t j = 0; j < core; ++j) {
        c->solve_start[j] = (int32_t)p;
        for (int32_t r = 0; r < core; ++r)
            if ((a[(size_t)j * words + (core + r) / 64] >> ((core + r) % 64)) & 1) c->solve[p++] = r;
    }
    c->solve_start[core] = (int32_t)p;
    free(a);
    return 0;
}

static int lifting_set_of(int32_t z) {
    for (int s = 0; s < 8; ++s)
        for (int q = 0; q < NR_LIFTING_SET_MAX; ++q)
            if (NR_LIFTING_SETS[s][q] == z) return s;
    return -1;
}

or_code* or_code_lifted(int bg, int32_t z) {
    int ils = lifting_set_of(z);
    if ((bg != 1 && bg != 2) || ils < 0) { set_err("not a base graph and lifting size of NR"); return NULL; }
    or_code* c = (or_code*)calloc(1, sizeof(or_code));
    c->bg = bg;
    c->z = z;
    c->ils = ils;
    lift(c, bg, z, ils);
    c->k = (bg == 1 ? 22 : 10) * z;
    c->k_info = c->k;
    c->core = 4 * z;
    c->e = c->n - 2 * z;
    if (core_inverse(c)) { or_code_free(c); return NULL; }
    /* transmitted bits: every column from 2Z, once */
    c->map = (int32_t*)malloc(sizeof(int32_t) * c->e);
    for (int32_t t = 0; t < c->e; ++t) c->map[t] = 2 * z + t;
    return c;
}

/* nr/ldpc.rs:78-81 */
static int nr_base_graph(int32_t k_info, int32_t e) {
    double r = (double)k_info / (double)e;
    return (k_info <= 292 || (k_info <= 3824 && r <= 0.67) || r <= 0.25) ? 2 : 1;
}

/* nr/ldpc.rs:101-162 */
or_code* or_nr_create(int32_t k_info, int32_t e, int bg_override, int rv, int32_t qm) {
    if (k_info < 1 || e < 1) { set_err("need K' >= 1 and E >= 1"); return NULL; }
    if (rv < 0 || rv > 3 || qm < 1 || e % qm != 0) { set_err("rv <= 3 and Qm | E"); return NULL; }
    int bg = bg_override ? bg_override : nr_base_graph(k_info, e);
    if (bg != 1 && bg != 2) { set_err("the base graph is 1 or 2"); return NULL; }
    int32_t k_cb = bg == 1 ? 8448 : 3840;
    if (k_info > k_cb) { set_err("the block exceeds one code block"); return NULL; }
    int kb = bg == 1 ? 22 : (k_info > 640 ? 10 : k_info > 560 ? 9 : k_info > 192 ? 8 : 6);
    int32_t z = 0;
    for (int s = 0; s < 8; ++s)
        for (int q = 0; q < NR_LIFTING_SET_MAX; ++q) {
            int32_t zz = NR_LIFTING_SETS[s][q];
            if (zz && kb * zz >= k_info && (z == 0 || zz < z)) z = zz;
        }
    if (!z) { set_err("no lifting size"); return NULL; }
    or_code* c = or_code_lifted(bg, z);
    if (!c) return NULL;
    c->k_info = k_info;
    c->e = e;
    int32_t ncb = c->n - 2 * z;
    int num;
    if (rv == 0) num = 0;
    else if (bg == 1) num = rv == 1 ? 17 : rv == 2 ? 33 : 56;
    else num = rv == 1 ? 13 : rv == 2 ? 25 : 43;
    int den = bg == 1 ? 66 : 50;
    int64_t k0 = (int64_t)num * ncb / ((int64_t)den * z) * z;
    int32_t* sel = (int32_t*)malloc(sizeof(int32_t) * e);
    int32_t len = 0;
    for (int64_t j = 0; len < e; ++j) {
        int32_t col = 2 * z + (int32_t)((k0 + j) % ncb);
        if (!(col >= k_info && col < c->k)) sel[len++] = col;
    }
    free(c->map);
    c->map = (int32_t*)malloc(sizeof(int32_t) * e);
    /* Interleaver::block(qm, e/qm) (interleaving.rs:58-60): output t takes input
       (t mod qm) (e/qm) + t / qm */
    int32_t cols = e / qm;
    for (int32_t t = 0; t < e; ++t) c->map[t] = sel[(t % qm) * cols + t / qm];
    free(sel);
    return c;
}

int64_t or_code_info(const or_code* c, int what) {
    switch (what) {
        case 0: return c->n;
        case 1: return c->m;
        case 2: return c->k;
        case 3: return c->k_info;
        case 4: return c->e;
        case 5: return c->bg;
        case 6: return c->z;
        case 7: return c->ils;
        case 8: return c->row_start[c->m];
        default: return -1;
    }
}
const int32_t* or_code_row_start(const or_code* c) { return c->row_start; }
const int32_t* or_code_cols(const or_code* c) { return c->cols; }
const int32_t* or_code_bit_map(const or_code* c) { return c->map; }

/* ldpc.rs:753-791, the Encoder::Core branch */
int or_encode(const or_code* c, const uint8_t* u, uint8_t* cw) {
    if (!c->solve) { set_err("no encoder"); return -1; }
    int32_t k = c->k;
    memset(cw, 0, (size_t)c->n);
    for (int32_t i = 0; i < k; ++i) cw[i] = u[i] & 1;
    uint8_t* syn = (uint8_t*)malloc((size_t)c->core);
    for (int32_t r = 0; r < c->core; ++r) {
        uint8_t a = 0;
        for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e)
            if (c->cols[e] < k) a ^= cw[c->cols[e]];
        syn[r] = a;
    }
    for (int32_t j = 0; j < c->core; ++j) {
        uint8_t a = 0;
        for (int32_t p = c->solve_start[j]; p < c->solve_start[j + 1]; ++p) a ^= syn[c->solve[p]];
        cw[k + j] = a;
    }
    for (int32_t r = c->core; r < c->m; ++r) {
        uint8_t a = 0;
        for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e)
            if (c->cols[e] != k + r) a ^= cw[c->cols[e]];
        cw[k + r] = a;
    }
    free(syn);
    return 0;
}

int or_is_codeword(const or_code* c, const uint8_t* cw) {
    for (int32_t r = 0; r < c->m; ++r) {
        uint8_t a = 0;
        for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e) a ^= cw[c->cols[e]];
        if (a) return 0;
    }
    return 1;
}

/* nr/ldpc.rs:208-214: fillers are zeros, then the bit selection */
int or_nr_encode(const or_code* c, const uint8_t* info, uint8_t* out) {
    uint8_t* u = (uint8_t*)calloc((size_t)c->k, 1);
    memcpy(u, info, (size_t)c->k_info);
    uint8_t* cw = (uint8_t*)malloc((size_t)c->n);
    int rc = or_encode(c, u, cw);
    if (!rc)
        for (int32_t t = 0; t < c->e; ++t) out[t] = cw[c->map[t]];
    free(u);
    free(cw);
    return rc;
}

/* nr/ldpc.rs:191-197 with BitSelection::accumulate_into (puncture.rs:157-165) and
   KNOWN_BIT_LLR = 30 (crates/com-core/src/signal.rs:37) */
int or_nr_full_llr(const or_code* c, const double* llr, double* full) {
    for (int32_t i = 0; i < c->n; ++i) full[i] = 0.0;
    for (int32_t t = 0; t < c->e; ++t) full[c->map[t]] += llr[t];
    for (int32_t i = c->k_info; i < c->k; ++i) full[i] = 30.0;
    return 0;
}

/* ---- decoder: the same body in f64 (the reference) and in f32 ------------------------ */

#define REAL double
#define SFX(name) name
#define FABS fabs
#define FMAX fmax
#define TANH tanh
#define ATANH atanh
#include "or_bp_body.h"
#undef REAL
#undef SFX
#undef FABS
#undef FMAX
#undef TANH
#undef ATANH

/* phi(x) = -ln tanh(x/2) = ln(1 + 2e^-x / (1 - e^-x)) for x >= 0, with its ends taken apart as
   the FP32 device kernel takes them: below 2^-12, ln 2 - ln x (tanh(x/2) = x/2 (1 - x^2/12 + ...)),
   phi(0) = inf; the argument capped at 128, where e^-x is 0 in FP32. */
static float or_spa_phi(float x) {
    if (x < 0x1p-12f) return x == 0.0f ? INFINITY : 0.693147180559945f - logf(x);
    x = fminf(x, 128.0f);
    return log1pf(2.0f * expf(-x) / -expm1f(-x));
}

#define REAL float
#define SFX(name) name##_f32
#define FABS fabsf
#define FMAX fmaxf
#define TANH tanhf
#define ATANH atanhf
#define SPA_GALLAGER
#include "or_bp_body.h"
#undef SPA_GALLAGER
#undef REAL
#undef SFX
#undef FABS
#undef FMAX
#undef TANH
#undef ATANH

/* The layered min-sum (NMS or OMS) of ldpc.rs:509-531,546-557 run in f64 and in FP32 side by
   side on the same channel LLRs (values exactly representable in FP32), each with its own
   syndrome stop, with a record of where the two can part: the smallest f64 magnitude of a
   check-node input or a posterior over the f64 run (a tie below it is a recorded tie), the
   first point where an FP32 sign differs from the f64 one (its f64 magnitude and
   iteration), the rows whose FP32 minimum is at another edge while the f64 minimum is
   strict (the smallest relative f64 gap m2 / m1 - 1 among them), and per iteration the
   largest |post32 - post64| / max(|post64|, 1). Layout of `out` (10 + max_iterations):
   it64, it32, min_abs, first_sign_abs, first_sign_it, sign_events, argmin_events,
   min_argmin_gap, first_argmin_it, decisions_differ, rel[1..max_iterations]. */
int or_trace_f32_f64(const or_code* c, const or_dec* d, const double* llr, double* out) {
    if (d->rule != OR_NMS && d->rule != OR_OMS) return -1;
    const int T = d->max_iterations;
    int32_t edges = c->row_start[c->m], n = c->n;
    double* p64 = (double*)malloc(sizeof(double) * (size_t)n);
    float* p32 = (float*)malloc(sizeof(float) * (size_t)n);
    double* c64 = (double*)calloc((size_t)edges, sizeof(double));
    float* c32 = (float*)calloc((size_t)edges, sizeof(float));
    int dm = 0;
    for (int32_t r = 0; r < c->m; ++r) {
        int deg = c->row_start[r + 1] - c->row_start[r];
        if (deg > dm) dm = deg;
    }
    double *in64 = (double*)malloc(sizeof(double) * (dm + 1)), *s64 = (double*)malloc(sizeof(double) * (dm + 1));
    float *in32 = (float*)malloc(sizeof(float) * (dm + 1)), *s32 = (float*)malloc(sizeof(float) * (dm + 1));
    for (int32_t v = 0; v < n; ++v) {
        p64[v] = llr[v];
        p32[v] = (float)llr[v];
    }
    for (int i = 0; i < 10 + T; ++i) out[i] = 0.0;
    double min_abs = INFINITY, first_sign = -1.0, min_gap = INFINITY;
    int first_sign_it = 0, first_argmin_it = 0, sign_events = 0, argmin_events = 0;
    for (int32_t v = 0; v < n; ++v) min_abs = fmin(min_abs, fabs(p64[v]));
    int done64 = is_codeword_post(c, p64), done32 = is_codeword_post_f32(c, p32);
    int it64 = done64 ? 0 : T, it32 = done32 ? 0 : T;
    for (int it = 1; it <= T && !(done64 && done32); ++it) {
        for (int32_t r = 0; r < c->m; ++r) {
            int32_t s = c->row_start[r];
            int deg = c->row_start[r + 1] - s;
            if (!done64) {
                for (int i = 0; i < deg; ++i) in64[i] = p64[c->cols[s + i]] - c64[s + i];
                check_message(d, in64, deg, c64 + s, s64);
                for (int i = 0; i < deg; ++i) p64[c->cols[s + i]] = in64[i] + c64[s + i];
            }
            if (!done32) {
                for (int i = 0; i < deg; ++i) in32[i] = p32[c->cols[s + i]] - c32[s + i];
                check_message_f32(d, in32, deg, c32 + s, s32);
                for (int i = 0; i < deg; ++i) p32[c->cols[s + i]] = in32[i] + c32[s + i];
            }
            if (done64 || done32) continue;
            int a64 = 0, a32 = 0;
            double m1 = INFINITY, m2 = INFINITY;
            float f1 = INFINITY;
            for (int i = 0; i < deg; ++i) {
                double x = in64[i], post = p64[c->cols[s + i]];
                min_abs = fmin(min_abs, fmin(fabs(x), fabs(post)));
                int flip = ((x < 0.0) != (in32[i] < 0.0f)) || ((post < 0.0) != (p32[c->cols[s + i]] < 0.0f));
                if (flip) {
                    if (sign_events == 0) {
                        first_sign = ((x < 0.0) != (in32[i] < 0.0f)) ? fabs(x) : fabs(post);
                        first_sign_it = it;
                    }
                    ++sign_events;
                }
                if (fabs(x) < m1) { m2 = m1; m1 = fabs(x); a64 = i; } else if (fabs(x) < m2) m2 = fabs(x);
                if (fabsf(in32[i]) < f1) { f1 = fabsf(in32[i]); a32 = i; }
            }
            if (a64 != a32 && m2 > m1) {
                if (argmin_events == 0) first_argmin_it = it;
                ++argmin_events;
                min_gap = fmin(min_gap, m1 > 0.0 ? m2 / m1 - 1.0 : INFINITY);
            }
        }
        if (!done64 && is_codeword_post(c, p64)) { done64 = 1; it64 = it; }
        if (!done32 && is_codeword_post_f32(c, p32)) { done32 = 1; it32 = it; }
        double rel = 0.0;
        for (int32_t v = 0; v < n; ++v) rel = fmax(rel, fabs((double)p32[v] - p64[v]) / fmax(fabs(p64[v]), 1.0));
        out[10 + it - 1] = rel;
    }
    int diff = 0;
    for (int32_t v = 0; v < n; ++v) diff += (p64[v] < 0.0) != (p32[v] < 0.0f);
    out[0] = it64;
    out[1] = it32;
    out[2] = min_abs;
    out[3] = first_sign;
    out[4] = first_sign_it;
    out[5] = sign_events;
    out[6] = argmin_events;
    out[7] = min_gap;
    out[8] = first_argmin_it;
    out[9] = diff;
    free(p64); free(p32); free(c64); free(c32); free(in64); free(s64); free(in32); free(s32);
    return 0;
}

int or_nr_decode_block(const or_code* c, const or_dec* d, const double* llr, uint8_t* bits) {
    double* full = (double*)malloc(sizeof(double) * c->n);
    double* post = (double*)malloc(sizeof(double) * c->n);
    or_nr_full_llr(c, llr, full);
    int it = or_decode_llr(c, d, full, post);
    for (int32_t i = 0; i < c->k_info; ++i) bits[i] = (uint8_t)(post[i] < 0.0);
    free(full);
    free(post);
    return it;
}

/* ---- trial chains -------------------------------------------------------------------- */

void or_trial_bpsk_lifted(const or_code* c, double ebn0_db, uint64_t seed, const char* name,
                          uint64_t cond, uint64_t trial, uint8_t* cw, double* llr) {
    or_rng r;
    or_stream(&r, seed, name, cond, trial);
    uint8_t* u = (uint8_t*)malloc((size_t)c->k);
    for (int32_t i = 0; i < c->k; ++i) u[i] = or_rng_bit(&r);
    or_encode(c, u, cw);
    int32_t tx = c->n - 2 * c->z;
    double rate = (double)c->k / (double)tx;
    double s2 = or_bpsk_noise_variance(ebn0_db, rate);
    for (int32_t i = 0; i < 2 * c->z; ++i) llr[i] = 0.0;
    or_bpsk_awgn_llr(cw + 2 * c->z, tx, s2, &r, llr + 2 * c->z);
    free(u);
}

void or_trial_nr_qam(const or_code* c, int m, double noise_var, uint64_t seed, const char* name,
                     uint64_t cond, uint64_t trial, uint8_t* info, double* llr) {
    or_rng r;
    or_stream(&r, seed, name, cond, trial);
    for (int32_t i = 0; i < c->k_info; ++i) info[i] = or_rng_bit(&r);
    uint8_t* tx = (uint8_t*)malloc((size_t)c->e);
    or_nr_encode(c, info, tx);
    or_qam_awgn_llr(m, tx, c->e, noise_var, &r, llr);
    free(tx);
}
