/*
 * Oracle (test infrastructure only): the body of the belief-propagation decoder of
 * crates/com-fec/src/ldpc.rs:449-557, written once for a message type REAL and included
 * by or_code.c twice: REAL = double (the reference's arithmetic) and REAL = float (the
 * same operations rounded to FP32, the exact target of the FP32 device kernels; with
 * SPA_GALLAGER defined, the FP32 sum-product check in the log form the FP32 kernel uses).
 * Expects REAL, SFX(name) (name with a type suffix), FABS, FMAX, TANH, ATANH.
 */

/* ldpc.rs:449-493 */
static void SFX(check_message)(const or_dec* d, const REAL* in, int deg, REAL* out, REAL* scratch) {
#ifdef SPA_GALLAGER
    if (d->rule == OR_SPA) {
        /* FP32: the same function as a sum of phi(x) = -ln tanh(x/2) = ln(1 + 2e^-x / (1 - e^-x)),
           |prod_j tanh(l_j/2)| = exp(-S), 2 atanh(exp(-S)) = phi(S), and the reference's clamp
           |p| <= 0.999999999999 as |out| <= 2 atanh(0.999999999999); the sign is the parity of
           the others' sign bits. The product form rounds tanh(l/2) to 1 in FP32 for |l| > 17. */
        const float cap = 28.324190418452805f;
        /* phi with its ends taken apart, as the FP32 kernel computes it: below 2^-12
           ln 2 - ln x (phi(0) = inf), the argument capped at 128 (phi = 0 past it) */
        float prefix = 0.0f;
        uint32_t parity = 0, bits;
        for (int i = 0; i < deg; ++i) {
            scratch[i] = or_spa_phi(fabsf(in[i]));
            out[i] = prefix;
            prefix += scratch[i];
            memcpy(&bits, &in[i], 4);
            parity ^= bits;
        }
        float suffix = 0.0f;
        for (int i = deg - 1; i >= 0; --i) {
            float sum = out[i] + suffix;
            float m = fminf(or_spa_phi(sum), cap);
            suffix += scratch[i];
            uint32_t mb, ib;
            memcpy(&mb, &m, 4);
            memcpy(&ib, &in[i], 4);
            mb |= (parity ^ ib) & 0x80000000u;
            memcpy(&out[i], &mb, 4);
        }
        return;
    }
#endif
    if (d->rule == OR_SPA) {
        for (int i = 0; i < deg; ++i) scratch[i] = TANH(in[i] / (REAL)2.0);
        REAL prefix = 1.0;
        for (int i = 0; i < deg; ++i) {
            out[i] = prefix;
            prefix *= scratch[i];
        }
        REAL suffix = 1.0;
        const REAL bound = (REAL)0.999999999999;
        for (int i = deg - 1; i >= 0; --i) {
            REAL p = out[i] * suffix;
            suffix *= scratch[i];
            if (p < -bound) p = -bound;
            if (p > bound) p = bound;
            out[i] = (REAL)2.0 * ATANH(p);
        }
        return;
    }
    REAL sign = 1.0;
    for (int i = 0; i < deg; ++i) sign *= in[i] < (REAL)0.0 ? (REAL)-1.0 : (REAL)1.0;
    REAL m1 = (REAL)INFINITY, m2 = (REAL)INFINITY;
    int i1 = 0;
    for (int i = 0; i < deg; ++i) {
        REAL a = FABS(in[i]);
        if (a < m1) {
            m2 = m1;
            m1 = a;
            i1 = i;
        } else if (a < m2) {
            m2 = a;
        }
    }
    const REAL alpha = (REAL)d->alpha;
    for (int i = 0; i < deg; ++i) {
        REAL mag = i == i1 ? m2 : m1;
        REAL s = sign * (in[i] < (REAL)0.0 ? (REAL)-1.0 : (REAL)1.0);
        if (d->rule == OR_NMS) mag = alpha * mag;
        else mag = FMAX(mag - alpha, (REAL)0.0);
        out[i] = s * mag;
    }
}

/* ldpc.rs:555-557 */
static int SFX(is_codeword_post)(const or_code* c, const REAL* post) {
    for (int32_t r = 0; r < c->m; ++r) {
        uint8_t a = 0;
        for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e)
            a ^= (uint8_t)(post[c->cols[e]] < (REAL)0.0);
        if (a) return 0;
    }
    return 1;
}

/* ldpc.rs:509-551 */
static int SFX(propagate)(const or_code* c, const or_dec* d, const REAL* channel, REAL* post) {
    if (SFX(is_codeword_post)(c, post)) return 0;
    int32_t edges = c->row_start[c->m];
    REAL* c2v = (REAL*)calloc((size_t)edges, sizeof(REAL));
    REAL* new_c2v = d->layered ? NULL : (REAL*)calloc((size_t)edges, sizeof(REAL));
    int dm = 0;
    for (int32_t r = 0; r < c->m; ++r) {
        int deg = c->row_start[r + 1] - c->row_start[r];
        if (deg > dm) dm = deg;
    }
    REAL* incoming = (REAL*)malloc(sizeof(REAL) * (dm + 1));
    REAL* scratch = (REAL*)malloc(sizeof(REAL) * (dm + 1));
    int result = d->max_iterations;
    for (int it = 1; it <= d->max_iterations; ++it) {
        if (d->layered) {
            for (int32_t r = 0; r < c->m; ++r) {
                int32_t s = c->row_start[r];
                int deg = c->row_start[r + 1] - s;
                REAL* msgs = c2v + s;
                for (int i = 0; i < deg; ++i) incoming[i] = post[c->cols[s + i]] - msgs[i];
                SFX(check_message)(d, incoming, deg, msgs, scratch);
                for (int i = 0; i < deg; ++i) post[c->cols[s + i]] = incoming[i] + msgs[i];
            }
        } else {
            for (int32_t r = 0; r < c->m; ++r) {
                int32_t s = c->row_start[r];
                int deg = c->row_start[r + 1] - s;
                for (int i = 0; i < deg; ++i) incoming[i] = post[c->cols[s + i]] - c2v[s + i];
                SFX(check_message)(d, incoming, deg, new_c2v + s, scratch);
            }
            memcpy(post, channel, sizeof(REAL) * c->n);
            for (int32_t r = 0; r < c->m; ++r)
                for (int32_t e = c->row_start[r]; e < c->row_start[r + 1]; ++e) post[c->cols[e]] += new_c2v[e];
            REAL* t = c2v;
            c2v = new_c2v;
            new_c2v = t;
        }
        if (SFX(is_codeword_post)(c, post)) { result = it; break; }
    }
    free(c2v);
    free(new_c2v);
    free(incoming);
    free(scratch);
    return result;
}

/* ldpc.rs:500-505 */
int SFX(or_decode_llr)(const or_code* c, const or_dec* d, const REAL* llr, REAL* post) {
    memcpy(post, llr, sizeof(REAL) * c->n);
    return SFX(propagate)(c, d, llr, post);
}

typedef struct {
    const or_code* c;
    const or_dec* d;
    const REAL* llr;
    REAL* post;
    int32_t* its;
    int64_t begin, end;
} SFX(many_job);

static void* SFX(many_worker)(void* arg) {
    SFX(many_job)* j = (SFX(many_job)*)arg;
    int32_t n = j->c->n;
    for (int64_t w = j->begin; w < j->end; ++w)
        j->its[w] = SFX(or_decode_llr)(j->c, j->d, j->llr + w * n, j->post + w * n);
    return NULL;
}

void SFX(or_decode_many)(const or_code* c, const or_dec* d, const REAL* llr, int64_t count, REAL* post,
                         int32_t* iterations, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    SFX(many_job) jobs[256];
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (SFX(many_job)){c, d, llr, post, iterations, count * t / threads, count * (t + 1) / threads};
        if (threads == 1) SFX(many_worker)(&jobs[t]);
        else pthread_create(&th[t], NULL, SFX(many_worker), &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}
