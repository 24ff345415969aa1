/*
 * CPU oracle of the NR LDPC decoder path: a plain-C, f64 restatement of the reference
 * algorithm, in the reference's operation order, compiled with -ffp-contract=off.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it, and only as the checker (or the timed CPU baseline),
 * never as part of the product path. The product (com-rs_amd/) never links it.
 *
 * Each function names the reference file:line it restates. Pins (what fixes each part
 * independently of this restatement) are listed in oracle/README.md.
 */
#ifndef COMFEC_ORACLE_H
#define COMFEC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random streams: crates/com-core/src/rng.rs ------------------------------------ */
typedef struct { uint64_t s[4]; } or_rng;
uint64_t or_splitmix64(uint64_t* state);                       /* rng.rs:49-55 */
uint64_t or_fnv1a64(const char* bytes, size_t n);              /* rng.rs:58-65 */
void or_rng_new(or_rng* r, uint64_t seed);                     /* rng.rs:75-79 */
uint64_t or_rng_next(or_rng* r);                               /* rng.rs:98-111 */
double or_rng_uniform(or_rng* r);                              /* rng.rs:115-117 */
uint8_t or_rng_bit(or_rng* r);                                 /* rng.rs:146-148 */
double or_rng_gaussian(or_rng* r);                             /* rng.rs:152-157 */
void or_rng_complex_gaussian(or_rng* r, double* re, double* im); /* rng.rs:161-167 */
void or_stream(or_rng* r, uint64_t base, const char* name, uint64_t cond, uint64_t trial); /* rng.rs:270-278 */

/* ---- codes: crates/com-fec/src/ldpc.rs, crates/com-fec/src/nr/ldpc.rs ---------------- */
typedef struct or_code or_code;
/* LdpcCode::from_rows without its encoder (rows given in CSR); k is unknown (0). */
or_code* or_code_from_rows(int32_t n, int32_t m, const int32_t* row_start, const int32_t* cols);
/* from_core_rows(N, lift(bg, z, ils), k = 22Z or 10Z, core = 4Z): ldpc.rs:270-287 on the rows of
   nr/ldpc.rs:124-133 (the way ldpc.rs:934-948 builds lifted matrices). */
or_code* or_code_lifted(int bg, int32_t z);
/* NrLdpc::with_options(k_info, e, bg or 0 for the rule of 7.2.2, rv, qm): nr/ldpc.rs:101-162. */
or_code* or_nr_create(int32_t k_info, int32_t e, int bg_override, int rv, int32_t qm);
void or_code_free(or_code* c);
const char* or_last_error(void);

/* sizes: 0 n, 1 m (checks), 2 k (systematic bits), 3 k_info, 4 e, 5 bg, 6 z, 7 ils, 8 edges */
int64_t or_code_info(const or_code* c, int what);
const int32_t* or_code_row_start(const or_code* c);
const int32_t* or_code_cols(const or_code* c);
const int32_t* or_code_bit_map(const or_code* c);   /* e positions (NrLdpc::bit_map) */

int or_encode(const or_code* c, const uint8_t* u, uint8_t* cw);              /* ldpc.rs:753-791 */
int or_nr_encode(const or_code* c, const uint8_t* info, uint8_t* out);       /* nr/ldpc.rs:208-214 */
int or_nr_full_llr(const or_code* c, const double* llr, double* full);       /* nr/ldpc.rs:191-197 */
int or_is_codeword(const or_code* c, const uint8_t* cw);                     /* ldpc.rs:396-398 */

/* ---- decoder: ldpc.rs:449-557 ------------------------------------------------------ */
enum { OR_SPA = 0, OR_NMS = 1, OR_OMS = 2 };
typedef struct { int rule; double alpha; int layered; int max_iterations; } or_dec;
/* LdpcCode::decode_llr: post receives the a posteriori LLRs; returns the iterations. */
int or_decode_llr(const or_code* c, const or_dec* d, const double* llr, double* post);
/* NrLdpc::decode_block (nr/ldpc.rs:216-220): k_info hard decisions; returns iterations. */
int or_nr_decode_block(const or_code* c, const or_dec* d, const double* llr, uint8_t* bits);
/* many codewords of n LLRs, on `threads` POSIX threads (1: serial) */
void or_decode_many(const or_code* c, const or_dec* d, const double* llr, int64_t count,
                    double* post, int32_t* iterations, int threads);

/* the same decoder with every message rounded to FP32 (min-sum rules: the exact target of
   the FP32 device kernels, whose operations are the same IEEE FP32 operations) */
int or_decode_llr_f32(const or_code* c, const or_dec* d, const float* llr, float* post);
void or_decode_many_f32(const or_code* c, const or_dec* d, const float* llr, int64_t count,
                        float* post, int32_t* iterations, int threads);

/* the layered min-sum in f64 and in FP32 side by side, with a record of where they part
   (see or_code.c): out holds 10 + d->max_iterations values; -1 for a rule other than NMS/OMS */
int or_trace_f32_f64(const or_code* c, const or_dec* d, const double* llr, double* out);

/* the layered min-sum (rules NMS and OMS) with the posteriors and check-to-variable
   messages stored in IEEE binary16 and the arithmetic in FP32 (the device FP16 mode);
   max_iterations, when not null, overrides d->max_iterations per codeword */
uint16_t or_f32_to_f16(float x);  /* round to nearest even */
float or_f16_to_f32(uint16_t h);
int or_decode_llr_f16(const or_code* c, const or_dec* d, const float* llr, float* post);
void or_decode_many_f16(const or_code* c, const or_dec* d, const float* llr, int64_t count,
                        const int32_t* max_iterations, float* post, int32_t* iterations, int threads);

/* learned offset/scale min-sum (no reference counterpart): layered schedule, iteration t
   and base row i use g(m) = alpha[t][i] * max(m - beta[t][i], 0); params t-major. */
int or_learned_decode(const or_code* c, int iterations, const double* alpha, const double* beta,
                      const double* llr, double* post);
/* loss = mean over the n bits of BCE(sigmoid(-post), bit) and its gradient wrt alpha, beta */
double or_learned_loss_grad(const or_code* c, int iterations, const double* alpha,
                            const double* beta, const double* llr, const uint8_t* cw,
                            double* g_alpha, double* g_beta);

/* ---- link: channel and LLRs --------------------------------------------------------- */
double or_db2lin(double x_db);                                   /* units.rs:38-40 */
double or_bpsk_noise_variance(double ebn0_db, double rate);      /* bounds.rs:58-60 */
double or_ebn0_noise_variance(double p, double ebn0_db, double bits, double sps); /* awgn.rs:139-141 */
/* validation/src/lib.rs:67-71: LLRs 2y/s2 of BPSK bits through real AWGN, one gaussian per bit */
void or_bpsk_awgn_llr(const uint8_t* c, int64_t n, double noise_var, or_rng* r, double* llr);
/* Gray square QAM of order m normalised to unit energy (constellation.rs:168-205, 396-398) */
void or_qam_points(int m, double* re, double* im);
/* map bits (MSB first) + add_noise (awgn.rs:94-99) + max-log LLRs (demapper.rs:270-297) */
void or_qam_awgn_llr(int m, const uint8_t* bits, int64_t nbits, double noise_var, or_rng* r,
                     double* llr);
void or_qam_llr(int m, const double* yre, const double* yim, int64_t nsym, double noise_var,
                double* llr);
/* LabelHalves::new(metric, shift, MaxLog).llr() (demapper.rs:94-146) and llr_from_metrics
   (demapper.rs:130-137) in max-log, the most significant bit first */
double or_label_halves_llr(const double* metric, int m, int shift);
void or_llr_from_metrics(const double* metric, int m, double* out);

/* Philox-4x32-10 and the information bits of the device trial generator */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_philox_info_bits(uint64_t seed, uint32_t point, uint64_t trial, int32_t k_info, uint8_t* bits);
/* its Box-Muller normals of draw `index` (purpose 1), and one whole trial of it in f64:
   information bits, encoder, BPSK (bps 1) or Gray QAM (bps 2, 4, 6), AWGN, LLRs */
void or_philox_normals(uint64_t seed, uint32_t point, uint64_t trial, uint32_t index, double* g0,
                       double* g1);
void or_philox_trial(const or_code* c, int bps, double noise_var, uint64_t seed, uint32_t point,
                     uint64_t trial, uint8_t* info, double* yre, double* yim, double* llr);

/* ---- statistics: crates/com-math/src/special.rs, crates/com-sim/src/stats.rs ---------- */
double or_erfc(double x);                       /* special.rs:180-255 (Cody) */
double or_erfcinv(double y);                    /* special.rs:265-273 */
double or_qfuncinv(double p);                   /* special.rs:291-293 */
double or_z_score(double confidence);           /* stats.rs:43-45 */
void or_wilson(uint64_t errors, uint64_t total, double confidence, double* lo, double* hi); /* stats.rs:74-87 */

/* ---- trial chains of the workloads ------------------------------------------------ */
/* One trial of the BPSK link on a lifted code (random info bits, encode, BPSK/AWGN LLRs with
   the first 2Z columns set to zero): validation/tests/aff3ct.rs:151-158 on from_core_rows. */
void or_trial_bpsk_lifted(const or_code* c, double ebn0_db, uint64_t seed, const char* name,
                          uint64_t cond, uint64_t trial, uint8_t* cw, double* llr);
/* One trial of the NR QAM link: info bits, NrLdpc::encode_block, Gray QAM, AWGN of variance
   noise_var per symbol, max-log LLRs (E values). */
void or_trial_nr_qam(const or_code* c, int m, double noise_var, uint64_t seed, const char* name,
                     uint64_t cond, uint64_t trial, uint8_t* info, double* llr);
/* the same for count trials from trial0 on `threads` POSIX threads */
void or_trial_nr_qam_many(const or_code* c, int m, double noise_var, uint64_t seed, const char* name,
                          uint64_t cond, uint64_t trial0, int64_t count, uint8_t* info, double* llr,
                          int threads);

#ifdef __cplusplus
}
#endif
#endif
