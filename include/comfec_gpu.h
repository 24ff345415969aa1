/*
 * comfec_gpu: the NR LDPC belief-propagation decoder of com-fec and the Monte Carlo trial
 * loop of com-sim, on one AMD Instinct MI355X (gfx950), behind a C ABI.
 *
 * Every entry point names the reference interface it replaces (paths relative to the
 * com-rs repository). Conventions, as in the reference:
 *   - LLRs are positive for a 0 (crates/com-core/src/signal.rs:1-49); a hard decision is
 *     1 for an LLR < 0.0 (so -0.0 decides 0).
 *   - Errors are the kinds of com_core::Error (crates/com-core/src/error.rs:25-41), plus
 *     CF_DEVICE for a HIP failure; cf_last_error() gives the message of the last failure
 *     of the calling thread, worded as the reference's ("inconsistent shapes: 10 LLRs
 *     for a block of 300").
 *   - The caller owns every buffer; the library owns the cf_code handles. A handle is
 *     immutable during a decode (the settings change only through cf_set_decoder), and
 *     calls on one handle from several threads are safe.
 *   - "_host" buffers are host memory; "d_" buffers are device memory of the current
 *     HIP device; `stream` is a hipStream_t (NULL: the default stream).
 */
#ifndef COMFEC_GPU_H
#define COMFEC_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cf_code cf_code;

typedef enum {
    CF_OK = 0,
    CF_CONFIG = 1,    /* Error::Config */
    CF_SHAPE = 2,     /* Error::Shape */
    CF_NUMERICAL = 3, /* Error::Numerical */
    CF_CONDITION = 4, /* Error::Condition */
    CF_IO = 5,        /* Error::Io */
    CF_DEVICE = 6     /* a HIP runtime failure (no reference counterpart) */
} cf_status;

/* CheckRule (crates/com-fec/src/ldpc.rs:95-104) */
enum { CF_RULE_SUM_PRODUCT = 0, CF_RULE_NORMALIZED_MIN_SUM = 1, CF_RULE_OFFSET_MIN_SUM = 2 };

/* Message arithmetic of the device decoder. F64 reproduces the reference's f64 decoder
   bit for bit (min-sum rules); F32 keeps FP32 messages; F16 (layered min-sum rules) stores
   the posteriors and the check-to-variable magnitudes in binary16 (round to nearest even,
   clamped to +-65504, a zero posterior stored as +0.0) and computes in FP32. */
enum { CF_MSG_F64 = 0, CF_MSG_F32 = 1, CF_MSG_F16 = 2 };

typedef struct {
    uint32_t k_info;    /* K': information bits (BlockCode::k) */
    uint32_t e;         /* E: transmitted bits (BlockCode::n) */
    uint32_t n;         /* N: bits of the lifted code (LdpcCode::n) */
    uint32_t k;         /* K: systematic columns, 22Z or 10Z (LdpcCode::k) */
    uint32_t m;         /* checks */
    uint32_t edges;     /* nonzero entries of H */
    uint32_t base_graph;
    uint32_t z;
    uint32_t ils;       /* lifting-set index */
    uint32_t qm;
    uint32_t rv;
} cf_code_dims;

/* The message of the last failed call of this thread ("" after a success). */
const char* cf_last_error(void);

/* NrLdpc::with_options (crates/com-fec/src/nr/ldpc.rs:101-162): base_graph 0 chooses it as
   Sec. 7.2.2 does (nr/ldpc.rs:78-81). Settings are LdpcDecoder::default() (ldpc.rs:171-175). */
cf_status cf_nr_create(uint32_t k_info, uint32_t e, int base_graph, int rv, uint32_t qm,
                       cf_code** out);

/* LdpcCode::from_core_rows (ldpc.rs:270-287) on the lifted base graph bg at lifting size z
   (the construction of ldpc.rs:934-948): K' = K = 22Z or 10Z, the N - 2Z columns from 2Z
   transmitted once, in order. */
cf_status cf_lifted_create(int base_graph, uint32_t z, cf_code** out);

/* LdpcCode::from_rows (crates/com-fec/src/ldpc.rs:216-231): the code of the parity-check
   matrix whose row r lists the variables cols[row_start[r] .. row_start[r+1]) (m + 1 row
   starts); K = N - rank H, the information positions are the columns that are not pivots of
   H's reduced form (pivots taken from the last column, gf2.rs:124-150), the encoder sums
   them (ldpc.rs:760-764), and the block is sent whole (E = N). Decoded by the general
   kernel (any rule, layered or flooding); rows of at most 19 variables, N, M and the edges
   at most 65535. cf_decode_block and cf_app_into return the information positions. */
cf_status cf_code_from_rows(uint32_t n, uint32_t m, const uint32_t* row_start_host,
                            const uint32_t* cols_host, cf_code** out);
void cf_code_destroy(cf_code* code);

/* NrLdpc::{k, n, base_graph, lifting_size} and LdpcCode::{n, k} (nr/ldpc.rs:164-187). */
cf_status cf_code_dims_get(const cf_code* code, cf_code_dims* out);

/* NrLdpc::bit_map (nr/ldpc.rs:175-177): the full-codeword bit of every transmitted bit. */
cf_status cf_bit_map(const cf_code* code, uint32_t* positions_host, size_t e);

/* The public fields of LdpcDecoder (ldpc.rs:164-179) through NrLdpc::decoder_mut
   (nr/ldpc.rs:184-187): rule, its alpha (scale of the normalized min-sum, offset of the
   offset min-sum), layered or flooding, maximum iterations (0: channel LLRs returned). */
cf_status cf_set_decoder(cf_code* code, int rule, double alpha, int layered,
                         uint32_t max_iterations);

/* BlockCode::encode_block (nr/ldpc.rs:208-214): K' bits in, E bits out (host). */
cf_status cf_encode_block(const cf_code* code, const uint8_t* info_host, size_t k_info,
                          uint8_t* out_host, size_t e);

/* LdpcCode::decode_llr (ldpc.rs:500-505): N channel LLRs in, N a posteriori LLRs and the
   iterations out. Runs on the GPU in f64, bit-identical to the reference for min-sum. */
cf_status cf_decode_llr(const cf_code* code, const double* llr_host, size_t n,
                        double* post_host, uint32_t* iterations);

/* BlockCode::decode_block (nr/ldpc.rs:216-220): E LLRs in, K' hard decisions out. */
cf_status cf_decode_block(const cf_code* code, const double* llr_host, size_t e,
                          uint8_t* bits_host, size_t k_info);

/* SoftDecode::app_into (nr/ldpc.rs:226-239): E LLRs and optional K' a priori LLRs in,
   K' a posteriori LLRs out; returns the iterations through *iterations. */
cf_status cf_app_into(const cf_code* code, const double* llr_host, size_t e,
                      const double* apriori_host, double* out_host, size_t k_info,
                      uint32_t* iterations);

/* BlockCode::decode_stream (crates/com-fec/src/block.rs:81-89) and the trial loop of
   crates/com-sim/src/runner.rs:435-446 as one launch: `count` codewords of the same code,
   E LLRs each (f32 for CF_MSG_F32/F16, f64 for CF_MSG_F64) resident in device memory.
   Outputs (device, any may be NULL): d_bits count x K' hard decisions; d_post count x N a
   posteriori LLRs (of the input type); d_iterations count iteration counts (0 when the
   input was a codeword). Asynchronous on `stream`. */
cf_status cf_decode_batch(const cf_code* code, int msg_type, const void* d_llr, size_t count,
                          uint8_t* d_bits, void* d_post, uint32_t* d_iterations, void* stream);

/* No counterpart in the reference: cf_decode_batch for callers that cannot allocate device
   memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). The
   same arguments, kernels and results; every scratch buffer of the decode (the layered
   kernel's task counter, the general kernels' check-to-variable messages) comes from one
   array of CF_STATIC_SCRATCH_BYTES that the library holds on each GPU it runs on until the
   process ends, and the call allocates and frees nothing. The code's tables must already be
   on the device (any earlier decode or generator call with the code on that device puts them
   there): otherwise CF_CONFIG. A code whose codeword needs more message scratch than the
   array holds (the general kernel's f64 and FP32 paths, codes given by their rows) gets
   CF_CONFIG; the general kernels run as many codewords side by side as the array holds.
   Every call on a GPU shares that GPU's array: calls that overlap on one GPU (several
   streams, or host threads) race and must be serialised by the caller. */
#define CF_STATIC_SCRATCH_BYTES (64u << 20)
cf_status cf_decode_batch_static(const cf_code* code, int msg_type, const void* d_llr, size_t count,
                                 uint8_t* d_bits, void* d_post, uint32_t* d_iterations, void* stream);


/* ---- batches of many codes: one persistent launch --------------------------------------- */

/* A decode plan for a fixed list of codewords of possibly different codes (any mix of the
   two base graphs and of lifting sizes up to 384, each codeword with its own iteration
   limit): BlockCode::decode_block_into (crates/com-fec/src/block.rs:14-90) over a stream of
   heterogeneous blocks, as the runner's trial bodies call it (crates/com-sim/src/
   runner.rs:435-446). The codes' decoder settings (layered normalized or offset min-sum) are
   read when the plan is made, and the handles must outlive the plan. Codeword i has the
   E_i = llr_offsets[i+1] - llr_offsets[i] f32 LLRs d_llr[llr_offsets[i] ..]; its K'_i
   decisions go to d_bits at the sum of the K'_j of the codewords before it; d_iterations
   and d_ok (may be NULL) get one entry per codeword. max_iterations: per codeword, or NULL
   for each code's own limit. A launch decodes the codewords of its codes: workgroups take
   tasks (up to floor(threads / Z) codewords of one code) from a counter, the most work
   first. By default one launch takes every code at the width of the largest Z rounded up to
   a multiple of 64 (a configuration may give each code its own width, one launch per width,
   in one stream). msg_type: CF_MSG_F32 or CF_MSG_F16 (the initial storage). A plan is used
   by one thread at a time. */
typedef struct cf_batch_plan cf_batch_plan;

typedef struct {
    uint64_t llr_total;     /* llr_offsets[count] */
    uint64_t bits_total;    /* the sum of K' */
    uint32_t codewords, classes, tasks;
    uint32_t threads;       /* lanes per workgroup (768 where two wave sets share a codeword) */
    uint32_t grid;          /* workgroups of the launch (resident ones: the launch is persistent) */
    uint32_t lds_bytes;     /* dynamic LDS per workgroup */
    int msg_type;
    uint32_t segments;      /* launches of a run; threads, grid and lds_bytes are the first's */
} cf_batch_plan_dims;

/* How a plan runs. The decisions, iteration counts and verdicts do not depend on threads,
   per_code_width or grid_per_cu; msg_type chooses the arithmetic of the storage (FP32, or
   binary16 posteriors and messages). */
typedef struct {
    uint32_t threads;         /* one launch at this width: a multiple of 64 from the largest Z
                                 (rounded up) to 384; ignored with per_code_width */
    uint32_t per_code_width;  /* 1: each code at the width (a multiple of 64 from its Z rounded
                                 up to 384) with the most codeword rows per busy SIMD at the
                                 device's occupancy, or, in FP32 where one codeword a task
                                 alone fills a CU at that width, on two wave sets of 384
                                 lanes; one launch per width, the widest first */
    uint32_t grid_per_cu;     /* resident workgroups per CU at most (0: as many as fit) */
    int msg_type;             /* CF_MSG_F32 or CF_MSG_F16 */
    uint32_t overlap;         /* 1: the launches of a run go on parallel streams forked from
                                 and joined to the caller's, so that one launch's workgroups
                                 take the CUs another's leave; 0: one after another */
} cf_batch_plan_config;

/* One launch of a plan's run: its lanes per workgroup (twice 384 where FP32 tasks of one
   codeword would leave one workgroup a CU: two wave sets on each codeword), grid and LDS. */
typedef struct {
    uint32_t threads, grid, lds_bytes;
    uint32_t classes, tasks, codewords;
} cf_batch_plan_segment;

/* A configuration the autotuner timed and its mean time per run. */
typedef struct {
    cf_batch_plan_config config;
    uint32_t segments;
    double ms;
} cf_batch_plan_trial;

#define CF_TUNE_STORAGE 1u /* the autotuner also tries the other storage type */

cf_status cf_batch_plan_create(const cf_code* const* codes, const uint64_t* llr_offsets,
                               const uint32_t* max_iterations, size_t count, int msg_type,
                               cf_batch_plan** out);
void cf_batch_plan_destroy(cf_batch_plan* plan);
cf_status cf_batch_plan_dims_get(const cf_batch_plan* plan, cf_batch_plan_dims* out);

/* One launch at this width (the other settings kept): a multiple of 64 from the largest Z
   (rounded up) to 384. The results do not depend on it. Waits for the device. */
cf_status cf_batch_plan_set_threads(cf_batch_plan* plan, uint32_t threads);

/* The configuration as set, and a new one (waits for the device). */
cf_status cf_batch_plan_config_get(const cf_batch_plan* plan, cf_batch_plan_config* out);
cf_status cf_batch_plan_configure(cf_batch_plan* plan, const cf_batch_plan_config* config);
/* Launch `index` (0 .. segments - 1) of a run. */
cf_status cf_batch_plan_segment_get(const cf_batch_plan* plan, uint32_t index, cf_batch_plan_segment* out);

/* The autotuner: times a run of every candidate configuration on these buffers (repeats
   times each, after a warm-up) and keeps the fastest. The candidates: one launch at every
   admissible width and the per-code widths (their launches one after another and side by
   side), each with as many resident workgroups per CU as fit and with one, on the plan's
   storage (flags CF_TUNE_STORAGE: on both); configurations that launch the same are timed
   once. The first min(trials_cap, candidates) trials go to
   `trials` (may be NULL) and their number to *trials_count, the chosen time to *best_ms
   (both may be NULL). The outputs hold a decode of the chosen configuration afterwards. */
cf_status cf_batch_plan_tune(cf_batch_plan* plan, const float* d_llr, uint8_t* d_bits,
                             uint32_t* d_iterations, uint8_t* d_ok, uint32_t repeats, uint32_t flags,
                             cf_batch_plan_trial* trials, size_t trials_cap, size_t* trials_count,
                             double* best_ms, void* stream);

/* Decode on `stream`. With use_graph the counter fill and the launches are captured once as
   a HIP graph and replayed while the buffers and the configuration stay the same. */
cf_status cf_batch_plan_run(cf_batch_plan* plan, const float* d_llr, uint8_t* d_bits,
                            uint32_t* d_iterations, uint8_t* d_ok, int use_graph, void* stream);

/* A plan made, run once without a graph and freed (waits for the stream). */
cf_status cf_decode_batch_uneven(const cf_code* const* codes, const uint64_t* llr_offsets,
                                 const uint32_t* max_iterations, size_t count, int msg_type,
                                 const float* d_llr, uint8_t* d_bits, uint32_t* d_iterations,
                                 uint8_t* d_ok, void* stream);


/* ---- the device trial generator ------------------------------------------------------ */

/* Modulations of the generator: BPSK on a real channel, or Gray square QAM (QPSK, 16-QAM,
   64-QAM) normalized to unit energy (crates/com-mapping/src/constellation.rs:168-205). */
enum { CF_BPSK = 1, CF_QPSK = 2, CF_QAM16 = 4, CF_QAM64 = 6 };
/* Constellation::new(Qam, 2^modulation).normalized().points() (crates/com-mapping/src/
   constellation.rs:168-205, 396-398): the m = 2^modulation points of the generator's Gray
   QAM (modulation CF_QPSK, CF_QAM16 or CF_QAM64), point i the symbol of label i, its bits
   read most significant first. */
cf_status cf_constellation_points(int modulation, double* re_host, double* im_host, size_t m);

/* How the noise is given: Eb/N0 in dB at the code rate K'/E, or Es/N0 in dB. */
enum { CF_EBN0_DB = 0, CF_ESN0_DB = 1 };

/* No reference counterpart as one call: the body of Scenario::trial of the coded links
   (validation/tests/aff3ct.rs:151-158 with source.rs:30-32, encode_block, bpsk_awgn_llr
   validation/src/lib.rs:67-71, awgn.rs:94-99 and the max-log LLRs demapper.rs:270-297) for
   trials trial0 .. trial0+count-1 of grid point `point`, keyed by Philox-4x32-10 counters
   under `seed`. Writes count x E float32 LLRs and count x ceil(K'/32) packed information
   bits (LSB first) to device memory. */
cf_status cf_generate_batch(const cf_code* code, int modulation, int snr_kind, double snr_db,
                            uint64_t seed, uint32_t point, uint64_t trial0, size_t count,
                            float* d_llr, uint32_t* d_info, void* stream);

/* ---- campaigns: crates/com-sim ------------------------------------------------------- */

/* Record (crates/com-sim/src/scenario.rs:66-198): the metrics of a batch of trials. */
typedef struct cf_record cf_record;
/* Record::errors (scenario.rs:90-97) and Record::value (scenario.rs:104-115) */
cf_status cf_record_errors(cf_record* rec, const char* name, uint64_t errors, uint64_t total);
cf_status cf_record_value(cf_record* rec, const char* name, double value);

/* Stop (crates/com-sim/src/stopping.rs:64-141) as its nodes in prefix order: an ANY (or)
   or ALL (and) node is followed by its n_children subtrees. */
enum { CF_STOP_MAX_TRIALS = 0, CF_STOP_MIN_TRIALS = 1, CF_STOP_MIN_ERRORS = 2,
       CF_STOP_PRECISION = 3, CF_STOP_MAX_TIME = 4, CF_STOP_ANY = 5, CF_STOP_ALL = 6 };
typedef struct {
    int kind;
    uint32_t n_children;  /* ANY / ALL */
    uint64_t count;       /* MAX_TRIALS, MIN_TRIALS: trials; MIN_ERRORS: errors */
    double value;         /* PRECISION: relative half-width; MAX_TIME: seconds */
    double confidence;    /* PRECISION */
    const char* metric;   /* MIN_ERRORS, PRECISION */
} cf_stop_node;

/* Grid::axis / Grid::fixed (crates/com-sim/src/grid.rs:231-290): axes in order, the last
   fastest; a fixed value is an axis of one value. */
typedef struct {
    const char* name;
    const double* values;
    size_t n_values;
} cf_grid_axis;

/* The callbacks and loggers of a run (Runner::callback, Runner::logger at runner.rs:266-276),
   in the order added; see "callbacks and loggers" below. */
typedef struct cf_hooks cf_hooks;

/* The Runner builder (crates/com-sim/src/runner.rs:224-276). */
typedef struct {
    uint64_t seed;
    uint64_t batch;
    double confidence;
    int fast_dev_run;
    uint64_t threads;         /* Runner::threads, reported in the run information (0: the host's cores) */
    const char* config_json;  /* Runner::config: a JSON document for the report, or NULL */
    cf_hooks* hooks;          /* NULL: none */
} cf_runner_config;

/* Scenario::setup + Scenario::trial for the trials [t0, t1) of condition `index` whose
   values are cond_values (in axis order): record their metrics in `rec` and return 0, or
   return nonzero after cf_record_fail to stop the run with that error. */
typedef int (*cf_batch_fn)(void* user, size_t index, const double* cond_values, size_t n_values,
                           uint64_t t0, uint64_t t1, cf_record* rec);
cf_status cf_record_fail(cf_record* rec, int kind, const char* message);

typedef struct cf_results cf_results;

/* Runner::run (runner.rs:309-473) on a scenario given as a callback. */
cf_status cf_run(const cf_runner_config* config, const cf_stop_node* stop, size_t n_stop,
                 const cf_grid_axis* axes, size_t n_axes, const char* scenario_name,
                 cf_batch_fn scenario, void* user, cf_results** out);

/* The NR link of a campaign on the GPU: per trial, K' random bits, the NR encoder, the
   modulation over AWGN at the condition's SNR (grid key "ebn0_db" or "esn0_db"), the
   device decoder with the code's settings (or the learned offset/scale min-sum when
   d_learned_params is set: (a, b) per iteration and base row, float32), and the records
   "ber" (bit errors over K' bits), "fer" (frame errors) and "iterations" (mean). */
typedef struct {
    int modulation;               /* CF_BPSK, CF_QPSK, CF_QAM16, CF_QAM64 */
    int snr_kind;                 /* CF_EBN0_DB or CF_ESN0_DB */
    const float* d_learned_params;
    uint32_t learned_iterations;  /* iterations of the learned parameters (0: the code's limit) */
    int fallback;                 /* learned: decode failed codewords again with the code's rule */
} cf_nr_link;

/* Runner::run with the NR link as the scenario: batches of `batch` trials on the GPU. */
cf_status cf_nr_campaign(const cf_code* code, const cf_nr_link* link, const cf_runner_config* config,
                         const cf_stop_node* stop, size_t n_stop, const cf_grid_axis* axes,
                         size_t n_axes, cf_results** out);

/* ConditionResult (runner.rs:71-99) and ConditionResult::metric (runner.rs:101-104). */
typedef struct {
    double value, lo, hi;   /* estimate and its interval at the requested confidence */
    uint64_t count, total;  /* errors and trials (rates), or values twice (means) */
    int is_rate;
    uint64_t dropped;
} cf_metric;
size_t cf_results_count(const cf_results* res);
/* status: 0 completed, 1 stopped, 2 skipped, 3 dev run (runner.rs:58-68) */
cf_status cf_results_condition(const cf_results* res, size_t i, int* status, uint64_t* trials,
                               double* wall_s);
/* ErrorRate::wilson (crates/com-sim/src/stats.rs:74-87): the Wilson score interval of
   `errors` among `total` at `confidence`, z from Cody's erfc (special.rs:180-293). */
cf_status cf_wilson(uint64_t errors, uint64_t total, double confidence, double* lo, double* hi);

cf_status cf_results_metric(const cf_results* res, size_t i, const char* name, double confidence,
                            cf_metric* out);
/* The grid index of result i and its condition's values in axis order (up to `cap` of them;
   *n_values = their number). */
cf_status cf_results_values(const cf_results* res, size_t i, size_t* index, double* values, size_t cap,
                            size_t* n_values);
/* The metrics of result i in name order (Record::iter, scenario.rs:189-192). */
cf_status cf_results_metric_count(const cf_results* res, size_t i, size_t* n);
cf_status cf_results_metric_name(const cf_results* res, size_t i, size_t j, const char** name);
void cf_results_free(cf_results* res);

/* ---- callbacks and loggers: crates/com-sim/src/callbacks.rs, loggers.rs ------------------- */

/* Action (callbacks.rs:50-61) */
enum { CF_CONTINUE = 0, CF_SKIP = 1, CF_STOP_CONDITION = 2, CF_STOP_RUN = 3 };

/* RunInfo (runner.rs:150-175); the strings live for the duration of the call. */
typedef struct {
    const char* scenario;
    uint64_t seed, batch, threads;
    const char* stop;      /* the rule's Display, or "fast_dev_run" */
    double confidence;
    uint64_t n_conditions;
    uint64_t started_unix_s;
    const char* version;
    const char* commit;    /* "<hash>", "<hash>-dirty" or "unknown" */
    const char* config;    /* the Runner::config JSON, or NULL */
} cf_run_info;

/* Progress (stopping.rs:48-59): the condition's merged record after a batch. */
typedef struct {
    uint64_t condition;
    uint64_t trials;
    double elapsed_s;
    const cf_record* record;
} cf_progress;
/* Record::get (scenario.rs:185-187) with the estimate's interval at `confidence`. */
cf_status cf_record_metric(const cf_record* rec, const char* name, double confidence, cf_metric* out);

/* Callback (callbacks.rs:64-81). Every function may be NULL (Continue). The results given to
   on_condition_end hold the one condition that ended (read it with cf_results_*, index 0);
   those of on_run_end hold every condition. Neither may be freed by the callee. */
typedef struct {
    void* user;
    void (*on_run_start)(void* user, const cf_run_info* info);
    int (*on_condition_start)(void* user, size_t index, const double* values, size_t n_values);
    int (*on_batch_end)(void* user, const cf_progress* progress);
    int (*on_condition_end)(void* user, const cf_results* result);
    void (*on_run_end)(void* user, const cf_results* results);
} cf_callback;

/* Logger (loggers.rs:53-82). A nonzero return ends the run with CF_IO. */
typedef struct {
    void* user;
    int (*start)(void* user, const cf_run_info* info);
    int (*log)(void* user, const cf_results* result, double confidence);
    int (*finish)(void* user, const cf_run_info* info, const cf_results* results, double wall_s);
} cf_logger;

cf_status cf_hooks_create(cf_hooks** out);
/* Runner::callback / Runner::logger with a caller's hooks (copied; `user` is not owned). */
cf_status cf_hooks_add_callback(cf_hooks* h, const cf_callback* callback);
cf_status cf_hooks_add_logger(cf_hooks* h, const cf_logger* logger);
/* ProgressBar::new (callbacks.rs:101-105): lines on the standard error. */
cf_status cf_hooks_progress_bar(cf_hooks* h, double every_s);
/* SkipAfter::below (callbacks.rs:178-183). */
cf_status cf_hooks_skip_after(cf_hooks* h, const char* metric, double threshold, const char* along);
/* CsvLogger::new (loggers.rs:117-128) and JsonLogger::new (loggers.rs:225-229). */
cf_status cf_hooks_csv_logger(cf_hooks* h, const char* path);
cf_status cf_hooks_json_logger(cf_hooks* h, const char* path);
/* A hooks object serves one run at a time (its callbacks keep state within a run). */
void cf_hooks_destroy(cf_hooks* h);

/* ---- the learned offset/scale min-sum (no reference counterpart) ---------------------------
   The layered min-sum of ldpc.rs:509-531 whose check-node magnitude map, at iteration t and
   base row i, is g(m) = a[t][i] max(m - b[t][i], 0); the parameters are float32 pairs (a, b)
   laid out [t][i][2] in device memory (a = alpha, b = 0 is the normalized min-sum). Training
   minimises the mean binary cross-entropy of sigmoid(-post) against the codeword bits over
   the unrolled iterations. */

/* Training data from the device trial generator: as cf_generate_batch, with each trial's
   SNR uniform in [snr_lo_db, snr_hi_db], and the count x ceil(N/32) packed codeword bits. */
cf_status cf_generate_training_batch(const cf_code* code, int modulation, int snr_kind,
                                     double snr_lo_db, double snr_hi_db, uint64_t seed,
                                     uint32_t point, uint64_t trial0, size_t count, float* d_llr,
                                     uint32_t* d_codeword, void* stream);

/* The loss over `count` codewords and its gradient with respect to the parameters
   (d_grad: iterations x rows x 2 floats), by the forward and backward kernels. */
cf_status cf_learned_loss_grad(const cf_code* code, uint32_t iterations, const float* d_params,
                               const float* d_llr, const uint32_t* d_codeword, size_t count,
                               float* d_grad, float* d_loss, void* stream);

/* One Adam step (Kingma & Ba, 2015) on n parameters; step counts from 1. */
cf_status cf_adam_step(float* d_params, const float* d_grad, float* d_m, float* d_v, size_t n,
                       uint32_t step, double lr, double beta1, double beta2, double eps, void* stream);

/* `steps` Adam steps (lr, beta 0.9 / 0.999) on batches of the device generator (trial t of
   step s is trial s*batch + t of grid point 0); the loss of every step to loss_history_host. */
cf_status cf_learned_train(const cf_code* code, uint32_t iterations, float* d_params, uint32_t steps,
                           uint32_t batch, double lr, int modulation, int snr_kind, double snr_lo_db,
                           double snr_hi_db, uint64_t seed, float* loss_history_host, void* stream);

/* Decode with the learned parameters (layered, syndrome stop). With fallback set, every
   codeword whose syndrome still fails is decoded again from its channel LLRs by the code's
   own min-sum settings, and its iteration count is the sum of both. d_ok: per codeword,
   whether the learned decoder alone ended on a codeword. */
cf_status cf_learned_decode_batch(const cf_code* code, uint32_t iterations, const float* d_params,
                                  int fallback, const float* d_llr, size_t count, uint8_t* d_bits,
                                  uint32_t* d_iterations, uint8_t* d_ok, void* stream);

#ifdef __cplusplus
}
#endif
#endif
