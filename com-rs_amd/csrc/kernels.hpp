// Device-side descriptions shared by the decoder kernels and their launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace comfec {

// LDS a workgroup may take for a codeword's posterior LLRs (gfx950: 160 KiB per CU).
constexpr size_t kLdsBudget = 150 * 1024;

// The lifted code as the kernels see it (tables in device memory).
struct DevCode {
    int32_t z, rows_b, cols_b, edges_b, n, k, k_info, e;
    const uint16_t* row_start;  // rows_b + 1
    const uint16_t* col;        // edges_b
    const uint16_t* shift;      // edges_b, V mod Z
    const uint16_t* col_start;  // cols_b + 1
    const uint16_t* col_edge;   // edges_b
    const uint32_t* src_start;  // n + 1
    const uint32_t* src;        // e
    const uint32_t* info;       // k_info information positions, or null: 0 .. k_info - 1
    const uint32_t* row_start32;  // row_start and col as 32-bit words
    const uint32_t* col32;
    int32_t max_deg;              // the widest check
    int32_t distinct;             // no check lists a variable twice
};

template <typename T>
struct GenericArgs {
    DevCode code;
    int rule;                 // 0 SPA, 1 NMS, 2 OMS
    double alpha;
    int layered;
    int max_iterations;
    int input_full;           // llr holds n values per codeword (decode_llr) instead of e
    const T* llr;
    const T* apriori;         // k_info per codeword or null
    int64_t count;
    uint8_t* bits;            // k_info per codeword or null
    T* post_out;              // n per codeword or null
    uint32_t* iterations;     // per codeword or null
    T* scratch;               // set by the launcher
    size_t scratch_stride;    // elements per workgroup
};

// The layered min-sum kernel (messages in registers, posteriors in LDS).
constexpr int kMaxThreads = 384;               // lanes of a workgroup: one per row of a layer
constexpr int kMaxGroup = kMaxThreads / 2;     // codewords of one task (the smallest Z is 2)
constexpr size_t kSharedBytes = 8192;          // a workgroup's tables and flags, after its posteriors
// what a two-set workgroup adds after its tables: per set and row, the two smallest magnitudes
// and the sign product of its part of a row split over both sets
constexpr size_t kExchangeBytes = 3 * 2 * kMaxThreads * 4;

// One code of a decode batch, as the layered kernel sees it.
struct BatchClass {
    int32_t bg, z, n, k, k_info, e;
    int32_t g;                  // codewords per task
    int32_t max_iterations;     // of the code's decoder settings
    float a, b;                 // its rule: g(m) = a max(m - b, 0)
    const uint32_t* shift32;    // edges_b shifts V mod Z, layer order
    const uint16_t* col;        // edges_b base columns
    const uint16_t* row_start;  // rows_b + 1
    const uint32_t* src_start;  // n + 1: rate recovery (the output bits of every codeword bit)
    const uint32_t* src;        // e
    const uint32_t* map;        // e: the codeword position of every output bit
    int32_t direct;             // no position is repeated: the LLRs can be scattered by map
};

// A workgroup's unit of work: codewords order[first .. first + count) of class `cls`.
struct BatchTask {
    uint32_t cls, first, count, unused;
};

struct LayeredArgs {
    BatchClass one;             // the class of a single-code batch (tasks == null)
    const BatchClass* classes;  // per class, or null
    const BatchTask* tasks;     // per task, or null: task t is codewords [t g, t g + g) of `one`
    int32_t n_tasks;
    uint32_t* queue;            // task counter, zero at launch (a fill kernel), or null: task t in block t mod grid
    int64_t count;              // codewords (tasks == null)
    const uint32_t* order;      // codeword ids in task order, or null (identity)
    const uint64_t* llr_off;    // first LLR of each codeword, or null: id e
    const uint64_t* bit_off;    // first decision of each codeword, or null: id K'
    const uint16_t* max_it;     // iteration limit of each codeword, or null: the class's
    const float* params;        // learned (a, b) per iteration and base row, or null
    int32_t learned_iterations;
    int32_t fallback;           // decode a failed learned codeword again with the class's rule
    const float* llr;
    uint8_t* bits;
    float* post_out;            // n per codeword id (single-code batches), or null
    uint32_t* iterations;       // per codeword id, or null
    uint8_t* ok;                // per codeword id: the syndrome held (learned: before any fallback), or null
    int32_t shared_off;         // byte offset of the tables in the dynamic LDS (after the posteriors)
};

struct LayeredLaunch {
    int bg_mask;                // 1: base graph 1 only, 2: base graph 2 only, 3: both
    bool half;                  // FP16 storage
    bool solo;                  // every class has G = 1
    int zc;                     // the Z of every class, when the kernel has it built in (384), or 0
    int threads;                // a multiple of 64, at most kMaxThreads, at least every class's Z
    int grid;
    size_t lds;                 // dynamic LDS: layered_shared_offset + kSharedBytes
    bool nms = false;           // every class the normalized min-sum and no learned parameters (solo, zc 384)
    bool dual = false;          // two wave sets a codeword (FP32, solo): 2 x 384 lanes
};

bool layered_supported(int bg, int z);
// LDS of a task of g codewords (posteriors and packed decisions)
size_t layered_lds_bytes(int bg, int z, int g, bool half);
// where the tables go: after the largest task's LDS, 16-byte aligned
inline size_t layered_shared_offset(size_t task_lds) { return (task_lds + 15) & ~size_t(15); }
hipError_t layered_occupancy(int bg_mask, bool half, bool solo, int zc, int threads, size_t lds, int* blocks_per_cu,
                             bool nms = false, bool dual = false);
// sets the kernel's dynamic LDS limit for the launch (not while a stream is being captured)
hipError_t layered_prepare(const LayeredLaunch& l);
hipError_t launch_layered(const LayeredArgs& a, const LayeredLaunch& l, hipStream_t stream, bool prepare);

// Training of the learned offset/scale min-sum (T unrolled layered iterations).
struct LearnArgs {
    DevCode code;
    int T;
    const float* params;     // (a, b) per iteration and base row
    const float* llr;        // count x e
    const uint32_t* cw;      // count x ceil(n/32) codeword bits
    int64_t count;
    float scale;             // 1 / (T n count)
    float* ckpt;             // count x T x edges: variable-to-check messages
    float* gpost;            // count x T x n: loss gradient wrt the posteriors of each iteration
    float* cbar;             // count x edges: backward scratch
    float* loss_out;         // count
    float* grad_partial;     // count x (T rows 2)
    float* grad;             // T rows 2
    float* loss;             // 1
};

hipError_t launch_learn_forward(const LearnArgs& a, hipStream_t stream);
hipError_t launch_learn_backward(const LearnArgs& a, hipStream_t stream);
hipError_t launch_adam(float* params, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t stream);

// The device trial generator: information bits, encoder, rate matching, channel, LLRs.
struct TrialGenArgs {
    int32_t z, n, k, k_info, e, rows_b, core_words;
    const uint16_t* row_start;
    const uint16_t* col;
    const uint16_t* shift;
    const uint32_t* map;       // e: codeword bit of every transmitted bit
    const uint64_t* core_inv;  // 4Z rows of core_words
    int bits_per_symbol;       // 1: BPSK on a real channel; 2, 4, 6: Gray QAM
    float noise_var;           // BPSK: per real dimension; QAM: N0 per complex symbol
    float noise_sd;            // BPSK: sqrt(noise_var); QAM: sqrt(N0 / 2) per component
    float log_prior;           // -ln M (the uniform prior of the max-log metric)
    float pts_re[64], pts_im[64];
    uint32_t k0, k1;           // the 64-bit seed
    uint32_t point;
    uint64_t trial0;
    int64_t count;
    float* llr;                // count x e
    uint32_t* info;            // count x ceil(K'/32), LSB first
    uint32_t* codeword;        // count x ceil(N/32) codeword bits, or null
    float snr_lo, snr_hi;      // per-trial SNR uniform in [lo, hi] dB when hi > lo
    int snr_kind;              // 0 Eb/N0, 1 Es/N0 (per-trial SNR)
    float rate;                // K'/E (per-trial SNR)
};

// The error counts of a decoded batch.
struct CountArgs {
    int32_t k_info;
    const uint8_t* bits;       // count x K'
    const uint32_t* info;      // count x ceil(K'/32)
    const uint32_t* iterations;
    int64_t count;
    uint64_t* counters;        // [bit errors, frame errors, sum of iterations, sum of their squares]
    uint32_t* per_trial_errors;
};

size_t generate_lds_bytes(int z, int n);
hipError_t launch_generate(const TrialGenArgs& a, hipStream_t stream);
hipError_t launch_count(const CountArgs& a, hipStream_t stream);
// a fill kernel (not a memset node, so that graph replays clear the counters too)
hipError_t launch_zero(uint64_t* p, int n, hipStream_t stream);

// A grow-only device buffer, one per (device, stream) in the C ABI layer; or, with `fixed`
// set, a buffer of `bytes` given whole that is never reallocated (the launchers then size
// their persistent grids to it).
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool fixed = false;
    void* get(size_t want) {
        if (want <= bytes) return ptr;
        if (fixed) return nullptr;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (hipMalloc(&ptr, want) != hipSuccess) {
            // the failure stays the thread's last error: clear it, so that a later launch
            // check (hipGetLastError) reports that launch and not this allocation
            (void)hipGetLastError();
            ptr = nullptr;
            return nullptr;
        }
        bytes = want;
        return ptr;
    }
};

hipError_t launch_generic(const GenericArgs<float>& a, int device, hipStream_t s, Workspace& ws);
hipError_t launch_generic(const GenericArgs<double>& a, int device, hipStream_t s, Workspace& ws);

}  // namespace comfec
