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

// The fast layered min-sum kernel (FP32 messages, check-to-variable messages in registers).
struct LayeredArgs {
    int32_t z, k, k_info, e;
    const uint16_t* row_start;  // base tables (device)
    const uint16_t* col;
    const uint16_t* shift;
    const uint32_t* shift32;    // the same shifts as 32-bit words
    const uint32_t* src_start;
    const uint32_t* src;
    float a, b;                 // g(m) = a * max(m - b, 0)
    const float* params;        // learned: (a, b) per iteration and layer, or null
    int max_iterations;
    const float* llr;
    int64_t count;
    uint8_t* bits;
    float* post_out;
    uint32_t* iterations;
    uint8_t* ok;                // per codeword: the syndrome holds at the end, or null
    const uint32_t* index;      // decode codewords index[0 .. *index_count) only, or null
    const uint32_t* index_count;
    int add_iterations;         // add to iterations[] instead of storing
};

bool layered_supported(int bg, int z);
// the codewords whose syndrome failed: out[0 .. *n) (n zeroed by the caller's fill kernel)
hipError_t launch_failed_list(const uint8_t* ok, int64_t count, uint32_t* out, uint32_t* n, hipStream_t stream);

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
hipError_t launch_layered(const LayeredArgs& a, int bg, hipStream_t stream);

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

// A grow-only device buffer, one per (device, stream) in the C ABI layer.
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t want) {
        if (want <= bytes) return ptr;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (hipMalloc(&ptr, want) != hipSuccess) return nullptr;
        bytes = want;
        return ptr;
    }
};

hipError_t launch_generic(const GenericArgs<float>& a, int device, hipStream_t s, Workspace& ws);
hipError_t launch_generic(const GenericArgs<double>& a, int device, hipStream_t s, Workspace& ws);

}  // namespace comfec
