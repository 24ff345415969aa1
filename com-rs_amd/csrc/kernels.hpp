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
};

bool layered_supported(int bg, int z);
hipError_t launch_layered(const LayeredArgs& a, int bg, hipStream_t stream);

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
