// Microbenchmarks of the resources that bound the decoder kernels, measured on the device
// the benchmark runs on (a separate library, libcomfec_peaks.so, loaded by bench.py only):
//   - HBM: a read-only stream (16-byte loads, a fold of what is read) and a copy;
//   - LDS: 4-byte and 16-byte reads at consecutive addresses (no bank conflicts) by every
//     wave a CU holds, the read width of the decoder's posterior accesses and the widest;
//   - VALU issue: independent chains of FP32 fused multiply-adds, 32-bit integer adds and
//     three-input bitwise operations on every SIMD.
// Each entry point times `launches` launches with HIP events after one warm-up launch and
// returns the mean time of a launch in milliseconds, and the work of one launch (bytes or
// wave instructions) through *work.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) hbm_read_kernel(const uint4* __restrict__ src, size_t n4,
                                                            uint32_t* __restrict__ sink) {
    uint32_t acc = 0;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    for (; i < n4; i += stride) {
        const uint4 a = src[i];
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    if (acc == 0x9E3779B9u) sink[blockIdx.x] = acc;  // keeps the loads; practically never stores
}

__global__ void __launch_bounds__(kThreads) hbm_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                            size_t n4) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

// Every lane reads 16 words (or 4 16-byte pieces) an iteration at consecutive addresses
// across the wavefront; the base address is made opaque each iteration, so the reads cannot
// be hoisted out of the loop, at two VALU instructions per 16 reads. Workgroups of 1024 lanes
// with 64 KiB of LDS each: two fit a CU, 32 waves.
constexpr int kLdsThreads = 1024;

template <int WIDTH>
__global__ void __launch_bounds__(kLdsThreads) lds_read_kernel(int iterations, uint32_t* __restrict__ sink) {
    constexpr int kWords = 16384;  // 64 KiB
    __shared__ uint32_t buf[kWords];
    for (int i = threadIdx.x; i < kWords; i += blockDim.x) buf[i] = uint32_t(i) * 2654435761u;
    __syncthreads();
    uint32_t acc[4] = {0u, 0u, 0u, 0u}, zero = 0;
    const uint32_t base = threadIdx.x;
    for (int it = 0; it < iterations; ++it) {
        asm volatile("" : "+v"(zero));
        if constexpr (WIDTH == 4) {
            // base + 15 * 1024 < 16384: every read in the buffer
            const uint32_t* p = buf + ((base + zero) & (kWords - 1));
            uint32_t x[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) x[k] = p[k * kLdsThreads];
#pragma unroll
            for (int k = 0; k < 16; ++k) acc[k & 3] ^= x[k];
        } else {
            // base + 3 * 1024 < 4096 pieces
            const uint4* p = reinterpret_cast<const uint4*>(buf) + ((base + zero) & (kWords / 4 - 1));
            uint4 x[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) x[k] = p[k * kLdsThreads];
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] ^= x[k].x ^ x[k].y ^ x[k].z ^ x[k].w;
        }
    }
    if ((acc[0] ^ acc[1] ^ acc[2] ^ acc[3]) == 0x9E3779B9u) sink[blockIdx.x] = 1u;
}

// Eight independent chains of FP32 fused multiply-adds per lane (one v_fma_f32 each; packed
// FP32 instructions are disabled for every translation unit).
__global__ void __launch_bounds__(kThreads) valu_kernel(int iterations, uint32_t* __restrict__ sink) {
    float r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = float(threadIdx.x) * 1e-3f + float(k);
    const float a = 0.999f, b = 1e-3f;
    for (int it = 0; it < iterations; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] = __builtin_fmaf(r[k], a, b);
    }
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += r[k];
    if (acc == -1.0f) sink[blockIdx.x] = 1u;
}

// Eight independent chains of a 32-bit integer instruction per lane: v_add_u32 (kind 0) or a
// three-input v_bitop3_b32 (kind 1); the operand is made opaque every iteration, so the chain
// cannot be folded.
template <int KIND>
__global__ void __launch_bounds__(kThreads) valu_int_kernel(int iterations, uint32_t* __restrict__ sink) {
    uint32_t r[8], inc = threadIdx.x | 1u;
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = threadIdx.x * 17u + uint32_t(k);
    for (int it = 0; it < iterations; ++it) {
        asm volatile("" : "+v"(inc));
#pragma unroll
        for (int k = 0; k < 8; ++k)
            r[k] = KIND == 0 ? r[k] + inc : __builtin_amdgcn_bitop3_b32(r[k], inc, 0x5bd1e995u, 0x96);
    }
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc ^= r[k];
    if (acc == 0x9E3779B9u) sink[blockIdx.x] = acc;
}

template <typename L>
double timed(int launches, L&& launch) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.0;
    launch();
    (void)hipEventRecord(a, nullptr);
    for (int i = 0; i < launches; ++i) launch();
    (void)hipEventRecord(b, nullptr);
    (void)hipEventSynchronize(b);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (hipGetLastError() != hipSuccess) return -1.0;
    return double(ms) / launches;
}

int cus() {
    int d = 0, n = 0;
    if (hipGetDevice(&d) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess) return 0;
    return n;
}

}  // namespace

extern "C" {

// HBM read of `bytes` (a multiple of 16) from d_src; *work = bytes a launch.
double cf_peak_hbm_read(const void* d_src, size_t bytes, uint32_t* d_sink, int launches, double* work) {
    const size_t n4 = bytes / 16;
    const unsigned grid = unsigned(cus() * 8);
    *work = double(n4 * 16);
    return timed(launches, [&] {
        hipLaunchKernelGGL(hbm_read_kernel, dim3(grid), dim3(kThreads), 0, nullptr,
                           static_cast<const uint4*>(d_src), n4, d_sink);
    });
}

// HBM copy of `bytes` from d_src to d_dst; *work = the bytes read and written a launch.
double cf_peak_hbm_copy(const void* d_src, void* d_dst, size_t bytes, int launches, double* work) {
    const size_t n4 = bytes / 16;
    const unsigned grid = unsigned(cus() * 8);
    *work = double(n4 * 32);
    return timed(launches, [&] {
        hipLaunchKernelGGL(hbm_copy_kernel, dim3(grid), dim3(kThreads), 0, nullptr, static_cast<const uint4*>(d_src),
                           static_cast<uint4*>(d_dst), n4);
    });
}

// LDS reads of `width` (4 or 16) bytes a lane; *work = the bytes read a launch.
double cf_peak_lds(int width, int iterations, uint32_t* d_sink, int launches, double* work) {
    const unsigned grid = unsigned(cus() * 2 * 4);  // two workgroups fit a CU at once; four rounds
    *work = double(grid) * kLdsThreads * double(iterations) * 64.0;
    if (width == 4)
        return timed(launches, [&] {
            hipLaunchKernelGGL(lds_read_kernel<4>, dim3(grid), dim3(kLdsThreads), 0, nullptr, iterations, d_sink);
        });
    return timed(launches, [&] {
        hipLaunchKernelGGL(lds_read_kernel<16>, dim3(grid), dim3(kLdsThreads), 0, nullptr, iterations, d_sink);
    });
}

// VALU issue; *work = the wave instructions of the chains a launch. kind 0: FP32 fused
// multiply-add; 1: 32-bit integer add; 2: three-input bitwise operation.
double cf_peak_valu(int kind, int iterations, uint32_t* d_sink, int launches, double* work) {
    const unsigned grid = unsigned(cus() * 8 * 4);
    *work = double(grid) * (kThreads / 64) * double(iterations) * 8.0;
    if (kind == 1)
        return timed(launches, [&] {
            hipLaunchKernelGGL(valu_int_kernel<0>, dim3(grid), dim3(kThreads), 0, nullptr, iterations, d_sink);
        });
    if (kind == 2)
        return timed(launches, [&] {
            hipLaunchKernelGGL(valu_int_kernel<1>, dim3(grid), dim3(kThreads), 0, nullptr, iterations, d_sink);
        });
    return timed(launches, [&] {
        hipLaunchKernelGGL(valu_kernel, dim3(grid), dim3(kThreads), 0, nullptr, iterations, d_sink);
    });
}

}  // extern "C"
