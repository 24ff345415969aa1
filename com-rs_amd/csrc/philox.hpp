// Philox-4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11): a counter-based generator, so that the draws of a
// trial depend only on (seed, grid point, trial, purpose, index) and never on the launch
// shape. Host and device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace comfec {

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = uint64_t(a) * uint64_t(b);
    hi = uint32_t(p >> 32);
    lo = uint32_t(p);
}

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c.x, hi0, lo0);
        mulhilo32(M1, c.z, hi1, lo1);
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// The counter of draw `index` of stream `purpose` of trial `trial` at grid point `point`.
// Purposes: 0 information bits, 1 channel noise.
__host__ __device__ inline U4 trial_counter(uint32_t point, uint64_t trial, uint32_t purpose, uint32_t index) {
    return U4{index, uint32_t(trial), uint32_t(trial >> 32), (point << 4) | (purpose & 15u)};
}

// uniform on (0, 1]: 24 bits
__host__ __device__ inline float unit_open0(uint32_t x) { return float((x >> 8) + 1u) * (1.0f / 16777216.0f); }

}  // namespace comfec
