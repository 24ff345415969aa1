// The Monte Carlo trial on the device: random information bits, the NR encoder
// (LdpcCode::encode_block, crates/com-fec/src/ldpc.rs:753-791, and the rate matching of
// nr/ldpc.rs:136-160), BPSK or Gray QAM over AWGN (awgn.rs:94-99) and the channel LLRs
// (bpsk_llr, demapper.rs:309-311; max-log, demapper.rs:270-297); then the error counts of
// the decoded batch (the records of validation/tests/aff3ct.rs:268-276, merged as
// runner.rs:448-450 does: integer sums, so the order of the trials does not matter).
//
// Trial t of grid point p draws from Philox-4x32-10 with counters keyed by (p, t, purpose,
// index) under the run's seed, so a trial's LLRs do not depend on the batch it runs in or
// on the launch shape.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "philox.hpp"

namespace comfec {

namespace {

constexpr int kGenThreads = 256;

__global__ void __launch_bounds__(kGenThreads) generate_kernel(TrialGenArgs a) {
    extern __shared__ uint32_t lds_words[];
    const int64_t w = blockIdx.x;
    if (w >= a.count) return;
    const uint64_t trial = a.trial0 + uint64_t(w);
    const int z = a.z, core = 4 * z;
    const int info_words = (a.k_info + 31) / 32;
    const int syn_words = (core + 31) / 32;
    uint32_t* syn = lds_words;                                         // syn_words
    uint8_t* cw = reinterpret_cast<uint8_t*>(lds_words + syn_words);  // n bytes

    // information bits, 128 per Philox block; fillers K'..K are zeros
    for (int blk = threadIdx.x; blk < (info_words + 3) / 4; blk += blockDim.x) {
        const U4 r = philox4x32_10(trial_counter(a.point, trial, 0, uint32_t(blk)), a.k0, a.k1);
        const uint32_t words[4] = {r.x, r.y, r.z, r.w};
        for (int q = 0; q < 4; ++q) {
            const int wi = blk * 4 + q;
            if (wi >= info_words) break;
            uint32_t word = words[q];
            const int valid = a.k_info - wi * 32;
            if (valid < 32) word &= (1u << valid) - 1u;
            if (a.info) a.info[w * info_words + wi] = word;
            for (int b = 0; b < 32 && wi * 32 + b < a.k_info; ++b) cw[wi * 32 + b] = uint8_t((word >> b) & 1u);
        }
    }
    for (int p = a.k_info + threadIdx.x; p < a.k; p += blockDim.x) cw[p] = 0;
    for (int wi = threadIdx.x; wi < syn_words; wi += blockDim.x) syn[wi] = 0;
    __syncthreads();

    // syndromes of the information bits on the 4Z core checks
    for (int r = threadIdx.x; r < core; r += blockDim.x) {
        const int i = r / z, rr = r - i * z;
        uint32_t par = 0;
        for (int eb = a.row_start[i]; eb < a.row_start[i + 1]; ++eb) {
            const int v = int(a.col[eb]) * z;
            if (v >= a.k) continue;
            int t = rr + a.shift[eb];
            t -= t >= z ? z : 0;
            par ^= cw[v + t];
        }
        if (par) atomicOr(&syn[r >> 5], 1u << (r & 31));
    }
    __syncthreads();

    // core parity bit j: the syndromes of the checks of row j of the inverse core block
    for (int j = threadIdx.x; j < core; j += blockDim.x) {
        const uint64_t* row = a.core_inv + size_t(j) * a.core_words;
        uint32_t acc = 0;
        for (int q = 0; q < a.core_words; ++q) {
            const uint64_t s = uint64_t(syn[2 * q]) | (2 * q + 1 < syn_words ? uint64_t(syn[2 * q + 1]) << 32 : 0);
            acc ^= uint32_t(__popcll(row[q] & s));
        }
        cw[a.k + j] = uint8_t(acc & 1u);
    }
    __syncthreads();

    // extension checks close on their own parity bit, the last of their row
    for (int rr0 = core + int(threadIdx.x); rr0 < a.rows_b * z; rr0 += blockDim.x) {
        const int i = rr0 / z, rr = rr0 - i * z;
        const int own = a.k + rr0;
        uint32_t par = 0;
        for (int eb = a.row_start[i]; eb < a.row_start[i + 1]; ++eb) {
            int t = rr + a.shift[eb];
            t -= t >= z ? z : 0;
            const int v = int(a.col[eb]) * z + t;
            if (v != own) par ^= cw[v];
        }
        cw[own] = uint8_t(par);
    }
    __syncthreads();

    if (a.codeword)
        for (int wi = threadIdx.x; wi < (a.n + 31) / 32; wi += blockDim.x) {
            uint32_t word = 0;
            for (int b = 0; b < 32 && wi * 32 + b < a.n; ++b) word |= uint32_t(cw[wi * 32 + b]) << b;
            a.codeword[w * ((a.n + 31) / 32) + wi] = word;
        }

    // modulation, noise and LLRs of the E transmitted bits
    const int bps = a.bits_per_symbol;
    float noise_var = a.noise_var, noise_sd = a.noise_sd;
    if (a.snr_hi > a.snr_lo) {
        // this trial's SNR, uniform in [lo, hi] dB (Philox purpose 2)
        const U4 u = philox4x32_10(trial_counter(a.point, trial, 2, 0u), a.k0, a.k1);
        const float snr = a.snr_lo + (a.snr_hi - a.snr_lo) * (float(u.x >> 8) * (1.0f / 16777216.0f));
        const float lin = exp10f(snr / 10.0f);
        if (bps == 1) {
            noise_var = a.snr_kind == 0 ? 1.0f / (2.0f * a.rate * lin) : 1.0f / (2.0f * lin);
            noise_sd = sqrtf(noise_var);
        } else {
            noise_var = a.snr_kind == 0 ? 1.0f / (float(bps) * a.rate * lin) : 1.0f / lin;
            noise_sd = sqrtf(noise_var / 2.0f);
        }
    }
    const int nsym = a.e / bps;
    float* out = a.llr + w * a.e;
    for (int s = threadIdx.x; s < nsym; s += blockDim.x) {
        const U4 r = philox4x32_10(trial_counter(a.point, trial, 1, uint32_t(s)), a.k0, a.k1);
        // Box-Muller: two independent standard normals
        const float rad = sqrtf(-2.0f * logf(unit_open0(r.x)));
        float sn, cs;
        sincosf(6.2831853071795864769f * unit_open0(r.y), &sn, &cs);
        const float g0 = rad * cs, g1 = rad * sn;
        if (bps == 1) {
            const float x = 1.0f - 2.0f * float(cw[a.map[s]]);
            const float y = x + noise_sd * g0;
            out[s] = 2.0f * y / noise_var;
            continue;
        }
        int label = 0;
        for (int l = 0; l < bps; ++l) label = (label << 1) | cw[a.map[s * bps + l]];
        const float yr = a.pts_re[label] + noise_sd * g0;
        const float yi = a.pts_im[label] + noise_sd * g1;
        const int m = 1 << bps;
        for (int l = 0; l < bps; ++l) {
            float best0 = -INFINITY, best1 = -INFINITY;
            for (int i = 0; i < m; ++i) {
                const float dr = yr - a.pts_re[i], di = yi - a.pts_im[i];
                const float metric = -(dr * dr + di * di) / noise_var + a.log_prior;
                if ((i >> (bps - 1 - l)) & 1) best1 = fmaxf(best1, metric);
                else best0 = fmaxf(best0, metric);
            }
            out[s * bps + l] = best0 - best1;
        }
    }
}

__global__ void __launch_bounds__(256) count_kernel(CountArgs a) {
    const int64_t w = blockIdx.x;
    if (w >= a.count) return;
    const int info_words = (a.k_info + 31) / 32;
    uint32_t errors = 0;
    for (int p = threadIdx.x; p < a.k_info; p += blockDim.x) {
        const uint32_t bit = (a.info[w * info_words + (p >> 5)] >> (p & 31)) & 1u;
        errors += uint32_t(a.bits[w * a.k_info + p] != bit);
    }
    // wavefront sums, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) errors += __shfl_down(errors, off, 64);
    __shared__ uint32_t per_wave[4];
    if ((threadIdx.x & 63) == 0) per_wave[threadIdx.x >> 6] = errors;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int q = 0; q < int(blockDim.x >> 6); ++q) total += per_wave[q];
        unsigned long long* c = reinterpret_cast<unsigned long long*>(a.counters);
        atomicAdd(&c[0], (unsigned long long)total);
        atomicAdd(&c[1], (unsigned long long)(total > 0));
        if (a.iterations) {
            const unsigned long long it = a.iterations[w];
            atomicAdd(&c[2], it);
            atomicAdd(&c[3], it * it);
        }
        if (a.per_trial_errors) a.per_trial_errors[w] = total;
    }
}

__global__ void zero_kernel(uint64_t* p, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}

}  // namespace

hipError_t launch_zero(uint64_t* p, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

size_t generate_lds_bytes(int z, int n) { return size_t((4 * z + 31) / 32) * 4 + size_t(n); }

hipError_t launch_generate(const TrialGenArgs& a, hipStream_t stream) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(generate_kernel, dim3(unsigned(a.count)), dim3(kGenThreads), generate_lds_bytes(a.z, a.n),
                       stream, a);
    return hipGetLastError();
}

hipError_t launch_count(const CountArgs& a, hipStream_t stream) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(count_kernel, dim3(unsigned(a.count)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace comfec
