// Training kernels of the learned offset/scale min-sum (no reference counterpart: the
// reference trains no decoder, README.md:109-111). The decoder is the layered schedule of
// crates/com-fec/src/ldpc.rs:509-531 with the check node of ldpc.rs:468-491 whose
// magnitude map, for iteration t and base row i, is g(m) = a[t][i] * max(m - b[t][i], 0).
//
// Loss: the mean over the T unrolled iterations and the N codeword bits of the binary
// cross-entropy of sigmoid(-post) against the bit, softplus(-s post), s = 1 - 2c.
//
// Forward: one workgroup per codeword, lane r on row r of every layer, posteriors in LDS.
// Every iteration's variable-to-check messages are checkpointed to HBM ([t][base edge][r],
// coalesced): they are all the backward pass needs, because a row's check-to-variable
// messages of iteration t are a function of its inputs of iteration t, so the forward
// recomputes the messages it subtracts from the checkpoint of the iteration before, and
// the backward recomputes each row's check node from the checkpoint of its iteration.
//
// Backward: layers and iterations in reverse. With O the adjoint of a row's outputs and
// I of its inputs: dL/da += O s max(m - b, 0), dL/db -= O s a [m > b], the output adjoints
// of the edges that took m1 (resp. m2) flow to the input that holds m1 (resp. m2) through
// d|x|/dx = sign(x); the posterior adjoint becomes I and the adjoint of the previous
// iteration's output on the edge becomes -I. Parameter gradients are reduced per wave
// (shuffles), per workgroup (fixed order) and across workgroups by a separate
// fixed-order kernel, so that a training step is deterministic.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace comfec {

namespace {

constexpr int kMaxDeg = 19;

struct RowScan {
    float m1, m2;
    int i1, i2;
    uint32_t neg;  // bit q: input q negative
    uint32_t parity;
};

__device__ inline RowScan scan_row(const float (&in)[kMaxDeg], int deg) {
    RowScan s{INFINITY, INFINITY, 0, 0, 0u, 0u};
#pragma unroll
    for (int q = 0; q < kMaxDeg; ++q)
        if (q < deg) {
            const float a = fabsf(in[q]);
            const uint32_t ng = in[q] < 0.0f;
            s.neg |= ng << q;
            s.parity ^= ng;
            if (a < s.m1) {
                s.m2 = s.m1;
                s.i2 = s.i1;
                s.m1 = a;
                s.i1 = q;
            } else if (a < s.m2) {
                s.m2 = a;
                s.i2 = q;
            }
        }
    return s;
}

__device__ inline float out_of(const RowScan& s, int q, float g1, float g2) {
    const float mag = q == s.i1 ? g2 : g1;
    return ((s.parity ^ (s.neg >> q)) & 1u) ? -mag : mag;
}

__device__ inline int wrap_add(int r, int v, int z) {
    const int t = r + v;
    return t >= z ? t - z : t;
}

__device__ inline float softplusf(float x) { return x > 0.0f ? x + log1pf(expf(-x)) : log1pf(expf(x)); }
__device__ inline float sigmoidf(float x) {
    return x >= 0.0f ? 1.0f / (1.0f + expf(-x)) : expf(x) / (1.0f + expf(x));
}

__global__ void __launch_bounds__(384) learn_forward_kernel(LearnArgs a) {
    extern __shared__ float post[];
    const DevCode& c = a.code;
    const int64_t w = blockIdx.x;
    if (w >= a.count) return;
    const int r = threadIdx.x;
    const size_t edges = size_t(c.edges_b) * c.z;
    float* ckpt = a.ckpt + size_t(w) * a.T * edges;
    const int cw_words = (c.n + 31) / 32;
    const uint32_t* cw = a.cw + size_t(w) * cw_words;

    for (int p = threadIdx.x; p < c.n; p += blockDim.x) {
        float v = 0.0f;
        for (uint32_t s = c.src_start[p]; s < c.src_start[p + 1]; ++s) v += a.llr[w * c.e + c.src[s]];
        if (p >= c.k_info && p < c.k) v = 30.0f;
        post[p] = v;
    }
    __syncthreads();
    float loss = 0.0f;
    // the checkpoint of iteration t-1 of row i (read one row ahead: nothing of this
    // iteration writes it, so its reads need not wait for the barrier)
    auto load_prev = [&](int t, int i, float (&dst)[kMaxDeg]) {
        const int e0 = c.row_start[i], deg = c.row_start[i + 1] - e0;
#pragma unroll
        for (int q = 0; q < kMaxDeg; ++q)
            if (q < deg) dst[q] = ckpt[size_t(t - 1) * edges + size_t(e0 + q) * c.z + r];
    };
    for (int t = 0; t < a.T; ++t) {
        float prev_next[kMaxDeg] = {};
        if (t > 0 && r < c.z) load_prev(t, 0, prev_next);
        for (int i = 0; i < c.rows_b; ++i) {
            if (r < c.z) {
                const int e0 = c.row_start[i], deg = c.row_start[i + 1] - e0;
                float old[kMaxDeg];
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q) old[q] = 0.0f;
                if (t > 0) {
                    // the row's messages of iteration t-1, recomputed from its checkpoint
                    float prev[kMaxDeg];
#pragma unroll
                    for (int q = 0; q < kMaxDeg; ++q) prev[q] = prev_next[q];
                    if (i + 1 < c.rows_b) load_prev(t, i + 1, prev_next);
                    const RowScan s = scan_row(prev, deg);
                    const float pa = a.params[((t - 1) * c.rows_b + i) * 2], pb = a.params[((t - 1) * c.rows_b + i) * 2 + 1];
                    const float g1 = pa * fmaxf(s.m1 - pb, 0.0f), g2 = pa * fmaxf(s.m2 - pb, 0.0f);
#pragma unroll
                    for (int q = 0; q < kMaxDeg; ++q)
                        if (q < deg) old[q] = out_of(s, q, g1, g2);
                }
                float in[kMaxDeg];
                int addr[kMaxDeg];
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) {
                        addr[q] = int(c.col[e0 + q]) * c.z + wrap_add(r, c.shift[e0 + q], c.z);
                        in[q] = post[addr[q]] - old[q];
                        ckpt[size_t(t) * edges + size_t(e0 + q) * c.z + r] = in[q];
                    }
                const RowScan s = scan_row(in, deg);
                const float pa = a.params[(t * c.rows_b + i) * 2], pb = a.params[(t * c.rows_b + i) * 2 + 1];
                const float g1 = pa * fmaxf(s.m1 - pb, 0.0f), g2 = pa * fmaxf(s.m2 - pb, 0.0f);
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) post[addr[q]] = in[q] + out_of(s, q, g1, g2);
            }
            __syncthreads();
        }
        // the loss of iteration t and its gradient with respect to the posteriors
        float* gpost = a.gpost + (size_t(w) * a.T + t) * c.n;
        for (int p = threadIdx.x; p < c.n; p += blockDim.x) {
            const float s = 1.0f - 2.0f * float((cw[p >> 5] >> (p & 31)) & 1u);
            const float x = post[p];
            loss += a.scale * softplusf(-s * x);
            gpost[p] = a.scale * (-s) * sigmoidf(-s * x);
        }
        __syncthreads();
    }
    if (a.loss_out) {
        // per-codeword loss, summed in a fixed order
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
        __shared__ float wl[8];
        if ((threadIdx.x & 63) == 0) wl[threadIdx.x >> 6] = loss;
        __syncthreads();
        if (threadIdx.x == 0) {
            float tot = 0.0f;
            for (int q = 0; q < int(blockDim.x >> 6); ++q) tot += wl[q];
            a.loss_out[w] = tot;
        }
    }
}

__global__ void __launch_bounds__(384) learn_backward_kernel(LearnArgs a) {
    extern __shared__ float lds[];
    const DevCode& c = a.code;
    const int64_t w = blockIdx.x;
    if (w >= a.count) return;
    const int r = threadIdx.x;
    const int waves = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int np = a.T * c.rows_b * 2;
    float* pbar = lds;                  // n
    float* wacc = lds + c.n;            // waves x np
    const size_t edges = size_t(c.edges_b) * c.z;
    const float* ckpt = a.ckpt + size_t(w) * a.T * edges;
    float* cbar = a.cbar + size_t(w) * edges;
    for (int p = threadIdx.x; p < c.n; p += blockDim.x) pbar[p] = 0.0f;
    for (int p = threadIdx.x; p < waves * np; p += blockDim.x) wacc[p] = 0.0f;
    __syncthreads();
    for (int t = a.T - 1; t >= 0; --t) {
        const float* gpost = a.gpost + (size_t(w) * a.T + t) * c.n;
        for (int p = threadIdx.x; p < c.n; p += blockDim.x) pbar[p] += gpost[p];
        __syncthreads();
        // a row's checkpoint and its message adjoints of iteration t+1 are read one row
        // ahead: nothing of this iteration before the row itself writes them
        float in_next[kMaxDeg] = {}, cb_next[kMaxDeg] = {};
        auto load_row = [&](int i) {
            const int e0 = c.row_start[i], deg = c.row_start[i + 1] - e0;
#pragma unroll
            for (int q = 0; q < kMaxDeg; ++q)
                if (q < deg) {
                    in_next[q] = ckpt[size_t(t) * edges + size_t(e0 + q) * c.z + r];
                    cb_next[q] = t < a.T - 1 ? cbar[size_t(e0 + q) * c.z + r] : 0.0f;
                }
        };
        if (r < c.z) load_row(c.rows_b - 1);
        for (int i = c.rows_b - 1; i >= 0; --i) {
            float ga = 0.0f, gb = 0.0f;
            if (r < c.z) {
                const int e0 = c.row_start[i], deg = c.row_start[i + 1] - e0;
                float in[kMaxDeg], cb[kMaxDeg];
                int addr[kMaxDeg];
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q) {
                    in[q] = in_next[q];
                    cb[q] = cb_next[q];
                }
                if (i > 0) load_row(i - 1);
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) addr[q] = int(c.col[e0 + q]) * c.z + wrap_add(r, c.shift[e0 + q], c.z);
                const RowScan s = scan_row(in, deg);
                const float pa = a.params[(t * c.rows_b + i) * 2], pb = a.params[(t * c.rows_b + i) * 2 + 1];
                float ibar[kMaxDeg];
                float m1bar = 0.0f, m2bar = 0.0f;
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) {
                        const float ob = pbar[addr[q]] + cb[q];
                        const float m = q == s.i1 ? s.m2 : s.m1;
                        const float sign = ((s.parity ^ (s.neg >> q)) & 1u) ? -1.0f : 1.0f;
                        const float active = m > pb ? 1.0f : 0.0f;
                        ga += ob * sign * fmaxf(m - pb, 0.0f);
                        gb += ob * sign * (-pa) * active;
                        if (q == s.i1) m2bar += ob * sign * pa * active;
                        else m1bar += ob * sign * pa * active;
                        ibar[q] = pbar[addr[q]];
                    }
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) {
                        if (q == s.i1) ibar[q] += m1bar * (in[q] > 0.0f ? 1.0f : in[q] < 0.0f ? -1.0f : 0.0f);
                        if (q == s.i2) ibar[q] += m2bar * (in[q] > 0.0f ? 1.0f : in[q] < 0.0f ? -1.0f : 0.0f);
                    }
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) {
                        pbar[addr[q]] = ibar[q];
                        cbar[size_t(e0 + q) * c.z + r] = -ibar[q];
                    }
            }
            // wave sums in a fixed order, one slot per wave
            for (int off = 32; off > 0; off >>= 1) {
                ga += __shfl_xor(ga, off, 64);
                gb += __shfl_xor(gb, off, 64);
            }
            if (lane == 0) {
                wacc[wave * np + (t * c.rows_b + i) * 2] += ga;
                wacc[wave * np + (t * c.rows_b + i) * 2 + 1] += gb;
            }
            __syncthreads();
        }
    }
    // the workgroup's gradient, waves summed in order
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
        float s = 0.0f;
        for (int q = 0; q < waves; ++q) s += wacc[q * np + p];
        a.grad_partial[size_t(w) * np + p] = s;
    }
}

// grad[p] = sum over codewords of grad_partial[cw][p], each parameter summed by one
// workgroup in a fixed tree order; loss likewise (parameter index np)
__global__ void __launch_bounds__(256) reduce_kernel(const float* partial, const float* loss_cw, int64_t count,
                                                     int np, float* grad, float* loss) {
    const int p = blockIdx.x;
    __shared__ float acc[256];
    float s = 0.0f;
    for (int64_t w = threadIdx.x; w < count; w += blockDim.x)
        s += p < np ? partial[size_t(w) * np + p] : loss_cw[w];
    acc[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) acc[threadIdx.x] += acc[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (p < np) grad[p] = acc[0];
        else if (loss) *loss = acc[0];
    }
}

__global__ void adam_kernel(float* params, const float* grad, float* m, float* v, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

}  // namespace

hipError_t launch_learn_forward(const LearnArgs& a, hipStream_t stream) {
    if (a.count <= 0) return hipSuccess;
    const int threads = ((a.code.z + 63) / 64) * 64;
    const size_t lds = size_t(a.code.n) * 4;
    hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(learn_forward_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(learn_forward_kernel, dim3(unsigned(a.count)), dim3(threads), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_learn_backward(const LearnArgs& a, hipStream_t stream) {
    if (a.count <= 0) return hipSuccess;
    const int threads = ((a.code.z + 63) / 64) * 64;
    const size_t np = size_t(a.T) * a.code.rows_b * 2;
    const size_t lds = size_t(a.code.n) * 4 + size_t(threads / 64) * np * 4;
    if (lds > kLdsBudget + 8 * 1024) return hipErrorInvalidValue;
    hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(learn_backward_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(learn_backward_kernel, dim3(unsigned(a.count)), dim3(threads), lds, stream, a);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(reduce_kernel, dim3(unsigned(np + 1)), dim3(256), 0, stream, a.grad_partial, a.loss_out,
                       a.count, int(np), a.grad, a.loss);
    return hipGetLastError();
}

hipError_t launch_adam(float* params, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t stream) {
    const float c1 = 1.0f - powf(b1, float(step)), c2 = 1.0f - powf(b2, float(step));
    hipLaunchKernelGGL(adam_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, params, grad, m, v, n, lr,
                       b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace comfec
