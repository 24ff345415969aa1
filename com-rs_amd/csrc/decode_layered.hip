// The fast layered min-sum decoder of the NR codes (FP32 messages): the layered branch
// of LdpcCode::propagate (crates/com-fec/src/ldpc.rs:509-531) with the min-sum check
// node of ldpc.rs:468-491 and the syndrome stop of ldpc.rs:546-557.
//
// One workgroup decodes one codeword; lane r owns lifted row r of every layer (base row),
// so a layer is Z rows updated at once, which is the reference's row-after-row order (the
// Z rows of a circulant layer share no variable). The base graph is a template parameter:
// the layer loop is unrolled at compile time, each layer with its own degree and columns,
// so that every lane keeps the check-to-variable messages of all its rows in registers, in
// the compressed form min-sum allows: the two output magnitudes g(m1), g(m2), the index of
// the first minimum and the output sign bits. The posterior LLRs stay in LDS; HBM sees the
// channel LLRs once and the decisions once.
//
// Check node: out_q = sign_q * g(q == i1 ? m2 : m1), g(m) = a * max(m - b, 0): a = alpha,
// b = 0 is the normalized min-sum, a = 1, b = beta the offset min-sum, and per-iteration,
// per-layer (a, b) the learned offset/scale min-sum. All operations are single IEEE FP32
// operations (no contraction), so the result equals the reference's arithmetic rounded to
// FP32 step by step.
//
// Syndrome: when Z is a multiple of 32 the hard decisions are packed (wavefront ballots)
// into a doubled bit vector per base column, and a row word of 32 parities is the XOR of
// funnel-shifted column words; otherwise every lane checks its rows edge by edge.
#include <hip/hip_runtime.h>

#define NR_TABLE_CONST constexpr
#include "kernels.hpp"
#include "nr_base_graphs.inc"

namespace comfec {

namespace {

template <int BG>
struct Graph;
template <>
struct Graph<1> {
    static constexpr int rows = 46, cols = 68, edges = NR_BG1_ENTRIES, max_deg = 19;
    static constexpr const nr_bg_entry* t = NR_BG1;
};
template <>
struct Graph<2> {
    static constexpr int rows = 42, cols = 52, edges = NR_BG2_ENTRIES, max_deg = 10;
    static constexpr const nr_bg_entry* t = NR_BG2;
};

template <int BG>
constexpr int row_begin(int i) {
    int s = 0;
    for (int e = 0; e < Graph<BG>::edges; ++e) s += Graph<BG>::t[e].i < i;
    return s;
}

// Whether layer I must wait for the layers before it: it shares a column with a layer
// processed since the last barrier (reads after writes; and writes after reads, which are
// the same shared columns). Layers on disjoint columns commute, so they run without one.
template <int BG>
constexpr bool shares_column(int i, int l) {
    for (int e = 0; e < Graph<BG>::edges; ++e)
        if (Graph<BG>::t[e].i == i)
            for (int f = 0; f < Graph<BG>::edges; ++f)
                if (Graph<BG>::t[f].i == l && Graph<BG>::t[f].j == Graph<BG>::t[e].j) return true;
    return false;
}

template <int BG>
constexpr bool barrier_before(int layer) {
    if (layer == 0) return false;
    int first = 0;  // first layer after the last barrier
    for (int l = 1; l <= layer; ++l) {
        bool conflict = false;
        for (int p = first; p < l; ++p) conflict = conflict || shares_column<BG>(l, p);
        if (conflict) {
            if (l == layer) return true;
            first = l;
        }
    }
    return false;
}

// The messages of one lane's rows: output magnitudes g(m1), g(m2) and, packed, the
// output sign bits (bit q of edge q) and the index of the first minimum (bits 24..28).
template <int ROWS>
struct Messages {
    float g1[ROWS], g2[ROWS];
    uint32_t pk[ROWS];
};

__device__ inline int wrap_add(int r, int v, int z) {
    int t = r + v;
    return t >= z ? t - z : t;
}

__device__ inline float with_sign(float mag, uint32_t bit) {
    return __uint_as_float(__float_as_uint(mag) | (bit << 31));
}

// ldpc.rs:522-531 for the row of lane r in layer I.
//
// Exactness of the shortcuts, given that no posterior or message is ever -0.0 (the LLRs
// are summed onto +0.0 when they are placed, and a difference or sum of two values that are
// not -0.0 is not -0.0 except for -0.0 - +0.0):
//  - the sign bit of a message is the reference's test `l < 0.0`, so the output sign of
//    edge q, the product of the other input signs, is bit 31 of (input_q XOR all inputs);
//  - the edge that takes the second minimum is told by |input_q| == m1: it is the first
//    minimum, or a tie with it, and a tie makes m2 = m1, so either way the magnitude is the
//    reference's g(q == i1 ? m2 : m1);
//  - m2 = med3(m1, m2, |input|) is the strict-< update of the scan of ldpc.rs:471-480.
// The stored state keeps the output sign bits (edge q at bit deg-1-q, shifted in with a
// funnel shift) and one index whose magnitude is g(m2).
template <int BG, int I>
__device__ __forceinline__ void layer(float* __restrict__ post, int r, int z, const uint32_t* __restrict__ shift,
                                      Messages<Graph<BG>::rows>& msg, float a, float b) {
    constexpr int e0 = row_begin<BG>(I);
    constexpr int deg = row_begin<BG>(I + 1) - e0;
    const float g1o = msg.g1[I], g2o = msg.g2[I];
    const uint32_t pko = msg.pk[I];
    const uint32_t i1o = pko >> 24;
    float v2c[deg];
    int addr[deg];
    float m1 = INFINITY, m2 = INFINITY;
    uint32_t parity = 0;
#pragma unroll
    for (int q = 0; q < deg; ++q) {
        // (r + shift) mod z: the unsigned min of t and t - z (which wraps when t < z)
        const uint32_t t = uint32_t(r) + shift[e0 + q];
        const uint32_t rr = min(t, t - uint32_t(z));
        addr[q] = int(Graph<BG>::t[e0 + q].j) * z + int(rr);
        const float mag = uint32_t(q) == i1o ? g2o : g1o;
        const float old = __uint_as_float(__float_as_uint(mag) ^ ((pko << (32 - deg + q)) & 0x80000000u));
        const float in = post[addr[q]] - old;
        v2c[q] = in;
        parity ^= __float_as_uint(in);
        const float m = fabsf(in);
        m2 = __builtin_amdgcn_fmed3f(m1, m2, m);
        m1 = fminf(m1, m);
    }
    const float g1 = a * fmaxf(m1 - b, 0.0f);
    const float g2 = a * fmaxf(m2 - b, 0.0f);
    parity &= 0x80000000u;
    uint32_t signs = 0, i1 = 0;
#pragma unroll
    for (int q = 0; q < deg; ++q) {
        const bool first = fabsf(v2c[q]) == m1;
        const float mag = first ? g2 : g1;
        i1 = first ? uint32_t(q) : i1;
        const uint32_t sb = __float_as_uint(v2c[q]) ^ parity;
        signs = __builtin_amdgcn_alignbit(signs, sb, 31u);
        const float out = __uint_as_float(__float_as_uint(mag) ^ (sb & 0x80000000u));
        post[addr[q]] = v2c[q] + out;
    }
    msg.g1[I] = g1;
    msg.g2[I] = g2;
    msg.pk[I] = signs | (i1 << 24);
}

template <int BG, int I>
__device__ __forceinline__ void sweep(float* post, int r, int z, bool active, const uint32_t* shift,
                                      Messages<Graph<BG>::rows>& msg, const float* ab, float a, float b) {
    if constexpr (I < Graph<BG>::rows) {
        if (ab) {
            a = ab[2 * I];
            b = ab[2 * I + 1];
        }
        if constexpr (barrier_before<BG>(I)) __syncthreads();
        if (active) layer<BG, I>(post, r, z, shift, msg, a, b);
        sweep<BG, I + 1>(post, r, z, active, shift, msg, ab, a, b);
    }
}

// whether a check fails (ldpc.rs:555-557), packed path: Z a multiple of 32
struct Tables {
    const uint32_t* shift;
    const uint16_t* col;
    const uint16_t* row_start;
};

template <int BG>
__device__ bool syndrome_fails_packed(const float* post, uint32_t* hd, const LayeredArgs& a, const Tables& tb) {
    const int z = a.z, w32 = z >> 5, n = Graph<BG>::cols * z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    for (int c = wave; c < (n >> 6); c += waves) {
        const uint64_t m = __ballot(post[c * 64 + lane] < 0.0f);
        if (lane < 2) {
            const int f = 2 * c + lane;
            const int j = f / w32, w = f - j * w32;
            const uint32_t word = lane ? uint32_t(m >> 32) : uint32_t(m);
            hd[j * 2 * w32 + w] = word;
            hd[j * 2 * w32 + w32 + w] = word;
        }
    }
    __syncthreads();
    uint32_t bad = 0;
    for (int task = threadIdx.x; task < Graph<BG>::rows * w32; task += blockDim.x) {
        const int i = task / w32, w = task - i * w32;
        uint32_t par = 0;
        for (int eb = tb.row_start[i]; eb < tb.row_start[i + 1]; ++eb) {
            const int v = int(tb.shift[eb]);
            const int base = tb.col[eb] * 2 * w32 + w + (v >> 5);
            par ^= __builtin_amdgcn_alignbit(hd[base + 1], hd[base], uint32_t(v & 31));
        }
        bad |= par;
    }
    return __syncthreads_or(bad != 0) != 0;
}

// general Z: every lane checks its rows edge by edge
template <int BG>
__device__ bool syndrome_fails_rows(const float* post, const LayeredArgs& a, const Tables& tb) {
    const int z = a.z;
    int bad = 0;
    for (int r = threadIdx.x; r < z; r += blockDim.x)
        for (int i = 0; i < Graph<BG>::rows && !bad; ++i) {
            int par = 0;
            for (int eb = tb.row_start[i]; eb < tb.row_start[i + 1]; ++eb)
                par ^= post[tb.col[eb] * z + wrap_add(r, int(tb.shift[eb]), z)] < 0.0f;
            bad |= par;
        }
    return __syncthreads_or(bad) != 0;
}

template <int BG>
__device__ __forceinline__ bool syndrome_fails(const float* post, uint32_t* hd, const LayeredArgs& a,
                                               const Tables& tb) {
    if ((a.z & 31) == 0) return syndrome_fails_packed<BG>(post, hd, a, tb);
    return syndrome_fails_rows<BG>(post, a, tb);
}

template <int BG>
__global__ void __launch_bounds__(384) decode_layered_kernel(LayeredArgs a) {
    extern __shared__ float lds[];
    const int z = a.z, n = Graph<BG>::cols * z;
    float* post = lds;
    uint32_t* hd = reinterpret_cast<uint32_t*>(lds + n);
    const int r = threadIdx.x;
    const bool active = r < z;
    int64_t w = blockIdx.x;
    if (a.index) {
        // a subset: the listed codewords (the count lives on the device)
        if (uint32_t(w) >= *a.index_count) return;
        w = a.index[w];
    } else if (w >= a.count) {
        return;
    }

    // the shifts of the layers (read with uniform addresses) and, for the syndrome, the
    // columns and row starts, in LDS
    __shared__ uint32_t shift_lds[Graph<BG>::edges];
    __shared__ uint16_t col_lds[Graph<BG>::edges];
    __shared__ uint16_t rs_lds[Graph<BG>::rows + 1];
    for (int e = threadIdx.x; e < Graph<BG>::edges; e += blockDim.x) {
        shift_lds[e] = a.shift32[e];
        col_lds[e] = a.col[e];
    }
    for (int i = threadIdx.x; i <= Graph<BG>::rows; i += blockDim.x) rs_lds[i] = a.row_start[i];

    // channel LLRs of the full codeword (nr/ldpc.rs:191-197): repeats summed in output
    // order, 0.0 on unsent bits, 30.0 on fillers
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        float v = 0.0f;
        for (uint32_t s = a.src_start[p]; s < a.src_start[p + 1]; ++s) v += a.llr[w * a.e + a.src[s]];
        if (p >= a.k_info && p < a.k) v = 30.0f;
        post[p] = v;
    }
    __syncthreads();

    Messages<Graph<BG>::rows> msg;
#pragma unroll
    for (int i = 0; i < Graph<BG>::rows; ++i) {
        msg.g1[i] = 0.0f;
        msg.g2[i] = 0.0f;
        msg.pk[i] = 0u;
    }
    const Tables tb{shift_lds, col_lds, rs_lds};
    int done = syndrome_fails<BG>(post, hd, a, tb) ? -1 : 0;
    for (int it = 1; it <= a.max_iterations && done < 0; ++it) {
        const float* ab = a.params ? a.params + size_t(it - 1) * 2 * Graph<BG>::rows : nullptr;
        sweep<BG, 0>(post, r, z, active, shift_lds, msg, ab, a.a, a.b);
        __syncthreads();
        if (!syndrome_fails<BG>(post, hd, a, tb)) done = it;
    }
    if (a.ok && threadIdx.x == 0) a.ok[w] = done >= 0;
    if (done < 0) done = a.max_iterations;
    if (threadIdx.x == 0 && a.iterations) a.iterations[w] = uint32_t(done) + (a.add_iterations ? a.iterations[w] : 0u);
    if (a.bits)
        for (int p = threadIdx.x; p < a.k_info; p += blockDim.x) a.bits[w * a.k_info + p] = post[p] < 0.0f;
    if (a.post_out)
        for (int p = threadIdx.x; p < n; p += blockDim.x) a.post_out[w * n + p] = post[p];
}

template <int BG>
hipError_t launch_bg(const LayeredArgs& a, hipStream_t stream) {
    const int threads = ((a.z + 63) / 64) * 64;
    const int n = Graph<BG>::cols * a.z;
    const size_t lds = size_t(n) * 4 + size_t(Graph<BG>::cols) * 2 * ((a.z + 31) / 32) * 4;
    if (lds > kLdsBudget + 8 * 1024) return hipErrorInvalidValue;
    // the dynamic LDS of this launch (the static tables come on top of it)
    hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(decode_layered_kernel<BG>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(decode_layered_kernel<BG>, dim3(unsigned(a.count)), dim3(threads), lds, stream, a);
    return hipGetLastError();
}

}  // namespace

bool layered_supported(int bg, int z) { return (bg == 1 || bg == 2) && z <= 384; }

hipError_t launch_layered(const LayeredArgs& a, int bg, hipStream_t stream) {
    if (a.count <= 0) return hipSuccess;
    return bg == 1 ? launch_bg<1>(a, stream) : launch_bg<2>(a, stream);
}

}  // namespace comfec
