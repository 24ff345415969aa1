// The fast layered min-sum decoder of the NR codes: the layered branch of
// LdpcCode::propagate (crates/com-fec/src/ldpc.rs:509-531) with the min-sum check node of
// ldpc.rs:468-491 and the syndrome stop of ldpc.rs:546-557, for batches of codewords of
// one code or of many codes (mixed base graphs, lifting sizes and iteration limits).
//
// Work is split into tasks: a task is up to G codewords of one code (a "class"), decoded
// side by side by one workgroup. Lane l of the workgroup owns row r = l mod Z of every
// layer (base row) of codeword g = l / Z, so a layer is Z rows of each codeword updated at
// once, which is the reference's row-after-row order (the Z rows of a circulant layer share
// no variable). Workgroups take tasks in order from a counter (or block by block).
//
// The base graph is a template parameter of the task body: the layer loop is unrolled at
// compile time, each layer with its own degree and columns, so that every lane keeps the
// check-to-variable messages of all its rows in registers in the compressed form min-sum
// allows: the two output magnitudes g(m1), g(m2), the index of the edge that takes g(m2)
// and the output sign bits. The posterior LLRs stay in LDS; HBM sees the channel LLRs once
// and the decisions once.
//
// Check node: out_q = sign_q * g(q == i1 ? m2 : m1), g(m) = a * max(m - b, 0): a = alpha,
// b = 0 is the normalized min-sum, a = 1, b = beta the offset min-sum, and per-iteration,
// per-layer (a, b) the learned offset/scale min-sum. All operations are single IEEE FP32
// operations (no contraction), so the result equals the reference's arithmetic rounded to
// FP32 step by step.
//
// FP16 storage (the HALF instantiation): the posteriors are stored in LDS as binary16 and
// the magnitudes g(m1), g(m2) as binary16 pairs; the arithmetic stays FP32. A posterior is
// stored as half(clamp(x, +-65504)) with a zero stored as +0.0, a magnitude as
// half(min(g, 65504)), both rounded to nearest even.
//
// Learned decoding: `params` gives (a, b) per iteration and layer for the first
// learned_iterations iterations; with `fallback` set, a codeword whose syndrome still fails
// is decoded again from its channel LLRs by the class's own rule, in the same workgroup,
// and its iteration count is the sum of both runs.
//
// Syndrome: when Z is a multiple of 32 the hard decisions are packed (wavefront ballots)
// into a doubled bit vector per base column, and a row word of 32 parities is the XOR of
// funnel-shifted column words; otherwise every lane checks its rows edge by edge.
#pragma once
#include <hip/hip_runtime.h>

#define NR_TABLE_CONST constexpr
#include "kernels.hpp"
#include "nr_base_graphs.inc"

namespace comfec {

namespace layered {

template <int BG>
struct Graph;
template <>
struct Graph<1> {
    static constexpr int rows = 46, cols = 68, edges = NR_BG1_ENTRIES;
    static constexpr const nr_bg_entry* t = NR_BG1;
};
template <>
struct Graph<2> {
    static constexpr int rows = 42, cols = 52, edges = NR_BG2_ENTRIES;
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

// The lifting set of Z and the shift V mod Z of base entry e, when Z is known at compile
// time (nr/ldpc.rs:117-133).
constexpr int lifting_set_of(int z) {
    for (int s = 0; s < 8; ++s)
        for (int k = 0; k < NR_LIFTING_SET_MAX; ++k)
            if (NR_LIFTING_SETS[s][k] == z) return s;
    return -1;
}

template <int BG, int ZC>
constexpr uint32_t shift_of(int e) {
    return uint32_t(Graph<BG>::t[e].v[lifting_set_of(ZC)] % ZC);
}

// ---- storage of the posteriors and of the message magnitudes ----------------------------

template <bool HALF>
struct Store;

// a pointer into LDS made from a byte address
template <typename T>
struct lds_ptr {
    typedef __attribute__((address_space(3))) T* type;
};

template <>
struct Store<false> {
    using T = float;
    static __device__ __forceinline__ float ld(const float* p, int i) { return p[i]; }
    static __device__ __forceinline__ void st(float* p, int i, float x) { p[i] = x; }
    static __device__ __forceinline__ uint32_t sign(const float* p, int i) { return __float_as_uint(p[i]) >> 31; }
    // at LDS byte address `off` (the kernel's posteriors start at LDS address 0)
    static __device__ __forceinline__ float ld_at(uint32_t off) { return *(const lds_ptr<float>::type)(uintptr_t)off; }
    static __device__ __forceinline__ void st_at(uint32_t off, float x) { *(lds_ptr<float>::type)(uintptr_t)off = x; }
};

__device__ __forceinline__ float h2f(uint32_t h) {
    return float(__builtin_bit_cast(_Float16, uint16_t(h)));
}
__device__ __forceinline__ uint32_t f2h(float x) {
    return uint32_t(__builtin_bit_cast(uint16_t, _Float16(x)));
}

template <>
struct Store<true> {
    using T = uint16_t;  // binary16 bits
    static __device__ __forceinline__ float ld(const uint16_t* p, int i) { return h2f(p[i]); }
    // half(clamp(x)), a zero stored as +0.0 (the half sum -0 + +0 is +0, and x + 0 is x
    // for every other half)
    static __device__ __forceinline__ uint16_t to_store(float x) {
        const _Float16 h = _Float16(__builtin_amdgcn_fmed3f(x, -65504.0f, 65504.0f));
        return __builtin_bit_cast(uint16_t, _Float16(h + _Float16(0.0f)));
    }
    static __device__ __forceinline__ void st(uint16_t* p, int i, float x) { p[i] = to_store(x); }
    static __device__ __forceinline__ uint32_t sign(const uint16_t* p, int i) { return uint32_t(p[i]) >> 15; }
    static __device__ __forceinline__ float ld_at(uint32_t off) {
        return h2f(*(const lds_ptr<uint16_t>::type)(uintptr_t)off);
    }
    static __device__ __forceinline__ void st_at(uint32_t off, float x) {
        *(lds_ptr<uint16_t>::type)(uintptr_t)off = to_store(x);
    }
};

// The messages of one lane's rows: output magnitudes g(m1), g(m2) and, packed, the
// output sign bits and the index of the edge that takes g(m2) (bits 24..28).
template <int ROWS, bool HALF>
struct Messages;

template <int ROWS>
struct Messages<ROWS, false> {
    float g1[ROWS], g2[ROWS];
    uint32_t pk[ROWS];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
            g1[i] = 0.0f;
            g2[i] = 0.0f;
            pk[i] = 0u;
        }
    }
    template <int I>
    __device__ __forceinline__ void get(float& a, float& b) const {
        a = g1[I];
        b = g2[I];
    }
    // keeps the magnitudes; a and b come back as the outputs use them
    template <int I>
    __device__ __forceinline__ void put(float& a, float& b) {
        g1[I] = a;
        g2[I] = b;
    }
};

template <int ROWS>
struct Messages<ROWS, true> {
    uint32_t gh[ROWS];  // half g(m1) | half g(m2) << 16
    uint32_t pk[ROWS];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
            gh[i] = 0u;
            pk[i] = 0u;
        }
    }
    template <int I>
    __device__ __forceinline__ void get(float& a, float& b) const {
        a = h2f(gh[I] & 0xffffu);
        b = h2f(gh[I] >> 16);
    }
    template <int I>
    __device__ __forceinline__ void put(float& a, float& b) {
        const uint32_t ha = f2h(fminf(a, 65504.0f)), hb = f2h(fminf(b, 65504.0f));
        gh[I] = ha | (hb << 16);
        a = h2f(ha);
        b = h2f(hb);
    }
};

// ldpc.rs:522-531 for the row of lane r in layer I of the codeword whose posteriors are
// `post`.
//
// Exactness of the shortcuts, given that no posterior is ever -0.0 (the LLRs are summed
// onto +0.0 when they are placed, FP16 stores a zero as +0.0, and a difference or sum of
// two values that are not -0.0 is not -0.0 except for -0.0 - +0.0):
//  - the sign bit of a message is the reference's test `l < 0.0`, so the output sign of
//    edge q, the product of the other input signs, is bit 31 of (input_q XOR all inputs);
//  - the edge that takes the second minimum is told by |input_q| == m1: it is the first
//    minimum, or a tie with it, and a tie makes m2 = m1, so either way the magnitude is the
//    reference's g(q == i1 ? m2 : m1);
//  - m2 = med3(m1, m2, |input|) is the strict-< update of the scan of ldpc.rs:471-480.
// The stored state keeps the output sign bits (edge q at bit deg-1-q, shifted in with a
// funnel shift) and one index whose magnitude is g(m2).
template <int BG, bool HALF, bool SOLO, int ZC, int I>
__device__ __forceinline__ void layer(typename Store<HALF>::T* __restrict__ post, int r, uint32_t gofs, int z,
                                      const uint32_t* __restrict__ shift, Messages<Graph<BG>::rows, HALF>& msg,
                                      float a, float b) {
    using S = Store<HALF>;
    constexpr int e0 = row_begin<BG>(I);
    constexpr int deg = row_begin<BG>(I + 1) - e0;
    // Z and the codeword's first posterior, opaque to the compiler: it would otherwise keep
    // one base address per column (68 of them) in registers through the whole iteration.
    // With Z known at compile time (ZC) the column bases are immediate offsets.
    uint32_t base = 0;
    if constexpr (ZC != 0) {
        // the shifts are constants too, so the row index is what keeps the addresses (all
        // 316 of them loop-invariant) from being computed once and held in registers
        z = ZC;
        asm volatile("" : "+v"(r));
    } else {
        asm volatile("" : "+s"(z));
    }
    if constexpr (!SOLO) {
        base = gofs;
        asm volatile("" : "+v"(base));
    }
    float g1o, g2o;
    msg.template get<I>(g1o, g2o);
    const uint32_t pko = msg.pk[I];
    const uint32_t i1o = pko >> 24;
    constexpr uint32_t es = sizeof(typename S::T);
    // (r + shift) mod z for edge q: the unsigned min of t and t - z (which wraps when t < z);
    // with ZC, the LDS byte address (the posteriors start at 0) so that the column base is
    // the LDS instruction's immediate offset, and a zero shift (every extension column)
    // needs no wrap; otherwise the element index
    auto address = [&](int q) -> int {
        if constexpr (ZC != 0) {
            const uint32_t s_q = shift_of<BG, ZC>(e0 + q) * es;
            const uint32_t t = uint32_t(r) * es + s_q;
            const uint32_t rr = s_q == 0 ? t : min(t, t - uint32_t(ZC) * es);
            return int(rr) + int(Graph<BG>::t[e0 + q].j) * ZC * int(es);
        } else {
            const uint32_t t = uint32_t(r) + shift[e0 + q];
            const uint32_t rr = min(t, t - uint32_t(z));
            return int(Graph<BG>::t[e0 + q].j) * z + int(rr) + int(base);
        }
    };
    // FP16 kernels run three waves a SIMD (168 registers): their widest layers compute the
    // addresses twice rather than hold them
    constexpr bool keep = !HALF || deg <= 10;
    float v2c[deg];
    int addr[keep ? deg : 1];
    float m1 = INFINITY, m2 = INFINITY;
    uint32_t parity = 0;
#pragma unroll
    for (int q = 0; q < deg; ++q) {
        const int ad = address(q);
        if constexpr (keep) addr[q] = ad;
        const float mag = uint32_t(q) == i1o ? g2o : g1o;
        const float old = __uint_as_float(__float_as_uint(mag) ^ ((pko << (32 - deg + q)) & 0x80000000u));
        const float in = (ZC != 0 ? S::ld_at(uint32_t(ad)) : S::ld(post, ad)) - old;
        v2c[q] = in;
        parity ^= __float_as_uint(in);
        const float m = fabsf(in);
        m2 = __builtin_amdgcn_fmed3f(m1, m2, m);
        m1 = fminf(m1, m);
    }
    float g1 = a * fmaxf(m1 - b, 0.0f);
    float g2 = a * fmaxf(m2 - b, 0.0f);
    msg.template put<I>(g1, g2);
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
        const int ad = keep ? addr[keep ? q : 0] : address(q);
        if constexpr (ZC != 0) S::st_at(uint32_t(ad), v2c[q] + out);
        else S::st(post, ad, v2c[q] + out);
    }
    msg.pk[I] = signs | (i1 << 24);
}

template <int BG, bool HALF, bool SOLO, int ZC, int I>
__device__ __forceinline__ void sweep(typename Store<HALF>::T* post, int r, uint32_t gofs, int z, bool active,
                                      const uint32_t* shift, Messages<Graph<BG>::rows, HALF>& msg, const float* ab,
                                      float a, float b) {
    if constexpr (I < Graph<BG>::rows) {
        if (ab) {
            a = ab[2 * I];
            b = ab[2 * I + 1];
        }
        if constexpr (barrier_before<BG>(I)) __syncthreads();
        // (the branch also bounds the scheduler's reach: without it the compiler interleaves
        // layers and spills hundreds of registers)
        if (active) layer<BG, HALF, SOLO, ZC, I>(post, r, gofs, z, shift, msg, a, b);
        sweep<BG, HALF, SOLO, ZC, I + 1>(post, r, gofs, z, active, shift, msg, ab, a, b);
    }
}

// The tables of the task's class in LDS, and the workgroup's scratch.
struct Shared {
    uint32_t shift[NR_BG1_ENTRIES];
    uint16_t col[NR_BG1_ENTRIES];
    uint16_t row_start[48];
    uint32_t synd[NR_BG1_ENTRIES];  // Z % 32 == 0: per edge, (word of its column, bit) in the packed decisions
    uint32_t fails[kMaxGroup];  // per codeword of the task: the syndrome stamp of a failing check
    uint32_t task;
    uint32_t vote[2];           // "some codeword still runs", alternate rounds
};

// Whether any lane of the workgroup has `pred` (one barrier). Two flags in turn: past the
// barrier every lane has finished the round before, so the other flag can be cleared for
// the round after, which the caller reaches only through further barriers.
__device__ __forceinline__ bool any_lane(Shared& sh, bool pred, uint32_t& round) {
    const uint32_t slot = round & 1u;
    ++round;
    if (pred) sh.vote[slot] = 1u;
    __syncthreads();
    if (threadIdx.x == 0) sh.vote[slot ^ 1u] = 0u;
    return sh.vote[slot] != 0u;
}
static_assert(sizeof(Shared) <= kSharedBytes, "the tables outgrow their LDS slot");

// The signs of 32 consecutive posteriors (16-byte aligned) as one word, bit k = posterior k:
// each funnel shift moves the word up one bit and brings in a sign bit at bit 0.
template <bool HALF>
__device__ __forceinline__ uint32_t pack_signs(const typename Store<HALF>::T* p) {
    const uint4* v = reinterpret_cast<const uint4*>(p);
    uint32_t acc = 0;
    if constexpr (!HALF) {
#pragma unroll
        for (int k = 7; k >= 0; --k) {
            const uint4 x = v[k];
            acc = __builtin_amdgcn_alignbit(acc, x.w, 31u);
            acc = __builtin_amdgcn_alignbit(acc, x.z, 31u);
            acc = __builtin_amdgcn_alignbit(acc, x.y, 31u);
            acc = __builtin_amdgcn_alignbit(acc, x.x, 31u);
        }
    } else {
        // two halves a word: the high one's sign is bit 31, the low one's bit 15
#pragma unroll
        for (int k = 3; k >= 0; --k) {
            const uint4 x = v[k];
            const uint32_t d[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int i = 3; i >= 0; --i) {
                acc = __builtin_amdgcn_alignbit(acc, d[i], 31u);
                acc = __builtin_amdgcn_alignbit(acc, d[i] << 16, 31u);
            }
        }
    }
    return acc;
}

// Whether the syndrome of the lane's codeword fails (ldpc.rs:555-557). Every codeword of
// the task is checked; a failing codeword has fails[g] == stamp afterwards.
template <int BG, bool HALF>
__device__ __forceinline__ bool syndrome_fails(const typename Store<HALF>::T* post_all, uint32_t* hd_all, Shared& sh,
                               uint32_t stamp, int z, int n, int cnt, int g, int r, bool lane_ok) {
    using S = Store<HALF>;
    if ((z & 31) == 0) {
        // packed: every lane packs the signs of 32 consecutive posteriors (one word of a
        // column) with funnel shifts, bit k = posterior k; each column's words doubled
        const int w32 = z >> 5, words = n >> 5, per_cw = Graph<BG>::cols * 2 * w32;
        for (int f = threadIdx.x; f < cnt * words; f += blockDim.x) {
            const int gg = f / words, fw = f - gg * words;
            const int j = fw / w32, w = fw - j * w32;
            const uint32_t word = pack_signs<HALF>(post_all + gg * n + fw * 32);
            uint32_t* hd = hd_all + gg * per_cw + j * 2 * w32;
            hd[w] = word;
            hd[w32 + w] = word;
        }
        __syncthreads();
        const int per_task = Graph<BG>::rows * w32;
        for (int task = threadIdx.x; task < cnt * per_task; task += blockDim.x) {
            const int gg = task / per_task, rem = task - gg * per_task;
            const int i = rem / w32, w = rem - i * w32;
            const uint32_t* hd = hd_all + gg * per_cw + w;
            uint32_t par = 0;
            for (int eb = sh.row_start[i]; eb < sh.row_start[i + 1]; ++eb) {
                const uint32_t x = sh.synd[eb];
                const uint32_t* q = hd + (x >> 5);
                par ^= __builtin_amdgcn_alignbit(q[1], q[0], x & 31u);
            }
            if (par) sh.fails[gg] = stamp;
        }
    } else if (lane_ok) {
        // every lane checks its rows edge by edge
        const typename S::T* post = post_all + g * n;
        uint32_t bad = 0;
        for (int i = 0; i < Graph<BG>::rows && !bad; ++i) {
            uint32_t par = 0;
            for (int eb = sh.row_start[i]; eb < sh.row_start[i + 1]; ++eb) {
                const uint32_t t = uint32_t(r) + sh.shift[eb];
                par ^= S::sign(post, sh.col[eb] * z + int(min(t, t - uint32_t(z))));
            }
            bad |= par;
        }
        if (bad) sh.fails[g] = stamp;
    }
    __syncthreads();
    return lane_ok && sh.fails[g] == stamp;
}

// full_llr (nr/ldpc.rs:191-197) of the lane's codeword: repeats summed in output order,
// 0.0 on unsent bits, 30.0 on fillers; lane r writes positions r, r + z, ...
template <bool HALF>
__device__ __forceinline__ void place_llrs(typename Store<HALF>::T* post, const float* llr, const BatchClass& c,
                                           int r) {
    for (int p = r; p < c.n; p += c.z) {
        float v = 0.0f;
        for (uint32_t s = c.src_start[p]; s < c.src_start[p + 1]; ++s) v += llr[c.src[s]];
        if (p >= c.k_info && p < c.k) v = 30.0f;
        Store<HALF>::st(post, p, v);
    }
}

// One task: codewords order[first .. first + cnt) of class c (base graph BG). SOLO: the
// task has one codeword (G = 1), whose posteriors start the LDS; ZC: Z, when every class of
// the launch has the same (0: read at run time).
template <int BG, bool HALF, bool SOLO, int ZC>
__device__ __forceinline__ void run_task(const LayeredArgs& a, const BatchClass& c, uint32_t first, int cnt, Shared& sh,
                         void* lds_dyn, uint32_t& stamp, uint32_t& round) {
    using S = Store<HALF>;
    using T = typename S::T;
    const int z = ZC ? ZC : c.z, n = ZC ? Graph<BG>::cols * ZC : c.n;
    T* post_all = static_cast<T*>(lds_dyn);
    uint32_t* hd_all = reinterpret_cast<uint32_t*>(post_all + c.g * n);  // g n is even: 4-byte aligned
    const int tid = threadIdx.x;
    const int g = SOLO ? 0 : tid / z, r = tid - g * z;
    // lanes beyond the task's codewords (SOLO: beyond Z) idle
    const bool lane_ok = SOLO ? tid < z : g < cnt;
    const uint32_t id = lane_ok ? (a.order ? a.order[first + g] : first + uint32_t(g)) : 0u;
    T* post = post_all + (lane_ok ? g : 0) * n;
    const float* llr = a.llr + (a.llr_off ? a.llr_off[id] : uint64_t(id) * uint64_t(c.e));
    const int limit = a.max_it ? int(a.max_it[id]) : c.max_iterations;

    if (lane_ok) place_llrs<HALF>(post, llr, c, r);
    __syncthreads();

    Messages<Graph<BG>::rows, HALF> msg;
    msg.zero();
    bool finished = !lane_ok, restarted = false;
    uint32_t its = 0, ok = 0;
    ++stamp;
    if (!syndrome_fails<BG, HALF>(post_all, hd_all, sh, stamp, z, n, cnt, g, r, lane_ok) && lane_ok) {
        finished = true;  // the channel LLRs are a codeword: 0 iterations
        ok = 1;
    }
    bool learned = a.params != nullptr;
    int it = 0, offset = 0;
    for (;;) {
        if (learned && it >= a.learned_iterations) {
            // the end of the learned iterations, the same for every codeword of the task
            if (!finished) {
                if (a.fallback) {
                    restarted = true;
                    place_llrs<HALF>(post, llr, c, r);
                    msg.zero();
                } else {
                    finished = true;
                    its = uint32_t(it);
                }
            }
            learned = false;
            offset = it;
            it = 0;
        }
        if (!learned && !finished && it >= limit) {
            finished = true;
            its = uint32_t(offset + it);
        }
        if (!any_lane(sh, !finished, round)) break;
        ++it;
        const float* ab = learned ? a.params + size_t(it - 1) * 2 * Graph<BG>::rows : nullptr;
        sweep<BG, HALF, SOLO, ZC, 0>(post_all, r, uint32_t(g * n), z, !finished, sh.shift, msg, ab, c.a, c.b);
        __syncthreads();
        ++stamp;
        const bool fails = syndrome_fails<BG, HALF>(post_all, hd_all, sh, stamp, z, n, cnt, g, r, lane_ok);
        if (!finished && !fails) {
            finished = true;
            its = uint32_t(offset + it);
            if (!restarted) ok = 1;
        }
    }

    if (lane_ok) {
        if (r == 0) {
            if (a.iterations) a.iterations[id] = its;
            if (a.ok) a.ok[id] = uint8_t(ok);
        }
        if (a.bits) {
            uint8_t* bits = a.bits + (a.bit_off ? a.bit_off[id] : uint64_t(id) * uint64_t(c.k_info));
            for (int p = r; p < c.k_info; p += z) bits[p] = uint8_t(S::sign(post, p));
        }
        if (a.post_out)
            for (int p = r; p < n; p += z) a.post_out[uint64_t(id) * uint64_t(n) + p] = S::ld(post, p);
    }
    __syncthreads();  // the next task overwrites the posteriors
}

// Values every lane holds alike, made scalar: loads from global memory inside the task
// loop come back in vector registers (the loop also stores to global memory), and a
// class's fields in vector registers would cost a register per column base in the sweep.
__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
template <typename T>
__device__ __forceinline__ const T* uni(const T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v)), hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<const T*>((uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ BatchClass uni(const BatchClass& c) {
    BatchClass u;
    u.bg = uni(c.bg);
    u.z = uni(c.z);
    u.n = uni(c.n);
    u.k = uni(c.k);
    u.k_info = uni(c.k_info);
    u.e = uni(c.e);
    u.g = uni(c.g);
    u.max_iterations = uni(c.max_iterations);
    u.a = uni(c.a);
    u.b = uni(c.b);
    u.shift32 = uni(c.shift32);
    u.col = uni(c.col);
    u.row_start = uni(c.row_start);
    u.src_start = uni(c.src_start);
    u.src = uni(c.src);
    return u;
}

// Workgroups take tasks until none is left. BGMASK: the base graphs of the classes; SOLO:
// every class has G = 1; ZC: every class has Z = ZC (0: any).
template <int BGMASK, bool HALF, bool SOLO, int ZC>
__global__ void __launch_bounds__(kMaxThreads, HALF ? 3 : 1) decode_layered_kernel(LayeredArgs a) {
    // one dynamic allocation and no static LDS, so that the posteriors start at address 0
    // and the column bases are immediate offsets of the LDS instructions
    extern __shared__ __align__(16) uint32_t lds_dyn[];
    Shared& sh = *reinterpret_cast<Shared*>(reinterpret_cast<char*>(lds_dyn) + a.shared_off);
    for (int i = threadIdx.x; i < kMaxGroup; i += blockDim.x) sh.fails[i] = 0u;
    if (threadIdx.x < 2) sh.vote[threadIdx.x] = 0u;
    uint32_t stamp = 0, round = 0;
    int loaded = -1;
    uint32_t t = blockIdx.x;
    for (;;) {
        if (a.queue) {
            if (threadIdx.x == 0) sh.task = atomicAdd(a.queue, 1u);
            __syncthreads();
            // the same value in every lane: uniform, so that the task's fields are scalars
            t = __builtin_amdgcn_readfirstlane(sh.task);
            __syncthreads();  // every lane has read it before the next task's update
        }
        if (t >= uint32_t(a.n_tasks)) break;
        BatchClass c;
        uint32_t first;
        int cnt, cls;
        if (a.tasks) {
            const BatchTask tk = a.tasks[t];
            cls = uni(int32_t(tk.cls));
            c = uni(a.classes[cls]);
            first = uint32_t(uni(int32_t(tk.first)));
            cnt = uni(int32_t(tk.count));
        } else {
            cls = 0;
            c = a.one;
            first = t * uint32_t(c.g);
            cnt = int(min(int64_t(c.g), a.count - int64_t(first)));
        }
        if (cls != loaded) {
            // the shifts (read with uniform addresses), columns and row starts in LDS
            const int edges = c.bg == 1 ? Graph<1>::edges : Graph<2>::edges;
            const int rows = c.bg == 1 ? Graph<1>::rows : Graph<2>::rows;
            const int w32 = c.z >> 5;
            for (int e = threadIdx.x; e < edges; e += blockDim.x) {
                const uint32_t v = c.shift32[e], j = c.col[e];
                sh.shift[e] = v;
                sh.col[e] = uint16_t(j);
                // the funnel shift of the packed syndrome: word j 2 w32 + v / 32, bit v % 32
                sh.synd[e] = ((j * 2u * uint32_t(w32) + (v >> 5)) << 5) | (v & 31u);
            }
            for (int i = threadIdx.x; i <= rows; i += blockDim.x) sh.row_start[i] = c.row_start[i];
            __syncthreads();
            loaded = cls;
        }
        if constexpr ((BGMASK & 1) != 0) {
            if (c.bg == 1) run_task<1, HALF, SOLO, ZC>(a, c, first, cnt, sh, lds_dyn, stamp, round);
        }
        if constexpr ((BGMASK & 2) != 0) {
            if (c.bg == 2) run_task<2, HALF, SOLO, ZC>(a, c, first, cnt, sh, lds_dyn, stamp, round);
        }
        if (!a.queue) t += gridDim.x;
    }
}

}  // namespace layered
}  // namespace comfec
