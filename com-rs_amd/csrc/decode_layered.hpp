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

// two 16-bit lanes of one register (packed integer arithmetic)
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));

// (rb + s) mod Zb for byte offsets rb, s < Zb <= 1536: rbp holds rb and rb - Zb (mod 2^16)
// in its halves and s2 holds s in both, so one 32-bit add forms t = rb + s and t - Zb at
// once (t < 3072: no carry from the low half; the high half's carry leaves the word), and
// their unsigned 16-bit min is the one in [0, Zb)
__device__ __forceinline__ uint32_t rotate_bytes(uint32_t rbp, uint32_t s2) {
    const u16x2 t = __builtin_bit_cast(u16x2, rbp + s2);
    return min(uint32_t(t.x), uint32_t(t.y));
}

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

// Whether layer l must wait for the layers before it: it shares a column with a layer
// processed since the last barrier (reads after writes, and writes after reads, which are
// the same shared columns). Layers on disjoint columns commute, so they run without one.
template <int BG>
struct Schedule {
    bool barrier[Graph<BG>::rows];
};

template <int BG>
constexpr Schedule<BG> make_schedule() {
    constexpr int R = Graph<BG>::rows, C = Graph<BG>::cols, E = Graph<BG>::edges;
    Schedule<BG> s{};
    bool has[R][C] = {};
    for (int e = 0; e < E; ++e) has[Graph<BG>::t[e].i][Graph<BG>::t[e].j] = true;
    int first = 0;  // first layer after the last barrier
    for (int l = 0; l < R; ++l) {
        bool conflict = false;
        for (int p = first; p < l; ++p)
            for (int c = 0; c < C; ++c) conflict = conflict || (has[l][c] && has[p][c]);
        s.barrier[l] = conflict;
        if (conflict) first = l;
    }
    return s;
}

template <int BG>
struct Sched {
    static constexpr Schedule<BG> s = make_schedule<BG>();
};

// Two sets of six waves on one codeword (the DUAL instantiations, 768 lanes): a CU places
// wave i and waves i + 4, i + 8 on one SIMD, so with set s = waves 6s .. 6s + 5 every SIMD
// holds waves of both sets. The layers between two barriers share no column, so the layers
// of such a group are spread over the sets (the widest first, each to the set with fewer
// edges of the group so far, ties to the set with fewer layers) and run side by side. A
// group of one wide layer (kSplitDegree edges or more) is split: each set updates part of
// every row's edges, and the two exchange their parts' two smallest magnitudes and sign
// products through LDS; a group of one narrower layer runs on one set while the other waits
// at the next barrier. A lane keeps the messages of its set's layers only: slot[s][l] is
// layer l's place among set s's (-1: not its layer), and slots the larger count.
constexpr int kSplitDegree = 12;

//
// Lookahead: while a group of one layer (owner set A) runs, the other set would only wait at
// the next barrier. If the next group holds a layer J of that set, the set reads J's inputs
// on the columns the running layer does not write (premask[J]: its edges there, most of J
// for BG1's extension rows, which share one or two columns with the layer before), forms
// their variable-to-check messages, sign product and two smallest magnitudes, and keeps them
// in registers over the barrier; J then reads only the rest. Every posterior is still read
// after the last write before it in the reference's layer order, and the two minima and the
// sign product do not depend on the order of their inputs, so the results are unchanged.
template <int BG>
struct DualPlan {
    int owner[Graph<BG>::rows];  // 0, 1, or 2 for a layer split over both sets
    int slot[2][Graph<BG>::rows];
    int slots;
    int pre_of[Graph<BG>::rows];        // the layer the idle set reads ahead during this one's group, or -1
    uint32_t premask[Graph<BG>::rows];  // of a layer read ahead: its edges read before its group's barrier
};

template <int BG>
constexpr int degree_of(int i) {
    return row_begin<BG>(i + 1) - row_begin<BG>(i);
}

template <int BG>
constexpr DualPlan<BG> make_dual() {
    constexpr int R = Graph<BG>::rows;
    DualPlan<BG> d{};
    int count[2] = {0, 0};
    for (int l = 0; l < R;) {
        int end = l + 1;
        while (end < R && !Sched<BG>::s.barrier[end]) ++end;
        if (end == l + 1 && degree_of<BG>(l) >= kSplitDegree) {
            d.owner[l] = 2;
            ++count[0];
            ++count[1];
            l = end;
            continue;
        }
        int load[2] = {0, 0};
        bool placed[R] = {};
        for (int n = l; n < end; ++n) {
            int best = -1;
            for (int k = l; k < end; ++k)
                if (!placed[k] && (best < 0 || degree_of<BG>(k) > degree_of<BG>(best))) best = k;
            const int set = load[0] != load[1] ? (load[0] < load[1] ? 0 : 1) : (count[0] <= count[1] ? 0 : 1);
            placed[best] = true;
            d.owner[best] = set;
            load[set] += degree_of<BG>(best);
            ++count[set];
        }
        l = end;
    }
    int next[2] = {0, 0};
    for (int i = 0; i < R; ++i)
        for (int s = 0; s < 2; ++s) d.slot[s][i] = d.owner[i] == s || d.owner[i] == 2 ? next[s]++ : -1;
    d.slots = count[0] > count[1] ? count[0] : count[1];
    // lookahead over groups of one layer
    for (int i = 0; i < R; ++i) {
        d.pre_of[i] = -1;
        d.premask[i] = 0u;
    }
    for (int l = 0; l < R; ++l) {
        const bool alone = (l + 1 == R || Sched<BG>::s.barrier[l + 1]) && (l == 0 || Sched<BG>::s.barrier[l]);
        if (!alone || d.owner[l] == 2 || l + 1 == R) continue;
        int end = l + 2;  // the next group: l + 1 .. end - 1
        while (end < R && !Sched<BG>::s.barrier[end]) ++end;
        for (int j = l + 1; j < end; ++j) {
            if (d.owner[j] != 1 - d.owner[l]) continue;
            uint32_t mask = 0u;
            for (int q = 0; q < degree_of<BG>(j); ++q) {
                bool shared = false;
                for (int e = row_begin<BG>(l); e < row_begin<BG>(l + 1); ++e)
                    shared = shared || Graph<BG>::t[e].j == Graph<BG>::t[row_begin<BG>(j) + q].j;
                if (!shared) mask |= 1u << q;
            }
            if (mask != 0u) {
                d.pre_of[l] = j;
                d.premask[j] = mask;
            }
            break;
        }
    }
    return d;
}

template <int BG>
struct Dual {
    static constexpr DualPlan<BG> p = make_dual<BG>();
};

// the message slots a lane holds
template <int BG, bool DUAL>
constexpr int msg_slots() {
    if constexpr (DUAL) return Dual<BG>::p.slots;
    else return Graph<BG>::rows;
}
template <int BG, bool DUAL>
constexpr int slot_of(int i) {
    if constexpr (DUAL) return Dual<BG>::p.slot[Dual<BG>::p.owner[i] == 1 ? 1 : 0][i];
    else return i;
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

// floats in global memory (address space 1)
typedef const float __attribute__((address_space(1))) gfloat;
__device__ __forceinline__ const gfloat* to_global(const float* p) {
    return reinterpret_cast<const gfloat*>(reinterpret_cast<uintptr_t>(p));
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
    static __device__ __forceinline__ uint32_t sign_at(uint32_t off) {
        return *(const lds_ptr<uint32_t>::type)(uintptr_t)off >> 31;
    }
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
    static __device__ __forceinline__ uint32_t sign_at(uint32_t off) {
        return uint32_t(*(const lds_ptr<uint16_t>::type)(uintptr_t)off) >> 15;
    }
};

// rbp of row r for lifting size z: r and r - z in bytes of the storage, in 16-bit halves
template <bool HALF>
__device__ __forceinline__ uint32_t row_pair(int r, int z) {
    constexpr uint32_t es = sizeof(typename Store<HALF>::T);
    const uint32_t rb = uint32_t(r) * es;
    return ((rb - uint32_t(z) * es) << 16) | rb;
}

// The messages of one lane's rows: output magnitudes g(m1), g(m2) and, packed, the
// index of the edge that takes g(m2) (bits 0..4) and the output sign bits (edge q of a
// row at bit 31 - q, so the first edge's is the sign bit itself and each doubling of the word
// brings the next one there). The index is compared with inline constants where messages
// are rebuilt, so no register holds an edge number across the sweep.
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
    // The widened magnitudes are made opaque: the compiler would otherwise move the
    // conversion past each edge's choice between them (one conversion per edge, not two
    // per row).
    template <int I>
    __device__ __forceinline__ void get(float& a, float& b) const {
        a = h2f(gh[I] & 0xffffu);
        b = h2f(gh[I] >> 16);
        asm volatile("" : "+v"(a), "+v"(b));
    }
    template <int I>
    __device__ __forceinline__ void put(float& a, float& b) {
        const uint32_t ha = f2h(fminf(a, 65504.0f)), hb = f2h(fminf(b, 65504.0f));
        gh[I] = ha | (hb << 16);
        a = h2f(ha);
        b = h2f(hb);
        asm volatile("" : "+v"(a), "+v"(b));
    }
};

// The inputs of a layer read ahead (two-set lookahead): per edge its variable-to-check
// message and posterior address, and the partial sign product and two smallest magnitudes
// over those edges.
constexpr int kAheadMax = 16;
struct Ahead {
    float v[kAheadMax];
    uint32_t ad[kAheadMax];
    float m1, m2;
    uint32_t par;
};

// Values the sweep of the two-set kernel keeps in vector registers, made once per task. At
// three waves a SIMD a VALU instruction with an operand in a scalar register or an inline
// constant often issues more slowly than one whose operands are all vector registers, so
// the constants the layers use on every row or edge are held rather than rebuilt or taken
// as scalars (each change timed faster on one device against the kernel without it). The
// other kernels take them as plain values, which the compiler folds into constants.
struct LaneRegs {
    uint32_t sgn;  // 0x80000000
    float zero;    // +0.0
    float a, b;    // the class's g(m) = a max(m - b, 0)
    uint32_t rbh;  // the lane's row offset + 64 KiB: a base for columns past the LDS offset field
    uint32_t one;  // 1: the edge counter's step
};

// Gives the read-ahead values of a layer (edges PRE) an undefined value with no instruction:
// an empty asm with them as outputs. On the path of the lanes that do not read ahead, the
// compiler would otherwise set each of them to zero, one move per register.
template <uint32_t PRE>
__device__ __forceinline__ void ahead_undefined(Ahead& ah) {
#pragma unroll
    for (int q = 0; q < kAheadMax; ++q)
        if ((PRE >> q) & 1u) asm volatile("" : "=v"(ah.v[q]), "=v"(ah.ad[q]));
    asm volatile("" : "=v"(ah.m1), "=v"(ah.m2), "=v"(ah.par));
}

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
//  - m1 and m2 are the smallest and the second smallest input magnitude counted with
//    multiplicity, which is what the strict-< scan of ldpc.rs:471-480 finds (a tie gives
//    m2 = m1); they are selected, never computed, so the order of the comparisons does not
//    matter: three inputs at a time give their min3 and med3, and a pair (t1 <= t2) merges
//    into (m1 <= m2) as min(m1, t1) and med3(m1, t1, min(m2, t2)).
// The stored state keeps the output sign bits (edge q at bit 31 - q, shifted in with a
// funnel shift) and one index whose magnitude is g(m2); a message is rebuilt as that
// magnitude with its sign bit replaced by the stored one, which is the output bit for bit.
//
// The lane updates edges Q0 .. Q1-1 of its row (all of them, or with SPLIT its set's part of
// a row split over both sets) and keeps their messages in slot SL. SPLIT: after its part's
// inputs, the lane writes their two smallest magnitudes and sign product to LDS at xbase,
// every lane of the workgroup (active or not) meets at one barrier, and the lane merges the
// other set's part into its own (the minima as three inputs are merged above, the sign
// products by XOR) before the outputs; its index of g(m2) is that of a part's edge whose
// |input| is the merged m1, or 31 (none) if the other part holds it.
//
// PRE (two-set lookahead): the edges whose inputs were read ahead into `ahead` (their
// messages, addresses, and partial sign product and minima are taken from there). AHEAD:
// read the inputs of the PRE edges into `ahead` and return (no outputs).
template <int BG, bool HALF, bool SOLO, int ZC, bool NMS, bool DUAL, int I, int SL, int Q0, int Q1, bool SPLIT,
          uint32_t PRE = 0u, bool AHEAD = false>
__device__ __forceinline__ void layer(uint32_t gofs, uint32_t& rbp, uint32_t& rb, const uint2* __restrict__ edge,
                                      Messages<msg_slots<BG, DUAL>(), HALF>& msg, float a, float b, bool active,
                                      uint32_t xbase, int set, LaneRegs& lr, Ahead& ahead) {
    using S = Store<HALF>;
    constexpr int e0 = row_begin<BG>(I) + Q0;
    constexpr int deg = Q1 - Q0;  // the edges the lane updates
    static_assert(ZC == 0 || SOLO, "a compile-time Z serves tasks of one codeword");
    static_assert(!SPLIT || (DUAL && !HALF), "split rows are a two-set FP32 layout");
    // Addresses are LDS byte addresses (the posteriors start at address 0). rbp: the row's
    // byte offset in a column and that minus Z, in 16-bit halves; rb: its low half; gofs: the
    // byte offset of the lane's codeword. With Z known at compile time (ZC) the shifts and
    // the column bases are constants too, so rbp and rb are made opaque, or the compiler
    // would compute all 316 loop-invariant addresses once and hold them in registers (they
    // pass from layer to layer by reference, so no copy is made); the column bases are then
    // the LDS instructions' immediate offsets.
    uint32_t base = 0;
    if constexpr (ZC != 0) asm volatile("" : "+v"(rbp), "+v"(rb));
    if constexpr (!SOLO) {
        base = gofs;
        asm volatile("" : "+v"(base));
    }
    float g1o, g2o;
    msg.template get<SL>(g1o, g2o);
    const uint32_t pko = msg.pk[SL];
    // extracted once a row (opaque, or each edge's compare becomes a byte-select form that
    // issues more slowly than the mask and plain compares together)
    uint32_t i1o = pko & 31u;
    asm volatile("" : "+v"(i1o));
    constexpr uint32_t es = sizeof(typename S::T);
    // The posterior of entry e on the lane's row: (r + shift) mod Z in column j. With ZC a
    // zero shift (every extension column) needs no wrap. With Z known only at run time the
    // layer's entries come from the LDS table, all of them before any posterior (the empty
    // asm keeps the compiler from interleaving the two kinds of reads, which would cost an
    // LDS round trip per pair of edges).
    uint2 ed[ZC != 0 ? 1 : deg];
    if constexpr (ZC == 0) {
#pragma unroll
        for (int q = 0; q < deg; ++q) ed[q] = edge[e0 + q];
        asm volatile("" ::: "memory");
    }
    // the column's first posterior: with ZC a constant, which becomes the LDS instruction's
    // immediate offset; the rest of the address is the rotated row
    // (a column whose base exceeds the 16-bit offset field, with no shift, takes its offset
    // from the lane's second base, 64 KiB above its row)
    // made opaque in place, as rb is
    if constexpr (ZC != 0) asm volatile("" : "+v"(lr.rbh));
    const uint32_t rbh = lr.rbh;
    auto high = [&](int q) -> bool {
        if constexpr (ZC != 0)
            return shift_of<BG, ZC>(e0 + q) == 0u && uint32_t(Graph<BG>::t[e0 + q].j) * uint32_t(ZC) * es >= 65536u;
        else return false;
    };
    auto colbase = [&](int q) -> uint32_t {
        if constexpr (ZC != 0) return uint32_t(Graph<BG>::t[e0 + q].j) * uint32_t(ZC) * es - (high(q) ? 65536u : 0u);
        else return 0u;
    };
    auto rotated = [&](int q) -> uint32_t {
        if constexpr (ZC != 0) {
            const uint32_t s_q = shift_of<BG, ZC>(e0 + q) * es;
            return s_q == 0 ? (high(q) ? rbh : rb) : rotate_bytes(rbp, s_q | (s_q << 16));
        } else {
            return rotate_bytes(rbp, ed[ZC != 0 ? 0 : q].x) + ed[ZC != 0 ? 0 : q].y + base;
        }
    };
    auto address = [&](int q) -> uint32_t { return rotated(q) + colbase(q); };
    auto load = [&](uint32_t ad) -> float { return S::ld_at(ad); };
    // FP16 kernels run three waves a SIMD (168 registers): their widest layers compute the
    // addresses twice rather than hold them
    constexpr bool keep = !HALF || deg <= 10;
    float v2c[deg];
    uint32_t addr[keep ? deg : 1];
    float m1 = INFINITY, m2 = INFINITY;
    uint32_t parity = 0;
    const uint32_t sgn = lr.sgn;
    static_assert(PRE == 0u || (DUAL && !HALF && !SPLIT && deg <= kAheadMax), "lookahead serves whole FP32 rows");
    static_assert(!AHEAD || PRE != 0u, "a lookahead reads some edges");
    if (!SPLIT || active) {
        // edge q's stored sign at bit 31 of the word shifted left q places; along a run of
        // edges the shift is one doubling more per edge (a 32-bit add of vector registers,
        // which issues at twice the rate of a shift left by an immediate)
        uint32_t w = 0u;
        int wq = -2;  // the edge whose sign w holds at bit 31
        float raw[AHEAD ? deg : 1];
        if constexpr (AHEAD) {
#pragma unroll
            for (int q = 0; q < deg; ++q) {
                if (((PRE >> q) & 1u) == 0u) continue;
                ahead.ad[q] = rotated(q);
                raw[q] = load(ahead.ad[q] + colbase(q));
            }
            asm volatile("" ::: "memory");
        }
#pragma unroll
        for (int q = 0; q < deg; ++q) {
            // AHEAD: the PRE edges only; otherwise the PRE edges come from `ahead`
            if (((PRE >> q) & 1u) != (AHEAD ? 1u : 0u)) {
                if constexpr (!AHEAD) {
                    if constexpr (keep) addr[q] = ahead.ad[q] + colbase(q);
                    v2c[q] = ahead.v[q];
                }
                continue;
            }
            uint32_t ad = 0u;
            if constexpr (!AHEAD) {
                ad = address(q);
                if constexpr (keep) addr[q] = ad;
            }
            const float mag = uint32_t(q) == i1o ? g2o : g1o;
            if (wq == q - 1) {
                uint32_t w2;
                asm("v_add_u32 %0, %1, %1" : "=v"(w2) : "v"(w));
                w = w2;
            } else {
                w = pko << q;
            }
            wq = q;
            // bit 31 from the stored word, the rest from the magnitude: (c ? a : b) bitwise
            const float old = __uint_as_float(__builtin_amdgcn_bitop3_b32(w, __float_as_uint(mag), sgn, 0xe4));
            if constexpr (AHEAD) {
                v2c[q] = raw[q] - old;
                ahead.v[q] = v2c[q];
            } else {
                v2c[q] = load(ad) - old;
            }
        }
        if constexpr (PRE == 0u) {
            // the row's sign product: three-input XORs from the first input on
            parity = __float_as_uint(v2c[0]);
#pragma unroll
            for (int q = 1; q < deg; q += 2)
                parity = q + 1 < deg ? __builtin_amdgcn_bitop3_b32(parity, __float_as_uint(v2c[q]),
                                                                   __float_as_uint(v2c[q + 1]), 0x96)
                                     : parity ^ __float_as_uint(v2c[q]);
            // the two smallest magnitudes, three inputs at a time
            constexpr int triples = deg / 3;
#pragma unroll
            for (int t = 0; t < triples; ++t) {
                const float x = fabsf(v2c[3 * t]), y = fabsf(v2c[3 * t + 1]), w = fabsf(v2c[3 * t + 2]);
                const float t1 = fminf(fminf(x, y), w), t2 = __builtin_amdgcn_fmed3f(x, y, w);
                if (t == 0) {
                    m1 = t1;
                    m2 = t2;
                } else {
                    m2 = __builtin_amdgcn_fmed3f(m1, t1, fminf(m2, t2));
                    m1 = fminf(m1, t1);
                }
            }
#pragma unroll
            for (int q = 3 * triples; q < deg; ++q) {
                const float m = fabsf(v2c[q]);
                m2 = __builtin_amdgcn_fmed3f(m1, m2, m);
                m1 = fminf(m1, m);
            }
        } else {
            // the inputs of one part of the edges (AHEAD: the PRE edges; otherwise the rest,
            // merged into the part read ahead), the PRE edges' list compacted at compile time
            constexpr uint32_t part = AHEAD ? PRE : (((1u << deg) - 1u) & ~PRE);
            constexpr int np = __builtin_popcount(part);
            float x[np > 0 ? np : 1];
            int k = 0;
#pragma unroll
            for (int q = 0; q < deg; ++q)
                if ((part >> q) & 1u) x[k++] = v2c[q];
            if constexpr (AHEAD) {
                parity = 0u;
                m1 = INFINITY;
                m2 = INFINITY;
            } else {
                parity = ahead.par;
                m1 = ahead.m1;
                m2 = ahead.m2;
            }
#pragma unroll
            for (int q = 0; q + 1 < np; q += 2)
                parity = __builtin_amdgcn_bitop3_b32(parity, __float_as_uint(x[q]), __float_as_uint(x[q + 1]), 0x96);
            if constexpr ((np & 1) != 0) parity ^= __float_as_uint(x[np - 1]);
            constexpr int triples = np / 3;
#pragma unroll
            for (int t = 0; t < triples; ++t) {
                const float p = fabsf(x[3 * t]), y = fabsf(x[3 * t + 1]), z = fabsf(x[3 * t + 2]);
                const float t1 = fminf(fminf(p, y), z), t2 = __builtin_amdgcn_fmed3f(p, y, z);
                m2 = __builtin_amdgcn_fmed3f(m1, t1, fminf(m2, t2));
                m1 = fminf(m1, t1);
            }
#pragma unroll
            for (int q = 3 * triples; q < np; ++q) {
                const float m = fabsf(x[q]);
                m2 = __builtin_amdgcn_fmed3f(m1, m2, m);
                m1 = fminf(m1, m);
            }
            if constexpr (AHEAD) {
                ahead.par = parity;
                ahead.m1 = m1;
                ahead.m2 = m2;
            }
        }
    }
    else {
        // SPLIT, lanes of a finished codeword: no inputs (their values are never used; an
        // empty asm defines them with no instruction, where the compiler would set each
        // register to zero)
#pragma unroll
        for (int q = 0; q < deg; ++q) asm volatile("" : "=v"(v2c[q]));
        if constexpr (keep) {
#pragma unroll
            for (int q = 0; q < deg; ++q) asm volatile("" : "=v"(addr[q]));
        }
        asm volatile("" : "=v"(m1), "=v"(m2), "=v"(parity));
    }
    if constexpr (AHEAD) return;
    if constexpr (SPLIT) {
        // [m1 | m2 | sign product] x [set 0 | set 1] x Z rows of 4 bytes
        constexpr uint32_t part = uint32_t(kMaxThreads) * 4u;
        const uint32_t mine = xbase + uint32_t(set) * part + rb, other = xbase + uint32_t(1 - set) * part + rb;
        if (active) {
            *(lds_ptr<float>::type)(uintptr_t)mine = m1;
            *(lds_ptr<float>::type)(uintptr_t)(mine + 2 * part) = m2;
            *(lds_ptr<uint32_t>::type)(uintptr_t)(mine + 4 * part) = parity;
        }
        __syncthreads();
        if (!active) return;
        const float om1 = *(const lds_ptr<float>::type)(uintptr_t)other;
        const float om2 = *(const lds_ptr<float>::type)(uintptr_t)(other + 2 * part);
        const uint32_t opar = *(const lds_ptr<uint32_t>::type)(uintptr_t)(other + 4 * part);
        m2 = __builtin_amdgcn_fmed3f(m1, om1, fminf(m2, om2));
        m1 = fminf(m1, om1);
        parity ^= opar;
    }
    // NMS: b = 0, and a max(m - 0, 0) is a m bit for bit for every magnitude m >= +0
    float g1 = NMS ? a * m1 : a * fmaxf(m1 - b, lr.zero);
    float g2 = NMS ? a * m2 : a * fmaxf(m2 - b, lr.zero);
    msg.template put<SL>(g1, g2);
    // the magnitudes with the row's sign product applied once: an edge's output is then
    // its choice XOR its own input's sign bit, and the stored bit is the output's sign
    // (parity & sign mask) ^ magnitude, one three-input operation each on vector registers
    uint32_t g1s = __builtin_amdgcn_bitop3_b32(parity, __float_as_uint(g1), sgn, 0x6c),
             g2s = __builtin_amdgcn_bitop3_b32(parity, __float_as_uint(g2), sgn, 0x6c);
    // opaque, or the compiler folds the sign product back into every edge
    asm volatile("" : "+v"(g1s), "+v"(g2s));
    // the edge number in a vector register (from the zero's bits, one more an edge): the
    // choice of i1 is then a select of two vector registers
    uint32_t qv = __float_as_uint(lr.zero);
    uint32_t signs = 0, i1 = SPLIT ? 31u : qv;
#pragma unroll
    for (int q = 0; q < deg; ++q) {
        const bool first = fabsf(v2c[q]) == m1;
        const uint32_t sel = first ? g2s : g1s;
        if (q > 0) qv += lr.one;
        i1 = first ? qv : i1;
        // (a & c) ^ b bitwise
        const uint32_t ob = __builtin_amdgcn_bitop3_b32(__float_as_uint(v2c[q]), sel, sgn, 0x6c);
        signs = __builtin_amdgcn_alignbit(signs, ob, 31u);
        const float out = __uint_as_float(ob);
        const uint32_t ad = keep ? addr[keep ? q : 0] : address(q);
        S::st_at(ad, v2c[q] + out);
    }
    msg.pk[SL] = (signs << (32 - deg)) | i1;
}

// the end of the barrier group that starts at layer I (its layers share no column)
template <int BG>
constexpr int group_end(int i) {
    int e = i + 1;
    while (e < Graph<BG>::rows && !Sched<BG>::s.barrier[e]) ++e;
    return e;
}

// the layer of group [I, E) whose inputs were read ahead, or -1
template <int BG>
constexpr int ahead_in(int i, int e) {
    for (int l = i; l < e; ++l)
        if (Dual<BG>::p.premask[l] != 0u) return l;
    return -1;
}

// One layer's update by the lanes that own it (no barrier before it); (a, b): its g(m).
template <int BG, bool HALF, bool SOLO, int ZC, bool NMS, bool DUAL, int I>
__device__ __forceinline__ void one_layer(uint32_t gofs, uint32_t& rbp, uint32_t& rb, bool active, int set,
                                          const uint2* edge, Messages<msg_slots<BG, DUAL>(), HALF>& msg, float a,
                                          float b, uint32_t xbase, LaneRegs& lr, Ahead& ahead) {
    constexpr int d = degree_of<BG>(I);
    if constexpr (DUAL && Dual<BG>::p.owner[I] == 2) {
        // split over both sets (each path holds one barrier, which every wave meets)
        constexpr int h = (d + 1) / 2;
        if (set == 0)
            layer<BG, HALF, SOLO, ZC, NMS, DUAL, I, Dual<BG>::p.slot[0][I], 0, h, true>(gofs, rbp, rb, edge, msg, a, b,
                                                                                     active, xbase, set, lr, ahead);
        else
            layer<BG, HALF, SOLO, ZC, NMS, DUAL, I, Dual<BG>::p.slot[1][I], h, d, true>(gofs, rbp, rb, edge, msg, a, b,
                                                                                     active, xbase, set, lr, ahead);
    } else {
        // DUAL: set is uniform over a wavefront, so a wave of the other set skips the layer
        // (the branch also bounds the scheduler's reach: without it the compiler interleaves
        // layers and spills hundreds of registers)
        constexpr uint32_t pre = DUAL ? Dual<BG>::p.premask[I] : 0u;
        if ((!DUAL || set == Dual<BG>::p.owner[I]) && active)
            layer<BG, HALF, SOLO, ZC, NMS, DUAL, I, slot_of<BG, DUAL>(I), 0, d, false, pre>(
                gofs, rbp, rb, edge, msg, a, b, true, xbase, set, lr, ahead);
    }
}

// The layers I .. E-1 of the group that starts at G, except P, in order; la, lb: the (a, b)
// of the group's layers.
template <int BG, bool HALF, bool SOLO, int ZC, bool NMS, bool DUAL, int G, int I, int E, int P>
__device__ __forceinline__ void group_rest(uint32_t gofs, uint32_t& rbp, uint32_t& rb, bool active, int set,
                                           const uint2* edge, Messages<msg_slots<BG, DUAL>(), HALF>& msg,
                                           const float* la, const float* lb, uint32_t xbase, LaneRegs& lr,
                                           Ahead& ahead) {
    if constexpr (I < E) {
        if constexpr (I != P)
            one_layer<BG, HALF, SOLO, ZC, NMS, DUAL, I>(gofs, rbp, rb, active, set, edge, msg, la[I - G], lb[I - G],
                                                        xbase, lr, ahead);
        group_rest<BG, HALF, SOLO, ZC, NMS, DUAL, G, I + 1, E, P>(gofs, rbp, rb, active, set, edge, msg, la, lb, xbase,
                                                                   lr, ahead);
    }
}

// The barrier groups from the one that starts at layer I to the end of the sweep. A layer read
// ahead goes first in its group, so that no other layer's code lies between the read-ahead
// inputs and their use (they would hold their registers through it).
template <int BG, bool HALF, bool SOLO, int ZC, bool NMS, bool DUAL, int I>
__device__ __forceinline__ void sweep(uint32_t gofs, uint32_t& rbp, uint32_t& rb, bool active, int set,
                                      const uint2* edge, Messages<msg_slots<BG, DUAL>(), HALF>& msg, const gfloat* ab,
                                      int st, float a, float b, uint32_t xbase, LaneRegs& lr, Ahead& ahead) {
    if constexpr (I < Graph<BG>::rows) {
        constexpr int E = group_end<BG>(I);
        constexpr int P = DUAL ? ahead_in<BG>(I, E) : -1;
        // the group's (a, b), read before the barrier so that their latency passes while the
        // waves wait there; every iteration reads them, as the learned ones or as the class's
        // own at stride 0, so that the loads are unconditional and need no default values
        float la[E - I], lb[E - I];
#pragma unroll
        for (int l = I; l < E; ++l) {
            const gfloat* p = NMS ? nullptr : uni(ab + st * l);
            la[l - I] = NMS ? a : p[0];
            lb[l - I] = NMS ? b : p[1];
        }
        if constexpr (Sched<BG>::s.barrier[I]) __syncthreads();
        if constexpr (P >= 0)
            one_layer<BG, HALF, SOLO, ZC, NMS, DUAL, P>(gofs, rbp, rb, active, set, edge, msg, la[P - I], lb[P - I],
                                                        xbase, lr, ahead);
        group_rest<BG, HALF, SOLO, ZC, NMS, DUAL, I, I, E, P>(gofs, rbp, rb, active, set, edge, msg, la, lb, xbase,
                                                              lr, ahead);
        if constexpr (DUAL && E == I + 1 && Dual<BG>::p.pre_of[I] >= 0) {
            // the other set reads its layer of the next group ahead
            constexpr int J = Dual<BG>::p.pre_of[I];
            if (set != Dual<BG>::p.owner[I] && active)
                layer<BG, HALF, SOLO, ZC, NMS, DUAL, J, slot_of<BG, DUAL>(J), 0, degree_of<BG>(J), false,
                      Dual<BG>::p.premask[J], true>(gofs, rbp, rb, edge, msg, a, b, true, xbase, set, lr, ahead);
            else
                ahead_undefined<Dual<BG>::p.premask[J]>(ahead);
        }
        sweep<BG, HALF, SOLO, ZC, NMS, DUAL, E>(gofs, rbp, rb, active, set, edge, msg, ab, st, a, b, xbase, lr, ahead);
    }
}

template <int BG>
constexpr int max_degree() {
    int m = 0;
    for (int i = 0; i < Graph<BG>::rows; ++i) m = row_begin<BG>(i + 1) - row_begin<BG>(i) > m ? row_begin<BG>(i + 1) - row_begin<BG>(i) : m;
    return m;
}
constexpr int kSyndSlots = Graph<1>::rows * max_degree<1>() > Graph<2>::rows * max_degree<2>()
                               ? Graph<1>::rows * max_degree<1>()
                               : Graph<2>::rows * max_degree<2>();

// The tables of the task's class in LDS, and the workgroup's scratch.
struct Shared {
    // per base entry: its shift in bytes in both 16-bit halves, and its column's first
    // posterior (the byte offset of column j, j Z times the element size)
    uint2 edge[NR_BG1_ENTRIES];
    uint16_t row_start[48];
    // Z % 32 == 0: per row, max_degree entries (word of the edge's column, bit) in the packed
    // decisions, the row's edges then entries of the zero column
    uint32_t synd[kSyndSlots];
    uint32_t fails[kMaxGroup];  // per codeword of the task: the syndrome stamp of a failing check
    uint32_t task;
    uint32_t vote[3];           // any_lane's flags, in turn
};

// Whether any lane of the workgroup has `pred` (one barrier). Three flags in turn: call k
// sets flag k mod 3 and, past its barrier, clears flag k + 2 mod 3, which call k + 2 sets
// only after the barrier of call k + 1; so calls may follow each other with no barrier
// between them.
__device__ __forceinline__ bool any_lane(Shared& sh, bool pred, uint32_t& round) {
    const uint32_t slot = round % 3u;
    ++round;
    if (pred) sh.vote[slot] = 1u;
    __syncthreads();
    if (threadIdx.x == 0) sh.vote[(slot + 2u) % 3u] = 0u;
    return sh.vote[slot] != 0u;
}
static_assert(sizeof(Shared) <= kSharedBytes, "the tables outgrow their LDS slot");

// The packed decisions: per codeword and base column j, the signs of the column's Z
// posteriors as Z/32 words (bit k of word w = posterior 32 w + k), stored twice in a row so
// that a funnel shift across two words reads any 32 consecutive positions of the circulant;
// after the last column a zero column (2 Z/32 zero words), where the padding entries of the
// row table point.
//
// A lane packs one word from its 16-byte pieces with funnel shifts, each shift bringing in
// one sign bit. Words are 128 (FP32) or 64 (FP16) bytes apart, so lanes reading piece k
// together would hit the same LDS banks; each lane starts at its own piece instead (lanes
// 2m, 2m+1 at piece m in FP32, lanes 4m..4m+3 at piece m in FP16), which puts the word
// together rotated by that many pieces, and one funnel shift turns it back.
template <bool HALF>
__device__ __forceinline__ uint32_t piece_signs(uint32_t acc, uint4 x) {
    if constexpr (!HALF) {
        acc = __builtin_amdgcn_alignbit(acc, x.w, 31u);
        acc = __builtin_amdgcn_alignbit(acc, x.z, 31u);
        acc = __builtin_amdgcn_alignbit(acc, x.y, 31u);
        acc = __builtin_amdgcn_alignbit(acc, x.x, 31u);
    } else {
        // two halves a word: the high one's sign is bit 31, the low one's bit 15
        const uint32_t d[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int i = 3; i >= 0; --i) {
            acc = __builtin_amdgcn_alignbit(acc, d[i], 31u);
            acc = __builtin_amdgcn_alignbit(acc, d[i] << 16, 31u);
        }
    }
    return acc;
}

template <bool HALF>
__device__ __forceinline__ uint32_t pack_word(const typename Store<HALF>::T* p) {
    constexpr int pieces = HALF ? 4 : 8, bits = 32 / pieces;
    const uint32_t rot = (threadIdx.x >> (HALF ? 2 : 1)) & uint32_t(pieces - 1);
    const char* w = reinterpret_cast<const char*>(p);
    uint32_t acc = 0;
    // pieces rot + pieces - 1, ..., rot + 1, rot (mod pieces): piece rot ends at bits 0..
#pragma unroll
    for (int k = pieces - 1; k >= 0; --k)
        acc = piece_signs<HALF>(acc, *reinterpret_cast<const uint4*>(w + ((rot + uint32_t(k)) & uint32_t(pieces - 1)) * 16u));
    // bit b of acc is bit (b + rot bits) mod 32 of the word: rotate left by rot bits (the
    // funnel shift takes its amount mod 32)
    return __builtin_amdgcn_alignbit(acc, acc, 32u - rot * uint32_t(bits));
}

// Whether the syndrome of the lane's codeword fails (ldpc.rs:555-557). Every codeword of
// the task is checked; a failing codeword has fails[g] == stamp afterwards.
template <int BG, bool HALF>
__device__ __forceinline__ bool syndrome_fails(const typename Store<HALF>::T* post_all, uint32_t* hd_all, Shared& sh,
                               uint32_t stamp, int z, int n, int cnt, int g, int r, bool lane_ok) {
    using S = Store<HALF>;
    if ((z & 31) == 0) {
        const int w32 = z >> 5, words = n >> 5, per_cw = (Graph<BG>::cols + 1) * 2 * w32;
        for (int f = threadIdx.x; f < cnt * words; f += blockDim.x) {
            const int gg = f / words, fw = f - gg * words;
            const int j = fw / w32, w = fw - j * w32;
            const uint32_t word = pack_word<HALF>(post_all + gg * n + fw * 32);
            uint32_t* hd = hd_all + gg * per_cw + j * 2 * w32;
            hd[w] = word;
            hd[w32 + w] = word;
        }
        __syncthreads();
        // a row word of 32 parities: the XOR of the funnel-shifted column words of the row's
        // edges, every row read through max_degree table entries so the reads go out together
        constexpr int maxd = max_degree<BG>();
        const int per_task = Graph<BG>::rows * w32;
        for (int task = threadIdx.x; task < cnt * per_task; task += blockDim.x) {
            const int gg = task / per_task, rem = task - gg * per_task;
            const int i = rem / w32, w = rem - i * w32;
            const uint32_t* hd = hd_all + gg * per_cw + w;
            const uint32_t* tab = sh.synd + i * maxd;
            uint32_t xs[maxd];
#pragma unroll
            for (int k = 0; k < maxd; ++k) xs[k] = tab[k];
            uint32_t par = 0;
#pragma unroll
            for (int k = 0; k < maxd; ++k) {
                const uint32_t* q = hd + (xs[k] >> 5);
                par ^= __builtin_amdgcn_alignbit(q[1], q[0], xs[k] & 31u);
            }
            if (par) sh.fails[gg] = stamp;
        }
    } else if (lane_ok) {
        // every lane checks its rows edge by edge
        const uint32_t rbp = row_pair<HALF>(r, z), gb = uint32_t(g * n) * uint32_t(sizeof(typename S::T));
        uint32_t bad = 0;
        for (int i = 0; i < Graph<BG>::rows && !bad; ++i) {
            uint32_t par = 0;
            for (int eb = sh.row_start[i]; eb < sh.row_start[i + 1]; ++eb) {
                const uint2 ed = sh.edge[eb];
                par ^= S::sign_at(rotate_bytes(rbp, ed.x) + ed.y + gb);
            }
            bad |= par;
        }
        if (bad) sh.fails[g] = stamp;
    }
    __syncthreads();
    return lane_ok && sh.fails[g] == stamp;
}

// A short exact test ahead of the full syndrome: whether a check of a few chosen layers
// fails on the lane's row (read directly, as the layer reads it). A failing check is
// enough to keep decoding; the full syndrome runs only when these all hold. The layers
// are the ones whose checks most often still fail on a codeword that does not yet satisfy
// every check (measured on the reference decoder's iterations: about 97 % of such
// iterations for base graph 1, 95 % for base graph 2, with the two layers below).
template <int BG>
struct Precheck;
template <>
struct Precheck<1> {
    static constexpr int layers[2] = {11, 42};
};
template <>
struct Precheck<2> {
    static constexpr int layers[2] = {1, 26};
};

template <int BG, bool HALF, int ZC, int L>
__device__ __forceinline__ uint32_t row_parity(uint32_t rbp, uint32_t gbytes, const uint2* edge) {
    using S = Store<HALF>;
    constexpr int e0 = row_begin<BG>(L), deg = row_begin<BG>(L + 1) - e0;
    constexpr uint32_t es = sizeof(typename S::T);
    // every sign first, then their parity, so that the reads go out together
    uint32_t sg[deg];
#pragma unroll
    for (int q = 0; q < deg; ++q) {
        const int e = e0 + q;
        if constexpr (ZC != 0) {
            const uint32_t s_q = shift_of<BG, ZC>(e) * es;
            const uint32_t rr = s_q == 0 ? (rbp & 0xffffu) : rotate_bytes(rbp, s_q | (s_q << 16));
            sg[q] = S::sign_at(rr + uint32_t(Graph<BG>::t[e].j) * ZC * es);
        } else {
            const uint2 ed = edge[e];
            sg[q] = S::sign_at(rotate_bytes(rbp, ed.x) + ed.y + gbytes);
        }
    }
    uint32_t par = 0;
#pragma unroll
    for (int q = 0; q < deg; ++q) par ^= sg[q];
    return par;
}

// Whether the syndrome of a task of one codeword fails: the short test, then (if no check of
// it fails) the full syndrome. The same answer on every lane.
template <int BG, bool HALF, int ZC>
__device__ __forceinline__ bool solo_syndrome_fails(const typename Store<HALF>::T* post_all, uint32_t* hd_all,
                                                    Shared& sh, uint32_t stamp, uint32_t& round, int z, int n,
                                                    int r, bool lane_ok) {
    const uint32_t rbp = row_pair<HALF>(r, z);
    const uint32_t bad = lane_ok ? row_parity<BG, HALF, ZC, Precheck<BG>::layers[0]>(rbp, 0u, sh.edge) |
                                       row_parity<BG, HALF, ZC, Precheck<BG>::layers[1]>(rbp, 0u, sh.edge)
                                 : 0u;
    if (any_lane(sh, bad != 0u, round)) return true;
    syndrome_fails<BG, HALF>(post_all, hd_all, sh, stamp, z, n, 1, 0, r, lane_ok);
    return sh.fails[0] == stamp;
}

// The same for a task of several codewords: each lane tests its codeword's row, a failing
// check marks the codeword, and the full syndrome runs only if a codeword still decoding
// passed the test. `finished`: the lane's codeword has stopped (its answer is not used).
template <int BG, bool HALF, int ZC>
__device__ __forceinline__ bool group_syndrome_fails(const typename Store<HALF>::T* post_all, uint32_t* hd_all,
                                                     Shared& sh, uint32_t stamp, uint32_t& round, int z, int n,
                                                     int cnt, int g, int r, bool lane_ok, bool finished) {
    const uint32_t rbp = row_pair<HALF>(r, z), gb = uint32_t(g * n) * uint32_t(sizeof(typename Store<HALF>::T));
    const uint32_t bad = lane_ok ? row_parity<BG, HALF, ZC, Precheck<BG>::layers[0]>(rbp, gb, sh.edge) |
                                       row_parity<BG, HALF, ZC, Precheck<BG>::layers[1]>(rbp, gb, sh.edge)
                                 : 0u;
    if (bad) sh.fails[g] = stamp;
    __syncthreads();
    const bool known = lane_ok && sh.fails[g] == stamp;
    if (any_lane(sh, lane_ok && !finished && !known, round))
        return syndrome_fails<BG, HALF>(post_all, hd_all, sh, stamp, z, n, cnt, g, r, lane_ok);
    return known;
}

// full_llr (nr/ldpc.rs:191-197): position p is 0.0 plus its received LLRs in order (none
// when p is not sent, several when E repeats it), or 30.0 on a filler. Lane r fills
// positions r, r + z, ..., one per base column, a quarter of the columns at a time so that
// the table and LLR loads of those positions go out together.
template <int BG, bool HALF>
__device__ __forceinline__ void place_llrs(typename Store<HALF>::T* post, const float* llr, const BatchClass& c,
                                           int r, int z) {
    constexpr int C = Graph<BG>::cols, U = (C + 3) / 4;
    // opaque, or the table addresses of a launch of one code (loop-invariant) would be
    // hoisted out of the task loop and held in registers through every sweep
    asm volatile("" : "+v"(r));
#pragma unroll
    for (int k0 = 0; k0 < C; k0 += U) {
        uint32_t a[U], b[U], s0[U];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k0 + u >= C) continue;
            const int p = r + (k0 + u) * z;
            a[u] = c.src_start[p];
            b[u] = c.src_start[p + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < C) s0[u] = c.src[a[u] < b[u] ? a[u] : 0u];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < C) v[u] = a[u] < b[u] ? 0.0f + llr[s0[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k0 + u >= C) continue;
            const int p = r + (k0 + u) * z;
            for (uint32_t s = a[u] + 1; s < b[u]; ++s) v[u] += llr[c.src[s]];
            Store<HALF>::st(post, p, p >= c.k_info && p < c.k ? 30.0f : v[u]);
        }
    }
}

// full_llr when no position is repeated: every position 0.0 (30.0 on a filler), then each
// received LLR, as 0.0 plus it, at its position map[t]. The codeword's lanes load the LLRs
// and the map as 16-byte vectors side by side, all before the first store. Every lane of
// the workgroup calls it (one barrier).
template <int BG, bool HALF>
__device__ __forceinline__ void place_direct(typename Store<HALF>::T* post, const float* llr, const BatchClass& c,
                                             int r, int z, bool lane_ok) {
    using S = Store<HALF>;
    constexpr int C = Graph<BG>::cols, V = (C + 3) / 4;
    // opaque, or the filler tests of a launch of one code (loop-invariant) would be hoisted
    // out of the task loop and held in registers through every sweep
    asm volatile("" : "+v"(r));
    if (lane_ok) {
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const int p = r + k * z;
            S::st(post, p, p >= c.k_info && p < c.k ? 30.0f : 0.0f);
        }
    }
    __syncthreads();
    if (!lane_ok) return;
    const int e = c.e;
    if ((e & 3) == 0 && ((reinterpret_cast<uintptr_t>(llr) | reinterpret_cast<uintptr_t>(c.map)) & 15u) == 0) {
        // at most V vectors a lane: e <= n - 2 z = (C - 2) z when nothing repeats
        const int e4 = e >> 2;
        const float4* l4 = reinterpret_cast<const float4*>(llr);
        const uint4* m4 = reinterpret_cast<const uint4*>(c.map);
        // in two halves, to stay within the registers
        constexpr int H = (V + 1) / 2;
#pragma unroll
        for (int v0 = 0; v0 < V; v0 += H) {
            float4 x[H];
            uint4 m[H];
#pragma unroll
            for (int v = 0; v < H; ++v) {
                const int i = min(r + (v0 + v) * z, e4 - 1);
                x[v] = l4[i];
                m[v] = m4[i];
            }
#pragma unroll
            for (int v = 0; v < H; ++v) {
                if (v0 + v >= V || r + (v0 + v) * z >= e4) continue;
                S::st(post, int(m[v].x), 0.0f + x[v].x);
                S::st(post, int(m[v].y), 0.0f + x[v].y);
                S::st(post, int(m[v].z), 0.0f + x[v].z);
                S::st(post, int(m[v].w), 0.0f + x[v].w);
            }
        }
    } else {
        // one LLR a load, a quarter of the lane's at a time (at most C - 2 of them)
        constexpr int Q = (C + 3) / 4;
#pragma unroll
        for (int u0 = 0; u0 < C; u0 += Q) {
            float x[Q];
            uint32_t m[Q];
#pragma unroll
            for (int u = 0; u < Q; ++u) {
                const int t = min(r + (u0 + u) * z, e - 1);
                x[u] = llr[t];
                m[u] = c.map[t];
            }
#pragma unroll
            for (int u = 0; u < Q; ++u)
                if (u0 + u < C && r + (u0 + u) * z < e) S::st(post, int(m[u]), 0.0f + x[u]);
        }
    }
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
    u.map = uni(c.map);
    u.direct = uni(c.direct);
    return u;
}

// One task: codewords order[first .. first + cnt) of class c (base graph BG). SOLO: the
// task has one codeword (G = 1), whose posteriors start the LDS; ZC: Z, when every class of
// the launch has the same (0: read at run time); DUAL: two sets of 384 lanes on a task of one
// codeword (FP32), each running its layers on rows 0 .. Z-1 (the second set places no LLRs
// and writes no outputs).
template <int BG, bool HALF, bool SOLO, int ZC, bool NMS, bool DUAL>
__device__ __forceinline__ void run_task(const LayeredArgs& a, const BatchClass& c, const float* own_ab, uint32_t first,
                                         int cnt, Shared& sh, void* lds_dyn, uint32_t& stamp, uint32_t& round) {
    using S = Store<HALF>;
    using T = typename S::T;
    const int z = ZC ? ZC : c.z, n = ZC ? Graph<BG>::cols * ZC : c.n;
    T* post_all = static_cast<T*>(lds_dyn);
    uint32_t* hd_all = reinterpret_cast<uint32_t*>(post_all + c.g * n);  // g n is even: 4-byte aligned
    static_assert(!DUAL || (SOLO && !HALF), "two wave sets serve FP32 tasks of one codeword");
    const int tid = threadIdx.x;
    // DUAL: the lane's set, uniform over a wavefront (384 lanes are six of them)
    const int set = DUAL ? int(__builtin_amdgcn_readfirstlane(uint32_t(tid >= kMaxThreads))) : 0;
    const int g = SOLO ? 0 : tid / z, r = tid - g * z - set * kMaxThreads;
    // lanes beyond the task's codewords (SOLO: beyond Z) idle
    const bool lane_ok = SOLO ? r < z : g < cnt;
    const bool places = !DUAL || set == 0;  // places the LLRs and writes the outputs
    const uint32_t id = lane_ok ? (a.order ? a.order[first + g] : first + uint32_t(g)) : 0u;
    T* post = post_all + (lane_ok ? g : 0) * n;
    const float* llr = a.llr + (a.llr_off ? a.llr_off[id] : uint64_t(id) * uint64_t(c.e));
    const int limit = a.max_it ? int(a.max_it[id]) : c.max_iterations;
    // the row's byte offset in a column and that minus Z (mod 2^16), in 16-bit halves, and
    // the byte offset of the lane's codeword (the posteriors start at LDS address 0)
    uint32_t rbp = row_pair<HALF>(r, z), rb = rbp & 0xffffu;
    const uint32_t gbytes = uint32_t(lane_ok ? g : 0) * uint32_t(n) * uint32_t(sizeof(T));

    if (c.direct) place_direct<BG, HALF>(post, llr, c, r, z, lane_ok && places);
    else if (lane_ok && places) place_llrs<BG, HALF>(post, llr, c, r, z);
    if ((z & 31) == 0) {
        // the zero column of every codeword's packed decisions
        const int w32 = z >> 5, per_cw = (Graph<BG>::cols + 1) * 2 * w32;
        for (int t = tid; t < cnt * 2 * w32; t += blockDim.x) {
            const int gg = t / (2 * w32);
            hd_all[gg * per_cw + Graph<BG>::cols * 2 * w32 + (t - gg * 2 * w32)] = 0u;
        }
    }
    __syncthreads();

    Messages<msg_slots<BG, DUAL>(), HALF> msg;
    msg.zero();
    LaneRegs lr{0x80000000u, 0.0f, c.a, c.b, rb + 65536u, 1u};
    // (only the two-set kernel, where each was timed faster: the FP16 kernels leave them to
    // the compiler as constants and scalars, since at 168 registers the registers they and
    // the edge numbers would hold cost spills, and the one-set FP32 kernels run as fast or
    // slightly faster without them)
    if constexpr (DUAL)
        asm volatile("" : "+v"(lr.sgn), "+v"(lr.zero), "+v"(lr.a), "+v"(lr.b), "+v"(lr.rbh), "+v"(lr.one));
    Ahead ahead;
    bool finished = !lane_ok, restarted = false;
    uint32_t its = 0, ok = 0;
    // the syndrome, the short test first
    auto fails_now = [&]() -> bool {
        ++stamp;
        if constexpr (SOLO) return solo_syndrome_fails<BG, HALF, ZC>(post_all, hd_all, sh, stamp, round, z, n, r,
                                                                     lane_ok);
        else return group_syndrome_fails<BG, HALF, ZC>(post_all, hd_all, sh, stamp, round, z, n, cnt, g, r, lane_ok,
                                                       finished);
    };
    if (!fails_now() && lane_ok) {
        finished = true;  // the channel LLRs are a codeword: 0 iterations
        ok = 1;
    }
    // NMS instantiations serve launches without learned parameters whose classes are all
    // normalized min-sum (b = 0)
    bool learned = !NMS && a.params != nullptr;
    int it = 0, offset = 0;
    // Z equal to the lanes of a set: every lane owns a row of the task's one codeword, and
    // `finished` is the same in every lane (it changes only on uniform conditions: the
    // syndrome's verdict, the iteration count, the codeword's limit), so the loop needs no
    // vote across the workgroup (a barrier an iteration)
    constexpr bool kUniform = SOLO && ZC == kMaxThreads;
    for (;;) {
        if (learned && it >= a.learned_iterations) {
            // the end of the learned iterations, the same for every codeword of the task
            if (!finished) {
                if (a.fallback) {
                    restarted = true;
                    if (places) place_llrs<BG, HALF>(post, llr, c, r, z);
                    msg.zero();
                    // no vote below: the placement's own barrier before the sweep reads it
                    if constexpr (kUniform) __syncthreads();
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
        // (uniform: read from the first lane, so that the loop's exit is a scalar branch)
        if (kUniform ? __builtin_amdgcn_readfirstlane(uint32_t(finished)) != 0u : !any_lane(sh, !finished, round)) break;
        ++it;
        // uniform, and a global-memory pointer: the sweep's reads of the parameters are then
        // vector-memory loads that only the parameters' own use waits for (a generic pointer
        // would make them flat loads, which every barrier's LDS wait waits for too)
        // (the class's own (a, b) for every layer: stride 0)
        const gfloat* ab = NMS ? nullptr
                               : to_global(uni(learned ? a.params + size_t(it - 1) * 2 * Graph<BG>::rows : own_ab));
        const int st = uni(int32_t(learned ? 2 : 0));
        sweep<BG, HALF, SOLO, ZC, NMS, DUAL, 0>(gbytes, rbp, rb, !finished, set, sh.edge, msg, ab, st, lr.a, lr.b,
                                                uint32_t(a.shared_off) + uint32_t(kSharedBytes), lr, ahead);
        __syncthreads();
        const bool fails = fails_now();
        if (!finished && !fails) {
            finished = true;
            its = uint32_t(offset + it);
            if (!restarted) ok = 1;
        }
    }

    if (lane_ok && places) {
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

// Workgroups take tasks until none is left. BGMASK: the base graphs of the classes; SOLO:
// every class has G = 1; ZC: every class has Z = ZC (0: any); NMS: every class runs the
// normalized min-sum and no learned parameters are given; DUAL: two wave sets a codeword.
// `a` is the calling kernel's only parameter: a launch of one class reads its (a, b) from
// the kernel's argument segment.
template <int BGMASK, bool HALF, bool SOLO, int ZC, bool NMS, bool DUAL>
__device__ __forceinline__ void layered_body(LayeredArgs a) {
    // one dynamic allocation and no static LDS, so that the posteriors start at address 0
    // and the column bases are immediate offsets of the LDS instructions
    extern __shared__ __align__(16) uint32_t lds_dyn[];
    Shared& sh = *reinterpret_cast<Shared*>(reinterpret_cast<char*>(lds_dyn) + a.shared_off);
    for (int i = threadIdx.x; i < kMaxGroup; i += blockDim.x) sh.fails[i] = 0u;
    if (threadIdx.x < 3) sh.vote[threadIdx.x] = 0u;


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
        // the class's (a, b) in global memory, which the sweep reads layer by layer when it has
        // no learned ones: for a launch of one class, in the kernel's argument segment (the
        // arguments are the kernel's one parameter, from the segment's first byte; taking the
        // parameter's own address would copy it to private memory)
        static_assert(offsetof(BatchClass, b) == offsetof(BatchClass, a) + sizeof(float),
                      "the sweep reads a class's (a, b) as two adjacent floats");
        const float* own_ab =
            a.tasks ? &a.classes[cls].a
                    : reinterpret_cast<const float*>(reinterpret_cast<uintptr_t>(__builtin_amdgcn_kernarg_segment_ptr()) +
                                                     offsetof(LayeredArgs, one) + offsetof(BatchClass, a));
        if (cls != loaded) {
            // the shifts (read with uniform addresses), columns and row starts in LDS
            const int edges = c.bg == 1 ? Graph<1>::edges : Graph<2>::edges;
            const int rows = c.bg == 1 ? Graph<1>::rows : Graph<2>::rows;
            const int cols = c.bg == 1 ? Graph<1>::cols : Graph<2>::cols;
            constexpr int maxd1 = max_degree<1>(), maxd2 = max_degree<2>();
            const int maxd = c.bg == 1 ? maxd1 : maxd2;
            const uint32_t w32 = uint32_t(c.z >> 5);
            for (int e = threadIdx.x; e < edges; e += blockDim.x) {
                const uint32_t sb = c.shift32[e] * uint32_t(HALF ? 2 : 4);
                sh.edge[e] = uint2{sb | (sb << 16), uint32_t(c.col[e]) * uint32_t(c.z) * uint32_t(HALF ? 2 : 4)};
            }
            for (int i = threadIdx.x; i <= rows; i += blockDim.x) sh.row_start[i] = c.row_start[i];
            // the funnel shifts of the packed syndrome, row by row: edge (j, v) reads word
            // j 2 w32 + v / 32 from bit v % 32; the padding reads the zero column
            for (int s = threadIdx.x; s < rows * maxd; s += blockDim.x) {
                const int i = s / maxd, k = s - i * maxd, e = int(c.row_start[i]) + k;
                uint32_t x = (uint32_t(cols) * 2u * w32) << 5;
                if (e < int(c.row_start[i + 1])) {
                    const uint32_t v = c.shift32[e], j = c.col[e];
                    x = ((j * 2u * w32 + (v >> 5)) << 5) | (v & 31u);
                }
                sh.synd[s] = x;
            }
            __syncthreads();
            loaded = cls;
        }
        if constexpr ((BGMASK & 1) != 0) {
            if (c.bg == 1) run_task<1, HALF, SOLO, ZC, NMS, DUAL>(a, c, own_ab, first, cnt, sh, lds_dyn, stamp, round);
        }
        if constexpr ((BGMASK & 2) != 0) {
            if (c.bg == 2) run_task<2, HALF, SOLO, ZC, NMS, DUAL>(a, c, own_ab, first, cnt, sh, lds_dyn, stamp, round);
        }
        if (!a.queue) t += gridDim.x;
    }
}

template <int BGMASK, bool HALF, bool SOLO, int ZC, bool NMS = false>
__global__ void __launch_bounds__(kMaxThreads, HALF ? 3 : 1) decode_layered_kernel(LayeredArgs a) {
    layered_body<BGMASK, HALF, SOLO, ZC, NMS, false>(a);
}

// 768 lanes: twelve waves a CU, three on every SIMD (at most 168 registers a lane)
template <int BGMASK, int ZC, bool NMS>
__global__ void __launch_bounds__(2 * kMaxThreads, 1) decode_layered_dual_kernel(LayeredArgs a) {
    layered_body<BGMASK, false, true, ZC, NMS, true>(a);
}

}  // namespace layered
}  // namespace comfec
