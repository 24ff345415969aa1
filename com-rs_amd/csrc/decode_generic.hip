// The general belief-propagation kernel of the lifted NR codes: any base graph and lifting
// size, layered or flooding schedule, normalized / offset min-sum or sum-product, f32 or
// f64 messages. It is the decoder of crates/com-fec/src/ldpc.rs:449-557 laid out for a
// 64-wide wavefront: lane r of a workgroup owns lifted row r of every base row (layer), so
// one layer is Z independent rows updated in parallel, which is exactly the reference's
// row-after-row order because the Z rows of a circulant layer share no variable.
//
// Per codeword: the posterior LLRs live in LDS when they fit (else in a per-workgroup
// scratch in HBM); the check-to-variable messages live in a per-workgroup scratch laid out
// [base edge][r], so lane r's accesses to one edge are contiguous across the wavefront.
// Workgroups are persistent: each walks codewords blockIdx.x, +gridDim.x, ...
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "sizing.hpp"

namespace comfec {

namespace {

constexpr int kMaxDeg = 19;  // largest row degree of BG1 (BG2: 10)

template <typename T>
__device__ inline T tanh_t(T x) { return tanh(x); }
template <typename T>
__device__ inline T atanh_t(T x) { return atanh(x); }

// The reference's clamp of the tanh product, |p| <= 0.999999999999 (ldpc.rs:93, 465), as a
// bound on the message: 2 atanh(0.999999999999).
constexpr float kSpaCap = 28.324190418452805f;

// phi(x) = -ln tanh(x/2) for x >= 0 (phi(phi(x)) = x), written as ln(1 + 2 e^-x / (1 - e^-x))
// with 1 - e^-x = -expm1(-x) so that FP32 keeps its relative precision at both ends. The
// ends are taken apart so that no zero or infinity reaches the library functions: below
// 2^-12, tanh(x/2) = x/2 (1 - x^2/12 + ...) gives ln 2 - ln x within 5e-9 (and phi(0) = inf);
// past 128 (also for x = inf), e^-x is 0 in FP32 and phi(x) = 0.
__device__ inline float spa_phi(float x) {
    if (x < 0x1p-12f) return x == 0.0f ? INFINITY : 0.693147180559945f - logf(x);
    x = fminf(x, 128.0f);
    return log1pf(2.0f * expf(-x) / -expm1f(-x));
}

// ldpc.rs:449-493 for one row of degree `deg`, messages in[] -> out[].
template <typename T>
__device__ inline void check_row(int rule, T alpha, const T (&in)[kMaxDeg], T (&out)[kMaxDeg], int deg) {
    if constexpr (sizeof(T) == 4) {
        if (rule == 0) {
            // sum-product in FP32, in Gallager's form: |prod_j tanh(l_j/2)| = exp(-S) with
            // S = sum_j phi(|l_j|), and 2 atanh(exp(-S)) = phi(S). The product form loses
            // everything in FP32 once tanh(l/2) rounds to 1 (|l| > 17): the clamp then
            // yields atanh(1) = inf. S over the others by prefix and suffix sums, the sign
            // the parity of the others' sign bits (tanh keeps the sign of -0.0 too).
            float s[kMaxDeg];
            float prefix = 0.0f;
            uint32_t parity = 0;
#pragma unroll
            for (int q = 0; q < kMaxDeg; ++q)
                if (q < deg) {
                    s[q] = spa_phi(fabsf(in[q]));
                    out[q] = prefix;
                    prefix += s[q];
                    parity ^= __float_as_uint(in[q]);
                }
            float suffix = 0.0f;
#pragma unroll
            for (int q = kMaxDeg - 1; q >= 0; --q)
                if (q < deg) {
                    float m = fminf(spa_phi(out[q] + suffix), kSpaCap);
                    suffix += s[q];
                    out[q] = __uint_as_float(__float_as_uint(m) |
                                             ((parity ^ __float_as_uint(in[q])) & 0x80000000u));
                }
            return;
        }
    }
    if (rule == 0) {
        // sum-product: exclusive prefix x suffix products of tanh(l/2), 2 atanh(clamp)
        T t[kMaxDeg];
        T prefix = T(1);
#pragma unroll
        for (int q = 0; q < kMaxDeg; ++q)
            if (q < deg) {
                t[q] = tanh_t(in[q] / T(2));
                out[q] = prefix;
                prefix *= t[q];
            }
        T suffix = T(1);
        const T bound = T(0.999999999999);
#pragma unroll
        for (int q = kMaxDeg - 1; q >= 0; --q)
            if (q < deg) {
                T p = out[q] * suffix;
                suffix *= t[q];
                p = p < -bound ? -bound : (p > bound ? bound : p);
                out[q] = T(2) * atanh_t(p);
            }
        return;
    }
    // min-sum: sign parity, first minimum (strict <) and second minimum in one scan
    bool neg = false;
    T m1 = T(INFINITY), m2 = T(INFINITY);
    int i1 = 0;
#pragma unroll
    for (int q = 0; q < kMaxDeg; ++q)
        if (q < deg) {
            T a = fabs(in[q]);
            neg ^= in[q] < T(0);
            if (a < m1) {
                m2 = m1;
                m1 = a;
                i1 = q;
            } else if (a < m2) {
                m2 = a;
            }
        }
    T g1, g2;
    if (rule == 1) {
        g1 = alpha * m1;
        g2 = alpha * m2;
    } else {
        g1 = fmax(m1 - alpha, T(0));
        g2 = fmax(m2 - alpha, T(0));
    }
#pragma unroll
    for (int q = 0; q < kMaxDeg; ++q)
        if (q < deg) {
            T g = q == i1 ? g2 : g1;
            bool s = neg ^ (in[q] < T(0));
            out[q] = s ? -g : g;
        }
}

template <typename T>
struct Block {
    const GenericArgs<T>& a;
    T* post;
    T* c2v;
    T* chan;
    int tid, nthreads;

    // whether some check fails on the hard decisions of post (ldpc.rs:555-557); the
    // lanes take the (base row, lifted row) pairs, so Z = 1 codes check rows in parallel
    __device__ bool syndrome_fails() const {
        const DevCode& c = a.code;
        int bad = 0;
        for (int task = tid; task < c.rows_b * c.z && !bad; task += nthreads) {
            const int i = task / c.z, r = task - i * c.z;
            int par = 0;
            for (int eb = c.row_start[i]; eb < c.row_start[i + 1]; ++eb) {
                int rr = r + c.shift[eb];
                rr -= rr >= c.z ? c.z : 0;
                par ^= post[c.col[eb] * c.z + rr] < T(0);
            }
            bad |= par;
        }
        return __syncthreads_or(bad) != 0;
    }

    // ldpc.rs:522-531 for every row of base row i
    __device__ void layer(int i, bool first) {
        const DevCode& c = a.code;
        const int e0 = c.row_start[i], deg = c.row_start[i + 1] - e0;
        for (int r = tid; r < c.z; r += nthreads) {
            T in[kMaxDeg], out[kMaxDeg];
            int idx[kMaxDeg];
#pragma unroll
            for (int q = 0; q < kMaxDeg; ++q)
                if (q < deg) {
                    int eb = e0 + q;
                    int rr = r + c.shift[eb];
                    rr -= rr >= c.z ? c.z : 0;
                    idx[q] = c.col[eb] * c.z + rr;
                    T old = first ? T(0) : c2v[size_t(eb) * c.z + r];
                    in[q] = post[idx[q]] - old;
                }
            check_row<T>(a.rule, T(a.alpha), in, out, deg);
#pragma unroll
            for (int q = 0; q < kMaxDeg; ++q)
                if (q < deg) {
                    post[idx[q]] = in[q] + out[q];
                    c2v[size_t(e0 + q) * c.z + r] = out[q];
                }
        }
    }

    // ldpc.rs:533-544: new messages of every row from the old posteriors, in place; the
    // rows are independent, so the lanes take (base row, lifted row) pairs
    __device__ void flood_checks(bool first) {
        const DevCode& c = a.code;
        for (int task = tid; task < c.rows_b * c.z; task += nthreads) {
            const int i = task / c.z, r = task - i * c.z;
            const int e0 = c.row_start[i], deg = c.row_start[i + 1] - e0;
            {
                T in[kMaxDeg], out[kMaxDeg];
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) {
                        int eb = e0 + q;
                        int rr = r + c.shift[eb];
                        rr -= rr >= c.z ? c.z : 0;
                        T old = first ? T(0) : c2v[size_t(eb) * c.z + r];
                        in[q] = post[c.col[eb] * c.z + rr] - old;
                    }
                check_row<T>(a.rule, T(a.alpha), in, out, deg);
#pragma unroll
                for (int q = 0; q < kMaxDeg; ++q)
                    if (q < deg) c2v[size_t(e0 + q) * c.z + r] = out[q];
            }
        }
    }

    // post = channel + the new messages of every check, in ascending check order
    __device__ void flood_variables() {
        const DevCode& c = a.code;
        for (int p = tid; p < c.n; p += nthreads) {
            int j = p / c.z, q = p - j * c.z;
            T v = chan[p];
            for (int s = c.col_start[j]; s < c.col_start[j + 1]; ++s) {
                int eb = c.col_edge[s];
                int r = q - int(c.shift[eb]);
                r += r < 0 ? c.z : 0;
                v += c2v[size_t(eb) * c.z + r];
            }
            post[p] = v;
        }
    }
};

template <typename T, bool kPostLds>
__global__ void __launch_bounds__(512) decode_generic_kernel(GenericArgs<T> a) {
    extern __shared__ unsigned char lds_raw[];
    const DevCode& c = a.code;
    T* scratch = a.scratch + size_t(blockIdx.x) * a.scratch_stride;
    Block<T> b{a, nullptr, scratch, nullptr, int(threadIdx.x), int(blockDim.x)};
    size_t off = size_t(c.edges_b) * c.z;
    if (kPostLds) {
        b.post = reinterpret_cast<T*>(lds_raw);
    } else {
        b.post = scratch + off;
        off += c.n;
    }
    b.chan = a.layered ? nullptr : scratch + off;

    for (int64_t w = blockIdx.x; w < a.count; w += gridDim.x) {
        // the channel LLRs of the full codeword (nr/ldpc.rs:191-197 or ldpc.rs:500-505)
        for (int p = b.tid; p < c.n; p += b.nthreads) {
            T v;
            if (a.input_full) {
                v = a.llr[w * c.n + p];
            } else {
                v = T(0);
                for (uint32_t s = c.src_start[p]; s < c.src_start[p + 1]; ++s) v += a.llr[w * c.e + c.src[s]];
                if (p >= c.k_info && p < c.k) v = T(30);
            }
            b.post[p] = v;
        }
        if (a.apriori) {
            // a priori LLRs add to the channel LLRs of the information positions
            __syncthreads();
            for (int i = b.tid; i < c.k_info; i += b.nthreads) {
                const int p = c.info ? int(c.info[i]) : i;
                b.post[p] += a.apriori[w * c.k_info + i];
            }
        }
        __syncthreads();
        if (!a.layered)
            for (int p = b.tid; p < c.n; p += b.nthreads) b.chan[p] = b.post[p];
        __syncthreads();
        int done = b.syndrome_fails() ? -1 : 0;
        for (int it = 1; it <= a.max_iterations && done < 0; ++it) {
            if (a.layered) {
                for (int i = 0; i < c.rows_b; ++i) {
                    b.layer(i, it == 1);
                    __syncthreads();
                }
            } else {
                b.flood_checks(it == 1);
                __syncthreads();
                b.flood_variables();
                __syncthreads();
            }
            if (!b.syndrome_fails()) done = it;
        }
        if (done < 0) done = a.max_iterations;
        if (b.tid == 0 && a.iterations) a.iterations[w] = uint32_t(done);
        if (a.bits)
            for (int i = b.tid; i < c.k_info; i += b.nthreads)
                a.bits[w * c.k_info + i] = b.post[c.info ? int(c.info[i]) : i] < T(0);
        if (a.post_out)
            for (int p = b.tid; p < c.n; p += b.nthreads) a.post_out[w * c.n + p] = b.post[p];
        __syncthreads();
    }
}

// Codes given by their rows (no lifting, Z = 1) under the layered schedule. The reference
// updates the checks one after the other (ldpc.rs:522-531) and consecutive checks share
// variables, so a codeword offers no parallelism inside a sweep. Each lane therefore
// decodes codewords of its own, walking the rows in the reference's order; the lanes of a
// wavefront walk the same row at once, and with the state laid out [index][lane] (the
// posteriors, then the check-to-variable messages) each access of a wavefront reads or
// writes 64 consecutive values. Lane g takes codewords g, g + lanes, ...; a codeword that
// stops early frees its lane for the next one at once (the loop below is flat: a lane
// loads, sweeps or finishes independently of the others).
template <typename T>
__global__ void __launch_bounds__(64) decode_rows_kernel(GenericArgs<T> a, int64_t lanes) {
    const DevCode& c = a.code;
    const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= lanes) return;
    const size_t L = size_t(lanes);
    // the row tables as 32-bit words: every lane reads the same entry, so these are scalar loads
    // (a column entry holds the variable in its low 16 bits)
    const uint32_t* __restrict__ row_start = c.row_start32;
    const uint32_t* __restrict__ col = c.col32;
    T* post = a.scratch + gid;                       // post[p * L], p < n
    T* c2v = a.scratch + size_t(c.n) * L + gid;      // c2v[e * L], e < edges
    // ldpc.rs:555-557: whether some check fails on the hard decisions
    auto fails = [&]() -> bool {
        for (int i = 0; i < c.rows_b; ++i) {
            uint32_t par = 0;
            for (uint32_t e = row_start[i]; e < row_start[i + 1]; ++e) par ^= post[size_t(col[e] & 0xffffu) * L] < T(0);
            if (par) return true;
        }
        return false;
    };
    auto finish = [&](int64_t w, int it) {
        if (a.iterations) a.iterations[w] = uint32_t(it);
        if (a.bits)
            for (int i = 0; i < c.k_info; ++i)
                a.bits[w * c.k_info + i] = post[size_t(c.info ? int(c.info[i]) : i) * L] < T(0);
        if (a.post_out)
            for (int p = 0; p < c.n; ++p) a.post_out[w * c.n + p] = post[size_t(p) * L];
    };
    int64_t w = gid;
    int it = -1;  // sweeps done on codeword w; -1: not loaded
    while (w < a.count) {
        if (it < 0) {
            // the channel LLRs of the full codeword, a priori LLRs on the information positions
            for (int p = 0; p < c.n; ++p) {
                T v;
                if (a.input_full) {
                    v = a.llr[w * c.n + p];
                } else {
                    v = T(0);
                    for (uint32_t s = c.src_start[p]; s < c.src_start[p + 1]; ++s) v += a.llr[w * c.e + c.src[s]];
                    if (p >= c.k_info && p < c.k) v = T(30);
                }
                post[size_t(p) * L] = v;
            }
            if (a.apriori)
                for (int i = 0; i < c.k_info; ++i) {
                    const int p = c.info ? int(c.info[i]) : i;
                    post[size_t(p) * L] += a.apriori[w * c.k_info + i];
                }
            it = 0;
            if (a.max_iterations == 0 || !fails()) {
                finish(w, 0);
                w += lanes;
                it = -1;
                continue;
            }
        }
        // one sweep: ldpc.rs:522-531 for rows 0, 1, ..., m - 1 (no messages yet on the first)
        for (int i = 0; i < c.rows_b; ++i) {
            const uint32_t e0 = row_start[i];
            const int deg = int(row_start[i + 1] - e0);
            T in[kMaxDeg], out[kMaxDeg];
            size_t at[kMaxDeg];
#pragma unroll
            for (int q = 0; q < kMaxDeg; ++q)
                if (q < deg) {
                    at[q] = size_t(col[e0 + q] & 0xffffu) * L;
                    const T old = it == 0 ? T(0) : c2v[size_t(e0 + q) * L];
                    in[q] = post[at[q]] - old;
                }
            check_row<T>(a.rule, T(a.alpha), in, out, deg);
#pragma unroll
            for (int q = 0; q < kMaxDeg; ++q)
                if (q < deg) {
                    post[at[q]] = in[q] + out[q];
                    c2v[size_t(e0 + q) * L] = out[q];
                }
        }
        ++it;
        if (it >= a.max_iterations || !fails()) {
            finish(w, it);
            w += lanes;
            it = -1;
        }
    }
}

// ---- codes given by their rows, a group of G lanes per codeword ------------------------
//
// The same schedule as decode_rows_kernel, with the edges of a check spread over a group of
// G = 8, 16 or 32 lanes (edge q on lane q of the group): G times as many codewords in flight
// for a batch, and a check's work done side by side. Groups of 8 and 16 sit inside the
// 16-lane rows of the wavefront, so their exchanges are DPP moves (groups of 32 cross one
// row boundary through ds_swizzle): the check's minimum, first minimum and
// sign parity by butterflies (order-free: a minimum is a selection), the sum-product's
// prefix and suffix by sequential shifts in the reference's order, so that every result is
// the one check_row computes. State laid out [index][group].

template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t x) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(x), CTRL, 0xf, 0xf, true));  // lanes without a source read 0
}
template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
    return __uint_as_float(dpp32<CTRL>(__float_as_uint(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double x) {
    const uint64_t b = __double_as_longlong(x);
    const uint64_t lo = dpp32<CTRL>(uint32_t(b)), hi = dpp32<CTRL>(uint32_t(b >> 32));
    return __longlong_as_double(int64_t(lo | (hi << 32)));
}
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t x) {
    return dpp32<CTRL>(x);
}

constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kHalfMirror = 0x141, kRowMirror = 0x140;
constexpr int kRowShr1 = 0x111, kRowShl1 = 0x101;

// Groups of 32 lanes span two DPP rows: the crossing goes through ds_swizzle within each
// 32 lanes (bit mode: lane' = ((lane & and) | or) ^ xor): lane ^ 16, lane 15, lane 16.
constexpr int kSwapRows = (0x10 << 10) | 0x1F, kFromLane15 = 15 << 5, kFromLane16 = 16 << 5;

template <int PATTERN>
__device__ __forceinline__ uint32_t swz32(uint32_t x) {
    return uint32_t(__builtin_amdgcn_ds_swizzle(int(x), PATTERN));
}
template <int PATTERN>
__device__ __forceinline__ float swz(float x) {
    return __uint_as_float(swz32<PATTERN>(__float_as_uint(x)));
}
template <int PATTERN>
__device__ __forceinline__ double swz(double x) {
    const uint64_t b = __double_as_longlong(x);
    const uint64_t lo = swz32<PATTERN>(uint32_t(b)), hi = swz32<PATTERN>(uint32_t(b >> 32));
    return __longlong_as_double(int64_t(lo | (hi << 32)));
}
template <int PATTERN>
__device__ __forceinline__ uint32_t swz(uint32_t x) {
    return swz32<PATTERN>(x);
}

// op over the G lanes of the group, the result in every lane of it
template <int G, typename V, typename F>
__device__ __forceinline__ V group_all(V x, F op) {
    x = op(x, dpp<kQuadXor1>(x));
    x = op(x, dpp<kQuadXor2>(x));
    x = op(x, dpp<kHalfMirror>(x));
    if constexpr (G >= 16) x = op(x, dpp<kRowMirror>(x));
    if constexpr (G == 32) x = op(x, swz<kSwapRows>(x));
    return x;
}

// exclusive scan in lane order: lane q gets ((id op v_0) op v_1) ... op v_{q-1}
template <int G, typename V, typename F>
__device__ __forceinline__ V group_prefix(V v, V id, int q, int deg, F op) {
    V pre = id;
    for (int k = 1; k < deg; ++k) {
        V cand;
        if (G == 32 && k == 16) cand = swz<kFromLane15>(op(pre, v));
        else cand = dpp<kRowShr1>(op(pre, v));
        pre = q == k ? cand : pre;
    }
    return pre;
}
// and from the end: lane q gets ((id op v_{deg-1}) op v_{deg-2}) ... op v_{q+1}
template <int G, typename V, typename F>
__device__ __forceinline__ V group_suffix(V v, V id, int q, int deg, F op) {
    V suf = id;
    for (int k = deg - 2; k >= 0; --k) {
        V cand;
        if (G == 32 && k == 15) cand = swz<kFromLane16>(op(suf, v));
        else cand = dpp<kRowShl1>(op(suf, v));
        suf = q == k ? cand : suf;
    }
    return suf;
}

// x of wavefront lane `lane` (ds_bpermute: any lane, chosen per lane)
__device__ __forceinline__ float lane_value(float x, int lane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(x)));
}
__device__ __forceinline__ double lane_value(double x, int lane) {
    const uint64_t b = __double_as_longlong(x);
    const uint64_t lo = uint32_t(__builtin_amdgcn_ds_bpermute(lane << 2, int(uint32_t(b))));
    const uint64_t hi = uint32_t(__builtin_amdgcn_ds_bpermute(lane << 2, int(uint32_t(b >> 32))));
    return __longlong_as_double(int64_t(lo | (hi << 32)));
}

// check_row for the edge of lane q (the lanes q >= deg take part with neutral values)
template <int G, typename T>
__device__ __forceinline__ T group_check(int rule, T alpha, T in, int q, int deg) {
    const bool act = q < deg;
    if (rule == 0) {
        if constexpr (sizeof(T) == 4) {
            const float s = act ? spa_phi(fabsf(in)) : 0.0f;
            const float pre = group_prefix<G>(s, 0.0f, q, deg, [](float x, float y) { return x + y; });
            const float suf = group_suffix<G>(s, 0.0f, q, deg, [](float x, float y) { return x + y; });
            const uint32_t parity = group_all<G>(act ? __float_as_uint(in) : 0u,
                                                 [](uint32_t x, uint32_t y) { return x ^ y; });
            const float m = fminf(spa_phi(pre + suf), kSpaCap);
            return __uint_as_float(__float_as_uint(m) | ((parity ^ __float_as_uint(in)) & 0x80000000u));
        } else {
            const T t = act ? tanh_t(in / T(2)) : T(1);
            const T pre = group_prefix<G>(t, T(1), q, deg, [](T x, T y) { return x * y; });
            const T suf = group_suffix<G>(t, T(1), q, deg, [](T x, T y) { return x * y; });
            const T bound = T(0.999999999999);
            T p = pre * suf;
            p = p < -bound ? -bound : (p > bound ? bound : p);
            return T(2) * atanh_t(p);
        }
    }
    const T a = act ? T(fabs(in)) : T(INFINITY);
    const T m1 = group_all<G>(a, [](T x, T y) { return x < y ? x : y; });
    const uint32_t i1 = group_all<G>(act && a == m1 ? uint32_t(q) : uint32_t(G),
                                     [](uint32_t x, uint32_t y) { return x < y ? x : y; });
    const T m2 = group_all<G>(uint32_t(q) == i1 ? T(INFINITY) : a, [](T x, T y) { return x < y ? x : y; });
    const uint32_t neg = group_all<G>(act && in < T(0) ? 1u : 0u, [](uint32_t x, uint32_t y) { return x ^ y; });
    T g1, g2;
    if (rule == 1) {
        g1 = alpha * m1;
        g2 = alpha * m2;
    } else {
        g1 = fmax(m1 - alpha, T(0));
        g2 = fmax(m2 - alpha, T(0));
    }
    const T g = uint32_t(q) == i1 ? g2 : g1;
    return (neg != 0) != (in < T(0)) ? -g : g;
}

constexpr uint32_t kNoColumn = 0xffffffffu;  // the column of a lane past the check's edges

// One check in flight: its edges (read three checks ahead), its posteriors and old messages
// (read two checks ahead) and what the checks in between wrote to its columns.
template <typename T>
struct Pending {
    uint32_t e0;
    int deg;
    uint32_t v;   // the lane's column, kNoColumn past the edges
    uint32_t from;  // where the two checks before list the column (bits 0..5, 6..11), 0 past the edges
    T p, old;     // the posterior as read, the message of the previous sweep
    uint32_t fresh;  // a check after the read wrote the column (1): its value is fv
    T fv;
};

template <typename T, int G>
__global__ void __launch_bounds__(64) decode_groups_kernel(GenericArgs<T> a, int64_t slots) {
    const DevCode& c = a.code;
    const int64_t slot = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int q = int(threadIdx.x) & (G - 1);
    if (slot >= slots) return;  // whole groups: G divides the workgroup
    const size_t L = size_t(slots);
    // element offsets in 32 bits: the scratch holds fewer than 2^32 values (at most 8 GiB)
    const uint32_t Lu = uint32_t(slots);
    auto at = [&](uint32_t index) -> size_t { return size_t(index * Lu); };
    const int rows = c.rows_b;
    const uint32_t* __restrict__ row_start = c.row_start32;
    const uint32_t* __restrict__ col = c.col32;
    T* post = a.scratch + slot;                      // post[p * L], p < n
    T* c2v = a.scratch + size_t(c.n) * L + slot;     // c2v[e * L], e < edges
    T* sink = c2v + at(uint32_t(c.edges_b));           // where the lanes past a check's edges store
    // a lane reads what another lane of its wavefront wrote: in order within a wavefront
    auto sync = []() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); };
    // Every read below is unconditional (a lane past its check's edges reads entry 0 and
    // drops the value), so that no branch separates a read from the next one and the reads
    // of several checks go out together. The check index is made wavefront-uniform and the
    // row starts read through the constant address space: scalar reads.
    using ConstU32 = const __attribute__((address_space(4))) uint32_t;
    ConstU32* row_start_s = (ConstU32*)row_start;  // the table is not written while the kernel runs
    auto row = [&](int i, uint32_t& e0, int& deg) {
        const int iu = __builtin_amdgcn_readfirstlane(i);
        const bool in_range = uint32_t(iu) < uint32_t(rows);  // checks before 0 and past the last: none
        const int ic = in_range ? iu : 0;
        e0 = row_start_s[ic];
        const uint32_t e1 = row_start_s[ic + 1];
        deg = in_range ? int(e1 - e0) : 0;
    };
    // ldpc.rs:555-557, the same answer in every lane of the group: 32 checks at a time, their
    // columns read together, then their posteriors, the parities packed one bit per check
    // (the XOR over the group reduces all 32 at once)
    auto fails = [&]() -> bool {
        for (int i0 = 0; i0 < rows; i0 += 32) {
            uint32_t e0[32], v[32];
            int deg[32];
#pragma unroll
            for (int j = 0; j < 32; ++j) row(i0 + j, e0[j], deg[j]);
#pragma unroll
            for (int j = 0; j < 32; ++j) v[j] = col[q < deg[j] ? e0[j] + q : 0u] & 0xffffu;
            uint32_t bits = 0;
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                const T pv = post[at(q < deg[j] ? v[j] : 0u)];
                bits |= q < deg[j] && pv < T(0) ? 1u << j : 0u;
            }
            if (group_all<G>(bits, [](uint32_t x, uint32_t y) { return x ^ y; })) return true;
        }
        return false;
    };
    auto finish = [&](int64_t w, int it) {
        if (a.iterations && q == 0) a.iterations[w] = uint32_t(it);
        if (a.bits)
            for (int i = q; i < c.k_info; i += G)
                a.bits[w * c.k_info + i] = post[size_t(c.info ? int(c.info[i]) : i) * L] < T(0);
        if (a.post_out)
            for (int p = q; p < c.n; p += G) a.post_out[w * c.n + p] = post[size_t(p) * L];
    };
    int64_t w = slot;
    int it = -1;  // sweeps done on codeword w; -1: not loaded (the same in every lane of the group)
    // A check reads the posteriors that the check before it has just written (always so in
    // a staircase), so a sweep that read them only then would wait a memory round trip per
    // check. They are read two checks ahead instead, before the two checks in between store
    // theirs; each of those then hands what it wrote to the lanes of the checks in flight
    // whose column it is, the newer over the older. The column table says, per edge, which
    // lane of each of the two checks before holds the same column (columns are distinct
    // within a check), so a hand-over is one lane permutation.
    const int group_lane0 = int(threadIdx.x) & 63 & ~(G - 1);
    auto head = [&](Pending<T>& s, int i) {
        row(i, s.e0, s.deg);
        const uint32_t cv = col[q < s.deg ? s.e0 + q : 0u];
        s.v = q < s.deg ? cv & 0xffffu : kNoColumn;
        s.from = q < s.deg ? cv >> 16 : 0u;
    };
    // the posterior and the message of the previous sweep (the latter unused on the first sweep)
    auto fetch = [&](Pending<T>& s) {
        s.fresh = 0;
        s.fv = T(0);
        s.p = post[at(q < s.deg ? s.v : 0u)];
        s.old = c2v[at(q < s.deg ? s.e0 + q : 0u)];
    };
    // check i: cur; i + 1 and i + 2 in flight (n1, n2: the reads of i + 2 go out here);
    // n3 takes the edges of i + 3
    auto step = [&](int i, Pending<T>& cur, Pending<T>& n1, Pending<T>& n2, Pending<T>& n3) {
        head(n3, i + 3);
        fetch(n2);
        const int deg = cur.deg;
        const T p = cur.fresh ? cur.fv : cur.p;
        const T in = q < deg ? p - (it > 0 ? cur.old : T(0)) : T(0);
        const T out = group_check<G, T>(a.rule, T(a.alpha), in, q, deg);
        const T np = in + out;
        // (stored by every lane, so that the wait for a later read counts the same stores on
        // every path)
        *(q < deg ? post + at(cur.v) : sink) = np;
        *(q < deg ? c2v + at(cur.e0 + q) : sink) = out;
        // to check i + 1 from the lane its entry names for the check before, to check i + 2
        // from the one named for the check two before
        const T x1 = lane_value(np, group_lane0 + int(n1.from & 31u));
        const T x2 = lane_value(np, group_lane0 + int((n2.from >> 6) & 31u));
        const bool h1 = (n1.from & 32u) != 0, h2 = (n2.from & 2048u) != 0;
        n1.fresh = h1 ? 1u : n1.fresh;
        n1.fv = h1 ? x1 : n1.fv;
        n2.fresh = h2 ? 1u : n2.fresh;
        n2.fv = h2 ? x2 : n2.fv;
        // kept apart from the reads in flight: folded into them, the selects would wait for them
        asm volatile("" : "+v"(n1.fresh), "+v"(n1.fv), "+v"(n2.fresh), "+v"(n2.fv));
        sync();
    };
    Pending<T> s0, s1, s2, s3;
    while (w < a.count) {
        if (it < 0) {
            for (int p = q; p < c.n; p += G) {
                T v;
                if (a.input_full) {
                    v = a.llr[w * c.n + p];
                } else {
                    v = T(0);
                    for (uint32_t s = c.src_start[p]; s < c.src_start[p + 1]; ++s) v += a.llr[w * c.e + c.src[s]];
                    if (p >= c.k_info && p < c.k) v = T(30);
                }
                post[size_t(p) * L] = v;
            }
            sync();
            if (a.apriori) {
                for (int i = q; i < c.k_info; i += G) {
                    const int p = c.info ? int(c.info[i]) : i;
                    post[size_t(p) * L] += a.apriori[w * c.k_info + i];
                }
                sync();
            }
            it = 0;
            if (a.max_iterations == 0 || !fails()) {
                finish(w, 0);
                sync();
                w += slots;
                it = -1;
                continue;
            }
        }
        // one sweep, ldpc.rs:522-531 for checks 0, 1, ..., m - 1 (the lanes past a check's
        // edges hold the column kNoColumn, which is no variable: what they take is unused)
        // The sweep starts 4 steps early, on checks -4 .. -1 that have no edges, so that its
        // first reads go out inside the loop as every later read does (the compiler's count of
        // the reads in flight is then the same on entry as around the loop), and ends on up to
        // 3 such steps past the last check; taking them keeps every slot in its registers, with
        // no copy that would wait for a read. Steps without edges read entry 0 and store to the
        // sink; no entry names them, so they hand nothing on.
        for (Pending<T>* e : {&s0, &s1, &s2, &s3}) {
            e->e0 = 0;
            e->deg = 0;
            e->v = kNoColumn;
            e->from = 0;
            e->p = e->old = e->fv = T(0);
            e->fresh = 0;
        }
        for (int i = -4; i < rows; i += 4) {
            step(i, s0, s1, s2, s3);
            step(i + 1, s1, s2, s3, s0);
            step(i + 2, s2, s3, s0, s1);
            step(i + 3, s3, s0, s1, s2);
        }
        ++it;
        if (it >= a.max_iterations || !fails()) {
            finish(w, it);
            sync();
            w += slots;
            it = -1;
        }
    }
}

// Scratch of the lane-per-codeword kernel at most (its lanes are sized to fit).
constexpr size_t kRowsScratchBytes = size_t(8) << 30;

template <typename T>
hipError_t launch_rows(const GenericArgs<T>& a0, Workspace& ws, hipStream_t stream) {
    GenericArgs<T> a = a0;
    const DevCode& c = a.code;
    const size_t per_lane = (size_t(c.n) + size_t(c.edges_b)) * sizeof(T);
    // fewer lanes (each then decodes more codewords) when the device cannot hold them all;
    // a buffer given whole: as many lanes as it holds
    const Sizing sz = size_scratch(a.count, per_lane, kRowsScratchBytes, 64, 64, ws.fixed, ws.bytes,
                                   [&](size_t bytes) { return ws.get(bytes); });
    if (sz.status == Sizing::kNothing) return hipSuccess;
    if (sz.status == Sizing::kOutOfMemory) return hipErrorOutOfMemory;
    const int64_t lanes = sz.units;
    a.scratch = static_cast<T*>(sz.scratch);
    a.scratch_stride = 0;
    // one wavefront per workgroup, so that few codewords still spread over many CUs
    const int threads = 64;
    const unsigned blocks = unsigned((lanes + threads - 1) / threads);
    hipLaunchKernelGGL((decode_rows_kernel<T>), dim3(blocks), dim3(threads), 0, stream, a, lanes);
    return hipGetLastError();
}

template <typename T, int G>
hipError_t launch_groups(const GenericArgs<T>& a0, Workspace& ws, hipStream_t stream) {
    GenericArgs<T> a = a0;
    const DevCode& c = a.code;
    // posteriors, messages and one value the lanes past a check's edges write
    const size_t per_slot = (size_t(c.n) + size_t(c.edges_b) + 1) * sizeof(T);
    // fewer groups (each then decodes more codewords) when the device cannot hold them all
    const Sizing sz = size_scratch(a.count, per_slot, kRowsScratchBytes, 1, 1, ws.fixed, ws.bytes,
                                   [&](size_t bytes) { return ws.get(bytes); });
    if (sz.status == Sizing::kNothing) return hipSuccess;
    if (sz.status == Sizing::kOutOfMemory) return hipErrorOutOfMemory;
    const int64_t slots = sz.units;
    a.scratch = static_cast<T*>(sz.scratch);
    a.scratch_stride = 0;
    const unsigned blocks = unsigned((slots * G + 63) / 64);
    hipLaunchKernelGGL((decode_groups_kernel<T, G>), dim3(blocks), dim3(64), 0, stream, a, slots);
    return hipGetLastError();
}

template <typename T>
hipError_t launch(const GenericArgs<T>& a0, int device, hipStream_t stream, Workspace& ws) {
    if (a0.layered && a0.code.z == 1) {
        // codes given by their rows: a group of lanes per codeword when a check fits in one
        // and lists no variable twice (else the edges of a check must be stored in order;
        // the group kernel also reads entry 0 of the edge tables, so it needs one edge)
        if (!a0.code.distinct || a0.code.edges_b == 0) return launch_rows(a0, ws, stream);
        if (a0.code.max_deg <= 8) return launch_groups<T, 8>(a0, ws, stream);
        if (a0.code.max_deg <= 16) return launch_groups<T, 16>(a0, ws, stream);
        if (a0.code.max_deg <= 32) return launch_groups<T, 32>(a0, ws, stream);
        return launch_rows(a0, ws, stream);
    }
    GenericArgs<T> a = a0;
    const DevCode& c = a.code;
    int threads = ((c.z + 63) / 64) * 64;
    if (threads > 512) threads = 512;
    const size_t post_bytes = size_t(c.n) * sizeof(T);
    const bool post_lds = post_bytes <= kLdsBudget;
    hipDeviceProp_t prop;
    hipError_t err = hipGetDeviceProperties(&prop, device);
    if (err != hipSuccess) return err;
    int per_cu = 0;
    if (post_lds)
        err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decode_generic_kernel<T, true>, threads,
                                                           post_bytes);
    else
        err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decode_generic_kernel<T, false>, threads, 0);
    if (err != hipSuccess) return err;
    if (per_cu < 1) per_cu = 1;
    int64_t grid = int64_t(per_cu) * prop.multiProcessorCount;
    if (grid > a.count) grid = a.count;
    if (grid < 1) return hipSuccess;
    a.scratch_stride = size_t(c.edges_b) * c.z + (post_lds ? 0 : c.n) + (a.layered ? 0 : c.n);
    a.scratch_stride = (a.scratch_stride + 63) & ~size_t(63);
    if (ws.fixed) {
        // a buffer given whole: as many workgroups as it holds
        const size_t per = a.scratch_stride * sizeof(T);
        if (ws.bytes < per) return hipErrorOutOfMemory;
        grid = std::min<int64_t>(grid, int64_t(ws.bytes / per));
    }
    T* scratch = static_cast<T*>(ws.get(size_t(grid) * a.scratch_stride * sizeof(T)));
    if (!scratch) return hipErrorOutOfMemory;
    a.scratch = scratch;
    if (post_lds)
        hipLaunchKernelGGL((decode_generic_kernel<T, true>), dim3(unsigned(grid)), dim3(threads), post_bytes, stream,
                           a);
    else
        hipLaunchKernelGGL((decode_generic_kernel<T, false>), dim3(unsigned(grid)), dim3(threads), 0, stream, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_generic(const GenericArgs<float>& a, int device, hipStream_t s, Workspace& ws) {
    return launch<float>(a, device, s, ws);
}
hipError_t launch_generic(const GenericArgs<double>& a, int device, hipStream_t s, Workspace& ws) {
    return launch<double>(a, device, s, ws);
}

}  // namespace comfec
