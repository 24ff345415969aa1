// Construction of the lifted NR LDPC codes (crates/com-fec/src/nr/ldpc.rs:101-162 and
// crates/com-fec/src/ldpc.rs:270-287), their encoder and their rate matching.
#include "nr_code.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "nr_base_graphs.inc"

namespace comfec {

void fail(Kind k, const std::string& msg) {
    const char* prefix = "";
    switch (k) {
        case Kind::Config: prefix = "invalid configuration: "; break;
        case Kind::Shape: prefix = "inconsistent shapes: "; break;
        case Kind::Numerical: prefix = "numerical failure: "; break;
        case Kind::Condition: prefix = "condition: "; break;
        case Kind::Io: prefix = "input/output: "; break;
        case Kind::Device: prefix = "device: "; break;
    }
    throw Error(k, prefix + msg);
}

void expect_len(size_t got, size_t want, const char* what) {
    if (got != want)
        fail(Kind::Shape, std::to_string(got) + " " + what + " for a block of " + std::to_string(want));
}

namespace {

int lifting_set(uint32_t z) {
    for (int s = 0; s < 8; ++s)
        for (int q = 0; q < NR_LIFTING_SET_MAX; ++q)
            if (NR_LIFTING_SETS[s][q] == z) return s;
    return -1;
}

// Gauss-Jordan over GF(2) on [B | I], B the core block (core checks x core parity bits);
// the right half ends as B^-1: row j lists the checks whose syndromes give parity bit k+j.
// The inverse is unique, so this is the matrix of ldpc.rs:568-584.
void invert_core(NrCode& c) {
    const uint32_t core = 4 * c.z;
    const uint32_t words = (2 * core + 63) / 64;
    std::vector<uint64_t> a(size_t(core) * words, 0);
    auto flip = [&](uint32_t r, uint32_t col) { a[size_t(r) * words + col / 64] ^= 1ull << (col % 64); };
    auto bit = [&](uint32_t r, uint32_t col) { return (a[size_t(r) * words + col / 64] >> (col % 64)) & 1; };
    // the core rows are the first 4 base rows
    for (uint32_t i = 0; i < 4; ++i)
        for (uint32_t eb = c.row_start[i]; eb < c.row_start[i + 1]; ++eb) {
            uint32_t j = c.col[eb];
            if (j * c.z < c.k) continue;
            for (uint32_t r = 0; r < c.z; ++r) flip(i * c.z + r, (j * c.z + (r + c.shift[eb]) % c.z) - c.k);
        }
    for (uint32_t r = 0; r < core; ++r) flip(r, core + r);
    for (uint32_t colx = 0; colx < core; ++colx) {
        uint32_t piv = core;
        for (uint32_t r = colx; r < core; ++r)
            if (bit(r, colx)) { piv = r; break; }
        if (piv == core) fail(Kind::Config, "the core block is singular");
        if (piv != colx)
            std::swap_ranges(a.begin() + size_t(piv) * words, a.begin() + size_t(piv + 1) * words,
                             a.begin() + size_t(colx) * words);
        for (uint32_t r = 0; r < core; ++r)
            if (r != colx && bit(r, colx))
                for (uint32_t w = 0; w < words; ++w) a[size_t(r) * words + w] ^= a[size_t(colx) * words + w];
    }
    c.core_words = (core + 63) / 64;
    c.core_inv.assign(size_t(core) * c.core_words, 0);
    for (uint32_t j = 0; j < core; ++j)
        for (uint32_t r = 0; r < core; ++r)
            if (bit(j, core + r)) c.core_inv[size_t(j) * c.core_words + r / 64] |= 1ull << (r % 64);
}

}  // namespace

NrCode::NrCode(NrCode&& o) noexcept
    : bg(o.bg), z(o.z), ils(o.ils), rows_b(o.rows_b), cols_b(o.cols_b), edges_b(o.edges_b), n(o.n), k(o.k),
      m(o.m), k_info(o.k_info), e(o.e), qm(o.qm), rv(o.rv), row_start(std::move(o.row_start)),
      col(std::move(o.col)), shift(std::move(o.shift)), col_start(std::move(o.col_start)),
      col_edge(std::move(o.col_edge)), map(std::move(o.map)), src_start(std::move(o.src_start)),
      src(std::move(o.src)), repeats(o.repeats), core_inv(std::move(o.core_inv)), core_words(o.core_words), info(std::move(o.info)),
      sum_col(std::move(o.sum_col)), sum_start(std::move(o.sum_start)), sum_src(std::move(o.sum_src)), dec(o.dec),
      dev_(std::move(o.dev_)) {
    o.dev_.clear();
}

NrCode::~NrCode() {
    for (auto& t : dev_)
        if (t.base) (void)hipFree(t.base);
}

NrCode NrCode::lifted(int bg, uint32_t z) {
    int s = lifting_set(z);
    if ((bg != 1 && bg != 2) || s < 0)
        fail(Kind::Config, "base graph " + std::to_string(bg) + " at Z = " + std::to_string(z) +
                               " is not an LDPC code of NR");
    NrCode c;
    c.bg = bg;
    c.z = z;
    c.ils = uint32_t(s);
    const nr_bg_entry* t = bg == 1 ? NR_BG1 : NR_BG2;
    c.edges_b = bg == 1 ? NR_BG1_ENTRIES : NR_BG2_ENTRIES;
    c.rows_b = bg == 1 ? 46 : 42;
    c.cols_b = bg == 1 ? 68 : 52;
    c.n = c.cols_b * z;
    c.m = c.rows_b * z;
    c.k = (bg == 1 ? 22 : 10) * z;
    c.row_start.assign(c.rows_b + 1, 0);
    for (uint32_t eb = 0; eb < c.edges_b; ++eb) {
        c.row_start[t[eb].i + 1]++;
        c.col.push_back(t[eb].j);
        c.shift.push_back(uint16_t(t[eb].v[s] % z));
    }
    for (uint32_t i = 0; i < c.rows_b; ++i) c.row_start[i + 1] += c.row_start[i];
    // by column, rows ascending
    c.col_start.assign(c.cols_b + 1, 0);
    for (uint32_t eb = 0; eb < c.edges_b; ++eb) c.col_start[c.col[eb] + 1]++;
    for (uint32_t j = 0; j < c.cols_b; ++j) c.col_start[j + 1] += c.col_start[j];
    std::vector<uint16_t> fill(c.col_start.begin(), c.col_start.end() - 1);
    c.col_edge.assign(c.edges_b, 0);
    for (uint32_t eb = 0; eb < c.edges_b; ++eb) c.col_edge[fill[c.col[eb]]++] = uint16_t(eb);
    invert_core(c);
    c.k_info = c.k;
    c.e = c.n - 2 * z;
    c.map.resize(c.e);
    for (uint32_t t2 = 0; t2 < c.e; ++t2) c.map[t2] = 2 * z + t2;
    c.finish_tables();
    return c;
}

NrCode NrCode::from_rows(uint32_t n, uint32_t m, const uint32_t* row_start, const uint32_t* cols) {
    if (!row_start || (!cols && m > 0)) fail(Kind::Config, "null rows");
    if (n < 1 || m < 1) fail(Kind::Config, std::to_string(m) + " checks for " + std::to_string(n) + " bits");
    if (row_start[0] != 0) fail(Kind::Config, "the row starts begin at 0");
    for (uint32_t r = 0; r < m; ++r)
        if (row_start[r + 1] < row_start[r]) fail(Kind::Config, "the row starts must not decrease");
    const uint32_t edges = row_start[m];
    for (uint32_t e = 0; e < edges; ++e)
        if (cols[e] >= n) fail(Kind::Config, "a check refers to a missing variable");
    // the device tables hold 16-bit indices; the generic kernel takes rows of up to 19 edges
    if (n > 65535 || edges > 65535 || m > 65535)
        fail(Kind::Config, "codes given by rows are limited to 65535 bits, checks and edges on the device");
    NrCode c;
    c.bg = 0;
    c.z = 1;
    c.ils = 0;
    c.rows_b = m;
    c.cols_b = n;
    c.edges_b = edges;
    c.n = n;
    c.m = m;
    c.row_start.resize(m + 1);
    for (uint32_t r = 0; r <= m; ++r) c.row_start[r] = uint16_t(row_start[r]);
    c.col.resize(edges);
    for (uint32_t e = 0; e < edges; ++e) c.col[e] = uint16_t(cols[e]);
    c.shift.assign(edges, 0);
    if (c.max_row_degree() > 19) fail(Kind::Config, "the device decoder takes checks of at most 19 variables");
    c.col_start.assign(n + 1, 0);
    for (uint32_t e = 0; e < edges; ++e) c.col_start[c.col[e] + 1]++;
    for (uint32_t j = 0; j < n; ++j) c.col_start[j + 1] += c.col_start[j];
    std::vector<uint16_t> fill(c.col_start.begin(), c.col_start.end() - 1);
    c.col_edge.assign(edges, 0);
    for (uint32_t e = 0; e < edges; ++e) c.col_edge[fill[c.col[e]]++] = uint16_t(e);
    // gf2::Matrix::reduce (crates/com-math/src/gf2.rs:124-150): the reduced row echelon form
    // with the pivots taken from the last column to the first (unique for that order)
    const uint32_t words = (n + 63) / 64;
    std::vector<uint64_t> h(size_t(m) * words, 0);
    for (uint32_t r = 0; r < m; ++r)
        for (uint32_t e = row_start[r]; e < row_start[r + 1]; ++e) h[size_t(r) * words + cols[e] / 64] ^= 1ull << (cols[e] % 64);
    auto bit = [&](uint32_t r, uint32_t col) { return (h[size_t(r) * words + col / 64] >> (col % 64)) & 1; };
    std::vector<uint32_t> pivots;
    uint32_t rank = 0;
    for (uint32_t colx = n; colx-- > 0 && rank < m;) {
        uint32_t p = rank;
        while (p < m && !bit(p, colx)) ++p;
        if (p == m) continue;
        if (p != rank)
            std::swap_ranges(h.begin() + size_t(p) * words, h.begin() + size_t(p + 1) * words,
                             h.begin() + size_t(rank) * words);
        // the rows not yet taken as pivots are zero past colx (every column after it is a pivot,
        // cleared from them, or was zero in all of them), so the pivot row ends at colx's word
        const uint32_t used = colx / 64 + 1;
        for (uint32_t r = 0; r < m; ++r)
            if (r != rank && bit(r, colx))
                for (uint32_t w = 0; w < used; ++w) h[size_t(r) * words + w] ^= h[size_t(rank) * words + w];
        pivots.push_back(colx);
        ++rank;
    }
    std::vector<uint8_t> is_pivot(n, 0);
    for (uint32_t p : pivots) is_pivot[p] = 1;
    for (uint32_t j = 0; j < n; ++j)
        if (!is_pivot[j]) c.info.push_back(j);
    if (c.info.empty()) fail(Kind::Config, "the code has no information bits");
    c.sum_start.push_back(0);
    for (uint32_t r = 0; r < rank; ++r) {
        c.sum_col.push_back(pivots[r]);
        for (uint32_t v : c.info)
            if (bit(r, v)) c.sum_src.push_back(v);
        c.sum_start.push_back(uint32_t(c.sum_src.size()));
    }
    c.k_info = uint32_t(c.info.size());
    c.k = c.k_info;  // no fillers: the kernels fix positions [K', K) to 30.0

    c.e = n;
    c.map.resize(n);
    for (uint32_t t = 0; t < n; ++t) c.map[t] = t;
    c.finish_tables();
    return c;
}

NrCode NrCode::with_options(uint32_t k_info, uint32_t e, int bg, int rv, uint32_t qm) {
    if (k_info < 1 || e < 1)
        fail(Kind::Config, "need K' ≥ 1 and E ≥ 1, got " + std::to_string(k_info) + " and " + std::to_string(e));
    if (rv < 0 || rv > 3 || qm < 1 || e % qm != 0)
        fail(Kind::Config, "rv ≤ 3 and Qm | E, got rv " + std::to_string(rv) + ", Qm " + std::to_string(qm) +
                               ", E " + std::to_string(e));
    if (bg == 0) {
        // nr/ldpc.rs:78-81
        double r = double(k_info) / double(e);
        bg = (k_info <= 292 || (k_info <= 3824 && r <= 0.67) || r <= 0.25) ? 2 : 1;
    }
    if (bg != 1 && bg != 2) fail(Kind::Config, "the base graph is 1 or 2, got " + std::to_string(bg));
    uint32_t k_cb = bg == 1 ? 8448 : 3840;
    if (k_info > k_cb)
        fail(Kind::Config, std::to_string(k_info) + " bits exceed one code block of base graph " +
                               std::to_string(bg) + " (" + std::to_string(k_cb) + "); segment the block first");
    uint32_t kb = bg == 1 ? 22 : k_info > 640 ? 10 : k_info > 560 ? 9 : k_info > 192 ? 8 : 6;
    uint32_t z = 0;
    for (int s = 0; s < 8; ++s)
        for (int q = 0; q < NR_LIFTING_SET_MAX; ++q) {
            uint32_t zz = NR_LIFTING_SETS[s][q];
            if (zz && kb * zz >= k_info && (z == 0 || zz < z)) z = zz;
        }
    if (!z) fail(Kind::Config, "no lifting size for K' = " + std::to_string(k_info));
    NrCode c = lifted(bg, z);
    c.k_info = k_info;
    c.e = e;
    c.qm = qm;
    c.rv = uint32_t(rv);
    // circular buffer from k0, skipping fillers (nr/ldpc.rs:136-157)
    const uint32_t ncb = c.n - 2 * z;
    static const int num1[4] = {0, 17, 33, 56}, num2[4] = {0, 13, 25, 43};
    uint64_t num = bg == 1 ? num1[rv] : num2[rv];
    uint64_t den = bg == 1 ? 66 : 50;
    uint64_t k0 = num * ncb / (den * z) * z;
    std::vector<uint32_t> sel;
    sel.reserve(e);
    for (uint64_t j = 0; sel.size() < e; ++j) {
        uint32_t colx = 2 * z + uint32_t((k0 + j) % ncb);
        if (!(colx >= k_info && colx < c.k)) sel.push_back(colx);
    }
    // the bit interleaver of Qm rows (interleaving.rs:58-60 applied at nr/ldpc.rs:160)
    const uint32_t cols = e / qm;
    c.map.resize(e);
    for (uint32_t t = 0; t < e; ++t) c.map[t] = sel[(t % qm) * cols + t / qm];
    c.finish_tables();
    return c;
}

void NrCode::finish_tables() {
    src_start.assign(n + 1, 0);
    for (uint32_t t = 0; t < e; ++t) src_start[map[t] + 1]++;
    for (uint32_t p = 0; p < n; ++p) src_start[p + 1] += src_start[p];
    std::vector<uint32_t> fill(src_start.begin(), src_start.end() - 1);
    src.assign(e, 0);
    for (uint32_t t = 0; t < e; ++t) src[fill[map[t]]++] = t;
    repeats = false;
    for (uint32_t p = 0; p < n; ++p) repeats = repeats || src_start[p + 1] - src_start[p] > 1;
}

uint32_t NrCode::max_row_degree() const {
    uint32_t d = 0;
    for (uint32_t i = 0; i < rows_b; ++i) d = std::max<uint32_t>(d, row_start[i + 1] - row_start[i]);
    return d;
}

bool NrCode::rows_distinct() const {
    std::vector<uint32_t> seen(cols_b, UINT32_MAX);
    for (uint32_t i = 0; i < rows_b; ++i)
        for (uint32_t e = row_start[i]; e < row_start[i + 1]; ++e) {
            if (seen[col[e]] == i) return false;
            seen[col[e]] = i;
        }
    return true;
}

bool NrCode::is_codeword(const uint8_t* cw) const {
    for (uint32_t i = 0; i < rows_b; ++i)
        for (uint32_t r = 0; r < z; ++r) {
            uint8_t a = 0;
            for (uint32_t eb = row_start[i]; eb < row_start[i + 1]; ++eb) a ^= cw[col[eb] * z + (r + shift[eb]) % z];
            if (a) return false;
        }
    return true;
}

void NrCode::encode_full(const uint8_t* u, uint8_t* cw) const {
    std::memset(cw, 0, n);
    for (uint32_t i = 0; i < k; ++i) cw[i] = u[i] & 1;
    const uint32_t core = 4 * z;
    // syndromes of the information bits on the core checks
    std::vector<uint64_t> syn(core_words, 0);
    for (uint32_t i = 0; i < 4; ++i)
        for (uint32_t r = 0; r < z; ++r) {
            uint8_t a = 0;
            for (uint32_t eb = row_start[i]; eb < row_start[i + 1]; ++eb) {
                uint32_t v = col[eb] * z + (r + shift[eb]) % z;
                if (v < k) a ^= cw[v];
            }
            if (a) syn[(i * z + r) / 64] |= 1ull << ((i * z + r) % 64);
        }
    for (uint32_t j = 0; j < core; ++j) {
        uint64_t acc = 0;
        for (uint32_t w = 0; w < core_words; ++w) acc ^= core_inv[size_t(j) * core_words + w] & syn[w];
        cw[k + j] = uint8_t(__builtin_popcountll(acc) & 1);
    }
    // extension checks close on their own parity bit, the last of their row
    for (uint32_t i = 4; i < rows_b; ++i)
        for (uint32_t r = 0; r < z; ++r) {
            uint32_t own = k + i * z + r;
            uint8_t a = 0;
            for (uint32_t eb = row_start[i]; eb < row_start[i + 1]; ++eb) {
                uint32_t v = col[eb] * z + (r + shift[eb]) % z;
                if (v != own) a ^= cw[v];
            }
            cw[own] = a;
        }
}

void NrCode::encode(const uint8_t* bits, uint8_t* out) const {
    std::vector<uint8_t> cw(n, 0);
    if (bg == 0) {
        // Encoder::Sums (ldpc.rs:760-764): the information bits in place, every pivot bit
        // the sum of the information bits of its reduced row
        for (uint32_t i = 0; i < k_info; ++i) cw[info[i]] = bits[i] & 1;
        for (size_t r = 0; r + 1 < sum_start.size(); ++r) {
            uint8_t a = 0;
            for (uint32_t s = sum_start[r]; s < sum_start[r + 1]; ++s) a ^= cw[sum_src[s]];
            cw[sum_col[r]] = a;
        }
    } else {
        std::vector<uint8_t> u(k, 0);
        std::memcpy(u.data(), bits, k_info);
        encode_full(u.data(), cw.data());
    }
    for (uint32_t t = 0; t < e; ++t) out[t] = cw[map[t]];
}

bool NrCode::tables_resident(int device) const {
    std::lock_guard<std::mutex> lock(dev_mu_);
    return device >= 0 && size_t(device) < dev_.size() && dev_[device].base != nullptr;
}

const NrCode::DeviceTables& NrCode::device_tables(int device) const {
    std::lock_guard<std::mutex> lock(dev_mu_);
    if (dev_.size() <= size_t(device)) dev_.resize(device + 1);
    DeviceTables& t = dev_[device];
    if (t.base) return t;
    // one buffer: u16 tables, then u32 tables from a 16-byte boundary, the map first (read as
    // 16-byte vectors)
    size_t n16 = row_start.size() + col.size() + shift.size() + col_start.size() + col_edge.size();
    n16 = (n16 + 7) & ~size_t(7);
    size_t n32 = src_start.size() + src.size() + map.size() + shift.size() + info.size() + row_start.size() +
                 col.size();
    n32 = (n32 + 1) & ~size_t(1);
    size_t bytes = n16 * 2 + n32 * 4 + core_inv.size() * 8;
    std::vector<uint8_t> host(bytes, 0);
    uint16_t* h16 = reinterpret_cast<uint16_t*>(host.data());
    size_t o = 0;
    auto put16 = [&](const std::vector<uint16_t>& v) {
        size_t at = o;
        std::memcpy(h16 + o, v.data(), v.size() * 2);
        o += v.size();
        return at;
    };
    size_t o_rs = put16(row_start), o_col = put16(col), o_sh = put16(shift), o_cs = put16(col_start),
           o_ce = put16(col_edge);
    uint32_t* h32 = reinterpret_cast<uint32_t*>(host.data() + n16 * 2);
    std::memcpy(h32, map.data(), map.size() * 4);
    std::memcpy(h32 + map.size(), src_start.data(), src_start.size() * 4);
    std::memcpy(h32 + map.size() + src_start.size(), src.data(), src.size() * 4);
    uint32_t* h_shift = h32 + src_start.size() + src.size() + map.size();
    for (size_t i = 0; i < shift.size(); ++i) h_shift[i] = shift[i];
    std::memcpy(h_shift + shift.size(), info.data(), info.size() * 4);
    uint32_t* h_rs = h_shift + shift.size() + info.size();
    for (size_t i = 0; i < row_start.size(); ++i) h_rs[i] = row_start[i];
    // the columns, each with where the two checks before list the same variable: bits 16..20
    // the edge's position in the check before and bit 21 set when it is there, bits 22..26
    // and 27 the same for the check two before (the last position of a variable listed twice)
    uint32_t* h_col = h_rs + row_start.size();
    std::vector<int> pos1(cols_b, -1), pos2(cols_b, -1);  // positions in checks i - 1 and i - 2
    auto mark = [&](std::vector<int>& pos, uint32_t r, bool set) {
        for (uint32_t e = row_start[r]; e < row_start[r + 1]; ++e) pos[col[e]] = set ? int(e - row_start[r]) : -1;
    };
    for (uint32_t i = 0; i < rows_b; ++i) {
        for (uint32_t e = row_start[i]; e < row_start[i + 1]; ++e) {
            uint32_t w = col[e];
            if (pos1[col[e]] >= 0) w |= (uint32_t(pos1[col[e]]) | 32u) << 16;
            if (pos2[col[e]] >= 0) w |= (uint32_t(pos2[col[e]]) | 32u) << 22;
            h_col[e] = w;
        }
        if (i >= 2) mark(pos2, i - 2, false);
        if (i >= 1) {
            mark(pos2, i - 1, true);
            mark(pos1, i - 1, false);
        }
        mark(pos1, i, true);
    }
    uint64_t* h64 = reinterpret_cast<uint64_t*>(host.data() + n16 * 2 + n32 * 4);
    std::memcpy(h64, core_inv.data(), core_inv.size() * 8);
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) fail(Kind::Device, "hipGetDevice failed");
    if (prev != device && hipSetDevice(device) != hipSuccess) fail(Kind::Device, "hipSetDevice failed");
    void* base = nullptr;
    hipError_t err = hipMalloc(&base, bytes);
    if (err == hipSuccess) err = hipMemcpy(base, host.data(), bytes, hipMemcpyHostToDevice);
    if (prev != device) (void)hipSetDevice(prev);
    if (err != hipSuccess) {
        if (base) (void)hipFree(base);
        fail(Kind::Device, std::string("code tables: ") + hipGetErrorString(err));
    }
    const uint16_t* d16 = static_cast<const uint16_t*>(base);
    const uint32_t* d32 = reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(base) + n16 * 2);
    t.base = base;
    t.row_start = d16 + o_rs;
    t.col = d16 + o_col;
    t.shift = d16 + o_sh;
    t.col_start = d16 + o_cs;
    t.col_edge = d16 + o_ce;
    t.map = d32;
    t.src_start = d32 + map.size();
    t.src = d32 + map.size() + src_start.size();
    t.shift32 = d32 + src_start.size() + src.size() + map.size();
    t.info = info.empty() ? nullptr : t.shift32 + shift.size();
    t.row_start32 = t.shift32 + shift.size() + info.size();
    t.col32 = t.row_start32 + row_start.size();
    t.core_inv = reinterpret_cast<const uint64_t*>(static_cast<const uint8_t*>(base) + n16 * 2 + n32 * 4);
    return t;
}

}  // namespace comfec
