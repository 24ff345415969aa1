// Host side of the NR LDPC code: construction of the lifted code, its encoder and its
// rate matching, and the compact quasi-cyclic description the device kernels walk.
#pragma once
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace comfec {

// The kinds of com_core::Error (crates/com-core/src/error.rs:25-41) plus a device failure.
enum class Kind { Config = 1, Shape = 2, Numerical = 3, Condition = 4, Io = 5, Device = 6 };

struct Error : std::runtime_error {
    Kind kind;
    Error(Kind k, const std::string& m) : std::runtime_error(m), kind(k) {}
};

// "invalid configuration: ..." and friends, the Display of com_core::Error.
[[noreturn]] void fail(Kind k, const std::string& msg);
// com_core::expect_len (error.rs:91-93): "{got} {what} for a block of {want}".
void expect_len(size_t got, size_t want, const char* what);

// LdpcDecoder (crates/com-fec/src/ldpc.rs:164-179); defaults of ldpc.rs:77-88.
struct DecoderSettings {
    int rule = 0;          // 0 sum-product, 1 normalized min-sum, 2 offset min-sum
    double alpha = 0.8;    // DEFAULT_ALPHA
    bool layered = true;   // DEFAULT_LAYERED
    uint32_t max_iterations = 50;
};

// One lifted base graph of NR, with everything the device needs to decode it.
struct NrCode {
    int bg = 0;                 // 1 or 2
    uint32_t z = 0, ils = 0;    // lifting size and its set
    uint32_t rows_b = 0;        // base rows (46 or 42): the layers
    uint32_t cols_b = 0;        // base columns (68 or 52)
    uint32_t edges_b = 0;       // nonzero base entries (316 or 197)
    uint32_t n = 0, k = 0, m = 0;
    uint32_t k_info = 0, e = 0, qm = 1, rv = 0;

    // base entries in layer order (row-major, columns ascending within a row)
    std::vector<uint16_t> row_start;  // rows_b + 1
    std::vector<uint16_t> col;        // edges_b
    std::vector<uint16_t> shift;      // edges_b: V mod Z
    // the same entries by column, rows ascending (the flooding variable pass)
    std::vector<uint16_t> col_start;  // cols_b + 1
    std::vector<uint16_t> col_edge;   // edges_b: index into the layer order
    // rate recovery: output bit t carries codeword bit map[t]; src lists, for every
    // codeword bit, the output bits that carry it in increasing t
    std::vector<uint32_t> map;        // e
    std::vector<uint32_t> src_start;  // n + 1
    std::vector<uint32_t> src;        // e
    bool repeats = false;             // some codeword bit is carried by more than one output bit
    // encoder: row j of the inverse of the core block, as a bit mask over the 4Z checks
    std::vector<uint64_t> core_inv;   // 4Z rows x words
    uint32_t core_words = 0;
    // a code given by its rows (bg 0): the information positions (the columns that are not
    // pivots of H's reduced form) and, per pivot column, the information bits it sums
    std::vector<uint32_t> info;
    std::vector<uint32_t> sum_col, sum_start, sum_src;

    DecoderSettings dec;

    static NrCode lifted(int bg, uint32_t z);
    // LdpcCode::from_rows (crates/com-fec/src/ldpc.rs:216-231): m checks listing their
    // variables, row r = cols[row_start[r] .. row_start[r+1]); sent whole (E = N)
    static NrCode from_rows(uint32_t n, uint32_t m, const uint32_t* row_start, const uint32_t* cols);
    static NrCode with_options(uint32_t k_info, uint32_t e, int bg, int rv, uint32_t qm);

    uint32_t max_row_degree() const;
    // no check lists a variable twice (from_rows accepts such rows, as the reference does)
    bool rows_distinct() const;
    // H c = 0 on the lifted matrix
    bool is_codeword(const uint8_t* cw) const;
    // LdpcCode::encode_block (ldpc.rs:753-791) on K systematic bits -> N bits
    void encode_full(const uint8_t* u, uint8_t* cw) const;
    // NrLdpc::encode_block (nr/ldpc.rs:208-214): K' bits -> E bits
    void encode(const uint8_t* info, uint8_t* out) const;

    // device copy of the tables, per HIP device
    struct DeviceTables {
        void* base = nullptr;
        const uint16_t* row_start = nullptr;
        const uint16_t* col = nullptr;
        const uint16_t* shift = nullptr;
        const uint16_t* col_start = nullptr;
        const uint16_t* col_edge = nullptr;
        const uint32_t* src_start = nullptr;
        const uint32_t* src = nullptr;
        const uint32_t* map = nullptr;
        const uint64_t* core_inv = nullptr;
        const uint32_t* shift32 = nullptr;  // the shifts as 32-bit words (scalar loads)
        const uint32_t* info = nullptr;     // K' information positions (codes given by rows), or null: 0..K'-1
        const uint32_t* row_start32 = nullptr;  // row_start and col as 32-bit words (scalar loads)
        const uint32_t* col32 = nullptr;
    };
    const DeviceTables& device_tables(int device) const;
    // whether the tables are already on `device` (then device_tables allocates nothing)
    bool tables_resident(int device) const;
    ~NrCode();
    NrCode() = default;
    NrCode(NrCode&&) noexcept;
    NrCode& operator=(NrCode&&) = delete;
    NrCode(const NrCode&) = delete;

  private:
    mutable std::mutex dev_mu_;
    mutable std::vector<DeviceTables> dev_;
    void finish_tables();
};

}  // namespace comfec
