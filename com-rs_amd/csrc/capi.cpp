// The C ABI of include/comfec_gpu.h: handles, error reporting, and the decode entry points.
#include "../../include/comfec_gpu.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"
#include "nr_code.hpp"

using namespace comfec;

namespace comfec {
namespace detail {

thread_local std::string g_error;

cf_status status_of(Kind k) { return static_cast<cf_status>(static_cast<int>(k)); }

void check_hip(hipError_t err, const char* what) {
    if (err != hipSuccess) fail(Kind::Device, std::string(what) + ": " + hipGetErrorString(err));
}

std::mutex g_ws_mu;
std::map<std::pair<std::pair<int, int>, hipStream_t>, std::unique_ptr<WorkspaceSlot>> g_ws;

WorkspaceSlot& workspace(int device, hipStream_t s, int purpose) {
    std::lock_guard<std::mutex> l(g_ws_mu);
    auto& slot = g_ws[{{device, purpose}, s}];
    if (!slot) slot = std::make_unique<WorkspaceSlot>();
    return *slot;
}

int current_device() {
    int d = 0;
    check_hip(hipGetDevice(&d), "hipGetDevice");
    return d;
}

DevCode dev_code(const NrCode& c, int device) {
    const NrCode::DeviceTables& t = c.device_tables(device);
    DevCode d;
    d.z = int32_t(c.z);
    d.rows_b = int32_t(c.rows_b);
    d.cols_b = int32_t(c.cols_b);
    d.edges_b = int32_t(c.edges_b);
    d.n = int32_t(c.n);
    d.k = int32_t(c.k);
    d.k_info = int32_t(c.k_info);
    d.e = int32_t(c.e);
    d.row_start = t.row_start;
    d.col = t.col;
    d.shift = t.shift;
    d.col_start = t.col_start;
    d.col_edge = t.col_edge;
    d.src_start = t.src_start;
    d.src = t.src;
    d.info = t.info;
    d.row_start32 = t.row_start32;
    d.col32 = t.col32;
    d.max_deg = int32_t(c.max_row_degree());
    d.distinct = c.rows_distinct() ? 1 : 0;
    return d;
}

template <typename T>
GenericArgs<T> generic_args(const cf_code* h, int device) {
    DecoderSettings s = h->settings();
    GenericArgs<T> a{};
    a.code = dev_code(h->code, device);
    a.rule = s.rule;
    a.alpha = s.alpha;
    a.layered = s.layered ? 1 : 0;
    a.max_iterations = int(s.max_iterations);
    return a;
}

template <typename T>
void run_generic(const GenericArgs<T>& a, int device, hipStream_t stream) {
    WorkspaceSlot& slot = workspace(device, stream, 0);
    std::lock_guard<std::mutex> l(slot.mu);
    check_hip(launch_generic(a, device, stream, slot.ws), "decode launch");
}

// The layered min-sum rules on a supported lifting size go to the register-resident kernel.
bool fast_layered(const cf_code* h) {
    DecoderSettings s = h->settings();
    return s.layered && (s.rule == 1 || s.rule == 2) && layered_supported(h->code.bg, int(h->code.z));
}

// What one host-buffer call needs on a device: its own stream (so that calls from several
// threads run side by side and wait only for their own work) and a grow-only buffer for the
// codeword's copies. Contexts are kept in a pool for the life of the process and handed to
// one call at a time, so that after the first calls no call allocates device memory.
struct HostCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    Workspace io;
};

std::mutex g_ctx_mu;
std::vector<HostCtx*> g_ctx_free;

class HostCtxLease {
public:
    explicit HostCtxLease(int device) {
        {
            std::lock_guard<std::mutex> l(g_ctx_mu);
            for (size_t i = 0; i < g_ctx_free.size(); ++i)
                if (g_ctx_free[i]->device == device) {
                    ctx_ = g_ctx_free[i];
                    g_ctx_free.erase(g_ctx_free.begin() + long(i));
                    break;
                }
        }
        if (!ctx_) {
            auto c = std::make_unique<HostCtx>();
            c->device = device;
            check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
            ctx_ = c.release();
        }
    }
    ~HostCtxLease() {
        std::lock_guard<std::mutex> l(g_ctx_mu);
        g_ctx_free.push_back(ctx_);
    }
    HostCtx& operator*() const { return *ctx_; }
    HostCtx* operator->() const { return ctx_; }

private:
    HostCtx* ctx_ = nullptr;
};

// One codeword from host f64 buffers, decoded on the GPU in f64.
void decode_host_f64(const cf_code* h, const double* llr, bool full, const double* apriori, double* post_host,
                     uint8_t* bits_host, uint32_t* iterations) {
    const NrCode& c = h->code;
    int device = current_device();
    HostCtxLease ctx(device);
    hipStream_t stream = ctx->stream;
    size_t in_len = full ? c.n : c.e;
    size_t bytes = (in_len + c.k_info + c.n) * sizeof(double) + c.k_info + sizeof(uint32_t) + 64;
    void* buf = ctx->io.get(bytes);
    if (!buf) fail(Kind::Device, "out of device memory for a codeword of " + std::to_string(in_len) + " LLRs");
    double* d_llr = static_cast<double*>(buf);
    double* d_apriori = d_llr + in_len;
    double* d_post = d_apriori + c.k_info;
    uint32_t* d_it = reinterpret_cast<uint32_t*>(d_post + c.n);
    uint8_t* d_bits = reinterpret_cast<uint8_t*>(d_it + 1);
    check_hip(hipMemcpyAsync(d_llr, llr, in_len * sizeof(double), hipMemcpyHostToDevice, stream), "copy in");
    if (apriori)
        check_hip(hipMemcpyAsync(d_apriori, apriori, c.k_info * sizeof(double), hipMemcpyHostToDevice, stream),
                  "copy in");
    GenericArgs<double> a = generic_args<double>(h, device);
    a.input_full = full ? 1 : 0;
    a.llr = d_llr;
    a.apriori = apriori ? d_apriori : nullptr;
    a.count = 1;
    a.bits = d_bits;
    a.post_out = d_post;
    a.iterations = d_it;
    run_generic(a, device, stream);
    uint32_t it = 0;
    check_hip(hipMemcpyAsync(&it, d_it, sizeof it, hipMemcpyDeviceToHost, stream), "copy out");
    if (post_host)
        check_hip(hipMemcpyAsync(post_host, d_post, c.n * sizeof(double), hipMemcpyDeviceToHost, stream), "copy out");
    if (bits_host) check_hip(hipMemcpyAsync(bits_host, d_bits, c.k_info, hipMemcpyDeviceToHost, stream), "copy out");
    check_hip(hipStreamSynchronize(stream), "decode");
    if (iterations) *iterations = it;
}

// cf_decode_batch, with its scratch (the layered kernel's task counter, the general kernel's
// messages) taken from the workspace of the stream, or, when `fixed` is given, from that
// buffer alone: then nothing is allocated, and the code's tables must already be on the device.
void decode_batch(const cf_code* h, int msg_type, const void* d_llr, size_t count, uint8_t* d_bits, void* d_post,
                  uint32_t* d_iterations, hipStream_t s, const FixedScratch* fixed) {
    if (!h) fail(Kind::Config, "null handle");
    if (count == 0) return;
    if (!d_llr) fail(Kind::Config, "null LLR buffer");
    int device = current_device();
    if (fixed && !h->code.tables_resident(device))
        fail(Kind::Config, "the code's tables are not on device " + std::to_string(device) +
                               " yet (an entry point that allocates nothing cannot upload them): make one decode "
                               "or generator call with this code on that device first");
    auto generic = [&](auto a) {
        if (!fixed) {
            run_generic(a, device, s);
            return;
        }
        Workspace ws;
        ws.ptr = fixed->messages;
        ws.bytes = fixed->message_bytes;
        ws.fixed = true;
        const hipError_t err = launch_generic(a, device, s, ws);
        if (err == hipErrorOutOfMemory)
            fail(Kind::Config, "a codeword of this code needs more message scratch than the " +
                                   std::to_string(fixed->message_bytes >> 20) + " MiB the library holds");
        check_hip(err, "decode launch");
    };
    if (msg_type == CF_MSG_F64) {
        GenericArgs<double> a = generic_args<double>(h, device);
        a.llr = static_cast<const double*>(d_llr);
        a.count = int64_t(count);
        a.bits = d_bits;
        a.post_out = static_cast<double*>(d_post);
        a.iterations = d_iterations;
        generic(a);
    } else if ((msg_type == CF_MSG_F32 || msg_type == CF_MSG_F16) && fast_layered(h)) {
        LayeredRun r;
        r.llr = static_cast<const float*>(d_llr);
        r.count = int64_t(count);
        r.bits = d_bits;
        r.post = static_cast<float*>(d_post);
        r.iterations = d_iterations;
        r.half = msg_type == CF_MSG_F16;
        run_layered(h, device, r, s, fixed ? fixed->counter : nullptr);
    } else if (msg_type == CF_MSG_F16) {
        fail(Kind::Config, "FP16 storage runs the layered normalized or offset min-sum");
    } else if (msg_type == CF_MSG_F32) {
        GenericArgs<float> a = generic_args<float>(h, device);
        a.llr = static_cast<const float*>(d_llr);
        a.count = int64_t(count);
        a.bits = d_bits;
        a.post_out = static_cast<float*>(d_post);
        a.iterations = d_iterations;
        generic(a);
    } else {
        fail(Kind::Config, "message type must be CF_MSG_F64, CF_MSG_F32 or CF_MSG_F16, got " +
                               std::to_string(msg_type));
    }
}

}  // namespace detail
}  // namespace comfec

using namespace comfec::detail;

extern "C" {

const char* cf_last_error(void) { return g_error.c_str(); }

cf_status cf_nr_create(uint32_t k_info, uint32_t e, int base_graph, int rv, uint32_t qm, cf_code** out) {
    return guarded([&] {
        if (!out) fail(Kind::Config, "null output handle");
        *out = new cf_code(NrCode::with_options(k_info, e, base_graph, rv, qm));
    });
}

cf_status cf_lifted_create(int base_graph, uint32_t z, cf_code** out) {
    return guarded([&] {
        if (!out) fail(Kind::Config, "null output handle");
        *out = new cf_code(NrCode::lifted(base_graph, z));
    });
}

cf_status cf_code_from_rows(uint32_t n, uint32_t m, const uint32_t* row_start, const uint32_t* cols, cf_code** out) {
    return guarded([&] {
        if (!out) fail(Kind::Config, "null output handle");
        *out = new cf_code(NrCode::from_rows(n, m, row_start, cols));
    });
}

void cf_code_destroy(cf_code* code) { delete code; }

cf_status cf_code_dims_get(const cf_code* h, cf_code_dims* out) {
    return guarded([&] {
        if (!h || !out) fail(Kind::Config, "null handle");
        const NrCode& c = h->code;
        *out = cf_code_dims{c.k_info, c.e, c.n, c.k, c.m, c.edges_b * c.z, uint32_t(c.bg), c.z, c.ils, c.qm, c.rv};
    });
}

cf_status cf_bit_map(const cf_code* h, uint32_t* positions, size_t e) {
    return guarded([&] {
        if (!h) fail(Kind::Config, "null handle");
        expect_len(e, h->code.e, "positions");
        std::memcpy(positions, h->code.map.data(), e * sizeof(uint32_t));
    });
}

cf_status cf_set_decoder(cf_code* h, int rule, double alpha, int layered, uint32_t max_iterations) {
    return guarded([&] {
        if (!h) fail(Kind::Config, "null handle");
        if (rule < 0 || rule > 2)
            fail(Kind::Config, "rule must be sum_product, min_sum, normalized_min_sum or offset_min_sum, got " +
                                   std::to_string(rule));
        std::lock_guard<std::mutex> l(h->mu);
        h->code.dec = DecoderSettings{rule, alpha, layered != 0, max_iterations};
    });
}

cf_status cf_encode_block(const cf_code* h, const uint8_t* info, size_t k_info, uint8_t* out, size_t e) {
    return guarded([&] {
        if (!h) fail(Kind::Config, "null handle");
        expect_len(k_info, h->code.k_info, "bits");
        expect_len(e, h->code.e, "coded bits");
        h->code.encode(info, out);
    });
}

cf_status cf_decode_llr(const cf_code* h, const double* llr, size_t n, double* post, uint32_t* iterations) {
    return guarded([&] {
        if (!h) fail(Kind::Config, "null handle");
        expect_len(n, h->code.n, "LLRs");
        decode_host_f64(h, llr, true, nullptr, post, nullptr, iterations);
    });
}

cf_status cf_decode_block(const cf_code* h, const double* llr, size_t e, uint8_t* bits, size_t k_info) {
    return guarded([&] {
        if (!h) fail(Kind::Config, "null handle");
        expect_len(e, h->code.e, "LLRs");
        expect_len(k_info, h->code.k_info, "decoded bits");
        decode_host_f64(h, llr, false, nullptr, nullptr, bits, nullptr);
    });
}

cf_status cf_app_into(const cf_code* h, const double* llr, size_t e, const double* apriori, double* out,
                      size_t k_info, uint32_t* iterations) {
    return guarded([&] {
        if (!h) fail(Kind::Config, "null handle");
        expect_len(e, h->code.e, "LLRs");
        expect_len(k_info, h->code.k_info, "a posteriori LLRs");
        std::vector<double> post(h->code.n);
        decode_host_f64(h, llr, false, apriori, post.data(), nullptr, iterations);
        const std::vector<uint32_t>& info = h->code.info;
        for (size_t i = 0; i < k_info; ++i) out[i] = post[info.empty() ? i : info[i]];
    });
}

cf_status cf_decode_batch(const cf_code* h, int msg_type, const void* d_llr, size_t count, uint8_t* d_bits,
                          void* d_post, uint32_t* d_iterations, void* stream) {
    return guarded([&] {
        decode_batch(h, msg_type, d_llr, count, d_bits, d_post, d_iterations, static_cast<hipStream_t>(stream),
                     nullptr);
    });
}

}  // extern "C"
