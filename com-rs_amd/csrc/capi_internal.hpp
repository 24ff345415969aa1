// Internals shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/comfec_gpu.h"
#include "kernels.hpp"
#include "nr_code.hpp"
#include "sim.hpp"

struct cf_code {
    comfec::NrCode code;
    mutable std::mutex mu;  // guards the decoder settings
    explicit cf_code(comfec::NrCode&& c) : code(std::move(c)) {}
    comfec::DecoderSettings settings() const {
        std::lock_guard<std::mutex> l(mu);
        return code.dec;
    }
};

// A scenario's record of a batch, or (view set) a read-only view of a condition's record.
struct cf_record {
    comfec::sim::Record rec;
    const comfec::sim::Record* view = nullptr;
    bool failed = false;
    int kind = 0;
    std::string message;
    const comfec::sim::Record& get() const { return view ? *view : rec; }
};

struct cf_results {
    std::vector<comfec::sim::ConditionResult> results;
};

struct cf_hooks {
    std::vector<std::shared_ptr<comfec::sim::Callback>> callbacks;
    std::vector<std::shared_ptr<comfec::sim::Logger>> loggers;
};

namespace comfec {
namespace detail {

extern thread_local std::string g_error;

// A grow-only device buffer per (device, stream, purpose); hold `mu` while using it.
struct WorkspaceSlot {
    std::mutex mu;
    Workspace ws;
};
WorkspaceSlot& workspace(int device, hipStream_t s, int purpose);
// workspace purposes: 0 the general kernel, 1 training (learned_api.cpp), 2 the layered
// kernel's task counter
enum { kQueueSlot = 2 };
DevCode dev_code(const NrCode& c, int device);
// the generator's arguments for a code, a modulation and one SNR (campaign.cpp)
TrialGenArgs make_generator_args(const NrCode& c, int device, int modulation, int snr_kind, double snr_db,
                                 uint64_t seed, uint32_t point);
cf_status status_of(Kind k);
void check_hip(hipError_t err, const char* what);
int current_device();
bool fast_layered(const cf_code* h);
// a launch of the layered min-sum kernel on `count` codewords of one code
struct LayeredRun {
    const float* llr = nullptr;
    int64_t count = 0;
    uint8_t* bits = nullptr;
    float* post = nullptr;
    uint32_t* iterations = nullptr;
    uint8_t* ok = nullptr;
    const float* params = nullptr;  // learned (a, b) per iteration and base row
    uint32_t learned_iterations = 0;
    bool fallback = false;
    bool half = false;
};
// counter: the task counter of the launch (device memory), or null for the stream's workspace
void run_layered(const cf_code* h, int device, const LayeredRun& r, hipStream_t stream, uint32_t* counter = nullptr);
// Scratch given whole to a decode: the layered kernel's task counter and the general
// kernels' message buffer (device memory of the current device).
struct FixedScratch {
    uint32_t* counter;
    void* messages;
    size_t message_bytes;
};
// cf_decode_batch; with `fixed`, no allocation (its scratch only, the code's tables resident)
void decode_batch(const cf_code* h, int msg_type, const void* d_llr, size_t count, uint8_t* d_bits, void* d_post,
                  uint32_t* d_iterations, hipStream_t s, const FixedScratch* fixed);
BatchClass batch_class(const NrCode& c, const DecoderSettings& s, int device, int g);

// Runs f, mapping a thrown Error to its status and the thread's last message.
template <typename F>
cf_status guarded(F&& f) {
    try {
        f();
        g_error.clear();
        return CF_OK;
    } catch (const Error& e) {
        g_error = e.what();
        return status_of(e.kind);
    } catch (const std::bad_alloc&) {
        g_error = "device: out of host memory";
        return CF_DEVICE;
    } catch (const std::exception& e) {
        g_error = std::string("device: ") + e.what();
        return CF_DEVICE;
    }
}

}  // namespace detail
}  // namespace comfec
