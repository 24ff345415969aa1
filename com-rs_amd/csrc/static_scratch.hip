// cf_decode_batch_static: cf_decode_batch for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops). Every scratch buffer of the decode
// (the layered kernel's task counter, the general kernels' messages) lies in one array of
// device memory that the library declares at file scope, so each GPU the library runs on
// holds one copy of it from the library's load to the end of the process; a call allocates
// nothing and frees nothing.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/comfec_gpu.h"
#include "capi_internal.hpp"

using namespace comfec;
using namespace comfec::detail;

namespace {

constexpr size_t kScratchBytes = CF_STATIC_SCRATCH_BYTES;
constexpr size_t kCounterBytes = 256;  // the task counter, then the messages from a 256-byte boundary
static_assert(kScratchBytes % 256 == 0 && kScratchBytes > kCounterBytes, "scratch layout");

// In host code this name is a host-side stand-in, not a device address: host code uses only
// the per-device address that hipGetSymbolAddress gives.
__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
std::vector<void*> g_ptr;  // the array's address on each device, looked up once per device

// The array's address on `device`, which must be the current device.
uint8_t* scratch_on(int device) {
    std::lock_guard<std::mutex> l(g_mu);
    if (g_ptr.size() <= size_t(device)) g_ptr.resize(size_t(device) + 1, nullptr);
    if (!g_ptr[size_t(device)]) {
        void* p = nullptr;
        check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress of the static scratch");
        if (!p) fail(Kind::Device, "hipGetSymbolAddress of the static scratch gave a null address");
        g_ptr[size_t(device)] = p;
    }
    return static_cast<uint8_t*>(g_ptr[size_t(device)]);
}

}  // namespace

extern "C" {

cf_status cf_decode_batch_static(const cf_code* code, int msg_type, const void* d_llr, size_t count,
                                 uint8_t* d_bits, void* d_post, uint32_t* d_iterations, void* stream) {
    return guarded([&] {
        if (!code) fail(Kind::Config, "null handle");
        uint8_t* p = scratch_on(current_device());
        const FixedScratch f{reinterpret_cast<uint32_t*>(p), p + kCounterBytes, kScratchBytes - kCounterBytes};
        decode_batch(code, msg_type, d_llr, count, d_bits, d_post, d_iterations, static_cast<hipStream_t>(stream),
                     &f);
    });
}

}  // extern "C"
