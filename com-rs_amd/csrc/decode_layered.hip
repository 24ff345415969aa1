// Launches of the layered min-sum kernels (decode_layered.hpp): kernel selection, LDS
// sizes, occupancy.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace comfec {

const void* layered_kernel_f32(int bg_mask, bool solo, int zc, bool nms, bool dual);
const void* layered_kernel_f16(int bg_mask, bool solo, int zc, bool nms);

namespace {

const void* select_kernel(int bg_mask, bool half, bool solo, int zc, bool nms, bool dual) {
    if (bg_mask < 1 || bg_mask > 3) return nullptr;
    if (dual && (half || !solo || (zc == 384 && bg_mask == 3))) return nullptr;
    return half ? layered_kernel_f16(bg_mask, solo, zc, nms) : layered_kernel_f32(bg_mask, solo, zc, nms, dual);
}

}  // namespace

bool layered_supported(int bg, int z) { return (bg == 1 || bg == 2) && z >= 1 && z <= kMaxThreads; }

size_t layered_lds_bytes(int bg, int z, int g, bool half) {
    const int cols = bg == 1 ? 68 : 52;  // base columns
    const size_t post = size_t(g) * size_t(cols) * size_t(z) * (half ? 2 : 4);
    // the packed decisions, with a zero column after the last
    const size_t hd = (z % 32 == 0) ? size_t(g) * size_t(cols + 1) * 2 * size_t(z / 32) * 4 : 0;
    return ((post + 3) & ~size_t(3)) + hd;
}

hipError_t layered_occupancy(int bg_mask, bool half, bool solo, int zc, int threads, size_t lds,
                             int* blocks_per_cu, bool nms, bool dual) {
    const void* k = select_kernel(bg_mask, half, solo, zc, nms, dual);
    if (!k) return hipErrorInvalidValue;
    hipError_t err = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (err != hipSuccess) return err;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, threads, lds);
}

hipError_t layered_prepare(const LayeredLaunch& l) {
    const void* k = select_kernel(l.bg_mask, l.half, l.solo, l.zc, l.nms, l.dual);
    if (!k || l.lds > kLdsBudget) return hipErrorInvalidValue;
    // every instantiation addresses the posteriors by LDS byte address from 0: no static LDS
    // before them
    hipFuncAttributes attr{};
    hipError_t err = hipFuncGetAttributes(&attr, k);
    if (err != hipSuccess) return err;
    if (attr.sharedSizeBytes != 0) return hipErrorInvalidDeviceFunction;
    return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(l.lds));
}

hipError_t launch_layered(const LayeredArgs& a, const LayeredLaunch& l, hipStream_t stream, bool prepare) {
    const void* k = select_kernel(l.bg_mask, l.half, l.solo, l.zc, l.nms, l.dual);
    // two wave sets take exactly 2 x 384 lanes; a compile-time Z of 384 exactly 384 a set (the
    // kernel then takes every lane to own a row of the task's codeword)
    if (!k || l.grid < 1 || (l.dual ? l.threads != 2 * kMaxThreads
                                    : (l.threads < 64 || l.threads > kMaxThreads || l.threads % 64 != 0)))
        return hipErrorInvalidValue;
    if (l.zc == kMaxThreads && !l.dual && l.threads != kMaxThreads) return hipErrorInvalidValue;
    if (l.lds > kLdsBudget) return hipErrorInvalidValue;
    if (prepare) {
        // the dynamic LDS of this launch (the static tables come on top of it)
        hipError_t err = layered_prepare(l);
        if (err != hipSuccess) return err;
    }
    void* args[] = {const_cast<LayeredArgs*>(&a)};
    hipError_t err = hipLaunchKernel(k, dim3(unsigned(l.grid)), dim3(unsigned(l.threads)), args, l.lds, stream);
    if (err != hipSuccess) return err;
    return hipGetLastError();
}

}  // namespace comfec
