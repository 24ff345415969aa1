// The layered min-sum kernels with FP32 storage (one translation unit per
// storage type, so that the builds run side by side).
#include "decode_layered.hpp"

namespace comfec {

const void* layered_kernel_f32(int bg_mask, bool solo, int zc, bool nms, bool dual) {
    using namespace layered;
    if (dual) {
        if (!solo) return nullptr;
        if (zc == 384) {
            if (bg_mask == 1) return nms ? reinterpret_cast<const void*>(decode_layered_dual_kernel<1, 384, true>)
                                         : reinterpret_cast<const void*>(decode_layered_dual_kernel<1, 384, false>);
            if (bg_mask == 2) return nms ? reinterpret_cast<const void*>(decode_layered_dual_kernel<2, 384, true>)
                                         : reinterpret_cast<const void*>(decode_layered_dual_kernel<2, 384, false>);
            return nullptr;
        }
        if (zc != 0 || nms) return nullptr;
        if (bg_mask == 1) return reinterpret_cast<const void*>(decode_layered_dual_kernel<1, 0, false>);
        if (bg_mask == 2) return reinterpret_cast<const void*>(decode_layered_dual_kernel<2, 0, false>);
        if (bg_mask == 3) return reinterpret_cast<const void*>(decode_layered_dual_kernel<3, 0, false>);
        return nullptr;
    }
    if (solo && zc == 384) {
        if (nms) {
            if (bg_mask == 1) return reinterpret_cast<const void*>(decode_layered_kernel<1, false, true, 384, true>);
            if (bg_mask == 2) return reinterpret_cast<const void*>(decode_layered_kernel<2, false, true, 384, true>);
        }
        if (bg_mask == 1) return reinterpret_cast<const void*>(decode_layered_kernel<1, false, true, 384>);
        if (bg_mask == 2) return reinterpret_cast<const void*>(decode_layered_kernel<2, false, true, 384>);
    }
    switch (bg_mask | (solo ? 4 : 0)) {
        case 1: return reinterpret_cast<const void*>(decode_layered_kernel<1, false, false, 0>);
        case 2: return reinterpret_cast<const void*>(decode_layered_kernel<2, false, false, 0>);
        case 3: return reinterpret_cast<const void*>(decode_layered_kernel<3, false, false, 0>);
        case 5: return reinterpret_cast<const void*>(decode_layered_kernel<1, false, true, 0>);
        case 6: return reinterpret_cast<const void*>(decode_layered_kernel<2, false, true, 0>);
        case 7: return reinterpret_cast<const void*>(decode_layered_kernel<3, false, true, 0>);
        default: return nullptr;
    }
}

}  // namespace comfec
