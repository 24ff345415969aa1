// How much scratch a launch of the general kernels takes: host-only arithmetic, with no HIP
// dependency, so that it builds and runs on a machine without a GPU.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace comfec {

// The result of sizing: `units` lanes (or codeword slots) at `scratch`.
struct Sizing {
    enum Status { kOk, kNothing, kOutOfMemory };
    Status status = kNothing;
    int64_t units = 0;
    void* scratch = nullptr;
};

// Units of per_unit bytes each for `count` codewords: as many as `budget` bytes hold (a
// multiple of `align`, at least `floor_units`), at most `count`, and with `fixed` at most
// what the caller's buffer of fixed_bytes holds (none fitting: out of memory). get(bytes)
// returns scratch of that size or null; on null the count halves (a multiple of `align`,
// at least 64) and get is asked again, down to 64 units; each unit then decodes more
// codewords, so the results do not change.
template <typename Get>
Sizing size_scratch(int64_t count, size_t per_unit, size_t budget, int64_t align, int64_t floor_units, bool fixed,
                    size_t fixed_bytes, Get&& get) {
    Sizing s;
    if (per_unit == 0 || align < 1) {
        s.status = Sizing::kOutOfMemory;
        return s;
    }
    int64_t units = int64_t(budget / per_unit) / align * align;
    units = std::max(units, floor_units);
    if (fixed) {
        if (fixed_bytes < per_unit) {
            s.status = Sizing::kOutOfMemory;
            return s;
        }
        units = std::min<int64_t>(units, int64_t(fixed_bytes / per_unit));
    }
    units = std::min(units, count);
    if (units < 1) return s;  // nothing to decode
    void* p = get(size_t(units) * per_unit);
    while (!p && units > 64) {
        units = std::max<int64_t>(64, (units / 2) / align * align);
        p = get(size_t(units) * per_unit);
    }
    if (!p) {
        s.status = Sizing::kOutOfMemory;
        return s;
    }
    s.status = Sizing::kOk;
    s.units = units;
    s.scratch = p;
    return s;
}

}  // namespace comfec
