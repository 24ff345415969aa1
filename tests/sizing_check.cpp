// CPU unit test of the general kernels' scratch sizing (com-rs_amd/csrc/sizing.hpp): built
// and run by tests/test_sizing_host.py with g++ and the address and undefined-behaviour
// sanitizers. A fake allocator refuses requests above a limit, as a device that cannot hold
// the whole grid does; every case prints one line and a failed case ends with status 1.
#include <cstdio>
#include <vector>

#include "../com-rs_amd/csrc/sizing.hpp"

using comfec::Sizing;
using comfec::size_scratch;

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
    if (!ok) ++failures;
}

// Records every request and refuses more than `limit` bytes; a granted request gets a
// token address (the sizing never reads or writes the scratch).
char token;
struct FakeDevice {
    size_t limit;
    std::vector<size_t> asked;
    void* get(size_t bytes) {
        asked.push_back(bytes);
        return bytes > limit ? nullptr : &token;
    }
};

}  // namespace

int main() {
    const size_t budget = size_t(8) << 30;
    {
        FakeDevice d{size_t(1) << 40, {}};
        const Sizing s = size_scratch(1000, 100, budget, 64, 64, false, 0, [&](size_t b) { return d.get(b); });
        expect(s.status == Sizing::kOk && s.units == 1000 && d.asked.size() == 1 && d.asked[0] == 100000,
               "a batch that fits: one unit per codeword, one request");
    }
    {
        FakeDevice d{size_t(1) << 40, {}};
        const Sizing s = size_scratch(0, 100, budget, 64, 64, false, 0, [&](size_t b) { return d.get(b); });
        expect(s.status == Sizing::kNothing && d.asked.empty(), "no codewords: nothing asked");
    }
    {
        // rows: the budget holds budget / per_unit lanes, rounded down to a multiple of 64
        FakeDevice d{size_t(1) << 40, {}};
        const size_t per = 1000003;
        const Sizing s = size_scratch(int64_t(1) << 40, per, budget, 64, 64, false, 0, [&](size_t b) { return d.get(b); });
        const int64_t want = int64_t(budget / per) / 64 * 64;
        expect(s.status == Sizing::kOk && s.units == want && s.units % 64 == 0, "the budget bounds a huge batch");
    }
    {
        // the device refuses: the lanes halve, multiples of 64, until the request fits
        FakeDevice d{size_t(300) * 64 * 1000, {}};
        const Sizing s = size_scratch(100000, 1000, budget, 64, 64, false, 0, [&](size_t b) { return d.get(b); });
        bool halving = d.asked.size() >= 2;
        for (size_t i = 1; i < d.asked.size(); ++i) halving = halving && d.asked[i] < d.asked[i - 1];
        expect(s.status == Sizing::kOk && s.units % 64 == 0 && size_t(s.units) * 1000 <= d.limit &&
                   size_t(s.units) * 2 * 1000 > d.limit && halving,
               "a refused request halves the lanes until one fits");
    }
    {
        // not even 64 lanes fit: out of memory, and the last request is for 64
        FakeDevice d{63 * 1000, {}};
        const Sizing s = size_scratch(100000, 1000, budget, 64, 64, false, 0, [&](size_t b) { return d.get(b); });
        expect(s.status == Sizing::kOutOfMemory && !d.asked.empty() && d.asked.back() == 64 * 1000 && !s.scratch,
               "out of memory when 64 lanes do not fit");
    }
    {
        // groups (align 1): halving stops at 64 slots, never below
        FakeDevice d{70 * 500, {}};
        const Sizing s = size_scratch(1000, 500, budget, 1, 1, false, 0, [&](size_t b) { return d.get(b); });
        expect(s.status == Sizing::kOk && s.units == 64 && d.asked.back() == 64 * 500,
               "groups halve to 64 slots and take them");
    }
    {
        // a fixed buffer: as many units as it holds, never a request beyond it
        FakeDevice d{10000, {}};
        const Sizing s = size_scratch(1000, 150, budget, 64, 64, true, 10000, [&](size_t b) { return d.get(b); });
        expect(s.status == Sizing::kOk && s.units == 10000 / 150 && d.asked.size() == 1 && d.asked[0] <= 10000,
               "a fixed buffer bounds the lanes");
        FakeDevice e{100, {}};
        const Sizing t = size_scratch(1000, 150, budget, 64, 64, true, 100, [&](size_t b) { return e.get(b); });
        expect(t.status == Sizing::kOutOfMemory && e.asked.empty(), "a fixed buffer smaller than one unit");
    }
    {
        FakeDevice d{size_t(1) << 40, {}};
        const Sizing s = size_scratch(10, 0, budget, 64, 64, false, 0, [&](size_t b) { return d.get(b); });
        expect(s.status == Sizing::kOutOfMemory && d.asked.empty(), "zero bytes per unit is refused");
    }
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
