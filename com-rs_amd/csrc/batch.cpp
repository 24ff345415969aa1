// Launches of the layered min-sum kernel: batches of one code (cf_decode_batch, the learned
// decoder, the campaigns) and plans for batches of many codes (cf_batch_plan_*): the task
// list, its device copy, the autotuned workgroup width and the captured HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"
#include "nr_code.hpp"

using namespace comfec;
using namespace comfec::detail;

namespace comfec {
namespace detail {

// The class of a code: its tables on the device and its min-sum rule.
BatchClass batch_class(const NrCode& c, const DecoderSettings& s, int device, int g) {
    const NrCode::DeviceTables& t = c.device_tables(device);
    BatchClass k{};
    k.bg = c.bg;
    k.z = int32_t(c.z);
    k.n = int32_t(c.n);
    k.k = int32_t(c.k);
    k.k_info = int32_t(c.k_info);
    k.e = int32_t(c.e);
    k.g = g;
    k.max_iterations = int32_t(s.max_iterations);
    // g(m) = a max(m - b, 0): (alpha, 0) normalized, (1, beta) offset min-sum
    k.a = s.rule == 1 ? float(s.alpha) : 1.0f;
    k.b = s.rule == 1 ? 0.0f : float(s.alpha);
    k.shift32 = t.shift32;
    k.col = t.col;
    k.row_start = t.row_start;
    k.src_start = t.src_start;
    k.src = t.src;
    k.map = t.map;
    k.direct = c.repeats ? 0 : 1;
    return k;
}

namespace {

int round64(int x) { return (x + 63) / 64 * 64; }

// codewords per task for lifting size z in a workgroup of `threads` lanes
int group_of(int z, int threads) { return std::max(1, std::min(threads / z, kMaxGroup)); }

// Whether a launch runs two wave sets a codeword: FP32 tasks of one codeword in workgroups of
// at least 256 lanes, of which the CU holds one (the two-set kernel holds one too)
bool two_sets(bool solo, bool half, int threads, int per_cu) { return solo && !half && threads >= 256 && per_cu == 1; }

// the per-code width of the codes that run on two wave sets: one codeword a task, 768 lanes
constexpr int kDualWidth = 2 * kMaxThreads;
// codewords per task at a per-code width
int group_at(int z, int w) { return w == kDualWidth ? 1 : group_of(z, w); }

}  // namespace

// The layered kernel on `count` codewords of one code: a workgroup of the next multiple of
// 64 lanes above Z takes floor(lanes / Z) codewords at a time.
void run_layered(const cf_code* h, int device, const LayeredRun& r, hipStream_t stream, uint32_t* counter) {
    const NrCode& c = h->code;
    if (!layered_supported(c.bg, int(c.z)))
        fail(Kind::Config, "the layered kernel needs a lifting size of at most " + std::to_string(kMaxThreads));
    if (r.count <= 0) return;
    const int threads = std::min(round64(int(c.z)), kMaxThreads);
    const int g = group_of(int(c.z), threads);
    const DecoderSettings st = h->settings();
    LayeredArgs a{};
    a.one = batch_class(c, st, device, g);
    a.n_tasks = int32_t((r.count + g - 1) / g);
    a.count = r.count;
    a.params = r.params;
    a.learned_iterations = int32_t(r.learned_iterations);
    a.fallback = r.fallback ? 1 : 0;
    a.llr = r.llr;
    a.bits = r.bits;
    a.post_out = r.post;
    a.iterations = r.iterations;
    a.ok = r.ok;
    const size_t off = layered_shared_offset(layered_lds_bytes(c.bg, int(c.z), g, r.half));
    a.shared_off = int32_t(off);
    // persistent: as many workgroups as fit on the device, taking tasks from a counter, so
    // that each loads the code's tables once and the codewords that stop early free their
    // workgroup for the next
    const bool solo = g == 1;
    const int zc = c.z == 384 && solo ? 384 : 0;
    // the normalized min-sum's own instantiation (no offset, no learned parameters)
    const bool nms = zc == 384 && !r.params && st.rule == 1;
    const size_t lds = off + kSharedBytes;
    int per_cu = 0, cus = 0;
    check_hip(layered_occupancy(c.bg, r.half, solo, zc, threads, lds, &per_cu, nms), "occupancy");
    // one FP32 codeword that fills a CU (four to six waves): two sets of six waves on it, the
    // layers of a barrier group side by side (twelve waves, three a SIMD, where one set alone
    // sits 2:2:1:1 or 1:1:1:1)
    const bool dual = two_sets(solo, r.half, threads, per_cu);
    const int lanes = dual ? 2 * kMaxThreads : threads;
    const size_t lds_run = dual ? lds + kExchangeBytes : lds;
    if (dual) check_hip(layered_occupancy(c.bg, r.half, solo, zc, lanes, lds_run, &per_cu, nms, true), "occupancy");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
    const int grid = int(std::min<int64_t>(a.n_tasks, int64_t(std::max(per_cu, 1)) * std::max(cus, 1)));
    LayeredLaunch l{c.bg, r.half, solo, zc, lanes, std::max(grid, 1), lds_run, nms, dual};
    if (counter) {
        // the caller's counter: a fill kernel (no memset node) zeroes it before the launch
        check_hip(launch_zero(reinterpret_cast<uint64_t*>(counter), 1, stream), "zero launch");
        a.queue = counter;
        check_hip(launch_layered(a, l, stream, true), "layered decode launch");
        return;
    }
    // the counter lives in the stream's workspace, held until both launches are queued
    WorkspaceSlot& slot = workspace(device, stream, kQueueSlot);
    std::lock_guard<std::mutex> lock(slot.mu);
    uint64_t* ws_counter = static_cast<uint64_t*>(slot.ws.get(sizeof(uint64_t)));
    if (!ws_counter) fail(Kind::Device, "task counter: out of device memory");
    check_hip(launch_zero(ws_counter, 1, stream), "zero launch");
    a.queue = reinterpret_cast<uint32_t*>(ws_counter);
    check_hip(launch_layered(a, l, stream, true), "layered decode launch");
}

}  // namespace detail
}  // namespace comfec

// ---- plans for batches of many codes ---------------------------------------------------

// One launch of a plan's run: the tasks of the classes at one width.
struct PlanSegment {
    int threads = 0;
    uint32_t classes = 0, codewords = 0;
    size_t first_task = 0, n_tasks = 0;
    LayeredLaunch launch{};
    LayeredArgs args{};
};

struct cf_batch_plan {
    std::mutex mu;
    int device = 0;
    size_t count = 0;
    uint64_t llr_total = 0, bits_total = 0;
    std::vector<const cf_code*> class_code;  // per class
    std::vector<DecoderSettings> class_set;
    std::vector<uint32_t> cls_of;            // per codeword
    std::vector<uint16_t> max_it;
    std::vector<uint64_t> llr_off, bit_off;
    int bg_mask = 0, max_z = 0;
    // the configuration as set, and what it resolved to
    cf_batch_plan_config cfg{};
    bool half = false;
    std::vector<int> class_width;  // per class
    std::vector<BatchClass> classes;
    std::vector<BatchTask> tasks;  // segment by segment
    std::vector<uint32_t> order;
    std::vector<PlanSegment> segs;
    // device copies: task counters, classes, tasks, order, offsets and limits
    void* dev = nullptr;
    size_t dev_bytes = 0;
    uint64_t* counters = nullptr;
    // overlap: a stream per launch after the first, the fork event and a join event each
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    // the captured graph and the buffers it was captured with
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* g_llr = nullptr;
    const void* g_bits = nullptr;
    const void* g_its = nullptr;
    const void* g_ok = nullptr;

    ~cf_batch_plan() {
        drop_graph();
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        for (hipStream_t q : side) (void)hipStreamDestroy(q);
        for (hipEvent_t e : joins) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (dev) (void)hipFree(dev);
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

constexpr size_t kMaxSegments = 256 / sizeof(uint64_t);  // the counters fill the buffer's first 256 bytes
static_assert(kMaxThreads / 64 <= int(kMaxSegments), "a counter per width");

size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

int multiprocessors(int device) {
    int cus = 0;
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
    return std::max(cus, 1);
}

// The width for one code under per_code_width: of the multiples of 64 from Z (rounded up)
// to 384, the one with the most codeword rows per busy SIMD, g Z x (resident workgroups) /
// max(ceil(waves on the CU / 4), 2), at the occupancy the device reports for that width's
// LDS. A SIMD with one wave hides little of the LDS latency, so it counts as no faster per
// wave than one with two; ties go to the wider width, which has fewer tasks.
int modelled_width(const NrCode& c, bool half, uint32_t cap_per_cu) {
    int best_w = 0;
    double best = -1.0;
    for (int w = kMaxThreads; w >= round64(int(c.z)); w -= 64) {
        const int g = group_of(int(c.z), w);
        const size_t lds = layered_shared_offset(layered_lds_bytes(c.bg, int(c.z), g, half)) + kSharedBytes;
        int per_cu = 0;
        if (layered_occupancy(c.bg, half, g == 1, 0, w, lds, &per_cu) != hipSuccess || per_cu < 1) {
            (void)hipGetLastError();
            continue;
        }
        if (cap_per_cu) per_cu = std::min(per_cu, int(cap_per_cu));
        const int busy = (per_cu * (w / 64) + 3) / 4;
        const double rate = double(per_cu) * double(g) * double(c.z) / double(std::max(busy, 2));
        if (rate > best * (1.0 + 1e-12)) {
            best = rate;
            best_w = w;
        }
    }
    if (!best_w) fail(Kind::Device, "no workgroup width fits the device for Z = " + std::to_string(c.z));
    // an FP32 codeword that alone fills the CU at that width runs on two wave sets instead
    const int g = group_of(int(c.z), best_w);
    if (g == 1 && !half) {
        const size_t lds = layered_shared_offset(layered_lds_bytes(c.bg, int(c.z), 1, half)) + kSharedBytes;
        int per_cu = 0;
        if (layered_occupancy(c.bg, half, true, 0, best_w, lds, &per_cu) == hipSuccess &&
            two_sets(true, half, best_w, per_cu))
            return kDualWidth;
        (void)hipGetLastError();
    }
    return best_w;
}

void check_config(const cf_batch_plan& p, const cf_batch_plan_config& cfg) {
    if (cfg.msg_type != CF_MSG_F32 && cfg.msg_type != CF_MSG_F16)
        fail(Kind::Config, "a batch plan decodes with CF_MSG_F32 or CF_MSG_F16 messages, got " +
                               std::to_string(cfg.msg_type));
    if (cfg.per_code_width > 1) fail(Kind::Config, "per_code_width is 0 or 1");
    if (cfg.overlap > 1) fail(Kind::Config, "overlap is 0 or 1");
    const int t = int(cfg.threads);
    if (!cfg.per_code_width && (t < round64(p.max_z) || t > kMaxThreads || t % 64 != 0))
        fail(Kind::Config, "the workgroup width must be a multiple of 64 from " + std::to_string(round64(p.max_z)) +
                               " to " + std::to_string(kMaxThreads) + ", got " + std::to_string(t));
}

// Tasks for a configuration: each code's codewords in groups of G = floor(width / Z), those
// with the larger iteration limits together; per width (the widest first) one launch whose
// tasks run in decreasing order of work (the largest first, so that its tail is short).
void configure(cf_batch_plan& p, const cf_batch_plan_config& cfg) {
    check_config(p, cfg);
    const size_t nc = p.class_code.size();
    const bool half = cfg.msg_type == CF_MSG_F16;
    const int cus = multiprocessors(p.device);
    std::vector<int> width(nc);
    for (size_t k = 0; k < nc; ++k)
        width[k] = cfg.per_code_width ? modelled_width(p.class_code[k]->code, half, cfg.grid_per_cu) : int(cfg.threads);
    std::vector<int> widths(width);
    std::sort(widths.begin(), widths.end(), std::greater<int>());
    widths.erase(std::unique(widths.begin(), widths.end()), widths.end());
    if (widths.size() > kMaxSegments) fail(Kind::Config, "too many launch widths");

    std::vector<BatchClass> classes(nc);
    for (size_t k = 0; k < nc; ++k) {
        const NrCode& c = p.class_code[k]->code;
        classes[k] = batch_class(c, p.class_set[k], p.device, group_at(int(c.z), width[k]));
    }
    std::vector<std::vector<uint32_t>> members(nc);
    for (size_t i = 0; i < p.count; ++i) members[p.cls_of[i]].push_back(uint32_t(i));

    struct Pending {
        BatchTask t;
        double cost;
        std::vector<uint32_t> ids;
    };
    std::vector<BatchTask> tasks;
    std::vector<uint32_t> order;
    std::vector<PlanSegment> segs;
    for (const int w : widths) {
        PlanSegment sg;
        sg.threads = w;
        sg.first_task = tasks.size();
        size_t lds = 0;
        bool solo = true, all384 = true, nms = true;
        int bgm = 0;
        std::vector<Pending> pend;
        for (size_t k = 0; k < nc; ++k) {
            if (width[k] != w) continue;
            const NrCode& c = p.class_code[k]->code;
            const int g = classes[k].g;
            lds = std::max(lds, layered_lds_bytes(c.bg, int(c.z), g, half));
            solo = solo && g == 1;
            all384 = all384 && c.z == 384;
            nms = nms && p.class_set[k].rule == 1;
            bgm |= c.bg;
            ++sg.classes;
            auto& m = members[k];
            sg.codewords += uint32_t(m.size());
            std::stable_sort(m.begin(), m.end(), [&](uint32_t x, uint32_t y) { return p.max_it[x] > p.max_it[y]; });
            for (size_t s = 0; s < m.size(); s += size_t(g)) {
                Pending q;
                q.t = BatchTask{uint32_t(k), 0, uint32_t(std::min(m.size() - s, size_t(g))), 0};
                q.ids.assign(m.begin() + long(s), m.begin() + long(s + q.t.count));
                // a workgroup runs until its slowest codeword stops: limit x edges x Z per codeword
                q.cost = double(std::max<uint16_t>(p.max_it[q.ids[0]], 1)) * double(c.edges_b) * double(c.z);
                pend.push_back(std::move(q));
            }
        }
        std::stable_sort(pend.begin(), pend.end(), [](const Pending& x, const Pending& y) { return x.cost > y.cost; });
        for (auto& q : pend) {
            q.t.first = uint32_t(order.size());
            tasks.push_back(q.t);
            order.insert(order.end(), q.ids.begin(), q.ids.end());
        }
        sg.n_tasks = tasks.size() - sg.first_task;
        // the normalized min-sum instantiation serves Z = 384 alone, one base graph
        const int zc = solo && all384 && bgm != 3 ? 384 : 0;
        nms = nms && zc == 384;
        const size_t shared_off = layered_shared_offset(lds);
        lds = shared_off + kSharedBytes;
        int per_cu = 0;
        bool dual = w == kDualWidth;
        if (!dual) {
            check_hip(layered_occupancy(bgm, half, solo, zc, w, lds, &per_cu, nms), "occupancy");
            dual = two_sets(solo, half, w, per_cu);
        }
        if (dual) {
            lds += kExchangeBytes;
            check_hip(layered_occupancy(bgm, half, solo, zc, kDualWidth, lds, &per_cu, nms, true), "occupancy");
        }
        per_cu = std::max(per_cu, 1);
        if (cfg.grid_per_cu) per_cu = std::min(per_cu, int(cfg.grid_per_cu));
        const int grid = int(std::min<size_t>(sg.n_tasks, size_t(per_cu) * size_t(cus)));
        sg.launch = LayeredLaunch{bgm, half, solo, zc, dual ? kDualWidth : w, std::max(grid, 1), lds, nms, dual};
        sg.args.shared_off = int32_t(shared_off);
        segs.push_back(sg);
    }

    // one device buffer: [counters | classes | tasks | order | llr_off | bit_off | max_it]
    const size_t o_cls = 256, o_task = o_cls + up256(nc * sizeof(BatchClass));
    const size_t o_order = o_task + up256(tasks.size() * sizeof(BatchTask));
    const size_t o_llr = o_order + up256(p.count * 4), o_bit = o_llr + up256(p.count * 8);
    const size_t o_it = o_bit + up256(p.count * 8), total = o_it + up256(p.count * 2);
    check_hip(hipDeviceSynchronize(), "synchronize");  // no launch still reads the old tables
    p.drop_graph();
    if (total > p.dev_bytes) {
        if (p.dev) (void)hipFree(p.dev);
        p.dev = nullptr;
        p.dev_bytes = 0;
        check_hip(hipMalloc(&p.dev, total), "hipMalloc");
        p.dev_bytes = total;
    }
    std::vector<uint8_t> host(total, 0);
    std::memcpy(host.data() + o_cls, classes.data(), nc * sizeof(BatchClass));
    std::memcpy(host.data() + o_task, tasks.data(), tasks.size() * sizeof(BatchTask));
    std::memcpy(host.data() + o_order, order.data(), p.count * 4);
    std::memcpy(host.data() + o_llr, p.llr_off.data(), p.count * 8);
    std::memcpy(host.data() + o_bit, p.bit_off.data(), p.count * 8);
    std::memcpy(host.data() + o_it, p.max_it.data(), p.count * 2);
    check_hip(hipMemcpy(p.dev, host.data(), total, hipMemcpyHostToDevice), "copy plan");
    uint8_t* d = static_cast<uint8_t*>(p.dev);
    p.counters = reinterpret_cast<uint64_t*>(d);
    for (size_t s = 0; s < segs.size(); ++s) {
        LayeredArgs& a = segs[s].args;
        const int32_t shared_off = a.shared_off;
        a = LayeredArgs{};
        a.classes = reinterpret_cast<const BatchClass*>(d + o_cls);
        a.tasks = reinterpret_cast<const BatchTask*>(d + o_task) + segs[s].first_task;
        a.n_tasks = int32_t(segs[s].n_tasks);
        a.queue = reinterpret_cast<uint32_t*>(p.counters + s);
        a.order = reinterpret_cast<const uint32_t*>(d + o_order);
        a.llr_off = reinterpret_cast<const uint64_t*>(d + o_llr);
        a.bit_off = reinterpret_cast<const uint64_t*>(d + o_bit);
        a.max_it = reinterpret_cast<const uint16_t*>(d + o_it);
        a.shared_off = shared_off;
    }
    // the side streams and events of an overlapped run (default priority)
    if (cfg.overlap)
        while (p.side.size() + 1 < segs.size()) {
            if (!p.fork) check_hip(hipEventCreateWithFlags(&p.fork, hipEventDisableTiming), "event");
            hipStream_t q = nullptr;
            check_hip(hipStreamCreateWithFlags(&q, hipStreamNonBlocking), "hipStreamCreate");
            p.side.push_back(q);
            hipEvent_t e = nullptr;
            check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            p.joins.push_back(e);
        }
    p.cfg = cfg;
    if (cfg.per_code_width) p.cfg.threads = 0;
    p.half = half;
    p.class_width = std::move(width);
    p.classes = std::move(classes);
    p.tasks = std::move(tasks);
    p.order = std::move(order);
    p.segs = std::move(segs);
}

// the counter fill and the decode launches
void enqueue(cf_batch_plan& p, const float* llr, uint8_t* bits, uint32_t* its, uint8_t* ok, hipStream_t s,
             bool set_attribute) {
    check_hip(launch_zero(p.counters, int(p.segs.size()), s), "zero launch");
    const bool overlap = p.cfg.overlap && p.segs.size() > 1;
    if (overlap) check_hip(hipEventRecord(p.fork, s), "fork");
    for (size_t i = 0; i < p.segs.size(); ++i) {
        LayeredArgs a = p.segs[i].args;
        a.llr = llr;
        a.bits = bits;
        a.iterations = its;
        a.ok = ok;
        hipStream_t q = s;
        if (overlap && i > 0) {
            q = p.side[i - 1];
            check_hip(hipStreamWaitEvent(q, p.fork, 0), "fork");
        }
        check_hip(launch_layered(a, p.segs[i].launch, q, set_attribute), "batch decode launch");
        if (q != s) check_hip(hipEventRecord(p.joins[i - 1], q), "join");
    }
    if (overlap)
        for (size_t i = 1; i < p.segs.size(); ++i) check_hip(hipStreamWaitEvent(s, p.joins[i - 1], 0), "join");
}

void run_plan(cf_batch_plan& p, const float* llr, uint8_t* bits, uint32_t* its, uint8_t* ok, bool use_graph,
              hipStream_t s) {
    if (p.count == 0) return;
    if (!llr) fail(Kind::Config, "null LLR buffer");
    if (!use_graph) {
        enqueue(p, llr, bits, its, ok, s, true);
        return;
    }
    if (!p.exec || p.g_llr != llr || p.g_bits != bits || p.g_its != its || p.g_ok != ok) {
        p.drop_graph();
        if (!p.capture_stream) check_hip(hipStreamCreate(&p.capture_stream), "hipStreamCreate");
        // kernel attributes are set outside the capture
        for (const PlanSegment& sg : p.segs) check_hip(layered_prepare(sg.launch), "kernel attribute");
        check_hip(hipStreamBeginCapture(p.capture_stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            enqueue(p, llr, bits, its, ok, p.capture_stream, false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(p.capture_stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check_hip(hipStreamEndCapture(p.capture_stream, &p.graph), "end capture");
        check_hip(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0), "graph instantiate");
        p.g_llr = llr;
        p.g_bits = bits;
        p.g_its = its;
        p.g_ok = ok;
    }
    check_hip(hipGraphLaunch(p.exec, s), "graph launch");
}

// What a configuration launches: per class its width, per launch its kernel and grid.
std::vector<int64_t> launch_signature(const cf_batch_plan& p) {
    std::vector<int64_t> v(p.class_width.begin(), p.class_width.end());
    v.push_back(p.half ? 1 : 0);
    v.push_back(p.cfg.overlap && p.segs.size() > 1 ? 1 : 0);
    for (const PlanSegment& sg : p.segs) {
        v.push_back(sg.threads);
        v.push_back(sg.launch.grid);
        v.push_back(int64_t(sg.launch.lds));
    }
    return v;
}

cf_batch_plan* create_plan(const cf_code* const* codes, const uint64_t* llr_offsets, const uint32_t* max_iterations,
                           size_t count, int msg_type) {
    if (!codes || !llr_offsets) fail(Kind::Config, "null code list or offsets");
    if (msg_type != CF_MSG_F32 && msg_type != CF_MSG_F16)
        fail(Kind::Config, "a batch plan decodes with CF_MSG_F32 or CF_MSG_F16 messages, got " +
                               std::to_string(msg_type));
    auto p = std::make_unique<cf_batch_plan>();
    p->device = current_device();
    p->half = msg_type == CF_MSG_F16;
    p->cfg.msg_type = msg_type;
    p->count = count;
    p->cls_of.resize(count);
    p->max_it.resize(count);
    p->llr_off.resize(count);
    p->bit_off.resize(count);
    if (llr_offsets[0] != 0) fail(Kind::Config, "the LLR offsets start at 0");
    for (size_t i = 0; i < count; ++i) {
        const cf_code* h = codes[i];
        if (!h) fail(Kind::Config, "null code of codeword " + std::to_string(i));
        auto it = std::find(p->class_code.begin(), p->class_code.end(), h);
        if (it == p->class_code.end()) {
            const DecoderSettings s = h->settings();
            if (!s.layered || (s.rule != 1 && s.rule != 2) || !layered_supported(h->code.bg, int(h->code.z)))
                fail(Kind::Config, "a batch plan runs the layered normalized or offset min-sum on lifting sizes "
                                   "up to " + std::to_string(kMaxThreads));
            p->class_code.push_back(h);
            p->class_set.push_back(s);
            p->bg_mask |= h->code.bg;
            p->max_z = std::max(p->max_z, int(h->code.z));
            it = p->class_code.end() - 1;
        }
        const size_t k = size_t(it - p->class_code.begin());
        const NrCode& c = h->code;
        if (llr_offsets[i + 1] < llr_offsets[i]) fail(Kind::Config, "the LLR offsets must not decrease");
        expect_len(size_t(llr_offsets[i + 1] - llr_offsets[i]), c.e, "LLRs");
        const uint32_t lim = max_iterations ? max_iterations[i] : p->class_set[k].max_iterations;
        if (lim > 65535) fail(Kind::Config, "an iteration limit of at most 65535, got " + std::to_string(lim));
        p->cls_of[i] = uint32_t(k);
        p->max_it[i] = uint16_t(lim);
        p->llr_off[i] = llr_offsets[i];
        p->bit_off[i] = p->bits_total;
        p->bits_total += c.k_info;
    }
    p->llr_total = count ? llr_offsets[count] : 0;
    p->cfg.threads = uint32_t(std::min(kMaxThreads, round64(p->max_z)));
    if (count) configure(*p, p->cfg);
    return p.release();
}

}  // namespace

extern "C" {

cf_status cf_batch_plan_create(const cf_code* const* codes, const uint64_t* llr_offsets,
                               const uint32_t* max_iterations, size_t count, int msg_type, cf_batch_plan** out) {
    return guarded([&] {
        if (!out) fail(Kind::Config, "null output handle");
        *out = create_plan(codes, llr_offsets, max_iterations, count, msg_type);
    });
}

void cf_batch_plan_destroy(cf_batch_plan* plan) { delete plan; }

cf_status cf_batch_plan_dims_get(const cf_batch_plan* plan, cf_batch_plan_dims* out) {
    return guarded([&] {
        if (!plan || !out) fail(Kind::Config, "null plan");
        const PlanSegment* s0 = plan->segs.empty() ? nullptr : &plan->segs[0];
        *out = cf_batch_plan_dims{plan->llr_total,
                                  plan->bits_total,
                                  uint32_t(plan->count),
                                  uint32_t(plan->class_code.size()),
                                  uint32_t(plan->tasks.size()),
                                  s0 ? uint32_t(s0->launch.threads) : plan->cfg.threads,
                                  s0 ? uint32_t(s0->launch.grid) : 0u,
                                  s0 ? uint32_t(s0->launch.lds) : 0u,
                                  plan->half ? CF_MSG_F16 : CF_MSG_F32,
                                  uint32_t(plan->segs.size())};
    });
}

cf_status cf_batch_plan_set_threads(cf_batch_plan* plan, uint32_t threads) {
    return guarded([&] {
        if (!plan) fail(Kind::Config, "null plan");
        std::lock_guard<std::mutex> l(plan->mu);
        cf_batch_plan_config c = plan->cfg;
        c.threads = threads;
        c.per_code_width = 0;
        check_config(*plan, c);
        if (plan->count) configure(*plan, c);
        else plan->cfg = c;
    });
}

cf_status cf_batch_plan_config_get(const cf_batch_plan* plan, cf_batch_plan_config* out) {
    return guarded([&] {
        if (!plan || !out) fail(Kind::Config, "null plan");
        *out = plan->cfg;
    });
}

cf_status cf_batch_plan_configure(cf_batch_plan* plan, const cf_batch_plan_config* config) {
    return guarded([&] {
        if (!plan || !config) fail(Kind::Config, "null plan or configuration");
        std::lock_guard<std::mutex> l(plan->mu);
        check_config(*plan, *config);
        if (plan->count) configure(*plan, *config);
        else {
            plan->cfg = *config;
            plan->half = config->msg_type == CF_MSG_F16;
        }
    });
}

cf_status cf_batch_plan_segment_get(const cf_batch_plan* plan, uint32_t index, cf_batch_plan_segment* out) {
    return guarded([&] {
        if (!plan || !out) fail(Kind::Config, "null plan");
        if (index >= plan->segs.size())
            fail(Kind::Config, "launch " + std::to_string(index) + " of " + std::to_string(plan->segs.size()));
        const PlanSegment& s = plan->segs[index];
        *out = cf_batch_plan_segment{uint32_t(s.launch.threads), uint32_t(s.launch.grid), uint32_t(s.launch.lds),
                                     s.classes,           uint32_t(s.n_tasks),     s.codewords};
    });
}

cf_status cf_batch_plan_tune(cf_batch_plan* plan, const float* d_llr, uint8_t* d_bits, uint32_t* d_iterations,
                             uint8_t* d_ok, uint32_t repeats, uint32_t flags, cf_batch_plan_trial* trials,
                             size_t trials_cap, size_t* trials_count, double* best_ms, void* stream) {
    return guarded([&] {
        if (!plan) fail(Kind::Config, "null plan");
        if (flags & ~CF_TUNE_STORAGE) fail(Kind::Config, "unknown tuning flags " + std::to_string(flags));
        std::lock_guard<std::mutex> l(plan->mu);
        if (trials_count) *trials_count = 0;
        if (plan->count == 0) return;
        hipStream_t s = static_cast<hipStream_t>(stream);
        repeats = std::max<uint32_t>(repeats, 1);
        hipEvent_t e0, e1;
        check_hip(hipEventCreate(&e0), "event");
        std::unique_ptr<void, void (*)(void*)> g0(e0, [](void* e) { (void)hipEventDestroy(hipEvent_t(e)); });
        check_hip(hipEventCreate(&e1), "event");
        std::unique_ptr<void, void (*)(void*)> g1(e1, [](void* e) { (void)hipEventDestroy(hipEvent_t(e)); });

        std::vector<int> storages{plan->cfg.msg_type};
        if (flags & CF_TUNE_STORAGE) storages.push_back(plan->cfg.msg_type == CF_MSG_F32 ? CF_MSG_F16 : CF_MSG_F32);
        std::vector<cf_batch_plan_config> cands;
        for (const int m : storages)
            for (const uint32_t per_cu : {0u, 1u}) {
                for (int t = round64(plan->max_z); t <= kMaxThreads; t += 64)
                    cands.push_back(cf_batch_plan_config{uint32_t(t), 0u, per_cu, m, 0u});
                cands.push_back(cf_batch_plan_config{0u, 1u, per_cu, m, 0u});
                cands.push_back(cf_batch_plan_config{0u, 1u, per_cu, m, 1u});
            }
        std::vector<std::vector<int64_t>> seen;
        size_t n = 0;
        bool have = false;
        cf_batch_plan_config best{};
        double best_t = 0.0;
        for (const cf_batch_plan_config& c : cands) {
            configure(*plan, c);
            std::vector<int64_t> sig = launch_signature(*plan);
            if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
            seen.push_back(std::move(sig));
            run_plan(*plan, d_llr, d_bits, d_iterations, d_ok, false, s);  // warm-up
            check_hip(hipEventRecord(e0, s), "event");
            for (uint32_t i = 0; i < repeats; ++i) run_plan(*plan, d_llr, d_bits, d_iterations, d_ok, false, s);
            check_hip(hipEventRecord(e1, s), "event");
            check_hip(hipEventSynchronize(e1), "event");
            float ms = 0.0f;
            check_hip(hipEventElapsedTime(&ms, e0, e1), "event");
            const double t = double(ms) / double(repeats);
            if (trials && n < trials_cap) trials[n] = cf_batch_plan_trial{plan->cfg, uint32_t(plan->segs.size()), t};
            ++n;
            if (!have || t < best_t) {
                have = true;
                best = plan->cfg;
                best_t = t;
            }
        }
        configure(*plan, best);
        run_plan(*plan, d_llr, d_bits, d_iterations, d_ok, false, s);
        check_hip(hipStreamSynchronize(s), "tuned decode");
        if (trials_count) *trials_count = n;
        if (best_ms) *best_ms = best_t;
    });
}

cf_status cf_batch_plan_run(cf_batch_plan* plan, const float* d_llr, uint8_t* d_bits, uint32_t* d_iterations,
                            uint8_t* d_ok, int use_graph, void* stream) {
    return guarded([&] {
        if (!plan) fail(Kind::Config, "null plan");
        std::lock_guard<std::mutex> l(plan->mu);
        run_plan(*plan, d_llr, d_bits, d_iterations, d_ok, use_graph != 0, static_cast<hipStream_t>(stream));
    });
}

cf_status cf_decode_batch_uneven(const cf_code* const* codes, const uint64_t* llr_offsets,
                                 const uint32_t* max_iterations, size_t count, int msg_type, const float* d_llr,
                                 uint8_t* d_bits, uint32_t* d_iterations, uint8_t* d_ok, void* stream) {
    return guarded([&] {
        std::unique_ptr<cf_batch_plan> p(create_plan(codes, llr_offsets, max_iterations, count, msg_type));
        hipStream_t s = static_cast<hipStream_t>(stream);
        run_plan(*p, d_llr, d_bits, d_iterations, d_ok, false, s);
        // the plan's tables must outlive the launch
        check_hip(hipStreamSynchronize(s), "batch decode");
    });
}

}  // extern "C"
