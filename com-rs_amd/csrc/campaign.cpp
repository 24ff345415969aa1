// The C ABI of the device trial generator and of the campaigns (include/comfec_gpu.h):
// the runner of crates/com-sim/src/runner.rs driving either a caller's scenario (a
// callback) or the NR link on the GPU, whose batches are one generator launch, one decoder
// launch and one error-count launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <iterator>
#include <memory>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"
#include "nr_code.hpp"
#include "sim.hpp"

using namespace comfec;
using namespace comfec::detail;

namespace {

// Gray square QAM of 2^bps points, unit average energy (constellation.rs:168-205, 396-398):
// the natural order walks the rows from the top, left to right on even rows and right to
// left on odd ones, and point gray(i) of the labelled constellation is natural point i.
void qam_points(int bps, double* re, double* im) {
    const int m = 1 << bps;
    const int l = int(std::lround(std::sqrt(double(m))));
    std::vector<double> nr(m), ni(m);
    int p = 0;
    for (int row = 0; row < l; ++row) {
        const double y = 2.0 * double(l - 1 - row) - (double(l) - 1.0);
        for (int c = 0; c < l; ++c) {
            const int cc = row % 2 == 1 ? l - 1 - c : c;
            nr[p] = 2.0 * double(cc) - (double(l) - 1.0);
            ni[p] = y;
            ++p;
        }
    }
    for (int i = 0; i < m; ++i) {
        re[i ^ (i >> 1)] = nr[i];
        im[i ^ (i >> 1)] = ni[i];
    }
    double energy = 0.0;
    for (int i = 0; i < m; ++i) energy += re[i] * re[i] + im[i] * im[i];
    const double s = 1.0 / std::sqrt(energy / double(m));
    for (int i = 0; i < m; ++i) {
        re[i] *= s;
        im[i] *= s;
    }
}

}  // namespace

namespace comfec {
namespace detail {

TrialGenArgs make_generator_args(const NrCode& c, int device, int modulation, int snr_kind, double snr_db,
                                 uint64_t seed, uint32_t point) {
    if (c.bg == 0) fail(Kind::Config, "the device trial generator draws NR codewords; this code is given by its rows");
    if (modulation != CF_BPSK && modulation != CF_QPSK && modulation != CF_QAM16 && modulation != CF_QAM64)
        fail(Kind::Config, "the modulation must be BPSK, QPSK, 16-QAM or 64-QAM, got " + std::to_string(modulation));
    if (snr_kind != CF_EBN0_DB && snr_kind != CF_ESN0_DB)
        fail(Kind::Config, "the SNR is Eb/N0 or Es/N0, got kind " + std::to_string(snr_kind));
    if (c.e % uint32_t(modulation) != 0)
        fail(Kind::Config, "E = " + std::to_string(c.e) + " is not a whole number of symbols of " +
                               std::to_string(modulation) + " bits");
    const NrCode::DeviceTables& t = c.device_tables(device);
    TrialGenArgs a{};
    a.z = int32_t(c.z);
    a.n = int32_t(c.n);
    a.k = int32_t(c.k);
    a.k_info = int32_t(c.k_info);
    a.e = int32_t(c.e);
    a.rows_b = int32_t(c.rows_b);
    a.core_words = int32_t(c.core_words);
    a.row_start = t.row_start;
    a.col = t.col;
    a.shift = t.shift;
    a.map = t.map;
    a.core_inv = t.core_inv;
    a.bits_per_symbol = modulation;
    const double lin = std::pow(10.0, snr_db / 10.0);  // db2lin (units.rs:38-40)
    const double rate = double(c.k_info) / double(c.e);
    if (modulation == CF_BPSK) {
        // bpsk_noise_variance (bounds.rs:58-60) per real dimension
        const double s2 = snr_kind == CF_EBN0_DB ? 1.0 / (2.0 * rate * lin) : 1.0 / (2.0 * lin);
        a.noise_var = float(s2);
        a.noise_sd = float(std::sqrt(s2));
    } else {
        // N0 of unit-energy symbols: ebn0_noise_variance (awgn.rs:139-141) with bits = R log2 M
        const double n0 = snr_kind == CF_EBN0_DB ? 1.0 / (double(modulation) * rate * lin) : 1.0 / lin;
        a.noise_var = float(n0);
        a.noise_sd = float(std::sqrt(n0 / 2.0));
        a.log_prior = float(-std::log(double(1 << modulation)));
        double re[64], im[64];
        qam_points(modulation, re, im);
        for (int i = 0; i < (1 << modulation); ++i) {
            a.pts_re[i] = float(re[i]);
            a.pts_im[i] = float(im[i]);
        }
    }
    a.k0 = uint32_t(seed);
    a.k1 = uint32_t(seed >> 32);
    a.point = point;
    return a;
}

}  // namespace detail
}  // namespace comfec

namespace {

cf_metric metric_of(const sim::Record& rec, const char* name, double confidence) {
    const sim::Estimate* e = rec.get(name);
    if (!e) fail(Kind::Condition, std::string("no metric `") + name + "` in the result");
    auto [lo, hi] = e->interval(confidence);
    auto d = rec.dropped.find(name);
    return cf_metric{e->value(), lo, hi, e->count(), e->total(), e->is_rate ? 1 : 0,
                     d == rec.dropped.end() ? 0 : d->second};
}

sim::Stop stop_from_nodes(const cf_stop_node* nodes, size_t n, size_t& pos) {
    if (pos >= n) fail(Kind::Config, "the stopping rule ends early");
    const cf_stop_node& s = nodes[pos++];
    switch (s.kind) {
        case CF_STOP_MAX_TRIALS: return sim::Stop::max_trials(s.count);
        case CF_STOP_MIN_TRIALS: return sim::Stop::min_trials(s.count);
        case CF_STOP_MIN_ERRORS: return sim::Stop::min_errors(s.metric ? s.metric : "", s.count);
        case CF_STOP_PRECISION: return sim::Stop::precision(s.metric ? s.metric : "", s.value, s.confidence);
        case CF_STOP_MAX_TIME: return sim::Stop::max_time(s.value);
        case CF_STOP_ANY:
        case CF_STOP_ALL: {
            sim::Stop out;
            out.kind = s.kind == CF_STOP_ANY ? sim::Stop::Any : sim::Stop::All;
            for (uint32_t i = 0; i < s.n_children; ++i) out.children.push_back(stop_from_nodes(nodes, n, pos));
            return out;
        }
        default: fail(Kind::Config, "unknown stopping rule kind " + std::to_string(s.kind));
    }
}

sim::Runner make_runner(const cf_runner_config* config, const cf_stop_node* stop, size_t n_stop) {
    if (!config) fail(Kind::Config, "null runner configuration");
    sim::Runner r;
    r.seed(config->seed).batch(config->batch).confidence(config->confidence).fast_dev_run(config->fast_dev_run != 0);
    r.threads(size_t(config->threads));
    if (config->config_json) r.config(config->config_json);
    if (config->hooks) {
        for (const auto& c : config->hooks->callbacks) r.callback(c);
        for (const auto& l : config->hooks->loggers) r.logger(l);
    }
    if (stop && n_stop) {
        size_t pos = 0;
        r.stop(stop_from_nodes(stop, n_stop, pos));
        if (pos != n_stop) fail(Kind::Config, "the stopping rule has nodes left over");
    }
    return r;
}

sim::Grid make_grid(const cf_grid_axis* axes, size_t n_axes) {
    sim::Grid g;
    for (size_t i = 0; i < n_axes; ++i)
        g.axis(axes[i].name ? axes[i].name : "", std::vector<double>(axes[i].values, axes[i].values + axes[i].n_values));
    return g;
}

// A scenario given as a C callback.
class CallbackScenario : public sim::BatchScenario {
  public:
    CallbackScenario(std::string name, cf_batch_fn fn, void* user) : name_(std::move(name)), fn_(fn), user_(user) {}
    std::string name() const override { return name_; }
    void setup(const sim::Condition&, size_t) override {}
    void run_batch(const sim::Condition& cond, size_t index, uint64_t t0, uint64_t t1, sim::Record& rec) override {
        std::vector<double> values;
        for (const auto& e : cond.entries) values.push_back(e.second);
        // the callback's record of the batch, merged as one
        cf_record r;
        const int rc = fn_(user_, index, values.data(), values.size(), t0, t1, &r);
        if (rc != 0 || r.failed) {
            const Kind k = r.kind >= 1 && r.kind <= 6 ? Kind(r.kind) : Kind::Config;
            fail(k, r.message.empty() ? "the scenario failed" : r.message);
        }
        rec.merge(r.rec);
    }

  private:
    std::string name_;
    cf_batch_fn fn_;
    void* user_;
};

template <typename T>
struct DeviceArray {
    T* p = nullptr;
    explicit DeviceArray(size_t n) { check_hip(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc"); }
    ~DeviceArray() {
        if (p) (void)hipFree(p);
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
};

// The NR link on the GPU, one batch of trials per launch sequence.
class NrLinkScenario : public sim::BatchScenario {
  public:
    NrLinkScenario(const cf_code* h, const cf_nr_link& link, uint64_t seed, uint64_t batch)
        : h_(h), link_(link), seed_(seed), cap_(std::min<uint64_t>(batch, 1u << 16)), device_(current_device()),
          llr_(size_t(cap_) * h->code.e), info_(size_t(cap_) * ((h->code.k_info + 31) / 32)),
          bits_(size_t(cap_) * h->code.k_info), its_(cap_), counters_(4) {
        check_hip(hipStreamCreate(&stream_), "hipStreamCreate");
        if (link_.d_learned_params && !layered_supported(h_->code.bg, int(h_->code.z)))
            fail(Kind::Config, "the learned decoder runs the layered min-sum schedule on a supported lifting size");
        if (link_.d_learned_params && link_.fallback && !fast_layered(h_))
            fail(Kind::Config, "the fallback runs the code's layered normalized or offset min-sum");
    }
    ~NrLinkScenario() override { (void)hipStreamDestroy(stream_); }
    std::string name() const override { return "nr_ldpc"; }
    void setup(const sim::Condition& cond, size_t) override {
        key_ = link_.snr_kind == CF_EBN0_DB ? "ebn0_db" : "esn0_db";
        snr_ = cond.f64(key_);
    }
    void run_batch(const sim::Condition&, size_t index, uint64_t t0, uint64_t t1, sim::Record& rec) override {
        const NrCode& c = h_->code;
        uint64_t counts[4] = {0, 0, 0, 0};
        std::vector<uint32_t> its(size_t(std::min<uint64_t>(cap_, t1 - t0)));
        for (uint64_t s = t0; s < t1; s += cap_) {
            const uint64_t n = std::min<uint64_t>(cap_, t1 - s);
            TrialGenArgs g = make_generator_args(c, device_, link_.modulation, link_.snr_kind, snr_, seed_, uint32_t(index));
            g.trial0 = s;
            g.count = int64_t(n);
            g.llr = llr_.p;
            g.info = info_.p;
            check_hip(launch_generate(g, stream_), "generator launch");
            if (fast_layered(h_) || link_.d_learned_params) {
                LayeredRun r;
                r.llr = llr_.p;
                r.count = int64_t(n);
                r.bits = bits_.p;
                r.iterations = its_.p;
                r.params = link_.d_learned_params;
                r.learned_iterations = link_.learned_iterations ? link_.learned_iterations
                                                                : h_->settings().max_iterations;
                r.fallback = link_.d_learned_params && link_.fallback;
                run_layered(h_, device_, r, stream_);
            } else {
                check_hip(cf_decode_batch(h_, CF_MSG_F32, llr_.p, n, bits_.p, nullptr, its_.p, stream_) == CF_OK
                              ? hipSuccess
                              : hipErrorUnknown,
                          g_error.c_str());
            }
            check_hip(launch_zero(counters_.p, 4, stream_), "zero launch");
            CountArgs ca{};
            ca.k_info = int32_t(c.k_info);
            ca.bits = bits_.p;
            ca.info = info_.p;
            ca.iterations = its_.p;
            ca.count = int64_t(n);
            ca.counters = counters_.p;
            check_hip(launch_count(ca, stream_), "count launch");
            uint64_t part[4];
            check_hip(hipMemcpyAsync(part, counters_.p, sizeof part, hipMemcpyDeviceToHost, stream_), "copy");
            check_hip(hipMemcpyAsync(its.data(), its_.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
                      "copy");
            check_hip(hipStreamSynchronize(stream_), "batch");
            for (int q = 0; q < 4; ++q) counts[q] += part[q];
            // the iteration count of every trial, merged in trial order as one-value means
            for (uint64_t i = 0; i < n; ++i) {
                sim::Record one;
                one.value("iterations", double(its[i]));
                rec.merge(one);
            }
        }
        const uint64_t trials = t1 - t0;
        // error counts are integer sums: the batch total is the sum of the per-trial records
        rec.errors("ber", counts[0], trials * c.k_info);
        rec.errors("fer", counts[1], trials);
    }

  private:
    const cf_code* h_;
    cf_nr_link link_;
    uint64_t seed_, cap_;
    int device_;
    std::string key_;
    double snr_ = 0.0;
    hipStream_t stream_ = nullptr;
    DeviceArray<float> llr_;
    DeviceArray<uint32_t> info_;
    DeviceArray<uint8_t> bits_;
    DeviceArray<uint32_t> its_;
    DeviceArray<uint64_t> counters_;
};

}  // namespace

extern "C" {

cf_status cf_constellation_points(int modulation, double* re_host, double* im_host, size_t m) {
    return guarded([&] {
        if (modulation != CF_QPSK && modulation != CF_QAM16 && modulation != CF_QAM64)
            fail(Kind::Config, "the constellation is QPSK, 16-QAM or 64-QAM, got " + std::to_string(modulation) +
                                   " bits per symbol");
        expect_len(m, size_t(1) << modulation, "points");
        if (!re_host || !im_host) fail(Kind::Config, "null output buffer");
        qam_points(modulation, re_host, im_host);
    });
}

cf_status cf_generate_batch(const cf_code* h, int modulation, int snr_kind, double snr_db, uint64_t seed,
                            uint32_t point, uint64_t trial0, size_t count, float* d_llr, uint32_t* d_info,
                            void* stream) {
    return guarded([&] {
        if (!h) fail(Kind::Config, "null handle");
        if (!d_llr) fail(Kind::Config, "null LLR buffer");
        TrialGenArgs a = make_generator_args(h->code, current_device(), modulation, snr_kind, snr_db, seed, point);
        a.trial0 = trial0;
        a.count = int64_t(count);
        a.llr = d_llr;
        a.info = d_info;
        check_hip(launch_generate(a, static_cast<hipStream_t>(stream)), "generator launch");
    });
}

cf_status cf_record_errors(cf_record* rec, const char* name, uint64_t errors, uint64_t total) {
    return guarded([&] {
        if (!rec || !name) fail(Kind::Config, "null record or name");
        if (rec->view) fail(Kind::Config, "the record of a progress report is read-only");
        rec->rec.errors(name, errors, total);
    });
}

cf_status cf_record_value(cf_record* rec, const char* name, double value) {
    return guarded([&] {
        if (!rec || !name) fail(Kind::Config, "null record or name");
        if (rec->view) fail(Kind::Config, "the record of a progress report is read-only");
        rec->rec.value(name, value);
    });
}

cf_status cf_record_fail(cf_record* rec, int kind, const char* message) {
    return guarded([&] {
        if (!rec) fail(Kind::Config, "null record");
        rec->failed = true;
        rec->kind = kind;
        rec->message = message ? message : "";
    });
}

cf_status cf_run(const cf_runner_config* config, const cf_stop_node* stop, size_t n_stop, const cf_grid_axis* axes,
                 size_t n_axes, const char* scenario_name, cf_batch_fn scenario, void* user, cf_results** out) {
    return guarded([&] {
        if (!scenario || !out) fail(Kind::Config, "null scenario or output");
        sim::Runner r = make_runner(config, stop, n_stop);
        CallbackScenario s(scenario_name ? scenario_name : "", scenario, user);
        auto res = std::make_unique<cf_results>();
        res->results = r.run(s, make_grid(axes, n_axes));
        *out = res.release();
    });
}

cf_status cf_nr_campaign(const cf_code* h, const cf_nr_link* link, const cf_runner_config* config,
                         const cf_stop_node* stop, size_t n_stop, const cf_grid_axis* axes, size_t n_axes,
                         cf_results** out) {
    return guarded([&] {
        if (!h || !link || !out) fail(Kind::Config, "null handle, link or output");
        sim::Runner r = make_runner(config, stop, n_stop);
        NrLinkScenario s(h, *link, config->seed, config->batch);
        auto res = std::make_unique<cf_results>();
        res->results = r.run(s, make_grid(axes, n_axes));
        *out = res.release();
    });
}

size_t cf_results_count(const cf_results* res) { return res ? res->results.size() : 0; }

cf_status cf_results_condition(const cf_results* res, size_t i, int* status, uint64_t* trials, double* wall_s) {
    return guarded([&] {
        if (!res || i >= res->results.size()) fail(Kind::Config, "no condition " + std::to_string(i));
        const auto& r = res->results[i];
        if (status) *status = int(r.status);
        if (trials) *trials = r.trials;
        if (wall_s) *wall_s = r.wall_s;
    });
}

cf_status cf_results_metric(const cf_results* res, size_t i, const char* name, double confidence, cf_metric* out) {
    return guarded([&] {
        if (!res || i >= res->results.size()) fail(Kind::Config, "no condition " + std::to_string(i));
        if (!name || !out) fail(Kind::Config, "null name or output");
        *out = metric_of(res->results[i].record, name, confidence);
    });
}

cf_status cf_record_metric(const cf_record* rec, const char* name, double confidence, cf_metric* out) {
    return guarded([&] {
        if (!rec || !name || !out) fail(Kind::Config, "null record, name or output");
        *out = metric_of(rec->get(), name, confidence);
    });
}

cf_status cf_results_values(const cf_results* res, size_t i, size_t* index, double* values, size_t cap,
                            size_t* n_values) {
    return guarded([&] {
        if (!res || i >= res->results.size()) fail(Kind::Config, "no condition " + std::to_string(i));
        const auto& r = res->results[i];
        if (index) *index = r.index;
        const auto& e = r.condition.entries;
        for (size_t j = 0; j < e.size() && j < cap && values; ++j) values[j] = e[j].second;
        if (n_values) *n_values = e.size();
    });
}

cf_status cf_results_metric_count(const cf_results* res, size_t i, size_t* n) {
    return guarded([&] {
        if (!res || i >= res->results.size()) fail(Kind::Config, "no condition " + std::to_string(i));
        if (!n) fail(Kind::Config, "null output");
        *n = res->results[i].record.metrics.size();
    });
}

cf_status cf_results_metric_name(const cf_results* res, size_t i, size_t j, const char** name) {
    return guarded([&] {
        if (!res || i >= res->results.size()) fail(Kind::Config, "no condition " + std::to_string(i));
        const auto& m = res->results[i].record.metrics;
        if (!name || j >= m.size()) fail(Kind::Config, "no metric " + std::to_string(j));
        *name = std::next(m.begin(), std::ptrdiff_t(j))->first.c_str();
    });
}

void cf_results_free(cf_results* res) { delete res; }

cf_status cf_wilson(uint64_t errors, uint64_t total, double confidence, double* lo, double* hi) {
    return guarded([&] {
        if (!lo || !hi) fail(Kind::Config, "null output");
        if (errors > total) fail(Kind::Config, "more errors than trials");
        sim::check_confidence(confidence, "the interval");
        sim::ErrorRate r{errors, total};
        auto [l, h] = r.wilson(confidence);
        *lo = l;
        *hi = h;
    });
}

}  // extern "C"
