// The C ABI of the runner's callbacks and loggers (include/comfec_gpu.h): the built-in ones
// of crates/com-sim/src/callbacks.rs and loggers.rs, and adapters that call a caller's
// function tables with C views of the run information, the progress and the results.
#include <memory>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace comfec;
using namespace comfec::detail;

namespace {

// RunInfo as a C struct whose strings point into `i`
cf_run_info c_info(const sim::RunInfo& i) {
    cf_run_info c{};
    c.scenario = i.scenario.c_str();
    c.seed = i.seed;
    c.batch = i.batch;
    c.threads = i.threads;
    c.stop = i.stop.c_str();
    c.confidence = i.confidence;
    c.n_conditions = i.n_conditions;
    c.started_unix_s = i.started_unix_s;
    c.version = i.version.c_str();
    c.commit = i.commit.c_str();
    c.config = i.has_config ? i.config.c_str() : nullptr;
    return c;
}

int c_action(int a) { return a >= CF_CONTINUE && a <= CF_STOP_RUN ? a : CF_CONTINUE; }

class CCallback : public sim::Callback {
  public:
    explicit CCallback(const cf_callback& f) : f_(f) {}
    void on_run_start(const sim::RunInfo& info) override {
        if (!f_.on_run_start) return;
        const cf_run_info c = c_info(info);
        f_.on_run_start(f_.user, &c);
    }
    sim::Action on_condition_start(size_t index, const sim::Condition& cond) override {
        if (!f_.on_condition_start) return sim::Action::Continue;
        std::vector<double> v;
        for (const auto& e : cond.entries) v.push_back(e.second);
        return sim::Action(c_action(f_.on_condition_start(f_.user, index, v.data(), v.size())));
    }
    sim::Action on_batch_end(const sim::Progress& p) override {
        if (!f_.on_batch_end) return sim::Action::Continue;
        cf_record view;
        view.view = p.record;
        const cf_progress c{p.condition, p.trials, p.elapsed_s, &view};
        return sim::Action(c_action(f_.on_batch_end(f_.user, &c)));
    }
    sim::Action on_condition_end(const sim::ConditionResult& r) override {
        if (!f_.on_condition_end) return sim::Action::Continue;
        cf_results one;
        one.results.push_back(r);
        return sim::Action(c_action(f_.on_condition_end(f_.user, &one)));
    }
    void on_run_end(const std::vector<sim::ConditionResult>& results) override {
        if (!f_.on_run_end) return;
        cf_results all;
        all.results = results;
        f_.on_run_end(f_.user, &all);
    }

  private:
    cf_callback f_;
};

class CLogger : public sim::Logger {
  public:
    explicit CLogger(const cf_logger& f) : f_(f) {}
    void start(const sim::RunInfo& info) override {
        if (!f_.start) return;
        const cf_run_info c = c_info(info);
        if (f_.start(f_.user, &c) != 0) fail(Kind::Io, "the logger failed to start");
    }
    void log(const sim::ConditionResult& r, double confidence) override {
        if (!f_.log) return;
        cf_results one;
        one.results.push_back(r);
        if (f_.log(f_.user, &one, confidence) != 0)
            fail(Kind::Io, "the logger failed on (" + r.condition.label() + ")");
    }
    void finish(const sim::RunReport& rep) override {
        if (!f_.finish) return;
        const cf_run_info c = c_info(rep.info);
        cf_results all;
        all.results = rep.results;
        if (f_.finish(f_.user, &c, &all, rep.wall_s) != 0) fail(Kind::Io, "the logger failed to finish");
    }

  private:
    cf_logger f_;
};

cf_hooks* checked(cf_hooks* h) {
    if (!h) fail(Kind::Config, "null hooks");
    return h;
}

}  // namespace

extern "C" {

cf_status cf_hooks_create(cf_hooks** out) {
    return guarded([&] {
        if (!out) fail(Kind::Config, "null output");
        *out = new cf_hooks();
    });
}

cf_status cf_hooks_add_callback(cf_hooks* h, const cf_callback* callback) {
    return guarded([&] {
        if (!callback) fail(Kind::Config, "null callback");
        checked(h)->callbacks.push_back(std::make_shared<CCallback>(*callback));
    });
}

cf_status cf_hooks_add_logger(cf_hooks* h, const cf_logger* logger) {
    return guarded([&] {
        if (!logger) fail(Kind::Config, "null logger");
        checked(h)->loggers.push_back(std::make_shared<CLogger>(*logger));
    });
}

cf_status cf_hooks_progress_bar(cf_hooks* h, double every_s) {
    return guarded([&] { checked(h)->callbacks.push_back(std::make_shared<sim::ProgressBar>(every_s)); });
}

cf_status cf_hooks_skip_after(cf_hooks* h, const char* metric, double threshold, const char* along) {
    return guarded([&] {
        if (!metric || !along) fail(Kind::Config, "null metric or key");
        checked(h)->callbacks.push_back(std::make_shared<sim::SkipAfter>(metric, threshold, along));
    });
}

cf_status cf_hooks_csv_logger(cf_hooks* h, const char* path) {
    return guarded([&] {
        if (!path) fail(Kind::Config, "null path");
        checked(h)->loggers.push_back(std::make_shared<sim::CsvLogger>(path));
    });
}

cf_status cf_hooks_json_logger(cf_hooks* h, const char* path) {
    return guarded([&] {
        if (!path) fail(Kind::Config, "null path");
        checked(h)->loggers.push_back(std::make_shared<sim::JsonLogger>(path));
    });
}

void cf_hooks_destroy(cf_hooks* h) { delete h; }

}  // extern "C"
