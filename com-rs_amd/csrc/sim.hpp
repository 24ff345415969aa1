// The Monte Carlo campaign engine of crates/com-sim, in C++: conditions of a grid, records
// of error rates and means with their intervals, stopping rules, and the runner that runs
// batches of trials of a scenario, merges their records and stops each condition by rule.
// A scenario runs a whole batch of trials at once (on the GPU for the NR link), trial t of
// condition c keyed by (seed, c, t), so the results never depend on the batch size.
#pragma once
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "nr_code.hpp"

namespace comfec {
namespace sim {

// Rust's Display of an f64: the shortest decimal that reads back to the value, with no
// exponent and no trailing ".0" (grid labels, stopping rules, JSON numbers).
std::string display(double v);

// z = Q^-1(alpha/2) of a two-sided level (stats.rs:43-45), through Cody's erfc
// (special.rs:180-255) and its bisection inverse (special.rs:265-273, search.rs:69-80).
double erfc_cody(double x);
double erfcinv(double y);
double qfuncinv(double p);
double z_score(double confidence);

// stats.rs:48-88
struct ErrorRate {
    uint64_t errors = 0, total = 0;
    void add(uint64_t e, uint64_t n) {
        errors += e;
        total += n;
    }
    double rate() const;
    std::pair<double, double> wilson(double confidence) const;
};

// stats.rs:90-162 (Welford, merged by Chan et al.)
struct Mean {
    uint64_t n = 0;
    double mean = 0.0, m2 = 0.0;
    void add(double x);
    void merge(const Mean& o);
    double value() const;
    double variance() const;
    std::pair<double, double> interval(double confidence) const;
};

// stats.rs:165-236
struct Estimate {
    bool is_rate = true;
    ErrorRate rate;
    Mean mean;
    double value() const { return is_rate ? rate.rate() : mean.value(); }
    std::pair<double, double> interval(double confidence) const {
        return is_rate ? rate.wilson(confidence) : mean.interval(confidence);
    }
    double relative_precision(double confidence) const;
    uint64_t count() const { return is_rate ? rate.errors : mean.n; }
    uint64_t total() const { return is_rate ? rate.total : mean.n; }
};

// scenario.rs:66-198: metrics by name (sorted), and the values dropped per name
struct Record {
    std::map<std::string, Estimate> metrics;
    std::map<std::string, uint64_t> dropped;
    void errors(const std::string& name, uint64_t e, uint64_t n);
    void value(const std::string& name, double v);
    void merge(const Record& o);
    bool knows(const std::string& name) const { return metrics.count(name) || dropped.count(name); }
    const Estimate* get(const std::string& name) const;
};

// grid.rs: a condition is an ordered list of named values (numbers here)
struct Condition {
    std::vector<std::pair<std::string, double>> entries;
    double f64(const std::string& key) const;
    std::string label() const;
};

// grid.rs:231-361: axes (the last one fastest), fixed values, explicit conditions
class Grid {
  public:
    Grid& axis(const std::string& name, std::vector<double> values);
    Grid& fixed(const std::string& name, double value);
    void check() const;
    std::vector<Condition> conditions() const;

  private:
    struct Part {
        std::string name;
        std::vector<double> values;
    };
    std::vector<Part> parts_;
};

// stopping.rs:64-251
struct Progress {
    size_t condition;
    uint64_t trials;
    const Record* record;
    double elapsed_s;
};

class Stop {
  public:
    enum Kind { MaxTrials, MinTrials, MinErrors, Precision, MaxTime, Any, All };
    static Stop max_trials(uint64_t n);
    static Stop min_trials(uint64_t n);
    static Stop min_errors(const std::string& metric, uint64_t errors);
    static Stop precision(const std::string& metric, double relative, double confidence = 0.95);
    static Stop max_time(double seconds);
    Stop operator|(const Stop& o) const;  // Stop::or
    Stop operator&(const Stop& o) const;  // Stop::and
    bool is_met(const Progress& p) const;
    bool is_bounded() const;
    bool trial_limit(uint64_t& n) const;
    void check_confidences() const;
    void check_metrics(const Record& first_batch) const;
    std::string str() const;

    Kind kind = MaxTrials;
    uint64_t count = 0;
    double value = 0.0, confidence = 0.95;
    std::string metric;
    std::vector<Stop> children;

  private:
    void targets(std::vector<std::pair<std::string, bool>>& out) const;
};

// The scenario of a campaign: setup once per condition, then batches of trials.
class BatchScenario {
  public:
    virtual ~BatchScenario() = default;
    virtual std::string name() const = 0;
    virtual void setup(const Condition& cond, size_t index) = 0;
    // trials [t0, t1) of condition `index`, merged into the condition's record `rec` in
    // trial order (runner.rs:448-450), so that means come out as the per-trial merges of the
    // reference whatever the batch size
    virtual void run_batch(const Condition& cond, size_t index, uint64_t t0, uint64_t t1, Record& rec) = 0;
};

enum class Status { Completed = 0, Stopped = 1, Skipped = 2, DevRun = 3 };
const char* status_str(Status s);  // Status::as_str (runner.rs:70-77)

// runner.rs:71-99
struct ConditionResult {
    size_t index = 0;
    Condition condition;
    Status status = Status::Completed;
    uint64_t trials = 0;
    double wall_s = 0.0;
    Record record;
};

// ConditionResult::row (runner.rs:106-137): index, the condition, status, trials, wall_s,
// then per metric (in name order) the estimate, its interval and its two counts
struct Cell {
    enum Type { F64, I64, Str } type = F64;
    double f = 0.0;
    int64_t i = 0;
    std::string s;
};
std::vector<std::pair<std::string, Cell>> row(const ConditionResult& r, double confidence);

// RunInfo (runner.rs:150-175)
struct RunInfo {
    std::string scenario;
    uint64_t seed = 0, batch = 0;
    size_t threads = 0;
    std::string stop;
    double confidence = 0.95;
    size_t n_conditions = 0;
    uint64_t started_unix_s = 0;
    std::string version, commit;
    bool has_config = false;
    std::string config;  // a JSON document
};

// RunReport (runner.rs:177-186)
struct RunReport {
    RunInfo info;
    std::vector<ConditionResult> results;
    double wall_s = 0.0;
};

// Action (callbacks.rs:50-61)
enum class Action { Continue = 0, Skip = 1, StopCondition = 2, StopRun = 3 };

// Callback (callbacks.rs:64-81): every hook optional
class Callback {
  public:
    virtual ~Callback() = default;
    virtual void on_run_start(const RunInfo&) {}
    virtual Action on_condition_start(size_t, const Condition&) { return Action::Continue; }
    virtual Action on_batch_end(const Progress&) { return Action::Continue; }
    virtual Action on_condition_end(const ConditionResult&) { return Action::Continue; }
    virtual void on_run_end(const std::vector<ConditionResult>&) {}
};

// Logger (loggers.rs:53-82): failures are Kind::Io errors that end the run
class Logger {
  public:
    virtual ~Logger() = default;
    virtual void start(const RunInfo&) {}
    virtual void log(const ConditionResult& r, double confidence) = 0;
    virtual void finish(const RunReport&) {}
};

// ProgressBar (callbacks.rs:85-166): a line per condition on the standard error, and a
// progress line within a condition every `every_s` seconds
class ProgressBar : public Callback {
  public:
    explicit ProgressBar(double every_s = 5.0, double confidence = 0.95) : every_s_(every_s), confidence_(confidence) {}
    void on_run_start(const RunInfo& info) override;
    Action on_condition_start(size_t, const Condition&) override;
    Action on_batch_end(const Progress& p) override;
    Action on_condition_end(const ConditionResult& r) override;

  private:
    double every_s_, confidence_;
    size_t n_ = 0;
    bool started_ = false;
    std::chrono::steady_clock::time_point last_;
};

// SkipAfter::below (callbacks.rs:168-203): once `metric` falls below `threshold`, the later
// conditions equal to that one on every key but `along` are skipped
class SkipAfter : public Callback {
  public:
    SkipAfter(std::string metric, double threshold, std::string along)
        : metric_(std::move(metric)), threshold_(threshold), along_(std::move(along)) {}
    void on_run_start(const RunInfo&) override { done_.clear(); }
    Action on_condition_start(size_t, const Condition& c) override;
    Action on_condition_end(const ConditionResult& r) override;

  private:
    std::vector<std::pair<std::string, double>> others(const Condition& c) const;
    std::string metric_;
    double threshold_;
    std::string along_;
    std::vector<std::vector<std::pair<std::string, double>>> done_;
};

// CsvLogger (loggers.rs:107-215): one row per condition as it ends, flushed
class CsvLogger : public Logger {
  public:
    explicit CsvLogger(std::string path) : path_(std::move(path)) {}
    ~CsvLogger() override;
    void start(const RunInfo&) override;
    void log(const ConditionResult& r, double confidence) override;
    void finish(const RunReport&) override;

  private:
    void write_header(const ConditionResult& r, double confidence);
    void write_row(const ConditionResult& r, double confidence);
    void put(const std::string& line);
    std::string path_;
    std::FILE* out_ = nullptr;
    bool header_ = false;
    std::vector<std::string> keys_, columns_;
    std::vector<std::pair<ConditionResult, double>> pending_;
};

// JsonLogger (loggers.rs:221-318): the whole report at the end of the run
class JsonLogger : public Logger {
  public:
    explicit JsonLogger(std::string path) : path_(std::move(path)) {}
    void log(const ConditionResult&, double) override {}
    void finish(const RunReport& rep) override;

  private:
    std::string path_;
};

// json.rs and loggers.rs helpers
std::string csv_field(const std::string& s);
std::string json_string(const std::string& s);
std::string json_number(double v);
// Rust's LowerExp of an f64 (`{:e}`: shortest digits; `{:.Ne}`: N digits after the point)
std::string lower_exp(double v, int precision = -1);

// The library's version and the commit it was built from (`<hash>`, `<hash>-dirty`, or
// `unknown`), recorded in every run (crates/com-sim/src/lib.rs:81-88)
extern const char* const kVersion;
extern const char* const kCommit;

// runner.rs:188-473
class Runner {
  public:
    Runner& seed(uint64_t s) {
        seed_ = s;
        return *this;
    }
    Runner& batch(uint64_t n) {
        batch_ = n < 1 ? 1 : n;
        return *this;
    }
    Runner& stop(Stop s) {
        stop_ = std::move(s);
        return *this;
    }
    Runner& confidence(double c) {
        confidence_ = c;
        return *this;
    }
    Runner& fast_dev_run(bool on) {
        fast_dev_run_ = on;
        return *this;
    }
    // worker threads reported in the run information (0: the host's cores); the trials of
    // a batch run wherever the scenario runs them
    Runner& threads(size_t n) {
        threads_ = n;
        return *this;
    }
    Runner& config(std::string json) {
        config_ = std::move(json);
        has_config_ = true;
        return *this;
    }
    Runner& callback(std::shared_ptr<Callback> c) {
        callbacks_.push_back(std::move(c));
        return *this;
    }
    Runner& logger(std::shared_ptr<Logger> l) {
        loggers_.push_back(std::move(l));
        return *this;
    }
    uint64_t seed_value() const { return seed_; }
    RunInfo run_info(const std::string& scenario, size_t n_conditions) const;  // runner.rs:278-295
    std::vector<ConditionResult> run(BatchScenario& scenario, const Grid& grid);

  private:
    ConditionResult run_condition(BatchScenario& scenario, size_t index, const Condition& cond);
    uint64_t seed_ = 0, batch_ = 32;
    Stop stop_ = Stop::max_trials(100);
    double confidence_ = 0.95;
    bool fast_dev_run_ = false;
    size_t threads_ = 0;
    bool has_config_ = false;
    std::string config_;
    std::vector<std::shared_ptr<Callback>> callbacks_;
    std::vector<std::shared_ptr<Logger>> loggers_;
};

void check_confidence(double c, const std::string& what);
// Error::context (crates/com-core/src/error.rs:50-58)
Error with_context(const Error& e, const std::string& context);

}  // namespace sim
}  // namespace comfec
