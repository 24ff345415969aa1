// The Monte Carlo campaign engine of crates/com-sim, in C++: conditions of a grid, records
// of error rates and means with their intervals, stopping rules, and the runner that runs
// batches of trials of a scenario, merges their records and stops each condition by rule.
// A scenario runs a whole batch of trials at once (on the GPU for the NR link), trial t of
// condition c keyed by (seed, c, t), so the results never depend on the batch size.
#pragma once
#include <chrono>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "nr_code.hpp"

namespace comfec {
namespace sim {

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

// runner.rs:71-99
struct ConditionResult {
    size_t index = 0;
    Condition condition;
    Status status = Status::Completed;
    uint64_t trials = 0;
    double wall_s = 0.0;
    Record record;
};

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
    uint64_t seed_value() const { return seed_; }
    std::vector<ConditionResult> run(BatchScenario& scenario, const Grid& grid);

  private:
    uint64_t seed_ = 0, batch_ = 32;
    Stop stop_ = Stop::max_trials(100);
    double confidence_ = 0.95;
    bool fast_dev_run_ = false;
};

void check_confidence(double c, const std::string& what);
// Error::context (crates/com-core/src/error.rs:50-58)
Error with_context(const Error& e, const std::string& context);

}  // namespace sim
}  // namespace comfec
