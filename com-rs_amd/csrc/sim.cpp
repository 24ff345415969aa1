// The campaign engine of crates/com-sim (see sim.hpp): statistics, records, grids,
// stopping rules and the runner loop.
#include "sim.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <ctime>
#include <limits>
#include <thread>

#include "nr_code.hpp"

namespace comfec {
namespace sim {

std::string display(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[400];
    auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::fixed);
    return std::string(buf, r.ptr);
}

// crates/com-math/src/special.rs:180-255: Cody's rational Chebyshev approximation (CALERF)
double erfc_cody(double x) {
    static const double A[5] = {3.16112374387056560e00, 1.13864154151050156e02, 3.77485237685302021e02,
                                3.20937758913846947e03, 1.85777706184603153e-1};
    static const double B[4] = {2.36012909523441209e01, 2.44024637934444173e02, 1.28261652607737228e03,
                                2.84423683343917062e03};
    static const double C[9] = {5.64188496988670089e-1, 8.88314979438837594e00, 6.61191906371416295e01,
                                2.98635138197400131e02, 8.81952221241769090e02, 1.71204761263407058e03,
                                2.05107837782607147e03, 1.23033935479799725e03, 2.15311535474403846e-8};
    static const double D[8] = {1.57449261107098347e01, 1.17693950891312499e02, 5.37181101862009858e02,
                                1.62138957456669019e03, 3.29079923573345963e03, 4.36261909014324716e03,
                                3.43936767414372164e03, 1.23033935480374942e03};
    static const double P[6] = {3.05326634961232344e-1, 3.60344899949804439e-1, 1.25781726111229246e-1,
                                1.60837851487422766e-2, 6.58749161529837803e-4, 1.63153871373020978e-2};
    static const double Q[5] = {2.56852019228982242e00, 1.87295284992346725e00, 5.27905102951428412e-1,
                                6.05183413124413191e-2, 2.33520497626869185e-3};
    const double y = std::fabs(x);
    if (y <= 0.46875) {
        const double ysq = y > 1.11e-16 ? y * y : 0.0;
        double num = A[4] * ysq, den = ysq;
        for (int i = 0; i < 3; ++i) {
            num = (num + A[i]) * ysq;
            den = (den + B[i]) * ysq;
        }
        return 1.0 - x * (num + A[3]) / (den + B[3]);
    }
    double r = 0.0;
    if (y <= 27.3) {
        double rr;
        if (y <= 4.0) {
            double num = C[8] * y, den = y;
            for (int i = 0; i < 7; ++i) {
                num = (num + C[i]) * y;
                den = (den + D[i]) * y;
            }
            rr = (num + C[7]) / (den + D[7]);
        } else {
            const double ysq = 1.0 / (y * y);
            double num = P[5] * ysq, den = ysq;
            for (int i = 0; i < 4; ++i) {
                num = (num + P[i]) * ysq;
                den = (den + Q[i]) * ysq;
            }
            rr = (5.6418958354775628695e-1 - ysq * (num + P[4]) / (den + Q[4])) / y;
        }
        const double u = std::trunc(y * 16.0) / 16.0;
        r = std::exp(-u * u) * std::exp(-(y - u) * (y + u)) * rr;
    }
    return x < 0.0 ? 2.0 - r : r;
}

double erfcinv(double y) {
    if (y <= 0.0) return std::numeric_limits<double>::infinity();
    if (y >= 2.0) return -std::numeric_limits<double>::infinity();
    double lo = -27.0, hi = 27.0;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (erfc_cody(mid) > y) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

double qfuncinv(double p) { return 1.41421356237309504880 * erfcinv(2.0 * p); }

double z_score(double confidence) {
    const double c = std::min(std::max(confidence, 0.0), 1.0 - 1e-16);
    return qfuncinv((1.0 - c) / 2.0);
}

double ErrorRate::rate() const {
    return total == 0 ? std::numeric_limits<double>::quiet_NaN() : double(errors) / double(total);
}

std::pair<double, double> ErrorRate::wilson(double confidence) const {
    if (total == 0) return {0.0, 1.0};
    const double n = double(total), p = rate(), z = z_score(confidence), z2 = z * z;
    const double centre = (p + z2 / (2.0 * n)) / (1.0 + z2 / n);
    const double half = z * std::sqrt(p * (1.0 - p) / n + z2 / (4.0 * n * n)) / (1.0 + z2 / n);
    const double lower = errors == 0 ? 0.0 : std::max(centre - half, 0.0);
    return {lower, std::min(centre + half, 1.0)};
}

void Mean::add(double x) {
    n += 1;
    const double d = x - mean;
    mean += d / double(n);
    m2 += d * (x - mean);
}

void Mean::merge(const Mean& o) {
    if (o.n == 0) return;
    if (n == 0) {
        *this = o;
        return;
    }
    const uint64_t t = n + o.n;
    const double d = o.mean - mean;
    mean += d * double(o.n) / double(t);
    m2 += o.m2 + d * d * double(n) * double(o.n) / double(t);
    n = t;
}

double Mean::value() const { return n == 0 ? std::numeric_limits<double>::quiet_NaN() : mean; }

double Mean::variance() const { return n < 2 ? std::numeric_limits<double>::quiet_NaN() : m2 / double(n - 1); }

std::pair<double, double> Mean::interval(double confidence) const {
    if (n < 2) return {-std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity()};
    const double h = z_score(confidence) * (std::sqrt(variance()) / std::sqrt(double(n)));
    return {mean - h, mean + h};
}

double Estimate::relative_precision(double confidence) const {
    auto [lo, hi] = interval(confidence);
    const double v = std::fabs(value());
    if (v == 0.0 || !std::isfinite(v)) return std::numeric_limits<double>::infinity();
    return 0.5 * (hi - lo) / v;
}

void Record::errors(const std::string& name, uint64_t e, uint64_t n) {
    auto it = metrics.find(name);
    if (it == metrics.end()) it = metrics.emplace(name, Estimate{}).first;
    if (it->second.is_rate) it->second.rate.add(e, n);
    else dropped[name] += 1;
}

void Record::value(const std::string& name, double v) {
    if (!std::isfinite(v)) {
        dropped[name] += 1;
        return;
    }
    auto it = metrics.find(name);
    if (it == metrics.end()) {
        Estimate e;
        e.is_rate = false;
        it = metrics.emplace(name, e).first;
    }
    if (!it->second.is_rate) it->second.mean.add(v);
    else dropped[name] += 1;
}

void Record::merge(const Record& o) {
    for (const auto& [k, v] : o.metrics) {
        auto it = metrics.find(k);
        if (it == metrics.end()) {
            metrics.emplace(k, v);
        } else if (it->second.is_rate && v.is_rate) {
            it->second.rate.add(v.rate.errors, v.rate.total);
        } else if (!it->second.is_rate && !v.is_rate) {
            it->second.mean.merge(v.mean);
        } else {
            dropped[k] += v.is_rate ? 1 : std::max<uint64_t>(v.mean.n, 1);
        }
    }
    for (const auto& [k, n] : o.dropped) dropped[k] += n;
}

const Estimate* Record::get(const std::string& name) const {
    auto it = metrics.find(name);
    return it == metrics.end() ? nullptr : &it->second;
}

double Condition::f64(const std::string& key) const {
    for (const auto& [k, v] : entries)
        if (k == key) return v;
    fail(Kind::Condition, "no key \"" + key + "\" in (" + label() + ")");
}

std::string Condition::label() const {
    std::string s;
    for (size_t i = 0; i < entries.size(); ++i) {
        if (i) s += ", ";
        s += entries[i].first + "=" + display(entries[i].second);
    }
    return s;
}

Grid& Grid::axis(const std::string& name, std::vector<double> values) {
    parts_.push_back(Part{name, std::move(values)});
    return *this;
}

Grid& Grid::fixed(const std::string& name, double value) {
    parts_.push_back(Part{name, {value}});
    return *this;
}

void Grid::check() const {
    std::vector<std::string> seen;
    for (const auto& p : parts_) {
        if (std::find(seen.begin(), seen.end(), p.name) != seen.end())
            fail(Kind::Config, "grid: the key \"" + p.name +
                                   "\" twice: the second value would replace the first and repeat conditions");
        seen.push_back(p.name);
    }
}

std::vector<Condition> Grid::conditions() const {
    std::vector<Condition> out(1);
    for (const auto& p : parts_) {
        std::vector<Condition> next;
        for (const auto& c : out)
            for (double v : p.values) {
                Condition d = c;
                d.entries.emplace_back(p.name, v);
                next.push_back(std::move(d));
            }
        out = std::move(next);
    }
    return out;
}

Stop Stop::max_trials(uint64_t n) {
    Stop s;
    s.kind = MaxTrials;
    s.count = n;
    return s;
}
Stop Stop::min_trials(uint64_t n) {
    Stop s;
    s.kind = MinTrials;
    s.count = n;
    return s;
}
Stop Stop::min_errors(const std::string& metric, uint64_t errors) {
    Stop s;
    s.kind = MinErrors;
    s.metric = metric;
    s.count = errors;
    return s;
}
Stop Stop::precision(const std::string& metric, double relative, double confidence) {
    Stop s;
    s.kind = Precision;
    s.metric = metric;
    s.value = relative;
    s.confidence = confidence;
    return s;
}
Stop Stop::max_time(double seconds) {
    Stop s;
    s.kind = MaxTime;
    s.value = seconds;
    return s;
}
Stop Stop::operator|(const Stop& o) const {
    Stop s;
    s.kind = Any;
    s.children = {*this, o};
    return s;
}
Stop Stop::operator&(const Stop& o) const {
    Stop s;
    s.kind = All;
    s.children = {*this, o};
    return s;
}

bool Stop::is_met(const Progress& p) const {
    switch (kind) {
        case MaxTrials:
        case MinTrials: return p.trials >= count;
        case MinErrors: {
            const Estimate* e = p.record->get(metric);
            return e && e->is_rate && e->rate.errors >= count;
        }
        case Precision: {
            const Estimate* e = p.record->get(metric);
            return e && e->relative_precision(confidence) <= value;
        }
        case MaxTime: return p.elapsed_s >= value;
        case Any:
            for (const auto& c : children)
                if (c.is_met(p)) return true;
            return false;
        case All:
            for (const auto& c : children)
                if (!c.is_met(p)) return false;
            return true;
    }
    return false;
}

bool Stop::is_bounded() const {
    switch (kind) {
        case MaxTrials:
        case MaxTime: return true;
        case MinTrials:
        case MinErrors:
        case Precision: return false;
        case Any:
            for (const auto& c : children)
                if (c.is_bounded()) return true;
            return false;
        case All:
            for (const auto& c : children)
                if (!c.is_bounded()) return false;
            return true;
    }
    return false;
}

bool Stop::trial_limit(uint64_t& n) const {
    switch (kind) {
        case MaxTrials:
        case MinTrials: n = count; return true;
        case MinErrors:
        case Precision:
        case MaxTime: return false;
        case Any: {
            bool any = false;
            for (const auto& c : children) {
                uint64_t m;
                if (c.trial_limit(m)) {
                    n = any ? std::min(n, m) : m;
                    any = true;
                }
            }
            return any;
        }
        case All: {
            uint64_t best = 0;
            for (const auto& c : children) {
                uint64_t m;
                if (!c.trial_limit(m)) return false;
                best = std::max(best, m);
            }
            n = best;
            return true;
        }
    }
    return false;
}

void Stop::targets(std::vector<std::pair<std::string, bool>>& out) const {
    if (kind == MinErrors) out.emplace_back(metric, true);
    if (kind == Precision) out.emplace_back(metric, false);
    for (const auto& c : children) c.targets(out);
}

void Stop::check_confidences() const {
    if (kind == Precision) check_confidence(confidence, "the stopping rule `" + str() + "`");
    for (const auto& c : children) c.check_confidences();
}

void Stop::check_metrics(const Record& record) const {
    std::vector<std::pair<std::string, bool>> t;
    targets(t);
    for (const auto& [metric, needs_rate] : t) {
        if (!record.knows(metric)) {
            std::string names;
            for (const auto& [k, v] : record.metrics) names += (names.empty() ? "" : ", ") + k;
            fail(comfec::Kind::Config, "the stopping rule `" + str() + "` targets the metric `" + metric +
                                   "`, which is not recorded (recorded: " + (names.empty() ? "none" : names) +
                                   "): it would never be met");
        }
        const Estimate* e = record.get(metric);
        if (needs_rate && e && !e->is_rate)
            fail(comfec::Kind::Config, "the stopping rule `" + str() + "`: min_errors needs an error rate, and `" + metric +
                                   "` is a mean (use precision)");
    }
}

std::string Stop::str() const {
    switch (kind) {
        case MaxTrials: return "max_trials(" + std::to_string(count) + ")";
        case MinTrials: return "min_trials(" + std::to_string(count) + ")";
        case MinErrors: return "errors(" + metric + ") ≥ " + std::to_string(count);
        case Precision: return "precision(" + metric + ") ≤ " + display(value) + " at " + display(confidence);
        case MaxTime: {
            char buf[64];
            std::snprintf(buf, sizeof buf, "time ≥ %.1f s", value);
            return buf;
        }
        case Any:
        case All: {
            std::string s = "(";
            for (size_t i = 0; i < children.size(); ++i) {
                if (i) s += kind == Any ? " or " : " and ";
                s += children[i].str();
            }
            return s + ")";
        }
    }
    return "";
}

Error with_context(const Error& e, const std::string& context) {
    // Error::context (crates/com-core/src/error.rs:50-58): the kind stays, the message gets
    // the context in front, after the kind's prefix
    const std::string m = e.what();
    const size_t at = m.find(": ");
    if (at == std::string::npos) return Error(e.kind, context + ": " + m);
    return Error(e.kind, m.substr(0, at + 2) + context + ": " + m.substr(at + 2));
}

void check_confidence(double c, const std::string& what) {
    if (!(c > 0.0 && c < 1.0))
        fail(Kind::Config,
             what + ": the confidence level must be between 0 and 1 (0.95 for 95 %), got " + display(c));
}

const char* status_str(Status s) {
    switch (s) {
        case Status::Completed: return "completed";
        case Status::Stopped: return "stopped";
        case Status::Skipped: return "skipped";
        case Status::DevRun: return "dev_run";
    }
    return "";
}

RunInfo Runner::run_info(const std::string& scenario, size_t n_conditions) const {
    RunInfo i;
    i.scenario = scenario;
    i.seed = seed_;
    i.batch = batch_;
    i.threads = threads_ ? threads_ : std::max(1u, std::thread::hardware_concurrency());
    i.stop = fast_dev_run_ ? "fast_dev_run" : stop_.str();
    i.confidence = confidence_;
    i.n_conditions = n_conditions;
    i.started_unix_s = uint64_t(std::max<std::time_t>(std::time(nullptr), 0));
    i.version = kVersion;
    i.commit = kCommit;
    i.has_config = has_config_;
    i.config = config_;
    return i;
}

// runner.rs:417-473: batches of trials until the rule is met (or a callback stops the
// condition), the last batch ending at the rule's bound on the trials
ConditionResult Runner::run_condition(BatchScenario& scenario, size_t index, const Condition& cond) {
    const auto t0 = std::chrono::steady_clock::now();
    try {
        scenario.setup(cond, index);
    } catch (const Error& e) {
        throw with_context(e, "setup of (" + cond.label() + ")");
    }
    ConditionResult res;
    res.index = index;
    res.condition = cond;
    uint64_t limit = 0;
    const bool bounded = !fast_dev_run_ && stop_.trial_limit(limit);
    uint64_t trials = 0;
    for (;;) {
        const uint64_t end = bounded ? std::max(std::min(trials + batch_, limit), trials + 1) : trials + batch_;
        try {
            scenario.run_batch(cond, index, trials, end, res.record);
        } catch (const Error& e) {
            throw with_context(e, "trials " + std::to_string(trials) + ".." + std::to_string(end) + " of (" +
                                      cond.label() + ")");
        }
        if (trials == 0) {
            try {
                stop_.check_metrics(res.record);
            } catch (const Error& e) {
                throw with_context(e, "(" + cond.label() + ")");
            }
        }
        trials = end;
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        Progress p{index, trials, &res.record, elapsed};
        // every callback sees the batch; Skip after a batch stops the condition too
        bool stopped = false;
        for (auto& c : callbacks_) {
            const Action a = c->on_batch_end(p);
            if (a == Action::StopCondition || a == Action::Skip) stopped = true;
        }
        if (fast_dev_run_) {
            res.status = Status::DevRun;
            break;
        }
        if (stop_.is_met(p)) {
            res.status = Status::Completed;
            break;
        }
        if (stopped) {
            res.status = Status::Stopped;
            break;
        }
    }
    res.trials = trials;
    res.wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return res;
}

// runner.rs:309-415: the callbacks, then the loggers, at the start of the run, at the end
// of every condition (run or skipped) and at the end of the run
std::vector<ConditionResult> Runner::run(BatchScenario& scenario, const Grid& grid) {
    check_confidence(confidence_, "the runner");
    stop_.check_confidences();
    grid.check();
    if (!fast_dev_run_ && !stop_.is_bounded())
        fail(Kind::Config, "the stopping rule `" + stop_.str() +
                               "` can run forever: add Stop::max_trials or Stop::max_time with `or`");
    const auto conds = grid.conditions();
    const RunInfo info = run_info(scenario.name(), conds.size());
    for (auto& c : callbacks_) c->on_run_start(info);
    for (auto& l : loggers_) l->start(info);
    const auto t_run = std::chrono::steady_clock::now();
    std::vector<ConditionResult> results;
    results.reserve(conds.size());
    bool stop_run = false;
    for (size_t index = 0; index < conds.size(); ++index) {
        const Condition& cond = conds[index];
        // every callback sees the start of the condition unless the run was stopped
        bool skip = stop_run;
        if (!stop_run)
            for (auto& c : callbacks_)
                if (c->on_condition_start(index, cond) == Action::Skip) skip = true;
        ConditionResult res;
        if (skip) {
            res.index = index;
            res.condition = cond;
            res.status = Status::Skipped;
        } else {
            res = run_condition(scenario, index, cond);
        }
        for (auto& c : callbacks_)
            if (c->on_condition_end(res) == Action::StopRun) stop_run = true;
        for (auto& l : loggers_) l->log(res, confidence_);
        results.push_back(std::move(res));
    }
    for (auto& c : callbacks_) c->on_run_end(results);
    RunReport report{info, std::move(results),
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t_run).count()};
    for (auto& l : loggers_) l->finish(report);
    return std::move(report.results);
}

}  // namespace sim
}  // namespace comfec
