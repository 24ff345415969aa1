// The observers of a campaign (crates/com-sim/src/callbacks.rs, loggers.rs, json.rs): the
// progress display, the skip-after-a-floor rule, the CSV rows written as conditions end and
// the JSON report of the whole run. Host code only.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <limits>

#include "sim.hpp"

#ifndef COMFEC_COMMIT
#define COMFEC_COMMIT "unknown"
#endif

namespace comfec {
namespace sim {

const char* const kVersion = "0.1.0";
const char* const kCommit = COMFEC_COMMIT;

namespace {

// printf's exponent ("e-05", "e+00") in Rust's form ("e-5", "e0")
std::string rust_exponent(const std::string& s) {
    const size_t at = s.find('e');
    if (at == std::string::npos) return s;
    const bool neg = s[at + 1] == '-';
    size_t d = at + 1;
    if (s[d] == '-' || s[d] == '+') ++d;
    while (d + 1 < s.size() && s[d] == '0') ++d;
    return s.substr(0, at) + "e" + (neg ? "-" : "") + s.substr(d);
}

std::string fixed(double v, int digits) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[400];
    std::snprintf(buf, sizeof buf, "%.*f", digits, v);
    return buf;
}

std::string cell_text(const Cell& c) {
    switch (c.type) {
        case Cell::F64: return display(c.f);
        case Cell::I64: return std::to_string(c.i);
        case Cell::Str: return c.s;
    }
    return "";
}

[[noreturn]] void io_fail(const std::string& path, int err) {
    fail(Kind::Io, path + ": " + std::strerror(err) + " (os error " + std::to_string(err) + ")");
}

// create_parent_dir (loggers.rs:92-105)
void create_parent_dir(const std::string& path) {
    const std::filesystem::path dir = std::filesystem::path(path).parent_path();
    if (dir.empty()) return;
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) io_fail(dir.string(), ec.value());
}

std::string json_object(const std::vector<std::pair<std::string, std::string>>& fields) {
    std::string s = "{";
    for (size_t i = 0; i < fields.size(); ++i) {
        if (i) s += ", ";
        s += json_string(fields[i].first) + ": " + fields[i].second;
    }
    return s + "}";
}

std::string json_array(const std::vector<std::string>& items) {
    std::string s = "[";
    for (size_t i = 0; i < items.size(); ++i) {
        if (i) s += ", ";
        s += items[i];
    }
    return s + "]";
}

// loggers.rs:254-269
std::string estimate_json(const Estimate& e, double confidence) {
    const auto [lo, hi] = e.interval(confidence);
    std::vector<std::pair<std::string, std::string>> f = {
        {"value", json_number(e.value())}, {"lo", json_number(lo)}, {"hi", json_number(hi)}};
    if (e.is_rate) {
        f.push_back({"kind", json_string("rate")});
        f.push_back({"errors", std::to_string(e.rate.errors)});
        f.push_back({"total", std::to_string(e.rate.total)});
    } else {
        f.push_back({"kind", json_string("mean")});
        f.push_back({"count", std::to_string(e.mean.n)});
        f.push_back({"std", json_number(std::sqrt(e.mean.variance()))});
    }
    return json_object(f);
}

int64_t count_cell(uint64_t n) {
    return n > uint64_t(std::numeric_limits<int64_t>::max()) ? std::numeric_limits<int64_t>::max() : int64_t(n);
}

}  // namespace

std::string lower_exp(double v, int precision) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    if (precision < 0) {
        auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
        return rust_exponent(std::string(buf, r.ptr));
    }
    std::snprintf(buf, sizeof buf, "%.*e", precision, v);
    return rust_exponent(buf);
}

// loggers.rs:84-90 (RFC 4180)
std::string csv_field(const std::string& s) {
    if (s.find_first_of(",\"\n") == std::string::npos) return s;
    std::string out = "\"";
    for (char c : s) {
        if (c == '"') out += '"';
        out += c;
    }
    return out + "\"";
}

// json.rs: string escapes and numbers (null when not finite)
std::string json_string(const std::string& s) {
    std::string out = "\"";
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", unsigned(c));
                    out += buf;
                } else {
                    out += char(c);
                }
        }
    }
    return out + "\"";
}

std::string json_number(double v) { return std::isfinite(v) ? display(v) : "null"; }

std::vector<std::pair<std::string, Cell>> row(const ConditionResult& r, double confidence) {
    auto f64 = [](double v) {
        Cell c;
        c.type = Cell::F64;
        c.f = v;
        return c;
    };
    auto i64 = [](uint64_t n) {
        Cell c;
        c.type = Cell::I64;
        c.i = count_cell(n);
        return c;
    };
    std::vector<std::pair<std::string, Cell>> out;
    out.push_back({"index", i64(r.index)});
    for (const auto& [k, v] : r.condition.entries) out.push_back({k, f64(v)});
    Cell st;
    st.type = Cell::Str;
    st.s = status_str(r.status);
    out.push_back({"status", st});
    out.push_back({"trials", i64(r.trials)});
    out.push_back({"wall_s", f64(r.wall_s)});
    for (const auto& [name, e] : r.record.metrics) {
        const auto [lo, hi] = e.interval(confidence);
        out.push_back({name, f64(e.value())});
        out.push_back({name + "_lo", f64(lo)});
        out.push_back({name + "_hi", f64(hi)});
        out.push_back({name + "_count", i64(e.count())});
        out.push_back({name + "_total", i64(e.total())});
    }
    return out;
}

// ---- ProgressBar (callbacks.rs:107-166) -----------------------------------------------

void ProgressBar::on_run_start(const RunInfo& info) {
    n_ = info.n_conditions;
    std::fprintf(stderr, "[%s] %zu conditions, seed %llu, batches of %llu, %zu threads, stop: %s\n",
                 info.scenario.c_str(), info.n_conditions, (unsigned long long)info.seed,
                 (unsigned long long)info.batch, info.threads, info.stop.c_str());
}

Action ProgressBar::on_condition_start(size_t, const Condition&) {
    last_ = std::chrono::steady_clock::now();
    started_ = true;
    return Action::Continue;
}

Action ProgressBar::on_batch_end(const Progress& p) {
    const auto now = std::chrono::steady_clock::now();
    // secs (runner.rs:476-480): no wait for a NaN or negative period, none ever for +inf
    const double every = std::isnan(every_s_) || every_s_ <= 0.0 ? 0.0 : every_s_;
    if (started_ && std::chrono::duration<double>(now - last_).count() >= every) {
        last_ = now;
        std::fprintf(stderr, "\r  condition %zu/%zu: %llu trials, %s s", p.condition + 1, n_,
                     (unsigned long long)p.trials, fixed(p.elapsed_s, 0).c_str());
        std::fflush(stderr);
    }
    return Action::Continue;
}

Action ProgressBar::on_condition_end(const ConditionResult& r) {
    std::string metrics;
    for (const auto& [k, e] : r.record.metrics) {
        const auto [lo, hi] = e.interval(confidence_);
        if (!metrics.empty()) metrics += ", ";
        metrics += k + " = " + lower_exp(e.value(), 4) + " [" + lower_exp(lo, 2) + ", " + lower_exp(hi, 2) + "]";
    }
    std::fprintf(stderr, "\r[%zu/%zu] (%s) %s: %llu trials, %s s%s%s\n", r.index + 1, n_,
                 r.condition.label().c_str(), status_str(r.status), (unsigned long long)r.trials,
                 fixed(r.wall_s, 2).c_str(), metrics.empty() ? "" : "  ", metrics.c_str());
    return Action::Continue;
}

// ---- SkipAfter (callbacks.rs:168-203) ---------------------------------------------------

std::vector<std::pair<std::string, double>> SkipAfter::others(const Condition& c) const {
    std::vector<std::pair<std::string, double>> out;
    for (const auto& e : c.entries)
        if (e.first != along_) out.push_back(e);
    return out;
}

Action SkipAfter::on_condition_start(size_t, const Condition& c) {
    return std::find(done_.begin(), done_.end(), others(c)) != done_.end() ? Action::Skip : Action::Continue;
}

Action SkipAfter::on_condition_end(const ConditionResult& r) {
    const Estimate* e = r.record.get(metric_);
    if (r.trials > 0 && e && e->value() < threshold_) {
        auto key = others(r.condition);
        if (std::find(done_.begin(), done_.end(), key) == done_.end()) done_.push_back(std::move(key));
    }
    return Action::Continue;
}

// ---- CsvLogger (loggers.rs:107-215) -----------------------------------------------------

CsvLogger::~CsvLogger() {
    if (out_) std::fclose(out_);
}

void CsvLogger::put(const std::string& line) {
    if (!out_) return;
    if (std::fputs(line.c_str(), out_) < 0 || std::fputc('\n', out_) == EOF || std::fflush(out_) != 0)
        io_fail(path_, errno);
}

void CsvLogger::start(const RunInfo&) {
    create_parent_dir(path_);
    if (out_) std::fclose(out_);
    out_ = std::fopen(path_.c_str(), "w");
    if (!out_) io_fail(path_, errno);
    header_ = false;
    pending_.clear();
}

void CsvLogger::write_header(const ConditionResult& r, double confidence) {
    keys_.clear();
    for (const auto& e : r.condition.entries) keys_.push_back(e.first);
    columns_.clear();
    std::string line;
    for (const auto& [k, v] : row(r, confidence)) {
        columns_.push_back(k);
        if (!line.empty()) line += ",";
        line += csv_field(k);
    }
    put(line);
    header_ = true;
}

void CsvLogger::write_row(const ConditionResult& r, double confidence) {
    // a value of the condition or of the head as it is, the wall time to the microsecond,
    // an estimate or a bound in scientific notation (empty when not finite)
    auto cell = [](const std::string& column, const Cell& v, bool metric) -> std::string {
        if (v.type == Cell::F64 && !metric && column == "wall_s") return fixed(v.f, 6);
        if (v.type == Cell::F64 && metric) return std::isfinite(v.f) ? lower_exp(v.f) : std::string();
        return csv_field(cell_text(v));
    };
    const auto cells = row(r, confidence);
    const size_t n_head = 4 + r.condition.entries.size(), n_cols_head = 4 + keys_.size();
    std::vector<std::string> out;
    out.push_back(cell("index", cells[0].second, false));
    // the condition by the header's keys (by name), the head by position, the metrics by name
    for (const auto& k : keys_) {
        std::string c;
        for (size_t i = 1; i + 3 < n_head; ++i)
            if (cells[i].first == k) {
                c = cell(k, cells[i].second, false);
                break;
            }
        out.push_back(c);
    }
    for (size_t i = n_head - 3; i < n_head; ++i) out.push_back(cell(cells[i].first, cells[i].second, false));
    for (size_t j = n_cols_head; j < columns_.size(); ++j) {
        std::string c;
        for (size_t i = n_head; i < cells.size(); ++i)
            if (cells[i].first == columns_[j]) {
                c = cell(columns_[j], cells[i].second, true);
                break;
            }
        out.push_back(c);
    }
    std::string line;
    for (size_t i = 0; i < out.size(); ++i) line += (i ? "," : "") + out[i];
    put(line);
}

void CsvLogger::log(const ConditionResult& r, double confidence) {
    if (!header_) {
        // skipped conditions before the first that ran wait for the header
        if (r.trials == 0) {
            pending_.push_back({r, confidence});
            return;
        }
        write_header(r, confidence);
        for (const auto& [p, c] : pending_) write_row(p, c);
        pending_.clear();
    }
    write_row(r, confidence);
}

void CsvLogger::finish(const RunReport&) {
    if (!header_) {
        if (!pending_.empty()) write_header(pending_.front().first, pending_.front().second);
        for (const auto& [p, c] : pending_) write_row(p, c);
        pending_.clear();
    }
    if (out_) {
        const bool bad = std::fclose(out_) != 0;
        out_ = nullptr;
        if (bad) io_fail(path_, errno);
    }
}

// ---- JsonLogger (loggers.rs:221-318) ----------------------------------------------------

void JsonLogger::finish(const RunReport& rep) {
    const RunInfo& i = rep.info;
    std::vector<std::string> conditions;
    for (const auto& r : rep.results) {
        std::vector<std::pair<std::string, std::string>> cond, metrics, dropped;
        for (const auto& [k, v] : r.condition.entries) cond.push_back({k, json_number(v)});
        for (const auto& [k, e] : r.record.metrics) metrics.push_back({k, estimate_json(e, i.confidence)});
        for (const auto& [k, n] : r.record.dropped) dropped.push_back({k, std::to_string(n)});
        conditions.push_back(json_object({{"index", std::to_string(r.index)},
                                          {"condition", json_object(cond)},
                                          {"status", json_string(status_str(r.status))},
                                          {"trials", std::to_string(r.trials)},
                                          {"wall_s", json_number(r.wall_s)},
                                          {"metrics", json_object(metrics)},
                                          {"dropped", json_object(dropped)}}));
    }
    // run_header (loggers.rs:271-288), the wall time and the conditions
    const std::string doc = json_object({{"scenario", json_string(i.scenario)},
                                         {"seed", std::to_string(i.seed)},
                                         {"batch", std::to_string(i.batch)},
                                         {"threads", std::to_string(i.threads)},
                                         {"stop", json_string(i.stop)},
                                         {"confidence", json_number(i.confidence)},
                                         {"version", json_string(i.version)},
                                         {"commit", json_string(i.commit)},
                                         {"config", i.has_config ? i.config : std::string("null")},
                                         {"started_unix_s", std::to_string(i.started_unix_s)},
                                         {"wall_s", json_number(rep.wall_s)},
                                         {"conditions", json_array(conditions)}});
    create_parent_dir(path_);
    std::FILE* f = std::fopen(path_.c_str(), "w");
    if (!f) io_fail(path_, errno);
    const bool bad = std::fputs(doc.c_str(), f) < 0 || std::fputc('\n', f) == EOF;
    const int err = errno;
    if (std::fclose(f) != 0 || bad) io_fail(path_, bad ? err : errno);
}

}  // namespace sim
}  // namespace comfec
