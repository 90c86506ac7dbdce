// cli_common.hpp -- pieces shared by the command-line tools in this folder:
// a table of options (each one a flag, a help line and a binder that stores
// its value), a usage text generated from that table, a CSV appender that
// writes its header once, and a stopwatch of named laps.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <optional>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace bfcli {

// How a flag that no option of the table names is treated.
enum class Unknown { Reject, Skip };

class OptionTable {
public:
    using Binder = std::function<void(const std::string& flag, const std::string& value)>;

    // An option that takes a value; `metavar` and `help` make its usage line.
    OptionTable& add(std::string flag, std::string metavar, std::string help, Binder bind,
                     bool required = false) {
        entries_.push_back({std::move(flag), std::move(metavar), std::move(help),
                            std::move(bind), required, false});
        return *this;
    }

    // Flags that print the usage text and stop the tool with status 0.
    OptionTable& help_flags(std::vector<std::string> flags) {
        help_flags_ = std::move(flags);
        return *this;
    }

    // Walks argv once.  Returns false when a help flag was seen (after
    // printing the usage to `out`).  With Unknown::Skip, unrecognised
    // arguments and a last flag without its value are passed over, which is
    // how the reference's tracker bridge treats them.
    bool parse(int argc, char** argv, Unknown policy, std::ostream& out) {
        for (int i = 1; i < argc; ++i) {
            const std::string arg = argv[i];
            if (is_help(arg)) {
                out << usage(argv[0]);
                return false;
            }
            Entry* e = find(arg);
            if (e == nullptr || i + 1 >= argc) {
                if (policy == Unknown::Skip) continue;
                if (e == nullptr) throw std::invalid_argument("unknown option: " + arg);
                throw std::invalid_argument("missing value after " + arg);
            }
            e->bind(arg, argv[++i]);
            e->seen = true;
        }
        return true;
    }

    // Names of the required options not given, joined by ", ".
    std::string missing_required() const {
        std::string names;
        for (const Entry& e : entries_)
            if (e.required && !e.seen) names += (names.empty() ? "" : ", ") + e.flag;
        return names;
    }

    std::string usage(const std::string& program) const {
        std::size_t width = 0;
        for (const Entry& e : entries_) width = std::max(width, e.flag.size() + e.metavar.size() + 1);
        std::ostringstream s;
        s << "Usage: " << program;
        for (const Entry& e : entries_)
            if (e.required) s << ' ' << e.flag << ' ' << e.metavar;
        s << " [options]\n";
        for (const Entry& e : entries_) {
            const std::string head = e.flag + ' ' + e.metavar;
            s << "  " << std::left << std::setw(static_cast<int>(width) + 2) << head << e.help
              << '\n';
        }
        return s.str();
    }

private:
    struct Entry {
        std::string flag, metavar, help;
        Binder bind;
        bool required;
        bool seen;
    };

    Entry* find(const std::string& flag) {
        for (Entry& e : entries_)
            if (e.flag == flag) return &e;
        return nullptr;
    }
    bool is_help(const std::string& arg) const {
        for (const std::string& h : help_flags_)
            if (h == arg) return true;
        return false;
    }

    std::vector<Entry> entries_;
    std::vector<std::string> help_flags_;
};

// ---- binders: each converts the text and stores it in the caller's field.
// `strict` binders reject trailing characters ("12x"); the tracker bridge
// uses the lenient forms, which read a leading number as std::stoull and
// std::stof do.
inline std::size_t to_size(const std::string& flag, const std::string& text, bool strict) {
    std::size_t used = 0;
    unsigned long long v = 0;
    try {
        v = std::stoull(text, &used);
    } catch (const std::exception&) {
        throw std::invalid_argument("invalid integer for " + flag + ": " + text);
    }
    if (strict && used != text.size())
        throw std::invalid_argument("invalid integer for " + flag + ": " + text);
    return static_cast<std::size_t>(v);
}

inline OptionTable::Binder size_into(std::size_t& dst, bool strict = true) {
    return [&dst, strict](const std::string& f, const std::string& v) { dst = to_size(f, v, strict); };
}
inline OptionTable::Binder size_into(std::optional<std::size_t>& dst) {
    return [&dst](const std::string& f, const std::string& v) { dst = to_size(f, v, true); };
}
inline OptionTable::Binder float_into(float& dst) {
    return [&dst](const std::string& f, const std::string& v) {
        try {
            dst = std::stof(v);
        } catch (const std::exception&) {
            throw std::invalid_argument("invalid number for " + f + ": " + v);
        }
    };
}
inline OptionTable::Binder text_into(std::string& dst) {
    return [&dst](const std::string&, const std::string& v) { dst = v; };
}
inline OptionTable::Binder path_into(std::filesystem::path& dst) {
    return [&dst](const std::string&, const std::string& v) { dst = v; };
}
inline OptionTable::Binder path_into(std::optional<std::filesystem::path>& dst) {
    return [&dst](const std::string&, const std::string& v) { dst = std::filesystem::path(v); };
}
// One of a fixed list of names, each with the value it stands for.
template <typename T>
OptionTable::Binder choice_into(T& dst, std::vector<std::pair<std::string, T>> names) {
    return [&dst, names](const std::string& f, const std::string& v) {
        std::string list;
        for (const auto& [name, value] : names) {
            if (name == v) {
                dst = value;
                return;
            }
            list += (list.empty() ? "" : ", ") + name;
        }
        throw std::invalid_argument(f + " must be one of " + list + " (got " + v + ")");
    };
}

// ---- a CSV file that gains one row per run; the header is written when the
// file is new or empty.
class CsvAppender {
public:
    explicit CsvAppender(std::filesystem::path path) : path_(std::move(path)) {}

    template <typename V>
    CsvAppender& column(const std::string& name, const V& value) {
        std::ostringstream s;
        s << std::fixed << std::setprecision(6) << value;
        cols_.emplace_back(name, s.str());
        return *this;
    }

    void append() const {
        std::error_code ec;
        const bool fresh = !std::filesystem::exists(path_, ec) || std::filesystem::file_size(path_, ec) == 0;
        if (ec) throw std::runtime_error("cannot inspect metrics file: " + path_.string());
        std::ofstream out(path_, std::ios::app);
        if (!out) throw std::runtime_error("cannot open metrics file: " + path_.string());
        auto row = [&](bool names) {
            for (std::size_t i = 0; i < cols_.size(); ++i)
                out << (i ? "," : "") << (names ? cols_[i].first : cols_[i].second);
            out << '\n';
        };
        if (fresh) row(true);
        row(false);
        if (!out) throw std::runtime_error("failed to write metrics file: " + path_.string());
    }

private:
    std::filesystem::path path_;
    std::vector<std::pair<std::string, std::string>> cols_;
};

// ---- wall-clock laps in milliseconds, named at the end of each lap.
class Laps {
public:
    Laps() : start_(Clock::now()), last_(start_) {}
    double lap() {
        const auto now = Clock::now();
        const double ms = std::chrono::duration<double, std::milli>(now - last_).count();
        last_ = now;
        return ms;
    }
    double total() const {
        return std::chrono::duration<double, std::milli>(last_ - start_).count();
    }

private:
    using Clock = std::chrono::steady_clock;
    Clock::time_point start_, last_;
};

}  // namespace bfcli
