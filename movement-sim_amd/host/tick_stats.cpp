// tick_stats.cpp — see tick_stats.hpp.
#include "tick_stats.hpp"

#include <cstdio>

namespace msim_host {

void TickRate::tick_at(clock::time_point now) {
    const std::chrono::nanoseconds since = now - last_;
    last_ = now;
    since_reset_ += since;
    cur_ += 1.0;
    if (since_reset_ >= std::chrono::seconds(1)) {
        const double factor = 1e9 / static_cast<double>(since_reset_.count());
        ticks_ = cur_ * factor;
        cur_ -= ticks_;
        since_reset_ -= std::chrono::seconds(1);
    }
}

void TickDurationHistory::add_time(std::chrono::nanoseconds t) {
    if (times_.size() >= kMax) {
        times_[index_++] = t;
        if (index_ >= kMax) index_ = 0;
    } else {
        times_.push_back(t);
    }
}

std::chrono::nanoseconds TickDurationHistory::get_avg_time() const {
    if (times_.empty()) return std::chrono::nanoseconds(0);
    std::chrono::nanoseconds sum(0);
    for (const auto& t : times_) sum += t;
    return std::chrono::nanoseconds(sum.count() / static_cast<long long>(times_.size()));
}

std::string TickDurationHistory::get_avg_time_str() const {
    double v = static_cast<double>(get_avg_time().count());
    const char* unit = "ns";
    if (v >= 1e3) {
        if (v >= 1e6) {
            if (v >= 1e9) {
                v /= 1e9;
                unit = "s";
            } else {
                v /= 1e6;
                unit = "ms";
            }
        } else {
            v /= 1e3;
            unit = "us";
        }
    }
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.2f%s", v, unit);
    return buf;
}

}  // namespace msim_host
