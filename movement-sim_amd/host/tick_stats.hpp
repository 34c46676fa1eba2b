// tick_stats.hpp — tick-rate and tick-duration statistics with the
// semantics of the reference's utils::TickRate (src/utils/TickRate.cpp) and
// utils::TickDurationHistory (src/utils/TickDurationHistory.cpp).
#pragma once

#include <chrono>
#include <cstddef>
#include <string>
#include <vector>

namespace msim_host {

// Ticks per second over the last full second: every tick adds one; once a
// second has accumulated, the count scaled to exactly one second becomes the
// rate and that second is taken off (the remainder carries over).
class TickRate {
 public:
    using clock = std::chrono::high_resolution_clock;
    TickRate() = default;
    explicit TickRate(clock::time_point start) : last_(start) {}
    void tick() { tick_at(clock::now()); }
    void tick_at(clock::time_point now);
    double get_ticks() const { return ticks_; }

 private:
    clock::time_point last_ = clock::now();
    double ticks_ = 0.0;
    double cur_ = 0.0;
    std::chrono::nanoseconds since_reset_{0};
};

// Mean of the last kMax durations (a ring once full).
class TickDurationHistory {
 public:
    static constexpr size_t kMax = 100;
    void add_time(std::chrono::nanoseconds t);
    std::chrono::nanoseconds get_avg_time() const;
    // "%.2f" with ns / us / ms / s, the largest unit the mean reaches.
    std::string get_avg_time_str() const;

 private:
    size_t index_ = 0;
    std::vector<std::chrono::nanoseconds> times_;
};

}  // namespace msim_host
