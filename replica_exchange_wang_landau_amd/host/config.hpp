// Run configuration of the wl_run command line (RunConfig, io.hpp:12-58) and its parser.
//
// Keys, defaults, units and error texts are the reference's interface (io.cpp:75-372); the
// parser itself is a table of setters shared by the INI and command-line front ends.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace rewl::host {

struct Vec3 {
    double x{}, y{}, z{};
};

struct Window {
    std::size_t begin{}, end{}; // [begin, end)
    bool contains(std::size_t b) const { return b >= begin && b < end; }
};

struct Grid {
    double minimum{}, maximum{}, width{};
    void validate() const;                 // wang_landau.cpp:13-20
    std::size_t bins() const;              // wang_landau.cpp:22-25
    long long index(double energy) const;  // wang_landau.cpp:27-34, -1 = none
    double center(std::size_t i) const;    // wang_landau.cpp:36-38
};

std::vector<Window> partition_windows(std::size_t bins, std::size_t count, double overlap);

struct WlParameters { // wang_landau.hpp:40-56
    double flatness{0.8};
    std::uint64_t minimum_visits{100};
    double final_factor{1e-8};
    std::uint64_t check_interval_attempts{10'000};
    std::uint64_t force_accept_after_attempts{};
    bool inverse_time_enabled{true};
    std::uint64_t initialization_max_attempts{1'000'000};
    double initialization_target_fraction{0.5};
    double initialization_temperature_fraction{0.05};
    std::uint64_t initialization_stall_attempts_per_spin{1'000};
    double initialization_temperature_multiplier{2.0};
    double initialization_max_temperature_fraction{0.5};
    bool collect_window_statistics{false};
    double round_trip_margin_fraction{0.1};
    bool nalivaiko_mod{false};
};

struct AdaptiveWindowParameters { // analysis.hpp:28-38
    bool enabled{false};
    std::size_t iterations{2};
    double pilot_mcs{10'000.0};
    double smoothing_width{};
    double minimum_width{};
    double diffusivity_floor_fraction{0.05};
    double curvature_weight{0.25};
    double round_trip_target{2.0};
    double maximum_round_trip_penalty{3.0};
};

// One attempt-based interval that may be given in Monte-Carlo sweeps (MCS) or in attempts.
struct Interval {
    double mcs{};
    std::uint64_t attempts{};
    bool uses_mcs{true};
};

struct RunConfig {
    std::size_t nx{2}, ny{2}, nz{2};
    double spacing{1.0};
    Vec3 axis{0.0, 0.0, 1.0};
    bool periodic{true};
    std::string geometry_file;
    std::string config_file;
    Vec3 box_lengths{2.0, 2.0, 2.0};
    double coupling_scale{1.0};
    double cutoff{0.0};
    Grid grid{-20.0, 20.0, 0.1};
    bool energy_grid_explicit{false};
    std::size_t windows{1};
    std::size_t walkers_per_rank{1};
    double overlap{0.75};
    std::vector<Window> explicit_windows;
    // starting configurations per walker (window-major; empty = generated), set between
    // adaptive pilots (io.hpp:28)
    std::vector<std::vector<std::int8_t>> initial_spins_by_walker;
    AdaptiveWindowParameters adaptive_windows;
    std::uint64_t seed{1};
    bool seed_explicit{false};
    WlParameters wl;
    Interval exchange{10.0, 1000, true};
    Interval check{100.0, 10'000, true};
    Interval force_accept{0.0, 0, true};
    Interval max_limit{1'000'000.0, 10'000'000, true};
    Interval checkpoint{10'000.0, 1'000'000, true};
    std::size_t resolved_spin_count{};
    bool complete_range{false};
    std::string output_prefix{"wl_output"};
    std::string checkpoint_path;
    std::vector<double> temperatures{0.5, 1.0, 2.0, 5.0};
    bool smoke_test{false};
    bool pilot{false};
    double progress_interval_seconds{};

    std::uint64_t exchange_interval_attempts() const { return exchange.attempts; }
    std::uint64_t max_attempts() const { return max_limit.attempts; }
    std::uint64_t checkpoint_interval_attempts() const { return checkpoint.attempts; }

    void resolve_mcs(std::size_t spin_count);      // io.cpp:210-236
    bool uses_legacy_attempt_units() const;        // io.cpp:238-241
    void validate(bool require_explicit_grid = true) const; // io.cpp:243-283
};

// --config files are applied first, in order; command-line options then override them.
RunConfig parse_arguments(int argc, char **argv);
std::string usage(const std::string &program);

} // namespace rewl::host
