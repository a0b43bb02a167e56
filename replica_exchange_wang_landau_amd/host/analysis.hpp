// Host post-processing of device walker states: per-walker histogram statistics, per-window
// DOS fragments, stitching into one DOS and thermodynamics. These run once per run on
// downloaded arrays; their arithmetic is the reference's (analysis.cpp, parallel.cpp:23-62),
// so the written files are byte-identical for identical walker states.
#pragma once

#include "config.hpp"

#include <cstdint>
#include <vector>

namespace rewl::host {

// A walker's state as downloaded from the device (WalkerSnapshot + driver bookkeeping).
struct WalkerState {
    std::uint64_t id{}, window{};
    Window bins_window{};
    std::vector<std::int8_t> spins;
    std::vector<double> fields;
    double energy{};
    std::vector<double> log_g;            // full grid
    std::vector<std::uint64_t> histogram; // full grid
    std::vector<std::uint8_t> active;     // full grid
    std::vector<std::uint8_t> refinement; // full grid (Nalivaiko mask)
    double factor{1.0};
    std::uint64_t attempted{}, accepted{}, forced_accepted{}, last_accepted_attempt{};
    int stage{};
    std::uint64_t active_bins{}, refinement_active_bins{};
    std::uint64_t rng_state[4]{};
    std::uint64_t last_flatness_check{}; // driver cadence (parallel.cpp:364-365)
    double last_flatness{};
    std::uint64_t clock_offset{}; // 1/t clock offset of a warm start (0 without one)
};

struct HistogramStatistics { // wang_landau.hpp:60-67
    std::size_t active_bins{}, covered_bins{};
    std::uint64_t minimum{};
    double mean{}, min_over_mean{}, coverage{};
};

HistogramStatistics histogram_statistics(const WalkerState &w, bool nalivaiko);

struct DosFragment { // analysis.hpp:12-18
    Window window;
    std::vector<double> log_g;
    std::vector<std::uint64_t> histogram;
    std::vector<double> standard_error;
    std::vector<std::uint8_t> valid;
};

struct DensityOfStates { // analysis.hpp:40-48
    Grid grid;
    std::vector<double> log_g;
    std::vector<std::uint64_t> histogram;
    std::vector<double> standard_error;
    std::vector<std::uint8_t> valid;
    std::vector<std::size_t> join_bins;
    bool fully_normalized{false};
};

struct ThermodynamicPoint { // analysis.hpp:50-57
    double temperature{}, log_partition{}, internal_energy{}, heat_capacity{}, free_energy{},
        entropy{};
};

// aligned log-mean of one window's walkers with the within-run SEM (parallel.cpp:23-62)
DosFragment summarize_walkers(Window window, const std::vector<const WalkerState *> &walkers,
                              std::size_t bins, bool allow_empty_intersection = false);
DensityOfStates stitch_dos(Grid grid, std::vector<DosFragment> fragments, bool complete_range,
                           std::size_t spin_count);
std::vector<ThermodynamicPoint> thermodynamics(const DensityOfStates &dos,
                                               const std::vector<double> &temperatures,
                                               double boltzmann = 1.0);

} // namespace rewl::host
