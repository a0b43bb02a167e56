// Adaptive energy-window planning between pilot runs (wl_run --adaptive-windows).
//
// A pilot run collects, per window, the squared energy displacement of every attempt and the
// walkers' round trips; the planner turns them into a local diffusivity and places the window
// cores at equal quantiles of a cost density ~ 1/sqrt(D(E)), raised where log g is curved or
// where a window's walkers completed few round trips (analysis.cpp:183-338). The next pilot
// (or the production run) starts its walkers from the pilot's representative configurations
// (analysis.cpp:341-416). Both run once per pilot on the host, on downloaded statistics.
#pragma once

#include "analysis.hpp"

#include <cstdint>
#include <vector>

namespace rewl::host {

struct WindowSamplingStatistics { // analysis.hpp:20-26
    Window window;
    std::vector<double> squared_energy_displacement; // full grid
    std::vector<std::uint64_t> displacement_samples;  // full grid
    std::uint64_t round_trips{};
    std::uint64_t walkers{};
};

struct EnergyRepresentative { // wang_landau.hpp:69-72
    double energy{};
    std::vector<std::int8_t> spins;
};

// summarize_sampling_impl (parallel.cpp:64-79): per-window sums over its walkers
struct WalkerSampling {
    std::vector<double> squared_energy_displacement; // full grid
    std::vector<std::uint64_t> displacement_samples;  // full grid
    std::uint64_t round_trips{};
    std::vector<EnergyRepresentative> representatives; // the recorded ones, in target order
};
WindowSamplingStatistics summarize_sampling(Window window,
                                            const std::vector<const WalkerSampling *> &walkers,
                                            std::size_t bins);

// adapt_energy_windows (analysis.cpp:183-338)
std::vector<Window> adapt_energy_windows(const Grid &grid,
                                         const std::vector<DosFragment> &fragments,
                                         const std::vector<WindowSamplingStatistics> &sampling,
                                         std::size_t window_count, double overlap,
                                         const AdaptiveWindowParameters &parameters);

// select_adaptive_initial_configurations on one process (analysis.cpp:341-416): one spin
// vector per walker (window-major), empty where no representative was available.
struct InitialConfigurations {
    std::vector<std::vector<std::int8_t>> spins;
    std::size_t missing{}, external_warm_starts{};
};
InitialConfigurations select_initial_configurations(
    const Grid &grid, const std::vector<Window> &windows,
    const std::vector<EnergyRepresentative> &representatives, std::size_t walkers_per_window);

} // namespace rewl::host
