// wl_run: the reference's command-line driver (apps/wl_run.cpp) on the MI355X engine.
// Same options, INI keys, stdout lines, output files and exit codes (0 converged or smoke
// test, 2 unconverged, 1 on error); the walkers run on the GPU.
#include "rewl.hpp"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <limits>
#include <random>
#include <string>

using namespace rewl::host;

namespace {

std::uint64_t scramble(std::uint64_t v)
{
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

// a fresh seed from the clock and the OS entropy source
std::uint64_t random_seed()
{
    std::uint64_t seed = static_cast<std::uint64_t>(
        std::chrono::high_resolution_clock::now().time_since_epoch().count());
    try {
        std::random_device device;
        for (int i = 0; i < 4; ++i) seed = scramble(seed ^ static_cast<std::uint64_t>(device()));
    } catch (const std::exception &) {
    }
    return scramble(seed);
}

int engine_rng_from_environment()
{
    const char *v = std::getenv("REWL_RNG");
    return (v && std::string(v) == "philox") ? WLX_RNG_PHILOX : WLX_RNG_XOSHIRO;
}

// REWL_COUPLINGS=fp32: couplings rounded to FP32 after the build (FP64 fields and energies)
int coupling_precision_from_environment()
{
    const char *v = std::getenv("REWL_COUPLINGS");
    return (v && std::string(v) == "fp32") ? WLX_COUPLINGS_FP32 : WLX_COUPLINGS_FP64;
}

// REWL_FIELDS=fp32: FP32 Kahan fields (needs REWL_RNG=philox); REWL_RESYNC_MCS: its FP64
// resync cadence in MCS (default 10)
void engine_fields_from_environment(EngineOptions &options)
{
    const char *v = std::getenv("REWL_FIELDS");
    if (v && std::string(v) == "fp32") options.fields = WLX_FIELDS_FP32_KAHAN;
    if (const char *r = std::getenv("REWL_RESYNC_MCS")) options.resync_interval_mcs = std::atof(r);
}

// The adaptive pilots (wl_run.cpp:105-174): each pilot runs REWL for pilot_mcs on the current
// windows with statistics collection and its own seed; the planner then moves the windows and
// picks the next starting configurations. On return config holds the final windows and starts.
void plan_adaptive_windows(RunConfig &config, std::shared_ptr<const Couplings> couplings,
                           std::size_t spins, EngineOptions options)
{
    options.fields = WLX_FIELDS_FP64; // the statistics are collected with FP64 fields
    auto windows = partition_windows(config.grid.bins(), config.windows, config.overlap);
    std::vector<std::vector<std::int8_t>> starts;
    const auto t0 = std::chrono::steady_clock::now();
    const std::size_t iterations = config.adaptive_windows.iterations;
    for (std::size_t it = 0; it < iterations; ++it) {
        std::cout << "adaptive_window_pilot=" << (it + 1) << '/' << iterations
                  << " pilot_mcs=" << config.adaptive_windows.pilot_mcs << '\n';
        RunConfig pilot = config;
        pilot.adaptive_windows.enabled = false;
        pilot.explicit_windows = windows;
        pilot.initial_spins_by_walker = starts;
        pilot.wl.collect_window_statistics = true;
        pilot.max_limit.mcs = config.adaptive_windows.pilot_mcs;
        pilot.max_limit.uses_mcs = true;
        pilot.checkpoint_path.clear();
        pilot.resolved_spin_count = 0;
        pilot.seed = config.seed ^ ((static_cast<std::uint64_t>(it) + 1) * 0x9e3779b97f4a7c15ULL);
        pilot.resolve_mcs(spins);
        const RewlResult r = run_rewl(couplings, pilot, options);

        windows = adapt_energy_windows(config.grid, r.fragments, r.sampling_statistics,
                                       config.windows, config.overlap, config.adaptive_windows);
        std::uint64_t round_trips = 0;
        for (const auto &s : r.sampling_statistics) round_trips += s.round_trips;
        double lowest = std::numeric_limits<double>::infinity();
        for (const auto &rep : r.representatives)
            if (!rep.spins.empty()) lowest = std::min(lowest, rep.energy);
        InitialConfigurations next = select_initial_configurations(
            config.grid, windows, r.representatives, config.walkers_per_rank);
        std::cout << "adaptive_window_round_trips=" << round_trips
                  << " seeded_walkers=" << next.spins.size() - next.missing << '/'
                  << next.spins.size() << " external_warm_starts=" << next.external_warm_starts;
        if (std::isfinite(lowest)) std::cout << " minimum_seed_energy=" << lowest;
        std::cout << " energy_ranges=";
        for (std::size_t w = 0; w < windows.size(); ++w) {
            if (w != 0) std::cout << ';';
            std::cout << '['
                      << config.grid.minimum + static_cast<double>(windows[w].begin) * config.grid.width
                      << ','
                      << config.grid.minimum + static_cast<double>(windows[w].end) * config.grid.width
                      << ')';
        }
        std::cout << '\n';
        starts = std::move(next.spins);
    }
    config.explicit_windows = std::move(windows);
    config.initial_spins_by_walker = std::move(starts);
    config.wl.collect_window_statistics = false;
    config.validate(!config.smoke_test);
    std::cout << "adaptive_window_elapsed_seconds="
              << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()
              << '\n';
}

} // namespace

int main(int argc, char **argv)
{
    try {
        for (int i = 1; i < argc; ++i)
            if (std::string(argv[i]) == "--help" || std::string(argv[i]) == "-h") {
                std::cout << usage(argv[0]);
                return 0;
            }
        RunConfig config = parse_arguments(argc, argv);
        if (!config.seed_explicit) config.seed = random_seed();
        std::cout << "seed=" << config.seed
                  << " source=" << (config.seed_explicit ? "configured" : "random") << '\n';
        const Geometry geometry = geometry_from_config(config);
        config.resolve_mcs(geometry.size());
        config.validate(!config.smoke_test && !config.pilot);
        if (config.pilot && config.max_attempts() == 0)
            throw std::invalid_argument("Pilot mode requires a positive max-mcs value");
        const auto &wl = config.wl;
        std::cout << "spins=" << geometry.size()
                  << " exchange_interval_mcs=" << config.exchange.mcs
                  << " check_interval_mcs=" << config.check.mcs
                  << " max_mcs=" << config.max_limit.mcs
                  << " checkpoint_interval_mcs=" << config.checkpoint.mcs
                  << " force_accept_after_mcs=" << config.force_accept.mcs
                  << " inverse_time=" << (wl.inverse_time_enabled ? "yes" : "no")
                  << " initialization_max_attempts=" << wl.initialization_max_attempts
                  << " initialization_target_fraction=" << wl.initialization_target_fraction
                  << " initialization_temperature_fraction="
                  << wl.initialization_temperature_fraction
                  << " initialization_stall_attempts_per_spin="
                  << wl.initialization_stall_attempts_per_spin
                  << " initialization_temperature_multiplier="
                  << wl.initialization_temperature_multiplier
                  << " initialization_max_temperature_fraction="
                  << wl.initialization_max_temperature_fraction
                  << " progress_interval_seconds=" << config.progress_interval_seconds << '\n';
        if (config.uses_legacy_attempt_units())
            std::cerr << "warning: legacy attempt-based interval option used; prefer the "
                         "corresponding *_mcs option\n";
        if (!config.pilot && wl.force_accept_after_attempts != 0)
            std::cerr << "warning: force_accept_after_mcs=" << config.force_accept.mcs
                      << " enables non-standard forced acceptance and may bias the DOS\n";
        auto couplings =
            std::make_shared<const Couplings>(geometry, config.coupling_scale, config.cutoff,
                                              coupling_precision_from_environment());
        if (config.pilot) { // wl_run.cpp:84-104
            const auto r = pilot_energy_range(*couplings, config.seed, config.max_attempts());
            const double padding = std::max(1e-12, 0.05 * (r.maximum - r.minimum));
            std::cout << "pilot_only=true attempted_flips=" << config.max_attempts()
                      << " mcs=" << config.max_limit.mcs << " sampled_min=" << r.minimum
                      << " sampled_max=" << r.maximum
                      << " suggested_emin=" << r.minimum - padding
                      << " suggested_emax=" << r.maximum + padding
                      << "\nwarning: sampled bounds do not prove the complete energy range\n";
            return 0;
        }
        EngineOptions options;
        options.rng = engine_rng_from_environment();
        engine_fields_from_environment(options);
        if (config.adaptive_windows.enabled)
            plan_adaptive_windows(config, couplings, geometry.size(), options);
        const auto start = std::chrono::steady_clock::now();
        const RewlResult result = run_rewl(couplings, config, options);
        const double elapsed =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        const auto dos =
            stitch_dos(config.grid, result.fragments, config.complete_range, geometry.size());
        write_dos_csv(config.output_prefix + "_dos.csv", dos);
        write_thermodynamics_csv(config.output_prefix + "_thermo.csv",
                                 thermodynamics(dos, config.temperatures));
        for (std::size_t i = 0; i < result.fragments.size(); ++i)
            write_fragment_csv(config.output_prefix + "_window_" + std::to_string(i) + ".csv",
                               config.grid, result.fragments[i], i);
        write_metadata_json(config.output_prefix + "_metadata.json", config, couplings->size(),
                            couplings->backend_name(), result.attempted, result.accepted,
                            result.forced_accepted, result.exchange_attempted,
                            result.exchange_accepted, result.converged, 1, 1, &result.engine);
        if (!result.converged)
            write_workers_stat_csv(config.output_prefix + "_workers_stat.csv",
                                   result.walker_statistics, geometry.size());
        std::cout << "attempted_flips=" << result.attempted << " aggregate_mcs="
                  << static_cast<double>(result.attempted) / static_cast<double>(geometry.size())
                  << " accepted=" << result.accepted << " forced=" << result.forced_accepted
                  << " exchanges=" << result.exchange_accepted << '/' << result.exchange_attempted
                  << " converged=" << (result.converged ? "yes" : "no")
                  << " elapsed_seconds=" << elapsed << '\n';
        if (!result.converged)
            std::cerr << "warning: max-mcs reached before final factor; walker statistics written to "
                      << config.output_prefix << "_workers_stat.csv\n";
        return result.converged || config.smoke_test ? 0 : 2;
    } catch (const std::exception &error) {
        std::cerr << "wl_run: " << error.what() << '\n';
        return 1;
    }
}
