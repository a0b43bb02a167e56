// wl_run output files and the WLCHKP3 walker checkpoint (io.cpp:398-572 of the reference):
// same file names, columns, number formatting (17 significant digits) and binary layout.
#pragma once

#include "analysis.hpp"

#include <string>
#include <vector>

namespace rewl::host {

struct WalkerStatistics { // io.hpp:60-75
    std::uint64_t walker_id{}, window_id{};
    int mpi_rank{};
    std::uint64_t attempted{}, accepted{}, forced_accepted{}, attempts_since_last_accepted{};
    double energy{}, factor{1.0};
    std::size_t active_bins{};
    std::uint64_t minimum_histogram{};
    double mean_histogram{}, min_over_mean{};
    std::uint64_t round_trips{};
};

// Run-level facts recorded in the metadata next to the reference's keys.
struct EngineInfo {
    std::string device;
    std::string rng;
    std::string fields;
    std::uint64_t walkers{};
    std::uint64_t epochs{};
    std::uint64_t ambiguous_decisions{};
};

void write_dos_csv(const std::string &path, const DensityOfStates &dos);
void write_fragment_csv(const std::string &path, const Grid &grid, const DosFragment &fragment,
                        std::size_t window_id);
void write_thermodynamics_csv(const std::string &path,
                              const std::vector<ThermodynamicPoint> &points);
void write_workers_stat_csv(const std::string &path, const std::vector<WalkerStatistics> &stats,
                            std::size_t spin_count);
void write_metadata_json(const std::string &path, const RunConfig &config, std::size_t spins,
                         const std::string &backend, std::uint64_t attempted,
                         std::uint64_t accepted, std::uint64_t forced_accepted,
                         std::uint64_t exchange_attempted, std::uint64_t exchange_accepted,
                         bool converged, int mpi_size, int openmp_threads,
                         const EngineInfo *engine);

// WLCHKP3: magic, walker id, spins, fields, energy, log_g, H, active, factor, attempted,
// accepted, forced, last accepted, stage, RNG words; vectors are u64 length + raw data.
void save_checkpoint(const std::string &path, const WalkerState &walker);
// Accepts WLCHKP1/2/3 (io.cpp:558-572).
WalkerState load_checkpoint(const std::string &path);

// Checkpoint of a whole run with any window x walker layout (the reference checkpoints 1x1 runs
// only): "WLCHKM2\0", u64 walker count, u64 epoch, u64 exchange attempts, u64 accepted
// exchanges, then per walker u64 last flatness check and f64 last flatness (the driver's
// cadence, parallel.cpp:364-365) and u64 1/t clock offset (a warm start's; 0 otherwise), then
// one WLCHKP3 record per walker in walker order. WLCHKM1 files (no clock offsets) still load,
// with offset 0.
struct RunCheckpoint {
    std::uint64_t epoch{}, exchange_attempted{}, exchange_accepted{};
    std::vector<WalkerState> walkers; // last_flatness_check / last_flatness filled
};
void save_run_checkpoint(const std::string &path, const RunCheckpoint &run);
RunCheckpoint load_run_checkpoint(const std::string &path);

} // namespace rewl::host
