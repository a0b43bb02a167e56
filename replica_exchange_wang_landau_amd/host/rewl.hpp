// run_rewl on the MI355X engine: the reference's driver signature and result semantics
// (parallel.hpp:23-36) on top of the C ABI of include/rewl.h. Walkers, sampling, the flatness
// cadence and the replica exchanges run on the device; this layer builds the configuration,
// restores/writes the 1x1 checkpoint, prints progress and summarises the downloaded walkers.
#pragma once

#include "adaptive.hpp"
#include "geometry.hpp"
#include "output.hpp"

#include "../../include/rewl.h"

#include <memory>
#include <string>
#include <vector>

namespace rewl::host {

// Couplings (physics.hpp:41-84) built by kernel (a): dense if cutoff <= 0, else CSR; precision
// is an enum wlx_couplings (FP64 as the reference, or rounded to FP32).
class Couplings {
  public:
    Couplings(const Geometry &geometry, double coupling_scale, double cutoff,
              int precision = WLX_COUPLINGS_FP64);
    ~Couplings();
    Couplings(const Couplings &) = delete;
    Couplings &operator=(const Couplings &) = delete;
    std::size_t size() const { return n_; }
    std::string backend_name() const { return csr_ ? "csr" : "dense"; }
    wlx_system *handle() const { return system_; }

  private:
    wlx_system *system_{};
    std::size_t n_{};
    bool csr_{};
};

struct RewlResult { // parallel.hpp:23-31
    std::vector<DosFragment> fragments;
    std::vector<WindowSamplingStatistics> sampling_statistics; // collect_window_statistics
    std::vector<WalkerStatistics> walker_statistics;
    std::vector<EnergyRepresentative> representatives;         // collect_window_statistics
    std::uint64_t attempted{}, accepted{}, forced_accepted{};
    std::uint64_t exchange_attempted{}, exchange_accepted{};
    bool converged{false};
    EngineInfo engine;
};

// Engine choices that have no INI/CLI key in the reference.
struct EngineOptions {
    int rng = WLX_RNG_XOSHIRO;
    int fields = WLX_FIELDS_FP64;
    double resync_interval_mcs = 10.0; // FP32 fields: FP64 resync cadence
};

RewlResult run_rewl(std::shared_ptr<const Couplings> couplings, const RunConfig &config,
                    const EngineOptions &options = {});

struct PilotRange {
    double minimum{}, maximum{};
};
// the sampled range of wl_run --pilot (wl_run.cpp:84-97)
PilotRange pilot_energy_range(const Couplings &couplings, std::uint64_t seed,
                              std::uint64_t attempts);

// throws the reference's exception class for a non-zero ABI status
void check(int status);

} // namespace rewl::host
