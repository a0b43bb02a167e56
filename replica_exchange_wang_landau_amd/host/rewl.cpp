#include "rewl.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace rewl::host {

void check(int status)
{
    if (status == WLX_OK) return;
    std::string message(wlx_last_error(nullptr, 0), '\0');
    wlx_last_error(message.data(), message.size() + 1);
    switch (status) {
    case WLX_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case WLX_LOGIC_ERROR: throw std::logic_error(message);
    case WLX_OUT_OF_RANGE: throw std::out_of_range(message);
    default: throw std::runtime_error(message);
    }
}

Couplings::Couplings(const Geometry &g, double scale, double cutoff, int precision)
    : n_(g.size()), csr_(cutoff > 0.0)
{
    g.validate();
    const auto p = g.flat_positions(), a = g.flat_axes();
    const double box[3] = {g.box.x, g.box.y, g.box.z};
    check(wlx_system_create(p.data(), a.data(), g.size(), box, g.periodic ? 1 : 0, scale, cutoff,
                            precision, &system_));
}

Couplings::~Couplings() { wlx_system_destroy(system_); }

PilotRange pilot_energy_range(const Couplings &c, std::uint64_t seed, std::uint64_t attempts)
{
    PilotRange r;
    check(wlx_system_pilot(c.handle(), seed, attempts, &r.minimum, &r.maximum));
    return r;
}

namespace {

class Engine {
  public:
    Engine(wlx_system *system, const wlx_config &config) { check(wlx_engine_create(system, &config, &e_)); }
    ~Engine() { wlx_engine_destroy(e_); }
    wlx_engine *get() const { return e_; }

  private:
    wlx_engine *e_{};
};

WalkerState download(wlx_engine *e, std::uint64_t k, std::size_t n, std::size_t bins)
{
    WalkerState w;
    w.spins.resize(n);
    w.fields.resize(n);
    w.log_g.resize(bins);
    w.histogram.resize(bins);
    w.active.resize(bins);
    w.refinement.resize(bins);
    wlx_walker_info info{};
    check(wlx_engine_read_walker(e, k, &info, w.spins.data(), w.fields.data(), w.log_g.data(),
                                 w.histogram.data(), w.active.data(), w.refinement.data()));
    w.id = info.walker_id;
    w.window = info.window;
    w.bins_window = {info.window_begin, info.window_end};
    w.energy = info.energy;
    w.factor = info.factor;
    w.attempted = info.attempted;
    w.accepted = info.accepted;
    w.forced_accepted = info.forced_accepted;
    w.last_accepted_attempt = info.last_accepted_attempt;
    w.stage = int(info.stage);
    w.active_bins = info.active_bins;
    w.refinement_active_bins = info.refinement_active_bins;
    std::memcpy(w.rng_state, info.rng_state, sizeof w.rng_state);
    w.last_flatness_check = info.last_flatness_check;
    w.last_flatness = info.last_flatness;
    w.clock_offset = info.clock_offset;
    return w;
}

// the pilot statistics of walker k (accessors of wang_landau.hpp:118-127)
WalkerSampling download_sampling(wlx_engine *e, std::uint64_t k, std::size_t n, std::size_t bins)
{
    WalkerSampling s;
    s.squared_energy_displacement.resize(bins);
    s.displacement_samples.resize(bins);
    double energies[9];
    std::uint8_t present[9];
    std::vector<std::int8_t> spins(9 * n);
    check(wlx_engine_read_sampling(e, k, s.squared_energy_displacement.data(),
                                   s.displacement_samples.data(), &s.round_trips, energies,
                                   present, spins.data()));
    for (int r = 0; r < 9; ++r)
        if (present[r])
            s.representatives.push_back(
                {energies[r], std::vector<std::int8_t>(spins.begin() + r * n,
                                                       spins.begin() + (r + 1) * n)});
    return s;
}

// WangLandauWalker::restore of a WLCHKP record into walker k (wang_landau.cpp:420-466)
void restore(wlx_engine *e, const WalkerState &s, std::size_t n, std::size_t bins,
             std::uint64_t k = 0)
{
    if (s.id != k || s.spins.size() != n || s.fields.size() != n || s.log_g.size() != bins ||
        s.histogram.size() != bins || s.active.size() != bins)
        throw std::invalid_argument("Checkpoint is incompatible with walker");
    wlx_walker_info info{};
    info.walker_id = s.id;
    info.energy = s.energy;
    info.factor = s.factor;
    info.attempted = s.attempted;
    info.accepted = s.accepted;
    info.forced_accepted = s.forced_accepted;
    info.last_accepted_attempt = s.last_accepted_attempt;
    info.stage = std::uint64_t(s.stage);
    std::memcpy(info.rng_state, s.rng_state, sizeof info.rng_state);
    info.clock_offset = s.clock_offset;
    check(wlx_engine_write_walker(e, k, &info, s.spins.data(), s.fields.data(), s.log_g.data(),
                                  s.histogram.data(), s.active.data()));
}

// every walker, the driver's cadence and the run counters from a WLCHKM1/WLCHKM2 file
void restore_run(wlx_engine *e, const RunCheckpoint &run, std::size_t walkers, std::size_t n,
                 std::size_t bins)
{
    if (run.walkers.size() != walkers)
        throw std::invalid_argument("Checkpoint holds a different number of walkers");
    for (std::size_t k = 0; k < walkers; ++k) {
        restore(e, run.walkers[k], n, bins, k);
        check(wlx_engine_set_walker_cadence(e, k, run.walkers[k].last_flatness_check,
                                            run.walkers[k].last_flatness));
    }
    check(wlx_engine_set_run_state(e, run.epoch, run.exchange_attempted, run.exchange_accepted));
}

std::string progress_line(const RunConfig &c, const std::vector<wlx_walker_info> &infos,
                          std::size_t per_window, std::size_t spins)
{ // parallel.cpp:468-519
    std::uint64_t attempted = 0;
    double factor = 0.0, flatness = 1.0;
    std::uint64_t stages[3]{};
    for (std::size_t k = 0; k < infos.size(); ++k) {
        const auto &i = infos[k];
        if (k % per_window == 0) attempted = std::max<std::uint64_t>(attempted, i.attempted);
        factor = std::max(factor, i.factor);
        ++stages[std::min<std::uint64_t>(i.stage, 2)];
        if (i.stage == WLX_STAGE_WANG_LANDAU) flatness = std::min(flatness, i.last_flatness);
    }
    std::ostringstream line;
    line << "progress mcs=" << static_cast<double>(attempted) / static_cast<double>(spins)
         << " attempted_flips_per_walker=" << attempted
         << (c.wl.inverse_time_enabled ? " coverage=" : " flatness=") << flatness
         << " factor=" << factor << " stages=" << stages[0] << '/' << stages[1] << '/'
         << stages[2];
    return line.str();
}

} // namespace

RewlResult run_rewl(std::shared_ptr<const Couplings> couplings, const RunConfig &c,
                    const EngineOptions &options)
{ // parallel.cpp:276-581, shared-memory path on one GPU
    if (!couplings) throw std::invalid_argument("Null couplings");
    c.validate(!c.smoke_test);
    const std::size_t n = couplings->size(), bins = c.grid.bins();
    const auto windows =
        c.explicit_windows.empty() ? partition_windows(bins, c.windows, c.overlap)
                                   : c.explicit_windows;
    const std::size_t walkers = c.windows * c.walkers_per_rank;
    std::vector<std::int8_t> initial;
    std::vector<std::uint8_t> supplied;
    if (!c.initial_spins_by_walker.empty()) { // parallel.cpp:286-293
        if (c.initial_spins_by_walker.size() != walkers)
            throw std::invalid_argument("Adaptive initial-configuration count mismatch");
        for (const auto &spins : c.initial_spins_by_walker)
            if (!spins.empty() && spins.size() != n)
                throw std::invalid_argument("Adaptive initial spin count mismatch");
        initial.assign(walkers * n, 1);
        supplied.assign(walkers, 0);
        for (std::size_t k = 0; k < walkers; ++k)
            if (!c.initial_spins_by_walker[k].empty()) {
                std::copy(c.initial_spins_by_walker[k].begin(), c.initial_spins_by_walker[k].end(),
                          initial.begin() + static_cast<std::ptrdiff_t>(k * n));
                supplied[k] = 1;
            }
    }
    std::vector<std::uint64_t> be;
    for (const auto &w : windows) be.insert(be.end(), {w.begin, w.end});

    wlx_config cfg{};
    cfg.emin = c.grid.minimum;
    cfg.emax = c.grid.maximum;
    cfg.bin_width = c.grid.width;
    cfg.windows = c.windows;
    cfg.walkers_per_window = c.walkers_per_rank;
    cfg.overlap = c.overlap;
    cfg.seed = c.seed;
    cfg.exchange_interval_attempts = c.exchange.attempts;
    cfg.max_attempts = c.max_limit.attempts;
    cfg.rng = std::uint64_t(options.rng);
    cfg.fields = std::uint64_t(options.fields);
    if (options.fields == WLX_FIELDS_FP32_KAHAN) // attempts_from_mcs (io.cpp:75-87)
        cfg.resync_interval_attempts = std::max<std::uint64_t>(
            1, static_cast<std::uint64_t>(std::floor(
                   static_cast<long double>(options.resync_interval_mcs) *
                       static_cast<long double>(n) +
                   0.5L)));
    cfg.explicit_windows = be.data();
    const auto &p = c.wl;
    cfg.wl = {p.flatness,
              p.minimum_visits,
              p.final_factor,
              p.check_interval_attempts,
              p.force_accept_after_attempts,
              p.inverse_time_enabled ? 1u : 0u,
              p.initialization_max_attempts,
              p.initialization_target_fraction,
              p.initialization_temperature_fraction,
              p.initialization_stall_attempts_per_spin,
              p.initialization_temperature_multiplier,
              p.initialization_max_temperature_fraction,
              p.nalivaiko_mod ? 1u : 0u,
              p.collect_window_statistics ? 1u : 0u,
              p.round_trip_margin_fraction};
    Engine engine(couplings->handle(), cfg);
    wlx_engine *e = engine.get();
    check(wlx_engine_initialize(e, initial.empty() ? nullptr : initial.data(),
                                supplied.empty() ? nullptr : supplied.data()));

    // 1x1 runs keep the reference's WLCHKP3 file (parallel.cpp:354-360, 522-526); other
    // layouts write a WLCHKM2 file of every walker at the same cadence (walker 0's attempts)
    const bool checkpointing = !c.checkpoint_path.empty() && walkers == 1;
    const bool run_checkpointing = !c.checkpoint_path.empty() && walkers > 1;
    if (checkpointing) {
        try {
            restore(e, load_checkpoint(c.checkpoint_path), n, bins);
        } catch (const std::exception &error) {
            std::cerr << "warning: checkpoint was not restored: " << error.what()
                      << "; starting a new run\n";
        }
    }
    std::uint64_t last_checkpoint = 0;
    if (run_checkpointing && std::filesystem::exists(c.checkpoint_path)) {
        try {
            const RunCheckpoint run = load_run_checkpoint(c.checkpoint_path);
            restore_run(e, run, walkers, n, bins);
            last_checkpoint = run.walkers[0].attempted;
        } catch (const std::exception &error) {
            std::cerr << "warning: checkpoint was not restored: " << error.what()
                      << "; starting a new run\n";
            check(wlx_engine_initialize(e, initial.empty() ? nullptr : initial.data(),
                                        supplied.empty() ? nullptr : supplied.data()));
        }
    }

    const auto period = std::chrono::duration_cast<std::chrono::steady_clock::duration>(
        std::chrono::duration<double>(c.progress_interval_seconds));
    auto next_progress = std::chrono::steady_clock::now() + period;
    std::size_t progress_width = 0;
    wlx_run_status st{};
    std::uint64_t chunk = 1;
    std::vector<wlx_walker_info> infos(walkers);
    if (run_checkpointing) check(wlx_engine_read_infos(e, infos.data()));
    for (;;) {
        // one epoch at a time while a 1x1 checkpoint may fall due, otherwise growing chunks;
        // a run checkpoint falls due after a known number of epochs (walker 0 gains one
        // exchange interval of attempts per epoch until its limit)
        std::uint64_t epochs = checkpointing ? 1 : chunk;
        if (run_checkpointing) {
            const std::uint64_t since = infos[0].attempted - last_checkpoint;
            const std::uint64_t due = since >= c.checkpoint.attempts ? 0 : c.checkpoint.attempts - since;
            epochs = std::min(epochs, std::max<std::uint64_t>(
                                          1, (due + c.exchange.attempts - 1) / c.exchange.attempts));
        }
        check(wlx_engine_run(e, epochs, &st));
        chunk = std::min<std::uint64_t>(chunk * 2, 4096);
        if (c.progress_interval_seconds > 0.0 && std::chrono::steady_clock::now() >= next_progress) {
            check(wlx_engine_read_infos(e, infos.data()));
            const std::string text = progress_line(c, infos, c.walkers_per_rank, n);
            std::cout << '\r' << text;
            if (progress_width > text.size())
                std::cout << std::string(progress_width - text.size(), ' ');
            std::cout << std::flush;
            progress_width = std::max(progress_width, text.size());
            next_progress = std::chrono::steady_clock::now() + period;
        }
        if (checkpointing) { // parallel.cpp:522-526
            check(wlx_engine_read_infos(e, infos.data()));
            if (infos[0].attempted - last_checkpoint >= c.checkpoint.attempts) {
                save_checkpoint(c.checkpoint_path, download(e, 0, n, bins));
                last_checkpoint = infos[0].attempted;
            }
        }
        if (run_checkpointing) {
            check(wlx_engine_read_infos(e, infos.data()));
            if (infos[0].attempted - last_checkpoint >= c.checkpoint.attempts) {
                RunCheckpoint run{st.epoch, st.exchange_attempted, st.exchange_accepted, {}};
                for (std::size_t k = 0; k < walkers; ++k) run.walkers.push_back(download(e, k, n, bins));
                save_run_checkpoint(c.checkpoint_path, run);
                last_checkpoint = infos[0].attempted;
            }
        }
        if (st.done) break;
    }
    if (c.progress_interval_seconds > 0.0 && progress_width != 0) std::cout << '\n';

    RewlResult result;
    std::vector<WalkerState> states;
    states.reserve(walkers);
    for (std::size_t k = 0; k < walkers; ++k) states.push_back(download(e, k, n, bins));
    const bool collect = c.wl.collect_window_statistics;
    std::vector<WalkerSampling> sampling(collect ? walkers : 0);
    for (std::size_t k = 0; k < sampling.size(); ++k) sampling[k] = download_sampling(e, k, n, bins);
    bool converged = true;
    for (std::size_t w = 0; w < c.windows; ++w) { // parallel.cpp:541-579
        std::vector<const WalkerState *> group;
        for (std::size_t l = 0; l < c.walkers_per_rank; ++l) {
            const WalkerState &s = states[w * c.walkers_per_rank + l];
            group.push_back(&s);
            converged &= s.stage == WLX_STAGE_FROZEN;
            const auto h = histogram_statistics(s, c.wl.nalivaiko_mod);
            const std::size_t k = w * c.walkers_per_rank + l;
            result.walker_statistics.push_back({s.id, w, 0, s.attempted, s.accepted,
                                                s.forced_accepted,
                                                s.attempted - s.last_accepted_attempt, s.energy,
                                                s.factor, h.active_bins, h.minimum, h.mean,
                                                h.min_over_mean,
                                                collect ? sampling[k].round_trips : 0});
            if (collect)
                for (const EnergyRepresentative &r : sampling[k].representatives)
                    result.representatives.push_back(r);
            result.attempted += s.attempted;
            result.accepted += s.accepted;
            result.forced_accepted += s.forced_accepted;
        }
        // a pilot window without a common active bin still contributes its statistics
        result.fragments.push_back(summarize_walkers(windows[w], group, bins, collect));
        if (collect) {
            std::vector<const WalkerSampling *> members;
            for (std::size_t l = 0; l < c.walkers_per_rank; ++l)
                members.push_back(&sampling[w * c.walkers_per_rank + l]);
            result.sampling_statistics.push_back(summarize_sampling(windows[w], members, bins));
        }
    }
    result.converged = converged;
    result.exchange_attempted = st.exchange_attempted;
    result.exchange_accepted = st.exchange_accepted;
    char device[256] = {0};
    std::uint64_t cus = 0;
    if (wlx_device_info(device, sizeof device, &cus) == WLX_OK) result.engine.device = device;
    result.engine.rng = options.rng == WLX_RNG_PHILOX ? "philox4x64-10" : "xoshiro256**";
    result.engine.fields = options.fields == WLX_FIELDS_FP32_KAHAN ? "fp32-kahan" : "fp64";
    result.engine.walkers = walkers;
    result.engine.epochs = st.epoch;
    result.engine.ambiguous_decisions = st.ambiguous_decisions;
    return result;
}

} // namespace rewl::host
