// ORACLE — test infrastructure, not product code.
//
// extern "C" driver around the compiled reference (Faikert/replica_exchange_wang_landau, built
// from its own sources by oracle/Makefile into oracle/_ref/). It exists only to pin the C
// oracle (wl_oracle.c) and to write the golden fixtures under tests/golden/. It drives the
// reference's public C++ API: Geometry + DenseCouplings/CsrCouplings (physics.hpp:22-84),
// WangLandauWalker (wang_landau.hpp:91-138) and run_rewl (parallel.hpp:35-36).
#include "wl/analysis.hpp"
#include "wl/parallel.hpp"

#include "../include/rewl.h"

#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

namespace {
thread_local std::string g_error;

int report(const std::exception &error)
{
    g_error = error.what();
    if (dynamic_cast<const std::invalid_argument *>(&error)) return WLX_INVALID_ARGUMENT;
    if (dynamic_cast<const std::out_of_range *>(&error)) return WLX_OUT_OF_RANGE;
    if (dynamic_cast<const std::logic_error *>(&error)) return WLX_LOGIC_ERROR;
    return WLX_RUNTIME_ERROR;
}

struct RefSystem {
    wl::Geometry geometry;
    std::shared_ptr<const wl::Couplings> couplings;
};

wl::WlParameters parameters(const wlx_params &p)
{
    wl::WlParameters q;
    q.flatness = p.flatness;
    q.minimum_visits = p.minimum_visits;
    q.final_factor = p.final_factor;
    q.check_interval_attempts = p.check_interval_attempts;
    q.force_accept_after_attempts = p.force_accept_after_attempts;
    q.inverse_time_enabled = p.inverse_time_enabled != 0;
    q.initialization_max_attempts = p.initialization_max_attempts;
    q.initialization_target_fraction = p.initialization_target_fraction;
    q.initialization_temperature_fraction = p.initialization_temperature_fraction;
    q.initialization_stall_attempts_per_spin = p.initialization_stall_attempts_per_spin;
    q.initialization_temperature_multiplier = p.initialization_temperature_multiplier;
    q.initialization_max_temperature_fraction = p.initialization_max_temperature_fraction;
    q.nalivaiko_mod = p.nalivaiko_mod != 0;
    q.collect_window_statistics = p.collect_window_statistics != 0;
    q.round_trip_margin_fraction = p.round_trip_margin_fraction;
    return q;
}

void fill_info(const wl::WangLandauWalker &w, wlx_walker_info *info)
{
    const auto s = w.snapshot();
    std::memset(info, 0, sizeof *info);
    info->walker_id = s.walker_id;
    info->window_begin = w.window().begin;
    info->window_end = w.window().end;
    info->energy = s.energy;
    info->factor = s.factor;
    info->attempted = s.attempted;
    info->accepted = s.accepted;
    info->forced_accepted = s.forced_accepted;
    info->last_accepted_attempt = s.last_accepted_attempt;
    info->stage = static_cast<uint64_t>(s.stage);
    info->active_bins = w.histogram_statistics().active_bins;
    for (int k = 0; k < 4; ++k) info->rng_state[k] = s.rng_state[k];
}
} // namespace

extern "C" {

size_t ref_last_error(char *buffer, size_t capacity)
{
    if (buffer && capacity) {
        const size_t n = std::min(g_error.size(), capacity - 1);
        std::memcpy(buffer, g_error.data(), n);
        buffer[n] = 0;
    }
    return g_error.size();
}

void ref_xoshiro(uint64_t seed, uint64_t stream, uint64_t count, uint64_t *out, uint64_t *state)
{
    wl::Xoshiro256StarStar rng(seed, stream);
    for (uint64_t k = 0; k < count; ++k) out[k] = rng();
    if (state)
        for (int k = 0; k < 4; ++k) state[k] = rng.state()[k];
}

void ref_xoshiro_bounded(uint64_t seed, uint64_t stream, uint64_t bound, uint64_t count,
                         uint64_t *out)
{
    wl::Xoshiro256StarStar rng(seed, stream);
    for (uint64_t k = 0; k < count; ++k) out[k] = rng.bounded(bound);
}

int ref_dipolar_pair(const double *d, const double *ai, const double *aj, double scale,
                     double *out)
{
    try {
        *out = wl::dipolar_pair({d[0], d[1], d[2]}, {ai[0], ai[1], ai[2]}, {aj[0], aj[1], aj[2]},
                                scale);
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

void ref_minimum_image(const double *d, const double *box, int periodic, double *out)
{
    const auto r = wl::minimum_image({d[0], d[1], d[2]}, {{box[0], box[1], box[2]}, periodic != 0});
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

int ref_simple_cubic(uint64_t nx, uint64_t ny, uint64_t nz, double spacing, const double *axis,
                     int periodic, double *positions, double *axes, double *box)
{
    try {
        const auto g = wl::Geometry::simple_cubic(nx, ny, nz, spacing,
                                                  {axis[0], axis[1], axis[2]}, periodic != 0);
        for (size_t i = 0; i < g.size(); ++i) {
            positions[3 * i] = g.positions[i].x;
            positions[3 * i + 1] = g.positions[i].y;
            positions[3 * i + 2] = g.positions[i].z;
            axes[3 * i] = g.axes[i].x;
            axes[3 * i + 1] = g.axes[i].y;
            axes[3 * i + 2] = g.axes[i].z;
        }
        box[0] = g.box.lengths.x;
        box[1] = g.box.lengths.y;
        box[2] = g.box.lengths.z;
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

// Geometry::load_xyz_axes (csv = 0, physics.cpp:77-97) or Geometry::load_csv (csv = 1,
// physics.cpp:99-149) of `path`: *n sites; positions/axes (N x 3) are written when *n <= capacity
int ref_load_geometry(const char *path, int csv, const double *box, int periodic,
                      uint64_t capacity, double *positions, double *axes, uint64_t *n)
{
    try {
        const wl::Box b{{box[0], box[1], box[2]}, periodic != 0};
        const auto g = csv ? wl::Geometry::load_csv(path, b) : wl::Geometry::load_xyz_axes(path, b);
        *n = g.size();
        if (g.size() <= capacity)
            for (size_t i = 0; i < g.size(); ++i) {
                const double p[3] = {g.positions[i].x, g.positions[i].y, g.positions[i].z};
                const double a[3] = {g.axes[i].x, g.axes[i].y, g.axes[i].z};
                std::memcpy(positions + 3 * i, p, sizeof p);
                std::memcpy(axes + 3 * i, a, sizeof a);
            }
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

void *ref_system_create(const double *positions, const double *axes, uint64_t n,
                        const double *box, int periodic, double scale, double cutoff)
{
    try {
        auto s = std::make_unique<RefSystem>();
        s->geometry.box = {{box[0], box[1], box[2]}, periodic != 0};
        for (uint64_t i = 0; i < n; ++i) {
            s->geometry.positions.push_back(
                {positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]});
            s->geometry.axes.push_back({axes[3 * i], axes[3 * i + 1], axes[3 * i + 2]});
        }
        if (cutoff > 0.0)
            s->couplings = std::make_shared<wl::CsrCouplings>(s->geometry, scale, cutoff);
        else
            s->couplings = std::make_shared<wl::DenseCouplings>(s->geometry, scale);
        return s.release();
    } catch (const std::exception &e) {
        report(e);
        return nullptr;
    }
}

void ref_system_destroy(void *system) { delete static_cast<RefSystem *>(system); }

void ref_system_dense(void *system, double *out)
{
    const auto &c = *static_cast<RefSystem *>(system)->couplings;
    const auto n = c.size();
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) out[i * n + j] = c.at(i, j);
}

void ref_system_fields(void *system, const int8_t *spins, double *fields, double *energy)
{
    const auto &c = *static_cast<RefSystem *>(system)->couplings;
    std::vector<int8_t> s(spins, spins + c.size());
    const auto h = wl::local_fields(c, s);
    std::copy(h.begin(), h.end(), fields);
    *energy = wl::total_energy(c, s);
}

void *ref_walker_create(void *system, uint64_t id, double emin, double emax, double width,
                        uint64_t begin, uint64_t end, const wlx_params *p, uint64_t seed,
                        const int8_t *initial)
{
    try {
        auto *s = static_cast<RefSystem *>(system);
        std::vector<int8_t> spins;
        if (initial) spins.assign(initial, initial + s->couplings->size());
        return new wl::WangLandauWalker(id, s->couplings, {emin, emax, width}, {begin, end},
                                        parameters(*p), seed, std::move(spins));
    } catch (const std::exception &e) {
        report(e);
        return nullptr;
    }
}

void ref_walker_destroy(void *walker) { delete static_cast<wl::WangLandauWalker *>(walker); }

// Runs `count` attempts; records per attempt the accepted flag and the bin after it.
int ref_walker_attempts(void *walker, uint64_t count, uint8_t *accepted, int64_t *bins)
{
    try {
        auto &w = *static_cast<wl::WangLandauWalker *>(walker);
        for (uint64_t k = 0; k < count; ++k) {
            const bool a = w.attempt_flip();
            if (accepted) accepted[k] = a ? 1 : 0;
            if (bins) bins[k] = static_cast<int64_t>(w.energy_bin().value_or(SIZE_MAX));
        }
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

int ref_walker_read(void *walker, wlx_walker_info *info, int8_t *spins, double *fields,
                    double *log_g, uint64_t *histogram, uint8_t *active)
{
    const auto &w = *static_cast<wl::WangLandauWalker *>(walker);
    const auto s = w.snapshot();
    fill_info(w, info);
    std::copy(s.spins.begin(), s.spins.end(), spins);
    std::copy(s.fields.begin(), s.fields.end(), fields);
    std::copy(s.log_g.begin(), s.log_g.end(), log_g);
    std::copy(s.histogram.begin(), s.histogram.end(), histogram);
    std::copy(s.active.begin(), s.active.end(), active);
    return WLX_OK;
}

int ref_walker_statistics(void *walker, wlx_histogram_stats *out)
{
    const auto st = static_cast<wl::WangLandauWalker *>(walker)->histogram_statistics();
    out->active_bins = st.active_bins;
    out->covered_bins = st.covered_bins;
    out->minimum = st.minimum;
    out->mean = st.mean;
    out->min_over_mean = st.min_over_mean;
    out->coverage = st.coverage;
    return WLX_OK;
}

int ref_walker_next_iteration(void *walker)
{
    try {
        static_cast<wl::WangLandauWalker *>(walker)->begin_next_iteration();
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

// the reference RunConfig of a wlx_config
static wl::RunConfig run_config(const wlx_config *c)
{
    wl::RunConfig config;
    config.grid = {c->emin, c->emax, c->bin_width};
    config.energy_grid_explicit = true;
    config.windows = c->windows;
    config.walkers_per_rank = c->walkers_per_window;
    config.overlap = c->overlap;
    config.seed = c->seed;
    config.seed_explicit = true;
    config.exchange_interval_attempts = c->exchange_interval_attempts;
    config.max_attempts = c->max_attempts;
    config.wl = parameters(c->wl);
    config.checkpoint_interval_attempts = 1;
    if (c->explicit_windows)
        for (uint64_t w = 0; w < c->windows; ++w)
            config.explicit_windows.push_back(
                {c->explicit_windows[2 * w], c->explicit_windows[2 * w + 1]});
    return config;
}

static wl::RewlResult run(void *system, const wl::RunConfig &config)
{
    auto *s = static_cast<RefSystem *>(system);
    int argc = 1;
    char name[] = "ref";
    char *args[] = {name, nullptr};
    char **argv = args;
    wl::ParallelContext context(argc, argv);
    return wl::run_rewl(context, s->couplings, config);
}

// run_rewl on the system (shared-memory path). Outputs, all caller-allocated:
//   fragments: log_g / standard_error [windows*bins], histogram [windows*bins], valid [...]
//   per-walker: energy, factor, attempted, accepted, forced, since_last, active_bins, min_h,
//               mean_h, min_over_mean  (walker_statistics order = window-major)
//   counters[5]: attempted, accepted, forced, exchange_attempted, exchange_accepted;
//   *converged.
int ref_run_rewl(void *system, const wlx_config *c, double *frag_log_g, double *frag_error,
                 uint64_t *frag_histogram, uint8_t *frag_valid, double *walker_doubles,
                 uint64_t *walker_integers, uint64_t *counters, uint64_t *converged)
{
    try {
        const wl::RunConfig config = run_config(c);
        const auto result = run(system, config);
        const auto bins = config.grid.bins();
        for (size_t w = 0; w < result.fragments.size(); ++w) {
            const auto &f = result.fragments[w];
            std::copy(f.log_g.begin(), f.log_g.end(), frag_log_g + w * bins);
            std::copy(f.standard_error.begin(), f.standard_error.end(), frag_error + w * bins);
            std::copy(f.histogram.begin(), f.histogram.end(), frag_histogram + w * bins);
            std::copy(f.valid.begin(), f.valid.end(), frag_valid + w * bins);
        }
        for (size_t k = 0; k < result.walker_statistics.size(); ++k) {
            const auto &st = result.walker_statistics[k];
            walker_doubles[4 * k] = st.energy;
            walker_doubles[4 * k + 1] = st.factor;
            walker_doubles[4 * k + 2] = st.mean_histogram;
            walker_doubles[4 * k + 3] = st.min_over_mean;
            uint64_t *wi = walker_integers + 8 * k;
            wi[0] = st.walker_id;
            wi[1] = st.window_id;
            wi[2] = st.attempted;
            wi[3] = st.accepted;
            wi[4] = st.forced_accepted;
            wi[5] = st.attempts_since_last_accepted;
            wi[6] = st.active_bins;
            wi[7] = st.minimum_histogram;
        }
        counters[0] = result.attempted;
        counters[1] = result.accepted;
        counters[2] = result.forced_accepted;
        counters[3] = result.exchange_attempted;
        counters[4] = result.exchange_accepted;
        *converged = result.converged ? 1 : 0;
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

// the pilot statistics of a reference walker built with collect_window_statistics
int ref_walker_sampling(void *walker, double *squared, uint64_t *samples, uint64_t *round_trips,
                        double *rep_energy, uint8_t *rep_present, int8_t *rep_spins, uint64_t n)
{
    try {
        const auto *w = static_cast<const wl::WangLandauWalker *>(walker);
        const auto &sq = w->squared_energy_displacement();
        const auto &ds = w->displacement_samples();
        std::copy(sq.begin(), sq.end(), squared);
        std::copy(ds.begin(), ds.end(), samples);
        *round_trips = w->round_trips();
        const auto &reps = w->representatives();
        for (size_t r = 0; r < reps.size(); ++r) {
            rep_present[r] = reps[r].spins.empty() ? 0 : 1;
            rep_energy[r] = reps[r].energy;
            if (!reps[r].spins.empty())
                std::copy(reps[r].spins.begin(), reps[r].spins.end(), rep_spins + r * n);
        }
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

// stitch_dos (analysis.cpp:45-130) over `count` fragments of `bins` bins each; fragment f
// covers [begin_end[2f], begin_end[2f+1]). Writes the stitched log_g and valid mask.
int ref_stitch(double emin, double emax, double width, uint64_t bins, uint64_t count,
               const uint64_t *begin_end, const double *log_g, const double *error,
               const uint64_t *histogram, const uint8_t *valid, int complete_range,
               uint64_t spins, double *out_log_g, uint8_t *out_valid)
{
    try {
        std::vector<wl::DosFragment> fragments;
        for (uint64_t f = 0; f < count; ++f) {
            wl::DosFragment d{{begin_end[2 * f], begin_end[2 * f + 1]},
                              std::vector<double>(log_g + f * bins, log_g + (f + 1) * bins),
                              std::vector<std::uint64_t>(histogram + f * bins,
                                                         histogram + (f + 1) * bins),
                              std::vector<double>(error + f * bins, error + (f + 1) * bins),
                              std::vector<std::uint8_t>(valid + f * bins, valid + (f + 1) * bins)};
            fragments.push_back(std::move(d));
        }
        const auto dos = wl::stitch_dos({emin, emax, width}, fragments, complete_range != 0,
                                        spins);
        std::copy(dos.log_g.begin(), dos.log_g.end(), out_log_g);
        std::copy(dos.valid.begin(), dos.valid.end(), out_valid);
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

// run_rewl with collect_window_statistics: per window the summed squared displacements and
// samples (windows x bins), round trips and walker counts; per walker its round trips; the
// recorded representatives in result order (energies, n spins each) and their count.
int ref_run_rewl_sampling(void *system, const wlx_config *c, double *squared, uint64_t *samples,
                          uint64_t *round_trips, uint64_t *walkers, uint64_t *walker_round_trips,
                          double *rep_energies, int8_t *rep_spins, uint64_t *rep_count)
{
    try {
        wl::RunConfig config = run_config(c);
        config.wl.collect_window_statistics = true;
        const auto result = run(system, config);
        const auto bins = config.grid.bins();
        for (size_t w = 0; w < result.sampling_statistics.size(); ++w) {
            const auto &s = result.sampling_statistics[w];
            std::copy(s.squared_energy_displacement.begin(), s.squared_energy_displacement.end(),
                      squared + w * bins);
            std::copy(s.displacement_samples.begin(), s.displacement_samples.end(),
                      samples + w * bins);
            round_trips[w] = s.round_trips;
            walkers[w] = s.walkers;
        }
        for (size_t k = 0; k < result.walker_statistics.size(); ++k)
            walker_round_trips[k] = result.walker_statistics[k].round_trips;
        size_t r = 0;
        for (const auto &rep : result.representatives) {
            rep_energies[r] = rep.energy;
            std::copy(rep.spins.begin(), rep.spins.end(), rep_spins + r * rep.spins.size());
            ++r;
        }
        *rep_count = r;
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

// adapt_energy_windows (analysis.cpp:183-338). Fragment and sampling arrays are `count`
// blocks of `bins` values; params = {smoothing_width, minimum_width,
// diffusivity_floor_fraction, curvature_weight, round_trip_target, maximum_round_trip_penalty}.
int ref_adapt_windows(double emin, double emax, double width, uint64_t bins, uint64_t count,
                      double overlap, const double *params, const uint64_t *begin_end,
                      const double *log_g, const uint8_t *valid, const double *squared,
                      const uint64_t *samples, const uint64_t *round_trips,
                      const uint64_t *walkers, uint64_t *out_begin_end)
{
    try {
        std::vector<wl::DosFragment> fragments;
        std::vector<wl::WindowSamplingStatistics> sampling;
        for (uint64_t f = 0; f < count; ++f) {
            const wl::EnergyWindow w{begin_end[2 * f], begin_end[2 * f + 1]};
            fragments.push_back({w, std::vector<double>(log_g + f * bins, log_g + (f + 1) * bins),
                                 std::vector<std::uint64_t>(bins, 0),
                                 std::vector<double>(bins, 0.0),
                                 std::vector<std::uint8_t>(valid + f * bins, valid + (f + 1) * bins)});
            sampling.push_back({w, std::vector<double>(squared + f * bins, squared + (f + 1) * bins),
                                std::vector<std::uint64_t>(samples + f * bins,
                                                           samples + (f + 1) * bins),
                                round_trips[f], walkers[f]});
        }
        wl::AdaptiveWindowParameters p;
        p.smoothing_width = params[0];
        p.minimum_width = params[1];
        p.diffusivity_floor_fraction = params[2];
        p.curvature_weight = params[3];
        p.round_trip_target = params[4];
        p.maximum_round_trip_penalty = params[5];
        const auto windows =
            wl::adapt_energy_windows({emin, emax, width}, fragments, sampling, count, overlap, p);
        for (std::size_t k = 0; k < windows.size(); ++k) {
            out_begin_end[2 * k] = windows[k].begin;
            out_begin_end[2 * k + 1] = windows[k].end;
        }
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

// select_adaptive_initial_configurations (analysis.cpp:341-416) on one process: representative
// r has energies[r] and spins[r*n..] (empty where present[r] == 0). Writes per walker its
// spins and whether it got any, plus the missing and external warm-start counts.
int ref_select_initial(double emin, double emax, double width, uint64_t count,
                       const uint64_t *begin_end, uint64_t reps, const double *energies,
                       const uint8_t *present, uint64_t n, const int8_t *spins,
                       uint64_t per_window, int8_t *out_spins, uint8_t *out_present,
                       uint64_t *missing, uint64_t *external)
{
    try {
        std::vector<wl::EnergyWindow> windows;
        for (uint64_t k = 0; k < count; ++k) windows.push_back({begin_end[2 * k], begin_end[2 * k + 1]});
        std::vector<wl::EnergyRepresentative> representatives(reps);
        for (uint64_t r = 0; r < reps; ++r) {
            representatives[r].energy = energies[r];
            if (present[r]) representatives[r].spins.assign(spins + r * n, spins + (r + 1) * n);
        }
        std::size_t m = 0, x = 0;
        const auto out = wl::select_adaptive_initial_configurations(
            {emin, emax, width}, windows, representatives, 1, per_window, m, x);
        for (std::size_t k = 0; k < out.size(); ++k) {
            out_present[k] = out[k].empty() ? 0 : 1;
            if (!out[k].empty()) std::copy(out[k].begin(), out[k].end(), out_spins + k * n);
        }
        *missing = m;
        *external = x;
        return WLX_OK;
    } catch (const std::exception &e) {
        return report(e);
    }
}

} // extern "C"
