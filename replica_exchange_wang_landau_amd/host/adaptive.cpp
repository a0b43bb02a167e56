#include "adaptive.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace rewl::host {

WindowSamplingStatistics summarize_sampling(Window window,
                                            const std::vector<const WalkerSampling *> &walkers,
                                            std::size_t bins)
{ // parallel.cpp:64-79
    WindowSamplingStatistics s{window, std::vector<double>(bins, 0.0),
                               std::vector<std::uint64_t>(bins, 0), 0, walkers.size()};
    for (const WalkerSampling *w : walkers) {
        if (w->squared_energy_displacement.empty()) continue;
        for (std::size_t i = window.begin; i < window.end; ++i) {
            s.squared_energy_displacement[i] += w->squared_energy_displacement[i];
            s.displacement_samples[i] += w->displacement_samples[i];
        }
        s.round_trips += w->round_trips;
    }
    return s;
}

namespace {

constexpr double kTiny = std::numeric_limits<double>::min();

void validate_inputs(const Grid &grid, const std::vector<DosFragment> &fragments,
                     const std::vector<WindowSamplingStatistics> &sampling, std::size_t count,
                     double overlap, const AdaptiveWindowParameters &p)
{ // analysis.cpp:188-198
    const std::size_t bins = grid.bins();
    const auto finite_at_least = [](double v, double lo) { return std::isfinite(v) && v >= lo; };
    const bool ok = count != 0 && count <= bins && fragments.size() == count &&
                    sampling.size() == count && !(overlap < 0.0) && !(overlap >= 1.0) &&
                    std::isfinite(p.diffusivity_floor_fraction) &&
                    p.diffusivity_floor_fraction > 0.0 &&
                    finite_at_least(p.curvature_weight, 0.0) &&
                    finite_at_least(p.round_trip_target, 0.0) &&
                    finite_at_least(p.maximum_round_trip_penalty, 1.0) &&
                    finite_at_least(p.smoothing_width, 0.0) &&
                    finite_at_least(p.minimum_width, 0.0);
    if (!ok) throw std::invalid_argument("Invalid adaptive-window parameters");
}

// The k-th smallest value (what nth_element places at position k).
double order_statistic(std::vector<double> values, std::size_t k)
{
    std::nth_element(values.begin(), values.begin() + static_cast<std::ptrdiff_t>(k),
                     values.end());
    return values[k];
}

// Pilot evidence pooled over windows on the global grid.
struct Evidence {
    std::vector<double> displacement;      // summed squared displacement per bin
    std::vector<std::uint64_t> samples;    // attempts per bin
    std::vector<double> penalty;           // round-trip penalty per bin (>= 1)
    std::vector<double> curvature;         // summed |d2 log g / dE2| per bin
    std::vector<std::uint64_t> curvatures; // curvature estimates per bin
};

Evidence pool_evidence(const Grid &grid, const std::vector<DosFragment> &fragments,
                       const std::vector<WindowSamplingStatistics> &sampling,
                       const AdaptiveWindowParameters &p, std::size_t radius)
{ // analysis.cpp:210-244
    const std::size_t bins = grid.bins();
    Evidence ev{std::vector<double>(bins, 0.0), std::vector<std::uint64_t>(bins, 0),
                std::vector<double>(bins, 1.0), std::vector<double>(bins, 0.0),
                std::vector<std::uint64_t>(bins, 0)};
    for (std::size_t w = 0; w < sampling.size(); ++w) {
        const WindowSamplingStatistics &s = sampling[w];
        const DosFragment &f = fragments[w];
        if (s.window.begin != f.window.begin || s.window.end != f.window.end ||
            s.squared_energy_displacement.size() != bins || s.displacement_samples.size() != bins ||
            f.log_g.size() != bins || f.valid.size() != bins)
            throw std::invalid_argument("Adaptive-window pilot fragments are incompatible");
        for (std::size_t i = s.window.begin; i < s.window.end; ++i) {
            ev.displacement[i] += s.squared_energy_displacement[i];
            ev.samples[i] += s.displacement_samples[i];
        }
        // windows whose walkers completed fewer round trips than the target weigh more
        const double expected = p.round_trip_target * static_cast<double>(s.walkers);
        const double observed = static_cast<double>(s.round_trips);
        const double penalty =
            expected <= 0.0 ? 1.0
                            : std::min(p.maximum_round_trip_penalty,
                                       std::sqrt(expected / std::max(1.0, observed)));
        for (std::size_t i = s.window.begin; i < s.window.end; ++i)
            ev.penalty[i] = std::max(ev.penalty[i], penalty);
        // second difference of log g over +-radius bins, where all three are valid
        const double step = static_cast<double>(radius) * grid.width;
        for (std::size_t i = f.window.begin + radius; i + radius < f.window.end; ++i) {
            if (!f.valid[i - radius] || !f.valid[i] || !f.valid[i + radius]) continue;
            const double c =
                std::abs(f.log_g[i + radius] - 2.0 * f.log_g[i] + f.log_g[i - radius]) /
                (step * step);
            if (std::isfinite(c)) {
                ev.curvature[i] += c;
                ++ev.curvatures[i];
            }
        }
    }
    return ev;
}

// D(E): mean squared displacement per attempt over +-radius bins (long double prefix sums)
std::vector<double> local_diffusivity(const Evidence &ev, std::size_t radius)
{ // analysis.cpp:246-259
    const std::size_t bins = ev.samples.size();
    std::vector<long double> sum(bins + 1, 0.0L);
    std::vector<std::uint64_t> count(bins + 1, 0);
    for (std::size_t i = 0; i < bins; ++i) {
        sum[i + 1] = sum[i] + ev.displacement[i];
        count[i + 1] = count[i] + ev.samples[i];
    }
    std::vector<double> d(bins, 0.0);
    for (std::size_t i = 0; i < bins; ++i) {
        const std::size_t lo = i > radius ? i - radius : 0;
        const std::size_t hi = std::min(bins, i + radius + 1);
        const std::uint64_t n = count[hi] - count[lo];
        if (n != 0) d[i] = static_cast<double>((sum[hi] - sum[lo]) / static_cast<long double>(n));
    }
    return d;
}

// core boundaries at equal quantiles of the cumulative cost, each core at least min_cells
std::vector<std::size_t> core_boundaries(const std::vector<long double> &cumulative,
                                         std::size_t count, std::size_t min_cells)
{ // analysis.cpp:309-320
    const std::size_t bins = cumulative.size() - 1;
    const long double total = cumulative.back();
    std::vector<std::size_t> core(count + 1, 0);
    core[count] = bins;
    for (std::size_t k = 1; k < count; ++k) {
        const long double target =
            total * static_cast<long double>(k) / static_cast<long double>(count);
        const std::size_t at = static_cast<std::size_t>(
            std::lower_bound(cumulative.begin(), cumulative.end(), target) - cumulative.begin());
        const std::size_t lowest = core[k - 1] + min_cells;
        const std::size_t highest = bins - (count - k) * min_cells;
        core[k] = at < lowest ? lowest : (highest < at ? highest : at);
    }
    return core;
}

} // namespace

std::vector<Window> adapt_energy_windows(const Grid &grid,
                                         const std::vector<DosFragment> &fragments,
                                         const std::vector<WindowSamplingStatistics> &sampling,
                                         std::size_t count, double overlap,
                                         const AdaptiveWindowParameters &p)
{ // analysis.cpp:183-338
    grid.validate();
    const std::size_t bins = grid.bins();
    validate_inputs(grid, fragments, sampling, count, overlap, p);
    if (count == 1) return {{0, bins}};

    const double automatic = (grid.maximum - grid.minimum) / (10.0 * static_cast<double>(count));
    const double smoothing =
        std::max(grid.width, p.smoothing_width > 0.0 ? p.smoothing_width : automatic);
    const std::size_t radius =
        std::max<std::size_t>(1, static_cast<std::size_t>(std::llround(smoothing / grid.width)));

    Evidence ev = pool_evidence(grid, fragments, sampling, p, radius);
    const std::vector<double> diffusivity = local_diffusivity(ev, radius);

    // floor on D: a fraction of the median positive diffusivity
    std::vector<double> positive;
    for (double d : diffusivity)
        if (d > 0.0 && std::isfinite(d)) positive.push_back(d);
    const double typical = positive.empty() ? 1.0 : order_statistic(positive, positive.size() / 2);
    const double floor = std::max(kTiny, p.diffusivity_floor_fraction * typical);

    // curvature per bin, normalised by its 90th percentile over the bins that have one
    std::vector<double> curved;
    for (std::size_t i = 0; i < bins; ++i)
        if (ev.curvatures[i] != 0) {
            ev.curvature[i] /= static_cast<double>(ev.curvatures[i]);
            if (ev.curvature[i] > 0.0 && std::isfinite(ev.curvature[i]))
                curved.push_back(ev.curvature[i]);
        }
    double curvature_scale = 1.0;
    if (!curved.empty()) {
        const std::size_t at = std::min(
            curved.size() - 1, static_cast<std::size_t>(0.9 * static_cast<double>(curved.size())));
        curvature_scale = std::max(order_statistic(curved, at), kTiny);
    }

    std::vector<long double> cumulative(bins + 1, 0.0L);
    for (std::size_t i = 0; i < bins; ++i) {
        const double curvature = std::min(1.0, ev.curvature[i] / curvature_scale);
        const double cost = ev.penalty[i] * (1.0 + p.curvature_weight * curvature) /
                            std::sqrt(std::max(diffusivity[i], floor));
        cumulative[i + 1] = cumulative[i] + static_cast<long double>(cost * grid.width);
    }
    if (!(cumulative.back() > 0.0L)) throw std::runtime_error("Adaptive-window weight is empty");

    const double span = grid.maximum - grid.minimum;
    const double min_width =
        p.minimum_width > 0.0 ? p.minimum_width : span / (4.0 * static_cast<double>(count));
    if (!(min_width > 0.0) || min_width * static_cast<double>(count) > span)
        throw std::invalid_argument("Adaptive minimum window width is too large");
    const std::size_t min_cells = std::max<std::size_t>(
        1, static_cast<std::size_t>(std::ceil(min_width / grid.width)));
    const std::vector<std::size_t> core = core_boundaries(cumulative, count, min_cells);

    // each core is widened on both sides so that neighbours overlap by the requested fraction
    const double extension = overlap == 0.0 ? 0.0 : overlap / (2.0 * (1.0 - overlap));
    std::vector<Window> out(count);
    for (std::size_t k = 0; k < count; ++k) {
        const std::size_t extra = static_cast<std::size_t>(
            std::ceil(extension * static_cast<double>(core[k + 1] - core[k])));
        out[k].begin = k == 0 ? 0 : (core[k] > extra ? core[k] - extra : 0);
        out[k].end =
            k + 1 == count ? bins : std::min(bins - (count - 1 - k), core[k + 1] + extra);
    }
    for (std::size_t k = 1; k < count; ++k) { // strictly increasing, overlapping
        out[k].begin = std::max(out[k].begin, out[k - 1].begin + 1);
        out[k].end = std::max(out[k].end, out[k - 1].end + 1);
        if (out[k].begin >= out[k - 1].end || out[k].begin >= out[k].end)
            throw std::runtime_error("Adaptive energy windows do not overlap");
    }
    return out;
}

InitialConfigurations select_initial_configurations(
    const Grid &grid, const std::vector<Window> &windows,
    const std::vector<EnergyRepresentative> &representatives, std::size_t per_window)
{ // analysis.cpp:341-416 with one process
    if (windows.empty() || per_window == 0)
        throw std::invalid_argument("Invalid adaptive initial-configuration layout");
    InitialConfigurations out;
    out.spins.resize(windows.size() * per_window);
    const EnergyRepresentative *lowest = nullptr;
    for (const EnergyRepresentative &r : representatives)
        if (!r.spins.empty() && (!lowest || r.energy < lowest->energy)) lowest = &r;
    const auto inside = [&](const EnergyRepresentative &r, const Window &w) {
        const long long b = grid.index(r.energy);
        return b >= 0 && w.contains(static_cast<std::size_t>(b));
    };

    for (std::size_t w = 0; w < windows.size(); ++w) {
        const Window &win = windows[w];
        const double lower = grid.minimum + static_cast<double>(win.begin) * grid.width;
        const double upper = grid.minimum + static_cast<double>(win.end) * grid.width;
        const double center = 0.5 * (lower + upper);
        // the lowest window: every walker starts from the lowest-energy configuration seen
        if (w == 0 && lowest) {
            const bool external = !inside(*lowest, win);
            for (std::size_t l = 0; l < per_window; ++l) {
                out.spins[w * per_window + l] = lowest->spins;
                if (external) ++out.external_warm_starts;
            }
            continue;
        }
        // otherwise the representatives inside the window (all of them if none is), nearest
        // to the window first, then nearest to its centre, then lowest energy
        std::vector<const EnergyRepresentative *> pool;
        for (const EnergyRepresentative &r : representatives)
            if (!r.spins.empty() && inside(r, win)) pool.push_back(&r);
        const bool external = pool.empty();
        if (external)
            for (const EnergyRepresentative &r : representatives)
                if (!r.spins.empty()) pool.push_back(&r);
        const auto gap = [lower, upper](double e) {
            return e < lower ? lower - e : (e >= upper ? e - upper : 0.0);
        };
        std::sort(pool.begin(), pool.end(), [&](const EnergyRepresentative *a,
                                                const EnergyRepresentative *b) {
            const double ga = gap(a->energy), gb = gap(b->energy);
            if (ga != gb) return ga < gb;
            const double da = std::abs(a->energy - center), db = std::abs(b->energy - center);
            return da == db ? a->energy < b->energy : da < db;
        });
        for (std::size_t l = 0; l < per_window; ++l) {
            if (pool.empty()) {
                ++out.missing;
                continue;
            }
            out.spins[w * per_window + l] = pool[l % pool.size()]->spins;
            if (external) ++out.external_warm_starts;
        }
    }
    return out;
}

} // namespace rewl::host
