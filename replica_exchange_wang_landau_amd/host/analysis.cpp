#include "analysis.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace rewl::host {
namespace {

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr double kInf = std::numeric_limits<double>::infinity();

// max-shifted log(sum exp) over the finite entries (analysis.cpp:12-20)
double log_sum_exp(const std::vector<double> &values)
{
    double top = -kInf;
    for (double v : values)
        if (std::isfinite(v)) top = std::max(top, v);
    if (!std::isfinite(top)) return top;
    double sum = 0.0;
    for (double v : values)
        if (std::isfinite(v)) sum += std::exp(v - top);
    return top + std::log(sum);
}

// least-squares slope of log g over up to five valid bins around `center`, per energy unit
// (analysis.cpp:22-41)
double local_slope(const std::vector<double> &values, const std::vector<std::uint8_t> &valid,
                   std::size_t center, std::size_t begin, std::size_t end, double width)
{
    const std::size_t lo = center > 2 ? std::max(begin, center - 2) : begin;
    const std::size_t hi = std::min(end, center + 3);
    std::size_t count = 0;
    double mx = 0.0, my = 0.0;
    for (std::size_t i = lo; i < hi; ++i)
        if (valid[i] != 0 && std::isfinite(values[i])) {
            mx += static_cast<double>(i);
            my += values[i];
            ++count;
        }
    if (count < 2) return kNaN;
    mx /= static_cast<double>(count);
    my /= static_cast<double>(count);
    double num = 0.0, den = 0.0;
    for (std::size_t i = lo; i < hi; ++i)
        if (valid[i] != 0 && std::isfinite(values[i])) {
            const double dx = static_cast<double>(i) - mx;
            num += dx * (values[i] - my);
            den += dx * dx;
        }
    return den > 0.0 ? num / den / width : kNaN;
}

} // namespace

HistogramStatistics histogram_statistics(const WalkerState &w, bool nalivaiko)
{ // wang_landau.cpp:344-367
    HistogramStatistics s;
    const auto &mask = nalivaiko ? w.refinement : w.active;
    s.active_bins = nalivaiko ? w.refinement_active_bins : w.active_bins;
    long double total = 0.0L;
    s.minimum = std::numeric_limits<std::uint64_t>::max();
    for (std::size_t i = w.bins_window.begin; i < w.bins_window.end; ++i) {
        if (!mask[i]) continue;
        if (w.histogram[i] != 0) ++s.covered_bins;
        s.minimum = std::min(s.minimum, w.histogram[i]);
        total += static_cast<long double>(w.histogram[i]);
    }
    if (s.active_bins == 0) {
        s.minimum = 0;
        return s;
    }
    s.mean = static_cast<double>(total / static_cast<long double>(s.active_bins));
    s.min_over_mean = s.mean > 0.0 ? static_cast<double>(s.minimum) / s.mean : 0.0;
    s.coverage = static_cast<double>(s.covered_bins) / static_cast<double>(s.active_bins);
    return s;
}

DosFragment summarize_walkers(Window window, const std::vector<const WalkerState *> &walkers,
                              std::size_t bins, bool allow_empty)
{
    if (walkers.empty()) throw std::invalid_argument("No walkers to summarize");
    for (const auto *w : walkers)
        if (!w || w->log_g.size() != bins || w->bins_window.begin != window.begin ||
            w->bins_window.end != window.end)
            throw std::invalid_argument("Cannot summarize incompatible walkers");
    DosFragment f{window, std::vector<double>(bins, kNaN), std::vector<std::uint64_t>(bins, 0),
                  std::vector<double>(bins, kNaN), std::vector<std::uint8_t>(bins, 0)};
    // a bin is valid when every walker has discovered it
    for (std::size_t i = window.begin; i < window.end; ++i) {
        bool valid = true;
        for (const auto *w : walkers) {
            valid = valid && w->active[i] != 0;
            f.histogram[i] += w->histogram[i];
        }
        f.valid[i] = valid ? 1 : 0;
    }
    std::size_t reference = window.end;
    for (std::size_t i = window.begin; i < window.end; ++i)
        if (f.valid[i]) {
            reference = i;
            break;
        }
    if (reference == window.end) {
        if (allow_empty) return f;
        throw std::runtime_error("Walkers in an energy window have no common active bin");
    }
    const double count = static_cast<double>(walkers.size());
    for (std::size_t i = window.begin; i < window.end; ++i) {
        if (!f.valid[i]) continue;
        double mean = 0.0;
        for (const auto *w : walkers) mean += w->log_g[i] - w->log_g[reference];
        mean /= count;
        f.log_g[i] = mean;
        if (walkers.size() > 1) {
            double squares = 0.0;
            for (const auto *w : walkers) {
                const double d = (w->log_g[i] - w->log_g[reference]) - mean;
                squares += d * d;
            }
            f.standard_error[i] = std::sqrt(squares / (count * (count - 1.0)));
        }
    }
    return f;
}

DensityOfStates stitch_dos(Grid grid, std::vector<DosFragment> fragments, bool complete_range,
                           std::size_t spin_count)
{ // analysis.cpp:45-130
    grid.validate();
    if (fragments.empty()) throw std::invalid_argument("No DOS fragments supplied");
    std::sort(fragments.begin(), fragments.end(),
              [](const DosFragment &a, const DosFragment &b) {
                  return a.window.begin < b.window.begin;
              });
    const std::size_t bins = grid.bins();
    DensityOfStates out{grid,
                        std::vector<double>(bins, kNaN),
                        std::vector<std::uint64_t>(bins, 0),
                        std::vector<double>(bins, kNaN),
                        std::vector<std::uint8_t>(bins, 0),
                        {},
                        false};
    auto sized = [bins](const DosFragment &f) {
        return f.log_g.size() == bins && f.histogram.size() == bins &&
               f.standard_error.size() == bins && f.valid.size() == bins;
    };
    const DosFragment &first = fragments.front();
    if (!sized(first)) throw std::invalid_argument("DOS fragment size mismatch");
    for (std::size_t i = first.window.begin; i < first.window.end; ++i) {
        out.histogram[i] = first.histogram[i];
        out.valid[i] = first.valid[i];
        if (first.valid[i]) {
            out.log_g[i] = first.log_g[i];
            out.standard_error[i] = first.standard_error[i];
        }
    }
    std::size_t covered_end = first.window.end;
    for (std::size_t f = 1; f < fragments.size(); ++f) {
        const DosFragment &right = fragments[f];
        if (!sized(right) || right.window.begin >= covered_end)
            throw std::invalid_argument("DOS fragments must overlap and match the global grid");
        const std::size_t ob = right.window.begin, oe = std::min(covered_end, right.window.end);
        // default join: the shared valid bin nearest the overlap midpoint
        const std::size_t midpoint = ob + (oe - ob) / 2;
        std::size_t join = oe, best_distance = std::numeric_limits<std::size_t>::max();
        for (std::size_t i = ob; i < oe; ++i)
            if (out.valid[i] && right.valid[i] && std::isfinite(out.log_g[i]) &&
                std::isfinite(right.log_g[i])) {
                const std::size_t d = i > midpoint ? i - midpoint : midpoint - i;
                if (d < best_distance) {
                    best_distance = d;
                    join = i;
                }
            }
        if (join == oe)
            throw std::runtime_error("Adjacent DOS fragments have no common valid overlap bin");
        // preferred join: where the two local inverse temperatures agree best
        double best = kInf;
        for (std::size_t i = ob + 2; i + 2 < oe; ++i) {
            if (!out.valid[i] || !right.valid[i]) continue;
            const double a = local_slope(out.log_g, out.valid, i, ob, oe, grid.width);
            const double b = local_slope(right.log_g, right.valid, i, ob, oe, grid.width);
            const double diff = std::abs(a - b);
            if (std::isfinite(diff) && diff < best) {
                best = diff;
                join = i;
            }
        }
        const double shift = out.log_g[join] - right.log_g[join];
        out.join_bins.push_back(join);
        for (std::size_t i = join; i < right.window.end; ++i) {
            out.histogram[i] = right.histogram[i];
            out.valid[i] = right.valid[i];
            if (right.valid[i]) {
                out.log_g[i] = right.log_g[i] + shift;
                out.standard_error[i] = right.standard_error[i];
            } else {
                out.log_g[i] = kNaN;
                out.standard_error[i] = kNaN;
            }
        }
        covered_end = std::max(covered_end, right.window.end);
    }
    if (complete_range) {
        if (std::any_of(out.valid.begin(), out.valid.end(), [](auto v) { return v == 0; }))
            throw std::runtime_error(
                "Complete-range normalization requires every DOS bin to be valid");
        const double shift =
            static_cast<double>(spin_count) * std::log(2.0) - log_sum_exp(out.log_g);
        for (std::size_t i = 0; i < bins; ++i)
            if (out.valid[i]) out.log_g[i] += shift;
        out.fully_normalized = true;
    } else {
        if (!std::isfinite(log_sum_exp(out.log_g)))
            throw std::runtime_error("Stitched DOS has no valid bins");
        double peak = -kInf;
        for (std::size_t i = 0; i < bins; ++i)
            if (out.valid[i]) peak = std::max(peak, out.log_g[i]);
        for (std::size_t i = 0; i < bins; ++i)
            if (out.valid[i]) out.log_g[i] -= peak;
    }
    return out;
}

std::vector<ThermodynamicPoint> thermodynamics(const DensityOfStates &dos,
                                               const std::vector<double> &temperatures,
                                               double k)
{ // analysis.cpp:132-160
    if (!(k > 0.0)) throw std::invalid_argument("Boltzmann constant must be positive");
    if (dos.valid.size() != dos.log_g.size())
        throw std::invalid_argument("DOS valid-mask size mismatch");
    std::vector<ThermodynamicPoint> out;
    out.reserve(temperatures.size());
    std::vector<double> weights(dos.log_g.size());
    for (double t : temperatures) {
        if (!(t > 0.0)) throw std::invalid_argument("Temperature must be positive");
        const double beta = 1.0 / (k * t);
        for (std::size_t i = 0; i < weights.size(); ++i)
            weights[i] = dos.valid[i] ? dos.log_g[i] - beta * dos.grid.center(i) : -kInf;
        const double log_z = log_sum_exp(weights);
        if (!std::isfinite(log_z))
            throw std::runtime_error("DOS has no valid bins for thermodynamics");
        double e1 = 0.0, e2 = 0.0;
        for (std::size_t i = 0; i < weights.size(); ++i) {
            if (!dos.valid[i]) continue;
            const double p = std::exp(weights[i] - log_z);
            const double e = dos.grid.center(i);
            e1 += p * e;
            e2 += p * e * e;
        }
        const double heat = (e2 - e1 * e1) / (k * t * t);
        const double free = -k * t * log_z;
        out.push_back({t, log_z, e1, heat, free, (e1 - free) / t});
    }
    return out;
}

} // namespace rewl::host
