// wl_run configuration: INI + command-line parser, MCS resolution and validation.
// Interface (keys, units, defaults, messages) follows io.cpp:75-396 of the reference.
#include "config.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>

namespace rewl::host {

// ---- grid and windows -------------------------------------------------------------------------
void Grid::validate() const
{
    if (!std::isfinite(minimum) || !std::isfinite(maximum) || !std::isfinite(width) ||
        !(maximum > minimum) || !(width > 0.0))
        throw std::invalid_argument("Invalid energy grid");
    const double count = (maximum - minimum) / width;
    if (std::abs(count - std::round(count)) > 1e-9 * std::max(1.0, count))
        throw std::invalid_argument("Energy range must be an integer multiple of bin width");
}

std::size_t Grid::bins() const
{
    validate();
    return static_cast<std::size_t>(std::llround((maximum - minimum) / width));
}

long long Grid::index(double energy) const
{
    if (!std::isfinite(energy) || energy < minimum || energy > maximum) return -1;
    const auto count = static_cast<std::size_t>(std::llround((maximum - minimum) / width));
    auto raw = static_cast<std::size_t>(std::floor((energy - minimum) / width));
    if (raw == count && energy == maximum) raw = count - 1;
    return raw < count ? static_cast<long long>(raw) : -1;
}

double Grid::center(std::size_t i) const
{
    return minimum + (static_cast<double>(i) + 0.5) * width;
}

std::vector<Window> partition_windows(std::size_t bins, std::size_t count, double overlap)
{
    if (bins == 0 || count == 0 || count > bins || overlap < 0.0 || overlap >= 1.0)
        throw std::invalid_argument("Invalid energy-window partition");
    if (count == 1) return {{0, bins}};
    const double denominator = 1.0 + static_cast<double>(count - 1) * (1.0 - overlap);
    std::size_t size =
        static_cast<std::size_t>(std::ceil(static_cast<double>(bins) / denominator));
    size = std::max<std::size_t>(2, std::min(size, bins));
    const std::size_t span = bins - size;
    std::vector<Window> out(count);
    for (std::size_t w = 0; w < count; ++w) {
        const auto begin = static_cast<std::size_t>(
            std::llround(static_cast<double>(w * span) / static_cast<double>(count - 1)));
        out[w] = {begin, begin + size};
    }
    out.back().end = bins;
    for (std::size_t w = 1; w < count; ++w)
        if (out[w].begin >= out[w - 1].end)
            throw std::invalid_argument("Energy windows do not overlap; increase overlap or bins");
    return out;
}

// ---- value parsing ------------------------------------------------------------------------
namespace {

template <class T> T parse_integer(const std::string &text, const std::string &option)
{
    T value{};
    const auto [end, error] = std::from_chars(text.data(), text.data() + text.size(), value);
    if (error != std::errc{} || end != text.data() + text.size())
        throw std::invalid_argument("Invalid value for " + option);
    return value;
}

double parse_double(const std::string &text, const std::string &option)
{
    std::size_t used = 0;
    const double value = std::stod(text, &used);
    if (used != text.size()) throw std::invalid_argument("Invalid value for " + option);
    return value;
}

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(),
                   [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
    return s;
}

bool parse_bool(const std::string &text, const std::string &option)
{
    const std::string v = lower(text);
    if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
    if (v == "0" || v == "false" || v == "no" || v == "off") return false;
    throw std::invalid_argument("Invalid boolean for " + option);
}

std::vector<double> parse_list(const std::string &text, const std::string &option);

std::string strip(const std::string &value, bool unquote)
{
    const auto first = value.find_first_not_of(" \t\r\n");
    if (first == std::string::npos) return {};
    const auto last = value.find_last_not_of(" \t\r\n");
    std::string s = value.substr(first, last - first + 1);
    if (unquote && s.size() >= 2 &&
        ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        s = s.substr(1, s.size() - 2);
    return s;
}

std::vector<double> parse_list(const std::string &text, const std::string &option)
{
    std::vector<double> out;
    std::stringstream stream(text);
    std::string item;
    while (std::getline(stream, item, ',')) out.push_back(parse_double(strip(item, true), option));
    if (out.empty()) throw std::invalid_argument("Empty list for " + option);
    return out;
}

// A setting receives the value text and the name under which it was given (for messages).
struct Context {
    RunConfig &c;
    bool &have_min, &have_max, &have_width;
    std::filesystem::path ini_directory; // empty on the command line
};
using Setter = std::function<void(Context &, const std::string &value, const std::string &name)>;

Setter size_value(std::size_t RunConfig::*field)
{
    return [field](Context &x, const std::string &v, const std::string &n) {
        x.c.*field = parse_integer<std::size_t>(v, n);
    };
}

template <class Member> Setter double_value(Member get)
{
    return [get](Context &x, const std::string &v, const std::string &n) {
        get(x.c) = parse_double(v, n);
    };
}

template <class Member> Setter u64_value(Member get)
{
    return [get](Context &x, const std::string &v, const std::string &n) {
        get(x.c) = parse_integer<std::uint64_t>(v, n);
    };
}

template <class Member> Setter bool_value(Member get)
{
    return [get](Context &x, const std::string &v, const std::string &n) {
        get(x.c) = parse_bool(v, n);
    };
}

Setter mcs_value(Interval RunConfig::*field)
{
    return [field](Context &x, const std::string &v, const std::string &n) {
        (x.c.*field).mcs = parse_double(v, n);
        (x.c.*field).uses_mcs = true;
    };
}

Setter attempts_value(Interval RunConfig::*field)
{
    return [field](Context &x, const std::string &v, const std::string &n) {
        (x.c.*field).attempts = parse_integer<std::uint64_t>(v, n);
        (x.c.*field).uses_mcs = false;
    };
}

#define FIELD(expr) [](RunConfig & c) -> auto & { return expr; }

// Settings shared by both front ends; each entry lists its INI keys and its CLI flag.
struct Setting {
    std::vector<std::string> ini;
    std::string flag; // empty: INI only
    Setter set;
};

const std::vector<Setting> &settings()
{
    static const std::vector<Setting> table = {
        {{"geometry.file", "geometry.csv"}, "--geometry",
         [](Context &x, const std::string &v, const std::string &) {
             std::filesystem::path p(v);
             if (!x.ini_directory.empty() && p.is_relative()) p = x.ini_directory / p;
             x.c.geometry_file = x.ini_directory.empty() ? v : p.lexically_normal().string();
         }},
        {{"geometry.nx"}, "--nx", size_value(&RunConfig::nx)},
        {{"geometry.ny"}, "--ny", size_value(&RunConfig::ny)},
        {{"geometry.nz"}, "--nz", size_value(&RunConfig::nz)},
        {{"geometry.spacing"}, "--spacing", double_value(FIELD(c.spacing))},
        {{"geometry.axis_x"}, "--axis-x", double_value(FIELD(c.axis.x))},
        {{"geometry.axis_y"}, "--axis-y", double_value(FIELD(c.axis.y))},
        {{"geometry.axis_z"}, "--axis-z", double_value(FIELD(c.axis.z))},
        {{"geometry.periodic"}, "--periodic", bool_value(FIELD(c.periodic))},
        {{"geometry.box_x"}, "--box-x", double_value(FIELD(c.box_lengths.x))},
        {{"geometry.box_y"}, "--box-y", double_value(FIELD(c.box_lengths.y))},
        {{"geometry.box_z"}, "--box-z", double_value(FIELD(c.box_lengths.z))},
        {{"physics.coupling", "physics.coupling_scale"}, "--coupling",
         double_value(FIELD(c.coupling_scale))},
        {{"physics.cutoff"}, "--cutoff", double_value(FIELD(c.cutoff))},
        {{"energy.emin", "energy.minimum"}, "--emin",
         [](Context &x, const std::string &v, const std::string &n) {
             x.c.grid.minimum = parse_double(v, n);
             x.have_min = true;
         }},
        {{"energy.emax", "energy.maximum"}, "--emax",
         [](Context &x, const std::string &v, const std::string &n) {
             x.c.grid.maximum = parse_double(v, n);
             x.have_max = true;
         }},
        {{"energy.bin_width", "energy.width"}, "--bin-width",
         [](Context &x, const std::string &v, const std::string &n) {
             x.c.grid.width = parse_double(v, n);
             x.have_width = true;
         }},
        {{"energy.complete_range"}, "--complete-range", bool_value(FIELD(c.complete_range))},
        {{"parallel.windows"}, "--windows", size_value(&RunConfig::windows)},
        {{"parallel.walkers", "parallel.walkers_per_rank"}, "--walkers",
         size_value(&RunConfig::walkers_per_rank)},
        {{"parallel.overlap"}, "--overlap", double_value(FIELD(c.overlap))},
        {{"parallel.exchange_interval"}, "--exchange-interval",
         attempts_value(&RunConfig::exchange)},
        {{"parallel.exchange_interval_mcs"}, "--exchange-interval-mcs",
         mcs_value(&RunConfig::exchange)},
        {{"adaptive_windows.enabled"}, "--adaptive-windows",
         bool_value(FIELD(c.adaptive_windows.enabled))},
        {{"adaptive_windows.iterations"}, "--adaptive-iterations",
         [](Context &x, const std::string &v, const std::string &n) {
             x.c.adaptive_windows.iterations = parse_integer<std::size_t>(v, n);
         }},
        {{"adaptive_windows.pilot_mcs"}, "--adaptive-pilot-mcs",
         double_value(FIELD(c.adaptive_windows.pilot_mcs))},
        {{"adaptive_windows.smoothing_width"}, "--adaptive-smoothing-width",
         double_value(FIELD(c.adaptive_windows.smoothing_width))},
        {{"adaptive_windows.minimum_width"}, "--adaptive-minimum-width",
         double_value(FIELD(c.adaptive_windows.minimum_width))},
        {{"adaptive_windows.diffusivity_floor_fraction"}, "--adaptive-diffusivity-floor",
         double_value(FIELD(c.adaptive_windows.diffusivity_floor_fraction))},
        {{"adaptive_windows.curvature_weight"}, "--adaptive-curvature-weight",
         double_value(FIELD(c.adaptive_windows.curvature_weight))},
        {{"adaptive_windows.round_trip_target"}, "--adaptive-round-trip-target",
         double_value(FIELD(c.adaptive_windows.round_trip_target))},
        {{"adaptive_windows.maximum_round_trip_penalty"}, "--adaptive-round-trip-penalty",
         double_value(FIELD(c.adaptive_windows.maximum_round_trip_penalty))},
        {{"adaptive_windows.round_trip_margin_fraction"}, "--adaptive-round-trip-margin",
         double_value(FIELD(c.wl.round_trip_margin_fraction))},
        {{"wl.flatness"}, "--flatness", double_value(FIELD(c.wl.flatness))},
        {{"wl.minimum_visits", "wl.min_visits"}, "--min-visits",
         u64_value(FIELD(c.wl.minimum_visits))},
        {{"wl.final_factor"}, "--final-factor", double_value(FIELD(c.wl.final_factor))},
        {{"wl.inverse_time"}, "--inverse-time", bool_value(FIELD(c.wl.inverse_time_enabled))},
        {{"wl.nalivaiko_mod"}, "", bool_value(FIELD(c.wl.nalivaiko_mod))},
        {{"wl.check_interval"}, "--check-interval", attempts_value(&RunConfig::check)},
        {{"wl.check_interval_mcs"}, "--check-interval-mcs", mcs_value(&RunConfig::check)},
        {{"wl.force_accept_after"}, "--force-accept-after",
         attempts_value(&RunConfig::force_accept)},
        {{"wl.force_accept_after_mcs"}, "--force-accept-after-mcs",
         mcs_value(&RunConfig::force_accept)},
        {{"initialization.max_attempts", "wl.initialization_max_attempts"},
         "--initialization-max-attempts", u64_value(FIELD(c.wl.initialization_max_attempts))},
        {{"initialization.target_fraction", "wl.initialization_target_fraction"},
         "--initialization-target-fraction",
         double_value(FIELD(c.wl.initialization_target_fraction))},
        {{"initialization.temperature_fraction", "wl.initialization_temperature_fraction"},
         "--initialization-temperature-fraction",
         double_value(FIELD(c.wl.initialization_temperature_fraction))},
        {{"initialization.stall_attempts_per_spin",
          "wl.initialization_stall_attempts_per_spin"},
         "--initialization-stall-attempts-per-spin",
         u64_value(FIELD(c.wl.initialization_stall_attempts_per_spin))},
        {{"initialization.temperature_multiplier", "wl.initialization_temperature_multiplier"},
         "--initialization-temperature-multiplier",
         double_value(FIELD(c.wl.initialization_temperature_multiplier))},
        {{"initialization.max_temperature_fraction",
          "wl.initialization_max_temperature_fraction"},
         "--initialization-max-temperature-fraction",
         double_value(FIELD(c.wl.initialization_max_temperature_fraction))},
        {{"run.seed"}, "--seed",
         [](Context &x, const std::string &v, const std::string &n) {
             x.c.seed = parse_integer<std::uint64_t>(v, n);
             x.c.seed_explicit = true;
         }},
        {{"run.max_attempts"}, "--max-attempts", attempts_value(&RunConfig::max_limit)},
        {{"run.max_mcs"}, "--max-mcs", mcs_value(&RunConfig::max_limit)},
        {{"run.checkpoint_interval"}, "--checkpoint-interval",
         attempts_value(&RunConfig::checkpoint)},
        {{"run.checkpoint_interval_mcs"}, "--checkpoint-interval-mcs",
         mcs_value(&RunConfig::checkpoint)},
        {{"run.checkpoint"}, "--checkpoint",
         [](Context &x, const std::string &v, const std::string &) { x.c.checkpoint_path = v; }},
        {{"run.temperatures"}, "--temperatures",
         [](Context &x, const std::string &v, const std::string &n) {
             x.c.temperatures = parse_list(v, n);
         }},
        {{"run.pilot"}, "", bool_value(FIELD(c.pilot))},
        {{"run.progress"}, "--progress", double_value(FIELD(c.progress_interval_seconds))},
        {{"output.prefix"}, "--output",
         [](Context &x, const std::string &v, const std::string &) { x.c.output_prefix = v; }},
    };
    return table;
}

#undef FIELD

const Setting *find_ini(const std::string &key)
{
    static const auto index = [] {
        std::map<std::string, const Setting *> m;
        for (const auto &s : settings())
            for (const auto &k : s.ini) m[k] = &s;
        return m;
    }();
    const auto it = index.find(key);
    return it == index.end() ? nullptr : it->second;
}

const Setting *find_flag(const std::string &flag)
{
    static const auto index = [] {
        std::map<std::string, const Setting *> m;
        for (const auto &s : settings())
            if (!s.flag.empty()) m[s.flag] = &s;
        return m;
    }();
    const auto it = index.find(flag);
    return it == index.end() ? nullptr : it->second;
}

void load_ini(Context &x, const std::string &path)
{ // io.cpp:167-187
    std::ifstream input(path);
    if (!input) throw std::runtime_error("Cannot open INI file: " + path);
    x.c.config_file = path;
    x.ini_directory = std::filesystem::absolute(path).parent_path();
    std::string line, section;
    std::size_t number = 0;
    while (std::getline(input, line)) {
        ++number;
        line = strip(line, true);
        if (line.empty() || line[0] == '#' || line[0] == ';') continue;
        if (line.front() == '[' && line.back() == ']') {
            section = lower(strip(line.substr(1, line.size() - 2), true));
            continue;
        }
        const auto equals = line.find('=');
        if (equals == std::string::npos)
            throw std::runtime_error("Invalid INI line " + std::to_string(number));
        try {
            const std::string key = lower(strip(line.substr(0, equals), true));
            const std::string full = section.empty() ? key : section + "." + key;
            const Setting *s = find_ini(full);
            if (!s) throw std::invalid_argument("Unknown INI setting: " + full);
            s->set(x, strip(line.substr(equals + 1), true), full);
        } catch (const std::exception &error) {
            throw std::runtime_error("INI line " + std::to_string(number) + ": " + error.what());
        }
    }
    x.ini_directory.clear();
}

std::uint64_t attempts_from_mcs(double mcs, std::size_t spins, bool allow_zero,
                                const std::string &name)
{ // io.cpp:75-87
    if (spins == 0) throw std::invalid_argument("Cannot resolve MCS for an empty system");
    if (!std::isfinite(mcs) || mcs < 0.0 || (!allow_zero && mcs == 0.0))
        throw std::invalid_argument("Invalid MCS value for " + name);
    if (mcs == 0.0) return 0;
    const long double raw = static_cast<long double>(mcs) * static_cast<long double>(spins);
    const long double maximum =
        static_cast<long double>(std::numeric_limits<std::uint64_t>::max());
    if (raw > maximum - 0.5L) throw std::overflow_error("MCS value is too large for " + name);
    return std::max<std::uint64_t>(1, static_cast<std::uint64_t>(std::floor(raw + 0.5L)));
}

} // namespace

void RunConfig::resolve_mcs(std::size_t spins)
{
    if (spins == 0) throw std::invalid_argument("MCS requires at least one spin");
    if (resolved_spin_count != 0 && resolved_spin_count != spins)
        throw std::logic_error("MCS intervals were resolved for a different spin count");
    struct Item {
        Interval *interval;
        bool allow_zero;
        const char *name;
    };
    const Item items[] = {{&exchange, false, "exchange_interval_mcs"},
                          {&check, false, "check_interval_mcs"},
                          {&force_accept, true, "force_accept_after_mcs"},
                          {&max_limit, true, "max_mcs"},
                          {&checkpoint, false, "checkpoint_interval_mcs"}};
    for (const auto &item : items) {
        Interval &v = *item.interval;
        if (v.uses_mcs)
            v.attempts = attempts_from_mcs(v.mcs, spins, item.allow_zero, item.name);
        else
            v.mcs = static_cast<double>(v.attempts) / static_cast<double>(spins);
    }
    wl.check_interval_attempts = check.attempts;
    wl.force_accept_after_attempts = force_accept.attempts;
    resolved_spin_count = spins;
}

bool RunConfig::uses_legacy_attempt_units() const
{
    return !exchange.uses_mcs || !check.uses_mcs || !force_accept.uses_mcs ||
           !max_limit.uses_mcs || !checkpoint.uses_mcs;
}

void RunConfig::validate(bool require_grid) const
{ // io.cpp:243-283
    if (geometry_file.empty() && (nx == 0 || ny == 0 || nz == 0))
        throw std::invalid_argument("Lattice dimensions must be positive");
    if (!(spacing > 0.0) || !std::isfinite(progress_interval_seconds) ||
        progress_interval_seconds < 0.0 || windows == 0 || walkers_per_rank == 0 ||
        overlap < 0.0 || overlap >= 1.0 || exchange.attempts == 0 ||
        wl.check_interval_attempts == 0 || checkpoint.attempts == 0 ||
        !std::isfinite(wl.round_trip_margin_fraction) || wl.round_trip_margin_fraction < 0.0 ||
        wl.round_trip_margin_fraction >= 0.5)
        throw std::invalid_argument("Invalid run configuration");
    if (require_grid && !energy_grid_explicit)
        throw std::invalid_argument("Production runs require --emin, --emax, and --bin-width");
    grid.validate();
    if (windows > grid.bins()) throw std::invalid_argument("More windows than energy bins");
    if (!explicit_windows.empty()) {
        if (explicit_windows.size() != windows || explicit_windows.front().begin != 0 ||
            explicit_windows.back().end != grid.bins())
            throw std::invalid_argument("Explicit energy windows must match the configured range");
        for (std::size_t i = 0; i < explicit_windows.size(); ++i) {
            const auto &w = explicit_windows[i];
            if (w.begin >= w.end || w.end > grid.bins() ||
                (i != 0 && (w.begin <= explicit_windows[i - 1].begin ||
                            w.begin >= explicit_windows[i - 1].end)))
                throw std::invalid_argument("Explicit energy windows must be ordered and overlap");
        }
    }
    const auto &a = adaptive_windows;
    auto finite = [](double v) { return std::isfinite(v); };
    if (a.enabled &&
        (a.iterations == 0 || !(a.pilot_mcs > 0.0) || !finite(a.pilot_mcs) ||
         a.smoothing_width < 0.0 || !finite(a.smoothing_width) || a.minimum_width < 0.0 ||
         !finite(a.minimum_width) || !(a.diffusivity_floor_fraction > 0.0) ||
         !finite(a.diffusivity_floor_fraction) || a.curvature_weight < 0.0 ||
         !finite(a.curvature_weight) || a.round_trip_target < 0.0 ||
         !finite(a.round_trip_target) || a.maximum_round_trip_penalty < 1.0 ||
         !finite(a.maximum_round_trip_penalty) ||
         (a.minimum_width > 0.0 &&
          a.minimum_width * static_cast<double>(windows) > grid.maximum - grid.minimum)))
        throw std::invalid_argument("Invalid adaptive-window configuration");
}

RunConfig parse_arguments(int argc, char **argv)
{ // io.cpp:285-372
    RunConfig c;
    bool have_min = false, have_max = false, have_width = false;
    Context x{c, have_min, have_max, have_width, {}};
    for (int i = 1; i < argc; ++i)
        if (std::string(argv[i]) == "--config") {
            if (++i >= argc) throw std::invalid_argument("Missing value for --config");
            load_ini(x, argv[i]);
        }
    for (int i = 1; i < argc; ++i) {
        const std::string key(argv[i]);
        if (key == "--config") {
            ++i;
            continue;
        }
        if (key == "--smoke-test") {
            c.smoke_test = true;
            continue;
        }
        if (key == "--pilot") {
            c.pilot = true;
            continue;
        }
        if (key == "--help" || key == "-h") continue;
        const Setting *s = find_flag(key);
        if (!s) throw std::invalid_argument("Unknown option: " + key);
        if (++i >= argc) throw std::invalid_argument("Missing value for " + key);
        s->set(x, argv[i], key);
    }
    c.energy_grid_explicit = have_min && have_max && have_width;
    if (c.smoke_test) { // io.cpp:365-370
        c.nx = 2;
        c.ny = 2;
        c.nz = 1;
        c.grid = {-8.0, 8.0, 0.25};
        c.energy_grid_explicit = true;
        c.max_limit.mcs = 2'500.0;
        c.max_limit.uses_mcs = true;
        c.wl.flatness = 0.1;
        c.wl.minimum_visits = 1;
        c.wl.final_factor = 0.5;
        c.windows = 1;
        c.walkers_per_rank = 1;
        c.output_prefix = "wl_smoke";
    }
    return c;
}

std::string usage(const std::string &program)
{ // the reference's option summary (io.cpp:374-396)
    return "Usage: " + program +
           " --emin E --emax E --bin-width dE [--nx N --ny N --nz N]\n"
           "  [--geometry file --box-x L --box-y L --box-z L] [--cutoff R]\n"
           "  [--config run.ini] (CLI options override INI values)\n"
           "  [--windows N --walkers N --overlap 0.75] [--seed N]\n"
           "  [--adaptive-windows true|false --adaptive-iterations N --adaptive-pilot-mcs MCS]\n"
           "  [--adaptive-smoothing-width dE --adaptive-minimum-width dE]\n"
           "  [--adaptive-diffusivity-floor F --adaptive-curvature-weight W]\n"
           "  [--adaptive-round-trip-target R --adaptive-round-trip-penalty P]\n"
           "  [--flatness 0.8 --min-visits 100 --final-factor 1e-8 --inverse-time true|false]\n"
           "  [--initialization-max-attempts N --initialization-target-fraction 0.5]\n"
           "  [--initialization-temperature-fraction 0.05]\n"
           "  [--initialization-stall-attempts-per-spin 1000]\n"
           "  [--initialization-temperature-multiplier 2 --initialization-max-temperature-fraction 0.5]\n"
           "  [--check-interval-mcs MCS --force-accept-after-mcs MCS]\n"
           "  [--exchange-interval-mcs MCS --max-mcs MCS --output prefix]\n"
           "  [--progress seconds] updates one console line; 0 disables progress.\n"
           "  MCS may be fractional; one MCS is N attempted single-spin flips per walker.\n"
           "  If --seed is omitted, a random seed is printed and saved in metadata.\n"
           "  Set --max-mcs 0 to run until convergence (not valid with --pilot).\n"
           "  Add --pilot to sample energies and print a non-rigorous grid suggestion.\n";
}

} // namespace rewl::host
