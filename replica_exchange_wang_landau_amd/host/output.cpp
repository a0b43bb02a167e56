#include "output.hpp"

#include <array>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string_view>

namespace rewl::host {
namespace {

// numbers as the reference's streams print them: doubles with 17 significant digits ("%.17g")
std::string num(double v)
{
    char buffer[40];
    std::snprintf(buffer, sizeof buffer, "%.17g", v);
    return buffer;
}
std::string num(std::uint64_t v) { return std::to_string(v); }
std::string num(int v) { return std::to_string(v); }
const char *flag(bool v) { return v ? "true" : "false"; }

std::string json_escape(std::string_view value)
{
    std::string out;
    for (char ch : value) {
        switch (ch) {
        case '\\': out += "\\\\"; break;
        case '"': out += "\\\""; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default: out += ch;
        }
    }
    return out;
}

std::ofstream open_text(const std::string &path)
{
    std::ofstream out(path);
    if (!out) throw std::runtime_error("Cannot write " + path);
    return out;
}

// one "key": value member of a flat JSON object written on one line
class Members {
  public:
    template <class T> Members &add(const char *key, const T &value)
    {
        text_ += (text_.empty() ? "" : ", ");
        text_ += std::string("\"") + key + "\": " + value;
        return *this;
    }
    Members &str(const char *key, const std::string &value)
    {
        return add(key, "\"" + json_escape(value) + "\"");
    }
    std::string object() const { return "{" + text_ + "}"; }

  private:
    std::string text_;
};

} // namespace

void write_dos_csv(const std::string &path, const DensityOfStates &dos)
{ // io.cpp:398-408
    const std::size_t n = dos.log_g.size();
    if (dos.histogram.size() != n || dos.standard_error.size() != n || dos.valid.size() != n)
        throw std::invalid_argument("DOS output arrays have different sizes");
    auto out = open_text(path);
    out << "bin,energy,log_g,histogram,standard_error,valid\n";
    for (std::size_t i = 0; i < n; ++i)
        out << i << ',' << num(dos.grid.center(i)) << ',' << num(dos.log_g[i]) << ','
            << dos.histogram[i] << ',' << num(dos.standard_error[i]) << ','
            << static_cast<unsigned>(dos.valid[i]) << '\n';
}

void write_fragment_csv(const std::string &path, const Grid &grid, const DosFragment &f,
                        std::size_t window_id)
{ // io.cpp:410-421
    const std::size_t n = f.log_g.size();
    if (f.histogram.size() != n || f.standard_error.size() != n || f.valid.size() != n ||
        f.window.end > n)
        throw std::invalid_argument("DOS fragment output arrays have different sizes");
    auto out = open_text(path);
    out << "window,bin,energy,log_g,histogram,standard_error,valid\n";
    for (std::size_t i = f.window.begin; i < f.window.end; ++i)
        out << window_id << ',' << i << ',' << num(grid.center(i)) << ',' << num(f.log_g[i]) << ','
            << f.histogram[i] << ',' << num(f.standard_error[i]) << ','
            << static_cast<unsigned>(f.valid[i]) << '\n';
}

void write_thermodynamics_csv(const std::string &path,
                              const std::vector<ThermodynamicPoint> &points)
{ // io.cpp:423-428
    auto out = open_text(path);
    out << "temperature,log_Z,U,C,F,S\n";
    for (const auto &p : points)
        out << num(p.temperature) << ',' << num(p.log_partition) << ',' << num(p.internal_energy)
            << ',' << num(p.heat_capacity) << ',' << num(p.free_energy) << ',' << num(p.entropy)
            << '\n';
}

void write_workers_stat_csv(const std::string &path, const std::vector<WalkerStatistics> &stats,
                            std::size_t spins)
{ // io.cpp:430-452
    if (spins == 0) throw std::invalid_argument("Cannot write MCS statistics for zero spins");
    auto out = open_text(path);
    out << "mpi_rank,window,walker_id,attempted_flips,attempted_mcs,accepted,forced_accepted,"
           "accepted_percent,attempts_since_last_accepted,mcs_since_last_accepted,energy,factor,"
           "active_bins,min_h,mean_h,min_over_mean,round_trips\n";
    const double n = static_cast<double>(spins);
    for (const auto &w : stats) {
        const double percent = w.attempted == 0 ? 0.0
                                                : 100.0 * static_cast<double>(w.accepted) /
                                                      static_cast<double>(w.attempted);
        out << w.mpi_rank << ',' << w.window_id << ',' << w.walker_id << ',' << w.attempted << ','
            << num(static_cast<double>(w.attempted) / n) << ',' << w.accepted << ','
            << w.forced_accepted << ',' << num(percent) << ',' << w.attempts_since_last_accepted
            << ',' << num(static_cast<double>(w.attempts_since_last_accepted) / n) << ','
            << num(w.energy) << ',' << num(w.factor) << ',' << w.active_bins << ','
            << w.minimum_histogram << ',' << num(w.mean_histogram) << ',' << num(w.min_over_mean)
            << ',' << w.round_trips << '\n';
    }
}

void write_metadata_json(const std::string &path, const RunConfig &c, std::size_t spins,
                         const std::string &backend, std::uint64_t attempted,
                         std::uint64_t accepted, std::uint64_t forced, std::uint64_t x_attempted,
                         std::uint64_t x_accepted, bool converged, int mpi_size, int threads,
                         const EngineInfo *engine)
{ // format_version 4 (io.cpp:454-537); the engine object is this build's addition
    const auto &wl = c.wl;
    const auto &a = c.adaptive_windows;
    const char *criterion =
        wl.inverse_time_enabled
            ? (wl.nalivaiko_mod ? "full_coverage_of_current_iteration_bins"
                                : "full_coverage_of_discovered_bins")
            : (wl.nalivaiko_mod ? "histogram_flatness_over_current_iteration_bins"
                                : "histogram_flatness");
    const auto windows = c.explicit_windows.empty()
                             ? partition_windows(c.grid.bins(), c.windows, c.overlap)
                             : c.explicit_windows;
    std::string ranges;
    for (std::size_t i = 0; i < windows.size(); ++i)
        ranges += (i ? ", [" : "[") +
                  num(c.grid.minimum + static_cast<double>(windows[i].begin) * c.grid.width) +
                  ", " +
                  num(c.grid.minimum + static_cast<double>(windows[i].end) * c.grid.width) + "]";
    Members physics;
    physics.str("model", "dipolar_ising")
        .add("spins", num(std::uint64_t(spins)))
        .add("coupling_scale", num(c.coupling_scale))
        .add("periodic", flag(c.periodic))
        .add("box", "[" + num(c.box_lengths.x) + ", " + num(c.box_lengths.y) + ", " +
                        num(c.box_lengths.z) + "]")
        .str("backend", backend)
        .add("cutoff", num(c.cutoff));
    Members grid;
    grid.add("minimum", num(c.grid.minimum))
        .add("maximum", num(c.grid.maximum))
        .add("bin_width", num(c.grid.width))
        .add("complete_range", flag(c.complete_range));
    Members algorithm;
    algorithm.str("variant", wl.inverse_time_enabled ? "REWL-1/t" : "REWL")
        .add("windows", num(std::uint64_t(c.windows)))
        .add("walkers_per_rank", num(std::uint64_t(c.walkers_per_rank)))
        .add("overlap", num(c.overlap))
        .add("flatness", num(wl.flatness))
        .str("flatness_scope", "walker_local")
        .str("initial_refinement_criterion", criterion)
        .add("nalivaiko_mod", flag(wl.nalivaiko_mod))
        .str("refinement_active_scope",
             wl.nalivaiko_mod ? "current_iteration" : "cumulative_discovered_bins")
        .str("inverse_time_clock", "walker_local_attempted_flips/active_bins")
        .str("dos_synchronization", "none")
        .str("refinement_schedule", "walker_independent")
        .str("final_dos_combination", "aligned_log_mean")
        .str("standard_error_kind", "within_run_walker_sem")
        .add("minimum_visits_per_walker", num(wl.minimum_visits))
        .add("check_interval_mcs", num(c.check.mcs))
        .add("check_interval_attempts", num(wl.check_interval_attempts))
        .add("final_factor", num(wl.final_factor))
        .add("force_accept_after_mcs", num(c.force_accept.mcs))
        .add("inverse_time_enabled", flag(wl.inverse_time_enabled))
        .add("force_accept_after_attempts", num(wl.force_accept_after_attempts));
    Members adaptive;
    adaptive.add("enabled", flag(a.enabled))
        .add("iterations", num(std::uint64_t(a.iterations)))
        .add("pilot_mcs", num(a.pilot_mcs))
        .add("smoothing_width", num(a.smoothing_width))
        .add("minimum_width", num(a.minimum_width))
        .add("diffusivity_floor_fraction", num(a.diffusivity_floor_fraction))
        .add("curvature_weight", num(a.curvature_weight))
        .add("round_trip_target", num(a.round_trip_target))
        .add("maximum_round_trip_penalty", num(a.maximum_round_trip_penalty))
        .add("round_trip_margin_fraction", num(wl.round_trip_margin_fraction))
        .add("external_warm_start_target_fraction", "1")
        .add("energy_ranges", "[" + ranges + "]");
    Members init;
    init.str("method", "adaptive_target_metropolis")
        .add("max_attempts", num(wl.initialization_max_attempts))
        .add("target_fraction", num(wl.initialization_target_fraction))
        .add("temperature_fraction", num(wl.initialization_temperature_fraction))
        .add("stall_attempts_per_spin", num(wl.initialization_stall_attempts_per_spin))
        .add("temperature_multiplier", num(wl.initialization_temperature_multiplier))
        .add("max_temperature_fraction", num(wl.initialization_max_temperature_fraction));
    Members parallel;
    parallel.add("mpi_size", num(mpi_size))
        .add("openmp_threads", num(threads))
        .add("exchange_interval_mcs", num(c.exchange.mcs))
        .add("exchange_interval_attempts", num(c.exchange.attempts));
    Members limits;
    limits.add("max_mcs_per_walker", num(c.max_limit.mcs))
        .add("max_attempts_per_walker", num(c.max_limit.attempts))
        .add("checkpoint_interval_mcs", num(c.checkpoint.mcs))
        .add("checkpoint_interval_attempts", num(c.checkpoint.attempts))
        .add("progress_interval_seconds", num(c.progress_interval_seconds));

    auto out = open_text(path);
    out << "{\n  \"format_version\": 4,\n  \"config_file\": \"" << json_escape(c.config_file)
        << "\",\n  \"geometry_file\": \"" << json_escape(c.geometry_file) << "\",\n"
        << "  \"physics\": " << physics.object() << ",\n"
        << "  \"energy_grid\": " << grid.object() << ",\n"
        << "  \"algorithm\": " << algorithm.object() << ",\n"
        << "  \"adaptive_windows\": " << adaptive.object() << ",\n"
        << "  \"initialization\": " << init.object() << ",\n"
        << "  \"parallel\": " << parallel.object() << ",\n"
        << "  \"run_limits\": " << limits.object() << ",\n"
        << "  \"seed\": " << c.seed << ",\n  \"seed_source\": \""
        << (c.seed_explicit ? "configured" : "random") << "\",\n  \"attempted\": " << attempted
        << ",\n  \"attempted_flips\": " << attempted
        << ",\n  \"aggregate_mcs\": "
        << num(static_cast<double>(attempted) / static_cast<double>(spins))
        << ",\n  \"accepted\": " << accepted << ",\n  \"forced_accepted\": " << forced
        << ",\n  \"exchange_attempted\": " << x_attempted
        << ",\n  \"exchange_accepted\": " << x_accepted
        << ",\n  \"converged\": " << flag(converged);
    if (engine) {
        Members e;
        e.str("implementation", "mi355x-hip")
            .str("device", engine->device)
            .str("rng", engine->rng)
            .str("fields", engine->fields)
            .add("walkers", num(engine->walkers))
            .add("epochs", num(engine->epochs))
            .add("ambiguous_log_decisions", num(engine->ambiguous_decisions));
        out << ",\n  \"engine\": " << e.object();
    }
    out << "\n}\n";
}

// ---- WLCHKP3 --------------------------------------------------------------------------------
namespace {

template <class T> void put(std::ostream &out, const T &v)
{
    out.write(reinterpret_cast<const char *>(&v), sizeof v);
}
template <class T> void put_vector(std::ostream &out, const std::vector<T> &v)
{
    put(out, static_cast<std::uint64_t>(v.size()));
    out.write(reinterpret_cast<const char *>(v.data()),
              static_cast<std::streamsize>(v.size() * sizeof(T)));
}
template <class T> void get(std::istream &in, T &v)
{
    in.read(reinterpret_cast<char *>(&v), sizeof v);
    if (!in) throw std::runtime_error("Truncated checkpoint");
}
template <class T> void get_vector(std::istream &in, std::vector<T> &v)
{
    std::uint64_t size{};
    get(in, size);
    if (size > 1'000'000'000ULL) throw std::runtime_error("Unreasonable checkpoint vector size");
    v.resize(static_cast<std::size_t>(size));
    in.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(size * sizeof(T)));
    if (!in) throw std::runtime_error("Truncated checkpoint vector");
}

// one WLCHKP3 record (io.cpp:543-551)
void put_record(std::ostream &out, const WalkerState &w)
{
    const std::array<char, 8> magic{'W', 'L', 'C', 'H', 'K', 'P', '3', '\0'};
    out.write(magic.data(), magic.size());
    put(out, w.id);
    put_vector(out, w.spins);
    put_vector(out, w.fields);
    put(out, w.energy);
    put_vector(out, w.log_g);
    put_vector(out, w.histogram);
    put_vector(out, w.active);
    put(out, w.factor);
    put(out, w.attempted);
    put(out, w.accepted);
    put(out, w.forced_accepted);
    put(out, w.last_accepted_attempt);
    put(out, static_cast<std::uint8_t>(w.stage));
    for (std::uint64_t word : w.rng_state) put(out, word);
}

// one WLCHKP1/2/3 record (io.cpp:558-572)
WalkerState get_record(std::istream &in)
{
    std::array<char, 8> magic{};
    in.read(magic.data(), magic.size());
    const std::string_view version(magic.data(), 7);
    if (!in || (version != "WLCHKP1" && version != "WLCHKP2" && version != "WLCHKP3"))
        throw std::runtime_error("Invalid checkpoint format");
    WalkerState w;
    get(in, w.id);
    get_vector(in, w.spins);
    get_vector(in, w.fields);
    get(in, w.energy);
    get_vector(in, w.log_g);
    get_vector(in, w.histogram);
    get_vector(in, w.active);
    get(in, w.factor);
    get(in, w.attempted);
    get(in, w.accepted);
    if (version == "WLCHKP3") get(in, w.forced_accepted);
    if (version == "WLCHKP2" || version == "WLCHKP3")
        get(in, w.last_accepted_attempt);
    else
        w.last_accepted_attempt = w.accepted == 0 ? 0 : w.attempted;
    std::uint8_t stage{};
    get(in, stage);
    w.stage = stage;
    for (auto &word : w.rng_state) get(in, word);
    return w;
}

// writes to path.tmp, then renames it over path (io.cpp:552-555)
template <class Write> void write_atomically(const std::string &path, Write &&write)
{
    const std::string temporary = path + ".tmp";
    {
        std::ofstream out(temporary, std::ios::binary | std::ios::trunc);
        if (!out) throw std::runtime_error("Cannot write checkpoint " + temporary);
        write(out);
        out.close();
        if (!out) throw std::runtime_error("Failed writing checkpoint");
    }
    std::error_code error;
    std::filesystem::remove(path, error);
    error.clear();
    std::filesystem::rename(temporary, path, error);
    if (error) throw std::runtime_error("Cannot atomically replace checkpoint: " + error.message());
}

} // namespace

void save_checkpoint(const std::string &path, const WalkerState &w)
{ // io.cpp:539-556
    write_atomically(path, [&](std::ostream &out) { put_record(out, w); });
}

WalkerState load_checkpoint(const std::string &path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open checkpoint " + path);
    return get_record(in);
}

void save_run_checkpoint(const std::string &path, const RunCheckpoint &run)
{
    write_atomically(path, [&](std::ostream &out) {
        const std::array<char, 8> magic{'W', 'L', 'C', 'H', 'K', 'M', '2', '\0'};
        out.write(magic.data(), magic.size());
        put(out, static_cast<std::uint64_t>(run.walkers.size()));
        put(out, run.epoch);
        put(out, run.exchange_attempted);
        put(out, run.exchange_accepted);
        for (const WalkerState &w : run.walkers) {
            put(out, w.last_flatness_check);
            put(out, w.last_flatness);
            put(out, w.clock_offset);
        }
        for (const WalkerState &w : run.walkers) put_record(out, w);
    });
}

RunCheckpoint load_run_checkpoint(const std::string &path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open checkpoint " + path);
    std::array<char, 8> magic{};
    in.read(magic.data(), magic.size());
    const std::string_view tag(magic.data(), 7);
    if (!in || (tag != "WLCHKM1" && tag != "WLCHKM2"))
        throw std::runtime_error("Invalid multi-walker checkpoint format");
    const bool offsets = tag == "WLCHKM2"; // WLCHKM1 files predate warm starts: offset 0
    RunCheckpoint run;
    std::uint64_t count{};
    get(in, count);
    if (count > (std::uint64_t(1) << 24)) throw std::runtime_error("Unreasonable walker count");
    get(in, run.epoch);
    get(in, run.exchange_attempted);
    get(in, run.exchange_accepted);
    std::vector<std::uint64_t> checks(count);
    std::vector<double> flatness(count);
    std::vector<std::uint64_t> clock(count, 0);
    for (std::uint64_t k = 0; k < count; ++k) {
        get(in, checks[k]);
        get(in, flatness[k]);
        if (offsets) get(in, clock[k]);
    }
    for (std::uint64_t k = 0; k < count; ++k) {
        run.walkers.push_back(get_record(in));
        run.walkers.back().last_flatness_check = checks[k];
        run.walkers.back().last_flatness = flatness[k];
        run.walkers.back().clock_offset = clock[k];
    }
    return run;
}

} // namespace rewl::host
