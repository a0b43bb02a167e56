// CPU self-test of the host layer (no GPU): configuration parsing and MCS resolution,
// windows, WLCHKP3 layout, summaries, stitching, thermodynamics and output formats. The
// expectations are those of the reference's tests (tests/test_main.cpp, cited per case).
//
// `host_selftest adapt FILE` instead runs the adaptive-window planner on the pilot data in
// FILE and prints its result, for comparison with the reference's planner.
#include "adaptive.hpp"
#include "geometry.hpp"
#include "output.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <iterator>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <vector>

using namespace rewl::host;
namespace fs = std::filesystem;

namespace {

void require(bool ok, const std::string &what)
{
    if (!ok) throw std::runtime_error(what);
}

void near(double a, double b, double tol, const std::string &what)
{
    if (std::abs(a - b) > tol * std::max({1.0, std::abs(a), std::abs(b)}))
        throw std::runtime_error(what);
}

RunConfig parse(std::vector<std::string> args)
{
    args.insert(args.begin(), "wl_run");
    std::vector<char *> argv;
    for (auto &a : args) argv.push_back(a.data());
    return parse_arguments(static_cast<int>(argv.size()), argv.data());
}

std::string slurp(const fs::path &p)
{
    std::ifstream in(p);
    return {std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>()};
}

void csv_geometry_and_ini()
{ // test_main.cpp:40-137
    const auto dir = fs::temp_directory_path() / "rewl_host_selftest";
    fs::create_directories(dir);
    const auto csv = dir / "system.csv", ini = dir / "run.ini";
    std::ofstream(csv) << "x,y,z,mx,my,mz\n0,0,0,0,0,1\n1,0,0,0,0,1\n";
    std::ofstream(ini) << "[geometry]\nfile=system.csv\nperiodic=false\n"
                       << "[energy]\nminimum=-4\nmaximum=4\nbin_width=0.25\n"
                       << "[parallel]\nexchange_interval_mcs=1.5\n"
                       << "[adaptive_windows]\nenabled=true\niterations=3\npilot_mcs=12\n"
                       << "smoothing_width=0.75\nminimum_width=1.25\n"
                       << "diffusivity_floor_fraction=0.08\ncurvature_weight=0.4\n"
                       << "round_trip_target=3\nmaximum_round_trip_penalty=2.5\n"
                       << "round_trip_margin_fraction=0.15\n"
                       << "[wl]\ncheck_interval_mcs=2.5\nforce_accept_after_mcs=3.5\n"
                       << "inverse_time=false\nnalivaiko_mod=true\n"
                       << "[initialization]\nmax_attempts=123\ntarget_fraction=0.4\n"
                       << "temperature_fraction=0.07\nstall_attempts_per_spin=17\n"
                       << "temperature_multiplier=3\nmax_temperature_fraction=0.4\n"
                       << "[run]\nseed=17\nmax_mcs=4.5\ncheckpoint_interval_mcs=5.5\nprogress=2\n"
                       << "temperatures=0.5, 1.5\n[output]\nprefix=from_ini\n";
    const auto geometry = Geometry::load_csv(csv.string(), {2, 2, 2}, false);
    require(geometry.size() == 2, "CSV geometry size");
    near(geometry.axes[1].z, 1, 1e-14, "CSV moment axis");
    auto c = parse({"--config", ini.string(), "--seed", "99"});
    c.resolve_mcs(geometry.size());
    require(c.geometry_file == csv.lexically_normal().string(), "INI-relative CSV path");
    require(c.seed == 99 && c.seed_explicit && c.energy_grid_explicit,
            "CLI override and INI energy grid");
    require(!parse({}).seed_explicit, "omitted seed detection");
    require(c.exchange.attempts == 3 && c.wl.check_interval_attempts == 5,
            "MCS exchange and check interval resolution");
    require(c.wl.force_accept_after_attempts == 7 && c.max_limit.attempts == 9 &&
                c.checkpoint.attempts == 11,
            "MCS run interval resolution");
    near(c.progress_interval_seconds, 2.0, 1e-14, "INI progress interval");
    require(!c.wl.inverse_time_enabled && c.wl.nalivaiko_mod, "INI WL switches");
    require(c.wl.initialization_max_attempts == 123 &&
                c.wl.initialization_stall_attempts_per_spin == 17,
            "INI initialization integers");
    near(c.wl.initialization_target_fraction, 0.4, 1e-14, "INI initialization target");
    near(c.wl.initialization_temperature_fraction, 0.07, 1e-14, "INI initialization temperature");
    near(c.wl.initialization_temperature_multiplier, 3.0, 1e-14, "INI temperature multiplier");
    near(c.wl.initialization_max_temperature_fraction, 0.4, 1e-14, "INI maximum temperature");
    require(c.adaptive_windows.enabled && c.adaptive_windows.iterations == 3, "INI adaptive");
    near(c.adaptive_windows.pilot_mcs, 12.0, 1e-14, "INI adaptive pilot length");
    near(c.adaptive_windows.maximum_round_trip_penalty, 2.5, 1e-14, "INI adaptive penalty");
    near(c.wl.round_trip_margin_fraction, 0.15, 1e-14, "INI adaptive round-trip margin");
    const auto progress = parse({"--progress", "3", "--inverse-time", "true"});
    near(progress.progress_interval_seconds, 3.0, 1e-14, "CLI progress interval");
    require(progress.wl.inverse_time_enabled && !progress.wl.nalivaiko_mod, "CLI switches");
    require(parse({"--progress", "0"}).progress_interval_seconds == 0.0, "zero progress");
    require(!c.uses_legacy_attempt_units(), "MCS options must not be marked legacy");
    auto legacy = parse({"--max-attempts", "9"});
    legacy.resolve_mcs(2);
    require(legacy.max_limit.attempts == 9 && legacy.max_limit.mcs == 4.5 &&
                legacy.uses_legacy_attempt_units(),
            "legacy attempts compatibility");
    require(c.temperatures.size() == 2 && c.output_prefix == "from_ini", "INI lists and output");
    // rejected inputs and their messages
    const auto bad = dir / "invalid.ini";
    std::ofstream(bad) << "[wl]\nnalivaiko_mod=maybe\n";
    bool rejected = false;
    try {
        parse({"--config", bad.string()});
    } catch (const std::runtime_error &e) {
        rejected = std::string(e.what()) == "INI line 2: Invalid boolean for wl.nalivaiko_mod";
    }
    require(rejected, "invalid Nalivaiko boolean must be rejected");
    std::ofstream(bad) << "[energy]\nbogus=1\n";
    rejected = false;
    try {
        parse({"--config", bad.string()});
    } catch (const std::runtime_error &e) {
        rejected = std::string(e.what()) == "INI line 2: Unknown INI setting: energy.bogus";
    }
    require(rejected, "unknown INI key");
    rejected = false;
    try {
        parse({"--bogus", "1"});
    } catch (const std::invalid_argument &e) {
        rejected = std::string(e.what()) == "Unknown option: --bogus";
    }
    require(rejected, "unknown CLI option");
    rejected = false;
    try {
        parse({"--nx"});
    } catch (const std::invalid_argument &e) {
        rejected = std::string(e.what()) == "Missing value for --nx";
    }
    require(rejected, "missing CLI value");
    rejected = false;
    try {
        parse({"--nx", "3x"});
    } catch (const std::invalid_argument &e) {
        rejected = std::string(e.what()) == "Invalid value for --nx";
    }
    require(rejected, "trailing characters in an integer");
    // metadata records the Nalivaiko mode
    const auto meta = dir / "metadata.json";
    write_metadata_json(meta.string(), c, 2, "dense", 0, 0, 0, 0, 0, false, 1, 1, nullptr);
    const auto text = slurp(meta);
    require(text.find("\"nalivaiko_mod\": true") != std::string::npos &&
                text.find("\"refinement_active_scope\": \"current_iteration\"") !=
                    std::string::npos,
            "Nalivaiko mode must be recorded in metadata");
    // smoke-test preset (io.cpp:365-370)
    auto smoke = parse({"--smoke-test"});
    require(smoke.nx == 2 && smoke.ny == 2 && smoke.nz == 1 && smoke.grid.bins() == 64 &&
                smoke.output_prefix == "wl_smoke" && smoke.wl.final_factor == 0.5,
            "smoke-test preset");
    fs::remove_all(dir);
}

void grid_and_windows()
{ // test_main.cpp:156-162
    const Grid grid{-2, 2, 0.5};
    require(grid.bins() == 8, "grid bin count");
    require(grid.index(-2) == 0 && grid.index(2) == 7 && grid.index(2.1) == -1,
            "grid boundary mapping");
    const auto w = partition_windows(100, 4, 0.75);
    require(w.size() == 4 && w.front().begin == 0 && w.back().end == 100, "window coverage");
    for (std::size_t i = 1; i < w.size(); ++i) require(w[i].begin < w[i - 1].end, "overlap");
    const auto c2 = partition_windows(1255, 8, 0.75);
    require(c2[1].begin == 114 && c2[7].begin == 798 && c2[0].end == 457, "C2 layout");
}

WalkerState synthetic_walker(std::size_t n, std::size_t bins)
{
    WalkerState w;
    w.bins_window = {0, bins};
    w.spins.assign(n, 1);
    w.fields.assign(n, 0.25);
    w.log_g.assign(bins, 0.0);
    w.histogram.assign(bins, 0);
    w.active.assign(bins, 0);
    w.refinement.assign(bins, 0);
    return w;
}

void checkpoint_layout()
{ // io.cpp:539-572, 2929 bytes for an 8-spin, 160-bin walker
    auto w = synthetic_walker(8, 160);
    w.energy = -9.25;
    w.log_g[3] = 1.5;
    w.histogram[3] = 7;
    w.active[3] = 1;
    w.factor = 0.125;
    w.attempted = 100;
    w.accepted = 40;
    w.forced_accepted = 2;
    w.last_accepted_attempt = 99;
    w.stage = 1;
    w.rng_state[0] = 1;
    w.rng_state[3] = 0xdeadbeef;
    const auto path = (fs::temp_directory_path() / "rewl_selftest.chk").string();
    save_checkpoint(path, w);
    require(fs::file_size(path) == 2929, "WLCHKP3 size");
    const auto r = load_checkpoint(path);
    require(r.spins == w.spins && r.fields == w.fields && r.log_g == w.log_g &&
                r.histogram == w.histogram && r.active == w.active && r.energy == w.energy &&
                r.factor == w.factor && r.attempted == 100 && r.accepted == 40 &&
                r.forced_accepted == 2 && r.last_accepted_attempt == 99 && r.stage == 1 &&
                r.rng_state[3] == 0xdeadbeef,
            "WLCHKP3 round trip");
    // a truncated file is rejected
    fs::resize_file(path, 100);
    bool rejected = false;
    try {
        (void)load_checkpoint(path);
    } catch (const std::runtime_error &e) {
        rejected = std::string(e.what()).find("Truncated checkpoint") == 0;
    }
    require(rejected, "truncated checkpoint");
    fs::remove(path);
}

void run_checkpoint_layout()
{ // WLCHKM1: 40-byte header, 16 bytes of cadence per walker, then WLCHKP3 records
    RunCheckpoint run{12, 30, 7, {}};
    for (std::uint64_t k = 0; k < 3; ++k) {
        auto w = synthetic_walker(8, 160);
        w.id = k;
        w.attempted = 100 + k;
        w.last_flatness_check = 50 + k;
        w.last_flatness = 0.25 * double(k);
        w.rng_state[2] = 0x1234 + k;
        run.walkers.push_back(w);
    }
    const auto path = (fs::temp_directory_path() / "rewl_selftest_run.chk").string();
    save_run_checkpoint(path, run);
    require(fs::file_size(path) == 40 + 3 * 16 + 3 * 2929, "WLCHKM1 size");
    const auto r = load_run_checkpoint(path);
    require(r.epoch == 12 && r.exchange_attempted == 30 && r.exchange_accepted == 7 &&
                r.walkers.size() == 3,
            "WLCHKM1 header");
    for (std::uint64_t k = 0; k < 3; ++k)
        require(r.walkers[k].id == k && r.walkers[k].attempted == 100 + k &&
                    r.walkers[k].last_flatness_check == 50 + k &&
                    r.walkers[k].last_flatness == 0.25 * double(k) &&
                    r.walkers[k].rng_state[2] == 0x1234 + k &&
                    r.walkers[k].spins == run.walkers[k].spins,
                "WLCHKM1 walker round trip");
    bool rejected = false; // a single-walker file is not a run checkpoint
    save_checkpoint(path, run.walkers[0]);
    try {
        (void)load_run_checkpoint(path);
    } catch (const std::runtime_error &e) {
        rejected = std::string(e.what()) == "Invalid multi-walker checkpoint format";
    }
    require(rejected, "WLCHKM1 magic");
    fs::remove(path);
}

void summary()
{ // test_main.cpp:730-752
    auto a = synthetic_walker(1, 4), b = synthetic_walker(1, 4);
    a.active = {1, 1, 1, 0};
    b.active = {0, 1, 1, 1};
    a.log_g[1] = 10.0;
    a.log_g[2] = 14.0;
    b.log_g[1] = -3.0;
    b.log_g[2] = 3.0;
    a.histogram[2] = 7;
    b.histogram[2] = 9;
    const auto f = summarize_walkers({0, 4}, {&a, &b}, 4);
    require(f.valid == std::vector<std::uint8_t>{0, 1, 1, 0}, "summary active-mask intersection");
    near(f.log_g[1], 0.0, 1e-14, "summary reference bin");
    near(f.log_g[2], 5.0, 1e-14, "summary aligned mean");
    near(f.standard_error[2], 1.0, 1e-14, "summary analytic SEM");
    require(f.histogram[2] == 16, "summary histogram sum");
    require(std::isnan(f.log_g[0]) && std::isnan(f.standard_error[0]), "invalid bins are NaN");
    require(std::isnan(summarize_walkers({0, 4}, {&a}, 4).standard_error[1]),
            "single-walker SEM is undefined");
    b.active = {0, 0, 0, 1};
    bool rejected = false;
    try {
        (void)summarize_walkers({0, 4}, {&a, &b}, 4);
    } catch (const std::runtime_error &) {
        rejected = true;
    }
    require(rejected, "summary must reject walkers without a common active bin");
}

void statistics_and_stat_row()
{ // test_main.cpp:534-565 and 351-362
    auto w = synthetic_walker(1, 4);
    w.active = {1, 1, 0, 0};
    w.histogram = {80, 120, 0, 0};
    w.active_bins = 2;
    auto s = histogram_statistics(w, false);
    require(s.active_bins == 2 && s.covered_bins == 2 && s.minimum == 80, "statistics");
    near(s.mean, 100.0, 1e-14, "mean");
    near(s.min_over_mean, 0.8, 1e-14, "min over mean");
    w.active[2] = 1;
    w.active_bins = 3;
    s = histogram_statistics(w, false);
    require(s.minimum == 0 && s.covered_bins == 2, "active zero bin");
    near(s.mean, 200.0 / 3.0, 1e-14, "mean with active zero bin");
    near(s.coverage, 2.0 / 3.0, 1e-14, "partial coverage");
    const auto path = fs::temp_directory_path() / "rewl_selftest_stat.csv";
    write_workers_stat_csv(path.string(), {{0, 0, 0, 5, 0, 0, 5, 1.0, 1.0, 1, 6, 6.0, 1.0, 0}}, 2);
    std::ifstream in(path);
    std::string header, row;
    std::getline(in, header);
    std::getline(in, row);
    require(header == "mpi_rank,window,walker_id,attempted_flips,attempted_mcs,accepted,"
                      "forced_accepted,accepted_percent,attempts_since_last_accepted,"
                      "mcs_since_last_accepted,energy,factor,active_bins,min_h,mean_h,"
                      "min_over_mean,round_trips",
            "walker statistics CSV header");
    require(row == "0,0,0,5,2.5,0,0,0,5,2.5,1,1,1,6,6,1,0", "walker statistics CSV row");
    fs::remove(path);
}

void stitching_and_thermodynamics()
{ // test_main.cpp:400-418 on the 2x2x1 open exact DOS (16 states)
    const Grid grid{-20, 20, 0.25};
    const std::size_t bins = grid.bins();
    // exact counts of the 2x2x1 open lattice, +z axes: E = -2*sqrt(2)/... (bins from the
    // reference's enumeration)
    DosFragment exact{{0, bins}, std::vector<double>(bins, -std::numeric_limits<double>::infinity()),
                      std::vector<std::uint64_t>(bins, 0), std::vector<double>(bins, 0.0),
                      std::vector<std::uint8_t>(bins, 1)};
    const std::pair<std::size_t, std::uint64_t> counts[] = {{61, 2}, {77, 8}, {82, 4}, {98, 2}};
    for (auto [b, k] : counts) {
        exact.histogram[b] = k;
        exact.log_g[b] = std::log(static_cast<double>(k));
    }
    const auto dos = stitch_dos(grid, {exact}, true, 4);
    near(std::log(std::exp(dos.log_g[61]) + std::exp(dos.log_g[77]) + std::exp(dos.log_g[82]) +
                  std::exp(dos.log_g[98])),
         4 * std::log(2.0), 1e-12, "complete-range normalisation");
    const auto thermo = thermodynamics(dos, {1.0, 2.0});
    require(thermo.size() == 2 && std::isfinite(thermo[0].heat_capacity), "finite thermo");
    auto incomplete = exact;
    incomplete.valid[0] = 0;
    incomplete.log_g[0] = std::numeric_limits<double>::quiet_NaN();
    bool rejected = false;
    try {
        (void)stitch_dos(grid, {incomplete}, true, 4);
    } catch (const std::runtime_error &) {
        rejected = true;
    }
    require(rejected, "complete normalisation rejects invalid bins");
    const auto relative = stitch_dos(grid, {incomplete}, false, 4);
    require(relative.valid[0] == 0 && std::isfinite(thermodynamics(relative, {1.0})[0].heat_capacity),
            "relative thermodynamics ignores invalid bins");
    // two overlapping fragments of a linear log g join with zero shift error
    DosFragment left{{0, 60}, std::vector<double>(bins, std::nan("")), std::vector<std::uint64_t>(bins, 1),
                     std::vector<double>(bins, 0.0), std::vector<std::uint8_t>(bins, 0)};
    DosFragment right = left;
    right.window = {40, bins};
    for (std::size_t i = 0; i < 60; ++i) {
        left.valid[i] = 1;
        left.log_g[i] = 0.5 * static_cast<double>(i);
    }
    for (std::size_t i = 40; i < bins; ++i) {
        right.valid[i] = 1;
        right.log_g[i] = 0.5 * static_cast<double>(i) - 7.0;
    }
    const auto joined = stitch_dos(grid, {right, left}, false, 4);
    require(joined.join_bins.size() == 1 && joined.join_bins[0] >= 40 && joined.join_bins[0] < 60,
            "join inside the overlap");
    for (std::size_t i = 1; i < bins; ++i)
        near(joined.log_g[i] - joined.log_g[i - 1], 0.5, 1e-12, "stitched slope continuity");
}

} // namespace

// FILE: whitespace-separated numbers (strtod syntax, so nan/inf and hex floats work):
//   bins count overlap emin emax width per_window n
//   smoothing_width minimum_width diffusivity_floor_fraction curvature_weight
//   round_trip_target maximum_round_trip_penalty
//   per window: begin end round_trips walkers, squared[bins] samples[bins] log_g[bins] valid[bins]
//   R, then per representative: present energy spins[n]
// Prints "windows b e ...", "missing M external X" and one line of spins (or "-") per walker,
// or "error <what>" when the planner throws.
int adapt_main(const char *path)
{
    std::ifstream in(path);
    std::vector<std::string> tokens{std::istream_iterator<std::string>(in),
                                    std::istream_iterator<std::string>()};
    std::size_t at = 0;
    const auto num = [&] {
        if (at >= tokens.size()) throw std::runtime_error("truncated input");
        return std::strtod(tokens[at++].c_str(), nullptr);
    };
    const auto count_of = [&] { return static_cast<std::size_t>(num()); };
    const std::size_t bins = count_of(), count = count_of();
    const double overlap = num();
    Grid grid;
    grid.minimum = num();
    grid.maximum = num();
    grid.width = num();
    const std::size_t per_window = count_of(), n = count_of();
    AdaptiveWindowParameters p;
    p.smoothing_width = num();
    p.minimum_width = num();
    p.diffusivity_floor_fraction = num();
    p.curvature_weight = num();
    p.round_trip_target = num();
    p.maximum_round_trip_penalty = num();
    std::vector<DosFragment> fragments;
    std::vector<WindowSamplingStatistics> sampling;
    for (std::size_t w = 0; w < count; ++w) {
        WindowSamplingStatistics s;
        s.window.begin = count_of();
        s.window.end = count_of();
        s.round_trips = static_cast<std::uint64_t>(num());
        s.walkers = static_cast<std::uint64_t>(num());
        DosFragment f{s.window, {}, std::vector<std::uint64_t>(bins, 0),
                      std::vector<double>(bins, 0.0), {}};
        for (std::size_t i = 0; i < bins; ++i) s.squared_energy_displacement.push_back(num());
        for (std::size_t i = 0; i < bins; ++i)
            s.displacement_samples.push_back(static_cast<std::uint64_t>(num()));
        for (std::size_t i = 0; i < bins; ++i) f.log_g.push_back(num());
        for (std::size_t i = 0; i < bins; ++i) f.valid.push_back(static_cast<std::uint8_t>(num()));
        fragments.push_back(std::move(f));
        sampling.push_back(std::move(s));
    }
    std::vector<EnergyRepresentative> reps(count_of());
    for (auto &r : reps) {
        const bool present = num() != 0.0;
        r.energy = num();
        for (std::size_t i = 0; i < n; ++i) {
            const auto v = static_cast<std::int8_t>(num());
            if (present) r.spins.push_back(v);
        }
    }
    try {
        const auto windows = adapt_energy_windows(grid, fragments, sampling, count, overlap, p);
        std::cout << "windows";
        for (const auto &w : windows) std::cout << ' ' << w.begin << ' ' << w.end;
        std::cout << '\n';
        const auto starts = select_initial_configurations(grid, windows, reps, per_window);
        std::cout << "missing " << starts.missing << " external " << starts.external_warm_starts
                  << '\n';
        for (const auto &s : starts.spins) {
            if (s.empty()) std::cout << '-';
            for (std::size_t i = 0; i < s.size(); ++i) std::cout << (i ? " " : "") << int(s[i]);
            std::cout << '\n';
        }
    } catch (const std::exception &e) {
        std::cout << "error " << e.what() << '\n';
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 3 && std::string(argv[1]) == "adapt") return adapt_main(argv[2]);
    const std::vector<std::pair<const char *, std::function<void()>>> tests{
        {"csv_geometry_ini", csv_geometry_and_ini},
        {"grid_windows", grid_and_windows},
        {"checkpoint_layout", checkpoint_layout},
        {"run_checkpoint_layout", run_checkpoint_layout},
        {"summary", summary},
        {"statistics_and_stat_row", statistics_and_stat_row},
        {"stitching_and_thermodynamics", stitching_and_thermodynamics}};
    int failed = 0;
    for (const auto &[name, test] : tests) try {
            test();
            std::cout << "PASS " << name << '\n';
        } catch (const std::exception &e) {
            ++failed;
            std::cout << "FAIL " << name << ": " << e.what() << '\n';
        }
    std::cout << (failed ? "FAILED" : "ALL PASS") << '\n';
    return failed ? 1 : 0;
}
