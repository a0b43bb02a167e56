// CPU checks of the host C++ layer (no GPU): run-configuration parsing and MCS resolution,
// grids and window partitions, the single- and multi-walker checkpoint files, per-window
// summaries, histogram statistics and the walker statistics file, stitching and
// thermodynamics. Test infrastructure: built by tests/host/Makefile, run by tests/test_host.py.
//
// Modes:
//   host_checks                       every check, one PASS/FAIL line each, then a verdict
//   host_checks adapt FILE            the adaptive-window planner on the pilot data in FILE
//   host_checks geometry FILE B0 B1 B2 PERIODIC
//                                     load FILE (.csv: CSV loader, else the whitespace .dat
//                                     loader) and print its sites as hex floats
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
#include <stdexcept>
#include <vector>

using namespace rewl::host;
namespace fs = std::filesystem;

namespace {

void expect(bool ok, const std::string &what)
{
    if (!ok) throw std::runtime_error(what);
}

void expect_close(double a, double b, double tol, const std::string &what)
{
    if (std::abs(a - b) > tol * std::max({1.0, std::abs(a), std::abs(b)}))
        throw std::runtime_error(what);
}

// the exception text of `fn`, or "" when it does not throw
std::string error_of(const std::function<void()> &fn)
{
    try {
        fn();
    } catch (const std::exception &e) {
        return e.what();
    }
    return {};
}

RunConfig cli(std::vector<std::string> args)
{
    args.insert(args.begin(), "wl_run");
    std::vector<char *> argv;
    for (auto &a : args) argv.push_back(a.data());
    return parse_arguments(static_cast<int>(argv.size()), argv.data());
}

std::string read_text(const fs::path &p)
{
    std::ifstream in(p);
    return {std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>()};
}

WalkerState blank_walker(std::size_t n, std::size_t bins)
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

// energy grid index and window partition (wang_landau.cpp:13-61)
void check_grid_and_partition()
{
    const Grid g{-2, 2, 0.5};
    expect(g.bins() == 8, "eight bins of 0.5 span [-2, 2]");
    expect(g.index(-2) == 0, "the lower edge opens bin 0");
    expect(g.index(2) == 7, "the upper edge folds into the last bin");
    expect(g.index(2.1) == -1, "energies above the grid have no bin");
    const auto four = partition_windows(100, 4, 0.75);
    expect(four.size() == 4 && four.front().begin == 0 && four.back().end == 100,
           "four windows cover all 100 bins");
    for (std::size_t i = 1; i < four.size(); ++i)
        expect(four[i].begin < four[i - 1].end, "neighbouring windows share bins");
    const auto eight = partition_windows(1255, 8, 0.75);
    expect(eight[0].end == 457 && eight[1].begin == 114 && eight[7].begin == 798,
           "1255 bins in 8 windows at overlap 0.75: width 457, starts 114 apart");
}

// the CSV loader, INI keys and command line, MCS -> attempts (io.cpp:75-87, 210-236)
void check_configuration()
{
    const auto dir = fs::temp_directory_path() / "rewl_host_checks";
    fs::create_directories(dir);
    const auto csv = dir / "two_sites.csv", ini = dir / "settings.ini";
    std::ofstream(csv) << "x,y,z,mx,my,mz\n0,0,0,0,0,1\n1,0,0,0,0,1\n";
    std::ofstream(ini) << "[geometry]\nfile=two_sites.csv\nperiodic=false\n"
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
    const auto sites = Geometry::load_csv(csv.string(), {2, 2, 2}, false);
    expect(sites.size() == 2, "two CSV rows give two sites");
    expect_close(sites.axes[1].z, 1, 1e-14, "the CSV axis is read as written");

    auto c = cli({"--config", ini.string(), "--seed", "99"});
    c.resolve_mcs(sites.size());
    expect(c.geometry_file == csv.lexically_normal().string(),
           "a geometry file named in the INI resolves next to the INI");
    expect(c.seed == 99 && c.seed_explicit, "the command line overrides the INI seed");
    expect(c.energy_grid_explicit, "an INI [energy] section counts as an explicit grid");
    expect(!cli({}).seed_explicit, "no --seed leaves the seed unset");
    // N = 2: attempts = floor(mcs * 2 + 0.5)
    expect(c.exchange.attempts == 3, "1.5 MCS of 2 spins is 3 attempts between exchanges");
    expect(c.wl.check_interval_attempts == 5, "2.5 MCS check cadence is 5 attempts");
    expect(c.wl.force_accept_after_attempts == 7, "3.5 MCS forcing threshold is 7 attempts");
    expect(c.max_limit.attempts == 9, "4.5 MCS limit is 9 attempts");
    expect(c.checkpoint.attempts == 11, "5.5 MCS checkpoint cadence is 11 attempts");
    expect_close(c.progress_interval_seconds, 2.0, 1e-14, "INI progress period");
    expect(!c.wl.inverse_time_enabled && c.wl.nalivaiko_mod, "INI booleans of [wl]");
    expect(c.wl.initialization_max_attempts == 123, "INI initialization attempt cap");
    expect(c.wl.initialization_stall_attempts_per_spin == 17, "INI stall length");
    expect_close(c.wl.initialization_target_fraction, 0.4, 1e-14, "INI target fraction");
    expect_close(c.wl.initialization_temperature_fraction, 0.07, 1e-14, "INI start temperature");
    expect_close(c.wl.initialization_temperature_multiplier, 3.0, 1e-14, "INI heating factor");
    expect_close(c.wl.initialization_max_temperature_fraction, 0.4, 1e-14, "INI hottest");
    expect(c.adaptive_windows.enabled && c.adaptive_windows.iterations == 3,
           "INI adaptive-window switch and passes");
    expect_close(c.adaptive_windows.pilot_mcs, 12.0, 1e-14, "INI pilot length");
    expect_close(c.adaptive_windows.maximum_round_trip_penalty, 2.5, 1e-14, "INI penalty cap");
    expect_close(c.wl.round_trip_margin_fraction, 0.15, 1e-14, "INI round-trip margin");
    expect(c.temperatures.size() == 2 && c.output_prefix == "from_ini",
           "INI temperature list and output prefix");
    expect(!c.uses_legacy_attempt_units(), "MCS-valued options are not the legacy form");

    const auto flags = cli({"--progress", "3", "--inverse-time", "true"});
    expect_close(flags.progress_interval_seconds, 3.0, 1e-14, "--progress seconds");
    expect(flags.wl.inverse_time_enabled && !flags.wl.nalivaiko_mod, "boolean flags");
    expect(cli({"--progress", "0"}).progress_interval_seconds == 0.0, "--progress 0 disables");
    auto legacy = cli({"--max-attempts", "9"});
    legacy.resolve_mcs(2);
    expect(legacy.max_limit.attempts == 9 && legacy.max_limit.mcs == 4.5,
           "--max-attempts 9 of 2 spins reads back as 4.5 MCS");
    expect(legacy.uses_legacy_attempt_units(), "--max-attempts is flagged as the legacy form");

    const auto bad = dir / "bad.ini";
    std::ofstream(bad) << "[wl]\nnalivaiko_mod=maybe\n";
    expect(error_of([&] { cli({"--config", bad.string()}); }) ==
               "INI line 2: Invalid boolean for wl.nalivaiko_mod",
           "a non-boolean switch names its line and key");
    std::ofstream(bad) << "[energy]\nbogus=1\n";
    expect(error_of([&] { cli({"--config", bad.string()}); }) ==
               "INI line 2: Unknown INI setting: energy.bogus",
           "an unknown INI key names its line");
    expect(error_of([] { cli({"--bogus", "1"}); }) == "Unknown option: --bogus",
           "an unknown flag is named");
    expect(error_of([] { cli({"--nx"}); }) == "Missing value for --nx", "a flag without value");
    expect(error_of([] { cli({"--nx", "3x"}); }) == "Invalid value for --nx",
           "an integer with trailing text");

    const auto meta = dir / "meta.json";
    write_metadata_json(meta.string(), c, 2, "dense", 0, 0, 0, 0, 0, false, 1, 1, nullptr);
    const auto text = read_text(meta);
    expect(text.find("\"nalivaiko_mod\": true") != std::string::npos,
           "metadata lists the Nalivaiko switch");
    expect(text.find("\"refinement_active_scope\": \"current_iteration\"") != std::string::npos,
           "metadata lists the refinement mask scope");
    const auto preset = cli({"--smoke-test"});
    expect(preset.nx == 2 && preset.ny == 2 && preset.nz == 1 && preset.grid.bins() == 64,
           "--smoke-test lattice and grid");
    expect(preset.output_prefix == "wl_smoke" && preset.wl.final_factor == 0.5,
           "--smoke-test prefix and final ln f");
    fs::remove_all(dir);
}

// one walker's WLCHKP3 file (io.cpp:539-572): 2929 bytes for 8 spins and 160 bins
void check_walker_checkpoint()
{
    auto w = blank_walker(8, 160);
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
    const auto path = (fs::temp_directory_path() / "rewl_host_checks_walker.chk").string();
    save_checkpoint(path, w);
    expect(fs::file_size(path) == 2929, "record length for 8 spins and 160 bins");
    const auto r = load_checkpoint(path);
    expect(r.spins == w.spins && r.fields == w.fields && r.energy == w.energy,
           "configuration read back");
    expect(r.log_g == w.log_g && r.histogram == w.histogram && r.active == w.active,
           "estimator arrays read back");
    expect(r.factor == w.factor && r.stage == 1 && r.attempted == 100 && r.accepted == 40 &&
               r.forced_accepted == 2 && r.last_accepted_attempt == 99,
           "scalars read back");
    expect(r.rng_state[0] == 1 && r.rng_state[3] == 0xdeadbeef, "generator words read back");
    fs::resize_file(path, 100);
    expect(error_of([&] { (void)load_checkpoint(path); }).rfind("Truncated checkpoint", 0) == 0,
           "a cut-off record is refused");
    fs::remove(path);
}

// a whole run's WLCHKM2 file: 40-byte header, 24 bytes per walker (cadence and 1/t clock
// offset), then one WLCHKP3 record per walker
void check_run_checkpoint()
{
    RunCheckpoint run{12, 30, 7, {}};
    for (std::uint64_t k = 0; k < 3; ++k) {
        auto w = blank_walker(8, 160);
        w.id = k;
        w.attempted = 100 + k;
        w.last_flatness_check = 50 + k;
        w.last_flatness = 0.25 * double(k);
        w.clock_offset = 1000 * k + 7;
        w.rng_state[2] = 0x1234 + k;
        run.walkers.push_back(w);
    }
    const auto path = (fs::temp_directory_path() / "rewl_host_checks_run.chk").string();
    save_run_checkpoint(path, run);
    expect(fs::file_size(path) == 40 + 3 * 24 + 3 * 2929, "run file length for 3 walkers");
    expect(read_text(path).substr(0, 8) == std::string("WLCHKM2\0", 8), "run file tag");
    const auto r = load_run_checkpoint(path);
    expect(r.epoch == 12 && r.exchange_attempted == 30 && r.exchange_accepted == 7,
           "run counters read back");
    expect(r.walkers.size() == 3, "walker count read back");
    for (std::uint64_t k = 0; k < 3; ++k) {
        const auto &x = r.walkers[k];
        expect(x.id == k && x.attempted == 100 + k && x.spins == run.walkers[k].spins &&
                   x.rng_state[2] == 0x1234 + k,
               "walker record " + std::to_string(k));
        expect(x.last_flatness_check == 50 + k && x.last_flatness == 0.25 * double(k),
               "driver cadence of walker " + std::to_string(k));
        expect(x.clock_offset == 1000 * k + 7, "1/t clock offset of walker " + std::to_string(k));
    }
    { // the previous tag (no clock offsets) still loads, with offset 0
        std::string bytes = read_text(path);
        std::string older = bytes.substr(0, 40);
        older[6] = '1';
        for (int k = 0; k < 3; ++k) older += bytes.substr(40 + 24 * k, 16);
        older += bytes.substr(40 + 3 * 24);
        std::ofstream(path, std::ios::binary) << older;
        const auto o = load_run_checkpoint(path);
        expect(o.walkers.size() == 3 && o.walkers[2].last_flatness_check == 52 &&
                   o.walkers[2].clock_offset == 0 && o.walkers[2].rng_state[2] == 0x1236,
               "a WLCHKM1 file loads with clock offset 0");
    }
    save_checkpoint(path, run.walkers[0]);
    expect(error_of([&] { (void)load_run_checkpoint(path); }) ==
               "Invalid multi-walker checkpoint format",
           "a single-walker record is not a run file");
    fs::remove(path);
}

// the per-window fragment of parallel.cpp:23-62 on two hand-made walkers
void check_window_summary()
{
    auto first = blank_walker(1, 4), second = blank_walker(1, 4);
    first.active = {1, 1, 1, 0};
    second.active = {0, 1, 1, 1};
    first.log_g[1] = 10.0;
    first.log_g[2] = 14.0;
    second.log_g[1] = -3.0;
    second.log_g[2] = 3.0;
    first.histogram[2] = 7;
    second.histogram[2] = 9;
    const auto f = summarize_walkers({0, 4}, {&first, &second}, 4);
    expect(f.valid == std::vector<std::uint8_t>{0, 1, 1, 0}, "valid = bins both walkers found");
    expect_close(f.log_g[1], 0.0, 1e-14, "the first valid bin is the zero of the fragment");
    // offsets 4 and 6 above the first valid bin: mean 5, standard error |6-4|/2 = 1
    expect_close(f.log_g[2], 5.0, 1e-14, "mean of the offsets");
    expect_close(f.standard_error[2], 1.0, 1e-14, "standard error of two offsets");
    expect(f.histogram[2] == 16, "histograms add");
    expect(std::isnan(f.log_g[0]) && std::isnan(f.standard_error[0]), "invalid bins hold NaN");
    expect(std::isnan(summarize_walkers({0, 4}, {&first}, 4).standard_error[1]),
           "one walker has no standard error");
    second.active = {0, 0, 0, 1};
    expect(!error_of([&] { (void)summarize_walkers({0, 4}, {&first, &second}, 4); }).empty(),
           "walkers without a shared bin cannot be summarised");
}

// histogram_statistics (wang_landau.cpp:344-367) and one row of the walker statistics file
void check_statistics()
{
    auto w = blank_walker(1, 4);
    w.active = {1, 1, 0, 0};
    w.histogram = {80, 120, 0, 0};
    w.active_bins = 2;
    auto s = histogram_statistics(w, false);
    expect(s.active_bins == 2 && s.covered_bins == 2 && s.minimum == 80, "two visited bins");
    expect_close(s.mean, 100.0, 1e-14, "mean of 80 and 120");
    expect_close(s.min_over_mean, 0.8, 1e-14, "flatness 80/100");
    w.active[2] = 1;
    w.active_bins = 3;
    s = histogram_statistics(w, false);
    expect(s.minimum == 0 && s.covered_bins == 2, "an active bin without visits this stage");
    expect_close(s.mean, 200.0 / 3.0, 1e-14, "mean over three active bins");
    expect_close(s.coverage, 2.0 / 3.0, 1e-14, "two of three active bins covered");
    const auto path = fs::temp_directory_path() / "rewl_host_checks_stat.csv";
    write_workers_stat_csv(path.string(), {{0, 0, 0, 5, 0, 0, 5, 1.0, 1.0, 1, 6, 6.0, 1.0, 0}}, 2);
    std::ifstream in(path);
    std::string header, row;
    std::getline(in, header);
    std::getline(in, row);
    expect(header == "mpi_rank,window,walker_id,attempted_flips,attempted_mcs,accepted,"
                     "forced_accepted,accepted_percent,attempts_since_last_accepted,"
                     "mcs_since_last_accepted,energy,factor,active_bins,min_h,mean_h,"
                     "min_over_mean,round_trips",
           "statistics file columns");
    // 5 attempts of 2 spins = 2.5 MCS, none accepted
    expect(row == "0,0,0,5,2.5,0,0,0,5,2.5,1,1,1,6,6,1,0", "statistics row of a 2-spin walker");
    fs::remove(path);
}

// stitch_dos and thermodynamics (analysis.cpp:45-160) on exact and synthetic fragments
void check_stitching()
{
    const Grid grid{-20, 20, 0.25};
    const std::size_t bins = grid.bins();
    // the 2x2x1 open +z lattice by enumeration: 16 states in four bins
    DosFragment exact{{0, bins},
                      std::vector<double>(bins, -std::numeric_limits<double>::infinity()),
                      std::vector<std::uint64_t>(bins, 0), std::vector<double>(bins, 0.0),
                      std::vector<std::uint8_t>(bins, 1)};
    const std::pair<std::size_t, std::uint64_t> levels[] = {{61, 2}, {77, 8}, {82, 4}, {98, 2}};
    for (auto [b, k] : levels) {
        exact.histogram[b] = k;
        exact.log_g[b] = std::log(static_cast<double>(k));
    }
    const auto dos = stitch_dos(grid, {exact}, true, 4);
    double total = 0.0;
    for (auto [b, k] : levels) total += std::exp(dos.log_g[b]);
    expect_close(std::log(total), 4 * std::log(2.0), 1e-12, "normalised to 2^4 states");
    const auto thermo = thermodynamics(dos, {1.0, 2.0});
    expect(thermo.size() == 2 && std::isfinite(thermo[0].heat_capacity),
           "one finite point per temperature");
    auto gap = exact;
    gap.valid[0] = 0;
    gap.log_g[0] = std::numeric_limits<double>::quiet_NaN();
    expect(!error_of([&] { (void)stitch_dos(grid, {gap}, true, 4); }).empty(),
           "full normalisation needs every bin");
    const auto relative = stitch_dos(grid, {gap}, false, 4);
    expect(relative.valid[0] == 0 &&
               std::isfinite(thermodynamics(relative, {1.0})[0].heat_capacity),
           "relative normalisation skips invalid bins");
    // two fragments of log g = i/2 with a constant offset between them join seamlessly
    DosFragment lower{{0, 60}, std::vector<double>(bins, std::nan("")),
                      std::vector<std::uint64_t>(bins, 1), std::vector<double>(bins, 0.0),
                      std::vector<std::uint8_t>(bins, 0)};
    DosFragment upper = lower;
    upper.window = {40, bins};
    for (std::size_t i = 0; i < 60; ++i) {
        lower.valid[i] = 1;
        lower.log_g[i] = 0.5 * static_cast<double>(i);
    }
    for (std::size_t i = 40; i < bins; ++i) {
        upper.valid[i] = 1;
        upper.log_g[i] = 0.5 * static_cast<double>(i) - 7.0;
    }
    const auto joined = stitch_dos(grid, {upper, lower}, false, 4);
    expect(joined.join_bins.size() == 1 && joined.join_bins[0] >= 40 && joined.join_bins[0] < 60,
           "one join, inside the shared bins");
    for (std::size_t i = 1; i < bins; ++i)
        expect_close(joined.log_g[i] - joined.log_g[i - 1], 0.5, 1e-12, "slope 1/2 everywhere");
}

// FILE: whitespace-separated numbers (strtod syntax, so nan/inf and hex floats work):
//   bins count overlap emin emax width per_window n
//   smoothing_width minimum_width diffusivity_floor_fraction curvature_weight
//   round_trip_target maximum_round_trip_penalty
//   per window: begin end round_trips walkers, squared[bins] samples[bins] log_g[bins] valid[bins]
//   R, then per representative: present energy spins[n]
// Prints "windows b e ...", "missing M external X" and one line of spins (or "-") per walker,
// or "error <what>" when the planner throws.
int plan_windows(const char *path)
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

// "sites N", then one line "x y z mx my mz" of hex floats per site, or "error <what>"
int load_sites(const char *path, const char *b0, const char *b1, const char *b2,
               const char *periodic)
{
    try {
        const Vec3 box{std::strtod(b0, nullptr), std::strtod(b1, nullptr), std::strtod(b2, nullptr)};
        const bool per = std::string(periodic) != "0";
        const Geometry g = fs::path(path).extension() == ".csv"
                               ? Geometry::load_csv(path, box, per)
                               : Geometry::load_xyz_axes(path, box, per);
        std::cout << "sites " << g.size() << '\n';
        for (std::size_t i = 0; i < g.size(); ++i) {
            const double v[6] = {g.positions[i].x, g.positions[i].y, g.positions[i].z,
                                 g.axes[i].x,      g.axes[i].y,      g.axes[i].z};
            for (int k = 0; k < 6; ++k) std::printf(k ? " %a" : "%a", v[k]);
            std::printf("\n");
        }
    } catch (const std::exception &e) {
        std::cout << "error " << e.what() << '\n';
    }
    return 0;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc == 3 && std::string(argv[1]) == "adapt") return plan_windows(argv[2]);
    if (argc == 7 && std::string(argv[1]) == "geometry")
        return load_sites(argv[2], argv[3], argv[4], argv[5], argv[6]);
    const std::vector<std::pair<const char *, std::function<void()>>> checks{
        {"grid_and_partition", check_grid_and_partition},
        {"configuration", check_configuration},
        {"walker_checkpoint", check_walker_checkpoint},
        {"run_checkpoint", check_run_checkpoint},
        {"window_summary", check_window_summary},
        {"statistics", check_statistics},
        {"stitching", check_stitching}};
    int failed = 0;
    for (const auto &[name, check] : checks) try {
            check();
            std::cout << "PASS " << name << '\n';
        } catch (const std::exception &e) {
            ++failed;
            std::cout << "FAIL " << name << ": " << e.what() << '\n';
        }
    std::cout << (failed ? "FAILED" : "ALL PASS") << '\n';
    return failed ? 1 : 0;
}
