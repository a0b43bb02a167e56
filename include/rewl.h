/*
 * rewl.h — C ABI of the MI355X replica-exchange Wang-Landau engine.
 *
 * The reference (Faikert/replica_exchange_wang_landau) has no FFI layer; its seams are C++ APIs.
 * Each entry point below replaces one of them and cites it (file:line in the reference tree).
 * A host C++20 shim (replica_exchange_wang_landau_amd/host) keeps the reference's `run_rewl`
 * signature and `wl_run` command line on top of this ABI.
 *
 * Conventions
 *   - Every call returns an int status (WLX_OK = 0). Codes map 1:1 to the exception classes the
 *     reference throws; the message of the last failure on the calling thread is returned by
 *     wlx_last_error() and reproduces the reference texts where its tests assert them.
 *   - The caller owns every host array; the library owns device buffers and one HIP stream per
 *     engine. Calls on one handle are not re-entrant: use one host thread per handle.
 *   - All structs hold only 8-byte members (no padding), so they bind directly from ctypes/cgo.
 *   - rng_state[4]: in WLX_RNG_XOSHIRO mode, the Xoshiro256** state words (rng.hpp:10-56), as in
 *     the WLCHKP3 checkpoint (io.cpp:539-556). In WLX_RNG_PHILOX mode production attempts draw
 *     Philox4x64-10 blocks keyed (seed, walker_id) at counter (attempt/2, 'WALK'); the stream
 *     position is the attempted count, and rng_state still carries the Xoshiro words used by
 *     the target-Metropolis initialisation.
 */
#ifndef REWL_H
#define REWL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum wlx_status {
    WLX_OK = 0,
    WLX_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    WLX_RUNTIME_ERROR = 2,    /* std::runtime_error */
    WLX_LOGIC_ERROR = 3,      /* std::logic_error */
    WLX_OUT_OF_RANGE = 4,     /* std::out_of_range */
    WLX_DEVICE_ERROR = 5      /* HIP runtime failure; no reference counterpart */
};

/* RefinementStage, wang_landau.hpp:58 */
enum wlx_stage { WLX_STAGE_WANG_LANDAU = 0, WLX_STAGE_INVERSE_TIME = 1, WLX_STAGE_FROZEN = 2 };
/* proposal/uniform stream: reference-compatible Xoshiro256** or counter-based Philox4x64-10 */
enum wlx_rng { WLX_RNG_XOSHIRO = 0, WLX_RNG_PHILOX = 1 };
/* local-field precision: FP64 (reference arithmetic), or FP32 fields with Kahan-compensated
 * incremental updates over FP32 couplings, resynchronised periodically from an FP64 H = J.S */
enum wlx_fields { WLX_FIELDS_FP64 = 0, WLX_FIELDS_FP32_KAHAN = 1 };
/* coupling precision of a system: FP64 as the reference builds it, or every J_ij rounded to the
 * nearest binary32 value after the FP64 build (CSR values then stream as 4-byte floats).
 * Fields, energies and log g stay FP64 either way, so an FP32-coupling run is the reference
 * algorithm on the rounded couplings. */
enum wlx_couplings { WLX_COUPLINGS_FP64 = 0, WLX_COUPLINGS_FP32 = 1 };
/* Walker kernel families (all give the same trajectory):
 * AUTO       lookahead for runs without statistics collection;
 * LOOKAHEAD  batches of proposals resolved together, fields in registers (N <= 4096);
 * SEQUENTIAL one attempt at a time, fields in registers;
 * LDS        one attempt at a time, fields and spins in LDS; a flip scatters the CSR row of a
 *            CSR system or streams the dense row otherwise. Serves any N in one kernel.
 * STAGED     lookahead with each batch's candidate rows copied into LDS while the previous
 *            batch decides (Philox stream, no trace, forcing or Nalivaiko mask, 65 to 512
 *            sites, two 16-row buffers and the window bins within a CU's LDS). AUTO picks it
 *            for such engines with at most one walker per CU, unless a lookahead width was
 *            set; batches never use it. */
enum wlx_walker_kernel {
    WLX_KERNEL_AUTO = 0,
    WLX_KERNEL_LOOKAHEAD = 1,
    WLX_KERNEL_SEQUENTIAL = 2,
    WLX_KERNEL_LDS = 3,
    WLX_KERNEL_STAGED = 4
};

/* trace flags, one byte per recorded attempt */
enum wlx_trace_flag {
    WLX_TRACE_ACCEPTED = 1,
    WLX_TRACE_FORCED = 2,
    WLX_TRACE_UNIFORM = 4,   /* a uniform was drawn for the decision */
    WLX_TRACE_IN_WINDOW = 8  /* the proposed bin lies inside the walker's window */
};

typedef struct wlx_system wlx_system;
typedef struct wlx_engine wlx_engine;

/* WlParameters, wang_landau.hpp:40-56 (booleans widened to uint64). The last two members
 * drive the pilot statistics of adaptive windows (wang_landau.cpp:104-118, 272-276, 317-342);
 * they sit at the end so that the reference's member order holds for everything before. */
typedef struct wlx_params {
    double flatness;
    uint64_t minimum_visits;
    double final_factor;
    uint64_t check_interval_attempts;
    uint64_t force_accept_after_attempts;
    uint64_t inverse_time_enabled;
    uint64_t initialization_max_attempts;
    double initialization_target_fraction;
    double initialization_temperature_fraction;
    uint64_t initialization_stall_attempts_per_spin;
    double initialization_temperature_multiplier;
    double initialization_max_temperature_fraction;
    uint64_t nalivaiko_mod;
    uint64_t collect_window_statistics;
    double round_trip_margin_fraction; /* in [0, 0.5) */
} wlx_params;

/* The part of RunConfig (io.hpp:12-58) that run_rewl (parallel.cpp:276-581) reads, with every
 * interval already resolved to attempts (RunConfig::resolve_mcs, io.cpp:210-236). */
typedef struct wlx_config {
    double emin;                         /* EnergyGrid::minimum */
    double emax;                         /* EnergyGrid::maximum */
    double bin_width;                    /* EnergyGrid::width */
    uint64_t windows;                    /* RunConfig::windows */
    uint64_t walkers_per_window;         /* RunConfig::walkers_per_rank */
    double overlap;                      /* RunConfig::overlap */
    uint64_t seed;                       /* RunConfig::seed */
    uint64_t exchange_interval_attempts; /* RunConfig::exchange_interval_attempts */
    uint64_t max_attempts;               /* RunConfig::max_attempts; 0 = until all frozen */
    uint64_t rng;                        /* enum wlx_rng */
    uint64_t fields;                     /* enum wlx_fields */
    const uint64_t *explicit_windows;    /* NULL or 2*windows words (begin, end) */
    wlx_params wl;
    uint64_t resync_interval_attempts;   /* FP32 fields: FP64 resync cadence; 0 = never */
} wlx_config;

/* Per-walker scalars (WalkerSnapshot, wang_landau.hpp:74-89, plus driver bookkeeping). */
typedef struct wlx_walker_info {
    uint64_t walker_id;
    uint64_t window;
    uint64_t window_begin;
    uint64_t window_end;
    double energy;
    double factor;
    uint64_t attempted;
    uint64_t accepted;
    uint64_t forced_accepted;
    uint64_t last_accepted_attempt;
    uint64_t stage; /* enum wlx_stage */
    uint64_t active_bins;
    uint64_t refinement_active_bins;
    uint64_t rng_state[4];
    uint64_t last_flatness_check; /* parallel.cpp:364 */
    double last_flatness;         /* parallel.cpp:365 */
    /* 1/t clock offset set by wlx_engine_warm_start (0 otherwise): the 1/t factor is
     * active / (attempted + clock_offset). No reference counterpart; restored by
     * wlx_engine_write_walker so that a resumed warm-started run continues exactly. */
    uint64_t clock_offset;
} wlx_walker_info;

/* HistogramStatistics, wang_landau.hpp:60-67 */
typedef struct wlx_histogram_stats {
    uint64_t active_bins;
    uint64_t covered_bins;
    uint64_t minimum;
    double mean;
    double min_over_mean;
    double coverage;
} wlx_histogram_stats;

/* Run counters (RewlResult, parallel.hpp:23-31) plus device diagnostics. */
typedef struct wlx_run_status {
    uint64_t epoch;               /* exchange epochs completed */
    uint64_t done;                /* all frozen or attempt limit reached (parallel.cpp:527-533) */
    uint64_t converged;           /* every walker frozen */
    uint64_t attempted;
    uint64_t accepted;
    uint64_t forced_accepted;
    uint64_t exchange_attempted;
    uint64_t exchange_accepted;
    uint64_t ambiguous_decisions; /* log(u) within 4 ulp of its threshold (libm-sensitive) */
} wlx_run_status;

/* ---- status ---------------------------------------------------------------------------- */

/* Copies the message of the last failed call on this thread (NUL-terminated, truncated to
 * capacity) and returns its full length. Replaces the what() texts of the reference. */
size_t wlx_last_error(char *buffer, size_t capacity);

/* ---- coupling system: kernel (a) ------------------------------------------------------- */

/* Builds J_ij on the device. cutoff <= 0 selects the dense N x N backend (DenseCouplings,
 * physics.cpp:172-184), cutoff > 0 the CSR backend with norm(d) <= cutoff (CsrCouplings,
 * physics.cpp:197-223). positions/axes are N x 3 row-major; box holds the three box lengths
 * (0 = open axis), periodic as in Box (physics.hpp:17-20). Validates like Geometry::validate
 * (physics.cpp:34-56). precision: enum wlx_couplings (the reference has FP64 only). */
int wlx_system_create(const double *positions, const double *axes, uint64_t n,
                      const double box[3], int periodic, double coupling_scale, double cutoff,
                      int precision, wlx_system **out);
int wlx_system_destroy(wlx_system *system);
/* Couplings::size (physics.hpp:44) and backend_name: 0 = dense, 1 = csr */
int wlx_system_info(const wlx_system *system, uint64_t *n, uint64_t *backend, uint64_t *nnz);
/* Copies J to the host: dense N*N row-major (DenseCouplings::at, physics.cpp:186-188). */
int wlx_system_read_dense(const wlx_system *system, double *out);
/* Copies the CSR arrays: offsets[N+1], columns[nnz], values[nnz] (physics.cpp:197-223). */
int wlx_system_read_csr(const wlx_system *system, uint64_t *offsets, uint64_t *columns,
                        double *values);
/* local_fields + total_energy (physics.cpp:238-259) for `count` configurations of N spins,
 * in the reference's summation order. fields may be NULL. */
int wlx_system_fields(const wlx_system *system, const int8_t *spins, uint64_t count,
                      double *fields, double *energies);
/* The same computation as wlx_system_fields without any device allocation. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops). Staging goes through a fixed 4 MiB device scratch array (one
 * per GPU, held until the process ends): count * (padded N) * 9 bytes + 8 bytes per
 * configuration must fit (e.g. up to 900 configurations of N = 512, 110 of N = 4096), larger
 * requests fail with WLX_INVALID_ARGUMENT. Calls that overlap on one GPU share the scratch
 * and race; the system must live on the current device. */
int wlx_system_fields_noalloc(const wlx_system *system, const int8_t *spins, uint64_t count,
                              double *fields, double *energies);
/* The wl_run --pilot sweep (wl_run.cpp:84-104): from all spins +1, `attempts` unconditional
 * single-spin flips drawn from Xoshiro256**(seed, 0), tracking the sampled energy range. */
int wlx_system_pilot(const wlx_system *system, uint64_t seed, uint64_t attempts,
                     double *minimum, double *maximum);
/* The same sweep for `count` systems in one launch (one workgroup each; attempts[k] flips for
 * system k), fields and spins in LDS: same results as wlx_system_pilot per system. No
 * reference counterpart (the reference pilots one system per run). */
int wlx_systems_pilot(wlx_system *const *systems, uint64_t count, uint64_t seed,
                      const uint64_t *attempts, double *minimum, double *maximum);
/* Name of the device the library runs on and its compute-unit count. No reference
 * counterpart. */
int wlx_device_info(char *name, size_t capacity, uint64_t *compute_units);

/* ---- engine: kernels (b) and the device-side replica exchange --------------------------- */

/* Allocates device-resident walkers for run_rewl (parallel.cpp:276-360): windows from
 * partition_windows (wang_landau.cpp:40-61) or explicit, walker id = window*walkers + local. */
int wlx_engine_create(wlx_system *system, const wlx_config *config, wlx_engine **out);
int wlx_engine_destroy(wlx_engine *engine);
/* Walker construction + target-Metropolis initialisation (wang_landau.cpp:63-242) of every
 * walker on the device. initial_spins: NULL, or walkers*N spins used where supplied[w] != 0
 * (RunConfig::initial_spins_by_walker, parallel.cpp:315-319). On failure returns
 * WLX_RUNTIME_ERROR with "Walker initialization failed: window W, local walker L: ..."
 * (parallel.cpp:323-352). */
int wlx_engine_initialize(wlx_engine *engine, const int8_t *initial_spins,
                          const uint8_t *supplied);
/* WangLandauWalker::run_attempts (wang_landau.cpp:286-288) on every walker. */
int wlx_engine_sample(wlx_engine *engine, uint64_t attempts);
/* The flatness/coverage check cadence of one epoch (parallel.cpp:395-407). */
int wlx_engine_check(wlx_engine *engine);
/* Even/odd replica exchange of one epoch (parallel.cpp:410-421, 157-166). */
int wlx_engine_exchange(wlx_engine *engine, uint64_t epoch, uint64_t *attempted,
                        uint64_t *accepted);
/* The run_rewl epoch loop (parallel.cpp:376-534) for at most max_epochs epochs; stops early
 * when every walker is frozen or the attempt limit is reached. Resumable. Returns after the
 * device has finished the epochs; `status` may be NULL, and is then not read back. */
int wlx_engine_run(wlx_engine *engine, uint64_t max_epochs, wlx_run_status *status);
int wlx_engine_status(wlx_engine *engine, wlx_run_status *status);
/* Replays `epochs_per_graph` epochs of wlx_engine_run (the sampling and exchange launches of
 * parallel.cpp:376-421) from one HIP graph, recaptured when the launch parameters or the
 * walker kernel change; 0 = plain launches (the default). Timed runs (wlx_engine_set_timing)
 * and FP32-field runs with a resync cadence launch plainly. No reference counterpart. */
int wlx_engine_set_graph(wlx_engine *engine, uint64_t epochs_per_graph);
/* Restores the run counters of a resumed run: the exchange epoch (which drives the even/odd
 * parity and the exchange random stream, parallel.cpp:410-421) and the exchange tallies, as
 * stored by a multi-walker checkpoint next to the per-walker WLCHKP3 records. Clears the
 * finished flag; the next epoch re-evaluates it (parallel.cpp:527-533). */
int wlx_engine_set_run_state(wlx_engine *engine, uint64_t epoch, uint64_t exchange_attempted,
                             uint64_t exchange_accepted);
int wlx_engine_walkers(const wlx_engine *engine, uint64_t *walkers, uint64_t *spins,
                       uint64_t *bins);
/* Device layout of an engine's walkers: the LDS bytes of one LDS-walker workgroup (fields,
 * spins and, unless bins_in_hbm, the window bins) and whether the window bins stay in HBM
 * because the widest window does not fit in LDS (then only the LDS walker samples). No
 * reference counterpart. */
int wlx_engine_layout(const wlx_engine *engine, uint64_t *lds_bytes, uint64_t *bins_in_hbm);
/* WangLandauWalker::snapshot (wang_landau.cpp:415-418). Arrays are full-grid (bins); any may
 * be NULL. */
int wlx_engine_read_walker(wlx_engine *engine, uint64_t walker, wlx_walker_info *info,
                           int8_t *spins, double *fields, double *log_g, uint64_t *histogram,
                           uint8_t *active, uint8_t *refinement_active);
/* Scalars of every walker at once (walkers entries), for progress lines and statistics. */
int wlx_engine_read_infos(wlx_engine *engine, wlx_walker_info *infos);
/* WangLandauWalker::restore (wang_landau.cpp:420-466): same validation and messages. */
int wlx_engine_write_walker(wlx_engine *engine, uint64_t walker, const wlx_walker_info *info,
                            const int8_t *spins, const double *fields, const double *log_g,
                            const uint64_t *histogram, const uint8_t *active);
/* The driver's flatness-check cadence of one walker (the last_flatness_check and
 * last_flatness entries of run_rewl, parallel.cpp:364-365, which restore leaves at 0 like the
 * reference's 1x1 resume); a multi-walker checkpoint restores them for an exact resume. */
int wlx_engine_set_walker_cadence(wlx_engine *engine, uint64_t walker,
                                  uint64_t last_flatness_check, double last_flatness);
/* FP32 fields: recomputes every configuration's fields H = J.S in FP64 (one matrix product on
 * the matrix cores over all walkers), measures the drift of the running FP32 fields and
 * energies against it, and restarts them from the FP64 values. Replaces local_fields and
 * total_energy (physics.cpp:238-259) as a periodic correction; WLX_LOGIC_ERROR in FP64 mode.
 * Also run by wlx_engine_run every resync_interval_attempts. */
int wlx_engine_resync(wlx_engine *engine);
/* Largest drift seen at any resync so far: max |h_fp32 - h_fp64| over all sites and
 * configurations, max |E - E_fp64| / bin_width over configurations (the energy drift in
 * bins: what decides whether a walker could be booked into a wrong bin), and the resync
 * count. */
int wlx_engine_drift(wlx_engine *engine, double *max_field_error, double *max_energy_bins,
                     uint64_t *resyncs);
/* The pilot statistics of a walker run with collect_window_statistics: squared accepted
 * energy displacements and attempt counts per bin (full grid, zero outside the window;
 * wang_landau.cpp:272-276), completed round trips (wang_landau.cpp:317-330), and the nine
 * representative configurations closest to the window's energy targets
 * (wang_landau.cpp:107-118, 332-342): energies, a present flag each, and N spins each. Any
 * pointer may be NULL. WLX_LOGIC_ERROR when the engine does not collect them. */
int wlx_engine_read_sampling(wlx_engine *engine, uint64_t walker, double *squared_displacement,
                             uint64_t *displacement_samples, uint64_t *round_trips,
                             double *representative_energies, uint8_t *representative_present,
                             int8_t *representative_spins);
/* Warm start (no reference counterpart: the reference starts WL from log g = 0, ln f = 1,
 * wang_landau.cpp:91-118). Presets every walker of an initialized engine: its window of
 * log_g (full grid, any offset, finite everywhere), ln f = factor (0 < factor <= 1) in the WL
 * stage, empty histograms and visited-bin masks. The 1/t clock starts at window_bins / factor,
 * so that after the switch to 1/t the factor continues near `factor` (active / (attempted +
 * window_bins / factor)). Configurations, counters and streams are kept. */
int wlx_engine_warm_start(wlx_engine *engine, const double *log_g, double factor);
/* Kernel selection. No reference counterpart: the reference has one CPU walker loop
 * (wang_landau.cpp:244-288). set_kernel forces a family (enum wlx_walker_kernel); kernel
 * reports the family and kernel name the next sampling launch uses; autotune runs `epochs`
 * epochs of the live run with each eligible family (the trajectory does not depend on it),
 * keeps the fastest and writes the seconds per family to seconds[5] (-1 = not tried). The
 * Philox lookahead family is tried at widths 16, 8 and 32 (wlx_engine_set_lookahead_width). */
int wlx_engine_set_kernel(wlx_engine *engine, uint64_t kernel);
int wlx_engine_kernel(const wlx_engine *engine, uint64_t *kernel, char *name, size_t capacity);
int wlx_engine_autotune(wlx_engine *engine, uint64_t epochs, uint64_t *chosen, double *seconds);
/* Lookahead width of the Philox common-configuration kernel: attempts resolved per batch and
 * rows streamed together after it (4, 8, 16 = default, 32; 32 only for N <= 1024). Every
 * width gives the same trajectory. No reference counterpart. */
int wlx_engine_set_lookahead_width(wlx_engine *engine, uint64_t width);
/* ---- batches of engines (uneven systems in one launch) -------------------------------------
 * No reference counterpart (the reference runs one system per process). A batch runs the
 * epochs of several initialized engines together: per epoch one sampling launch over every
 * walker of every engine (one wave per walker) and one exchange launch (one workgroup per
 * engine). The sampling launch runs the kernel every engine would run alone when that is one
 * register kernel for all of them (one family and padded N, e.g. lookahead_fast_kernel<8> for
 * replicas of one system), else the LDS walker. Each engine keeps its own schedule,
 * gate, counters and results, identical to running it alone. The engines must share the
 * proposal stream and use FP64 fields without statistics or traces; the batch captures their
 * parameters at creation, and an engine must not be run on its own while the batch runs. */
typedef struct wlx_batch wlx_batch;
int wlx_batch_create(wlx_engine *const *engines, uint64_t count, wlx_batch **out);
int wlx_batch_destroy(wlx_batch *batch);
/* Captures `epochs_per_graph` epochs in a HIP graph that run replays (0 = plain launches). */
int wlx_batch_set_graph(wlx_batch *batch, uint64_t epochs_per_graph);
/* Up to max_epochs epochs (engines that finish stop on the device); returns the epochs
 * enqueued and how many engines are done. */
int wlx_batch_run(wlx_batch *batch, uint64_t max_epochs, uint64_t *epochs,
                  uint64_t *done_engines);
/* Name of the batch's sampling kernel ("lookahead_fast_kernel<8>", "lds_sample_kernel<true>",
 * as wlx_engine_kernel names them). */
int wlx_batch_kernel(const wlx_batch *batch, char *name, size_t capacity);
/* The LDS walker's launches are grouped by the LDS one walker needs (classes of 16, 32, 64,
 * 96 and 160 KiB), one stream, table and graph per group, so that small systems pack many
 * walkers per CU; groups run their epochs concurrently. Engines whose fields and window bins
 * need more than `bytes` of LDS keep their bins in HBM in this batch (default: 160 KiB, only
 * what cannot fit). Results do not depend on it. */
int wlx_batch_set_lds_budget(wlx_batch *batch, uint64_t bytes);
/* Launch groups, engines with bins in HBM, and the largest per-walker LDS of any group. */
int wlx_batch_layout(const wlx_batch *batch, uint64_t *groups, uint64_t *bins_in_hbm,
                     uint64_t *max_lds_bytes);
/* Runs `epochs` epochs of the live batch under each LDS budget of {160, 64, 32} KiB (the
 * trajectories do not depend on it), keeps the fastest, and writes the seconds per budget to
 * seconds[3] (-1 = not tried). */
int wlx_batch_autotune(wlx_batch *batch, uint64_t epochs, uint64_t *chosen, double *seconds);
/* ---- warm-start network (no reference counterpart) ---------------------------------------
 * A 4-layer, 128-wide MLP, 6 features -> 128 -> 128 -> 128 -> 1 with SiLU hidden
 * activations, that maps (E/N, N, density, dilution, slab thickness, axis mode), scaled by the
 * caller, to log g / N. Matrix products take BF16 operands on the matrix cores with FP32
 * accumulation; biases, the output layer, the mean-squared-error loss and Adam (0.9, 0.999,
 * 1e-8) are FP32 on FP32 master weights. Parameters are one FP32 array of wlx_mlp_params()
 * numbers: W0 [128 x 16] (columns 6..15 unused), b0, W1 [128 x 128], b1, W2, b2, w3 [128],
 * b3. x is row-major [count x 6]. */
typedef struct wlx_mlp wlx_mlp;
int wlx_mlp_params(uint64_t *count);
int wlx_mlp_create(uint64_t seed, wlx_mlp **out);
int wlx_mlp_destroy(wlx_mlp *mlp);
int wlx_mlp_get_params(wlx_mlp *mlp, float *params);
int wlx_mlp_set_params(wlx_mlp *mlp, const float *params);
int wlx_mlp_predict(wlx_mlp *mlp, const float *x, uint64_t count, float *y);
/* Loss and gradient of the mean squared error over the samples, without an update. */
int wlx_mlp_gradient(wlx_mlp *mlp, const float *x, const float *y, uint64_t count, float *loss,
                     float *gradient);
/* `steps` Adam steps on minibatches of `batch` samples (epoch-wise shuffles from `seed`);
 * losses[k] = the minibatch loss before step k (NULL to skip). */
int wlx_mlp_train(wlx_mlp *mlp, const float *x, const float *y, uint64_t count, uint64_t steps,
                  uint64_t batch, float learning_rate, uint64_t seed, float *losses);
/* WangLandauWalker::histogram_statistics (wang_landau.cpp:344-367), computed on the device. */
int wlx_engine_histogram_statistics(wlx_engine *engine, uint64_t walker,
                                    wlx_histogram_stats *out);
/* ready_for_iteration (wang_landau.cpp:384-387); *ready = 0/1. */
int wlx_engine_ready(wlx_engine *engine, uint64_t walker, uint64_t *ready);
/* begin_next_iteration (wang_landau.cpp:389-409); WLX_LOGIC_ERROR if not ready. */
int wlx_engine_next_iteration(wlx_engine *engine, uint64_t walker);
/* Records the next `capacity` attempts of every walker (site, bin after the attempt, flags). */
int wlx_engine_trace(wlx_engine *engine, uint64_t capacity);
int wlx_engine_read_trace(wlx_engine *engine, uint64_t walker, uint64_t *count,
                          uint32_t *sites, uint32_t *bins, uint8_t *flags);
/* Times sampling-kernel launches of wlx_engine_run with HIP events on the engine stream and
 * resets the accumulators: enable = k > 0 times every k-th launch (1: every launch), 0 stops.
 * wlx_engine_kernel_time returns the mean duration of the timed launches in ms and their
 * number. No reference counterpart. */
int wlx_engine_set_timing(wlx_engine *engine, int enable);
int wlx_engine_kernel_time(wlx_engine *engine, double *sample_ms, uint64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* REWL_H */
