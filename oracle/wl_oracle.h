/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain-C restatement of the reference's REWL walker-step path
 * (Faikert/replica_exchange_wang_landau: include/wl/rng.hpp, src/physics.cpp,
 * src/wang_landau.cpp, src/parallel.cpp) used only as the checker of the HIP engine by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. Its handle API mirrors the
 * engine's C ABI (include/rewl.h) so one harness can drive both and compare arrays.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   - Xoshiro256** stream seeding: golden words written by the compiled reference
 *     (tests/golden/), plus the published splitmix64/xoshiro256** test values.
 *   - Philox4x64-10: numpy.random.Philox known answers.
 *   - J build, dense and CSR: reference closed forms (test_main.cpp:24-38), bitwise symmetry,
 *     dense == CSR at large cutoff, an independent numpy FP64 recompute, and bit equality
 *     with the compiled reference (oracle/_ref) where it builds.
 *   - walker, exchange and run_rewl schedule: bit equality of full runs (log_g, histograms,
 *     counters, spins, energies, RNG states) against golden runs of the compiled reference;
 *     exact-enumeration DOS; invariants (H gains one per attempt, E == FP64 recompute).
 *   - log() at the acceptance threshold uses the host libm, like the reference; the bit-level
 *     result there is "parity unpinned" (no reference test fixes it), see wlx ambiguous count.
 */
#ifndef WL_ORACLE_H
#define WL_ORACLE_H

#include "../include/rewl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system orc_system;
typedef struct orc_engine orc_engine;

/* rng.hpp:10-56 */
void orc_xoshiro_seed(uint64_t seed, uint64_t stream, uint64_t state[4]);
uint64_t orc_xoshiro_next(uint64_t state[4]);
uint64_t orc_xoshiro_bounded(uint64_t state[4], uint64_t bound);
double orc_xoshiro_uniform(uint64_t state[4]);
/* Philox4x64-10 block (Salmon et al., SC'11) */
void orc_philox4x64(const uint64_t counter[4], const uint64_t key[2], uint64_t out[4]);

/* physics.cpp:151-170 */
int orc_minimum_image(const double d[3], const double box[3], int periodic, double out[3]);
int orc_dipolar_pair(const double d[3], const double ai[3], const double aj[3], double scale,
                     double *out);

/* EnergyGrid (wang_landau.cpp:13-38) and partition_windows (wang_landau.cpp:40-61) */
int orc_grid_bins(double emin, double emax, double width, uint64_t *bins);
int64_t orc_grid_index(double emin, double emax, double width, double energy);
int orc_partition_windows(uint64_t bins, uint64_t count, double overlap, uint64_t *begin_end);

orc_system *orc_system_create(const double *positions, const double *axes, uint64_t n,
                              const double box[3], int periodic, double scale, double cutoff,
                              int precision);
void orc_system_destroy(orc_system *system);
int orc_system_info(const orc_system *system, uint64_t *n, uint64_t *backend, uint64_t *nnz);
int orc_system_read_dense(const orc_system *system, double *out);
int orc_system_read_csr(const orc_system *system, uint64_t *offsets, uint64_t *columns,
                        double *values);
int orc_system_fields(const orc_system *system, const int8_t *spins, uint64_t count,
                      double *fields, double *energies);

orc_engine *orc_engine_create(const orc_system *system, const wlx_config *config);
void orc_engine_destroy(orc_engine *engine);
int orc_engine_initialize(orc_engine *engine, const int8_t *initial_spins,
                          const uint8_t *supplied);
int orc_engine_sample(orc_engine *engine, uint64_t attempts);
int orc_engine_check(orc_engine *engine);
int orc_engine_exchange(orc_engine *engine, uint64_t epoch, uint64_t *attempted,
                        uint64_t *accepted);
int orc_engine_run(orc_engine *engine, uint64_t max_epochs, wlx_run_status *status);
int orc_engine_status(orc_engine *engine, wlx_run_status *status);
int orc_engine_read_sampling(orc_engine *engine, uint64_t walker, double *squared,
                             uint64_t *samples, uint64_t *round_trips, double *rep_energy,
                             uint8_t *rep_present, int8_t *rep_spins);
int orc_engine_warm_start(orc_engine *engine, const double *log_g, double factor);
int orc_engine_set_walker_cadence(orc_engine *engine, uint64_t walker,
                                  uint64_t last_flatness_check, double last_flatness);
int orc_engine_set_run_state(orc_engine *engine, uint64_t epoch, uint64_t exchange_attempted,
                             uint64_t exchange_accepted);
int orc_engine_walkers(const orc_engine *engine, uint64_t *walkers, uint64_t *spins,
                       uint64_t *bins);
int orc_engine_read_walker(orc_engine *engine, uint64_t walker, wlx_walker_info *info,
                           int8_t *spins, double *fields, double *log_g, uint64_t *histogram,
                           uint8_t *active, uint8_t *refinement_active);
int orc_engine_write_walker(orc_engine *engine, uint64_t walker, const wlx_walker_info *info,
                            const int8_t *spins, const double *fields, const double *log_g,
                            const uint64_t *histogram, const uint8_t *active);
int orc_engine_histogram_statistics(orc_engine *engine, uint64_t walker,
                                    wlx_histogram_stats *out);
int orc_engine_ready(orc_engine *engine, uint64_t walker, uint64_t *ready);
int orc_engine_next_iteration(orc_engine *engine, uint64_t walker);
int orc_engine_trace(orc_engine *engine, uint64_t capacity);
int orc_engine_read_trace(orc_engine *engine, uint64_t walker, uint64_t *count,
                          uint32_t *sites, uint32_t *bins, uint8_t *flags);

/* last error text and status of the calling thread */
size_t orc_last_error(char *buffer, size_t capacity);
int orc_last_status(void);

#ifdef __cplusplus
}
#endif

#endif
