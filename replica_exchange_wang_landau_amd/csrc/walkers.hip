// Kernel (b): device-resident Wang-Landau walkers with on-device replica exchange.
//
// Replaces WangLandauWalker (wang_landau.cpp:63-508), the OpenMP walker loop of run_rewl
// (parallel.cpp:376-421) and its exchanges (parallel.cpp:157-166). One wavefront owns one
// walker for a whole batch of attempts:
//   - the N local fields live in registers, site = k*64 + lane (k < npl), and a flip streams
//     row J[i, :] from HBM/L2 with coalesced 512-byte accesses (J is symmetric, so the row is
//     the reference's column walk of physics.cpp:190-195, element for element);
//   - spins are one bit per (lane, k) in a 64-bit register;
//   - log_g / H / active masks of the walker's window live in LDS during the batch;
//   - the RNG (Xoshiro256** or Philox4x64-10) and all decisions are wave-uniform.
// The arithmetic is the reference's, operation for operation (compiled with
// -ffp-contract=off), so in Xoshiro mode a walker's trajectory, spins, fields, energy and
// estimator are bit-identical to the reference walker's.
//
// Configurations (spins, fields, energy) sit in "slots"; a walker points at its slot, so an
// accepted replica exchange (swap_configuration, wang_landau.cpp:483-495) is an O(1) swap of
// two slot indices on the device.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <memory>
#include <type_traits>

namespace rewl {

struct WalkerState {
    uint64_t rng[4];
    double factor;
    double last_flatness;
    uint64_t attempted, accepted, forced, last_accepted;
    uint64_t last_check;
    uint64_t ambiguous;
    uint32_t stage, slot;
    uint32_t active_count, refine_count;
    uint32_t begin, end;
    uint32_t trace_count, escaped;
    // 1/t clock offset of a warm start (wlx_engine_warm_start): the 1/t factor is
    // active / (attempted + clock_offset); 0 gives the reference's active / attempted
    uint64_t clock_offset;
};
static_assert(sizeof(WalkerState) == 136, "walker state layout");

struct Globals {
    uint64_t epoch;
    uint64_t done;
    uint64_t exchange_attempted;
    uint64_t exchange_accepted;
    uint64_t exchange_ambiguous;
    uint64_t escaped;
    uint64_t pad[2];
};

struct InitReport {
    uint64_t status; // 0 ok, 1 target band not reached, 2 spins not +-1
    double min_reached, max_reached, closest, closest_distance, highest;
    uint64_t increases, restarts;
};

struct InitParams {
    uint64_t max_attempts;
    double target_fraction;
    double temperature_fraction;
    uint64_t stall_per_spin;
    double temperature_multiplier;
    double max_temperature_fraction;
};

struct SamplingState; // pilot statistics, defined with their kernels

struct Args {
    const double *J;
    // pilot statistics (collect_window_statistics); collect_offset: their LDS position
    double *sq_disp;            // walkers x ws_stride
    uint64_t *disp_samples;     // walkers x ws_stride
    SamplingState *sampling;    // walkers
    int8_t *rep_spins;          // walkers x 9 x np
    uint32_t collect, collect_offset;
    double rt_margin;
    // CSR couplings of a CSR system (physics.cpp:197-223): row offsets, columns, values
    const uint64_t *csr_offsets;
    const uint32_t *csr_columns;
    const double *csr_values;
    const float *csr_values32; // the same values as floats when the system's couplings are FP32
    const float *J32;   // FP32 couplings (FP32-field engines)
    float2 *fields32;   // FP32 fields as (h, c) pairs, value h - c (FP32-field engines)
    uint32_t n, np;
    Bounded bounded;
    Grid grid;
    WalkerState *state;
    int8_t *spins;
    double *fields;
    double *energy;
    double *logg;
    uint64_t *hist;
    uint8_t *active;
    uint8_t *refine;
    uint32_t ws_stride;
    uint32_t walkers, windows, per_window;
    uint32_t *trace_site;
    uint32_t *trace_bin;
    uint8_t *trace_flag;
    uint64_t trace_cap;
    Globals *glob;
    uint64_t seed;
    uint64_t exchange_interval, max_attempts, check_interval, force_after, min_visits;
    double final_factor, flatness;
    uint32_t inverse_time, nalivaiko, philox, unlimited;
    // window bins too wide for LDS: the walker works on its bins in HBM (the LDS walker, the
    // initialisation and the harness operations), with the same arithmetic
    uint32_t bins_hbm, pad0;
};

// one wave of a sampling launch: the engine (index into the launch's Args table) and walker.
// An engine alone launches with the identity list {0, w}; a batch with its sorted list.
struct BatchSlot {
    uint32_t engine, walker;
};

enum : uint32_t { kGate = 1, kCheck = 2, kTrace = 4 };

// ---- wave helpers -------------------------------------------------------------------------
// lane within the wave (walker kernels run one wave per workgroup, the staged walker two)
__device__ inline uint32_t lane_id() { return threadIdx.x & (kWave - 1); }

__device__ inline double readlane_f64(double v, uint32_t lane)
{
    const uint64_t bits = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(bits), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(bits >> 32), lane);
    return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

__device__ inline uint64_t readlane_u64(uint64_t v, uint32_t lane)
{
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
    return (uint64_t(hi) << 32) | lo;
}

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int mask)
{
    const uint32_t lo = __shfl_xor(uint32_t(v), mask);
    const uint32_t hi = __shfl_xor(uint32_t(v >> 32), mask);
    return (uint64_t(hi) << 32) | lo;
}

// h[k] for a wave-uniform k, as a binary tree of selects on the bits of k (an equality
// chain is turned into a stack lookup table by the compiler, which costs scratch traffic).
template <int NPL> __device__ inline double select_slot(const double (&h)[NPL], uint32_t k)
{
    double level[NPL];
#pragma unroll
    for (int j = 0; j < NPL; ++j) level[j] = h[j];
#pragma unroll
    for (int width = NPL, bit = 0; width > 1; width >>= 1, ++bit) {
        const bool upper = (k >> bit) & 1;
#pragma unroll
        for (int j = 0; j < width / 2; ++j) level[j] = upper ? level[2 * j + 1] : level[2 * j];
    }
    return level[0];
}

// ---- window bins in LDS -------------------------------------------------------------------
struct Bins {
    double *lg;
    uint64_t *H;
    uint8_t *act;
    uint8_t *ref;
};

__host__ __device__ inline size_t bins_lds_bytes(uint32_t stride) { return size_t(stride) * 18; }

__device__ inline Bins bins_view(unsigned char *smem, uint32_t stride)
{
    return {reinterpret_cast<double *>(smem), reinterpret_cast<uint64_t *>(smem + 8 * size_t(stride)),
            smem + 16 * size_t(stride), smem + 17 * size_t(stride)};
}

// walker w's bins in place in HBM (Args::bins_hbm). The lanes of one wave see each other's
// stores to these addresses in program order (one vector L1 per CU), as with LDS.
__device__ inline Bins bins_hbm_view(const Args &a, uint32_t w)
{
    const size_t base = size_t(w) * a.ws_stride;
    return {a.logg + base, a.hist + base, a.active + base, a.refine + base};
}

__device__ inline void load_bins(const Args &a, uint32_t w, uint32_t ws, Bins v)
{
    const size_t base = size_t(w) * a.ws_stride;
    for (uint32_t b = lane_id(); b < ws; b += kWave) {
        v.lg[b] = a.logg[base + b];
        v.H[b] = a.hist[base + b];
        v.act[b] = a.active[base + b];
        v.ref[b] = a.refine[base + b];
    }
}

__device__ inline void store_bins(const Args &a, uint32_t w, uint32_t ws, Bins v)
{
    const size_t base = size_t(w) * a.ws_stride;
    for (uint32_t b = lane_id(); b < ws; b += kWave) {
        a.logg[base + b] = v.lg[b];
        a.hist[base + b] = v.H[b];
        a.active[base + b] = v.act[b];
        a.refine[base + b] = v.ref[b];
    }
}

// ---- walker primitives (uniform across the wave) --------------------------------------------

// update_current_bin (wang_landau.cpp:290-303); `b` is window-relative
__device__ inline void update_current_bin(const Args &a, WalkerState &st, Bins v, uint32_t b)
{
    const uint64_t count = v.H[b];
    const uint8_t seen = v.act[b];
    if (seen == 0) ++st.active_count;
    uint8_t refined = 1;
    if (a.nalivaiko) {
        refined = v.ref[b];
        if (refined == 0) ++st.refine_count;
    }
    double lg = v.lg[b];
    if (st.stage != WLX_STAGE_FROZEN) lg += st.factor;
    if (lane_id() == 0) {
        v.H[b] = count + 1;
        v.act[b] = 1;
        if (a.nalivaiko) v.ref[b] = 1;
        v.lg[b] = lg;
    }
}

// ---- pilot statistics (collect_window_statistics, wang_landau.cpp:104-118, 272-282, 317-342)
constexpr int kRepresentatives = 9;

struct SamplingState {
    uint64_t round_trips;
    uint32_t rt_state, present; // present: bit r set once representative r is recorded
    double target[kRepresentatives], dist[kRepresentatives], energy[kRepresentatives];
};

struct SamplingView { // the walker's statistics staged in LDS
    double *sq;
    uint64_t *cnt;
    SamplingState *st;
};

__device__ inline SamplingView sampling_view(const Args &a, unsigned char *smem)
{
    double *sq = reinterpret_cast<double *>(smem + a.collect_offset);
    uint64_t *cnt = reinterpret_cast<uint64_t *>(sq + a.ws_stride);
    return {sq, cnt, reinterpret_cast<SamplingState *>(cnt + a.ws_stride)};
}

__device__ inline void load_sampling(const Args &a, uint32_t w, uint32_t ws, SamplingView c)
{
    const size_t base = size_t(w) * a.ws_stride;
    for (uint32_t b = lane_id(); b < ws; b += kWave) {
        c.sq[b] = a.sq_disp[base + b];
        c.cnt[b] = a.disp_samples[base + b];
    }
    if (lane_id() == 0) *c.st = a.sampling[w];
}

__device__ inline void store_sampling(const Args &a, uint32_t w, uint32_t ws, SamplingView c)
{
    const size_t base = size_t(w) * a.ws_stride;
    for (uint32_t b = lane_id(); b < ws; b += kWave) {
        a.sq_disp[base + b] = c.sq[b];
        a.disp_samples[base + b] = c.cnt[b];
    }
    if (lane_id() == 0) a.sampling[w] = *c.st;
}

// squared accepted displacement and one sample at the bin the attempt started from
__device__ inline void record_displacement(SamplingView c, uint32_t old_rel, double displacement)
{
    if (lane_id() == 0) {
        c.sq[old_rel] += displacement * displacement;
        c.cnt[old_rel] += 1;
    }
}

// update_round_trip_state (wang_landau.cpp:317-330): state 0 = not yet at the lower margin,
// 1 = last at the lower margin, 2 = reached the upper margin since; a return to the lower
// margin from 2 completes a round trip
__device__ inline void round_trip_step(const Args &a, uint32_t begin, uint32_t end, double energy,
                                       uint32_t &state, uint64_t &trips)
{
    const double lower = a.grid.minimum + double(begin) * a.grid.width;
    const double upper = a.grid.minimum + double(end) * a.grid.width;
    const double margin = a.rt_margin * (upper - lower);
    if (energy <= lower + margin) {
        if (state == 2) {
            ++trips;
            state = 1;
        } else if (state == 0) {
            state = 1;
        }
    } else if (energy >= upper - margin && state == 1) {
        state = 2;
    }
}

// update_round_trip_state and update_representatives (wang_landau.cpp:332-342) of a walker
// wave after an accepted flip or its placement
template <int NPL>
__device__ inline void round_trip_and_representatives(const Args &a, SamplingView c, uint32_t w,
                                                      uint32_t begin, uint32_t end,
                                                      double energy, uint64_t neg)
{
    const uint32_t lane = lane_id();
    uint32_t state = c.st->rt_state;
    uint64_t trips = c.st->round_trips;
    round_trip_step(a, begin, end, energy, state, trips);
    const uint32_t present = c.st->present;
    uint32_t improved = 0;
    for (int r = 0; r < kRepresentatives; ++r) {
        const double d = fabs(energy - c.st->target[r]);
        if (d < c.st->dist[r]) {
            improved |= 1u << r;
            int8_t *out = a.rep_spins + (size_t(w) * kRepresentatives + r) * a.np;
#pragma unroll
            for (int k = 0; k < NPL; ++k)
                out[k * kWave + lane] = ((neg >> k) & 1) ? int8_t(-1) : int8_t(1);
        }
    }
    if (lane == 0) {
        c.st->rt_state = state;
        c.st->round_trips = trips;
        for (int r = 0; r < kRepresentatives; ++r)
            if ((improved >> r) & 1) {
                c.st->dist[r] = fabs(energy - c.st->target[r]);
                c.st->energy[r] = energy;
            }
        c.st->present = present | improved;
    }
}

// the same two updates by one thread for a walker that received configuration `slot` in an
// accepted replica exchange (swap_configuration, wang_landau.cpp:483-495)
__device__ inline void exchange_statistics(const Args &a, uint32_t w, const WalkerState &s,
                                           uint32_t slot, double energy)
{
    SamplingState &z = a.sampling[w];
    uint32_t state = z.rt_state;
    uint64_t trips = z.round_trips;
    round_trip_step(a, s.begin, s.end, energy, state, trips);
    z.rt_state = state;
    z.round_trips = trips;
    for (int r = 0; r < kRepresentatives; ++r) {
        const double d = fabs(energy - z.target[r]);
        if (d < z.dist[r]) {
            z.dist[r] = d;
            z.energy[r] = energy;
            z.present |= 1u << r;
            const int8_t *from = a.spins + size_t(slot) * a.np;
            int8_t *to = a.rep_spins + (size_t(w) * kRepresentatives + r) * a.np;
            for (uint32_t i = 0; i < a.np; ++i) to[i] = from[i];
        }
    }
}

struct Stats {
    uint64_t active, covered, minimum;
    double mean, min_over_mean, coverage;
};

// histogram_statistics (wang_landau.cpp:344-367): lanes stride the window, then reduce
__device__ inline Stats statistics(const Args &a, const WalkerState &st, Bins v, uint32_t ws)
{
    const uint8_t *mask = a.nalivaiko ? v.ref : v.act;
    uint64_t covered = 0, total = 0, minimum = ~uint64_t(0);
    for (uint32_t b = lane_id(); b < ws; b += kWave) {
        if (!mask[b]) continue;
        const uint64_t h = v.H[b];
        covered += h != 0;
        minimum = h < minimum ? h : minimum;
        total += h;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        covered += shfl_xor_u64(covered, off);
        total += shfl_xor_u64(total, off);
        const uint64_t other = shfl_xor_u64(minimum, off);
        minimum = other < minimum ? other : minimum;
    }
    Stats s{};
    s.active = a.nalivaiko ? st.refine_count : st.active_count;
    s.covered = covered;
    if (s.active == 0) return s;
    s.minimum = minimum;
    s.mean = extended_quotient(total, s.active);
    s.min_over_mean = s.mean > 0.0 ? double(minimum) / s.mean : 0.0;
    s.coverage = double(covered) / double(s.active);
    return s;
}

// flat / covered / ready_for_iteration (wang_landau.cpp:369-387)
__device__ inline bool ready_for_iteration(const Args &a, const WalkerState &st, const Stats &s)
{
    if (st.stage != WLX_STAGE_WANG_LANDAU) return false;
    if (a.inverse_time) return s.active != 0 && s.covered == s.active;
    return s.active != 0 && s.minimum >= a.min_visits && s.min_over_mean >= a.flatness;
}

// begin_next_iteration + freeze_if_finished (wang_landau.cpp:389-413), readiness checked
__device__ inline void next_iteration(const Args &a, WalkerState &st, Bins v, uint32_t ws)
{
    st.factor *= 0.5;
    for (uint32_t b = lane_id(); b < ws; b += kWave) v.H[b] = 0;
    if (a.nalivaiko) {
        for (uint32_t b = lane_id(); b < ws; b += kWave) v.ref[b] = 0;
        st.refine_count = 0;
    }
    if (a.inverse_time) {
        const uint64_t clock = st.attempted + st.clock_offset;
        const double inverse_time = clock == 0 ? 1.0 : double(st.active_count) / double(clock);
        if (st.factor <= inverse_time) {
            st.stage = WLX_STAGE_INVERSE_TIME;
            st.factor = inverse_time;
        }
    }
    if (st.factor <= a.final_factor) st.stage = WLX_STAGE_FROZEN;
}

__device__ inline bool near_threshold(double x, double threshold)
{
    return fabs(x - threshold) <= 8.9e-16 * fmax(fabs(x), fabs(threshold));
}

// local_fields (physics.cpp:238-251) for site k*64+lane: sum over j ascending of J[j][i]*s_j
// (J is bitwise symmetric); the products are exact, so the sum is the reference's.
template <int NPL>
__device__ inline void compute_fields(const Args &a, uint64_t neg, double (&h)[NPL])
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int k = 0; k < NPL; ++k) h[k] = 0.0;
    for (uint32_t j = 0; j < a.n; ++j) {
        const uint64_t bits = readlane_u64(neg, j & (kWave - 1));
        const double s = ((bits >> (j >> 6)) & 1) ? -1.0 : 1.0;
        const double *row = a.J + size_t(j) * a.np;
#pragma unroll
        for (int k = 0; k < NPL; ++k) h[k] += row[k * kWave + lane] * s;
    }
}

// total_energy (physics.cpp:253-259): sequential sum over i ascending, via LDS scratch
template <int NPL>
__device__ inline double total_energy(const Args &a, uint64_t neg, const double (&h)[NPL],
                                      double *scratch)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        const double s = ((neg >> k) & 1) ? -1.0 : 1.0;
        scratch[k * kWave + lane] = 0.5 * s * h[k];
    }
    __syncthreads();
    double e = 0.0;
    if (lane == 0)
        for (uint32_t i = 0; i < a.n; ++i) e += scratch[i];
    e = readlane_f64(e, 0);
    __syncthreads();
    return e;
}

template <int NPL> __device__ inline void load_config(const Args &a, uint32_t slot, uint64_t &neg,
                                                      double (&h)[NPL])
{
    const uint32_t lane = lane_id();
    const size_t base = size_t(slot) * a.np;
    neg = 0;
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        h[k] = a.fields[base + k * kWave + lane];
        if (a.spins[base + k * kWave + lane] < 0) neg |= uint64_t(1) << k;
    }
}

template <int NPL> __device__ inline void store_config(const Args &a, uint32_t slot, uint64_t neg,
                                                       const double (&h)[NPL])
{
    const uint32_t lane = lane_id();
    const size_t base = size_t(slot) * a.np;
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        a.fields[base + k * kWave + lane] = h[k];
        a.spins[base + k * kWave + lane] = ((neg >> k) & 1) ? int8_t(-1) : int8_t(1);
    }
}

// the flip of an accepted proposal: h += (-2 s_old) J[site, :] (physics.cpp:190-195), spin bit
template <int NPL>
__device__ inline void apply_flip(const Args &a, uint32_t site, double sigma, uint64_t &neg,
                                  double (&h)[NPL])
{
    const uint32_t lane = lane_id();
    const double m = -2.0 * sigma;
    const double *row = a.J + size_t(site) * a.np;
    constexpr int kChunk = NPL < 16 ? NPL : 16; // loads in flight per lane
#pragma unroll
    for (int c = 0; c < NPL; c += kChunk) {
        double r[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) r[k] = row[(c + k) * kWave + lane];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) h[c + k] += m * r[k];
    }
    if (lane == (site & (kWave - 1))) neg ^= uint64_t(1) << (site >> 6);
}

// attempts of this launch: `request`, or with kGate run_rewl's min(exchange_interval,
// max - attempted) (parallel.cpp:383-389)
__device__ inline uint64_t batch_count(const Args &a, const WalkerState &st, uint64_t request,
                                       uint32_t flags)
{
    if (!(flags & kGate)) return request;
    uint64_t count = a.exchange_interval;
    if (!a.unlimited) {
        const uint64_t remaining = st.attempted < a.max_attempts ? a.max_attempts - st.attempted : 0;
        count = remaining < count ? remaining : count;
    }
    return count;
}

// flatness cadence after the batch (parallel.cpp:395-407), then the walker's bins, energy and
// state back to HBM (the configuration's spins and fields are stored by the caller)
__device__ inline void finish_walker_state(const Args &a, uint32_t w, WalkerState &st, Bins v,
                                           uint32_t ws, double energy, uint32_t flags)
{
    if ((flags & kCheck) && !st.escaped && st.stage == WLX_STAGE_WANG_LANDAU) {
        __syncthreads();
        const uint64_t interval = a.check_interval > 1 ? a.check_interval : 1;
        const bool at_limit = !a.unlimited && st.attempted >= a.max_attempts;
        if (st.attempted - st.last_check >= interval || at_limit) {
            st.last_check = st.attempted;
            const Stats s = statistics(a, st, v, ws);
            st.last_flatness = a.inverse_time ? s.coverage : s.min_over_mean;
            if (ready_for_iteration(a, st, s)) next_iteration(a, st, v, ws);
        }
    }
    __syncthreads();
    if (!a.bins_hbm) store_bins(a, w, ws, v);
    if (lane_id() == 0) {
        a.energy[st.slot] = energy;
        a.state[w] = st;
        if (st.escaped) atomicOr(reinterpret_cast<unsigned long long *>(&a.glob->escaped), 1ull);
    }
}

template <int NPL>
__device__ inline void finish_walker(const Args &a, uint32_t w, WalkerState &st, Bins v,
                                     uint32_t ws, uint64_t neg, const double (&h)[NPL],
                                     double energy, uint32_t flags)
{
    store_config<NPL>(a, st.slot, neg, h);
    finish_walker_state(a, w, st, v, ws, energy, flags);
}

// ---------------------------------------------------------------------------------------------
// sample: `count` attempt_flip calls (wang_landau.cpp:244-284) per walker; with kGate the
// batch is run_rewl's min(exchange_interval, max - attempted) and a finished run is skipped;
// with kCheck the flatness cadence of parallel.cpp:395-407 follows the batch.
// ---------------------------------------------------------------------------------------------
template <int NPL, bool PHILOX>
__global__ __launch_bounds__(kWave) void sample_kernel(const Args *table,
                                                  const BatchSlot *slots,
                                                  uint64_t request, uint32_t flags)
{
    const BatchSlot slot = slots[blockIdx.x];
    const Args a = table[slot.engine];
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t w = slot.walker, lane = lane_id();
    if ((flags & kGate) && a.glob->done) return;
    WalkerState st = a.state[w];
    if (st.escaped) return;
    const uint32_t ws = st.end - st.begin;
    Bins v = bins_view(smem, a.ws_stride);
    load_bins(a, w, ws, v);
    uint64_t neg;
    double h[NPL];
    load_config<NPL>(a, st.slot, neg, h);
    double energy = a.energy[st.slot];
    __syncthreads();

    const uint64_t count = batch_count(a, st, request, flags);
    Xoshiro rng{st.rng[0], st.rng[1], st.rng[2], st.rng[3]};
    const uint64_t trace_base = uint64_t(w) * a.trace_cap;
    int64_t cur = grid_index(a.grid, energy);
    Philox4 block{};
    const SamplingView col = sampling_view(a, smem);
    if (a.collect) {
        load_sampling(a, w, ws, col);
        __syncthreads();
    }
    for (uint64_t t = 0; t < count; ++t) {
        uint64_t site;
        if (PHILOX) {
            if (t == 0 || (st.attempted & 1) == 0)
                block = philox4x64(st.attempted >> 1, kPhiloxWalkTag, 0, 0, a.seed,
                                   uint64_t(w));
            site = __umul64hi((st.attempted & 1) ? block.v[2] : block.v[0], uint64_t(a.n));
        } else {
            site = draw_bounded(rng, a.bounded);
        }
        if (cur < int64_t(st.begin) || cur >= int64_t(st.end)) { // wang_landau.cpp:247-248
            st.escaped = 1;
            break;
        }
        const uint32_t s = uint32_t(site);
        const double hs = readlane_f64(select_slot<NPL>(h, s >> 6), s & (kWave - 1));
        const uint64_t bits = readlane_u64(neg, s & (kWave - 1));
        const double sigma = ((bits >> (s >> 6)) & 1) ? -1.0 : 1.0;
        const double delta = -2.0 * sigma * hs; // flip_delta, physics.cpp:261-266
        const double proposed = energy + delta;
        const int64_t next = grid_index(a.grid, proposed);
        bool accepted = false, forced = false, drew = false, inside = false;
        if (next >= int64_t(st.begin) && next < int64_t(st.end)) {
            inside = true;
            const double log_ratio = v.lg[cur - st.begin] - v.lg[next - st.begin];
            const bool force_due =
                a.force_after != 0 && st.attempted - st.last_accepted >= a.force_after;
            if (log_ratio >= 0.0)
                accepted = true;
            else if (force_due)
                accepted = forced = true;
            else {
                const double u =
                    PHILOX ? double(((st.attempted & 1) ? block.v[3] : block.v[1]) >> 11) *
                                 0x1.0p-53
                           : rng.uniform();
                drew = true;
                const double lu = log(fmax(u, 0x1.0p-53));
                accepted = lu < log_ratio;
                if (near_threshold(lu, log_ratio)) ++st.ambiguous;
            }
        }
        ++st.attempted;
        const uint32_t old_rel = uint32_t(cur - st.begin);
        if (accepted) {
            apply_flip<NPL>(a, s, sigma, neg, h);
            energy = proposed;
            ++st.accepted;
            if (forced) ++st.forced;
            st.last_accepted = st.attempted;
            cur = next;
        }
        if (a.collect) record_displacement(col, old_rel, accepted ? delta : 0.0);
        if (st.stage == WLX_STAGE_INVERSE_TIME) { // update_inverse_time_factor, 309-315
            if (st.active_count != 0 && st.attempted != 0)
                st.factor = double(st.active_count) / double(st.attempted + st.clock_offset);
            if (st.factor <= a.final_factor) st.stage = WLX_STAGE_FROZEN;
        }
        update_current_bin(a, st, v, uint32_t(cur - st.begin));
        if (a.collect && accepted)
            round_trip_and_representatives<NPL>(a, col, w, st.begin, st.end, energy, neg);
        if ((flags & kTrace) && st.trace_count < a.trace_cap) {
            if (lane == 0) {
                const uint64_t k = trace_base + st.trace_count;
                a.trace_site[k] = s;
                a.trace_bin[k] = uint32_t(cur);
                a.trace_flag[k] = uint8_t((accepted ? WLX_TRACE_ACCEPTED : 0) |
                                          (forced ? WLX_TRACE_FORCED : 0) |
                                          (drew ? WLX_TRACE_UNIFORM : 0) |
                                          (inside ? WLX_TRACE_IN_WINDOW : 0));
            }
            ++st.trace_count;
        }
    }
    st.rng[0] = rng.s0;
    st.rng[1] = rng.s1;
    st.rng[2] = rng.s2;
    st.rng[3] = rng.s3;
    if (a.collect) {
        __syncthreads();
        store_sampling(a, w, ws, col);
    }
    finish_walker<NPL>(a, w, st, v, ws, neg, h, energy, flags);
}

// ---------------------------------------------------------------------------------------------
// LDS walker (WLX_KERNEL_LDS): the same attempt sequence as sample_kernel with the
// configuration's fields and spins staged in LDS instead of registers, so one kernel serves
// every N (no per-size register layout) and a batch of uneven systems shares a launch. An
// accepted flip on a CSR system scatters its CSR row, h[col] += m val (physics.cpp:231-236),
// one column per lane: the columns of a row are distinct, so the lanes never collide. On a
// dense system it streams the row, h[j] += m J[s][j] (physics.cpp:190-195). fma(m, J, h) is the
// reference's rounding because m J is exact. LDS: bins | fields (np doubles) | spins (np bytes).
// ---------------------------------------------------------------------------------------------
// fields and spins of n sites, rounded up to 16 sites
__host__ __device__ inline uint32_t lds_sites(uint32_t n) { return (n + 15u) & ~15u; }

__host__ __device__ inline size_t lds_walker_bytes(uint32_t ws_stride, uint32_t n,
                                                   bool bins_hbm = false)
{
    return (bins_hbm ? 0 : bins_lds_bytes(ws_stride)) + size_t(lds_sites(n)) * 9;
}

// up to 64 (column, value) entries of a CSR row, one per lane, loaded ahead of their use
struct RowAhead {
    uint64_t begin, end;
    uint32_t column;
    double value;
};

// CSR value k: 8 bytes per entry, or 4 when the system's couplings are FP32 (the float is the
// rounded coupling exactly, so the field update is the same FP64 fma)
__device__ inline double csr_value(const Args &a, uint64_t k)
{
    return a.csr_values32 ? double(a.csr_values32[k]) : a.csr_values[k];
}

__device__ inline RowAhead load_row_ahead(const Args &a, uint32_t site)
{
    RowAhead r;
    r.begin = a.csr_offsets[site];
    r.end = a.csr_offsets[site + 1];
    const uint64_t k = r.begin + lane_id();
    r.column = k < r.end ? a.csr_columns[k] : 0u;
    r.value = k < r.end ? csr_value(a, k) : 0.0;
    return r;
}

template <bool PHILOX, bool HBM_BINS>
__device__ inline void lds_walk(const Args &a, uint32_t w, uint64_t request, uint32_t flags,
                                unsigned char *smem)
{
    const uint32_t lane = lane_id();
    if ((flags & kGate) && a.glob->done) return;
    WalkerState st = a.state[w];
    if (st.escaped) return;
    const uint32_t ws = st.end - st.begin;
    // window bins in LDS, or in place in HBM for windows too wide for LDS
    Bins v = HBM_BINS ? bins_hbm_view(a, w) : bins_view(smem, a.ws_stride);
    double *hf = reinterpret_cast<double *>(smem + (HBM_BINS ? 0 : bins_lds_bytes(a.ws_stride)));
    int8_t *sp = reinterpret_cast<int8_t *>(hf + lds_sites(a.n));
    if (!HBM_BINS) load_bins(a, w, ws, v);
    const size_t base = size_t(st.slot) * a.np;
    for (uint32_t i = lane; i < a.n; i += kWave) {
        hf[i] = a.fields[base + i];
        sp[i] = a.spins[base + i];
    }
    double energy = a.energy[st.slot];
    __syncthreads();

    const uint64_t count = batch_count(a, st, request, flags);
    Xoshiro rng{st.rng[0], st.rng[1], st.rng[2], st.rng[3]};
    const uint64_t trace_base = uint64_t(w) * a.trace_cap;
    int64_t cur = grid_index(a.grid, energy);
    // Philox: lane j draws the proposal site and log(u) of attempt first + j for 64 attempts
    // at a time, so the proposal of the next attempt is known and its CSR row is loaded one
    // attempt ahead, behind the current decision
    const bool ahead = PHILOX && a.csr_offsets != nullptr;
    uint64_t first = st.attempted; // attempt of lane 0 in p_site / p_lu
    uint32_t p_site = 0;
    double p_lu = 0.0;
    auto draw_block = [&](uint64_t from) {
        const uint64_t t = from + lane;
        const Philox4 blk = philox4x64(t >> 1, kPhiloxWalkTag, 0, 0, a.seed, uint64_t(w));
        const bool odd = t & 1;
        p_site = uint32_t(__umul64hi(odd ? blk.v[2] : blk.v[0], uint64_t(a.n)));
        const double u = double((odd ? blk.v[3] : blk.v[1]) >> 11) * 0x1.0p-53;
        p_lu = log(fmax(u, 0x1.0p-53));
        first = from;
    };
    uint32_t next_site = 0;
    RowAhead row{};
    if (PHILOX && count > 0) {
        draw_block(st.attempted);
        next_site = __builtin_amdgcn_readlane(p_site, 0);
        if (ahead) row = load_row_ahead(a, next_site);
    }
    for (uint64_t t = 0; t < count; ++t) {
        uint64_t site;
        double lu_philox = 0.0;
        RowAhead this_row{};
        if (PHILOX) {
            site = next_site;
            lu_philox = readlane_f64(p_lu, uint32_t(st.attempted - first));
            this_row = row;
            if (t + 1 < count) { // the next attempt's proposal and, on CSR systems, its row
                const uint64_t na = st.attempted + 1;
                if (na - first == kWave) draw_block(na);
                next_site = __builtin_amdgcn_readlane(p_site, uint32_t(na - first));
                if (ahead) row = load_row_ahead(a, next_site);
            }
        } else {
            site = draw_bounded(rng, a.bounded);
        }
        if (cur < int64_t(st.begin) || cur >= int64_t(st.end)) { // wang_landau.cpp:247-248
            st.escaped = 1;
            break;
        }
        const uint32_t s = uint32_t(site);
        const double hs = hf[s];
        const double sigma = sp[s] < 0 ? -1.0 : 1.0;
        const double delta = -2.0 * sigma * hs; // flip_delta, physics.cpp:261-266
        const double proposed = energy + delta;
        const int64_t next = grid_index(a.grid, proposed);
        bool accepted = false, forced = false, drew = false, inside = false;
        if (next >= int64_t(st.begin) && next < int64_t(st.end)) {
            inside = true;
            const double log_ratio = v.lg[cur - st.begin] - v.lg[next - st.begin];
            const bool force_due =
                a.force_after != 0 && st.attempted - st.last_accepted >= a.force_after;
            if (log_ratio >= 0.0)
                accepted = true;
            else if (force_due)
                accepted = forced = true;
            else {
                drew = true;
                const double lu = PHILOX ? lu_philox : log(fmax(rng.uniform(), 0x1.0p-53));
                accepted = lu < log_ratio;
                if (near_threshold(lu, log_ratio)) ++st.ambiguous;
            }
        }
        ++st.attempted;
        if (accepted) {
            const double m = -2.0 * sigma;
            if (ahead) {
                if (this_row.begin + lane < this_row.end)
                    hf[this_row.column] = fma(m, this_row.value, hf[this_row.column]);
                for (uint64_t k = this_row.begin + kWave + lane; k < this_row.end; k += kWave) {
                    const uint32_t c = a.csr_columns[k];
                    hf[c] = fma(m, csr_value(a, k), hf[c]);
                }
            } else if (a.csr_offsets) {
                const uint64_t row_end = a.csr_offsets[s + 1];
                for (uint64_t k = a.csr_offsets[s] + lane; k < row_end; k += kWave) {
                    const uint32_t c = a.csr_columns[k];
                    hf[c] = fma(m, csr_value(a, k), hf[c]);
                }
            } else {
                const double *jrow = a.J + size_t(s) * a.np;
                for (uint32_t j = lane; j < a.n; j += kWave) hf[j] = fma(m, jrow[j], hf[j]);
            }
            if (lane == 0) sp[s] = sigma < 0.0 ? int8_t(1) : int8_t(-1);
            __syncthreads(); // the scattered fields before the next attempt reads one
            energy = proposed;
            ++st.accepted;
            if (forced) ++st.forced;
            st.last_accepted = st.attempted;
            cur = next;
        }
        if (st.stage == WLX_STAGE_INVERSE_TIME) { // update_inverse_time_factor, 309-315
            if (st.active_count != 0 && st.attempted != 0)
                st.factor = double(st.active_count) / double(st.attempted + st.clock_offset);
            if (st.factor <= a.final_factor) st.stage = WLX_STAGE_FROZEN;
        }
        update_current_bin(a, st, v, uint32_t(cur - st.begin));
        if ((flags & kTrace) && st.trace_count < a.trace_cap) {
            if (lane == 0) {
                const uint64_t k = trace_base + st.trace_count;
                a.trace_site[k] = s;
                a.trace_bin[k] = uint32_t(cur);
                a.trace_flag[k] = uint8_t((accepted ? WLX_TRACE_ACCEPTED : 0) |
                                          (forced ? WLX_TRACE_FORCED : 0) |
                                          (drew ? WLX_TRACE_UNIFORM : 0) |
                                          (inside ? WLX_TRACE_IN_WINDOW : 0));
            }
            ++st.trace_count;
        }
    }
    st.rng[0] = rng.s0;
    st.rng[1] = rng.s1;
    st.rng[2] = rng.s2;
    st.rng[3] = rng.s3;
    __syncthreads();
    for (uint32_t i = lane; i < a.n; i += kWave) {
        a.fields[base + i] = hf[i];
        a.spins[base + i] = sp[i];
    }
    finish_walker_state(a, w, st, v, ws, energy, flags);
}

template <bool PHILOX>
__global__ __launch_bounds__(kWave) void lds_sample_kernel(const Args *table,
                                                  const BatchSlot *slots,
                                                  uint64_t request, uint32_t flags)
{
    const BatchSlot slot = slots[blockIdx.x];
    const Args a = table[slot.engine];
    extern __shared__ __align__(16) unsigned char smem[];
    if (a.bins_hbm)
        lds_walk<PHILOX, true>(a, slot.walker, request, flags, smem);
    else
        lds_walk<PHILOX, false>(a, slot.walker, request, flags, smem);
}

// ---------------------------------------------------------------------------------------------
// lookahead sampling (WLX_RNG_PHILOX): counter-based proposals make the sites and uniforms of
// future attempts known in advance, so a wave
//   1. draws the proposals of 64 attempts at once (one Philox block and one log per lane),
//   2. for the next KB attempts, gathers the field and spin at each proposal site and the
//      couplings J[s_a][s_b] among those sites in one memory round trip,
//   3. resolves the KB accept/reject decisions with a short serial chain: the field seen by
//      attempt b is h_{s_b} plus m_a J[s_a][s_b] for every earlier accepted a, added in
//      acceptance order, exactly the sequence of additions the element receives in the
//      one-flip-at-a-time walker,
//   4. streams the accepted rows J[s_a, :] into the register-resident fields together.
// The result is bit-identical to the sequential walker on the same Philox stream.
// ---------------------------------------------------------------------------------------------
template <int NPL> __device__ inline double gather_field(const double (&h)[NPL], uint32_t site)
{
    double hv[NPL];
#pragma unroll
    for (int k = 0; k < NPL; ++k) hv[k] = __shfl(h[k], int(site & (kWave - 1)));
    return select_slot<NPL>(hv, site >> 6);
}

// EnergyGrid::index (wang_landau.cpp:27-34) for a proposed energy: `in` is false outside
// [minimum, maximum] (and for NaN); otherwise `idx` is the bin, with E == maximum folded into
// the last bin. Same result as grid_index.
__device__ inline uint32_t grid_bin(const Grid &g, double e, bool &in)
{
    in = (e >= g.minimum) && (e <= g.maximum); // min/max are finite: excludes +-inf, NaN
    const double f = bin_floor(g, e);
    uint32_t idx = in ? uint32_t(f) : 0u;
    if (idx == uint32_t(g.count) && e == g.maximum) idx = uint32_t(g.count) - 1;
    in = in && idx < uint32_t(g.count);
    return idx;
}

__device__ inline double hi_double(uint32_t hi)
{
    return __hiloint2double(int(hi), 0);
}

// A value every active lane holds, marked wave-uniform for the compiler: what is computed from
// it stays uniform, so its branches and selects are scalar instead of exec-mask sequences.
__device__ inline double uniform_f64(double x)
{
    const uint64_t bits = __double_as_longlong(x);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(bits));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(bits >> 32));
    return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}
__device__ inline uint32_t uniform_u32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// h += m_b J[s_b, :] for the proposals of the batch, in order. m_hi holds the high word of
// m_b = -2 sigma_b (+-2.0; the low word is zero) in the lane of every accepted proposal and 0
// elsewhere. m_b J is exact (a power-of-two scaling), so fma(m_b, J, h) rounds exactly like
// the reference's h + (m_b J) (physics.cpp:190-195); m = 0 leaves a field unchanged (fields
// are never -0.0).
template <int NPL, int KB>
__device__ inline void apply_rows(const Args &a, uint32_t p_site, uint32_t m_hi, uint32_t first,
                                  uint32_t count, uint64_t &neg, double (&h)[NPL])
{
    const uint32_t lane = lane_id();
    constexpr int kRows = (64 / NPL) < 1 ? 1 : ((64 / NPL) > KB ? KB : (64 / NPL));
#pragma unroll
    for (int g = 0; g < KB; g += kRows) {
        if (uint32_t(g) >= count) break;
        double r[kRows][NPL];
        uint32_t site[kRows];
        double m[kRows];
#pragma unroll
        for (int j = 0; j < kRows; ++j) {
            const uint32_t want = first + uint32_t(g + j);
            const uint32_t L = min(want, uint32_t(kWave - 1));
            site[j] = __builtin_amdgcn_readlane(p_site, L);
            const uint32_t mh = __builtin_amdgcn_readlane(m_hi, L);
            m[j] = hi_double(want < uint32_t(kWave) ? mh : 0u);
            const double *row = a.J + size_t(site[j]) * a.np;
#pragma unroll
            for (int k = 0; k < NPL; ++k) r[j][k] = row[k * kWave + lane];
        }
#pragma unroll
        for (int j = 0; j < kRows; ++j) {
#pragma unroll
            for (int k = 0; k < NPL; ++k) h[k] = fma(m[j], r[j][k], h[k]);
            if (m[j] != 0.0 && lane == (site[j] & (kWave - 1)))
                neg ^= uint64_t(1) << (site[j] >> 6);
        }
    }
}

constexpr int kLookaheadBatch = 16;

template <int NPL>
__global__ __launch_bounds__(kWave) void lookahead_kernel(const Args *table,
                                                  const BatchSlot *slots,
                                                  uint64_t request, uint32_t flags)
{
    const BatchSlot slot = slots[blockIdx.x];
    const Args a = table[slot.engine];
    constexpr int KB = kLookaheadBatch;
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t w = slot.walker, lane = lane_id();
    if ((flags & kGate) && a.glob->done) return;
    WalkerState st = a.state[w];
    if (st.escaped) return;
    const uint32_t ws = st.end - st.begin;
    Bins v = bins_view(smem, a.ws_stride);
    load_bins(a, w, ws, v);
    uint64_t neg;
    double h[NPL];
    load_config<NPL>(a, st.slot, neg, h);
    double energy = a.energy[st.slot];
    __syncthreads();
    const uint64_t count = batch_count(a, st, request, flags);
    const uint64_t trace_base = uint64_t(w) * a.trace_cap;
    const int64_t start_bin = grid_index(a.grid, energy);
    // the current bin can only leave the window through an accepted move, which requires the
    // proposed bin to lie inside: the check of wang_landau.cpp:247-248 is decided up front
    if (count > 0 && (start_bin < int64_t(st.begin) || start_bin >= int64_t(st.end))) {
        st.escaped = 1;
    } else if (count > 0) {
        const uint32_t begin = st.begin, end = st.end;
        uint32_t cur = uint32_t(start_bin);
        double lg_cur = v.lg[cur - begin];
        bool cur_seen = v.act[cur - begin] != 0;
        bool cur_ref = !a.nalivaiko || v.ref[cur - begin] != 0;
        uint32_t p_site = 0, used = kWave;
        double p_lu = 0.0;
        uint64_t remaining = count;
        while (remaining > 0) {
            if (used == kWave) { // proposals of attempts attempted + lane
                const uint64_t t = st.attempted + lane;
                const Philox4 blk = philox4x64(t >> 1, kPhiloxWalkTag, 0, 0, a.seed, uint64_t(w));
                const bool odd = t & 1;
                p_site = uint32_t(__umul64hi(odd ? blk.v[2] : blk.v[0], uint64_t(a.n)));
                const double u = double((odd ? blk.v[3] : blk.v[1]) >> 11) * 0x1.0p-53;
                p_lu = log(fmax(u, 0x1.0p-53));
                used = 0;
            }
            const uint32_t K =
                uint32_t(min(min(uint64_t(KB), remaining), uint64_t(kWave - used)));
            // field and -2 sigma at each proposal site; couplings among the batch's sites
            double h_run = gather_field<NPL>(h, p_site);
            const uint32_t src = p_site & (kWave - 1);
            const uint64_t bits =
                (uint64_t(uint32_t(__shfl(int(uint32_t(neg >> 32)), int(src)))) << 32) |
                uint32_t(__shfl(int(uint32_t(neg)), int(src)));
            uint32_t dsig_hi = ((bits >> (p_site >> 6)) & 1) ? 0x40000000u : 0xC0000000u;
            double jc[KB];
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const uint32_t sq = __builtin_amdgcn_readlane(
                    p_site, min(used + uint32_t(q), uint32_t(kWave - 1)));
                jc[q] = a.J[size_t(sq) * a.np + p_site];
            }
            // 1/t factors (update_inverse_time_factor, wang_landau.cpp:309-315): lane x holds
            // active / attempted for the attempt of lane x, computed for the batch at once
            double fac_lane = 0.0;
            if (st.stage == WLX_STAGE_INVERSE_TIME)
                fac_lane = double(st.active_count) /
                           double(st.attempted + st.clock_offset + 1 + uint64_t(lane) -
                                  uint64_t(used));
            uint32_t m_hi = 0;
            double lr_lane = 0.0;
            uint64_t drew_mask = 0;
            const uint64_t batch_start = st.attempted;
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                if (uint32_t(b) >= K) break;
                const uint32_t L = used + uint32_t(b);
                const double hs = readlane_f64(h_run, L);
                const uint32_t ms_hi = __builtin_amdgcn_readlane(dsig_hi, L);
                const double ms = hi_double(ms_hi); // -2 sigma
                const double proposed = energy + ms * hs; // flip_delta, physics.cpp:261-266
                bool in;
                const uint32_t idx = grid_bin(a.grid, proposed, in);
                const bool inside = in && idx >= begin && idx < end;
                const uint32_t rel = inside ? idx - begin : 0u;
                const double lgn = v.lg[rel];
                const bool seen_n = v.act[rel] != 0;
                const bool ref_n = !a.nalivaiko || v.ref[rel] != 0;
                const double log_ratio = lg_cur - lgn;
                const bool force_due =
                    a.force_after != 0 && st.attempted - st.last_accepted >= a.force_after;
                const double lu = readlane_f64(p_lu, L);
                const bool up = log_ratio >= 0.0;
                const bool drew = inside && !up && !force_due;
                const bool forced = inside && !up && force_due;
                const bool accepted = inside && (up || force_due || lu < log_ratio);
                ++st.attempted;
                drew_mask |= drew ? uint64_t(1) << L : 0;
                lr_lane = lane == L ? log_ratio : lr_lane;
                m_hi = lane == L ? (accepted ? ms_hi : 0u) : m_hi;
                h_run = fma(accepted ? ms : 0.0, jc[b], h_run);
                const uint32_t s = __builtin_amdgcn_readlane(p_site, L);
                dsig_hi ^= (accepted && p_site == s) ? 0x80000000u : 0u;
                energy = accepted ? proposed : energy;
                cur = accepted ? idx : cur;
                lg_cur = accepted ? lgn : lg_cur;
                cur_seen = accepted ? seen_n : cur_seen;
                cur_ref = accepted ? ref_n : cur_ref;
                st.accepted += accepted ? 1 : 0;
                st.forced += forced ? 1 : 0;
                st.last_accepted = accepted ? st.attempted : st.last_accepted;
                if (st.stage == WLX_STAGE_INVERSE_TIME) {
                    if (st.active_count != 0) st.factor = readlane_f64(fac_lane, L);
                    if (st.factor <= a.final_factor) st.stage = WLX_STAGE_FROZEN;
                }
                // update_current_bin (wang_landau.cpp:290-303)
                const uint32_t crel = cur - begin;
                if (!cur_seen || !cur_ref) {
                    if (!cur_seen) {
                        ++st.active_count;
                        cur_seen = true;
                        if (lane == 0) v.act[crel] = 1;
                        if (st.stage == WLX_STAGE_INVERSE_TIME)
                            fac_lane = double(st.active_count) /
                                       double(st.attempted + st.clock_offset + uint64_t(lane) -
                                              uint64_t(L));
                    }
                    if (!cur_ref) {
                        ++st.refine_count;
                        cur_ref = true;
                        if (lane == 0) v.ref[crel] = 1;
                    }
                }
                lg_cur = lg_cur + (st.stage != WLX_STAGE_FROZEN ? st.factor : 0.0);
                if (lane == 0) {
                    v.lg[crel] = lg_cur;
                    __hip_atomic_fetch_add(&v.H[crel], uint64_t(1), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if ((flags & kTrace) && st.trace_count < a.trace_cap) {
                    if (lane == 0) {
                        const uint64_t k = trace_base + st.trace_count;
                        a.trace_site[k] = s;
                        a.trace_bin[k] = cur;
                        a.trace_flag[k] = uint8_t((accepted ? WLX_TRACE_ACCEPTED : 0) |
                                                  (forced ? WLX_TRACE_FORCED : 0) |
                                                  (drew ? WLX_TRACE_UNIFORM : 0) |
                                                  (inside ? WLX_TRACE_IN_WINDOW : 0));
                    }
                    ++st.trace_count;
                }
            }
            if (drew_mask) { // decisions whose log(u) lay within 4 ulp of the threshold
                const bool amb = ((drew_mask >> lane) & 1) && near_threshold(p_lu, lr_lane);
                st.ambiguous += __popcll(__ballot(amb));
            }
            const uint32_t done = uint32_t(st.attempted - batch_start);
            apply_rows<NPL, KB>(a, p_site, m_hi, used, done, neg, h);
            used += done;
            remaining -= done;
        }
    }
    __syncthreads();
    finish_walker<NPL>(a, w, st, v, ws, neg, h, energy, flags);
}


// h += m_b J[s_b, :] for the `count` proposals of a batch (lanes first..), without the spin
// update (see apply_rows for the exactness argument)
template <int NPL, int KB>
__device__ inline void rows_fma(const Args &a, uint32_t p_site, uint32_t m_hi, uint32_t first,
                                uint32_t count, double (&h)[NPL])
{
    const uint32_t lane = lane_id();
    constexpr int kRows = (64 / NPL) < 1 ? 1 : ((64 / NPL) > KB ? KB : (64 / NPL));
#pragma unroll
    for (int g = 0; g < KB; g += kRows) {
        if (uint32_t(g) >= count) break;
        double r[kRows][NPL];
        double m[kRows];
#pragma unroll
        for (int j = 0; j < kRows; ++j) {
            const uint32_t want = first + uint32_t(g + j);
            const uint32_t L = min(want, uint32_t(kWave - 1));
            const uint32_t site = __builtin_amdgcn_readlane(p_site, L);
            const uint32_t mh = __builtin_amdgcn_readlane(m_hi, L);
            m[j] = hi_double(want < uint32_t(kWave) ? mh : 0u);
            const double *row = a.J + size_t(site) * a.np;
#pragma unroll
            for (int k = 0; k < NPL; ++k) r[j][k] = row[k * kWave + lane];
        }
#pragma unroll
        for (int j = 0; j < kRows; ++j)
#pragma unroll
            for (int k = 0; k < NPL; ++k) h[k] = fma(m[j], r[j][k], h[k]);
    }
}

// ---------------------------------------------------------------------------------------------
// lookahead sampling, common configuration: no trace, no forced acceptance, no Nalivaiko mask
// (the general kernel above covers those). Same decisions in the same order, with less work on
// the serial chain:
//   - the log_g increment of every attempt of a batch (the WL factor, the 1/t factors
//     active/attempted with the attempt at which the stage freezes, or 0 when frozen) is
//     prepared lane-parallel before the chain runs;
//   - H[cur] += 1 of each attempt is recorded per lane and applied by one LDS atomic per batch
//     (nothing reads H inside a batch);
//   - spin flips go through a 64-word LDS box (one XOR per accepted proposal, one read per
//     lane) instead of a lane test per row;
//   - counters within a call are 32-bit.
// ---------------------------------------------------------------------------------------------
constexpr size_t kFlipBoxBytes = kWave * sizeof(uint64_t);

// `value` in lane `lane` of `old`, the other lanes unchanged
__device__ inline uint32_t writelane_u32(uint32_t old, uint32_t value, uint32_t lane)
{
    return lane_id() == lane ? value : old;
}

// Per-lane factor for the attempts of lanes >= first, whose attempted count (after its
// increment) is base + (lane - first) + 1: the WL factor, the 1/t factor active/attempted
// (update_inverse_time_factor, wang_landau.cpp:309-315; unchanged while no bin is active), or
// 0 once frozen, so that update_current_bin's log_g increment is this value.
__device__ inline double batch_factors(uint32_t stage, double factor, uint32_t active,
                                       uint64_t base, uint32_t first)
{
    if (stage == WLX_STAGE_FROZEN) return 0.0;
    if (stage == WLX_STAGE_WANG_LANDAU || active == 0) return factor;
    const uint64_t t = base + 1 + uint64_t(lane_id()) - uint64_t(first); // >= 1 from `first`
    return double(active) / double(t);
}

template <int NPL, int KB>
__device__ __attribute__((always_inline)) inline void fast_walk(const Args &a, uint32_t w,
                                                                uint64_t request, uint32_t flags,
                                                                unsigned char *smem)
{
    static_assert(KB == 4 || KB == 8 || KB == 16 || KB == 32, "lookahead widths 4, 8, 16, 32");
    const uint32_t lane = lane_id();
    if ((flags & kGate) && a.glob->done) return;
    WalkerState st = a.state[w];
    if (st.escaped) return;
    const uint32_t ws = st.end - st.begin;
    Bins v = bins_view(smem, a.ws_stride);
    uint64_t *flipbox = reinterpret_cast<uint64_t *>(smem + bins_lds_bytes(a.ws_stride));
    load_bins(a, w, ws, v);
    flipbox[lane] = 0;
    uint64_t neg;
    double h[NPL];
    load_config<NPL>(a, st.slot, neg, h);
    double energy = uniform_f64(a.energy[st.slot]);
    __syncthreads();
    const uint64_t count = batch_count(a, st, request, flags);
    const int64_t start_bin = grid_index(a.grid, energy);
    if (count > 0 && (start_bin < int64_t(st.begin) || start_bin >= int64_t(st.end))) {
        st.escaped = 1; // wang_landau.cpp:247-248, decided up front as in the general kernel
    } else if (count > 0) {
        const uint32_t begin = st.begin, end_bin = st.end;
        const uint64_t att0 = st.attempted, clk0 = att0 + st.clock_offset;
        uint32_t cur = uint32_t(start_bin);
        double lg_cur = uniform_f64(v.lg[cur - begin]);
        bool cur_seen = uniform_u32(v.act[cur - begin]) != 0;
        uint32_t stage = uniform_u32(st.stage), active = uniform_u32(st.active_count);
        double factor = uniform_f64(st.factor);
        uint32_t done_total = 0, n_acc = 0, last_acc = 0; // last_acc: 1 + index, 0 = none
        uint32_t p_site = 0, used = kWave;
        double p_lu = 0.0;
        uint64_t remaining = count;
        while (remaining > 0) {
            if (used == kWave) { // proposals of attempts att0 + done_total + lane
                const uint64_t t = att0 + done_total + lane;
                const Philox4 blk = philox4x64(t >> 1, kPhiloxWalkTag, 0, 0, a.seed, uint64_t(w));
                const bool odd = t & 1;
                p_site = uint32_t(__umul64hi(odd ? blk.v[2] : blk.v[0], uint64_t(a.n)));
                const double u = double((odd ? blk.v[3] : blk.v[1]) >> 11) * 0x1.0p-53;
                p_lu = log(fmax(u, 0x1.0p-53));
                used = 0;
            }
            const uint32_t rem32 = remaining < uint64_t(kWave) ? uint32_t(remaining) : kWave;
            const uint32_t K = min(min(uint32_t(KB), rem32), kWave - used);
            double h_run = gather_field<NPL>(h, p_site);
            const uint32_t src = p_site & (kWave - 1);
            const uint64_t bits =
                (uint64_t(uint32_t(__shfl(int(uint32_t(neg >> 32)), int(src)))) << 32) |
                uint32_t(__shfl(int(uint32_t(neg)), int(src)));
            uint32_t dsig_hi = ((bits >> (p_site >> 6)) & 1) ? 0x40000000u : 0xC0000000u;
            double jc[KB];
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const uint32_t sq = __builtin_amdgcn_readlane(
                    p_site, min(used + uint32_t(q), uint32_t(kWave - 1)));
                jc[q] = a.J[size_t(sq) * a.np + p_site];
            }
            // log_g increments of the batch's attempts, lane-parallel: the WL factor, the 1/t
            // factors, or 0 from the attempt whose 1/t factor reaches final_factor (it freezes
            // the stage, wang_landau.cpp:309-315) on. A first visit recomputes them for the
            // later attempts with the new active count.
            const bool inverse_time = stage == WLX_STAGE_INVERSE_TIME;
            double inc = batch_factors(stage, factor, active, clk0 + done_total, used);
            uint32_t freeze = kWave; // lane of the freezing attempt; kWave: none
            if (inverse_time) {
                const uint64_t fz = __ballot(lane >= used && inc <= a.final_factor);
                freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : kWave;
            }
            double incz = lane >= freeze ? 0.0 : inc;
            uint32_t m_hi = 0, cur_lane = 0;
            double lr_lane = 0.0;
            const uint32_t used_s = uniform_u32(used);
            // EnergyGrid::index (wang_landau.cpp:27-34) folded into the window test: energies
            // below the grid give negative bins, and above it bins >= count >= end, both outside
            // [begin, end); only a window that ends at the grid's last bin tests the maximum
            // (above it: outside; equal to it: the last bin)
            const bool top = end_bin == uint32_t(a.grid.count);
            const double count_bins = double(a.grid.count);
            // +inf unless the window ends at the grid's last bin: the tests below are then false
            const double top_max = top ? a.grid.maximum : __builtin_huge_val();
            auto window_bin = [&](double proposed, double fb) __attribute__((always_inline)) {
                // far outside the grid stays outside after the conversion
                int32_t idx = int32_t(fmin(fmax(fb, -1.0), count_bins));
                idx = proposed > top_max ? -1 : idx;
                idx = proposed == top_max ? int32_t(a.grid.count) - 1 : idx;
                return uint32_t(idx) - begin;
            };
            auto decide = [&](uint32_t b) __attribute__((always_inline)) {
                const uint32_t L = used_s + b;
                const double hs = readlane_f64(h_run, L);
                const uint32_t ms_hi = __builtin_amdgcn_readlane(dsig_hi, L);
                const double ms = hi_double(ms_hi); // -2 sigma
                const double proposed = energy + ms * hs; // flip_delta, physics.cpp:261-266
                const double lu = readlane_f64(p_lu, L);
                // bin from the reciprocal product; the LDS reads go out before the rare check
                // that the product's floor is the correctly rounded quotient's (bin_floor)
                const double x = proposed - a.grid.minimum;
                const double q = x * a.grid.inverse_width;
                const double fq = floor(q);
                uint32_t rel = window_bin(proposed, fq);
                bool inside = rel < ws;
                uint32_t rr = inside ? rel : 0u;
                double lgn = v.lg[rr];
                bool seen_n = v.act[rr] != 0;
                const double fr = q - fq; // exact: fq <= q < fq + 1
                if (fr < a.grid.margin || fr > a.grid.margin_hi) {
                    rel = window_bin(proposed, floor(x / a.grid.width));
                    inside = rel < ws;
                    rr = inside ? rel : 0u;
                    lgn = v.lg[rr];
                    seen_n = v.act[rr] != 0;
                }
                const uint32_t idx = rel + begin;
                const double log_ratio = lg_cur - lgn;
                const bool up = log_ratio >= 0.0;
                const bool accepted = inside && (up || lu < log_ratio);
                // log ratio of a drawn decision, +inf when no uniform was drawn (outside)
                lr_lane = lane == L ? (inside ? log_ratio : __builtin_huge_val()) : lr_lane;

                m_hi = writelane_u32(m_hi, accepted ? ms_hi : 0u, L);
                h_run = fma(accepted ? ms : 0.0, jc[b], h_run);
                const uint32_t s = __builtin_amdgcn_readlane(p_site, L);
                dsig_hi ^= (accepted && p_site == s) ? 0x80000000u : 0u;
                energy = accepted ? proposed : energy;
                cur = accepted ? idx : cur;
                lg_cur = accepted ? lgn : lg_cur;
                cur_seen = accepted ? seen_n : cur_seen;
                const uint32_t crel = cur - begin;
                if (!cur_seen) { // first visit: update_current_bin (wang_landau.cpp:294-297)
                    ++active;
                    cur_seen = true;
                    v.act[crel] = 1; // every lane writes the same byte
                    if (inverse_time && L < freeze) { // later attempts see the new count
                        inc = lane > L ? batch_factors(WLX_STAGE_INVERSE_TIME, factor, active,
                                                       clk0 + done_total + b + 1, L + 1)
                                       : inc;
                        const uint64_t fz = __ballot(lane > L && inc <= a.final_factor);
                        freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : kWave;
                        incz = lane > L ? (lane >= freeze ? 0.0 : inc) : incz;
                    }
                }
                lg_cur = lg_cur + readlane_f64(incz, L);
                v.lg[crel] = lg_cur; // every lane writes the same value
                cur_lane = writelane_u32(cur_lane, crel, L);
            };
            if (K == uint32_t(KB)) {
#pragma unroll
                for (int b = 0; b < KB; ++b) decide(uint32_t(b));
            } else { // the last batch of a call
#pragma unroll 1
                for (uint32_t b = 0; b < K; ++b) decide(b);
            }
            if (inverse_time) { // the factor of the last 1/t attempt, and a freeze
                const uint32_t last_lane = used + K - 1;
                factor = readlane_f64(inc, min(freeze, last_lane));
                if (freeze <= last_lane) stage = WLX_STAGE_FROZEN;
            }
            const uint32_t last = used + K - 1;
            const bool mine = lane >= used && lane <= last;
            { // accepted proposals of the batch: the count and the last one
                const uint64_t acc_lanes = __ballot(mine && m_hi != 0u);
                n_acc += uint32_t(__popcll(acc_lanes));
                if (acc_lanes)
                    last_acc = done_total + (63u - uint32_t(__clzll(acc_lanes))) - used + 1u;
            }
            if (mine) // H[cur] += 1 of every attempt of the batch
                __hip_atomic_fetch_add(&v.H[cur_lane], uint64_t(1), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            { // drawn decisions (log ratio < 0) whose log(u) lay within 4 ulp of the threshold
                const bool amb = mine && lr_lane < 0.0 && near_threshold(p_lu, lr_lane);
                st.ambiguous += __popcll(__ballot(amb));
            }
            // spin flips of the accepted proposals, in any order (XOR commutes)
            if (mine && m_hi != 0)
                __hip_atomic_fetch_xor(&flipbox[p_site & (kWave - 1)], uint64_t(1) << (p_site >> 6),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __syncthreads();
            const uint64_t flips = flipbox[lane];
            neg ^= flips;
            flipbox[lane] = 0;
            rows_fma<NPL, KB>(a, p_site, m_hi, used, K, h);
            used += K;
            remaining -= K;
            done_total += K;
        }
        // reload the fields the loop did not touch instead of keeping them in registers
        asm volatile("" ::: "memory");
        const uint64_t accepted0 = st.accepted, ambiguous = st.ambiguous;
        st = a.state[w];
        st.attempted = att0 + done_total;
        st.accepted = accepted0 + n_acc;
        if (last_acc) st.last_accepted = att0 + last_acc;
        st.stage = stage;
        st.factor = factor;
        st.active_count = active;
        st.ambiguous = ambiguous;
    }
    __syncthreads();
    finish_walker<NPL>(a, w, st, v, ws, neg, h, energy, flags);
}

// KB: attempts resolved per batch (the lookahead width), which is also the number of
// accepted rows streamed together after a batch
template <int NPL, int KB>
__global__ __launch_bounds__(kWave) void lookahead_fast_kernel(const Args *table,
                                                  const BatchSlot *slots,
                                                  uint64_t request, uint32_t flags)
{
    const BatchSlot slot = slots[blockIdx.x];
    const Args a = table[slot.engine];
    extern __shared__ __align__(16) unsigned char smem[];
    fast_walk<NPL, KB>(a, slot.walker, request, flags, smem);
}


// ---------------------------------------------------------------------------------------------
// lookahead sampling with the candidate rows staged in LDS, two waves per walker (common
// configuration, Philox stream, 65 to 512 sites). The decisions are those of fast_walk, in the
// same order. A wave issues about one vector instruction every 4.6 cycles, in order, whether or
// not it depends on the previous one, so a walker's speed is set by the instructions its
// decision loop holds; here the loop holds little besides the arithmetic of each decision and
// one LDS read, and everything else runs on a second wave of the same workgroup, on another
// SIMD of the CU:
//   - wave A (the decider) owns the walker: energy, current bin, log_g, visited flags and spin
//     bits. Per batch of 16 attempts it gathers the field at each proposal site, decides the
//     batch and writes one record per decision (bin, accepted m, log ratio, replaced log_g);
//   - wave B (the worker) owns the register fields. It applies the previous batch's records
//     (H, ambiguity count, accepted count), FMA's its accepted rows into the fields and mirrors
//     them in LDS, draws the Philox proposals one block ahead, copies the next batch's rows
//     J[s_q, :] from L2 into LDS with global_load_lds_dwordx4 (1 KiB per instruction) and
//     flags a batch that proposes a site twice;
//   - two workgroup barriers per batch. B's field mirror lags A by one batch, so A's gather adds
//     the previous batch's accepted rows itself, at the new proposal sites only: J[s_a][s_j] is
//     row a of the previous batch's staged rows at column s_j, added in acceptance order, the
//     same additions in the same order as the full row update. The barrier after A's gather
//     holds B's overwrite of those rows until A has read them;
//   - A holds the batch replicated in the four 16-lane rows of its wave (lane l: proposal
//     l % 16), so a DPP row broadcast gives every lane decision b's values without a scalar
//     round trip;
//   - the window's log_g has a +inf entry on either side (relative bins -1 and ws), and the
//     proposed bin is clamped onto them, so a proposal outside the window gives log ratio -inf
//     and is rejected by the same comparison log(u) < log ratio that accepts the others (log u
//     is finite and negative, so that comparison also accepts every log ratio >= 0);
//   - the bin comes from the reciprocal product, which is exact unless the product lies near
//     an integer, or, in the window that ends at the grid's last bin, lies between the grid
//     maximum's product and max(that product, bin count) (rounding is monotonic, so the
//     maximum itself lands there, and a product above the band floors past the last bin);
//   - a batch is first decided without branches, every decision selecting its outcome, and
//     with the visited flags and the active count left for after the batch: one atomic OR per
//     decision's bin then finds the batch's first visits. A decision whose product is not known
//     exact, or a first visit in the 1/t stage (it changes the 1/t factors of later attempts),
//     sends the batch back: registers from a snapshot, LDS log_g (from each decision's record
//     of the value it replaced, in reverse order) and visited flags are restored, and the batch
//     is decided again with the exact quotient, first visits and the 1/t refresh in line;
//   - a batch without a repeated proposal site runs a loop without the sign bookkeeping of
//     repeats, and outside the 1/t stage the log_g increment is one value for the batch.
// ---------------------------------------------------------------------------------------------
// f(std::integral_constant<int, k>) for k = 0 .. N - 1, in order (a compile-time index for
// the DPP broadcasts below)
template <int N, int K = 0, class F> __device__ __attribute__((always_inline)) inline void static_for(F &&f)
{
    if constexpr (K < N) {
        f(std::integral_constant<int, K>{});
        static_for<N, K + 1>(f);
    }
}

// lane B of the value's 16-lane row, in every lane of that row (DPP row_newbcast)
template <int B> __device__ inline uint32_t row_bcast_u32(uint32_t v)
{
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x150 + B, 0xF, 0xF, true));
}
// one v_mov_b64_dpp: gfx950's 64-bit DPP move takes row_newbcast, so a double moves in one
// instruction instead of two 32-bit moves
template <int B> __device__ inline double row_bcast_f64(double v)
{
    const long long bits = __double_as_longlong(v);
    return __longlong_as_double(__builtin_amdgcn_update_dpp(bits, bits, 0x150 + B, 0xF, 0xF, true));
}

constexpr int kStagedBatch = 16;
constexpr int kStagedThreads = 2 * kWave;

struct StagedLayout {
    uint32_t lg, H, act, ref, rows, fields, spins, dups, psite, plu, rec_cm, rec_lr, rec_old, fin, rflag;
    uint32_t total;
};

// LDS offsets (bytes) of the staged walker: log_g with one entry below the window (lg - 8) and
// one past it, H, the visited flags (4-byte aligned, with spare bytes on either side), the
// refine flags, two row buffers of kStagedBatch rows each, the field and spin-bit mirrors, a
// site table for finding repeated proposals, two blocks of 64 proposals (sites, log u), two
// batches of decision records, the log_g each decision replaced, the waves' final values, and
// two batches' repeated-site flags
__host__ __device__ inline StagedLayout staged_layout(uint32_t stride, uint32_t np)
{
    StagedLayout s;
    s.lg = 8;
    s.H = 8 * (stride + 2);
    s.act = s.H + 8 * stride + 4;
    s.ref = s.act + (stride + 4 + 3) / 4 * 4;
    s.rows = (s.ref + stride + 15) / 16 * 16;
    s.fields = s.rows + 2 * kStagedBatch * np * 8;
    s.spins = s.fields + np * 8;
    s.dups = s.spins + kWave * 4;
    s.psite = s.dups + np * 4;
    s.plu = s.psite + 2 * kWave * 4;
    s.rec_cm = s.plu + 2 * kWave * 8;
    s.rec_lr = s.rec_cm + 2 * kStagedBatch * 8;
    s.rec_old = s.rec_lr + 2 * kStagedBatch * 8;
    s.fin = s.rec_old + kStagedBatch * 8;
    s.rflag = s.fin + 8 * 8;
    s.total = s.rflag + 2 * 4;
    return s;
}

// v_cvt_i32_f64 with the hardware's saturation for values outside the int32 range
__device__ inline int32_t cvt_i32_sat(double x)
{
    int32_t r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// copies the row at `row` (64 * NPL doubles) into LDS at byte address `dst`: each
// global_load_lds_dwordx4 moves 64 lanes x 16 B = 1 KiB, lane l's 16 B landing at
// M0 + offset + 16 l; the instruction offset advances the global and the LDS address together.
// Counted by vmcnt; stage_wait waits for them. M0 is written here only: the staged walker's
// own code uses no M0.
template <int NPL>
__device__ inline void stage_row(const double *row_any, uint32_t dst_any)
{
    // both are wave-uniform; readfirstlane places them in scalar registers for the "s" operands
    // (the builtin returns int: each half goes through uint32_t, so none is sign-extended)
    const uint32_t dst = uint32_t(__builtin_amdgcn_readfirstlane(dst_any));
    const uint64_t ra = reinterpret_cast<uint64_t>(row_any);
    const uint32_t ra_hi = uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(ra >> 32)));
    const uint32_t ra_lo = uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(ra)));
    const double *row = reinterpret_cast<const double *>((uint64_t(ra_hi) << 32) | ra_lo);
    const uint32_t voff = lane_id() * 16;
    if constexpr (NPL == 2) {
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\t"
                     "global_load_lds_dwordx4 %1, %2" :: "s"(dst), "v"(voff), "s"(row));
    } else if constexpr (NPL == 4) {
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\t"
                     "global_load_lds_dwordx4 %1, %2\n\t"
                     "global_load_lds_dwordx4 %1, %2 offset:1024" :: "s"(dst), "v"(voff), "s"(row));
    } else {
        static_assert(NPL == 8, "staged rows: 1, 2 or 4 KiB");
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\t"
                     "global_load_lds_dwordx4 %1, %2\n\t"
                     "global_load_lds_dwordx4 %1, %2 offset:1024\n\t"
                     "global_load_lds_dwordx4 %1, %2 offset:2048\n\t"
                     "global_load_lds_dwordx4 %1, %2 offset:3072" :: "s"(dst), "v"(voff), "s"(row));
    }
}

// the median of x, lo and hi: x clamped to [lo, hi] for lo <= hi. The compiler turns this
// form of the median into one v_med3_i32 (and, unlike inline assembly, needs no wait state
// after it)
__device__ inline int32_t med3_i32(int32_t x, int32_t lo, int32_t hi)
{
    return max(min(x, lo), min(max(x, lo), hi));
}

__device__ inline void stage_wait()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int NPL>
__device__ __attribute__((always_inline)) inline void staged_walk(const Args &a, uint32_t w,
                                                                  uint64_t request, uint32_t flags,
                                                                  unsigned char *smem)
{
    static_assert(NPL >= 2 && NPL <= 8, "staged rows: 1, 2 or 4 KiB per row");
    constexpr int KB = kStagedBatch;
    constexpr uint32_t kRowBytes = 64 * NPL * 8;
    constexpr uint32_t kRowDoubles = kRowBytes / 8;
    constexpr uint32_t kBufBytes = KB * kRowBytes;
    const uint32_t lane = lane_id();
    // wave A decides, wave B works; the role is uniform in each wave, which readfirstlane tells
    // the compiler (its code then keeps wave-uniform values scalar)
    const bool decider = __builtin_amdgcn_readfirstlane(threadIdx.x) < uint32_t(kWave);
    if ((flags & kGate) && a.glob->done) return;
    WalkerState st = a.state[w];
    if (st.escaped) return;
    const uint32_t ws = st.end - st.begin;
    const StagedLayout lay = staged_layout(a.ws_stride, a.np);
    Bins v = {reinterpret_cast<double *>(smem + lay.lg), reinterpret_cast<uint64_t *>(smem + lay.H),
              smem + lay.act, smem + lay.ref};
    double *fmirror = reinterpret_cast<double *>(smem + lay.fields);
    uint32_t *smirror = reinterpret_cast<uint32_t *>(smem + lay.spins);
    uint32_t *duptab = reinterpret_cast<uint32_t *>(smem + lay.dups);
    uint32_t *psite = reinterpret_cast<uint32_t *>(smem + lay.psite);
    double *plu = reinterpret_cast<double *>(smem + lay.plu);
    uint64_t *rec_cm = reinterpret_cast<uint64_t *>(smem + lay.rec_cm);
    double *rec_lr = reinterpret_cast<double *>(smem + lay.rec_lr);
    double *rec_old = reinterpret_cast<double *>(smem + lay.rec_old);
    uint64_t *fin = reinterpret_cast<uint64_t *>(smem + lay.fin);
    uint32_t *rflag = reinterpret_cast<uint32_t *>(smem + lay.rflag);
    const uint32_t rows_lds = uint32_t(reinterpret_cast<uintptr_t>(smem)) + lay.rows;
    const double *J = a.J;
    const uint64_t count = batch_count(a, st, request, flags);
    // proposals of attempts first + lane (one Philox4x64 block per two attempts)
    auto draw = [&](uint64_t first, uint32_t &site, double &lu) {
        const uint64_t t = first + lane;
        const Philox4 blk = philox4x64(t >> 1, kPhiloxWalkTag, 0, 0, a.seed, uint64_t(w));
        const bool odd = t & 1;
        site = uint32_t(__umul64hi(odd ? blk.v[2] : blk.v[0], uint64_t(a.n)));
        const double u = double((odd ? blk.v[3] : blk.v[1]) >> 11) * 0x1.0p-53;
        lu = fmin(log(fmax(u, 0x1.0p-53)), -0x1.0p-1074); // < 0: see the header
    };
    // wave B: whether the batch at lanes [used, used + KB) of `sites` proposes a site twice (a
    // later proposal's sign then depends on the earlier one's outcome), into rflag[slot]; one
    // wave's LDS accesses complete in order, so its read sees its own writes
    auto flag_repeats = [&](uint32_t sites, uint32_t used, uint32_t slot) {
        const bool in = lane >= used && lane < used + uint32_t(KB);
        if (in) duptab[sites] = lane;
        asm volatile("" ::: "memory"); // the other lanes' stores are read back
        const bool dup = __ballot(in && duptab[sites] != lane) != 0;
        if (lane == 0) rflag[slot] = dup ? 1u : 0u;
    };
    // wave B requests block 0 of the proposals and the first batch's rows before the walker's
    // bins and configuration are loaded, so that their latencies overlap
    uint32_t b_site = 0; // B: the proposal sites of the current block
    if (!decider && count > 0) {
        double lu;
        draw(st.attempted, b_site, lu);
        psite[lane] = b_site;
        plu[lane] = lu;
#pragma unroll
        for (int q = 0; q < KB; ++q)
            stage_row<NPL>(J + size_t(uint32_t(__builtin_amdgcn_readlane(b_site, uint32_t(q)))) * kRowDoubles,
                           rows_lds + uint32_t(q) * kRowBytes);
        flag_repeats(b_site, 0, 0);
    }
    // the window's bins by both waves, four bins per thread with every load of a round issued
    // before its LDS stores; the configuration by both (the stores agree)
    {
        const size_t base = size_t(w) * a.ws_stride;
        for (uint32_t b0 = threadIdx.x; b0 < ws; b0 += 4 * kStagedThreads) {
            double lg4[4] = {0.0, 0.0, 0.0, 0.0};
            uint64_t h4[4] = {0, 0, 0, 0};
            uint8_t act4[4] = {0, 0, 0, 0}, ref4[4] = {0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t b = b0 + uint32_t(j) * kStagedThreads;
                if (b < ws) {
                    lg4[j] = a.logg[base + b];
                    h4[j] = a.hist[base + b];
                    act4[j] = a.active[base + b];
                    ref4[j] = a.refine[base + b];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t b = b0 + uint32_t(j) * kStagedThreads;
                if (b < ws) {
                    v.lg[b] = lg4[j];
                    v.H[b] = h4[j];
                    v.act[b] = act4[j];
                    v.ref[b] = ref4[j];
                }
            }
        }
    }
    uint64_t neg;
    double h[NPL];
    load_config<NPL>(a, st.slot, neg, h);
    if (decider) {
        if (lane == 0) {
            v.lg[-1] = __builtin_huge_val();
            v.lg[ws] = __builtin_huge_val();
        }
        smirror[lane] = uint32_t(neg);
    } else {
#pragma unroll
        for (int k = 0; k < NPL; ++k) fmirror[k * kWave + lane] = h[k];
    }
    double energy = uniform_f64(a.energy[st.slot]);
    const int64_t start_bin = grid_index(a.grid, energy);
    stage_wait(); // B: the first batch's rows are in LDS
    __syncthreads();
    if (count > 0 && (start_bin < int64_t(st.begin) || start_bin >= int64_t(st.end))) {
        st.escaped = 1; // wang_landau.cpp:247-248, decided up front as in the general kernel
    } else if (count > 0) {
        const uint32_t begin = st.begin;
        const uint64_t att0 = st.attempted, clk0 = att0 + st.clock_offset;
        const uint32_t nb = uint32_t((count + KB - 1) / KB); // batches, the last one partial
        uint32_t n_acc = 0, last_acc = 0; // B
        uint64_t ambiguous = 0;           // B
        // ---- wave A: the walker ----
        const double gmin = a.grid.minimum, ginv = a.grid.inverse_width;
        const double margin = a.grid.margin, margin_hi = a.grid.margin_hi;
        // the product of the grid maximum (top window only, +inf elsewhere): at or above it the
        // exact path takes the quotient
        const double q_top = st.end == uint32_t(a.grid.count)
                                 ? (a.grid.maximum - a.grid.minimum) * a.grid.inverse_width
                                 : __builtin_huge_val();
        // products in [q_top, max(q_top, count)] are the grid maximum itself (the last bin) or
        // energies above it whose product's floor is still inside the grid: the quotient decides
        // them. Above that band the floor is at least count and the proposal lands on the +inf
        // entry, rejected as an energy outside the grid. The fast path tests the distance to the
        // band's centre against its half width plus `margin` (at least 4 ulp of count, more
        // than the rounding of the centre and of the distance); outside the top window the centre
        // is +inf and no distance is within it.
        const double band_hi = fmax(q_top, double(a.grid.count));
        const double band_c = st.end == uint32_t(a.grid.count) ? 0.5 * (q_top + band_hi)
                                                                : __builtin_huge_val();
        const double band_r = 0.5 * (band_hi - q_top) + margin;
        const int32_t lo = int32_t(begin) - 1, hi = int32_t(st.end);
        double *lgm = v.lg - begin; // by absolute bin
        uint8_t *actm = v.act - begin;
        int32_t cur = int32_t(start_bin); // absolute bin
        double lgc = 0.0;
        uint32_t cur_seen = 0;
        if (decider) {
            lgc = uniform_f64(lgm[cur]);
            cur_seen = uniform_u32(actm[cur]) != 0 ? 1u : 0u;
        }
        uint32_t stage = uniform_u32(st.stage), active = uniform_u32(st.active_count);
        double factor = uniform_f64(st.factor);
        uint32_t m_prev = 0; // A: the previous batch's accepted signs (lane l: proposal l % 16)
        for (uint32_t i = 0; i <= nb; ++i) {
            // ---- wave A: gather batch i, before B may overwrite batch i - 1's rows ----
            // A holds the batch replicated in the four 16-lane rows of the wave: lane l carries
            // proposal l % 16, so lane b of every row is decision b, and a DPP row broadcast
            // (row_newbcast) gives every lane its value without a scalar round trip
            const uint32_t used = (i * KB) & (kWave - 1);
            const uint32_t blk = ((i * KB) / kWave) & 1;
            const uint32_t done_total = i * KB;
            const uint32_t jl = lane & (KB - 1); // this lane's proposal in the batch
            uint32_t p_site = 0, mv_hi = 0, rep_flag = 0;
            double p_lu = 0.0, hv = 0.0;
            double jc[KB];
            if (decider && i < nb) {
                p_site = psite[blk * kWave + used + jl];
                p_lu = plu[blk * kWave + used + jl];
                const double *rows =
                    reinterpret_cast<const double *>(smem + lay.rows + (i & 1) * kBufBytes);
                // B's mirror holds the fields after batch i - 2; batch i - 1's rows are added at
                // this batch's sites, in order (m = 0 for a rejection), from batch i - 1's staged
                // rows: J[s_a][s_j] is row a of that batch at column s_j
                hv = fmirror[p_site];
                double pr[KB]; // batch i - 1's rows at this site
                const double *prow = reinterpret_cast<const double *>(
                                         smem + lay.rows + ((i - 1) & 1) * kBufBytes) + p_site;
                if (i > 0) {
#pragma unroll
                    for (int q = 0; q < KB; ++q) pr[q] = prow[q * kRowDoubles];
                }
                const uint32_t bits = smirror[p_site & (kWave - 1)];
                rep_flag = rflag[i & 1];
#pragma unroll
                for (int q = 0; q < KB; ++q) jc[q] = rows[q * kRowDoubles + p_site];
                // every read above is issued before the additions below wait for the first
                // (the scheduler would otherwise keep two reads in flight)
                __builtin_amdgcn_sched_barrier(0);
                if (i > 0) {
                    // the accepted signs of batch i - 1, kept by this wave in every 16-lane row
                    static_for<KB>([&](auto qc) __attribute__((always_inline)) {
                        constexpr int q = decltype(qc)::value;
                        hv = fma(hi_double(row_bcast_u32<q>(m_prev)), pr[q], hv);
                    });
                }
                mv_hi = ((bits >> (p_site >> 6)) & 1) ? 0x40000000u : 0xC0000000u; // -2 sigma
            }
            __syncthreads(); // B applies and overwrites batch i - 1's rows after this
            if (decider && i < nb) {
                // ================= wave A: decide batch i =================
                const uint64_t left = count - done_total;
                const uint32_t K = left < uint64_t(KB) ? uint32_t(left) : uint32_t(KB);
                const bool inverse_time = stage == WLX_STAGE_INVERSE_TIME;
                // log_g increment of attempt done_total + jl (update_inverse_time_factor,
                // wang_landau.cpp:309-315): the WL factor, the 1/t factor active / attempted
                // (unchanged while no bin is active), or 0 once frozen; `from` is the first
                // proposal whose factor is (re)computed
                auto factor_at = [&](uint32_t stage_now, uint32_t active_now, uint32_t from) {
                    if (stage_now == WLX_STAGE_FROZEN) return 0.0;
                    if (stage_now == WLX_STAGE_WANG_LANDAU || active_now == 0) return factor;
                    const uint64_t t = clk0 + done_total + 1 + uint64_t(jl >= from ? jl : from);
                    return double(active_now) / double(t);
                };
                double inc = factor_at(stage, active, 0);
                uint32_t freeze = KB; // proposal of the freezing attempt; KB: none
                if (inverse_time) {
                    const uint64_t fz = __ballot(lane < uint32_t(KB) && inc <= a.final_factor);
                    freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : uint32_t(KB);
                }
                double incz = jl >= freeze ? 0.0 : inc;
                double dv = hi_double(mv_hi) * hv; // flip_delta of each proposal (physics.cpp:261-266)
                uint64_t *rcm = rec_cm + (i & 1) * KB;
                double *rlr = rec_lr + (i & 1) * KB;
                // the exact decision: the correctly rounded quotient where the product is not
                // known exact, first visits in line, refreshing the 1/t factors of later attempts
                auto decide_exact = [&](uint32_t b) __attribute__((always_inline)) {
                    const double proposed = energy + readlane_f64(dv, b);
                    const double q = (proposed - gmin) * ginv;
                    const double fq = floor(q);
                    int32_t idx = min(max(cvt_i32_sat(fq), lo), hi);
                    const double fr = q - fq; // exact: fq <= q < fq + 1
                    if (fr < margin || fr > margin_hi || q >= q_top) {
                        bool in;
                        const uint32_t g = grid_bin(a.grid, proposed, in);
                        idx = in ? min(max(int32_t(g), lo), hi) : lo;
                    }
                    const double lgn = lgm[idx];
                    const uint32_t actn = actm[idx];
                    const double lr = lgc - lgn; // -inf outside the window
                    const double luL = readlane_f64(p_lu, b);
                    const uint32_t mL = __builtin_amdgcn_readlane(mv_hi, b);
                    const uint32_t sL = __builtin_amdgcn_readlane(p_site, b);
                    const bool accepted = luL < lr;
                    const uint32_t c_hi = accepted ? mL : 0u;
                    hv = fma(hi_double(c_hi), jc[b], hv);
                    const uint32_t t = p_site == sL ? c_hi : 0u;
                    mv_hi += (t & 0x40000000u) << 1; // a later proposal of the same site: -m
                    dv = hi_double(mv_hi) * hv;
                    energy = accepted ? proposed : energy;
                    cur = accepted ? idx : cur;
                    lgc = accepted ? lgn : lgc;
                    cur_seen = accepted ? (actn != 0 ? 1u : 0u) : cur_seen;
                    if (!cur_seen) { // first visit: update_current_bin (wang_landau.cpp:294-297)
                        ++active;
                        cur_seen = 1;
                        actm[cur] = 1; // every lane writes the same byte
                        if (inverse_time && b < freeze) { // later attempts see the new count
                            inc = jl > b ? factor_at(WLX_STAGE_INVERSE_TIME, active, b + 1) : inc;
                            const uint64_t fz =
                                __ballot(lane < uint32_t(KB) && lane > b && inc <= a.final_factor);
                            freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : uint32_t(KB);
                            incz = jl > b ? (jl >= freeze ? 0.0 : inc) : incz;
                        }
                    }
                    lgc = lgc + readlane_f64(incz, b);
                    lgm[cur] = lgc; // every lane writes the same value
                    rcm[b] = (uint64_t(c_hi) << 32) | uint32_t(cur);
                    rlr[b] = lr;
                };
                if (K == uint32_t(KB)) {
                    // snapshot for a replay
                    const double energy0 = energy, lgc0 = lgc, hv0 = hv, dv0 = dv;
                    const int32_t cur0 = cur;
                    const uint32_t seen0 = cur_seen, active0 = active, mv0 = mv_hi;
                    // a site proposed twice in the batch (found by wave B when it staged the
                    // batch's rows)
                    const bool repeats = uniform_u32(rep_flag) != 0;
                    const double inc_u = uniform_f64(inc); // WL or frozen: one increment
                    // the batch's smallest and largest fractional parts and the smallest
                    // distance to the top window's band:
                    // a lane-wise minimum and maximum per decision (a comparison per decision
                    // would hand its result to the scalar unit, which waits for it)
                    double fr_min = 1.0, fr_max = 0.0, band_d = __builtin_huge_val();
                    auto fast_batch = [&](auto repeat_tag, auto per_lane_tag)
                        __attribute__((always_inline)) {
                        constexpr bool REPEATS = decltype(repeat_tag)::value;
                        constexpr bool PER_LANE = decltype(per_lane_tag)::value;
                        // the records of decision b - 1 are written while decision b's LDS
                        // read is in flight, with the work that does not need its result (the
                        // compiler would otherwise leave the read's latency uncovered)
                        uint64_t cm_prev = 0;
                        double lr_prev = 0.0, old_prev = 0.0;
                        static_for<KB>([&](auto bc) __attribute__((always_inline)) {
                            constexpr int b = decltype(bc)::value; // branch-free decision b
                            const double proposed = energy + row_bcast_f64<b>(dv);
                            const double q = (proposed - gmin) * ginv;
                            const double fq = floor(q);
                            const int32_t idx = med3_i32(cvt_i32_sat(fq), lo, hi);
                            double lgn = lgm[idx];
                            asm volatile("" ::: "memory"); // the read is issued first
                            if constexpr (b > 0) {
                                rcm[b - 1] = cm_prev;
                                rlr[b - 1] = lr_prev;
                                rec_old[b - 1] = old_prev;
                            }
                            const double fr = q - fq;
                            fr_min = fmin(fr_min, fr);
                            fr_max = fmax(fr_max, fr);
                            band_d = fmin(band_d, fabs(q - band_c));
                            const double luL = row_bcast_f64<b>(p_lu);
                            const uint32_t mL = row_bcast_u32<b>(mv_hi);
                            const double incL = PER_LANE ? row_bcast_f64<b>(incz) : inc_u;
                            uint32_t sL = 0;
                            if constexpr (REPEATS) sL = row_bcast_u32<b>(p_site);
                            asm volatile("" : "+v"(lgn) : "v"(fr_min), "v"(fr_max), "v"(band_d), "v"(luL),
                                         "v"(mL), "v"(incL), "v"(sL) : "memory");
                            const double lr = lgc - lgn;
                            const bool accepted = luL < lr;
                            const uint32_t c_hi = accepted ? mL : 0u;
                            hv = fma(hi_double(c_hi), jc[b], hv);
                            if constexpr (REPEATS) {
                                const uint32_t t = p_site == sL ? c_hi : 0u;
                                mv_hi += (t & 0x40000000u) << 1;
                            }
                            dv = hi_double(mv_hi) * hv;
                            energy = accepted ? proposed : energy;
                            const double lg_old = accepted ? lgn : lgc;
                            cur = accepted ? idx : cur;
                            lgc = lg_old + incL;
                            lgm[cur] = lgc;
                            cm_prev = (uint64_t(c_hi) << 32) | uint32_t(cur);
                            lr_prev = lr;
                            old_prev = lg_old;
                        });
                        rcm[KB - 1] = cm_prev;
                        rlr[KB - 1] = lr_prev;
                        rec_old[KB - 1] = old_prev;
                    };
                    using yes = std::integral_constant<bool, true>;
                    using no = std::integral_constant<bool, false>;
                    if (repeats) {
                        if (inverse_time) fast_batch(yes{}, yes{});
                        else fast_batch(yes{}, no{});
                    } else {
                        if (inverse_time) fast_batch(no{}, yes{});
                        else fast_batch(no{}, no{});
                    }
                    // first visits: an atomic OR per decision's bin into its visited flag, from
                    // all four copies of the batch (exactly one lane finds each newly visited
                    // bin unset, so the count and the restore below are those of one copy)
                    const uint32_t bin = uint32_t(rcm[jl]) - begin;
                    uint32_t *word = reinterpret_cast<uint32_t *>(v.act + (bin & ~3u));
                    const uint32_t shift = 8 * (bin & 3u);
                    const bool first = ((__hip_atomic_fetch_or(word, 1u << shift, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_WORKGROUP) >>
                                         shift) & 0xFFu) == 0;
                    const uint64_t first_lanes = __ballot(first);
                    active += uint32_t(__popcll(first_lanes));
                    cur_seen = 1;
                    const bool redo = fr_min < margin || fr_max > margin_hi || band_d <= band_r;
                    if (redo || (inverse_time && first_lanes)) {
                        // restore the batch's start: visited flags, log_g in reverse order,
                        // registers; then decide the batch exactly
                        if (first)
                            __hip_atomic_fetch_and(word, ~(0xFFu << shift), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll 1
                        for (int b = KB - 1; b >= 0; --b)
                            lgm[int32_t(uint32_t(rcm[b]))] = rec_old[b];
                        energy = energy0;
                        lgc = lgc0;
                        hv = hv0;
                        dv = dv0;
                        cur = cur0;
                        cur_seen = seen0;
                        active = active0;
                        mv_hi = mv0;
#pragma unroll
                        for (int b = 0; b < KB; ++b) decide_exact(uint32_t(b));
                    }
                } else { // the last batch of a call
#pragma unroll
                    for (int b = 0; b < KB; ++b)
                        if (uint32_t(b) < K) decide_exact(uint32_t(b));
                }
                if (inverse_time) { // the factor of the last 1/t attempt, and a freeze
                    factor = readlane_f64(inc, min(freeze, K - 1));
                    if (freeze <= K - 1) stage = WLX_STAGE_FROZEN;
                }
                // spin flips of the accepted proposals (XOR commutes); every row keeps the
                // batch's signs for the next gather, and row 0 flips
                const uint32_t m_row = jl < K ? uint32_t(rcm[jl] >> 32) : 0u;
                m_prev = m_row;
                const uint32_t m_hi = lane < K ? m_row : 0u;
                if (m_hi != 0)
                    __hip_atomic_fetch_xor(&smirror[p_site & (kWave - 1)], 1u << (p_site >> 6),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else if (!decider) {
                // ================= wave B: apply batch i - 1, stage batch i + 1 =================
                if (i > 0) {
                    const uint32_t bi = i - 1;
                    const uint32_t used = (bi * KB) & (kWave - 1);
                    const uint32_t blk = ((bi * KB) / kWave) & 1;
                    const uint64_t left = count - uint64_t(bi) * KB;
                    const uint32_t K = left < uint64_t(KB) ? uint32_t(left) : uint32_t(KB);
                    // lane l holds record l % 16 in every 16-lane row (the signs are then
                    // broadcast by DPP); row 0's first K lanes count the batch once
                    const uint32_t r = lane & (KB - 1);
                    const bool mine = lane < K;
                    const uint64_t cm = rec_cm[(bi & 1) * KB + r];
                    const double lr = rec_lr[(bi & 1) * KB + r];
                    const double lu = plu[blk * kWave + used + r];
                    const uint32_t m_row = r < K ? uint32_t(cm >> 32) : 0u;
                    const uint32_t m_hi = mine ? m_row : 0u;
                    const uint64_t acc_lanes = __ballot(m_hi != 0u);
                    n_acc += uint32_t(__popcll(acc_lanes));
                    if (acc_lanes) last_acc = bi * KB + (63u - uint32_t(__clzll(acc_lanes))) + 1u;
                    if (mine) // H[cur] += 1 of every attempt of the batch
                        __hip_atomic_fetch_add(&v.H[uint32_t(cm) - begin], uint64_t(1),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    // drawn decisions (finite log ratio < 0) whose log(u) lay within 4 ulp of it
                    const bool amb = mine && lr < 0.0 && lr > -__builtin_huge_val() &&
                                     near_threshold(lu, lr);
                    ambiguous += __popcll(__ballot(amb));
                    // h += m_q J[s_q, :] for the batch's proposals in order, m_q = 0 when
                    // rejected (the product is exact and h + 0 = h: fields are never -0.0)
                    const double *rowl =
                        reinterpret_cast<const double *>(smem + lay.rows + (bi & 1) * kBufBytes) + lane;
                    static_for<KB / 4>([&](auto gc) __attribute__((always_inline)) {
                        constexpr int g = 4 * decltype(gc)::value;
                        double rr[4][NPL];
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                            for (int k = 0; k < NPL; ++k)
                                rr[jj][k] = rowl[(g + jj) * kRowDoubles + k * kWave];
                        static_for<4>([&](auto jc_) __attribute__((always_inline)) {
                            constexpr int jj = decltype(jc_)::value;
                            const double m = hi_double(row_bcast_u32<g + jj>(m_row));
#pragma unroll
                            for (int k = 0; k < NPL; ++k) h[k] = fma(m, rr[jj][k], h[k]);
                        });
                    });
#pragma unroll
                    for (int k = 0; k < NPL; ++k) fmirror[k * kWave + lane] = h[k];
                }
                if (i + 1 < nb) { // the rows of batch i + 1, and its block of proposals
                    const uint32_t nused = ((i + 1) * KB) & (kWave - 1);
                    if (nused == 0) {
                        double lu;
                        draw(att0 + uint64_t(i + 1) * KB, b_site, lu);
                        const uint32_t nblk = (((i + 1) * KB) / kWave) & 1;
                        psite[nblk * kWave + lane] = b_site;
                        plu[nblk * kWave + lane] = lu;
                    }
                    const uint32_t nrows = rows_lds + ((i + 1) & 1) * kBufBytes;
                    const uint32_t nused_s = uniform_u32(nused);
#pragma unroll
                    for (int q = 0; q < KB; ++q)
                        stage_row<NPL>(J + size_t(uint32_t(__builtin_amdgcn_readlane(b_site, nused_s + uint32_t(q)))) *
                                               kRowDoubles,
                                       nrows + uint32_t(q) * kRowBytes);
                    flag_repeats(b_site, nused_s, (i + 1) & 1);
                    stage_wait();
                }
            }
            __syncthreads();
        }
        // the waves' results to both waves
        if (decider && lane == 0) {
            fin[0] = __double_as_longlong(energy);
            fin[1] = stage;
            fin[2] = __double_as_longlong(factor);
            fin[3] = active;
        }
        if (!decider && lane == 0) {
            fin[4] = n_acc;
            fin[5] = last_acc;
            fin[6] = ambiguous;
        }
        __syncthreads();
        energy = __longlong_as_double((long long)fin[0]);
        neg = smirror[lane];
#pragma unroll
        for (int k = 0; k < NPL; ++k) h[k] = fmirror[k * kWave + lane];
        const uint64_t done_total = count;
        const uint64_t accepted0 = st.accepted, ambiguous0 = st.ambiguous;
        st = a.state[w];
        st.attempted = att0 + done_total;
        st.accepted = accepted0 + fin[4];
        if (fin[5]) st.last_accepted = att0 + fin[5];
        st.stage = uint32_t(fin[1]);
        st.factor = __longlong_as_double((long long)fin[2]);
        st.active_count = uint32_t(fin[3]);
        st.ambiguous = ambiguous0 + fin[6];
    }
    __syncthreads();
    if (decider) {
        finish_walker<NPL>(a, w, st, v, ws, neg, h, energy, flags);
    } else { // wave B meets finish_walker_state's barriers (same state, same branch)
        if ((flags & kCheck) && !st.escaped && st.stage == WLX_STAGE_WANG_LANDAU) __syncthreads();
        __syncthreads();
    }
}

template <int NPL>
__global__ __launch_bounds__(kStagedThreads) void lookahead_staged_kernel(const Args *table,
                                                                          const BatchSlot *slots,
                                                                          uint64_t request,
                                                                          uint32_t flags)
{
    const BatchSlot slot = slots[blockIdx.x];
    const Args a = table[slot.engine];
    extern __shared__ __align__(16) unsigned char smem[];
    staged_walk<NPL>(a, slot.walker, request, flags, smem);
}


// ---------------------------------------------------------------------------------------------
// lookahead sampling on the reference's Xoshiro256** stream (common configuration: no trace,
// forcing or Nalivaiko mask). The walker consumes its stream word by word: a site draw (one
// word, repeated while bounded() rejects it, rng.hpp:29-36), then a uniform only when the
// decision needs one (wang_landau.cpp:254-261). The words themselves are a fixed sequence, so
// a wave generates the next 64 on the scalar unit and gives word l to lane l. A batch covers
// 16 stream positions: every position is gathered as a potential site (field, spin and the
// couplings among them), and the serial chain then walks the positions. A position is an
// attempt when it is the next site draw, a uniform when the previous attempt drew one, or a
// rejected draw. The words consumed set the walker's final Xoshiro state.
// ---------------------------------------------------------------------------------------------
__device__ inline void xoshiro_advance(Xoshiro &r, uint32_t steps)
{
    for (uint32_t k = 0; k < steps; ++k) (void)r.next();
}

template <int NPL>
__global__ __launch_bounds__(kWave) void lookahead_xo_kernel(const Args *table,
                                                  const BatchSlot *slots,
                                                  uint64_t request, uint32_t flags)
{
    const BatchSlot slot = slots[blockIdx.x];
    const Args a = table[slot.engine];
    constexpr int KB = kLookaheadBatch;
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t w = slot.walker, lane = lane_id();
    if ((flags & kGate) && a.glob->done) return;
    WalkerState st = a.state[w];
    if (st.escaped) return;
    const uint32_t ws = st.end - st.begin;
    Bins v = bins_view(smem, a.ws_stride);
    uint64_t *flipbox = reinterpret_cast<uint64_t *>(smem + bins_lds_bytes(a.ws_stride));
    load_bins(a, w, ws, v);
    flipbox[lane] = 0;
    uint64_t neg;
    double h[NPL];
    load_config<NPL>(a, st.slot, neg, h);
    double energy = a.energy[st.slot];
    __syncthreads();
    const uint64_t count = batch_count(a, st, request, flags);
    const int64_t start_bin = grid_index(a.grid, energy);
    Xoshiro rng{st.rng[0], st.rng[1], st.rng[2], st.rng[3]};
    if (count > 0 && (start_bin < int64_t(st.begin) || start_bin >= int64_t(st.end))) {
        st.escaped = 1; // wang_landau.cpp:247-248
    } else if (count > 0) {
        const uint32_t begin = st.begin;
        const uint64_t att0 = st.attempted, clk0 = att0 + st.clock_offset;
        uint32_t cur = uint32_t(start_bin);
        double lg_cur = v.lg[cur - begin];
        bool cur_seen = v.act[cur - begin] != 0;
        uint32_t stage = st.stage, active = st.active_count;
        double factor = st.factor;
        uint32_t done_total = 0, n_acc = 0, last_acc = 0;
        Xoshiro block_start = rng; // state before word 0 of the current block
        uint32_t p_site = 0, p_rej = 0, next = 0; // next: block position of the next site draw
        bool have_block = false;
        double p_lu = 0.0;
        uint64_t remaining = count;
        while (remaining > 0) {
            if (!have_block || next >= kWave - 1) { // 64 words from the current position
                if (have_block) xoshiro_advance(block_start, next);
                have_block = true;
                Xoshiro gen = block_start;
                uint64_t word = 0;
                for (uint32_t l = 0; l < kWave; ++l) {
                    const uint64_t x = gen.next();
                    word = lane == l ? x : word;
                }
                const bool rejected = word < a.bounded.threshold;
                const uint64_t site = a.bounded.mask ? (word & a.bounded.mask)
                                      : (a.bounded.n == 1 ? 0 : word % a.bounded.n);
                p_site = uint32_t(site);
                p_rej = rejected ? 1u : 0u;
                const double u = double(word >> 11) * 0x1.0p-53;
                p_lu = log(fmax(u, 0x1.0p-53));
                next = 0;
            }
            // positions next .. next + B - 1 of the block are site candidates; word 63 only
            // ever serves as a uniform of the block's last attempt
            const uint32_t pos = next;
            const uint32_t B = min(uint32_t(KB), kWave - 1 - pos);
            double h_run = gather_field<NPL>(h, p_site);
            const uint32_t src = p_site & (kWave - 1);
            const uint64_t bits =
                (uint64_t(uint32_t(__shfl(int(uint32_t(neg >> 32)), int(src)))) << 32) |
                uint32_t(__shfl(int(uint32_t(neg)), int(src)));
            uint32_t dsig_hi = ((bits >> (p_site >> 6)) & 1) ? 0x40000000u : 0xC0000000u;
            double jc[KB];
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const uint32_t sq = __builtin_amdgcn_readlane(
                    p_site, min(pos + uint32_t(q), uint32_t(kWave - 1)));
                jc[q] = a.J[size_t(sq) * a.np + p_site];
            }
            // lane j: the log_g increment of the batch's j-th attempt; from the attempt whose
            // 1/t factor reaches final_factor (it freezes the stage) on, 0
            const bool inverse_time = stage == WLX_STAGE_INVERSE_TIME;
            double inc = batch_factors(stage, factor, active, clk0 + done_total, 0);
            uint32_t freeze = kWave; // attempt that freezes the stage; kWave: none
            if (inverse_time) {
                const uint64_t fz = __ballot(inc <= a.final_factor);
                freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : kWave;
            }
            double incz = lane >= freeze ? 0.0 : inc;
            uint32_t m_hi = 0, cur_lane = 0, attempts = 0;
            double lr_lane = 0.0;
            uint64_t attempt_lanes = 0;
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                const uint32_t L = pos + uint32_t(b);
                if (uint32_t(b) >= B || L != next || remaining == attempts) continue;
                if (__builtin_amdgcn_readlane(p_rej, L)) { // bounded() draws again
                    next = L + 1;
                    continue;
                }
                const double hs = readlane_f64(h_run, L);
                const uint32_t ms_hi = __builtin_amdgcn_readlane(dsig_hi, L);
                const double ms = hi_double(ms_hi); // -2 sigma
                const double proposed = energy + ms * hs; // flip_delta, physics.cpp:261-266
                const bool in = proposed >= a.grid.minimum && proposed <= a.grid.maximum;
                const double fb = bin_floor(a.grid, proposed);
                int32_t idx = int32_t(in ? fb : 0.0);
                if (proposed == a.grid.maximum) idx = min(idx, int32_t(a.grid.count) - 1);
                const uint32_t rel = uint32_t(idx) - begin;
                const bool inside = in && rel < ws;
                const uint32_t rr = inside ? rel : 0u;
                const double lgn = v.lg[rr];
                const bool seen_n = v.act[rr] != 0;
                const double log_ratio = lg_cur - lgn;
                const bool up = log_ratio >= 0.0;
                const bool drew = inside && !up; // a uniform is drawn: the next word
                const double lu = readlane_f64(p_lu, L + 1);
                const bool accepted = inside && (up || lu < log_ratio);
                lr_lane = lane == L ? (drew ? log_ratio : __builtin_huge_val()) : lr_lane;
                lr_lane = lane == L + 1 && drew ? lu : lr_lane; // keep log(u) beside it
                next = L + 1 + (drew ? 1u : 0u);
                m_hi = writelane_u32(m_hi, accepted ? ms_hi : 0u, L);
                h_run = fma(accepted ? ms : 0.0, jc[b], h_run);
                const uint32_t s = __builtin_amdgcn_readlane(p_site, L);
                dsig_hi ^= (accepted && p_site == s) ? 0x80000000u : 0u;
                energy = accepted ? proposed : energy;
                cur = accepted ? uint32_t(idx) : cur;
                lg_cur = accepted ? lgn : lg_cur;
                cur_seen = accepted ? seen_n : cur_seen;
                const uint32_t crel = cur - begin;
                if (!cur_seen) { // first visit: update_current_bin (wang_landau.cpp:294-297)
                    ++active;
                    cur_seen = true;
                    v.act[crel] = 1; // every lane writes the same byte
                    if (inverse_time && attempts < freeze) { // later attempts see the new count
                        inc = lane > attempts
                                  ? batch_factors(WLX_STAGE_INVERSE_TIME, factor, active,
                                                  clk0 + done_total + attempts + 1, attempts + 1)
                                  : inc;
                        const uint64_t fz = __ballot(lane > attempts && inc <= a.final_factor);
                        freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : kWave;
                        incz = lane > attempts ? (lane >= freeze ? 0.0 : inc) : incz;
                    }
                }
                lg_cur = lg_cur + readlane_f64(incz, attempts);
                v.lg[crel] = lg_cur; // every lane writes the same value
                cur_lane = writelane_u32(cur_lane, crel, L);
                attempt_lanes |= uint64_t(1) << L;
                ++attempts;
            }
            if (inverse_time && attempts > 0) { // the factor of the last 1/t attempt, a freeze
                factor = readlane_f64(inc, min(freeze, attempts - 1));
                if (freeze < attempts) stage = WLX_STAGE_FROZEN;
            }
            // the batch consumed the words before `next`; from position 63 a new block follows
            const bool mine = (attempt_lanes >> lane) & 1;
            { // accepted proposals: the count and the last one (its attempt index in the batch)
                const uint64_t acc_lanes = __ballot(mine && m_hi != 0u);
                n_acc += uint32_t(__popcll(acc_lanes));
                if (acc_lanes) {
                    const uint32_t top = 63u - uint32_t(__clzll(acc_lanes));
                    const uint64_t below = top ? attempt_lanes & ((uint64_t(1) << top) - 1) : 0;
                    last_acc = done_total + uint32_t(__popcll(below)) + 1u;
                }
            }
            if (mine)
                __hip_atomic_fetch_add(&v.H[cur_lane], uint64_t(1), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            { // drawn decisions whose log(u) (kept in the next lane) lay within 4 ulp
                const double lu_next = __shfl_down(lr_lane, 1);
                const bool amb = mine && lr_lane < 0.0 && near_threshold(lu_next, lr_lane);
                st.ambiguous += __popcll(__ballot(amb));
            }
            if (mine && m_hi != 0)
                __hip_atomic_fetch_xor(&flipbox[p_site & (kWave - 1)], uint64_t(1) << (p_site >> 6),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __syncthreads();
            neg ^= flipbox[lane];
            flipbox[lane] = 0;
            rows_fma<NPL, KB>(a, p_site, m_hi, pos, B, h);
            remaining -= attempts;
            done_total += attempts;
        }
        rng = block_start;
        xoshiro_advance(rng, next);
        asm volatile("" ::: "memory");
        const uint64_t accepted0 = st.accepted, ambiguous = st.ambiguous;
        st = a.state[w];
        st.attempted = att0 + done_total;
        st.accepted = accepted0 + n_acc;
        if (last_acc) st.last_accepted = att0 + last_acc;
        st.stage = stage;
        st.factor = factor;
        st.active_count = active;
        st.ambiguous = ambiguous;
    }
    st.rng[0] = rng.s0;
    st.rng[1] = rng.s1;
    st.rng[2] = rng.s2;
    st.rng[3] = rng.s3;
    __syncthreads();
    finish_walker<NPL>(a, w, st, v, ws, neg, h, energy, flags);
}

// ---------------------------------------------------------------------------------------------
// FP32 local fields (WLX_FIELDS_FP32_KAHAN): the field of site j is a pair (h_j, c_j) of FP32
// numbers with value h_j - c_j. A flip with multiplier m = -2 sigma adds m J32[i][j] to every
// field by Kahan's compensated step (Kahan 1965): y = m J32 - c; t = h + y; c = (t - h) - y;
// h = t. m J32 is exact (a power-of-two scaling). Decisions use the FP64 value
// (double)h - (double)c and energies stay FP64, as do log_g and the histogram. The lookahead
// applies the same step to the pair gathered at each proposal site, in acceptance order, so
// the result equals a one-flip-at-a-time FP32 Kahan walker on the same stream. The drift
// against FP64 is removed by the periodic resync below.
// ---------------------------------------------------------------------------------------------
// FP32 walker layout: lane l holds in slot k the site f32_site(l, k). Each lane owns runs of
// G = min(NPL, 4) consecutive sites, one run in every 64·G sites, so that a row's values and a
// configuration's pairs load G per lane at a time (one dwordx4 per lane per 256 sites once
// NPL >= 4). Only the register placement changes: HBM keeps fields, spins and rows in site order.
template <int NPL> constexpr uint32_t f32_group() { return NPL < 4 ? uint32_t(NPL) : 4u; }
template <int NPL> __device__ inline uint32_t f32_site(uint32_t lane, uint32_t k)
{
    constexpr uint32_t G = f32_group<NPL>();
    return (k / G) * (kWave * G) + lane * G + (k % G);
}
template <int NPL> __device__ inline uint32_t f32_lane(uint32_t site)
{
    constexpr uint32_t G = f32_group<NPL>();
    return (site % (kWave * G)) / G;
}
template <int NPL> __device__ inline uint32_t f32_slot(uint32_t site)
{
    constexpr uint32_t G = f32_group<NPL>();
    return (site / (kWave * G)) * G + site % G;
}

template <int NPL> __device__ inline float select_slot_f32(const float (&h)[NPL], uint32_t k)
{
    float level[NPL];
#pragma unroll
    for (int j = 0; j < NPL; ++j) level[j] = h[j];
#pragma unroll
    for (int width = NPL, bit = 0; width > 1; width >>= 1, ++bit) {
        const bool upper = (k >> bit) & 1;
#pragma unroll
        for (int j = 0; j < width / 2; ++j) level[j] = upper ? level[2 * j + 1] : level[2 * j];
    }
    return level[0];
}

template <int NPL> __device__ inline float gather_f32(const float (&h)[NPL], uint32_t site)
{
    float hv[NPL];
    const int src = int(f32_lane<NPL>(site));
#pragma unroll
    for (int k = 0; k < NPL; ++k) hv[k] = __shfl(h[k], src);
    return select_slot_f32<NPL>(hv, f32_slot<NPL>(site));
}

__device__ inline float readlane_f32(float v, uint32_t lane)
{
    return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), lane));
}

// one compensated addition of `add` to the pair (h, c)
__device__ inline void kahan_add(float &h, float &c, float add)
{
    const float y = add - c;
    const float t = h + y;
    c = (t - h) - y;
    h = t;
}

// kahan_add(h, c, m * r) for m = +-2: the product is exact (a power-of-two scaling), so the fused
// y = m * r - c rounds once, exactly as the separate product and difference do
__device__ inline void kahan_add2(float &h, float &c, float m, float r)
{
    const float y = __builtin_fmaf(m, r, -c);
    const float t = h + y;
    c = (t - h) - y;
    h = t;
}

// slots [k0, k0 + S) of this lane from a row (or configuration) stored in site order, G at a time
template <int NPL, int S, class T>
__device__ inline void load_slots_f32(const T *base, uint32_t lane, uint32_t k0, T (&r)[S])
{
    constexpr uint32_t G = f32_group<NPL>();
    static_assert(S % G == 0, "whole runs");
    if constexpr (G == 4 && sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < S / 4; ++q) {
            const float4 x = *reinterpret_cast<const float4 *>(base + f32_site<NPL>(lane, k0 + 4 * q));
            r[4 * q] = x.x;
            r[4 * q + 1] = x.y;
            r[4 * q + 2] = x.z;
            r[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < S; ++k) r[k] = base[f32_site<NPL>(lane, k0 + uint32_t(k))];
    }
}

template <int NPL>
__device__ inline void load_config32(const Args &a, uint32_t slot, uint64_t &neg,
                                     float (&h)[NPL], float (&c)[NPL])
{
    const uint32_t lane = lane_id();
    const size_t base = size_t(slot) * a.np;
    neg = 0;
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        const uint32_t site = f32_site<NPL>(lane, uint32_t(k));
        const float2 p = a.fields32[base + site];
        h[k] = p.x;
        c[k] = p.y;
        if (a.spins[base + site] < 0) neg |= uint64_t(1) << k;
    }
}

template <int NPL>
__device__ inline void store_config32(const Args &a, uint32_t slot, uint64_t neg,
                                      const float (&h)[NPL], const float (&c)[NPL])
{
    const uint32_t lane = lane_id();
    const size_t base = size_t(slot) * a.np;
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        const uint32_t site = f32_site<NPL>(lane, uint32_t(k));
        a.fields32[base + site] = make_float2(h[k], c[k]);
        a.spins[base + site] = ((neg >> k) & 1) ? int8_t(-1) : int8_t(1);
    }
}

// compensated row updates of the `count` proposals of a batch, in order; rejected ones skipped
template <int NPL, int KB>
__device__ inline void rows_kahan(const Args &a, uint32_t p_site, uint32_t m_hi, uint32_t first,
                                  uint32_t count, float (&h)[NPL], float (&c)[NPL])
{
    const uint32_t lane = lane_id();
    if constexpr (NPL >= 32) {
        // The accepted rows stream in four parts of S slots. Each part's registers are refilled
        // with the next accepted row's part as soon as they are added, so the next row's loads
        // are in flight during this row's additions; after the last row they load that row
        // again, which keeps the loads free of branches. The rows are unrolled up to the batch
        // width. Every site still takes the rows in acceptance order.
        constexpr int P = 4, S = NPL / P;
        uint64_t acc = __ballot(lane >= first && lane < first + count && m_hi != 0u);
        if (acc == 0) return;
        auto pop = [&](const float *&row, float &m) {
            const uint32_t L = uint32_t(__ffsll((unsigned long long)acc) - 1);
            acc &= acc - 1;
            row = a.J32 + size_t(__builtin_amdgcn_readlane(p_site, L)) * a.np;
            m = float(hi_double(__builtin_amdgcn_readlane(m_hi, L)));
        };
        uint32_t rest = uint32_t(__popcll(acc)) - 1; // rows after the current one
        const float *row;
        float m;
        pop(row, m);
        float buf[P][S];
#pragma unroll
        for (int p = 0; p < P; ++p) load_slots_f32<NPL, S>(row, lane, uint32_t(p * S), buf[p]);
        auto step = [&](const float *next) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
#pragma unroll
                for (int s = 0; s < S; ++s) kahan_add2(h[p * S + s], c[p * S + s], m, buf[p][s]);
                load_slots_f32<NPL, S>(next, lane, uint32_t(p * S), buf[p]);
                // the scheduler would otherwise sink the loads below the remaining parts
                __builtin_amdgcn_sched_barrier(0);
            }
        };
#pragma unroll
        for (int r = 0; r < KB; ++r) { // straight-line code, so loads are waited for one by one
            const float *next = row;
            float mn = m;
            if (rest) pop(next, mn);
            step(next);
            if (rest == 0) break;
            --rest;
            row = next;
            m = mn;
        }
    } else {
        constexpr int kRows = (64 / NPL) > 8 ? 8 : (64 / NPL);
#pragma unroll
        for (int g = 0; g < KB; g += kRows) {
            if (uint32_t(g) >= count) break;
            float r[kRows][NPL];
            float m[kRows];
#pragma unroll
            for (int j = 0; j < kRows; ++j) {
                const uint32_t want = first + uint32_t(g + j);
                const uint32_t L = min(want, uint32_t(kWave - 1));
                const uint32_t site = __builtin_amdgcn_readlane(p_site, L);
                const uint32_t mh = __builtin_amdgcn_readlane(m_hi, L);
                m[j] = float(hi_double(want < uint32_t(kWave) && uint32_t(g + j) < count ? mh : 0u));
                load_slots_f32<NPL, NPL>(a.J32 + size_t(site) * a.np, lane, 0u, r[j]);
            }
#pragma unroll
            for (int j = 0; j < kRows; ++j) {
                if (m[j] == 0.0f) continue; // a zero increment still moves c into h: skip it
#pragma unroll
                for (int k = 0; k < NPL; ++k) kahan_add2(h[k], c[k], m[j], r[j][k]);
            }
        }
    }
}

// Parameters by value (the kernel-argument segment), not from a table like the other sampling
// kernels: FP32-field engines never join a batch, and at 64 register slots per lane the table
// copy costs registers (N = 4096 with 256 walkers ran 5 % slower with it).
template <int NPL>
__global__ __launch_bounds__(kWave) void lookahead_f32_kernel(Args a, uint64_t request,
                                                              uint32_t flags)
{
    constexpr int KB = kLookaheadBatch;
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t w = blockIdx.x, lane = lane_id();
    if ((flags & kGate) && a.glob->done) return;
    WalkerState st = a.state[w];
    if (st.escaped) return;
    const uint32_t ws = st.end - st.begin;
    Bins v = bins_view(smem, a.ws_stride);
    uint64_t *flipbox = reinterpret_cast<uint64_t *>(smem + bins_lds_bytes(a.ws_stride));
    load_bins(a, w, ws, v);
    flipbox[lane] = 0;
    uint64_t neg;
    float h[NPL], c[NPL];
    load_config32<NPL>(a, st.slot, neg, h, c);
    double energy = a.energy[st.slot];
    __syncthreads();
    const uint64_t count = batch_count(a, st, request, flags);
    const int64_t start_bin = grid_index(a.grid, energy);
    if (count > 0 && (start_bin < int64_t(st.begin) || start_bin >= int64_t(st.end))) {
        st.escaped = 1; // wang_landau.cpp:247-248
    } else if (count > 0) {
        const uint32_t begin = st.begin;
        const uint64_t att0 = st.attempted, clk0 = att0 + st.clock_offset;
        uint32_t cur = uint32_t(start_bin);
        double lg_cur = v.lg[cur - begin];
        bool cur_seen = v.act[cur - begin] != 0;
        uint32_t stage = st.stage, active = st.active_count;
        double factor = st.factor;
        uint32_t done_total = 0, n_acc = 0, last_acc = 0;
        uint32_t p_site = 0, used = kWave;
        double p_lu = 0.0;
        uint64_t remaining = count;
        while (remaining > 0) {
            if (used == kWave) {
                const uint64_t t = att0 + done_total + lane;
                const Philox4 blk = philox4x64(t >> 1, kPhiloxWalkTag, 0, 0, a.seed, uint64_t(w));
                const bool odd = t & 1;
                p_site = uint32_t(__umul64hi(odd ? blk.v[2] : blk.v[0], uint64_t(a.n)));
                const double u = double((odd ? blk.v[3] : blk.v[1]) >> 11) * 0x1.0p-53;
                p_lu = log(fmax(u, 0x1.0p-53));
                used = 0;
            }
            const uint32_t rem32 = remaining < uint64_t(kWave) ? uint32_t(remaining) : kWave;
            const uint32_t K = min(min(uint32_t(KB), rem32), kWave - used);
            float hr = gather_f32<NPL>(h, p_site), cr = gather_f32<NPL>(c, p_site);
            const uint32_t src = f32_lane<NPL>(p_site);
            const uint64_t bits =
                (uint64_t(uint32_t(__shfl(int(uint32_t(neg >> 32)), int(src)))) << 32) |
                uint32_t(__shfl(int(uint32_t(neg)), int(src)));
            uint32_t dsig_hi = ((bits >> f32_slot<NPL>(p_site)) & 1) ? 0x40000000u : 0xC0000000u;
            float jc[KB];
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const uint32_t sq = __builtin_amdgcn_readlane(
                    p_site, min(used + uint32_t(q), uint32_t(kWave - 1)));
                jc[q] = a.J32[size_t(sq) * a.np + p_site];
            }
            // log_g increments, lane-parallel, 0 from the freezing 1/t attempt on (fast_walk)
            const bool inverse_time = stage == WLX_STAGE_INVERSE_TIME;
            double inc = batch_factors(stage, factor, active, clk0 + done_total, used);
            uint32_t freeze = kWave;
            if (inverse_time) {
                const uint64_t fz = __ballot(lane >= used && inc <= a.final_factor);
                freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : kWave;
            }
            double incz = lane >= freeze ? 0.0 : inc;
            uint32_t m_hi = 0, cur_lane = 0;
            double lr_lane = 0.0;
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                if (uint32_t(b) >= K) continue;
                const uint32_t L = used + uint32_t(b);
                const double hs =
                    double(readlane_f32(hr, L)) - double(readlane_f32(cr, L)); // field value
                const uint32_t ms_hi = __builtin_amdgcn_readlane(dsig_hi, L);
                const double ms = hi_double(ms_hi); // -2 sigma
                const double proposed = energy + ms * hs;
                const bool in = proposed >= a.grid.minimum && proposed <= a.grid.maximum;
                const double fb = bin_floor(a.grid, proposed);
                int32_t idx = int32_t(in ? fb : 0.0);
                if (proposed == a.grid.maximum) idx = min(idx, int32_t(a.grid.count) - 1);
                const uint32_t rel = uint32_t(idx) - begin;
                const bool inside = in && rel < ws;
                const uint32_t rr = inside ? rel : 0u;
                const double lgn = v.lg[rr];
                const bool seen_n = v.act[rr] != 0;
                const double log_ratio = lg_cur - lgn;
                const double lu = readlane_f64(p_lu, L);
                const bool up = log_ratio >= 0.0;
                const bool accepted = inside && (up || lu < log_ratio);
                lr_lane = lane == L ? (inside ? log_ratio : __builtin_huge_val()) : lr_lane;
                m_hi = writelane_u32(m_hi, accepted ? ms_hi : 0u, L);
                if (accepted) { // the flip's compensated step on every proposal's pair
                    float hn = hr, cn = cr;
                    kahan_add(hn, cn, float(ms) * jc[b]);
                    hr = hn;
                    cr = cn;
                }
                const uint32_t s = __builtin_amdgcn_readlane(p_site, L);
                dsig_hi ^= (accepted && p_site == s) ? 0x80000000u : 0u;
                energy = accepted ? proposed : energy;
                cur = accepted ? uint32_t(idx) : cur;
                lg_cur = accepted ? lgn : lg_cur;
                cur_seen = accepted ? seen_n : cur_seen;
                const uint32_t crel = cur - begin;
                if (!cur_seen) { // first visit: update_current_bin (wang_landau.cpp:294-297)
                    ++active;
                    cur_seen = true;
                    v.act[crel] = 1; // every lane writes the same byte
                    if (inverse_time && L < freeze) { // later attempts see the new count
                        inc = lane > L ? batch_factors(WLX_STAGE_INVERSE_TIME, factor, active,
                                                       clk0 + done_total + b + 1, L + 1)
                                       : inc;
                        const uint64_t fz = __ballot(lane > L && inc <= a.final_factor);
                        freeze = fz ? uint32_t(__ffsll((unsigned long long)fz) - 1) : kWave;
                        incz = lane > L ? (lane >= freeze ? 0.0 : inc) : incz;
                    }
                }
                lg_cur = lg_cur + readlane_f64(incz, L);
                v.lg[crel] = lg_cur; // every lane writes the same value
                cur_lane = writelane_u32(cur_lane, crel, L);
            }
            if (inverse_time) { // the factor of the last 1/t attempt, and a freeze
                const uint32_t last_lane = used + K - 1;
                factor = readlane_f64(inc, min(freeze, last_lane));
                if (freeze <= last_lane) stage = WLX_STAGE_FROZEN;
            }
            const uint32_t last = used + K - 1;
            const bool mine = lane >= used && lane <= last;
            { // accepted proposals of the batch: the count and the last one
                const uint64_t acc_lanes = __ballot(mine && m_hi != 0u);
                n_acc += uint32_t(__popcll(acc_lanes));
                if (acc_lanes)
                    last_acc = done_total + (63u - uint32_t(__clzll(acc_lanes))) - used + 1u;
            }
            if (mine)
                __hip_atomic_fetch_add(&v.H[cur_lane], uint64_t(1), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            {
                const bool amb = mine && lr_lane < 0.0 && near_threshold(p_lu, lr_lane);
                st.ambiguous += __popcll(__ballot(amb));
            }
            if (mine && m_hi != 0)
                __hip_atomic_fetch_xor(&flipbox[f32_lane<NPL>(p_site)],
                                       uint64_t(1) << f32_slot<NPL>(p_site), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            __syncthreads();
            neg ^= flipbox[lane];
            flipbox[lane] = 0;
            rows_kahan<NPL, KB>(a, p_site, m_hi, used, K, h, c);
            used += K;
            remaining -= K;
            done_total += K;
        }
        asm volatile("" ::: "memory");
        const uint64_t accepted0 = st.accepted, ambiguous = st.ambiguous;
        st = a.state[w];
        st.attempted = att0 + done_total;
        st.accepted = accepted0 + n_acc;
        if (last_acc) st.last_accepted = att0 + last_acc;
        st.stage = stage;
        st.factor = factor;
        st.active_count = active;
        st.ambiguous = ambiguous;
    }
    __syncthreads();
    store_config32<NPL>(a, st.slot, neg, h, c);
    finish_walker_state(a, w, st, v, ws, energy, flags);
}

// ---------------------------------------------------------------------------------------------
// FP64 resync of FP32 fields (kernel c): H = J.S for every configuration slot on the matrix
// cores, energies E = 1/2 sum_i s_i H_i, the drift of the running FP32 state against them, and
// a restart of the pairs from H (h = fl32(H), c = fl32(h - H)).
// ---------------------------------------------------------------------------------------------
struct Drift {
    unsigned long long max_field;  // bit pattern of a non-negative double
    unsigned long long max_energy; // bit pattern of a non-negative double (energy drift in bins)
};

// S[k][s] = spin of site k in slot s, as +-1.0 (zero for padding slots s >= slots)
__global__ void resync_spins_kernel(const int8_t *spins, uint32_t np, uint32_t slots,
                                    uint32_t slots_padded, double *S)
{
    const size_t total = size_t(np) * slots_padded;
    for (size_t x = size_t(blockIdx.x) * blockDim.x + threadIdx.x; x < total;
         x += size_t(gridDim.x) * blockDim.x) {
        const uint32_t k = uint32_t(x / slots_padded), s = uint32_t(x % slots_padded);
        S[x] = s < slots ? (spins[size_t(s) * np + k] < 0 ? -1.0 : 1.0) : 0.0;
    }
}

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

// the resync's K steps are split in this many parts, each summed into its own partial of H;
// a wave computes 16·kResyncTileX sites x 16·kResyncTileY slots, kResyncSteps K steps a round
constexpr uint32_t kResyncParts = 8;
constexpr int kResyncTileX = 4, kResyncTileY = 4, kResyncSteps = 2;

// H[s][i] = sum_k J[k][i] S[k][s] with v_mfma_f64_16x16x4_f64: one wave per 16·kResyncTileX
// sites x 16·kResyncTileY slots, kResyncTileX x kResyncTileY accumulator blocks of 16x16, K
// stepped by 4. J is bitwise symmetric, so the A operand J[i][k] is read as J[k][i]
// (consecutive lanes, consecutive addresses). Operand layout of the instruction on gfx950: lane
// l holds A[l % 16][l / 16], B[l / 16][l % 16] and D[4 r + l / 16][l % 16] in accumulator
// register r (rows interleaved across the four lane groups; checked with constant-row and
// constant-column operands). J has n rows of np columns: a K step k >= n reads row n - 1 and
// enters with zero spins.
__global__ __launch_bounds__(kWave) void resync_mfma_kernel(const double *J, const double *S,
                                                           uint32_t n, uint32_t np,
                                                           uint32_t slots_padded, uint32_t slots,
                                                           double *H)
{
    constexpr int TX = kResyncTileX, TY = kResyncTileY;
    const uint32_t lane = threadIdx.x;
    const uint32_t tiles_i = np / (16 * TX), tiles = tiles_i * (slots_padded / (16 * TY));
    const uint32_t part = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const uint32_t ti = tile % tiles_i, ts = tile / tiles_i;
    const uint32_t i0 = ti * 16 * TX, s0 = ts * 16 * TY;
    const uint32_t row = lane % 16, kk = lane / 16;
    mfma_d4 acc[TX][TY];
#pragma unroll
    for (int x = 0; x < TX; ++x)
#pragma unroll
        for (int y = 0; y < TY; ++y) acc[x][y] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    // kResyncSteps K steps per round. Rounds are double-buffered: the loads of round r + 1 are
    // issued before the MFMAs of round r, so the matrix core keeps busy while the next operands
    // arrive. Every load is unconditional (a conditional one compiles to a branch, after which
    // the loads in flight are all waited for): a step outside [kb, ke) reads a valid row and
    // multiplies its spins by 0, so it adds nothing. Block part p sums the K steps of its part
    // into its own partial of H (kResyncParts partials, added in order by the finish kernel),
    // which puts two waves on every SIMD at N = 4096 with 256 configurations.
    constexpr int kSteps = kResyncSteps;
    constexpr uint32_t kRound = 4 * kSteps;
    const uint32_t span = (n + kResyncParts * 2 * kRound - 1) / (kResyncParts * 2 * kRound) * 2 * kRound;
    const uint32_t kb = min(n, part * span), ke = min(n, kb + span);
    double av[2][TX][kSteps], bv[2][TY][kSteps];
    auto load = [&](uint32_t k0, int buf) {
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
            const uint32_t k = k0 + 4 * uint32_t(u) + kk;
            const uint32_t kc = min(k, n - 1);
            const double *jrow = J + size_t(kc) * np + i0 + row;
            const double *srow = S + size_t(kc) * slots_padded + s0 + row;
#pragma unroll
            for (int x = 0; x < TX; ++x) av[buf][x][u] = jrow[16 * x];
#pragma unroll
            for (int y = 0; y < TY; ++y) bv[buf][y][u] = srow[16 * y];
        }
    };
    // the zeroing of steps outside [kb, ke) happens here, after the scheduling barrier, so that
    // it does not wait for the loads just issued
    auto multiply = [&](uint32_t k0, int buf) {
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
            const double keep = k0 + 4 * uint32_t(u) + kk < ke ? 1.0 : 0.0;
            double c[TY];
#pragma unroll
            for (int y = 0; y < TY; ++y) c[y] = bv[buf][y][u] * keep;
#pragma unroll
            for (int x = 0; x < TX; ++x)
#pragma unroll
                for (int y = 0; y < TY; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[buf][x][u], c[y], acc[x][y],
                                                                     0, 0, 0);
        }
    };
    // whole pairs of rounds, the steps past ke zeroed by `keep`: no exit between the two halves,
    // which would make the compiler keep a second copy of the accumulators
    load(kb, 0);
    for (uint32_t k0 = kb; k0 < ke; k0 += 2 * kRound) {
        load(k0 + kRound, 1);
        __builtin_amdgcn_sched_barrier(0); // the loads stay ahead of the MFMAs
        multiply(k0, 0);
        load(k0 + 2 * kRound, 0);
        __builtin_amdgcn_sched_barrier(0);
        multiply(k0 + kRound, 1);
    }
#pragma unroll
    for (int x = 0; x < TX; ++x)
#pragma unroll
        for (int y = 0; y < TY; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t i = i0 + 16 * x + 4 * r + kk, s = s0 + 16 * y + row;
                if (s < slots) H[(size_t(part) * slots + s) * np + i] = acc[x][y][r];
            }
}

// one wave per slot: drift of the running pairs and energy, then restart from H
// kResyncFinishThreads threads per slot (four waves, so that the partial sums' loads of many
// sites are in flight at once)
constexpr uint32_t kResyncFinishThreads = 256;

__global__ __launch_bounds__(kResyncFinishThreads) void resync_finish_kernel(Args a,
                                                                            const double *H,
                                                                            Drift *drift)
{
    __shared__ double part[kResyncFinishThreads];
    __shared__ double err[kResyncFinishThreads / kWave];
    const uint32_t s = blockIdx.x, t = threadIdx.x;
    const size_t base = size_t(s) * a.np;
    double field_err = 0.0, e = 0.0;
    const size_t stride = size_t(a.walkers) * a.np; // between the partials of H
    for (uint32_t i = t; i < a.n; i += kResyncFinishThreads) {
        double hv = H[base + i];
#pragma unroll
        for (uint32_t q = 1; q < kResyncParts; ++q) hv += H[q * stride + base + i];
        const float2 p = a.fields32[base + i];
        field_err = fmax(field_err, fabs((double(p.x) - double(p.y)) - hv));
        const double sigma = a.spins[base + i] < 0 ? -1.0 : 1.0;
        e += 0.5 * sigma * hv;
        const float hf = float(hv);
        a.fields32[base + i] = make_float2(hf, float(double(hf) - hv));
    }
    part[t] = e;
    for (int m = 1; m < kWave; m <<= 1) field_err = fmax(field_err, __shfl_xor(field_err, m));
    if (lane_id() == 0) err[t / kWave] = field_err;
    __syncthreads();
    for (uint32_t width = kResyncFinishThreads / 2; width > 0; width >>= 1) {
        if (t < width) part[t] += part[t + width];
        __syncthreads();
    }
    if (t == 0) {
        for (uint32_t w = 1; w < kResyncFinishThreads / kWave; ++w) field_err = fmax(field_err, err[w]);
        const double e64 = part[0];
        const double rel = fabs(a.energy[s] - e64) / a.grid.width; // drift in bins
        atomicMax(&drift->max_field, (unsigned long long)__double_as_longlong(field_err));
        atomicMax(&drift->max_energy, (unsigned long long)__double_as_longlong(rel));
        a.energy[s] = e64;
    }
}

// FP64 fields (from the initial local_fields) into (h, c) pairs
__global__ void narrow_fields_kernel(const double *fields, float2 *fields32, size_t count)
{
    for (size_t x = size_t(blockIdx.x) * blockDim.x + threadIdx.x; x < count;
         x += size_t(gridDim.x) * blockDim.x) {
        const double hv = fields[x];
        const float hf = float(hv);
        fields32[x] = make_float2(hf, float(double(hf) - hv));
    }
}

// ---------------------------------------------------------------------------------------------
// init: walker construction + target-Metropolis initialisation (wang_landau.cpp:91-242)
// ---------------------------------------------------------------------------------------------
template <int NPL>
__global__ __launch_bounds__(kWave) void init_kernel(Args a, const int8_t *initial,
                                                     const uint8_t *supplied_mask, InitParams ip,
                                                     InitReport *report)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t w = blockIdx.x, lane = lane_id();
    WalkerState st = a.state[w];
    const uint32_t ws = st.end - st.begin;
    Bins v = a.bins_hbm ? bins_hbm_view(a, w) : bins_view(smem, a.ws_stride);
    double *scratch =
        reinterpret_cast<double *>(smem + (a.bins_hbm ? 0 : bins_lds_bytes(a.ws_stride)));
    for (uint32_t b = lane; b < ws; b += kWave) {
        v.lg[b] = 0.0;
        v.H[b] = 0;
        v.act[b] = 0;
        v.ref[b] = 0;
    }
    InitReport rep{};
    const bool supplied = supplied_mask && supplied_mask[w];
    uint64_t neg = 0;
    bool bad_spin = false;
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        const uint32_t i = k * kWave + lane;
        const int8_t value = (supplied && i < a.n) ? initial[size_t(w) * a.n + i] : int8_t(1);
        if (value != 1 && value != -1) bad_spin = true;
        if (value < 0) neg |= uint64_t(1) << k;
    }
    if (__any(bad_spin)) {
        if (lane == 0) {
            rep.status = 2;
            report[w] = rep;
        }
        return;
    }
    Xoshiro rng = Xoshiro::seeded(a.seed, w); // rng_(master_seed, walker_id)
    st.factor = 1.0;
    st.last_flatness = 0.0;
    st.attempted = st.accepted = st.forced = st.last_accepted = 0;
    st.last_check = 0;
    st.ambiguous = 0;
    st.stage = WLX_STAGE_WANG_LANDAU;
    st.active_count = st.refine_count = 0;
    st.trace_count = 0;
    st.escaped = 0;
    st.clock_offset = 0;
    double h[NPL];
    compute_fields<NPL>(a, neg, h);
    double energy = total_energy<NPL>(a, neg, h, scratch);

    const int64_t begin = st.begin, end = st.end;
    auto in_window = [&](int64_t b) { return b >= begin && b < end; };
    bool placed = false;
    if (supplied && in_window(grid_index(a.grid, energy))) {
        placed = true;
    } else {
        const double lower = a.grid.minimum + double(st.begin) * a.grid.width;
        const double upper = a.grid.minimum + double(st.end) * a.grid.width;
        const double span = upper - lower;
        const double target = 0.5 * (lower + upper);
        const double target_fraction = supplied ? 1.0 : ip.target_fraction;
        const double half = 0.5 * target_fraction * span;
        const double tf = ip.temperature_fraction * span;
        const double t0 = a.grid.width < tf ? tf : a.grid.width;
        const double tm = ip.max_temperature_fraction * span;
        const double tmax = t0 < tm ? tm : t0;
        double temperature = t0;
        const uint64_t stall = ip.stall_per_spin > ~uint64_t(0) / a.n ? ~uint64_t(0)
                                                                       : ip.stall_per_spin * a.n;
        uint64_t since = 0;
        rep.highest = temperature;
        rep.min_reached = rep.max_reached = rep.closest = energy;
        rep.closest_distance = fabs(energy - target);
        auto record = [&](double c) {
            if (c < rep.min_reached) rep.min_reached = c;
            if (rep.max_reached < c) rep.max_reached = c;
            const double d = fabs(c - target);
            if (d < rep.closest_distance) {
                rep.closest_distance = d;
                rep.closest = c;
                return true;
            }
            return false;
        };
        auto in_band = [&](double c) {
            return in_window(grid_index(a.grid, c)) && fabs(c - target) <= half;
        };
        const Bounded two = make_bounded(2);
        for (uint64_t step = 0; step < ip.max_attempts && !placed; ++step) {
            if (in_band(energy)) {
                placed = true;
                break;
            }
            if (since >= stall) {
                if (temperature < tmax) {
                    const double raised = temperature * ip.temperature_multiplier;
                    temperature = raised < tmax ? raised : tmax;
                    if (rep.highest < temperature) rep.highest = temperature;
                    ++rep.increases;
                } else { // reproducible random restart, wang_landau.cpp:187-193
                    neg = 0;
                    for (uint32_t i = 0; i < a.n; ++i) {
                        const bool down = draw_bounded(rng, two) == 0;
                        if (down && lane == (i & (kWave - 1))) neg |= uint64_t(1) << (i >> 6);
                    }
                    compute_fields<NPL>(a, neg, h);
                    energy = total_energy<NPL>(a, neg, h, scratch);
                    (void)record(energy);
                    temperature = t0;
                    ++rep.restarts;
                }
                since = 0;
                if (in_band(energy)) {
                    placed = true;
                    break;
                }
            }
            const uint32_t s = uint32_t(draw_bounded(rng, a.bounded));
            const double hs = readlane_f64(select_slot<NPL>(h, s >> 6), s & (kWave - 1));
            const uint64_t bits = readlane_u64(neg, s & (kWave - 1));
            const double sigma = ((bits >> (s >> 6)) & 1) ? -1.0 : 1.0;
            const double delta = -2.0 * sigma * hs;
            const double proposed = energy + delta;
            const double log_acceptance =
                (fabs(energy - target) - fabs(proposed - target)) / temperature;
            bool improved = false;
            if (log_acceptance >= 0.0 ||
                log(fmax(rng.uniform(), 0x1.0p-53)) < log_acceptance) {
                apply_flip<NPL>(a, s, sigma, neg, h);
                energy = proposed;
                improved = record(energy);
            }
            if (improved)
                since = 0;
            else if (since < ~uint64_t(0))
                ++since;
        }
        if (!placed && in_band(energy)) placed = true;
    }
    if (placed) {
        __syncthreads();
        update_current_bin(a, st, v, uint32_t(grid_index(a.grid, energy) - begin));
        if (a.collect) { // wang_landau.cpp:123-127, 177, 197, 219
            const SamplingView col = sampling_view(a, smem);
            load_sampling(a, w, ws, col);
            __syncthreads();
            round_trip_and_representatives<NPL>(a, col, w, st.begin, st.end, energy, neg);
            __syncthreads();
            store_sampling(a, w, ws, col);
        }
    } else {
        rep.status = 1;
    }
    st.rng[0] = rng.s0;
    st.rng[1] = rng.s1;
    st.rng[2] = rng.s2;
    st.rng[3] = rng.s3;
    __syncthreads();
    if (!a.bins_hbm) store_bins(a, w, ws, v);
    store_config<NPL>(a, st.slot, neg, h);
    if (lane == 0) {
        a.energy[st.slot] = energy;
        a.state[w] = st;
        report[w] = rep;
    }
}

// local_fields + total_energy for stored configurations (one wave per configuration)
template <int NPL>
__global__ __launch_bounds__(kWave) void fields_kernel(Args a, const int8_t *spins,
                                                       double *fields, double *energies)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t c = blockIdx.x, lane = lane_id();
    uint64_t neg = 0;
#pragma unroll
    for (int k = 0; k < NPL; ++k)
        if (spins[size_t(c) * a.np + k * kWave + lane] < 0) neg |= uint64_t(1) << k;
    double h[NPL];
    compute_fields<NPL>(a, neg, h);
    const double e = total_energy<NPL>(a, neg, h, reinterpret_cast<double *>(smem));
#pragma unroll
    for (int k = 0; k < NPL; ++k) fields[size_t(c) * a.np + k * kWave + lane] = h[k];
    if (lane == 0) energies[c] = e;
}

// wl_run's --pilot sweep (wl_run.cpp:86-97): unconditional flips from all spins +1
template <int NPL>
__global__ __launch_bounds__(kWave) void pilot_kernel(Args a, uint64_t seed, uint64_t attempts,
                                                      double *range)
{
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t neg = 0;
    double h[NPL];
    compute_fields<NPL>(a, neg, h);
    double energy = total_energy<NPL>(a, neg, h, reinterpret_cast<double *>(smem));
    double lo = energy, hi = energy;
    Xoshiro rng = Xoshiro::seeded(seed, 0);
    for (uint64_t t = 0; t < attempts; ++t) {
        const uint32_t s = uint32_t(draw_bounded(rng, a.bounded));
        const double hs = readlane_f64(select_slot<NPL>(h, s >> 6), s & (kWave - 1));
        const uint64_t bits = readlane_u64(neg, s & (kWave - 1));
        const double sigma = ((bits >> (s >> 6)) & 1) ? -1.0 : 1.0;
        energy += -2.0 * sigma * hs;
        apply_flip<NPL>(a, s, sigma, neg, h);
        lo = energy < lo ? energy : lo;
        hi = hi < energy ? energy : hi;
    }
    if (lane_id() == 0) {
        range[0] = lo;
        range[1] = hi;
    }
}

// The same sweep for many systems at once, one workgroup per system, with the fields and spins
// in LDS (any N): a flip scatters the CSR row of a CSR system or streams the dense row.
struct PilotJob {
    const double *J;
    const uint64_t *offsets; // null for dense systems
    const uint32_t *columns;
    const double *values;
    uint32_t n, np;
    Bounded bounded;
    uint64_t seed, attempts;
    double *range;
};

__global__ __launch_bounds__(kWave) void pilot_lds_kernel(const PilotJob *jobs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const PilotJob j = jobs[blockIdx.x];
    const uint32_t lane = lane_id();
    double *hf = reinterpret_cast<double *>(smem);
    int8_t *sp = reinterpret_cast<int8_t *>(hf + j.np);
    // local_fields of all spins +1 (physics.cpp:238-251): h_i = sum_j J_ij in column order; the
    // CSR row holds the nonzero columns in that order, and the skipped zeros add nothing
    for (uint32_t i = lane; i < j.n; i += kWave) {
        double h = 0.0;
        if (j.offsets) {
            for (uint64_t k = j.offsets[i]; k < j.offsets[i + 1]; ++k) h += j.values[k];
        } else {
            for (uint32_t c = 0; c < j.n; ++c) h += j.J[size_t(i) * j.np + c];
        }
        hf[i] = h;
        sp[i] = 1;
    }
    __syncthreads();
    double energy = 0.0; // total_energy (physics.cpp:253-259), site order
    for (uint32_t i = 0; i < j.n; ++i) energy += 0.5 * hf[i];
    double lo = energy, hi = energy;
    Xoshiro rng = Xoshiro::seeded(j.seed, 0);
    for (uint64_t t = 0; t < j.attempts; ++t) {
        const uint32_t s = uint32_t(draw_bounded(rng, j.bounded));
        const double sigma = sp[s] < 0 ? -1.0 : 1.0;
        energy += -2.0 * sigma * hf[s];
        const double m = -2.0 * sigma;
        if (j.offsets) {
            const uint64_t row_end = j.offsets[s + 1];
            for (uint64_t k = j.offsets[s] + lane; k < row_end; k += kWave) {
                const uint32_t c = j.columns[k];
                hf[c] = fma(m, j.values[k], hf[c]);
            }
        } else {
            const double *row = j.J + size_t(s) * j.np;
            for (uint32_t c = lane; c < j.n; c += kWave) hf[c] = fma(m, row[c], hf[c]);
        }
        if (lane == 0) sp[s] = sigma < 0.0 ? int8_t(1) : int8_t(-1);
        __syncthreads();
        lo = energy < lo ? energy : lo;
        hi = hi < energy ? energy : hi;
    }
    if (lane == 0) {
        j.range[0] = lo;
        j.range[1] = hi;
    }
}

// ---------------------------------------------------------------------------------------------
// exchange: the even/odd schedule of parallel.cpp:410-421. Each walker is in at most one pair
// per epoch, so all pairs are independent; a swap exchanges two slot indices.
// ---------------------------------------------------------------------------------------------
constexpr int kExchangeBlock = 256;

__device__ inline void exchange_epoch(const Args &a, uint32_t gated, uint64_t epoch_arg)
{
    // termination flags (parallel.cpp:391-393, 527-530): a swap moves configurations, not
    // stages or attempt counts, so they are read first, alongside the gate and epoch
    uint64_t not_frozen = 0, below_limit = 0;
    for (uint32_t k = threadIdx.x; k < a.walkers; k += blockDim.x) {
        not_frozen += a.state[k].stage != WLX_STAGE_FROZEN;
        below_limit += a.state[k].attempted < a.max_attempts;
    }
    if (gated && a.glob->done) return;
    const uint64_t epoch = gated ? a.glob->epoch : epoch_arg;
    const uint32_t parity = uint32_t(epoch & 1);
    const uint32_t P = a.per_window;
    const uint32_t boundaries = a.windows > parity + 1 ? (a.windows - parity) / 2 : 0;
    const uint64_t total = uint64_t(boundaries) * P;
    uint64_t taken = 0, ambiguous = 0;
    for (uint64_t t = threadIdx.x; t < total; t += blockDim.x) {
        const uint32_t b = parity + 2 * uint32_t(t / P), p = uint32_t(t % P);
        const uint32_t L = b * P + p, R = (b + 1) * P + uint32_t((p + epoch) % P);
        WalkerState &l = a.state[L];
        WalkerState &r = a.state[R];
        const double el = a.energy[l.slot], er = a.energy[r.slot];
        const int64_t x = grid_index(a.grid, el), y = grid_index(a.grid, er);
        auto inside = [](const WalkerState &s, int64_t bin) {
            return bin >= int64_t(s.begin) && bin < int64_t(s.end);
        };
        // exchange_log_probability (wang_landau.cpp:497-508)
        if (!inside(l, x) || !inside(l, y) || !inside(r, x) || !inside(r, y)) continue;
        const double *gl = a.logg + size_t(L) * a.ws_stride - l.begin;
        const double *gr = a.logg + size_t(R) * a.ws_stride - r.begin;
        const double lnp = gl[x] - gl[y] + gr[y] - gr[x];
        if (!(lnp - lnp == 0.0)) continue; // not finite
        bool accept = lnp >= 0.0;
        if (!accept) { // local_exchange (parallel.cpp:157-166)
            double u;
            if (a.philox) {
                const Philox4 blk = philox4x64(epoch, kPhiloxExchangeTag, 0, 0, a.seed, b);
                u = double(blk.v[0] >> 11) * 0x1.0p-53;
            } else {
                Xoshiro rng = Xoshiro::seeded(a.seed ^ (epoch * kGolden), b);
                u = rng.uniform();
            }
            const double lu = log(fmax(u, 0x1.0p-53));
            accept = lu < lnp;
            ambiguous += near_threshold(lu, lnp);
        }
        if (accept) {
            const uint32_t slot = l.slot;
            l.slot = r.slot;
            r.slot = slot;
            ++taken;
            if (a.collect) { // the walkers now hold energies er (L) and el (R)
                exchange_statistics(a, L, l, l.slot, er);
                exchange_statistics(a, R, r, r.slot, el);
            }
        }
    }
    uint64_t vals[4] = {taken, ambiguous, not_frozen, below_limit};
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < 4; ++q) vals[q] += shfl_xor_u64(vals[q], off);
    __shared__ uint64_t sums[kExchangeBlock / kWave][4];
    const uint32_t wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0)
        for (int q = 0; q < 4; ++q) sums[wave][q] = vals[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t[4] = {0, 0, 0, 0};
        for (int wv = 0; wv < kExchangeBlock / kWave; ++wv)
            for (int q = 0; q < 4; ++q) t[q] += sums[wv][q];
        a.glob->exchange_attempted += total;
        a.glob->exchange_accepted += t[0];
        a.glob->exchange_ambiguous += t[1];
        if (gated) {
            const bool frozen = t[2] == 0;
            const bool limit = !a.unlimited && t[3] == 0;
            a.glob->done = (frozen || limit) ? 1 : 0;
            a.glob->epoch = epoch + 1;
        }
    }
}

__global__ __launch_bounds__(kExchangeBlock) void exchange_kernel(Args a, uint32_t gated,
                                                                  uint64_t epoch_arg)
{
    exchange_epoch(a, gated, epoch_arg);
}

// ---------------------------------------------------------------------------------------------
// batches (wlx_batch): the walkers of several engines in one sampling launch (the sampling
// kernels above with the batch's Args table and slot list) and their replica exchanges in one
// launch (one workgroup per engine). Each engine keeps its own gate, epoch and counters.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kExchangeBlock) void batch_exchange_kernel(const Args *table)
{
    const Args a = table[blockIdx.x];
    exchange_epoch(a, 1, 0);
}

// per-walker operations for the harness: 0 statistics, 1 ready, 2 begin_next_iteration
__global__ __launch_bounds__(kWave) void walker_op_kernel(Args a, uint32_t w, uint32_t op,
                                                          uint64_t *out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    WalkerState st = a.state[w];
    const uint32_t ws = st.end - st.begin;
    Bins v = a.bins_hbm ? bins_hbm_view(a, w) : bins_view(smem, a.ws_stride);
    if (!a.bins_hbm) load_bins(a, w, ws, v);
    __syncthreads();
    const Stats s = statistics(a, st, v, ws);
    const bool ready = ready_for_iteration(a, st, s);
    if (op == 2 && ready) next_iteration(a, st, v, ws);
    __syncthreads();
    if (op == 2 && ready && !a.bins_hbm) store_bins(a, w, ws, v);
    if (lane_id() == 0) {
        if (op == 2 && ready) a.state[w] = st;
        out[0] = s.active;
        out[1] = s.covered;
        out[2] = s.minimum;
        out[3] = __double_as_longlong(s.mean);
        out[4] = __double_as_longlong(s.min_over_mean);
        out[5] = __double_as_longlong(s.coverage);
        out[6] = ready ? 1 : 0;
    }
}

// ---- dispatch ------------------------------------------------------------------------------
using SampleFn = void (*)(const Args *, const BatchSlot *, uint64_t, uint32_t);
using F32Fn = void (*)(Args, uint64_t, uint32_t);
using InitFn = void (*)(Args, const int8_t *, const uint8_t *, InitParams, InitReport *);
using FieldsFn = void (*)(Args, const int8_t *, double *, double *);
using PilotFn = void (*)(Args, uint64_t, uint64_t, double *);

// lookahead widths of the Philox common-configuration kernel, and the index of the default
constexpr uint32_t kLookaheadWidths[] = {4, 8, 16, 32};
constexpr int kDefaultWidth = 2;

struct KernelSet {
    SampleFn sample_xo, sample_xo_fast, sample_ph, sample_ph_seq;
    SampleFn ph_fast_width[4]; // lookahead_fast_kernel<NPL, kLookaheadWidths[k]>
    SampleFn lds_xo, lds_ph; // any N
    SampleFn staged; // lookahead_staged_kernel<NPL> (NPL 2, 4, 8), else null
    F32Fn sample_f32;
    InitFn init;
    FieldsFn fields;
    PilotFn pilot;
};

template <int NPL> KernelSet kernel_set()
{
    return {sample_kernel<NPL, false>, lookahead_xo_kernel<NPL>, lookahead_kernel<NPL>,
            sample_kernel<NPL, true>,
            {lookahead_fast_kernel<NPL, 4>, lookahead_fast_kernel<NPL, 8>,
             lookahead_fast_kernel<NPL, 16>, lookahead_fast_kernel<NPL, 32>},
            lds_sample_kernel<false>, lds_sample_kernel<true>,
            (NPL >= 2 && NPL <= 8) ? lookahead_staged_kernel<(NPL >= 2 && NPL <= 8 ? NPL : 2)>
                                   : nullptr,
            lookahead_f32_kernel<NPL>, init_kernel<NPL>, fields_kernel<NPL>, pilot_kernel<NPL>};
}

KernelSet kernels_for(uint32_t npl)
{
    switch (npl) {
    case 1: return kernel_set<1>();
    case 2: return kernel_set<2>();
    case 4: return kernel_set<4>();
    case 8: return kernel_set<8>();
    case 16: return kernel_set<16>();
    case 32: return kernel_set<32>();
    default: return kernel_set<64>();
    }
}

// raises a kernel's dynamic-LDS limit to `bytes`; never lowers it, since the kernels are
// shared by every engine and batch of the process
int set_lds_limit(const void *fn, size_t bytes)
{
    if (bytes <= 64 * 1024) return WLX_OK;
    static std::mutex guard;
    static std::map<const void *, size_t> limits;
    std::lock_guard<std::mutex> lock(guard);
    size_t &limit = limits[fn];
    if (bytes <= limit) return WLX_OK;
    REWL_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
    limit = bytes;
    return WLX_OK;
}

int system_fields_device(wlx_system *s, const int8_t *d_spins, uint64_t count, double *d_fields,
                         double *d_energy, hipStream_t stream)
{
    Args a{};
    a.J = s->d_J64;
    a.n = uint32_t(s->n);
    a.np = uint32_t(s->np);
    const KernelSet k = kernels_for(s->npl);
    const size_t lds = s->np * sizeof(double);
    if (int status = set_lds_limit(reinterpret_cast<const void *>(k.fields), lds)) return status;
    hipLaunchKernelGGL(k.fields, dim3(unsigned(count)), dim3(kWave), lds, stream, a, d_spins,
                       d_fields, d_energy);
    REWL_HIP(hipGetLastError());
    return WLX_OK;
}

int system_pilot_device(wlx_system *s, uint64_t seed, uint64_t attempts, double range[2])
{
    Args a{};
    a.J = s->d_J64;
    a.n = uint32_t(s->n);
    a.np = uint32_t(s->np);
    a.bounded = make_bounded(s->n);
    const KernelSet k = kernels_for(s->npl);
    const size_t lds = s->np * sizeof(double);
    if (int status = set_lds_limit(reinterpret_cast<const void *>(k.pilot), lds)) return status;
    double *d_range = nullptr;
    REWL_HIP(hipMalloc(&d_range, 2 * sizeof(double)));
    hipLaunchKernelGGL(k.pilot, dim3(1), dim3(kWave), lds, s->stream, a, seed, attempts, d_range);
    hipError_t launch = hipGetLastError();
    hipError_t copy = hipMemcpyAsync(range, d_range, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                     s->stream);
    hipError_t sync = hipStreamSynchronize(s->stream);
    (void)hipFree(d_range);
    REWL_HIP(launch);
    REWL_HIP(copy);
    REWL_HIP(sync);
    return WLX_OK;
}

template <class T> int device_alloc(T **p, size_t count)
{
    REWL_HIP(hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(1, count) * sizeof(T)));
    return WLX_OK;
}

} // namespace rewl

using namespace rewl;

struct wlx_engine {
    wlx_system *sys = nullptr;
    wlx_config cfg{};
    std::vector<uint64_t> window_be;
    Grid grid{};
    uint64_t bins = 0;
    uint32_t windows = 0, per_window = 0, walkers = 0, ws_stride = 8;
    KernelSet kernels{};
    WalkerState *d_state = nullptr;
    int8_t *d_spins = nullptr;
    double *d_fields = nullptr;
    double *d_energy = nullptr;
    double *d_logg = nullptr;
    uint64_t *d_hist = nullptr;
    uint8_t *d_active = nullptr;
    uint8_t *d_refine = nullptr;
    uint32_t *d_tsite = nullptr;
    uint32_t *d_tbin = nullptr;
    uint8_t *d_tflag = nullptr;
    uint64_t trace_cap = 0;
    Globals *d_glob = nullptr;
    Args *d_args = nullptr;       // the sampling kernels' parameters (a table of one)
    BatchSlot *d_slots = nullptr; // {0, w} for every walker w
    Args uploaded{};
    bool args_uploaded = false;
    InitReport *d_report = nullptr;
    uint64_t *d_op = nullptr;
    hipStream_t stream = nullptr;
    bool initialized = false;
    size_t lds_sample = 0, lds_init = 0, lds_walker = 0;
    size_t lds_staged = 0; // the staged lookahead walker's LDS (0: not available)
    uint32_t compute_units = 0; // of the engine's device
    // windows too wide for LDS: bins stay in HBM and only the LDS walker samples
    bool bins_hbm = false;
    uint64_t kernel_choice = WLX_KERNEL_AUTO;
    int width_index = kDefaultWidth; // lookahead width of the Philox common-configuration kernel
    bool width_forced = false;       // set by wlx_engine_set_lookahead_width
    bool timing = false;
    uint64_t timing_every = 1, timing_count = 0; // time launches timing_count % every == 0
    std::vector<hipEvent_t> events;
    double sample_ms = 0.0;
    uint64_t launches = 0;
    uint64_t epochs_run = 0;
    // HIP graph of graph_epochs epochs (sampling and exchange launches) that run replays;
    // recaptured when the launch parameters or the sampling kernel change
    uint64_t graph_epochs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    Args graph_args{};
    SampleFn graph_fn = nullptr;
    // FP32 fields (WLX_FIELDS_FP32_KAHAN)
    bool f32 = false;
    float2 *d_fields32 = nullptr;
    double *d_S = nullptr, *d_H = nullptr;
    Drift *d_drift = nullptr;
    uint32_t slots_padded = 0;
    uint64_t resync_every = 0, since_resync = 0, resyncs = 0;
    // pilot statistics (collect_window_statistics)
    bool collect = false;
    uint32_t collect_offset = 0;
    double *d_sq = nullptr;
    uint64_t *d_cnt = nullptr;
    SamplingState *d_sampling = nullptr;
    int8_t *d_rep = nullptr;

    Args args(uint32_t flags_trace = 0) const
    {
        Args a;
        std::memset(&a, 0, sizeof a); // padding too: upload_args compares bytes
        a.J = sys->d_J64;
        a.J32 = sys->d_J32;
        a.fields32 = d_fields32;
        a.n = uint32_t(sys->n);
        a.np = uint32_t(sys->np);
        a.bounded = make_bounded(sys->n);
        a.grid = grid;
        a.state = d_state;
        a.spins = d_spins;
        a.fields = d_fields;
        a.energy = d_energy;
        a.logg = d_logg;
        a.hist = d_hist;
        a.active = d_active;
        a.refine = d_refine;
        a.ws_stride = ws_stride;
        a.walkers = walkers;
        a.windows = windows;
        a.per_window = per_window;
        a.trace_site = d_tsite;
        a.trace_bin = d_tbin;
        a.trace_flag = d_tflag;
        a.trace_cap = trace_cap;
        a.glob = d_glob;
        a.seed = cfg.seed;
        a.exchange_interval = cfg.exchange_interval_attempts;
        a.max_attempts = cfg.max_attempts;
        a.check_interval = cfg.wl.check_interval_attempts;
        a.force_after = cfg.wl.force_accept_after_attempts;
        a.min_visits = cfg.wl.minimum_visits;
        a.final_factor = cfg.wl.final_factor;
        a.flatness = cfg.wl.flatness;
        a.inverse_time = cfg.wl.inverse_time_enabled ? 1 : 0;
        a.nalivaiko = cfg.wl.nalivaiko_mod ? 1 : 0;
        a.philox = cfg.rng == WLX_RNG_PHILOX ? 1 : 0;
        a.unlimited = cfg.max_attempts == 0 ? 1 : 0;
        a.csr_offsets = sys->csr ? sys->d_offsets : nullptr; // null: dense rows
        a.csr_columns = sys->d_columns;
        a.csr_values = sys->d_values;
        a.csr_values32 = sys->csr ? sys->d_values32 : nullptr;
        a.bins_hbm = bins_hbm ? 1 : 0;
        a.sq_disp = d_sq;
        a.disp_samples = d_cnt;
        a.sampling = d_sampling;
        a.rep_spins = d_rep;
        a.collect = collect ? 1 : 0;
        a.collect_offset = collect_offset;
        a.rt_margin = cfg.wl.round_trip_margin_fraction;
        (void)flags_trace;
        return a;
    }

    // the sampling kernels read Args from d_args, refreshed when the parameters change
    int upload_args()
    {
        const Args a = args();
        if (args_uploaded && std::memcmp(&a, &uploaded, sizeof a) == 0) return WLX_OK;
        REWL_HIP(hipStreamSynchronize(stream)); // the previous copy has left `uploaded`
        std::memcpy(&uploaded, &a, sizeof a);
        REWL_HIP(hipMemcpyAsync(d_args, &uploaded, sizeof uploaded, hipMemcpyHostToDevice, stream));
        args_uploaded = true;
        return WLX_OK;
    }

    // the staged lookahead walker serves this engine: Philox common configuration, 2 to 8
    // sites per lane, its row buffers and window bins within a CU's LDS
    bool staged_eligible() const
    {
        return kernels.staged && lds_staged && !f32 && !collect && !bins_hbm &&
               cfg.rng == WLX_RNG_PHILOX && trace_cap == 0 &&
               cfg.wl.force_accept_after_attempts == 0 && !cfg.wl.nalivaiko_mod;
    }

    // the sampling kernel of the next launch; `batched`: the kernel this engine contributes
    // to a wlx_batch, where the staged walker (most of a CU's LDS per walker) is not used
    SampleFn sample_fn(bool batched = false) const
    {
        if (f32) return nullptr; // lookahead_f32_kernel, launched by value (launch_sample)
        const bool philox = cfg.rng == WLX_RNG_PHILOX;
        // the statistics are recorded by the one-flip-at-a-time kernels
        if (collect) return philox ? kernels.sample_ph_seq : kernels.sample_xo;
        if (kernel_choice == WLX_KERNEL_LDS || bins_hbm)
            return philox ? kernels.lds_ph : kernels.lds_xo;
        if (kernel_choice == WLX_KERNEL_SEQUENTIAL)
            return philox ? kernels.sample_ph_seq : kernels.sample_xo;
        const bool common =
            trace_cap == 0 && cfg.wl.force_accept_after_attempts == 0 && !cfg.wl.nalivaiko_mod;
        if (cfg.rng != WLX_RNG_PHILOX) return common ? kernels.sample_xo_fast : kernels.sample_xo;
        if (trace_cap == 0 && cfg.wl.force_accept_after_attempts == 0 && !cfg.wl.nalivaiko_mod) {
            // AUTO: while every walker can have a CU's LDS to itself (the staged walker holds
            // one workgroup per CU; more walkers queue behind each other)
            const bool staged = kernel_choice == WLX_KERNEL_STAGED ||
                                (kernel_choice == WLX_KERNEL_AUTO && !width_forced &&
                                 walkers <= compute_units);
            if (staged && !batched && staged_eligible()) return kernels.staged;
            return kernels.ph_fast_width[width_index];
        }
        return kernels.sample_ph;
    }

    int launch_sample(uint64_t request, uint32_t flags)
    {
        if (trace_cap) flags |= kTrace;
        if (f32) {
            hipLaunchKernelGGL(kernels.sample_f32, dim3(walkers), dim3(kWave), lds_sample, stream,
                               args(), request, flags);
            REWL_HIP(hipGetLastError());
            return WLX_OK;
        }
        const SampleFn fn = sample_fn();
        const bool lds_fn = fn == kernels.lds_xo || fn == kernels.lds_ph;
        const size_t lds = lds_fn ? lds_walker : fn == kernels.staged ? lds_staged : lds_sample;
        if (int status = upload_args()) return status;
        const uint32_t threads = fn == kernels.staged ? kStagedThreads : kWave;
        hipLaunchKernelGGL(fn, dim3(walkers), dim3(threads), lds, stream, d_args, d_slots, request,
                           flags);
        REWL_HIP(hipGetLastError());
        return WLX_OK;
    }

    // kernel c: FP64 H = J.S over all configuration slots, drift, restart of the FP32 pairs
    int launch_resync()
    {
        const uint32_t np = uint32_t(sys->np);
        hipLaunchKernelGGL(resync_spins_kernel, dim3(256), dim3(256), 0, stream, d_spins, np,
                           walkers, slots_padded, d_S);
        REWL_HIP(hipGetLastError());
        hipLaunchKernelGGL(resync_mfma_kernel,
                           dim3(kResyncParts * (np / (16 * kResyncTileX)) *
                                (slots_padded / (16 * kResyncTileY))),
                           dim3(kWave),
                           0, stream, sys->d_J64, d_S, uint32_t(sys->n), np, slots_padded,
                           walkers, d_H);
        REWL_HIP(hipGetLastError());
        hipLaunchKernelGGL(resync_finish_kernel, dim3(walkers), dim3(kResyncFinishThreads), 0,
                           stream, args(), d_H, d_drift);
        REWL_HIP(hipGetLastError());
        ++resyncs;
        return WLX_OK;
    }

    int launch_exchange(uint32_t gated, uint64_t epoch)
    {
        hipLaunchKernelGGL(exchange_kernel, dim3(1), dim3(kExchangeBlock), 0, stream, args(),
                           gated, epoch);
        REWL_HIP(hipGetLastError());
        return WLX_OK;
    }

    void release_graph()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_fn = nullptr;
    }

    // the graph of graph_epochs epochs for the current parameters and kernel: kernel nodes
    // only (the parameter table is uploaded before the capture, outside the graph)
    int ensure_graph()
    {
        if (int status = upload_args()) return status;
        const Args a = args();
        const SampleFn fn = sample_fn();
        if (graph_exec && fn == graph_fn && std::memcmp(&a, &graph_args, sizeof a) == 0)
            return WLX_OK;
        release_graph();
        REWL_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int status = WLX_OK;
        for (uint64_t k = 0; k < graph_epochs && status == WLX_OK; ++k) {
            status = launch_sample(0, kGate | kCheck);
            if (status == WLX_OK) status = launch_exchange(1, 0);
        }
        hipGraph_t captured = nullptr;
        const hipError_t end = hipStreamEndCapture(stream, &captured);
        if (status) {
            if (captured) (void)hipGraphDestroy(captured);
            return status;
        }
        REWL_HIP(end);
        graph = captured;
        REWL_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
        std::memcpy(&graph_args, &a, sizeof a);
        graph_fn = fn;
        return WLX_OK;
    }

    int check_escape()
    {
        Globals g{};
        REWL_HIP(hipMemcpyAsync(&g, d_glob, sizeof g, hipMemcpyDeviceToHost, stream));
        REWL_HIP(hipStreamSynchronize(stream));
        if (g.escaped) return set_error(WLX_RUNTIME_ERROR, "Walker escaped its energy window");
        return WLX_OK;
    }
};

namespace {

int require_engine(const wlx_engine *e, bool initialized = true)
{
    if (!e) return set_error(WLX_INVALID_ARGUMENT, "Null engine");
    if (initialized && !e->initialized)
        return set_error(WLX_LOGIC_ERROR, "Engine walkers are not initialized");
    return WLX_OK;
}

int require_walker(const wlx_engine *e, uint64_t w)
{
    if (int status = require_engine(e)) return status;
    if (w >= e->walkers) return set_error(WLX_OUT_OF_RANGE, "Walker index out of range");
    return WLX_OK;
}

// WlParameters validation (wang_landau.cpp:74-90)
int validate_parameters(const wlx_params &p)
{
    if (!(p.flatness > 0.0 && p.flatness <= 1.0) || !(p.final_factor > 0.0) ||
        p.check_interval_attempts == 0 || p.initialization_max_attempts == 0 ||
        !(p.initialization_target_fraction > 0.0 && p.initialization_target_fraction <= 1.0) ||
        !std::isfinite(p.initialization_temperature_fraction) ||
        !(p.initialization_temperature_fraction > 0.0) ||
        p.initialization_stall_attempts_per_spin == 0 ||
        !std::isfinite(p.initialization_temperature_multiplier) ||
        !(p.initialization_temperature_multiplier > 1.0) ||
        !std::isfinite(p.initialization_max_temperature_fraction) ||
        p.initialization_max_temperature_fraction < p.initialization_temperature_fraction)
        return set_error(WLX_INVALID_ARGUMENT, "Invalid WL parameters");
    return WLX_OK;
}

// partition_windows (wang_landau.cpp:40-61)
int partition_windows(uint64_t bins, uint64_t count, double overlap, std::vector<uint64_t> &be)
{
    if (bins == 0 || count == 0 || count > bins || overlap < 0.0 || overlap >= 1.0)
        return set_error(WLX_INVALID_ARGUMENT, "Invalid energy-window partition");
    be.assign(2 * count, 0);
    if (count == 1) {
        be[1] = bins;
        return WLX_OK;
    }
    const double denominator = 1.0 + double(count - 1) * (1.0 - overlap);
    uint64_t size = uint64_t(std::ceil(double(bins) / denominator));
    size = std::max<uint64_t>(2, std::min(size, bins));
    const uint64_t span = bins - size;
    for (uint64_t w = 0; w < count; ++w) {
        const uint64_t begin = uint64_t(std::llround(double(w * span) / double(count - 1)));
        be[2 * w] = begin;
        be[2 * w + 1] = begin + size;
    }
    be[2 * count - 1] = bins;
    for (uint64_t w = 1; w < count; ++w)
        if (be[2 * w] >= be[2 * w - 1])
            return set_error(WLX_INVALID_ARGUMENT,
                             "Energy windows do not overlap; increase overlap or bins");
    return WLX_OK;
}

std::string format17(double x)
{
    char buffer[64];
    std::snprintf(buffer, sizeof buffer, "%.17g", x);
    return buffer;
}

} // namespace

extern "C" {

int wlx_system_pilot(const wlx_system *s, uint64_t seed, uint64_t attempts, double *minimum,
                     double *maximum)
{
    if (!s || !minimum || !maximum) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    double range[2];
    if (int status = system_pilot_device(const_cast<wlx_system *>(s), seed, attempts, range))
        return status;
    *minimum = range[0];
    *maximum = range[1];
    return WLX_OK;
}

int wlx_systems_pilot(wlx_system *const *systems, uint64_t count, uint64_t seed,
                      const uint64_t *attempts, double *minimum, double *maximum)
{
    if (!systems || !attempts || !minimum || !maximum || count == 0)
        return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (count > (uint64_t(1) << 20)) return set_error(WLX_INVALID_ARGUMENT, "Too many systems");
    std::vector<PilotJob> jobs(count);
    size_t lds = 0;
    for (uint64_t k = 0; k < count; ++k) {
        const wlx_system *s = systems[k];
        if (!s) return set_error(WLX_INVALID_ARGUMENT, "Null system");
        jobs[k] = PilotJob{s->d_J64, s->csr ? s->d_offsets : nullptr, s->d_columns, s->d_values,
                           uint32_t(s->n), uint32_t(s->np), make_bounded(s->n), seed,
                           attempts[k], nullptr};
        lds = std::max(lds, size_t(s->np) * 9);
    }
    hipStream_t stream = systems[0]->stream;
    PilotJob *d_jobs = nullptr;
    double *d_range = nullptr;
    REWL_HIP(hipMalloc(&d_jobs, count * sizeof(PilotJob)));
    if (hipMalloc(&d_range, 2 * count * sizeof(double)) != hipSuccess) {
        (void)hipFree(d_jobs);
        return set_error(WLX_DEVICE_ERROR, "HIP error: out of memory for pilot ranges");
    }
    for (uint64_t k = 0; k < count; ++k) jobs[k].range = d_range + 2 * k;
    std::vector<double> range(2 * count);
    hipError_t err = set_lds_limit(reinterpret_cast<const void *>(pilot_lds_kernel), lds) == WLX_OK
                         ? hipSuccess
                         : hipErrorInvalidValue;
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_jobs, jobs.data(), count * sizeof(PilotJob),
                             hipMemcpyHostToDevice, stream);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(pilot_lds_kernel, dim3(uint32_t(count)), dim3(kWave), lds, stream,
                           d_jobs);
        err = hipGetLastError();
    }
    if (err == hipSuccess)
        err = hipMemcpyAsync(range.data(), d_range, 2 * count * sizeof(double),
                             hipMemcpyDeviceToHost, stream);
    const hipError_t sync = hipStreamSynchronize(stream);
    (void)hipFree(d_jobs);
    (void)hipFree(d_range);
    REWL_HIP(err);
    REWL_HIP(sync);
    for (uint64_t k = 0; k < count; ++k) {
        minimum[k] = range[2 * k];
        maximum[k] = range[2 * k + 1];
    }
    return WLX_OK;
}

int wlx_device_info(char *name, size_t capacity, uint64_t *compute_units)
{
    int device = 0;
    REWL_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop{};
    REWL_HIP(hipGetDeviceProperties(&prop, device));
    if (name && capacity) {
        std::snprintf(name, capacity, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = uint64_t(prop.multiProcessorCount);
    return WLX_OK;
}

int wlx_engine_create(wlx_system *sys, const wlx_config *c, wlx_engine **out)
{
    if (!sys || !c || !out) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    *out = nullptr;
    // EnergyGrid::validate / bins (wang_landau.cpp:13-25)
    if (!std::isfinite(c->emin) || !std::isfinite(c->emax) || !std::isfinite(c->bin_width) ||
        !(c->emax > c->emin) || !(c->bin_width > 0.0))
        return set_error(WLX_INVALID_ARGUMENT, "Invalid energy grid");
    const double count = (c->emax - c->emin) / c->bin_width;
    if (std::abs(count - std::round(count)) > 1e-9 * std::max(1.0, count))
        return set_error(WLX_INVALID_ARGUMENT,
                         "Energy range must be an integer multiple of bin width");
    const uint64_t bins = uint64_t(std::llround(count));
    // RunConfig::validate subset (io.cpp:243-268)
    if (c->windows == 0 || c->walkers_per_window == 0 || c->overlap < 0.0 || c->overlap >= 1.0 ||
        c->exchange_interval_attempts == 0 || c->wl.check_interval_attempts == 0)
        return set_error(WLX_INVALID_ARGUMENT, "Invalid run configuration");
    if (c->windows > bins) return set_error(WLX_INVALID_ARGUMENT, "More windows than energy bins");
    if (c->rng != WLX_RNG_XOSHIRO && c->rng != WLX_RNG_PHILOX)
        return set_error(WLX_INVALID_ARGUMENT, "Unknown RNG mode");
    if (c->fields != WLX_FIELDS_FP64 && c->fields != WLX_FIELDS_FP32_KAHAN)
        return set_error(WLX_INVALID_ARGUMENT, "Unknown field precision mode");
    if (c->fields == WLX_FIELDS_FP32_KAHAN &&
        (c->rng != WLX_RNG_PHILOX || c->wl.force_accept_after_attempts != 0 || c->wl.nalivaiko_mod))
        return set_error(WLX_INVALID_ARGUMENT,
                         "FP32 fields run on the Philox stream without forced acceptance or the "
                         "Nalivaiko mask");
    if (c->fields == WLX_FIELDS_FP32_KAHAN && c->wl.collect_window_statistics)
        return set_error(WLX_INVALID_ARGUMENT,
                         "Window statistics are collected with FP64 fields only");
    if (int status = validate_parameters(c->wl)) return status;
    auto e = std::make_unique<wlx_engine>();
    e->sys = sys;
    e->cfg = *c;
    e->cfg.explicit_windows = nullptr;
    e->bins = bins;
    if (bins >= (uint64_t(1) << 31))
        return set_error(WLX_INVALID_ARGUMENT, "Energy grid too large for the device engine");
    e->grid = make_grid(c->emin, c->emax, c->bin_width, int64_t(bins));
    if (c->explicit_windows) {
        // per-walker window check of the WangLandauWalker constructor (wang_landau.cpp:72-73);
        // RunConfig's range/ordering rules (io.cpp:257-267) are the host shim's
        e->window_be.assign(c->explicit_windows, c->explicit_windows + 2 * c->windows);
        for (uint64_t w = 0; w < c->windows; ++w)
            if (e->window_be[2 * w] >= e->window_be[2 * w + 1] || e->window_be[2 * w + 1] > bins)
                return set_error(WLX_INVALID_ARGUMENT, "Invalid walker energy window");
    } else if (int status = partition_windows(bins, c->windows, c->overlap, e->window_be)) {
        return status;
    }
    if (c->windows * c->walkers_per_window > (uint64_t(1) << 24))
        return set_error(WLX_INVALID_ARGUMENT, "Too many walkers");
    e->windows = uint32_t(c->windows);
    e->per_window = uint32_t(c->walkers_per_window);
    e->walkers = e->windows * e->per_window;
    uint64_t ws_max = 0;
    for (uint32_t w = 0; w < e->windows; ++w)
        ws_max = std::max(ws_max, e->window_be[2 * w + 1] - e->window_be[2 * w]);
    e->ws_stride = uint32_t((ws_max + 7) / 8 * 8);
    e->kernels = kernels_for(sys->npl);
    e->lds_sample = bins_lds_bytes(e->ws_stride) + kFlipBoxBytes;
    e->lds_init = e->lds_sample + sys->np * sizeof(double);
    if (c->wl.collect_window_statistics) { // per-bin sums and the walker's state after init's scratch
        e->collect = true;
        e->collect_offset = uint32_t((e->lds_init + 15) / 16 * 16);
        const size_t end = e->collect_offset + 16 * size_t(e->ws_stride) + sizeof(SamplingState);
        e->lds_sample = end;
        e->lds_init = end;
    }
    if (e->lds_init > kLdsBytes) {
        // windows too wide for LDS: the bins stay in HBM, where the initialisation, the LDS
        // walker and the harness operations work on them in place
        if (e->collect || c->fields == WLX_FIELDS_FP32_KAHAN)
            return set_error(WLX_INVALID_ARGUMENT, "Energy windows too wide for LDS residency");
        e->bins_hbm = true;
        e->lds_sample = 0;
        e->lds_init = sys->np * sizeof(double); // total_energy's scratch
    }
    if (!e->bins_hbm) {
        for (SampleFn fn : {e->kernels.sample_xo, e->kernels.sample_xo_fast, e->kernels.sample_ph,
                            e->kernels.sample_ph_seq, e->kernels.ph_fast_width[0],
                            e->kernels.ph_fast_width[1], e->kernels.ph_fast_width[2],
                            e->kernels.ph_fast_width[3]})
            if (int status = set_lds_limit(reinterpret_cast<const void *>(fn), e->lds_sample))
                return status;
        if (int status = set_lds_limit(reinterpret_cast<const void *>(e->kernels.sample_f32),
                                       e->lds_sample))
            return status;
    }
    if (int status = set_lds_limit(reinterpret_cast<const void *>(e->kernels.init), e->lds_init))
        return status;
    {
        int device = 0;
        REWL_HIP(hipGetDevice(&device));
        int cus = 0;
        REWL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        e->compute_units = uint32_t(cus);
    }
    if (e->kernels.staged && !e->bins_hbm) { // candidate rows staged in LDS
        const size_t bytes = staged_layout(e->ws_stride, uint32_t(sys->np)).total;
        if (bytes <= kLdsBytes) {
            if (int status = set_lds_limit(reinterpret_cast<const void *>(e->kernels.staged), bytes))
                return status;
            e->lds_staged = bytes;
        }
    }
    { // the LDS walker stages the configuration in LDS beside the bins (or alone)
        e->lds_walker = lds_walker_bytes(e->ws_stride, uint32_t(sys->n), e->bins_hbm);
        if (e->lds_walker <= kLdsBytes)
            for (SampleFn fn : {e->kernels.lds_xo, e->kernels.lds_ph})
                if (int status = set_lds_limit(reinterpret_cast<const void *>(fn), e->lds_walker))
                    return status;
    }
    if (int status = set_lds_limit(reinterpret_cast<const void *>(walker_op_kernel),
                                   e->lds_sample))
        return status;

    const size_t W = e->walkers, np = sys->np, S = e->ws_stride;
    REWL_HIP(hipStreamCreate(&e->stream));
    if (int st = device_alloc(&e->d_state, W)) return st;
    if (int st = device_alloc(&e->d_spins, W * np)) return st;
    if (int st = device_alloc(&e->d_fields, W * np)) return st;
    if (int st = device_alloc(&e->d_energy, W)) return st;
    if (int st = device_alloc(&e->d_logg, W * S)) return st;
    if (int st = device_alloc(&e->d_hist, W * S)) return st;
    if (int st = device_alloc(&e->d_active, W * S)) return st;
    if (int st = device_alloc(&e->d_refine, W * S)) return st;
    if (int st = device_alloc(&e->d_glob, 1)) return st;
    if (int st = device_alloc(&e->d_args, 1)) return st;
    if (int st = device_alloc(&e->d_slots, e->walkers)) return st;
    {
        std::vector<BatchSlot> identity(e->walkers);
        for (uint32_t w = 0; w < e->walkers; ++w) identity[w] = BatchSlot{0, w};
        REWL_HIP(hipMemcpyAsync(e->d_slots, identity.data(), identity.size() * sizeof(BatchSlot),
                                hipMemcpyHostToDevice, e->stream));
        REWL_HIP(hipStreamSynchronize(e->stream));
    }
    if (int st = device_alloc(&e->d_report, W)) return st;
    if (int st = device_alloc(&e->d_op, 8)) return st;
    if (c->fields == WLX_FIELDS_FP32_KAHAN) {
        e->f32 = true;
        if (int st = system_j32(sys)) return st;
        // whole slot tiles of the resync
        e->slots_padded = (W + 16 * kResyncTileY - 1) / (16 * kResyncTileY) * (16 * kResyncTileY);
        if (int st = device_alloc(&e->d_fields32, W * np)) return st;
        if (int st = device_alloc(&e->d_S, size_t(np) * e->slots_padded)) return st;
        if (int st = device_alloc(&e->d_H, kResyncParts * W * np)) return st;
        if (int st = device_alloc(&e->d_drift, 1)) return st;
        REWL_HIP(hipMemsetAsync(e->d_fields32, 0, W * np * sizeof(float2), e->stream));
        REWL_HIP(hipMemsetAsync(e->d_drift, 0, sizeof(Drift), e->stream));
        const uint64_t x = c->exchange_interval_attempts;
        e->resync_every = c->resync_interval_attempts ? (c->resync_interval_attempts + x - 1) / x : 0;
    }
    if (e->collect) {
        if (int st = device_alloc(&e->d_sq, W * S)) return st;
        if (int st = device_alloc(&e->d_cnt, W * S)) return st;
        if (int st = device_alloc(&e->d_sampling, W)) return st;
        if (int st = device_alloc(&e->d_rep, W * kRepresentatives * np)) return st;
    }
    std::vector<WalkerState> states(W);
    for (uint32_t k = 0; k < W; ++k) {
        WalkerState &s = states[k];
        std::memset(&s, 0, sizeof s);
        const uint32_t window = k / e->per_window;
        s.begin = uint32_t(e->window_be[2 * window]);
        s.end = uint32_t(e->window_be[2 * window + 1]);
        s.slot = k;
        s.factor = 1.0;
    }
    REWL_HIP(hipMemcpyAsync(e->d_state, states.data(), W * sizeof(WalkerState),
                            hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemsetAsync(e->d_spins, 1, W * np, e->stream));
    REWL_HIP(hipMemsetAsync(e->d_fields, 0, W * np * sizeof(double), e->stream));
    REWL_HIP(hipMemsetAsync(e->d_glob, 0, sizeof(Globals), e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    *out = e.release();
    return WLX_OK;
}

int wlx_engine_destroy(wlx_engine *e)
{
    if (!e) return WLX_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    void *buffers[] = {e->d_state, e->d_spins,  e->d_fields, e->d_energy, e->d_logg,
                       e->d_hist,  e->d_active, e->d_refine, e->d_tsite,  e->d_tbin,
                       e->d_tflag, e->d_glob,   e->d_report, e->d_op,     e->d_fields32,
                       e->d_S,     e->d_H,      e->d_drift,  e->d_sq,     e->d_cnt,
                       e->d_sampling, e->d_rep, e->d_args, e->d_slots};
    for (void *p : buffers)
        if (p) (void)hipFree(p);
    e->release_graph();
    for (hipEvent_t ev : e->events) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return WLX_OK;
}

int wlx_engine_walkers(const wlx_engine *e, uint64_t *walkers, uint64_t *spins, uint64_t *bins)
{
    if (!e) return set_error(WLX_INVALID_ARGUMENT, "Null engine");
    if (walkers) *walkers = e->walkers;
    if (spins) *spins = e->sys->n;
    if (bins) *bins = e->bins;
    return WLX_OK;
}

int wlx_engine_layout(const wlx_engine *e, uint64_t *lds_bytes, uint64_t *bins_in_hbm)
{
    if (!e) return set_error(WLX_INVALID_ARGUMENT, "Null engine");
    if (lds_bytes) *lds_bytes = e->lds_walker;
    if (bins_in_hbm) *bins_in_hbm = e->bins_hbm ? 1 : 0;
    return WLX_OK;
}

int wlx_engine_initialize(wlx_engine *e, const int8_t *initial, const uint8_t *supplied)
{
    if (int status = require_engine(e, false)) return status;
    const wlx_system *s = e->sys;
    const size_t W = e->walkers;
    int8_t *d_initial = nullptr;
    uint8_t *d_supplied = nullptr;
    if (initial && supplied) {
        if (int st = device_alloc(&d_initial, W * s->n)) return st;
        if (int st = device_alloc(&d_supplied, W)) return st;
        REWL_HIP(hipMemcpyAsync(d_initial, initial, W * s->n, hipMemcpyHostToDevice, e->stream));
        REWL_HIP(hipMemcpyAsync(d_supplied, supplied, W, hipMemcpyHostToDevice, e->stream));
    }
    // every walker starts in its own slot
    std::vector<WalkerState> states(W);
    REWL_HIP(hipMemcpyAsync(states.data(), e->d_state, W * sizeof(WalkerState),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    for (uint32_t k = 0; k < W; ++k) states[k].slot = k;
    REWL_HIP(hipMemcpyAsync(e->d_state, states.data(), W * sizeof(WalkerState),
                            hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemsetAsync(e->d_glob, 0, sizeof(Globals), e->stream));
    InitParams ip{e->cfg.wl.initialization_max_attempts,
                  e->cfg.wl.initialization_target_fraction,
                  e->cfg.wl.initialization_temperature_fraction,
                  e->cfg.wl.initialization_stall_attempts_per_spin,
                  e->cfg.wl.initialization_temperature_multiplier,
                  e->cfg.wl.initialization_max_temperature_fraction};
    if (e->collect) { // statistics restart with the walker (wang_landau.cpp:104-118)
        REWL_HIP(hipMemsetAsync(e->d_sq, 0, W * e->ws_stride * sizeof(double), e->stream));
        REWL_HIP(hipMemsetAsync(e->d_cnt, 0, W * e->ws_stride * sizeof(uint64_t), e->stream));
        std::vector<SamplingState> sampling(W);
        for (uint32_t k = 0; k < W; ++k) {
            SamplingState &z = sampling[k];
            std::memset(&z, 0, sizeof z);
            const double lo = e->grid.minimum + double(states[k].begin) * e->grid.width;
            const double hi = e->grid.minimum + double(states[k].end) * e->grid.width;
            z.target[0] = lo;
            for (int r = 1; r < kRepresentatives; ++r)
                z.target[r] = lo + (hi - lo) * (double(r) - 0.5) / double(kRepresentatives - 1);
            for (int r = 0; r < kRepresentatives; ++r) z.dist[r] = HUGE_VAL;
        }
        REWL_HIP(hipMemcpyAsync(e->d_sampling, sampling.data(), W * sizeof(SamplingState),
                                hipMemcpyHostToDevice, e->stream));
        REWL_HIP(hipStreamSynchronize(e->stream));
    }
    hipLaunchKernelGGL(e->kernels.init, dim3(e->walkers), dim3(kWave), e->lds_init, e->stream,
                       e->args(), d_initial, d_supplied, ip, e->d_report);
    REWL_HIP(hipGetLastError());
    std::vector<InitReport> reports(W);
    REWL_HIP(hipMemcpyAsync(reports.data(), e->d_report, W * sizeof(InitReport),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    if (d_initial) (void)hipFree(d_initial);
    if (d_supplied) (void)hipFree(d_supplied);
    e->epochs_run = 0;
    if (e->f32) { // the FP32 pairs start from the FP64 local fields
        hipLaunchKernelGGL(narrow_fields_kernel, dim3(256), dim3(256), 0, e->stream, e->d_fields,
                           e->d_fields32, size_t(W) * s->np);
        REWL_HIP(hipGetLastError());
        REWL_HIP(hipStreamSynchronize(e->stream));
        e->since_resync = 0;
    }
    for (uint32_t k = 0; k < W; ++k) {
        const InitReport &r = reports[k];
        if (r.status == 0) continue;
        std::string what;
        if (r.status == 2) {
            what = "Spins must be +/-1";
        } else {
            const uint64_t begin = e->window_be[2 * (k / e->per_window)];
            const uint64_t end = e->window_be[2 * (k / e->per_window) + 1];
            const wlx_params &p = e->cfg.wl;
            const double lower = e->cfg.emin + double(begin) * e->cfg.bin_width;
            const double upper = e->cfg.emin + double(end) * e->cfg.bin_width;
            const double span = upper - lower, target = 0.5 * (lower + upper);
            const bool sup = initial && supplied && supplied[k];
            const double half = 0.5 * (sup ? 1.0 : p.initialization_target_fraction) * span;
            const uint64_t stall = p.initialization_stall_attempts_per_spin > ~uint64_t(0) / s->n
                                       ? ~uint64_t(0)
                                       : p.initialization_stall_attempts_per_spin * s->n;
            const int64_t cb = grid_index(e->grid, r.closest);
            what = "Target Metropolis initialization did not reach target band [" +
                   format17(target - half) + ", " + format17(target + half) +
                   "] for energy window bins [" + std::to_string(begin) + ", " +
                   std::to_string(end) + "), energy bounds [" + format17(lower) + ", " +
                   format17(upper) + (end == e->bins ? "]" : ")") + " after " +
                   std::to_string(p.initialization_max_attempts) +
                   " attempts; reached energy range [" + format17(r.min_reached) + ", " +
                   format17(r.max_reached) + "], closest energy " + format17(r.closest) +
                   (cb >= 0 ? " (bin " + std::to_string(cb) + ")"
                            : std::string(" (outside the energy grid)")) +
                   ", distance to target band " +
                   format17(std::max(0.0, r.closest_distance - half)) +
                   "; adaptive search used " + std::to_string(r.increases) +
                   " temperature increases up to " + format17(r.highest) + " and " +
                   std::to_string(r.restarts) + " random restarts (stall interval " +
                   std::to_string(stall) + " attempts)";
        }
        return set_error(WLX_RUNTIME_ERROR,
                         "Walker initialization failed: window " +
                             std::to_string(k / e->per_window) + ", local walker " +
                             std::to_string(k % e->per_window) + ": " + what);
    }
    e->initialized = true;
    return WLX_OK;
}

int wlx_engine_sample(wlx_engine *e, uint64_t attempts)
{
    if (int status = require_engine(e)) return status;
    if (int status = e->launch_sample(attempts, 0)) return status;
    return e->check_escape();
}

int wlx_engine_check(wlx_engine *e)
{
    if (int status = require_engine(e)) return status;
    if (int status = e->launch_sample(0, kCheck)) return status;
    return e->check_escape();
}

int wlx_engine_exchange(wlx_engine *e, uint64_t epoch, uint64_t *attempted, uint64_t *accepted)
{
    if (int status = require_engine(e)) return status;
    Globals before{}, after{};
    REWL_HIP(hipMemcpyAsync(&before, e->d_glob, sizeof before, hipMemcpyDeviceToHost, e->stream));
    if (int status = e->launch_exchange(0, epoch)) return status;
    REWL_HIP(hipMemcpyAsync(&after, e->d_glob, sizeof after, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    if (attempted) *attempted = after.exchange_attempted - before.exchange_attempted;
    if (accepted) *accepted = after.exchange_accepted - before.exchange_accepted;
    return WLX_OK;
}

int wlx_engine_resync(wlx_engine *e)
{
    if (int status = require_engine(e)) return status;
    if (!e->f32) return set_error(WLX_LOGIC_ERROR, "Resync applies to FP32 fields only");
    if (int status = e->launch_resync()) return status;
    REWL_HIP(hipStreamSynchronize(e->stream));
    return WLX_OK;
}

int wlx_engine_warm_start(wlx_engine *e, const double *log_g, double factor)
{
    if (int status = require_engine(e)) return status;
    if (!log_g) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (!(factor > 0.0) || !(factor <= 1.0))
        return set_error(WLX_INVALID_ARGUMENT, "The initial ln f must lie in (0, 1]");
    for (uint64_t b = 0; b < e->bins; ++b)
        if (!std::isfinite(log_g[b]))
            return set_error(WLX_INVALID_ARGUMENT, "The warm-start log g must be finite on every bin");
    const size_t W = e->walkers, S = e->ws_stride;
    std::vector<WalkerState> states(W);
    REWL_HIP(hipMemcpyAsync(states.data(), e->d_state, W * sizeof(WalkerState),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    std::vector<double> lg(W * S, 0.0);
    for (size_t w = 0; w < W; ++w) {
        WalkerState &st = states[w];
        const uint32_t ws = st.end - st.begin;
        double lowest = log_g[st.begin];
        for (uint32_t b = 1; b < ws; ++b) lowest = std::min(lowest, log_g[st.begin + b]);
        for (uint32_t b = 0; b < ws; ++b) lg[w * S + b] = log_g[st.begin + b] - lowest;
        st.factor = factor;
        st.stage = WLX_STAGE_WANG_LANDAU;
        st.active_count = st.refine_count = 0;
        st.last_check = st.attempted;
        // as if 1/t had run for window_bins / factor attempts: once the stage switches, the
        // factor continues from about `factor` instead of restarting at active / attempted
        st.clock_offset = uint64_t(std::ceil(double(ws) / factor));
    }
    REWL_HIP(hipMemcpyAsync(e->d_logg, lg.data(), W * S * sizeof(double), hipMemcpyHostToDevice,
                            e->stream));
    REWL_HIP(hipMemsetAsync(e->d_hist, 0, W * S * sizeof(uint64_t), e->stream));
    REWL_HIP(hipMemsetAsync(e->d_active, 0, W * S, e->stream));
    REWL_HIP(hipMemsetAsync(e->d_refine, 0, W * S, e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_state, states.data(), W * sizeof(WalkerState),
                            hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    return WLX_OK;
}

int wlx_engine_set_kernel(wlx_engine *e, uint64_t kernel)
{
    if (int status = require_engine(e, false)) return status;
    if (kernel > WLX_KERNEL_STAGED) return set_error(WLX_INVALID_ARGUMENT, "Unknown walker kernel");
    if (kernel == WLX_KERNEL_STAGED && !e->staged_eligible())
        return set_error(WLX_INVALID_ARGUMENT,
                         "The staged walker needs the Philox stream in the common configuration, "
                         "2 to 8 sites per lane and its rows and window bins within LDS");
    if (kernel == WLX_KERNEL_LDS && e->lds_walker > kLdsBytes)
        return set_error(WLX_INVALID_ARGUMENT, "Fields and window bins exceed the LDS of a CU");
    if (e->bins_hbm && kernel != WLX_KERNEL_LDS && kernel != WLX_KERNEL_AUTO)
        return set_error(WLX_INVALID_ARGUMENT,
                         "Windows too wide for LDS run on the LDS walker only");
    e->kernel_choice = kernel;
    return WLX_OK;
}

namespace {
// family and name of sampling kernel `fn` of engine e
void kernel_name(const wlx_engine *e, SampleFn fn, uint64_t *kernel, char *name, size_t capacity)
{
    const KernelSet &k = e->kernels;
    uint64_t family = WLX_KERNEL_LOOKAHEAD;
    const char *base = "lookahead_kernel";
    if (fn == k.sample_xo || fn == k.sample_ph_seq) {
        family = WLX_KERNEL_SEQUENTIAL;
        base = "sample_kernel";
    } else if (fn == k.lds_xo || fn == k.lds_ph) {
        family = WLX_KERNEL_LDS;
        base = "lds_sample_kernel";
    } else if (fn == k.sample_xo_fast) {
        base = "lookahead_xo_kernel";
    } else if (fn == k.ph_fast_width[0] || fn == k.ph_fast_width[1] || fn == k.ph_fast_width[2] ||
               fn == k.ph_fast_width[3]) {
        base = "lookahead_fast_kernel";
    } else if (k.staged && fn == k.staged) {
        family = WLX_KERNEL_STAGED;
        base = "lookahead_staged_kernel";
    } else if (fn == nullptr) { // FP32 fields
        base = "lookahead_f32_kernel";
    }
    if (kernel) *kernel = family;
    if (name && capacity) {
        if (family == WLX_KERNEL_LDS)
            std::snprintf(name, capacity, "%s<%s>", base,
                          e->cfg.rng == WLX_RNG_PHILOX ? "true" : "false");
        else if (family == WLX_KERNEL_SEQUENTIAL)
            std::snprintf(name, capacity, "%s<%u,%s>", base, unsigned(e->sys->npl),
                          e->cfg.rng == WLX_RNG_PHILOX ? "true" : "false");
        else if (fn == k.ph_fast_width[e->width_index] && e->width_index != kDefaultWidth)
            std::snprintf(name, capacity, "%s<%u,%u>", base, unsigned(e->sys->npl),
                          kLookaheadWidths[e->width_index]);
        else
            std::snprintf(name, capacity, "%s<%u>", base, unsigned(e->sys->npl));
    }
}
} // namespace

int wlx_engine_kernel(const wlx_engine *e, uint64_t *kernel, char *name, size_t capacity)
{
    if (!e) return set_error(WLX_INVALID_ARGUMENT, "Null engine");
    kernel_name(e, e->sample_fn(), kernel, name, capacity);
    return WLX_OK;
}

int wlx_engine_set_lookahead_width(wlx_engine *e, uint64_t width)
{
    if (int status = require_engine(e, false)) return status;
    // width 32 with more than 16 sites per lane gave wrong histogram counts on the GPU (at 64
    // sites per lane: 506 registers and 304 bytes of scratch per lane), cause not found: it is
    // only offered where it was shown exact
    if (width == 32 && e->sys->npl > 16)
        return set_error(WLX_INVALID_ARGUMENT,
                         "Lookahead width 32 is limited to systems of at most 1024 sites");
    for (int k = 0; k < 4; ++k)
        if (kLookaheadWidths[k] == width) {
            e->width_index = k;
            e->width_forced = true;
            return WLX_OK;
        }
    return set_error(WLX_INVALID_ARGUMENT, "Lookahead width must be 4, 8, 16 or 32");
}

int wlx_engine_autotune(wlx_engine *e, uint64_t epochs, uint64_t *chosen, double *seconds)
{
    if (int status = require_engine(e)) return status;
    if (epochs == 0) return set_error(WLX_INVALID_ARGUMENT, "Autotuning needs at least one epoch");
    if (seconds)
        for (int k = 0; k <= int(WLX_KERNEL_STAGED); ++k) seconds[k] = -1.0;
    // (family, lookahead width index); the Philox common-configuration lookahead kernel is
    // tried at each width shown exact for this size
    std::vector<std::pair<uint64_t, int>> candidates;
    if (!e->bins_hbm) {
        const bool widths = e->cfg.rng == WLX_RNG_PHILOX && !e->f32 && !e->collect &&
                            e->trace_cap == 0 && e->cfg.wl.force_accept_after_attempts == 0 &&
                            !e->cfg.wl.nalivaiko_mod;
        if (widths) {
            if (e->staged_eligible()) candidates.push_back({WLX_KERNEL_STAGED, kDefaultWidth});
            for (int k : {kDefaultWidth, 1, 3})
                if (kLookaheadWidths[k] != 32 || e->sys->npl <= 16)
                    candidates.push_back({WLX_KERNEL_LOOKAHEAD, k});
        } else {
            candidates.push_back({WLX_KERNEL_LOOKAHEAD, e->width_index});
        }
        candidates.push_back({WLX_KERNEL_SEQUENTIAL, e->width_index});
    }
    if (e->lds_walker <= kLdsBytes) candidates.push_back({WLX_KERNEL_LDS, e->width_index});
    uint64_t best = e->kernel_choice == WLX_KERNEL_AUTO ? WLX_KERNEL_LOOKAHEAD : e->kernel_choice;
    int best_width = e->width_index;
    double best_s = -1.0;
    // every variant yields the same trajectory, so the trials are part of the run itself
    for (const auto &[variant, width] : candidates) {
        e->kernel_choice = variant;
        e->width_index = width;
        wlx_run_status st{};
        if (int status = wlx_engine_run(e, 1, &st)) return status; // first launch: code load
        if (st.done) break;
        const auto t0 = std::chrono::steady_clock::now();
        if (int status = wlx_engine_run(e, epochs, &st)) return status;
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (seconds && (seconds[variant] < 0.0 || s < seconds[variant])) seconds[variant] = s;
        if (best_s < 0.0 || s < best_s) {
            best_s = s;
            best = variant;
            best_width = width;
        }
        if (st.done) break;
    }
    e->kernel_choice = best;
    e->width_index = best_width;
    if (chosen) *chosen = best;
    return WLX_OK;
}

int wlx_engine_drift(wlx_engine *e, double *max_field_error, double *max_energy_bins,
                     uint64_t *resyncs)
{
    if (int status = require_engine(e, false)) return status;
    Drift d{};
    if (e->d_drift) {
        REWL_HIP(hipMemcpyAsync(&d, e->d_drift, sizeof d, hipMemcpyDeviceToHost, e->stream));
        REWL_HIP(hipStreamSynchronize(e->stream));
    }
    double f, en;
    std::memcpy(&f, &d.max_field, sizeof f);
    std::memcpy(&en, &d.max_energy, sizeof en);
    if (max_field_error) *max_field_error = f;
    if (max_energy_bins) *max_energy_bins = en;
    if (resyncs) *resyncs = e->resyncs;
    return WLX_OK;
}

int wlx_engine_read_sampling(wlx_engine *e, uint64_t w, double *squared_displacement,
                             uint64_t *displacement_samples, uint64_t *round_trips,
                             double *representative_energies, uint8_t *representative_present,
                             int8_t *representative_spins)
{
    if (int status = require_walker(e, w)) return status;
    if (!e->collect) return set_error(WLX_LOGIC_ERROR, "Window statistics are not collected");
    WalkerState st{};
    SamplingState z{};
    REWL_HIP(hipMemcpyAsync(&st, e->d_state + w, sizeof st, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipMemcpyAsync(&z, e->d_sampling + w, sizeof z, hipMemcpyDeviceToHost, e->stream));
    const size_t S = e->ws_stride, np = e->sys->np, n = e->sys->n;
    std::vector<double> sq(S);
    std::vector<uint64_t> cnt(S);
    std::vector<int8_t> spins(kRepresentatives * np);
    REWL_HIP(hipMemcpyAsync(sq.data(), e->d_sq + w * S, S * sizeof(double), hipMemcpyDeviceToHost,
                            e->stream));
    REWL_HIP(hipMemcpyAsync(cnt.data(), e->d_cnt + w * S, S * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, e->stream));
    if (representative_spins)
        REWL_HIP(hipMemcpyAsync(spins.data(), e->d_rep + w * kRepresentatives * np, spins.size(),
                                hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    const uint32_t ws = st.end - st.begin;
    if (squared_displacement) {
        std::fill(squared_displacement, squared_displacement + e->bins, 0.0);
        std::copy(sq.begin(), sq.begin() + ws, squared_displacement + st.begin);
    }
    if (displacement_samples) {
        std::fill(displacement_samples, displacement_samples + e->bins, uint64_t(0));
        std::copy(cnt.begin(), cnt.begin() + ws, displacement_samples + st.begin);
    }
    if (round_trips) *round_trips = z.round_trips;
    for (int r = 0; r < kRepresentatives; ++r) {
        const bool present = (z.present >> r) & 1;
        if (representative_energies) representative_energies[r] = present ? z.energy[r] : 0.0;
        if (representative_present) representative_present[r] = present ? 1 : 0;
        if (representative_spins && present)
            std::memcpy(representative_spins + r * n, spins.data() + r * np, n);
    }
    return WLX_OK;
}

int wlx_engine_set_walker_cadence(wlx_engine *e, uint64_t w, uint64_t last_flatness_check,
                                  double last_flatness)
{
    if (int status = require_walker(e, w)) return status;
    WalkerState st{};
    REWL_HIP(hipMemcpyAsync(&st, e->d_state + w, sizeof st, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    st.last_check = last_flatness_check;
    st.last_flatness = last_flatness;
    REWL_HIP(hipMemcpyAsync(e->d_state + w, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    return WLX_OK;
}

int wlx_engine_set_run_state(wlx_engine *e, uint64_t epoch, uint64_t exchange_attempted,
                             uint64_t exchange_accepted)
{
    if (int status = require_engine(e, false)) return status;
    Globals g{};
    REWL_HIP(hipMemcpyAsync(&g, e->d_glob, sizeof g, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    g.epoch = epoch;
    g.done = 0;
    g.exchange_attempted = exchange_attempted;
    g.exchange_accepted = exchange_accepted;
    REWL_HIP(hipMemcpyAsync(e->d_glob, &g, sizeof g, hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    return WLX_OK;
}

int wlx_engine_status(wlx_engine *e, wlx_run_status *st)
{
    if (int status = require_engine(e, false)) return status;
    if (!st) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    Globals g{};
    std::vector<WalkerState> states(e->walkers);
    REWL_HIP(hipMemcpyAsync(&g, e->d_glob, sizeof g, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipMemcpyAsync(states.data(), e->d_state, states.size() * sizeof(WalkerState),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    std::memset(st, 0, sizeof *st);
    st->epoch = g.epoch;
    st->done = g.done;
    st->exchange_attempted = g.exchange_attempted;
    st->exchange_accepted = g.exchange_accepted;
    st->ambiguous_decisions = g.exchange_ambiguous;
    bool converged = true;
    for (const auto &s : states) {
        st->attempted += s.attempted;
        st->accepted += s.accepted;
        st->forced_accepted += s.forced;
        st->ambiguous_decisions += s.ambiguous;
        converged &= s.stage == WLX_STAGE_FROZEN;
    }
    st->converged = converged ? 1 : 0;
    return WLX_OK;
}

int wlx_engine_run(wlx_engine *e, uint64_t max_epochs, wlx_run_status *status)
{
    if (int st = require_engine(e)) return st;
    // the device flags (done, escaped) are polled after every chunk of epochs; a run that
    // finishes inside a chunk leaves the rest of its launches gated (they return at once)
    constexpr uint64_t kChunk = 64;
    // timed launches and the host-side resync cadence run as plain launches
    const bool graphs = e->graph_epochs > 0 && !e->timing && e->resync_every == 0;
    const uint64_t chunk = std::max(kChunk, graphs ? e->graph_epochs : 0);
    if (e->timing && e->events.size() < 2 * kChunk)
        return set_error(WLX_LOGIC_ERROR, "Timing events were not created");
    uint64_t remaining = max_epochs;
    while (remaining > 0) {
        uint64_t n = std::min(remaining, chunk);
        remaining -= n;
        if (graphs && n >= e->graph_epochs) {
            if (int st = e->ensure_graph()) return st;
            for (; n >= e->graph_epochs; n -= e->graph_epochs)
                REWL_HIP(hipGraphLaunch(e->graph_exec, e->stream));
        }
        uint64_t timed = 0; // event pairs recorded in this chunk
        for (uint64_t k = 0; k < n; ++k) {
            const bool time_it = e->timing && e->timing_count++ % e->timing_every == 0;
            if (time_it) REWL_HIP(hipEventRecord(e->events[2 * timed], e->stream));
            if (int st = e->launch_sample(0, kGate | kCheck)) return st;
            if (time_it) REWL_HIP(hipEventRecord(e->events[2 * timed++ + 1], e->stream));
            if (e->resync_every && ++e->since_resync >= e->resync_every) {
                if (int st = e->launch_resync()) return st;
                e->since_resync = 0;
            }
            if (int st = e->launch_exchange(1, 0)) return st;
        }
        Globals g{};
        REWL_HIP(hipMemcpyAsync(&g, e->d_glob, sizeof g, hipMemcpyDeviceToHost, e->stream));
        REWL_HIP(hipStreamSynchronize(e->stream));
        for (uint64_t k = 0; k < timed; ++k) {
            float ms = 0.0f;
            REWL_HIP(hipEventElapsedTime(&ms, e->events[2 * k], e->events[2 * k + 1]));
            e->sample_ms += ms;
            ++e->launches;
        }
        if (g.escaped) return set_error(WLX_RUNTIME_ERROR, "Walker escaped its energy window");
        if (g.done) break;
    }
    if (status) return wlx_engine_status(e, status);
    return WLX_OK;
}

int wlx_engine_set_graph(wlx_engine *e, uint64_t epochs_per_graph)
{
    if (!e) return set_error(WLX_INVALID_ARGUMENT, "Null engine");
    if (epochs_per_graph > 4096)
        return set_error(WLX_INVALID_ARGUMENT, "At most 4096 epochs per graph");
    if (e->stream) REWL_HIP(hipStreamSynchronize(e->stream));
    e->release_graph();
    e->graph_epochs = epochs_per_graph;
    return WLX_OK;
}

int wlx_engine_read_walker(wlx_engine *e, uint64_t w, wlx_walker_info *info, int8_t *spins,
                           double *fields, double *log_g, uint64_t *histogram, uint8_t *active,
                           uint8_t *refinement_active)
{
    if (int status = require_walker(e, w)) return status;
    const wlx_system *s = e->sys;
    WalkerState st{};
    REWL_HIP(hipMemcpyAsync(&st, e->d_state + w, sizeof st, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    const uint32_t ws = st.end - st.begin;
    const size_t wb = size_t(w) * e->ws_stride;
    double energy = 0.0;
    std::vector<int8_t> sp(s->np);
    std::vector<double> fl(s->np), lg(ws);
    std::vector<uint64_t> hs(ws);
    std::vector<uint8_t> ac(ws), rf(ws);
    REWL_HIP(hipMemcpyAsync(&energy, e->d_energy + st.slot, sizeof energy, hipMemcpyDeviceToHost,
                            e->stream));
    REWL_HIP(hipMemcpyAsync(sp.data(), e->d_spins + size_t(st.slot) * s->np, s->np,
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipMemcpyAsync(fl.data(), e->d_fields + size_t(st.slot) * s->np,
                            s->np * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipMemcpyAsync(lg.data(), e->d_logg + wb, ws * sizeof(double), hipMemcpyDeviceToHost,
                            e->stream));
    REWL_HIP(hipMemcpyAsync(hs.data(), e->d_hist + wb, ws * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipMemcpyAsync(ac.data(), e->d_active + wb, ws, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipMemcpyAsync(rf.data(), e->d_refine + wb, ws, hipMemcpyDeviceToHost, e->stream));
    if (e->f32) { // FP32 fields: the value h - c of each pair
        std::vector<float2> pairs(s->np);
        REWL_HIP(hipMemcpyAsync(pairs.data(), e->d_fields32 + size_t(st.slot) * s->np,
                                s->np * sizeof(float2), hipMemcpyDeviceToHost, e->stream));
        REWL_HIP(hipStreamSynchronize(e->stream));
        for (uint64_t i = 0; i < s->np; ++i) fl[i] = double(pairs[i].x) - double(pairs[i].y);
    }
    REWL_HIP(hipStreamSynchronize(e->stream));
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->walker_id = w;
        info->window = w / e->per_window;
        info->window_begin = st.begin;
        info->window_end = st.end;
        info->energy = energy;
        info->factor = st.factor;
        info->attempted = st.attempted;
        info->accepted = st.accepted;
        info->forced_accepted = st.forced;
        info->last_accepted_attempt = st.last_accepted;
        info->stage = st.stage;
        info->active_bins = st.active_count;
        info->refinement_active_bins = st.refine_count;
        for (int k = 0; k < 4; ++k) info->rng_state[k] = st.rng[k];
        info->last_flatness_check = st.last_check;
        info->last_flatness = st.last_flatness;
        info->clock_offset = st.clock_offset;
    }
    if (spins) std::memcpy(spins, sp.data(), s->n);
    if (fields) std::memcpy(fields, fl.data(), s->n * sizeof(double));
    if (log_g) {
        std::fill(log_g, log_g + e->bins, 0.0);
        std::copy(lg.begin(), lg.end(), log_g + st.begin);
    }
    if (histogram) {
        std::fill(histogram, histogram + e->bins, uint64_t(0));
        std::copy(hs.begin(), hs.end(), histogram + st.begin);
    }
    if (active) {
        std::fill(active, active + e->bins, uint8_t(0));
        std::copy(ac.begin(), ac.end(), active + st.begin);
    }
    if (refinement_active) {
        std::fill(refinement_active, refinement_active + e->bins, uint8_t(0));
        std::copy(rf.begin(), rf.end(), refinement_active + st.begin);
    }
    return WLX_OK;
}

int wlx_engine_read_infos(wlx_engine *e, wlx_walker_info *infos)
{
    if (int status = require_engine(e)) return status;
    if (!infos) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    std::vector<WalkerState> states(e->walkers);
    std::vector<double> energies(e->walkers);
    REWL_HIP(hipMemcpyAsync(states.data(), e->d_state, states.size() * sizeof(WalkerState),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipMemcpyAsync(energies.data(), e->d_energy, energies.size() * sizeof(double),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    for (uint32_t w = 0; w < e->walkers; ++w) {
        const WalkerState &st = states[w];
        wlx_walker_info &info = infos[w];
        std::memset(&info, 0, sizeof info);
        info.walker_id = w;
        info.window = w / e->per_window;
        info.window_begin = st.begin;
        info.window_end = st.end;
        info.energy = energies[st.slot];
        info.factor = st.factor;
        info.attempted = st.attempted;
        info.accepted = st.accepted;
        info.forced_accepted = st.forced;
        info.last_accepted_attempt = st.last_accepted;
        info.stage = st.stage;
        info.active_bins = st.active_count;
        info.refinement_active_bins = st.refine_count;
        for (int k = 0; k < 4; ++k) info.rng_state[k] = st.rng[k];
        info.last_flatness_check = st.last_check;
        info.last_flatness = st.last_flatness;
        info.clock_offset = st.clock_offset;
    }
    return WLX_OK;
}

int wlx_engine_write_walker(wlx_engine *e, uint64_t w, const wlx_walker_info *info,
                            const int8_t *spins, const double *fields, const double *log_g,
                            const uint64_t *histogram, const uint8_t *active)
{
    if (int status = require_walker(e, w)) return status;
    const wlx_system *s = e->sys;
    // restore validation (wang_landau.cpp:421-430)
    if (!info || info->walker_id != w || !spins || !fields || !log_g || !histogram || !active)
        return set_error(WLX_INVALID_ARGUMENT, "Checkpoint is incompatible with walker");
    if (info->last_accepted_attempt > info->attempted)
        return set_error(WLX_INVALID_ARGUMENT, "Checkpoint has an invalid last accepted attempt");
    if (info->forced_accepted > info->accepted)
        return set_error(WLX_INVALID_ARGUMENT,
                         "Checkpoint has an invalid forced acceptance count");
    if (!e->cfg.wl.inverse_time_enabled && info->stage == WLX_STAGE_INVERSE_TIME)
        return set_error(WLX_INVALID_ARGUMENT, "Checkpoint uses disabled inverse-time refinement");
    for (uint64_t i = 0; i < s->n; ++i)
        if (spins[i] != 1 && spins[i] != -1)
            return set_error(WLX_INVALID_ARGUMENT, "Spins must be +/-1");
    WalkerState st{};
    REWL_HIP(hipMemcpyAsync(&st, e->d_state + w, sizeof st, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    const uint32_t ws = st.end - st.begin;
    std::vector<uint8_t> ac(ws), rf(ws, 0);
    uint32_t active_count = 0, refine_count = 0;
    for (uint32_t b = 0; b < ws; ++b) {
        ac[b] = active[st.begin + b] != 0 ? 1 : 0;
        active_count += ac[b];
        if (e->cfg.wl.nalivaiko_mod) {
            rf[b] = histogram[st.begin + b] != 0 ? 1 : 0;
            refine_count += rf[b];
        }
    }
    st.factor = info->factor;
    st.attempted = info->attempted;
    st.accepted = info->accepted;
    st.forced = info->forced_accepted;
    st.last_accepted = info->last_accepted_attempt;
    st.stage = uint32_t(info->stage);
    st.active_count = active_count;
    st.refine_count = refine_count;
    for (int k = 0; k < 4; ++k) st.rng[k] = info->rng_state[k];
    st.escaped = 0;
    st.clock_offset = info->clock_offset;
    std::vector<int8_t> sp(s->np, 1);
    std::vector<double> fl(s->np, 0.0);
    std::memcpy(sp.data(), spins, s->n);
    std::memcpy(fl.data(), fields, s->n * sizeof(double));
    const size_t wb = size_t(w) * e->ws_stride;
    const double energy = info->energy;
    REWL_HIP(hipMemcpyAsync(e->d_state + w, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_energy + st.slot, &energy, sizeof energy, hipMemcpyHostToDevice,
                            e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_spins + size_t(st.slot) * s->np, sp.data(), s->np,
                            hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_fields + size_t(st.slot) * s->np, fl.data(),
                            s->np * sizeof(double), hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_logg + wb, log_g + st.begin, ws * sizeof(double),
                            hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_hist + wb, histogram + st.begin, ws * sizeof(uint64_t),
                            hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_active + wb, ac.data(), ws, hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipMemcpyAsync(e->d_refine + wb, rf.data(), ws, hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    // re-verify energy and fields (wang_landau.cpp:459-465)
    std::vector<double> exact(s->n);
    double exact_energy = 0.0;
    if (int status = wlx_system_fields(s, spins, 1, exact.data(), &exact_energy)) return status;
    if (e->f32) { // FP32 fields restart from the FP64 recompute of the restored spins
        std::vector<float2> pairs(s->np, make_float2(0.0f, 0.0f));
        for (uint64_t i = 0; i < s->n; ++i) {
            const float hf = float(exact[i]);
            pairs[i] = make_float2(hf, float(double(hf) - exact[i]));
        }
        REWL_HIP(hipMemcpyAsync(e->d_fields32 + size_t(st.slot) * s->np, pairs.data(),
                                s->np * sizeof(float2), hipMemcpyHostToDevice, e->stream));
        REWL_HIP(hipMemcpyAsync(e->d_energy + st.slot, &exact_energy, sizeof exact_energy,
                                hipMemcpyHostToDevice, e->stream));
        REWL_HIP(hipStreamSynchronize(e->stream));
        return WLX_OK;
    }
    if (std::abs(exact_energy - energy) > 1e-9 * std::max(1.0, std::abs(exact_energy)))
        return set_error(WLX_RUNTIME_ERROR, "Checkpoint energy does not match spin configuration");
    for (uint64_t i = 0; i < s->n; ++i)
        if (std::abs(exact[i] - fields[i]) > 1e-9 * std::max(1.0, std::abs(exact[i])))
            return set_error(WLX_RUNTIME_ERROR,
                             "Checkpoint local fields do not match spin configuration");
    return WLX_OK;
}

static int walker_op(wlx_engine *e, uint64_t w, uint32_t op, uint64_t out[8])
{
    if (int status = require_walker(e, w)) return status;
    hipLaunchKernelGGL(walker_op_kernel, dim3(1), dim3(kWave), e->lds_sample, e->stream, e->args(),
                       uint32_t(w), op, e->d_op);
    REWL_HIP(hipGetLastError());
    REWL_HIP(hipMemcpyAsync(out, e->d_op, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    return WLX_OK;
}

int wlx_engine_histogram_statistics(wlx_engine *e, uint64_t w, wlx_histogram_stats *out)
{
    if (!out) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    uint64_t r[8];
    if (int status = walker_op(e, w, 0, r)) return status;
    out->active_bins = r[0];
    out->covered_bins = r[1];
    out->minimum = r[2];
    std::memcpy(&out->mean, &r[3], 8);
    std::memcpy(&out->min_over_mean, &r[4], 8);
    std::memcpy(&out->coverage, &r[5], 8);
    return WLX_OK;
}

int wlx_engine_ready(wlx_engine *e, uint64_t w, uint64_t *ready)
{
    if (!ready) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    uint64_t r[8];
    if (int status = walker_op(e, w, 1, r)) return status;
    *ready = r[6];
    return WLX_OK;
}

int wlx_engine_next_iteration(wlx_engine *e, uint64_t w)
{
    uint64_t r[8];
    if (int status = walker_op(e, w, 2, r)) return status;
    if (!r[6]) return set_error(WLX_LOGIC_ERROR, "WL iteration is not complete");
    return WLX_OK;
}

int wlx_engine_trace(wlx_engine *e, uint64_t capacity)
{
    if (int status = require_engine(e, false)) return status;
    if (e->f32 && capacity)
        return set_error(WLX_INVALID_ARGUMENT, "Attempt traces are not recorded with FP32 fields");
    (void)hipFree(e->d_tsite);
    (void)hipFree(e->d_tbin);
    (void)hipFree(e->d_tflag);
    e->d_tsite = nullptr;
    e->d_tbin = nullptr;
    e->d_tflag = nullptr;
    e->trace_cap = capacity;
    if (capacity) {
        const size_t total = size_t(e->walkers) * capacity;
        if (int st = device_alloc(&e->d_tsite, total)) return st;
        if (int st = device_alloc(&e->d_tbin, total)) return st;
        if (int st = device_alloc(&e->d_tflag, total)) return st;
    }
    // reset per-walker trace counters
    std::vector<WalkerState> states(e->walkers);
    REWL_HIP(hipMemcpyAsync(states.data(), e->d_state, states.size() * sizeof(WalkerState),
                            hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    for (auto &s : states) s.trace_count = 0;
    REWL_HIP(hipMemcpyAsync(e->d_state, states.data(), states.size() * sizeof(WalkerState),
                            hipMemcpyHostToDevice, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    return WLX_OK;
}

int wlx_engine_read_trace(wlx_engine *e, uint64_t w, uint64_t *count, uint32_t *sites,
                          uint32_t *bins, uint8_t *flags)
{
    if (int status = require_walker(e, w)) return status;
    if (!count) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    WalkerState st{};
    REWL_HIP(hipMemcpyAsync(&st, e->d_state + w, sizeof st, hipMemcpyDeviceToHost, e->stream));
    REWL_HIP(hipStreamSynchronize(e->stream));
    *count = st.trace_count;
    const size_t base = size_t(w) * e->trace_cap;
    if (st.trace_count) {
        if (sites)
            REWL_HIP(hipMemcpyAsync(sites, e->d_tsite + base, st.trace_count * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, e->stream));
        if (bins)
            REWL_HIP(hipMemcpyAsync(bins, e->d_tbin + base, st.trace_count * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, e->stream));
        if (flags)
            REWL_HIP(hipMemcpyAsync(flags, e->d_tflag + base, st.trace_count,
                                    hipMemcpyDeviceToHost, e->stream));
        REWL_HIP(hipStreamSynchronize(e->stream));
    }
    return WLX_OK;
}

int wlx_engine_set_timing(wlx_engine *e, int enable)
{
    if (!e) return set_error(WLX_INVALID_ARGUMENT, "Null engine");
    if (enable < 0) return set_error(WLX_INVALID_ARGUMENT, "Timing period must be >= 0");
    // one event pair per launch of a polling chunk (64 epochs), created here so that no
    // timed region pays for their creation
    while (enable > 0 && e->events.size() < 2 * 64) {
        hipEvent_t ev;
        REWL_HIP(hipEventCreate(&ev));
        e->events.push_back(ev);
    }
    e->timing = enable != 0;
    e->timing_every = enable > 0 ? uint64_t(enable) : 1;
    e->timing_count = 0;
    e->sample_ms = 0.0;
    e->launches = 0;
    return WLX_OK;
}

int wlx_engine_kernel_time(wlx_engine *e, double *sample_ms, uint64_t *launches)
{
    if (!e) return set_error(WLX_INVALID_ARGUMENT, "Null engine");
    if (sample_ms) *sample_ms = e->launches ? e->sample_ms / double(e->launches) : 0.0;
    if (launches) *launches = e->launches;
    return WLX_OK;
}

} // extern "C"

// ---- batches --------------------------------------------------------------------------------
// A batch runs the epochs of several engines together. Replicas of one register layout share
// that register kernel in one launch. Any other mix runs the LDS walker, whose engines are
// split into launch groups by the LDS one walker needs (classes of 16, 32, 64, 96 and 160 KiB):
// a launch's dynamic LDS is that of its largest walker, so a group of small systems packs many
// walkers per CU instead of inheriting the largest system's allocation. Each group has its own
// stream, parameter table, slot list (heaviest systems first) and HIP graph; an engine belongs
// to one group, so the groups' epochs run concurrently and independently. Engines whose
// window bins would need more LDS than the batch's budget keep their bins in HBM.
struct BatchGroup {
    std::vector<uint32_t> members; // indices into wlx_batch::engines
    Args *d_table = nullptr;
    BatchSlot *d_slots = nullptr;
    uint32_t slots = 0;
    size_t lds = 0;
    SampleFn sample = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    int enqueue_epochs(uint64_t count)
    {
        for (uint64_t k = 0; k < count; ++k) {
            hipLaunchKernelGGL(sample, dim3(slots), dim3(kWave), lds, stream, d_table, d_slots,
                               uint64_t(0), uint32_t(kGate | kCheck));
            REWL_HIP(hipGetLastError());
            hipLaunchKernelGGL(batch_exchange_kernel, dim3(uint32_t(members.size())),
                               dim3(kExchangeBlock), 0, stream, d_table);
            REWL_HIP(hipGetLastError());
        }
        return WLX_OK;
    }

    void release_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        release_graph();
        if (d_table) (void)hipFree(d_table);
        if (d_slots) (void)hipFree(d_slots);
        if (stream) (void)hipStreamDestroy(stream);
        d_table = nullptr;
        d_slots = nullptr;
        stream = nullptr;
    }
};

struct wlx_batch {
    std::vector<wlx_engine *> engines;
    std::vector<BatchGroup> groups;
    std::vector<uint32_t> group_of; // engine -> group
    bool philox = false;
    SampleFn common = nullptr; // the register kernel every engine runs alone, if one
    size_t common_lds = 0;
    size_t lds_budget = kLdsBytes;
    uint64_t graph_epochs = 0;
    std::string kernel_name;
    std::vector<Globals> globs;

    ~wlx_batch()
    {
        for (BatchGroup &g : groups) g.release();
    }

    int capture_graphs()
    {
        for (BatchGroup &g : groups) {
            g.release_graph();
            if (graph_epochs == 0) continue;
            REWL_HIP(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            const int status = g.enqueue_epochs(graph_epochs);
            hipGraph_t graph = nullptr;
            const hipError_t end = hipStreamEndCapture(g.stream, &graph);
            if (status) {
                if (graph) (void)hipGraphDestroy(graph);
                return status;
            }
            REWL_HIP(end);
            g.graph = graph;
            REWL_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        }
        return WLX_OK;
    }

    // (re)builds the launch groups for the current LDS budget
    int build_groups()
    {
        for (BatchGroup &g : groups) g.release();
        groups.clear();
        group_of.assign(engines.size(), 0);
        std::vector<Args> table(engines.size());
        std::vector<size_t> need(engines.size());
        static const size_t kClasses[] = {16 * 1024, 32 * 1024, 64 * 1024, 96 * 1024, kLdsBytes};
        std::map<size_t, uint32_t> by_class;
        for (size_t k = 0; k < engines.size(); ++k) {
            wlx_engine *e = engines[k];
            table[k] = e->args();
            size_t cls = 0;
            if (common) {
                need[k] = common_lds;
            } else {
                const bool hbm = e->bins_hbm || e->lds_walker > lds_budget;
                table[k].bins_hbm = hbm ? 1 : 0;
                need[k] = lds_walker_bytes(e->ws_stride, uint32_t(e->sys->n), hbm);
                if (need[k] > kLdsBytes)
                    return set_error(WLX_INVALID_ARGUMENT,
                                     "Fields and window bins exceed the LDS of a CU");
                for (size_t c : kClasses)
                    if (need[k] <= c) {
                        cls = c;
                        break;
                    }
            }
            auto it = by_class.find(cls);
            if (it == by_class.end()) {
                it = by_class.emplace(cls, uint32_t(groups.size())).first;
                groups.emplace_back();
            }
            group_of[k] = it->second;
            groups[it->second].members.push_back(uint32_t(k));
        }
        for (BatchGroup &g : groups) {
            std::vector<Args> sub;
            std::vector<std::pair<uint64_t, BatchSlot>> order; // heaviest systems first
            for (uint32_t m : g.members) {
                const uint32_t local = uint32_t(sub.size());
                sub.push_back(table[m]);
                g.lds = std::max(g.lds, need[m]);
                for (uint32_t w = 0; w < engines[m]->walkers; ++w)
                    order.push_back({engines[m]->sys->n, BatchSlot{local, w}});
            }
            std::stable_sort(order.begin(), order.end(),
                             [](const auto &x, const auto &y) { return x.first > y.first; });
            std::vector<BatchSlot> slots;
            for (const auto &o : order) slots.push_back(o.second);
            if (slots.size() >= (size_t(1) << 31))
                return set_error(WLX_INVALID_ARGUMENT, "Too many walkers");
            g.slots = uint32_t(slots.size());
            g.sample = common ? common
                              : (philox ? engines[0]->kernels.lds_ph : engines[0]->kernels.lds_xo);
            REWL_HIP(hipStreamCreate(&g.stream));
            if (int st = device_alloc(&g.d_table, sub.size())) return st;
            if (int st = device_alloc(&g.d_slots, slots.size())) return st;
            REWL_HIP(hipMemcpyAsync(g.d_table, sub.data(), sub.size() * sizeof(Args),
                                    hipMemcpyHostToDevice, g.stream));
            REWL_HIP(hipMemcpyAsync(g.d_slots, slots.data(), slots.size() * sizeof(BatchSlot),
                                    hipMemcpyHostToDevice, g.stream));
            REWL_HIP(hipStreamSynchronize(g.stream));
            if (int status = set_lds_limit(reinterpret_cast<const void *>(g.sample), g.lds))
                return status;
        }
        return capture_graphs();
    }

    int poll(uint64_t &done)
    {
        for (size_t k = 0; k < engines.size(); ++k)
            REWL_HIP(hipMemcpyAsync(&globs[k], engines[k]->d_glob, sizeof(Globals),
                                    hipMemcpyDeviceToHost, groups[group_of[k]].stream));
        for (BatchGroup &g : groups) REWL_HIP(hipStreamSynchronize(g.stream));
        done = 0;
        for (const Globals &g : globs) {
            if (g.escaped) return set_error(WLX_RUNTIME_ERROR, "Walker escaped its energy window");
            done += g.done ? 1 : 0;
        }
        return WLX_OK;
    }
};

extern "C" {

int wlx_batch_create(wlx_engine *const *engines, uint64_t count, wlx_batch **out)
{
    if (!engines || !out || count == 0) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    *out = nullptr;
    if (count > (uint64_t(1) << 20)) return set_error(WLX_INVALID_ARGUMENT, "Too many engines");
    auto b = std::make_unique<wlx_batch>();
    // every engine on one register kernel (one family and register layout): that kernel runs
    // the batch; any other mix runs the LDS walker (the trajectories are the same)
    SampleFn common = engines[0] ? engines[0]->sample_fn(true) : nullptr;
    size_t common_lds = 0;
    for (uint64_t k = 0; k < count; ++k) {
        wlx_engine *e = engines[k];
        if (int status = require_engine(e)) return status;
        for (uint64_t j = 0; j < k; ++j)
            if (engines[j] == e) return set_error(WLX_INVALID_ARGUMENT, "Engine listed twice");
        const bool philox = e->cfg.rng == WLX_RNG_PHILOX;
        if (k == 0) b->philox = philox;
        if (philox != b->philox)
            return set_error(WLX_INVALID_ARGUMENT, "Batched engines share one proposal stream");
        if (e->f32 || e->collect || e->trace_cap)
            return set_error(WLX_INVALID_ARGUMENT,
                             "Batched engines run FP64 fields without statistics or traces");
        if (e->sample_fn(true) != common) common = nullptr;
        REWL_HIP(hipStreamSynchronize(e->stream));
        common_lds = std::max(common_lds, e->lds_sample);
        b->engines.push_back(e);
    }
    if (common && common != engines[0]->kernels.lds_ph && common != engines[0]->kernels.lds_xo) {
        b->common = common;
        b->common_lds = common_lds;
        char name[64];
        kernel_name(engines[0], common, nullptr, name, sizeof name);
        b->kernel_name = name;
    } else {
        b->kernel_name = b->philox ? "lds_sample_kernel<true>" : "lds_sample_kernel<false>";
    }
    b->globs.resize(count);
    if (int status = b->build_groups()) return status;
    *out = b.release();
    return WLX_OK;
}

int wlx_batch_destroy(wlx_batch *b)
{
    delete b;
    return WLX_OK;
}

int wlx_batch_set_graph(wlx_batch *b, uint64_t epochs_per_graph)
{
    if (!b) return set_error(WLX_INVALID_ARGUMENT, "Null batch");
    if (epochs_per_graph > 4096)
        return set_error(WLX_INVALID_ARGUMENT, "At most 4096 epochs per graph");
    b->graph_epochs = epochs_per_graph;
    return b->capture_graphs();
}

int wlx_batch_set_lds_budget(wlx_batch *b, uint64_t bytes)
{
    if (!b) return set_error(WLX_INVALID_ARGUMENT, "Null batch");
    if (bytes == 0 || bytes > kLdsBytes)
        return set_error(WLX_INVALID_ARGUMENT, "The LDS budget must lie in (0, 160 KiB]");
    for (BatchGroup &g : b->groups) REWL_HIP(hipStreamSynchronize(g.stream));
    b->lds_budget = bytes;
    return b->build_groups();
}

int wlx_batch_layout(const wlx_batch *b, uint64_t *groups, uint64_t *bins_in_hbm,
                     uint64_t *max_lds_bytes)
{
    if (!b) return set_error(WLX_INVALID_ARGUMENT, "Null batch");
    uint64_t hbm = 0;
    size_t lds = 0;
    for (const BatchGroup &g : b->groups) lds = std::max(lds, g.lds);
    if (!b->common)
        for (const wlx_engine *e : b->engines) hbm += (e->bins_hbm || e->lds_walker > b->lds_budget);
    if (groups) *groups = b->groups.size();
    if (bins_in_hbm) *bins_in_hbm = hbm;
    if (max_lds_bytes) *max_lds_bytes = lds;
    return WLX_OK;
}

int wlx_batch_kernel(const wlx_batch *b, char *name, size_t capacity)
{
    if (!b || !name || capacity == 0) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    std::snprintf(name, capacity, "%s", b->kernel_name.c_str());
    return WLX_OK;
}

int wlx_batch_run(wlx_batch *b, uint64_t max_epochs, uint64_t *epochs, uint64_t *done_engines)
{
    if (!b) return set_error(WLX_INVALID_ARGUMENT, "Null batch");
    uint64_t ran = 0, done = 0;
    if (int status = b->poll(done)) return status;
    // one graph replay, or a growing chunk of plain launches, between polls of the engines'
    // flags: engines that finish stop on the device, and the call returns soon after the last
    uint64_t chunk = 1;
    const bool graphs = b->graph_epochs > 0 && !b->groups.empty() && b->groups[0].exec;
    while (ran < max_epochs && done < b->engines.size()) {
        uint64_t n;
        if (graphs && max_epochs - ran >= b->graph_epochs) {
            n = b->graph_epochs;
            for (BatchGroup &g : b->groups) REWL_HIP(hipGraphLaunch(g.exec, g.stream));
        } else {
            n = std::min(max_epochs - ran, chunk);
            for (BatchGroup &g : b->groups)
                if (int status = g.enqueue_epochs(n)) return status;
            chunk = std::min<uint64_t>(chunk * 2, 64);
        }
        ran += n;
        if (int status = b->poll(done)) return status;
    }
    if (epochs) *epochs = ran;
    if (done_engines) *done_engines = done;
    return WLX_OK;
}

int wlx_batch_autotune(wlx_batch *b, uint64_t epochs, uint64_t *chosen, double *seconds)
{
    if (!b) return set_error(WLX_INVALID_ARGUMENT, "Null batch");
    if (epochs == 0) return set_error(WLX_INVALID_ARGUMENT, "Autotuning needs at least one epoch");
    static const uint64_t kBudgets[] = {kLdsBytes, 64 * 1024, 32 * 1024};
    uint64_t best = b->lds_budget;
    double best_s = -1.0;
    for (int k = 0; k < 3; ++k) {
        if (seconds) seconds[k] = -1.0;
        if (b->common) continue; // the register kernels hold no bins in LDS to trade
        if (int status = wlx_batch_set_lds_budget(b, kBudgets[k])) return status;
        uint64_t ran = 0, done = 0;
        // the first epochs load the kernels; every variant continues the same trajectories
        if (int status = wlx_batch_run(b, b->graph_epochs ? b->graph_epochs : 1, &ran, &done))
            return status;
        if (done == b->engines.size()) break;
        const auto t0 = std::chrono::steady_clock::now();
        if (int status = wlx_batch_run(b, epochs, &ran, &done)) return status;
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (seconds) seconds[k] = s;
        if (best_s < 0.0 || s < best_s) {
            best_s = s;
            best = kBudgets[k];
        }
        if (done == b->engines.size()) break;
    }
    if (!b->common && best != b->lds_budget)
        if (int status = wlx_batch_set_lds_budget(b, best)) return status;
    if (chosen) *chosen = best;
    return WLX_OK;
}

} // extern "C"
