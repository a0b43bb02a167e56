/*
 * ORACLE — test infrastructure, not product code. See wl_oracle.h for what it is, what may
 * use it (tests/, smoke(), bench.py's cpu_baseline leg only) and how each function is pinned.
 *
 * Every function cites the reference file:line it restates
 * (Faikert/replica_exchange_wang_landau). Compile with -ffp-contract=off: the J build is
 * FMA-sensitive, and the reference is built for baseline x86-64 (no FMA instructions).
 */
#include "wl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------
 * errors: one message per thread, status codes of include/rewl.h
 * ---------------------------------------------------------------------------------------- */
static _Thread_local char g_error[2048];
static _Thread_local int g_status;

static int fail(int status, const char *format, ...)
{
    g_status = status;
    va_list args;
    va_start(args, format);
    vsnprintf(g_error, sizeof g_error, format, args);
    va_end(args);
    return status;
}

int orc_last_status(void) { return g_status; }

size_t orc_last_error(char *buffer, size_t capacity)
{
    size_t length = strlen(g_error);
    if (buffer && capacity) {
        size_t n = length < capacity - 1 ? length : capacity - 1;
        memcpy(buffer, g_error, n);
        buffer[n] = 0;
    }
    return length;
}

/* ------------------------------------------------------------------------------------------
 * RNG — rng.hpp:10-56
 * ---------------------------------------------------------------------------------------- */
static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t splitmix_step(uint64_t *x) /* rng.hpp:47-51 */
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void orc_xoshiro_seed(uint64_t seed, uint64_t stream, uint64_t s[4]) /* rng.hpp:52-55 */
{
    uint64_t x = seed ^ (0xd2b74407b1ce6e93ULL * (stream + 1));
    for (int k = 0; k < 4; ++k) s[k] = splitmix_step(&x);
}

uint64_t orc_xoshiro_next(uint64_t s[4]) /* rng.hpp:16-23 */
{
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

double orc_xoshiro_uniform(uint64_t s[4]) /* rng.hpp:25-27 */
{
    return (double)(orc_xoshiro_next(s) >> 11) * 0x1.0p-53;
}

uint64_t orc_xoshiro_bounded(uint64_t s[4], uint64_t bound) /* rng.hpp:29-36 */
{
    if (bound == 0) return 0;
    uint64_t threshold = (0 - bound) % bound;
    for (;;) {
        uint64_t x = orc_xoshiro_next(s);
        if (x >= threshold) return x % bound;
    }
}

/* Philox4x64-10: round function and Weyl key schedule of Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3" (SC'11). */
void orc_philox4x64(const uint64_t counter[4], const uint64_t key[2], uint64_t out[4])
{
    uint64_t c0 = counter[0], c1 = counter[1], c2 = counter[2], c3 = counter[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ULL * c0;
        unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ULL * c2;
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

#define PHILOX_WALK_TAG 0x57414c4bULL /* 'WALK' */
#define PHILOX_EXCH_TAG 0x45584348ULL /* 'EXCH' */

/* ------------------------------------------------------------------------------------------
 * physics — physics.cpp:16-30, 151-266
 * ---------------------------------------------------------------------------------------- */
static double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

int orc_minimum_image(const double d[3], const double box[3], int periodic, double out[3])
{ /* physics.cpp:151-160 */
    for (int k = 0; k < 3; ++k) {
        out[k] = d[k];
        if (periodic && box[k] > 0.0) out[k] -= box[k] * nearbyint(d[k] / box[k]);
    }
    return WLX_OK;
}

int orc_dipolar_pair(const double d[3], const double ai[3], const double aj[3], double scale,
                     double *out)
{ /* physics.cpp:162-170 */
    double r2 = dot3(d, d);
    if (!(r2 > 0.0)) return fail(WLX_INVALID_ARGUMENT, "Coincident dipole sites");
    double inv_r = 1.0 / sqrt(r2);
    double rhat[3] = {d[0] * inv_r, d[1] * inv_r, d[2] * inv_r};
    double angular = dot3(ai, aj) - 3.0 * dot3(ai, rhat) * dot3(aj, rhat);
    *out = scale * angular * inv_r * inv_r * inv_r;
    return WLX_OK;
}

static int validate_geometry(const double *pos, const double *axes, uint64_t n,
                             const double box[3], int periodic)
{ /* physics.cpp:34-56 */
    if (n == 0) return fail(WLX_INVALID_ARGUMENT, "Geometry contains no sites");
    if (periodic) {
        for (int k = 0; k < 3; ++k)
            if (!isfinite(box[k]) || !(box[k] >= 0.0))
                return fail(WLX_INVALID_ARGUMENT,
                            "Periodic box lengths must be finite and non-negative");
        if (box[0] == 0.0 && box[1] == 0.0 && box[2] == 0.0)
            return fail(WLX_INVALID_ARGUMENT, "At least one periodic box length must be positive");
    }
    for (uint64_t i = 0; i < n; ++i) {
        const double *p = pos + 3 * i;
        if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2]))
            return fail(WLX_INVALID_ARGUMENT, "Site coordinates must be finite");
        double length = sqrt(dot3(axes + 3 * i, axes + 3 * i));
        if (fabs(length - 1.0) > 1e-10)
            return fail(WLX_INVALID_ARGUMENT, "Ising axes must be normalized");
    }
    return WLX_OK;
}

struct orc_system {
    uint64_t n;
    int csr;
    double *dense;     /* n*n, row-major, zero diagonal */
    uint64_t *offsets; /* csr: n+1 */
    uint64_t *columns;
    double *values;
    uint64_t nnz;
};

/* pair value J_ij for i<j: d = p_j - p_i, minimum image, dipolar pair (physics.cpp:26-30) */
static int pair_value(const double *pos, const double *axes, uint64_t i, uint64_t j,
                      const double box[3], int periodic, double scale, double d_out[3],
                      double *value)
{
    double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                   pos[3 * j + 2] - pos[3 * i + 2]};
    orc_minimum_image(d, box, periodic, d_out);
    if (!value) return WLX_OK;
    return orc_dipolar_pair(d_out, axes + 3 * i, axes + 3 * j, scale, value);
}

orc_system *orc_system_create(const double *pos, const double *axes, uint64_t n,
                              const double box[3], int periodic, double scale, double cutoff,
                              int precision)
{
    if (validate_geometry(pos, axes, n, box, periodic) != WLX_OK) return NULL;
    if (precision != WLX_COUPLINGS_FP64 && precision != WLX_COUPLINGS_FP32) {
        fail(WLX_INVALID_ARGUMENT, "Unknown coupling precision");
        return NULL;
    }
    orc_system *s = calloc(1, sizeof *s);
    s->n = n;
    if (!(cutoff > 0.0)) { /* DenseCouplings, physics.cpp:172-184 */
        s->dense = calloc(n * n, sizeof(double));
        for (uint64_t i = 0; i < n; ++i)
            for (uint64_t j = i + 1; j < n; ++j) {
                double d[3], v;
                if (pair_value(pos, axes, i, j, box, periodic, scale, d, &v) != WLX_OK) {
                    orc_system_destroy(s);
                    return NULL;
                }
                if (precision == WLX_COUPLINGS_FP32) v = (double)(float)v;
                s->dense[i * n + j] = v;
                s->dense[j * n + i] = v;
            }
        return s;
    }
    /* CsrCouplings, physics.cpp:197-223: rows hold every j != i with norm(d) <= cutoff in
     * ascending column order; the value is computed from the (min, max) orientation. */
    s->csr = 1;
    s->offsets = calloc(n + 1, sizeof(uint64_t));
    uint64_t capacity = 1024, nnz = 0;
    s->columns = malloc(capacity * sizeof(uint64_t));
    s->values = malloc(capacity * sizeof(double));
    for (uint64_t i = 0; i < n; ++i) {
        for (uint64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            uint64_t lo = i < j ? i : j, hi = i < j ? j : i;
            double d[3], v;
            pair_value(pos, axes, lo, hi, box, periodic, scale, d, NULL);
            if (!(sqrt(dot3(d, d)) <= cutoff)) continue;
            if (orc_dipolar_pair(d, axes + 3 * lo, axes + 3 * hi, scale, &v) != WLX_OK) {
                orc_system_destroy(s);
                return NULL;
            }
            if (nnz == capacity) {
                capacity *= 2;
                s->columns = realloc(s->columns, capacity * sizeof(uint64_t));
                s->values = realloc(s->values, capacity * sizeof(double));
            }
            s->columns[nnz] = j;
            s->values[nnz] = precision == WLX_COUPLINGS_FP32 ? (double)(float)v : v;
            ++nnz;
        }
        s->offsets[i + 1] = nnz;
    }
    s->nnz = nnz;
    return s;
}

void orc_system_destroy(orc_system *s)
{
    if (!s) return;
    free(s->dense);
    free(s->offsets);
    free(s->columns);
    free(s->values);
    free(s);
}

int orc_system_info(const orc_system *s, uint64_t *n, uint64_t *backend, uint64_t *nnz)
{
    if (n) *n = s->n;
    if (backend) *backend = (uint64_t)s->csr;
    if (nnz) *nnz = s->csr ? s->nnz : s->n * s->n;
    return WLX_OK;
}

int orc_system_read_dense(const orc_system *s, double *out)
{
    if (!s->csr) {
        memcpy(out, s->dense, s->n * s->n * sizeof(double));
        return WLX_OK;
    }
    memset(out, 0, s->n * s->n * sizeof(double));
    for (uint64_t i = 0; i < s->n; ++i)
        for (uint64_t k = s->offsets[i]; k < s->offsets[i + 1]; ++k)
            out[i * s->n + s->columns[k]] = s->values[k];
    return WLX_OK;
}

int orc_system_read_csr(const orc_system *s, uint64_t *offsets, uint64_t *columns,
                        double *values)
{
    if (!s->csr) return fail(WLX_INVALID_ARGUMENT, "System uses the dense backend");
    memcpy(offsets, s->offsets, (s->n + 1) * sizeof(uint64_t));
    memcpy(columns, s->columns, s->nnz * sizeof(uint64_t));
    memcpy(values, s->values, s->nnz * sizeof(double));
    return WLX_OK;
}

/* local_fields (physics.cpp:238-251): h_i = sum_j at(i,j)*s_j, j ascending from 0.0. Zero
 * entries leave a sum unchanged, so the CSR row walk gives the dense result bit for bit. */
static void local_fields(const orc_system *s, const int8_t *spins, double *h)
{
    for (uint64_t i = 0; i < s->n; ++i) {
        double sum = 0.0;
        if (!s->csr) {
            const double *row = s->dense + i * s->n;
            for (uint64_t j = 0; j < s->n; ++j) sum += row[j] * (double)spins[j];
        } else {
            for (uint64_t k = s->offsets[i]; k < s->offsets[i + 1]; ++k)
                sum += s->values[k] * (double)spins[s->columns[k]];
        }
        h[i] = sum;
    }
}

/* total_energy (physics.cpp:253-259) */
static double total_energy(const orc_system *s, const int8_t *spins, const double *h)
{
    double energy = 0.0;
    for (uint64_t i = 0; i < s->n; ++i) energy += 0.5 * (double)spins[i] * h[i];
    return energy;
}

int orc_system_fields(const orc_system *s, const int8_t *spins, uint64_t count, double *fields,
                      double *energies)
{
    double *h = malloc(s->n * sizeof(double));
    for (uint64_t c = 0; c < count; ++c) {
        local_fields(s, spins + c * s->n, h);
        if (fields) memcpy(fields + c * s->n, h, s->n * sizeof(double));
        if (energies) energies[c] = total_energy(s, spins + c * s->n, h);
    }
    free(h);
    return WLX_OK;
}

/* add_flip_delta: dense column walk (physics.cpp:190-195), CSR row scatter (231-236) */
static void add_flip_delta(const orc_system *s, uint64_t flipped, int8_t old_spin, double *h)
{
    double multiplier = -2.0 * (double)old_spin;
    if (!s->csr) {
        for (uint64_t j = 0; j < s->n; ++j) h[j] += multiplier * s->dense[j * s->n + flipped];
    } else {
        for (uint64_t k = s->offsets[flipped]; k < s->offsets[flipped + 1]; ++k)
            h[s->columns[k]] += multiplier * s->values[k];
    }
}

/* ------------------------------------------------------------------------------------------
 * energy grid and windows — wang_landau.cpp:13-61
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    double minimum, maximum, width;
    uint64_t count;
} grid_t;

static int grid_validate(double mn, double mx, double w)
{ /* wang_landau.cpp:13-20 */
    if (!isfinite(mn) || !isfinite(mx) || !isfinite(w) || !(mx > mn) || !(w > 0.0))
        return fail(WLX_INVALID_ARGUMENT, "Invalid energy grid");
    double count = (mx - mn) / w;
    if (fabs(count - round(count)) > 1e-9 * fmax(1.0, count))
        return fail(WLX_INVALID_ARGUMENT, "Energy range must be an integer multiple of bin width");
    return WLX_OK;
}

int orc_grid_bins(double mn, double mx, double w, uint64_t *bins)
{ /* wang_landau.cpp:22-25 */
    int status = grid_validate(mn, mx, w);
    if (status) return status;
    *bins = (uint64_t)llround((mx - mn) / w);
    return WLX_OK;
}

/* EnergyGrid::index (wang_landau.cpp:27-34); -1 = nullopt */
static int64_t grid_index(const grid_t *g, double e)
{
    if (!isfinite(e) || e < g->minimum || e > g->maximum) return -1;
    uint64_t raw = (uint64_t)floor((e - g->minimum) / g->width);
    if (raw == g->count && e == g->maximum) raw = g->count - 1;
    if (raw >= g->count) return -1;
    return (int64_t)raw;
}

int64_t orc_grid_index(double mn, double mx, double w, double e)
{
    grid_t g = {mn, mx, w, (uint64_t)llround((mx - mn) / w)};
    return grid_index(&g, e);
}

int orc_partition_windows(uint64_t bins, uint64_t count, double overlap, uint64_t *be)
{ /* wang_landau.cpp:40-61 */
    if (bins == 0 || count == 0 || count > bins || overlap < 0.0 || overlap >= 1.0)
        return fail(WLX_INVALID_ARGUMENT, "Invalid energy-window partition");
    if (count == 1) {
        be[0] = 0;
        be[1] = bins;
        return WLX_OK;
    }
    double denominator = 1.0 + (double)(count - 1) * (1.0 - overlap);
    uint64_t size = (uint64_t)ceil((double)bins / denominator);
    if (size > bins) size = bins;
    if (size < 2) size = 2;
    uint64_t span = bins - size;
    for (uint64_t w = 0; w < count; ++w) {
        uint64_t begin = (uint64_t)llround((double)(w * span) / (double)(count - 1));
        be[2 * w] = begin;
        be[2 * w + 1] = begin + size;
    }
    be[2 * count - 1] = bins;
    for (uint64_t w = 1; w < count; ++w)
        if (be[2 * w] >= be[2 * w - 1])
            return fail(WLX_INVALID_ARGUMENT,
                        "Energy windows do not overlap; increase overlap or bins");
    return WLX_OK;
}

/* ------------------------------------------------------------------------------------------
 * walker — wang_landau.cpp:63-508
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    uint64_t id, window, begin, end;
    uint64_t rng[4];
    int8_t *spins;
    double *fields;
    double energy;
    double *log_g;      /* full grid, as the reference */
    uint64_t *histogram;
    uint8_t *active;
    uint8_t *refine;
    uint64_t active_count, refine_count;
    double factor;
    uint64_t attempted, accepted, forced, last_accepted;
    int stage;
    uint64_t last_check;
    double last_flatness;
    /* trace */
    uint64_t trace_count;
    uint32_t *trace_sites, *trace_bins;
    uint8_t *trace_flags;
    int escaped;
    /* 1/t clock offset of a warm start: factor = active / (attempted + clock_offset) */
    uint64_t clock_offset;
    /* pilot statistics (collect_window_statistics), wang_landau.cpp:104-118 */
    double *sq_disp;         /* full grid */
    uint64_t *disp_samples;  /* full grid */
    uint64_t round_trips;
    int rt_state;
    double rep_target[9], rep_dist[9], rep_energy[9];
    uint8_t rep_present[9];
    int8_t *rep_spins;       /* 9 x n */
} walker_t;

#define REPRESENTATIVES 9

struct orc_engine {
    const orc_system *sys;
    wlx_config cfg;
    grid_t grid;
    uint64_t windows, per_window, walkers;
    uint64_t *window_be;
    walker_t *w;
    uint64_t epoch;
    uint64_t done, converged;
    uint64_t exchange_attempted, exchange_accepted;
    uint64_t trace_capacity;
};

static int in_window(const walker_t *w, int64_t bin)
{
    return bin >= 0 && (uint64_t)bin >= w->begin && (uint64_t)bin < w->end;
}

/* update_current_bin (wang_landau.cpp:290-303) */
static int update_current_bin(orc_engine *e, walker_t *w)
{
    int64_t bin = grid_index(&e->grid, w->energy);
    if (!in_window(w, bin)) return fail(WLX_RUNTIME_ERROR, "Invalid current energy bin");
    ++w->histogram[bin];
    if (w->active[bin] == 0) {
        w->active[bin] = 1;
        ++w->active_count;
    }
    if (e->cfg.wl.nalivaiko_mod && w->refine[bin] == 0) {
        w->refine[bin] = 1;
        ++w->refine_count;
    }
    if (w->stage != WLX_STAGE_FROZEN) w->log_g[bin] += w->factor;
    return WLX_OK;
}

/* update_inverse_time_factor (wang_landau.cpp:309-315) */
static void update_inverse_time_factor(orc_engine *e, walker_t *w)
{
    if (w->stage != WLX_STAGE_INVERSE_TIME) return;
    if (w->active_count != 0 && w->attempted != 0)
        w->factor = (double)w->active_count / (double)(w->attempted + w->clock_offset);
    if (w->factor <= e->cfg.wl.final_factor) w->stage = WLX_STAGE_FROZEN;
}

static void flip(orc_engine *e, walker_t *w, uint64_t site, double proposed)
{
    int8_t old = w->spins[site];
    add_flip_delta(e->sys, site, old, w->fields);
    w->spins[site] = (int8_t)-old;
    w->energy = proposed;
}

/* attempt_flip (wang_landau.cpp:244-284); returns 1 accepted, 0 rejected, -1 escaped */
/* update_round_trip_state (wang_landau.cpp:317-330) */
static void update_round_trip_state(orc_engine *e, walker_t *w)
{
    if (!e->cfg.wl.collect_window_statistics) return;
    const double lower = e->grid.minimum + (double)w->begin * e->grid.width;
    const double upper = e->grid.minimum + (double)w->end * e->grid.width;
    const double margin = e->cfg.wl.round_trip_margin_fraction * (upper - lower);
    if (w->energy <= lower + margin) {
        if (w->rt_state == 2) {
            ++w->round_trips;
            w->rt_state = 1;
        } else if (w->rt_state == 0)
            w->rt_state = 1;
    } else if (w->energy >= upper - margin && w->rt_state == 1) {
        w->rt_state = 2;
    }
}

/* update_representatives (wang_landau.cpp:332-342) */
static void update_representatives(orc_engine *e, walker_t *w)
{
    if (!e->cfg.wl.collect_window_statistics) return;
    for (int r = 0; r < REPRESENTATIVES; ++r) {
        const double distance = fabs(w->energy - w->rep_target[r]);
        if (distance < w->rep_dist[r]) {
            w->rep_dist[r] = distance;
            w->rep_energy[r] = w->energy;
            memcpy(w->rep_spins + (size_t)r * e->sys->n, w->spins, e->sys->n);
            w->rep_present[r] = 1;
        }
    }
}

/* the end of every initialisation path (wang_landau.cpp:123-127, 177, 197, 219) */
static int finish_initialization(orc_engine *e, walker_t *w)
{
    int status = update_current_bin(e, w);
    update_round_trip_state(e, w);
    update_representatives(e, w);
    return status;
}

static int attempt_flip(orc_engine *e, walker_t *w)
{
    const uint64_t n = e->sys->n;
    uint64_t site;
    uint64_t philox_words[4];
    int philox = e->cfg.rng == WLX_RNG_PHILOX;
    if (philox) {
        uint64_t ctr[4] = {w->attempted >> 1, PHILOX_WALK_TAG, 0, 0};
        uint64_t key[2] = {e->cfg.seed, w->id};
        orc_philox4x64(ctr, key, philox_words);
        uint64_t x = philox_words[(w->attempted & 1) * 2];
        site = (uint64_t)(((unsigned __int128)x * n) >> 64);
    } else {
        site = orc_xoshiro_bounded(w->rng, n);
    }
    int64_t old_bin = grid_index(&e->grid, w->energy);
    if (!in_window(w, old_bin)) {
        w->escaped = 1;
        return -1;
    }
    int8_t old_spin = w->spins[site];
    double delta = -2.0 * (double)old_spin * w->fields[site]; /* flip_delta, physics.cpp:261-266 */
    double proposed = w->energy + delta;
    int64_t new_bin = grid_index(&e->grid, proposed);
    int accepted = 0, forced = 0, drew = 0, inside = 0;
    if (in_window(w, new_bin)) {
        inside = 1;
        double log_ratio = w->log_g[old_bin] - w->log_g[new_bin];
        int force_due = e->cfg.wl.force_accept_after_attempts != 0 &&
                        w->attempted - w->last_accepted >= e->cfg.wl.force_accept_after_attempts;
        if (log_ratio >= 0.0)
            accepted = 1;
        else if (force_due)
            accepted = forced = 1;
        else {
            double u = philox ? (double)(philox_words[(w->attempted & 1) * 2 + 1] >> 11) * 0x1.0p-53
                              : orc_xoshiro_uniform(w->rng);
            drew = 1;
            accepted = log(fmax(u, 0x1.0p-53)) < log_ratio;
        }
    }
    ++w->attempted;
    if (accepted) {
        flip(e, w, site, proposed);
        ++w->accepted;
        if (forced) ++w->forced;
        w->last_accepted = w->attempted;
    }
    if (e->cfg.wl.collect_window_statistics) { /* wang_landau.cpp:272-276 */
        const double displacement = accepted ? delta : 0.0;
        w->sq_disp[old_bin] += displacement * displacement;
        ++w->disp_samples[old_bin];
    }
    update_inverse_time_factor(e, w);
    update_current_bin(e, w);
    if (accepted) { /* wang_landau.cpp:278-282 */
        update_round_trip_state(e, w);
        update_representatives(e, w);
    }
    if (w->trace_count < e->trace_capacity) {
        uint64_t k = w->trace_count++;
        w->trace_sites[k] = (uint32_t)site;
        w->trace_bins[k] = (uint32_t)(accepted ? new_bin : old_bin);
        w->trace_flags[k] = (uint8_t)((accepted ? WLX_TRACE_ACCEPTED : 0) |
                                      (forced ? WLX_TRACE_FORCED : 0) |
                                      (drew ? WLX_TRACE_UNIFORM : 0) |
                                      (inside ? WLX_TRACE_IN_WINDOW : 0));
    }
    return accepted;
}

/* histogram_statistics (wang_landau.cpp:344-367) */
static wlx_histogram_stats histogram_statistics(const orc_engine *e, const walker_t *w)
{
    wlx_histogram_stats st = {0};
    const uint8_t *mask = e->cfg.wl.nalivaiko_mod ? w->refine : w->active;
    st.active_bins = e->cfg.wl.nalivaiko_mod ? w->refine_count : w->active_count;
    long double total = 0.0L;
    st.minimum = UINT64_MAX;
    for (uint64_t i = w->begin; i < w->end; ++i) {
        if (!mask[i]) continue;
        if (w->histogram[i] != 0) ++st.covered_bins;
        if (w->histogram[i] < st.minimum) st.minimum = w->histogram[i];
        total += (long double)w->histogram[i];
    }
    if (st.active_bins == 0) {
        st.minimum = 0;
        return st;
    }
    st.mean = (double)(total / (long double)st.active_bins);
    st.min_over_mean = st.mean > 0.0 ? (double)st.minimum / st.mean : 0.0;
    st.coverage = (double)st.covered_bins / (double)st.active_bins;
    return st;
}

/* flat / covered / ready_for_iteration (wang_landau.cpp:369-387) */
static int ready_for_iteration(const orc_engine *e, const walker_t *w)
{
    if (w->stage != WLX_STAGE_WANG_LANDAU) return 0;
    wlx_histogram_stats st = histogram_statistics(e, w);
    if (e->cfg.wl.inverse_time_enabled)
        return st.active_bins != 0 && st.covered_bins == st.active_bins;
    return st.active_bins != 0 && st.minimum >= e->cfg.wl.minimum_visits &&
           st.min_over_mean >= e->cfg.wl.flatness;
}

/* begin_next_iteration + freeze_if_finished (wang_landau.cpp:389-413) */
static int begin_next_iteration(orc_engine *e, walker_t *w)
{
    if (!ready_for_iteration(e, w)) return fail(WLX_LOGIC_ERROR, "WL iteration is not complete");
    w->factor *= 0.5;
    for (uint64_t i = w->begin; i < w->end; ++i) w->histogram[i] = 0;
    if (e->cfg.wl.nalivaiko_mod) {
        for (uint64_t i = w->begin; i < w->end; ++i) w->refine[i] = 0;
        w->refine_count = 0;
    }
    if (e->cfg.wl.inverse_time_enabled) {
        const uint64_t clock = w->attempted + w->clock_offset;
        double inverse_time = clock == 0 ? 1.0 : (double)w->active_count / (double)clock;
        if (w->factor <= inverse_time) {
            w->stage = WLX_STAGE_INVERSE_TIME;
            w->factor = inverse_time;
        }
    }
    if (w->factor <= e->cfg.wl.final_factor) w->stage = WLX_STAGE_FROZEN;
    return WLX_OK;
}

/* WangLandauWalker construction + target-Metropolis initialisation (wang_landau.cpp:63-242).
 * Parameter validation happens once per engine (orc_engine_create). */
static int initialize_walker(orc_engine *e, walker_t *w, const int8_t *initial)
{
    const orc_system *s = e->sys;
    const wlx_params *p = &e->cfg.wl;
    const uint64_t n = s->n;
    int supplied = initial != NULL;
    for (uint64_t i = 0; i < n; ++i) {
        int8_t v = supplied ? initial[i] : 1;
        if (v != -1 && v != 1) return fail(WLX_INVALID_ARGUMENT, "Spins must be +/-1");
        w->spins[i] = v;
    }
    local_fields(s, w->spins, w->fields);
    w->energy = total_energy(s, w->spins, w->fields);
    memset(w->log_g, 0, e->grid.count * sizeof(double));
    memset(w->histogram, 0, e->grid.count * sizeof(uint64_t));
    memset(w->active, 0, e->grid.count);
    memset(w->refine, 0, e->grid.count);
    w->active_count = w->refine_count = 0;
    w->factor = 1.0;
    w->attempted = w->accepted = w->forced = w->last_accepted = 0;
    w->stage = WLX_STAGE_WANG_LANDAU;
    w->last_check = 0;
    w->clock_offset = 0;
    w->last_flatness = 0.0;
    orc_xoshiro_seed(e->cfg.seed, w->id, w->rng);
    if (p->collect_window_statistics) { /* wang_landau.cpp:104-118 */
        memset(w->sq_disp, 0, e->grid.count * sizeof(double));
        memset(w->disp_samples, 0, e->grid.count * sizeof(uint64_t));
        w->round_trips = 0;
        w->rt_state = 0;
        const double lo = e->grid.minimum + (double)w->begin * e->grid.width;
        const double hi = e->grid.minimum + (double)w->end * e->grid.width;
        w->rep_target[0] = lo;
        for (int r = 1; r < REPRESENTATIVES; ++r)
            w->rep_target[r] = lo + (hi - lo) * ((double)r - 0.5) / (double)(REPRESENTATIVES - 1);
        for (int r = 0; r < REPRESENTATIVES; ++r) {
            w->rep_dist[r] = INFINITY;
            w->rep_energy[r] = 0.0;
            w->rep_present[r] = 0;
        }
    }

    int64_t initial_bin = grid_index(&e->grid, w->energy);
    if (supplied && in_window(w, initial_bin)) return finish_initialization(e, w);

    const double lower = e->grid.minimum + (double)w->begin * e->grid.width;
    const double upper = e->grid.minimum + (double)w->end * e->grid.width;
    const double span = upper - lower;
    const double target = 0.5 * (lower + upper);
    const double target_fraction = supplied ? 1.0 : p->initialization_target_fraction;
    const double half = 0.5 * target_fraction * span;
    double t0 = p->initialization_temperature_fraction * span;
    if (e->grid.width >= t0) t0 = e->grid.width; /* std::max(width, fraction*span) */
    double tmax = p->initialization_max_temperature_fraction * span;
    if (t0 >= tmax) tmax = t0;
    double temperature = t0;
    const uint64_t stall = p->initialization_stall_attempts_per_spin > UINT64_MAX / n
                               ? UINT64_MAX
                               : p->initialization_stall_attempts_per_spin * n;
    uint64_t since = 0, increases = 0, restarts = 0;
    double highest = temperature;
    double min_reached = w->energy, max_reached = w->energy, closest = w->energy;
    double closest_distance = fabs(w->energy - target);

#define RECORD(c)                                                                              \
    ({                                                                                         \
        double c_ = (c);                                                                       \
        if (c_ < min_reached) min_reached = c_;                                                \
        if (max_reached < c_) max_reached = c_;                                                \
        double dist_ = fabs(c_ - target);                                                      \
        int better_ = dist_ < closest_distance;                                                \
        if (better_) {                                                                         \
            closest_distance = dist_;                                                          \
            closest = c_;                                                                      \
        }                                                                                      \
        better_;                                                                               \
    })
#define IN_BAND(c)                                                                             \
    (in_window(w, grid_index(&e->grid, (c))) && fabs((c) - target) <= half)

    for (uint64_t step = 0; step < p->initialization_max_attempts; ++step) {
        if (IN_BAND(w->energy)) return finish_initialization(e, w);
        if (since >= stall) {
            if (temperature < tmax) {
                double raised = temperature * p->initialization_temperature_multiplier;
                temperature = raised < tmax ? raised : tmax;
                if (highest < temperature) highest = temperature;
                ++increases;
            } else {
                for (uint64_t i = 0; i < n; ++i)
                    w->spins[i] = orc_xoshiro_bounded(w->rng, 2) == 0 ? (int8_t)-1 : (int8_t)1;
                local_fields(s, w->spins, w->fields);
                w->energy = total_energy(s, w->spins, w->fields);
                (void)RECORD(w->energy);
                temperature = t0;
                ++restarts;
            }
            since = 0;
            if (IN_BAND(w->energy)) return finish_initialization(e, w);
        }
        uint64_t i = orc_xoshiro_bounded(w->rng, n);
        double delta = -2.0 * (double)w->spins[i] * w->fields[i];
        double proposed = w->energy + delta;
        double log_acceptance =
            (fabs(w->energy - target) - fabs(proposed - target)) / temperature;
        int improved = 0;
        if (log_acceptance >= 0.0 ||
            log(fmax(orc_xoshiro_uniform(w->rng), 0x1.0p-53)) < log_acceptance) {
            flip(e, w, i, proposed);
            improved = RECORD(w->energy);
        }
        if (improved)
            since = 0;
        else if (since < UINT64_MAX)
            ++since;
    }
    if (IN_BAND(w->energy)) return finish_initialization(e, w);
#undef RECORD
#undef IN_BAND
    char closest_bin[64];
    int64_t cb = grid_index(&e->grid, closest);
    if (cb >= 0)
        snprintf(closest_bin, sizeof closest_bin, " (bin %lld)", (long long)cb);
    else
        snprintf(closest_bin, sizeof closest_bin, " (outside the energy grid)");
    double band_distance = closest_distance - half;
    if (band_distance < 0.0) band_distance = 0.0;
    return fail(WLX_RUNTIME_ERROR,
                "Target Metropolis initialization did not reach target band [%.17g, %.17g] for "
                "energy window bins [%llu, %llu), energy bounds [%.17g, %.17g%c after %llu "
                "attempts; reached energy range [%.17g, %.17g], closest energy %.17g%s, "
                "distance to target band %.17g; adaptive search used %llu temperature "
                "increases up to %.17g and %llu random restarts (stall interval %llu attempts)",
                target - half, target + half, (unsigned long long)w->begin,
                (unsigned long long)w->end, lower, upper, w->end == e->grid.count ? ']' : ')',
                (unsigned long long)p->initialization_max_attempts, min_reached, max_reached,
                closest, closest_bin, band_distance, (unsigned long long)increases, highest,
                (unsigned long long)restarts, (unsigned long long)stall);
}

/* ------------------------------------------------------------------------------------------
 * engine — run_rewl (parallel.cpp:276-581), shared-memory path
 * ---------------------------------------------------------------------------------------- */
static int validate_parameters(const wlx_params *p)
{ /* wang_landau.cpp:74-90 */
    if (!(p->flatness > 0.0 && p->flatness <= 1.0) || !(p->final_factor > 0.0) ||
        p->check_interval_attempts == 0 || p->initialization_max_attempts == 0 ||
        !(p->initialization_target_fraction > 0.0 && p->initialization_target_fraction <= 1.0) ||
        !isfinite(p->initialization_temperature_fraction) ||
        !(p->initialization_temperature_fraction > 0.0) ||
        p->initialization_stall_attempts_per_spin == 0 ||
        !isfinite(p->initialization_temperature_multiplier) ||
        !(p->initialization_temperature_multiplier > 1.0) ||
        !isfinite(p->initialization_max_temperature_fraction) ||
        p->initialization_max_temperature_fraction < p->initialization_temperature_fraction ||
        !isfinite(p->round_trip_margin_fraction) || p->round_trip_margin_fraction < 0.0 ||
        p->round_trip_margin_fraction >= 0.5)
        return fail(WLX_INVALID_ARGUMENT, "Invalid WL parameters");
    return WLX_OK;
}

orc_engine *orc_engine_create(const orc_system *s, const wlx_config *c)
{
    uint64_t bins;
    if (orc_grid_bins(c->emin, c->emax, c->bin_width, &bins) != WLX_OK) return NULL;
    if (c->windows == 0 || c->walkers_per_window == 0 || c->exchange_interval_attempts == 0) {
        fail(WLX_INVALID_ARGUMENT, "Invalid run configuration");
        return NULL;
    }
    if (validate_parameters(&c->wl) != WLX_OK) return NULL;
    if (c->fields != WLX_FIELDS_FP64) { /* the reference's arithmetic only */
        fail(WLX_INVALID_ARGUMENT, "The oracle restates FP64 fields only");
        return NULL;
    }
    orc_engine *e = calloc(1, sizeof *e);
    e->sys = s;
    e->cfg = *c;
    e->cfg.explicit_windows = NULL;
    e->grid = (grid_t){c->emin, c->emax, c->bin_width, bins};
    e->windows = c->windows;
    e->per_window = c->walkers_per_window;
    e->walkers = e->windows * e->per_window;
    e->window_be = malloc(2 * e->windows * sizeof(uint64_t));
    if (c->explicit_windows)
        memcpy(e->window_be, c->explicit_windows, 2 * e->windows * sizeof(uint64_t));
    else if (orc_partition_windows(bins, e->windows, c->overlap, e->window_be) != WLX_OK) {
        orc_engine_destroy(e);
        return NULL;
    }
    for (uint64_t k = 0; k < e->windows; ++k)
        if (e->window_be[2 * k] >= e->window_be[2 * k + 1] || e->window_be[2 * k + 1] > bins) {
            fail(WLX_INVALID_ARGUMENT, "Invalid walker energy window");
            orc_engine_destroy(e);
            return NULL;
        }
    e->w = calloc(e->walkers, sizeof(walker_t));
    for (uint64_t k = 0; k < e->walkers; ++k) {
        walker_t *w = &e->w[k];
        w->id = k;
        w->window = k / e->per_window;
        w->begin = e->window_be[2 * w->window];
        w->end = e->window_be[2 * w->window + 1];
        w->spins = calloc(s->n, 1);
        w->fields = calloc(s->n, sizeof(double));
        w->log_g = calloc(bins, sizeof(double));
        w->histogram = calloc(bins, sizeof(uint64_t));
        w->active = calloc(bins, 1);
        w->refine = calloc(bins, 1);
        w->factor = 1.0;
        if (c->wl.collect_window_statistics) {
            w->sq_disp = calloc(bins, sizeof(double));
            w->disp_samples = calloc(bins, sizeof(uint64_t));
            w->rep_spins = calloc(REPRESENTATIVES * s->n, 1);
        }
    }
    return e;
}

void orc_engine_destroy(orc_engine *e)
{
    if (!e) return;
    for (uint64_t k = 0; e->w && k < e->walkers; ++k) {
        walker_t *w = &e->w[k];
        free(w->spins);
        free(w->fields);
        free(w->log_g);
        free(w->histogram);
        free(w->active);
        free(w->refine);
        free(w->trace_sites);
        free(w->trace_bins);
        free(w->trace_flags);
        free(w->sq_disp);
        free(w->disp_samples);
        free(w->rep_spins);
    }
    free(e->w);
    free(e->window_be);
    free(e);
}

int orc_engine_initialize(orc_engine *e, const int8_t *initial, const uint8_t *supplied)
{ /* parallel.cpp:304-352: serial construction, first failure reported with its position */
    int first_status = WLX_OK;
    uint64_t first = UINT64_MAX;
    char *message = NULL;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < (int64_t)e->walkers; ++k) {
        const int8_t *spins = (initial && supplied && supplied[k]) ? initial + k * e->sys->n : NULL;
        int status = initialize_walker(e, &e->w[k], spins);
        if (status != WLX_OK) {
#pragma omp critical
            if ((uint64_t)k < first) {
                first = (uint64_t)k;
                first_status = status;
                free(message);
                message = strdup(g_error);
            }
        }
    }
    if (first_status != WLX_OK) {
        int status = fail(WLX_RUNTIME_ERROR,
                          "Walker initialization failed: window %llu, local walker %llu: %s",
                          (unsigned long long)(first / e->per_window),
                          (unsigned long long)(first % e->per_window), message);
        free(message);
        (void)first_status;
        return status;
    }
    return WLX_OK;
}

static int run_batch(orc_engine *e, walker_t *w, uint64_t count)
{
    for (uint64_t i = 0; i < count; ++i)
        if (attempt_flip(e, w) < 0) return -1;
    return 0;
}

/* attempts per parallel region below which the walkers run on one thread: the fork/join of
 * a region costs more than a few thousand attempts (C1: 160 attempts per epoch) */
#define PARALLEL_ATTEMPTS 32768u

int orc_engine_sample(orc_engine *e, uint64_t attempts)
{
    int escaped = 0;
#pragma omp parallel for schedule(static) reduction(| : escaped) \
    if (e->walkers * attempts >= PARALLEL_ATTEMPTS)
    for (int64_t k = 0; k < (int64_t)e->walkers; ++k) escaped |= run_batch(e, &e->w[k], attempts) != 0;
    if (escaped) return fail(WLX_RUNTIME_ERROR, "Walker escaped its energy window");
    return WLX_OK;
}

/* flatness check cadence (parallel.cpp:395-407) */
static void check_walker(orc_engine *e, walker_t *w)
{
    if (w->stage != WLX_STAGE_WANG_LANDAU) return;
    const uint64_t interval =
        e->cfg.wl.check_interval_attempts > 1 ? e->cfg.wl.check_interval_attempts : 1;
    const int unlimited = e->cfg.max_attempts == 0;
    int at_limit = !unlimited && w->attempted >= e->cfg.max_attempts;
    if (w->attempted - w->last_check < interval && !at_limit) return;
    w->last_check = w->attempted;
    wlx_histogram_stats st = histogram_statistics(e, w);
    w->last_flatness = e->cfg.wl.inverse_time_enabled ? st.coverage : st.min_over_mean;
    if (ready_for_iteration(e, w)) begin_next_iteration(e, w);
}

int orc_engine_check(orc_engine *e)
{
    for (uint64_t k = 0; k < e->walkers; ++k) check_walker(e, &e->w[k]);
    return WLX_OK;
}

/* exchange_log_probability (wang_landau.cpp:497-508) */
static double exchange_log_probability(const orc_engine *e, const walker_t *a,
                                       const walker_t *b)
{
    int64_t x_here = grid_index(&e->grid, a->energy);
    int64_t y_here = grid_index(&e->grid, b->energy);
    if (!in_window(a, x_here) || !in_window(a, y_here) || !in_window(b, x_here) ||
        !in_window(b, y_here))
        return -INFINITY;
    return a->log_g[x_here] - a->log_g[y_here] + b->log_g[y_here] - b->log_g[x_here];
}

/* swap_configuration (wang_landau.cpp:483-495): estimators stay with their walkers; with
 * collect_window_statistics both walkers update their round-trip state and representatives
 * from the configuration they received */
static void swap_configuration(orc_engine *e, walker_t *a, walker_t *b)
{
    int8_t *spins = a->spins;
    a->spins = b->spins;
    b->spins = spins;
    double *fields = a->fields;
    a->fields = b->fields;
    b->fields = fields;
    double energy = a->energy;
    a->energy = b->energy;
    b->energy = energy;
    update_round_trip_state(e, a);
    update_round_trip_state(e, b);
    update_representatives(e, a);
    update_representatives(e, b);
}

/* local_exchange (parallel.cpp:157-166) */
static int local_exchange(orc_engine *e, walker_t *left, walker_t *right, uint64_t epoch,
                          uint64_t boundary)
{
    double lnp = exchange_log_probability(e, left, right);
    if (!isfinite(lnp)) return 0;
    int accept = lnp >= 0.0;
    if (!accept) {
        double u;
        if (e->cfg.rng == WLX_RNG_PHILOX) {
            uint64_t ctr[4] = {epoch, PHILOX_EXCH_TAG, 0, 0}, key[2] = {e->cfg.seed, boundary};
            uint64_t out[4];
            orc_philox4x64(ctr, key, out);
            u = (double)(out[0] >> 11) * 0x1.0p-53;
        } else {
            uint64_t rng[4];
            orc_xoshiro_seed(e->cfg.seed ^ (epoch * 0x9e3779b97f4a7c15ULL), boundary, rng);
            u = orc_xoshiro_uniform(rng);
        }
        accept = log(fmax(u, 0x1.0p-53)) < lnp;
    }
    if (accept) swap_configuration(e, left, right);
    return accept;
}

int orc_engine_exchange(orc_engine *e, uint64_t epoch, uint64_t *attempted, uint64_t *accepted)
{ /* parallel.cpp:410-421 */
    uint64_t tried = 0, taken = 0;
    const uint64_t pairs = e->per_window;
    for (uint64_t b = epoch % 2; b + 1 < e->windows; b += 2) {
        walker_t *left = e->w + b * e->per_window;
        walker_t *right = e->w + (b + 1) * e->per_window;
        for (uint64_t p = 0; p < pairs; ++p) {
            ++tried;
            taken += (uint64_t)local_exchange(e, &left[p], &right[(p + epoch) % pairs], epoch, b);
        }
    }
    e->exchange_attempted += tried;
    e->exchange_accepted += taken;
    if (attempted) *attempted = tried;
    if (accepted) *accepted = taken;
    return WLX_OK;
}

int orc_engine_status(orc_engine *e, wlx_run_status *st)
{
    memset(st, 0, sizeof *st);
    st->epoch = e->epoch;
    st->done = e->done;
    int converged = 1;
    for (uint64_t k = 0; k < e->walkers; ++k) {
        st->attempted += e->w[k].attempted;
        st->accepted += e->w[k].accepted;
        st->forced_accepted += e->w[k].forced;
        converged &= e->w[k].stage == WLX_STAGE_FROZEN;
    }
    st->converged = (uint64_t)converged;
    st->exchange_attempted = e->exchange_attempted;
    st->exchange_accepted = e->exchange_accepted;
    return WLX_OK;
}

int orc_engine_read_sampling(orc_engine *e, uint64_t k, double *squared, uint64_t *samples,
                             uint64_t *round_trips, double *rep_energy, uint8_t *rep_present,
                             int8_t *rep_spins)
{
    if (k >= e->walkers) return fail(WLX_OUT_OF_RANGE, "Walker index out of range");
    if (!e->cfg.wl.collect_window_statistics)
        return fail(WLX_LOGIC_ERROR, "Window statistics are not collected");
    const walker_t *w = &e->w[k];
    const uint64_t bins = e->grid.count;
    if (squared) memcpy(squared, w->sq_disp, bins * sizeof(double));
    if (samples) memcpy(samples, w->disp_samples, bins * sizeof(uint64_t));
    if (round_trips) *round_trips = w->round_trips;
    for (int r = 0; r < REPRESENTATIVES; ++r) {
        if (rep_energy) rep_energy[r] = w->rep_energy[r];
        if (rep_present) rep_present[r] = w->rep_present[r];
        if (rep_spins && w->rep_present[r])
            memcpy(rep_spins + (size_t)r * e->sys->n, w->rep_spins + (size_t)r * e->sys->n,
                   e->sys->n);
    }
    return WLX_OK;
}

/* wlx_engine_warm_start: window log g from `log_g`, ln f = factor in the WL stage, empty
 * histograms and masks, 1/t clock offset ceil(window_bins / factor) (no reference counterpart) */
int orc_engine_warm_start(orc_engine *e, const double *log_g, double factor)
{
    if (!log_g) return fail(WLX_INVALID_ARGUMENT, "Null argument");
    if (!(factor > 0.0) || !(factor <= 1.0))
        return fail(WLX_INVALID_ARGUMENT, "The initial ln f must lie in (0, 1]");
    for (uint64_t b = 0; b < e->grid.count; ++b)
        if (!isfinite(log_g[b]))
            return fail(WLX_INVALID_ARGUMENT, "The warm-start log g must be finite on every bin");
    for (uint64_t k = 0; k < e->walkers; ++k) {
        walker_t *w = &e->w[k];
        double lowest = log_g[w->begin];
        for (uint64_t b = w->begin + 1; b < w->end; ++b)
            if (log_g[b] < lowest) lowest = log_g[b];
        for (uint64_t b = 0; b < e->grid.count; ++b) {
            w->log_g[b] = (b >= w->begin && b < w->end) ? log_g[b] - lowest : 0.0;
            w->histogram[b] = 0;
            w->active[b] = 0;
            w->refine[b] = 0;
        }
        w->factor = factor;
        w->stage = WLX_STAGE_WANG_LANDAU;
        w->active_count = w->refine_count = 0;
        w->last_check = w->attempted;
        w->clock_offset = (uint64_t)ceil((double)(w->end - w->begin) / factor);
    }
    return WLX_OK;
}

int orc_engine_set_walker_cadence(orc_engine *e, uint64_t k, uint64_t last_flatness_check,
                                  double last_flatness)
{
    if (k >= e->walkers) return fail(WLX_OUT_OF_RANGE, "Walker index out of range");
    e->w[k].last_check = last_flatness_check;
    e->w[k].last_flatness = last_flatness;
    return WLX_OK;
}

int orc_engine_set_run_state(orc_engine *e, uint64_t epoch, uint64_t exchange_attempted,
                             uint64_t exchange_accepted)
{
    e->epoch = epoch;
    e->done = 0;
    e->exchange_attempted = exchange_attempted;
    e->exchange_accepted = exchange_accepted;
    return WLX_OK;
}

int orc_engine_run(orc_engine *e, uint64_t max_epochs, wlx_run_status *status)
{ /* parallel.cpp:376-534 */
    const int unlimited = e->cfg.max_attempts == 0;
    for (uint64_t done_epochs = 0; !e->done && done_epochs < max_epochs; ++done_epochs) {
        int limit_reached = !unlimited, escaped = 0;
#pragma omp parallel for schedule(static) reduction(| : escaped) \
    if (e->walkers * e->cfg.exchange_interval_attempts >= PARALLEL_ATTEMPTS)
        for (int64_t k = 0; k < (int64_t)e->walkers; ++k) {
            walker_t *w = &e->w[k];
            uint64_t batch = e->cfg.exchange_interval_attempts;
            if (!unlimited) {
                uint64_t remaining =
                    w->attempted < e->cfg.max_attempts ? e->cfg.max_attempts - w->attempted : 0;
                if (remaining < batch) batch = remaining;
            }
            escaped |= run_batch(e, w, batch) != 0;
        }
        if (escaped) return fail(WLX_RUNTIME_ERROR, "Walker escaped its energy window");
        for (uint64_t k = 0; k < e->walkers; ++k) {
            if (!unlimited && e->w[k].attempted < e->cfg.max_attempts) limit_reached = 0;
            check_walker(e, &e->w[k]);
        }
        orc_engine_exchange(e, e->epoch, NULL, NULL);
        ++e->epoch;
        int frozen = 1;
        for (uint64_t k = 0; k < e->walkers; ++k) frozen &= e->w[k].stage == WLX_STAGE_FROZEN;
        e->done = (uint64_t)(frozen || limit_reached);
    }
    if (status) orc_engine_status(e, status);
    return WLX_OK;
}

int orc_engine_walkers(const orc_engine *e, uint64_t *walkers, uint64_t *spins, uint64_t *bins)
{
    if (walkers) *walkers = e->walkers;
    if (spins) *spins = e->sys->n;
    if (bins) *bins = e->grid.count;
    return WLX_OK;
}

static int walker_index(const orc_engine *e, uint64_t k)
{
    if (k >= e->walkers) return fail(WLX_OUT_OF_RANGE, "Walker index out of range");
    return WLX_OK;
}

int orc_engine_read_walker(orc_engine *e, uint64_t k, wlx_walker_info *info, int8_t *spins,
                           double *fields, double *log_g, uint64_t *histogram, uint8_t *active,
                           uint8_t *refine)
{
    if (walker_index(e, k)) return WLX_OUT_OF_RANGE;
    const walker_t *w = &e->w[k];
    const uint64_t n = e->sys->n, bins = e->grid.count;
    if (info) {
        memset(info, 0, sizeof *info);
        info->walker_id = w->id;
        info->window = w->window;
        info->window_begin = w->begin;
        info->window_end = w->end;
        info->energy = w->energy;
        info->factor = w->factor;
        info->attempted = w->attempted;
        info->accepted = w->accepted;
        info->forced_accepted = w->forced;
        info->last_accepted_attempt = w->last_accepted;
        info->stage = (uint64_t)w->stage;
        info->active_bins = w->active_count;
        info->refinement_active_bins = w->refine_count;
        memcpy(info->rng_state, w->rng, sizeof w->rng);
        info->last_flatness_check = w->last_check;
        info->last_flatness = w->last_flatness;
        info->clock_offset = w->clock_offset;
    }
    if (spins) memcpy(spins, w->spins, n);
    if (fields) memcpy(fields, w->fields, n * sizeof(double));
    if (log_g) memcpy(log_g, w->log_g, bins * sizeof(double));
    if (histogram) memcpy(histogram, w->histogram, bins * sizeof(uint64_t));
    if (active) memcpy(active, w->active, bins);
    if (refine) memcpy(refine, w->refine, bins);
    return WLX_OK;
}

int orc_engine_write_walker(orc_engine *e, uint64_t k, const wlx_walker_info *s,
                            const int8_t *spins, const double *fields, const double *log_g,
                            const uint64_t *histogram, const uint8_t *active)
{ /* restore (wang_landau.cpp:420-466) */
    if (walker_index(e, k)) return WLX_OUT_OF_RANGE;
    walker_t *w = &e->w[k];
    const uint64_t n = e->sys->n, bins = e->grid.count;
    if (s->walker_id != w->id || !spins || !fields || !log_g || !histogram || !active)
        return fail(WLX_INVALID_ARGUMENT, "Checkpoint is incompatible with walker");
    if (s->last_accepted_attempt > s->attempted)
        return fail(WLX_INVALID_ARGUMENT, "Checkpoint has an invalid last accepted attempt");
    if (s->forced_accepted > s->accepted)
        return fail(WLX_INVALID_ARGUMENT, "Checkpoint has an invalid forced acceptance count");
    if (!e->cfg.wl.inverse_time_enabled && s->stage == WLX_STAGE_INVERSE_TIME)
        return fail(WLX_INVALID_ARGUMENT, "Checkpoint uses disabled inverse-time refinement");
    memcpy(w->spins, spins, n);
    memcpy(w->fields, fields, n * sizeof(double));
    w->energy = s->energy;
    memcpy(w->log_g, log_g, bins * sizeof(double));
    memcpy(w->histogram, histogram, bins * sizeof(uint64_t));
    memcpy(w->active, active, bins);
    w->factor = s->factor;
    w->attempted = s->attempted;
    w->accepted = s->accepted;
    w->active_count = 0;
    memset(w->refine, 0, bins);
    w->refine_count = 0;
    for (uint64_t i = w->begin; i < w->end; ++i) {
        w->active[i] = w->active[i] != 0 ? 1 : 0;
        w->active_count += w->active[i];
        if (e->cfg.wl.nalivaiko_mod) {
            w->refine[i] = (uint8_t)(w->histogram[i] != 0);
            w->refine_count += w->refine[i];
        }
    }
    w->forced = s->forced_accepted;
    w->last_accepted = s->last_accepted_attempt;
    w->stage = (int)s->stage;
    w->clock_offset = s->clock_offset;
    memcpy(w->rng, s->rng_state, sizeof w->rng);
    double *exact = malloc(n * sizeof(double));
    local_fields(e->sys, w->spins, exact);
    double energy = total_energy(e->sys, w->spins, exact);
    int status = WLX_OK;
    if (fabs(energy - w->energy) > 1e-9 * fmax(1.0, fabs(energy)))
        status = fail(WLX_RUNTIME_ERROR, "Checkpoint energy does not match spin configuration");
    else
        for (uint64_t i = 0; i < n; ++i)
            if (fabs(exact[i] - w->fields[i]) > 1e-9 * fmax(1.0, fabs(exact[i]))) {
                status = fail(WLX_RUNTIME_ERROR,
                              "Checkpoint local fields do not match spin configuration");
                break;
            }
    free(exact);
    return status;
}

int orc_engine_histogram_statistics(orc_engine *e, uint64_t k, wlx_histogram_stats *out)
{
    if (walker_index(e, k)) return WLX_OUT_OF_RANGE;
    *out = histogram_statistics(e, &e->w[k]);
    return WLX_OK;
}

int orc_engine_ready(orc_engine *e, uint64_t k, uint64_t *ready)
{
    if (walker_index(e, k)) return WLX_OUT_OF_RANGE;
    *ready = (uint64_t)ready_for_iteration(e, &e->w[k]);
    return WLX_OK;
}

int orc_engine_next_iteration(orc_engine *e, uint64_t k)
{
    if (walker_index(e, k)) return WLX_OUT_OF_RANGE;
    return begin_next_iteration(e, &e->w[k]);
}

int orc_engine_trace(orc_engine *e, uint64_t capacity)
{
    e->trace_capacity = capacity;
    for (uint64_t k = 0; k < e->walkers; ++k) {
        walker_t *w = &e->w[k];
        free(w->trace_sites);
        free(w->trace_bins);
        free(w->trace_flags);
        w->trace_sites = calloc(capacity ? capacity : 1, sizeof(uint32_t));
        w->trace_bins = calloc(capacity ? capacity : 1, sizeof(uint32_t));
        w->trace_flags = calloc(capacity ? capacity : 1, 1);
        w->trace_count = 0;
    }
    return WLX_OK;
}

int orc_engine_read_trace(orc_engine *e, uint64_t k, uint64_t *count, uint32_t *sites,
                          uint32_t *bins, uint8_t *flags)
{
    if (walker_index(e, k)) return WLX_OUT_OF_RANGE;
    const walker_t *w = &e->w[k];
    *count = w->trace_count;
    if (sites) memcpy(sites, w->trace_sites, w->trace_count * sizeof(uint32_t));
    if (bins) memcpy(bins, w->trace_bins, w->trace_count * sizeof(uint32_t));
    if (flags) memcpy(flags, w->trace_flags, w->trace_count);
    return WLX_OK;
}
