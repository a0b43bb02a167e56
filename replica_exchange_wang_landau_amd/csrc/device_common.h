// Device/host helpers shared by the coupling build and the walker kernels.
//
// Everything here must reproduce the reference's host arithmetic bit for bit: the whole
// library is compiled with -ffp-contract=off, divisions and square roots are the IEEE
// correctly rounded ones, and integer streams follow rng.hpp exactly.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace rewl {

// ---------------------------------------------------------------------------------------
// Xoshiro256** with splitmix64 seeding: rng.hpp:10-56 (Blackman & Vigna 2018).
// ---------------------------------------------------------------------------------------
struct Xoshiro {
    uint64_t s0, s1, s2, s3;

    __host__ __device__ static inline uint64_t rotl(uint64_t x, int k)
    {
        return (x << k) | (x >> (64 - k));
    }
    __host__ __device__ static inline uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    // rng.hpp:52-55: stream-separated seed
    __host__ __device__ static inline Xoshiro seeded(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (0xd2b74407b1ce6e93ULL * (stream + 1));
        Xoshiro r;
        r.s0 = splitmix(x);
        r.s1 = splitmix(x);
        r.s2 = splitmix(x);
        r.s3 = splitmix(x);
        return r;
    }
    __host__ __device__ inline uint64_t next()
    {
        const uint64_t result = rotl(s1 * 5, 7) * 9;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl(s3, 45);
        return result;
    }
    __host__ __device__ inline double uniform() { return double(next() >> 11) * 0x1.0p-53; }
};

// bounded(n) of rng.hpp:29-36 with the rejection threshold (-n) % n and the modulus
// precomputed on the host; `mask` != 0 marks a power-of-two n.
struct Bounded {
    uint64_t n;
    uint64_t threshold;
    uint64_t mask;
};

__host__ __device__ inline Bounded make_bounded(uint64_t n)
{
    Bounded b;
    b.n = n;
    b.threshold = n ? (0 - n) % n : 0;
    b.mask = (n && (n & (n - 1)) == 0) ? n - 1 : 0;
    return b;
}

__host__ __device__ inline uint64_t draw_bounded(Xoshiro &r, const Bounded &b)
{
    for (;;) {
        const uint64_t x = r.next();
        if (x >= b.threshold) return b.mask ? (x & b.mask) : (b.n == 1 ? 0 : x % b.n);
    }
}

// ---------------------------------------------------------------------------------------
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based stream of the
// WLX_RNG_PHILOX mode.
// ---------------------------------------------------------------------------------------
__host__ __device__ inline void mulhilo64(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo)
{
#if defined(__HIP_DEVICE_COMPILE__)
    lo = a * b;
    hi = __umul64hi(a, b);
#else
    const unsigned __int128 p = (unsigned __int128)a * b;
    hi = uint64_t(p >> 64);
    lo = uint64_t(p);
#endif
}

struct Philox4 {
    uint64_t v[4];
};

__host__ __device__ inline Philox4 philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                              uint64_t k0, uint64_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, hi0, lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, hi1, lo1);
        const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
    }
    return {{c0, c1, c2, c3}};
}

constexpr uint64_t kPhiloxWalkTag = 0x57414c4bULL;     // 'WALK'
constexpr uint64_t kPhiloxExchangeTag = 0x45584348ULL; // 'EXCH'
constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;    // exchange seed mixing, parallel.cpp:160

// ---------------------------------------------------------------------------------------
// EnergyGrid::index (wang_landau.cpp:27-34); -1 = nullopt.
// ---------------------------------------------------------------------------------------
struct Grid {
    double minimum, maximum, width;
    int64_t count;
    double inverse_width; // 1 / width, for the division-free fast path
    double margin;        // distance to an integer below which the fast path defers to division
    double margin_hi;     // 1 - margin
};

// The reciprocal product q = (e - minimum) * (1 / width) and the correctly rounded quotient
// Q = (e - minimum) / width both lie within 1.5 and 0.5 ulp(q) of the exact quotient, so their
// floors can differ only when an integer lies within 2 ulp(q) of q. `margin` is at least
// 4 ulp of the largest quotient (count), and never below 1e-9.
__host__ __device__ inline Grid make_grid(double minimum, double maximum, double width,
                                          int64_t count)
{
    double margin = 1e-9;
    const double ulp_bound = ldexp(4.0, ilogb(double(count) + 1.0) - 52);
    if (ulp_bound > margin) margin = ulp_bound;
    return {minimum, maximum, width, count, 1.0 / width, margin, 1.0 - margin};
}

// floor((e - minimum) / width) with the correctly rounded quotient: the reciprocal product's
// floor unless its fractional part lies within `margin` of 0 or 1 (then the division decides).
__host__ __device__ inline double bin_floor(const Grid &g, double e)
{
    const double x = e - g.minimum;
    const double q = x * g.inverse_width;
    const double f = floor(q);
    const double fr = q - f; // exact: f <= q < f + 1
    if (fr < g.margin || fr > g.margin_hi) return floor(x / g.width);
    return f;
}

__host__ __device__ inline int64_t grid_index(const Grid &g, double e)
{
    if (!(e >= g.minimum) || !(e <= g.maximum)) return -1; // also rejects NaN
    if (e - e != 0.0) return -1;                             // +-inf
    const double q = bin_floor(g, e);
    uint64_t raw = uint64_t(q);
    if (raw == uint64_t(g.count) && e == g.maximum) raw = uint64_t(g.count) - 1;
    if (raw >= uint64_t(g.count)) return -1;
    return int64_t(raw);
}

// ---------------------------------------------------------------------------------------
// (double)((long double)a / (long double)b) of histogram_statistics (wang_landau.cpp:361):
// the quotient is rounded to a 64-bit significand (x87 extended precision), then to a
// double, both round-to-nearest-even. Exact for a < 2^64, 0 < b < 2^62.
// ---------------------------------------------------------------------------------------
__host__ __device__ inline double extended_quotient(uint64_t a, uint64_t b)
{
    if (a == 0) return 0.0;
    uint64_t integer = a / b, r = a % b;
    // collect 65 significant quotient bits into (hi:1 bit, lo:64 bits) and a sticky bit
    int exponent; // weight of the leading bit is 2^exponent
    uint64_t v = 0; // top 64 bits of the 65-bit window
    int have = 0;   // significant bits collected
    uint64_t round_bit = 0;
    if (integer) {
        const int bits = 64 - __builtin_clzll(integer);
        exponent = bits - 1;
        v = integer << (64 - bits) >> 0;
        have = bits;
        if (bits == 64) v = integer;
    } else {
        exponent = 0;
        // skip leading zero fraction bits
        for (;;) {
            r <<= 1;
            --exponent;
            if (r >= b) {
                r -= b;
                break;
            }
        }
        v = uint64_t(1) << 63;
        have = 1;
    }
    // v holds `have` bits left-aligned; append fraction bits up to 64, then the round bit
    while (have < 64) {
        r <<= 1;
        const uint64_t bit = r >= b;
        if (bit) r -= b;
        v |= bit << (63 - have);
        ++have;
    }
    {
        r <<= 1;
        round_bit = r >= b;
        if (round_bit) r -= b;
    }
    const bool sticky = r != 0;
    // round to 64 bits, nearest-even
    if (round_bit && (sticky || (v & 1))) {
        ++v;
        if (v == 0) {
            v = uint64_t(1) << 63;
            ++exponent;
        }
    }
    // round the 64-bit significand to 53 bits, nearest-even (no sticky from the first step)
    const uint64_t low = v & 0x7FF;
    uint64_t m = v >> 11;
    if (low > 0x400 || (low == 0x400 && (m & 1))) {
        ++m;
        if (m == (uint64_t(1) << 53)) {
            m >>= 1;
            ++exponent;
        }
    }
    return ldexp(double(m), exponent - 52);
}

} // namespace rewl
