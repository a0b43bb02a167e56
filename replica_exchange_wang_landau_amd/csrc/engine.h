// Internal structures shared by the coupling build (couplings.hip) and the walker engine
// (walkers.hip). Not part of the public ABI (include/rewl.h).
#pragma once

#include "../../include/rewl.h"
#include "device_common.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace rewl {

// ---- error plumbing --------------------------------------------------------------------
int set_error(int status, const std::string &message);
int device_error(hipError_t error, const char *what);
#define REWL_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t rewl_e_ = (call);                                                            \
        if (rewl_e_ != hipSuccess) return ::rewl::device_error(rewl_e_, #call);                 \
    } while (0)

// ---- wavefront geometry ------------------------------------------------------------------
constexpr int kWave = 64;
// LDS of one CU on gfx950 (160 KiB), the most one workgroup can allocate
constexpr size_t kLdsBytes = 160 * 1024;
// Sites per lane in the register-resident field layout: site = k*64 + lane, k < npl.
constexpr int kMaxSitesPerLane = 64; // N <= 4096 on the register path

inline uint32_t sites_per_lane(uint64_t n)
{
    uint32_t npl = 1;
    while (uint64_t(npl) * kWave < n) npl <<= 1;
    return npl;
}

} // namespace rewl

// Coupling system: J on the device, row stride `np` = npl*64 with zero padding, zero diagonal.
struct wlx_system {
    int device = 0;
    uint64_t n = 0;
    uint32_t npl = 1;
    uint64_t np = 64;
    bool csr = false;
    uint64_t nnz = 0;
    double *d_J64 = nullptr;      // dense n x np (CSR systems are expanded with exact zeros)
    uint64_t *d_offsets = nullptr; // CSR: n+1
    uint32_t *d_columns = nullptr; // CSR: nnz
    double *d_values = nullptr;    // CSR: nnz
    int precision = WLX_COUPLINGS_FP64;
    float *d_values32 = nullptr;   // CSR with FP32 couplings: the values as floats (nnz)
    float *d_J32 = nullptr;        // FP32 copy of d_J64 for FP32-field engines (made on demand)
    hipStream_t stream = nullptr;
};

namespace rewl {
// Runs local_fields + total_energy for `count` configurations (device); spins are np-strided.
int system_fields_device(wlx_system *s, const int8_t *d_spins, uint64_t count, double *d_fields,
                         double *d_energy, hipStream_t stream);
// Creates d_J32 (each coupling rounded to nearest FP32) if it does not exist yet.
int system_j32(wlx_system *s);
} // namespace rewl
