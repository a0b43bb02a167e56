// Kernel (a): minimum-image dipolar coupling build, dense or CSR by cutoff, FP64.
//
// Replaces the OpenMP row loop of DenseCouplings (physics.cpp:172-184) and the serial O(N^2)
// scan + sort of CsrCouplings (physics.cpp:197-223). The pair value follows pair_for /
// minimum_image / dipolar_pair (physics.cpp:26-30, 151-170) operation for operation; the
// library is compiled with -ffp-contract=off so no product is fused, and FP64 division and
// square root are IEEE correctly rounded, which makes J bit-identical to the reference.
//
// Layout in HBM: J is n rows of np = npl*64 doubles (zero padded, zero diagonal), so a
// wavefront streams row J[i, :] with fully coalesced 512-byte accesses per site slot. A CSR
// system keeps offsets/columns/values and is also expanded into that dense layout with exact
// zeros; adding +-0.0 to a field never changes it, so the dense walker reproduces the CSR
// scatter (physics.cpp:231-236) bit for bit.
#include "engine.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>

namespace rewl {

// FP32 copy of the dense couplings (round to nearest), for FP32-field sampling
__global__ void narrow_couplings_kernel(const double *J64, float *J32, size_t count)
{
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < count;
         k += size_t(gridDim.x) * blockDim.x)
        J32[k] = float(J64[k]);
}

// FP32 couplings: each value rounded to the nearest binary32 in place (the FP64 array then holds
// exactly representable FP32 values), with an optional float copy
__global__ void round_couplings_kernel(double *values, float *copy, size_t count)
{
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < count;
         k += size_t(gridDim.x) * blockDim.x) {
        const float v = float(values[k]);
        values[k] = double(v);
        if (copy) copy[k] = v;
    }
}

int system_j32(wlx_system *s)
{
    if (s->d_J32) return WLX_OK;
    const size_t count = size_t(s->n) * s->np;
    REWL_HIP(hipMalloc(reinterpret_cast<void **>(&s->d_J32), count * sizeof(float)));
    hipLaunchKernelGGL(narrow_couplings_kernel, dim3(1024), dim3(256), 0, s->stream, s->d_J64,
                       s->d_J32, count);
    REWL_HIP(hipGetLastError());
    REWL_HIP(hipStreamSynchronize(s->stream));
    return WLX_OK;
}


namespace {
thread_local std::string g_error;
}

int set_error(int status, const std::string &message)
{
    g_error = message;
    return status;
}

int device_error(hipError_t error, const char *what)
{
    return set_error(WLX_DEVICE_ERROR,
                     std::string("HIP error ") + hipGetErrorString(error) + " in " + what);
}

// Site data in SoA form for the build kernels.
struct SiteArrays {
    const double *px, *py, *pz;
    const double *ax, *ay, *az;
};

struct BoxArgs {
    double lx, ly, lz;
    int periodic;
};

// minimum_image (physics.cpp:151-160): d -= L * nearbyint(d / L) per periodic axis with L > 0
__device__ inline double image(double d, double length, int periodic)
{
    if (periodic && length > 0.0) d -= length * rint(d / length);
    return d;
}

// displacement p_j - p_i for i < j after the minimum image (physics.cpp:26-28)
__device__ inline void displacement(const SiteArrays &s, const BoxArgs &b, uint32_t i, uint32_t j,
                                    double &dx, double &dy, double &dz)
{
    dx = image(s.px[j] - s.px[i], b.lx, b.periodic);
    dy = image(s.py[j] - s.py[i], b.ly, b.periodic);
    dz = image(s.pz[j] - s.pz[i], b.lz, b.periodic);
}

// dipolar_pair (physics.cpp:162-170), evaluated left to right without contraction;
// returns false for coincident sites
__device__ inline bool dipolar(const SiteArrays &s, uint32_t i, uint32_t j, double dx, double dy,
                               double dz, double scale, double &value)
{
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 > 0.0)) return false;
    const double inv_r = 1.0 / sqrt(r2);
    const double rx = dx * inv_r, ry = dy * inv_r, rz = dz * inv_r;
    const double ij = s.ax[i] * s.ax[j] + s.ay[i] * s.ay[j] + s.az[i] * s.az[j];
    const double ir = s.ax[i] * rx + s.ay[i] * ry + s.az[i] * rz;
    const double jr = s.ax[j] * rx + s.ay[j] * ry + s.az[j] * rz;
    const double angular = ij - 3.0 * ir * jr;
    value = scale * angular * inv_r * inv_r * inv_r;
    return true;
}

// One thread per (row i, column j); the row-major write is coalesced. The value of (i, j) and
// (j, i) comes from the same (min, max) evaluation, so J is bitwise symmetric.
__global__ void build_dense_kernel(SiteArrays s, BoxArgs b, uint32_t n, uint32_t np, double scale,
                                   double *J, int *coincident)
{
    const uint32_t i = blockIdx.y;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    double value = 0.0;
    if (j < n && j != i) {
        const uint32_t lo = min(i, j), hi = max(i, j);
        double dx, dy, dz;
        displacement(s, b, lo, hi, dx, dy, dz);
        if (!dipolar(s, lo, hi, dx, dy, dz, scale, value)) {
            atomicExch(coincident, 1);
            value = 0.0;
        }
    }
    J[size_t(i) * np + j] = value;
}

constexpr int kBuildBlock = 256;

// CSR pass 1: neighbours per row, norm(d) <= cutoff (physics.cpp:206)
__global__ __launch_bounds__(kBuildBlock) void csr_count_kernel(SiteArrays s, BoxArgs b,
                                                                uint32_t n, double cutoff,
                                                                uint64_t *counts)
{
    const uint32_t i = blockIdx.x;
    uint32_t count = 0;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        if (j == i) continue;
        const uint32_t lo = min(i, j), hi = max(i, j);
        double dx, dy, dz;
        displacement(s, b, lo, hi, dx, dy, dz);
        if (sqrt(dx * dx + dy * dy + dz * dz) <= cutoff) ++count;
    }
    __shared__ uint32_t partial[kBuildBlock];
    partial[threadIdx.x] = count;
    __syncthreads();
    for (int stride = kBuildBlock / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) partial[threadIdx.x] += partial[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[i] = partial[0];
}

// CSR pass 2: ordered compaction of each row (ascending column, as the (row, column) sort of
// physics.cpp:213-215 yields), values from the (min, max) orientation.
__global__ __launch_bounds__(kBuildBlock) void csr_fill_kernel(SiteArrays s, BoxArgs b, uint32_t n,
                                                               double cutoff, double scale,
                                                               const uint64_t *offsets,
                                                               uint32_t *columns, double *values,
                                                               int *coincident)
{
    const uint32_t i = blockIdx.x;
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    __shared__ uint32_t wave_counts[kBuildBlock / kWave];
    uint64_t base = offsets[i];
    for (uint32_t start = 0; start < n; start += kBuildBlock) {
        const uint32_t j = start + threadIdx.x;
        bool keep = false;
        double value = 0.0;
        if (j < n && j != i) {
            const uint32_t lo = min(i, j), hi = max(i, j);
            double dx, dy, dz;
            displacement(s, b, lo, hi, dx, dy, dz);
            if (sqrt(dx * dx + dy * dy + dz * dz) <= cutoff) {
                keep = true;
                if (!dipolar(s, lo, hi, dx, dy, dz, scale, value)) atomicExch(coincident, 1);
            }
        }
        const uint64_t ballot = __ballot(keep);
        const uint32_t rank = __popcll(ballot & ((uint64_t(1) << lane) - 1));
        if (lane == 0) wave_counts[wave] = __popcll(ballot);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < kBuildBlock / kWave; ++w) {
            if (w < wave) before += wave_counts[w];
            total += wave_counts[w];
        }
        if (keep) {
            columns[base + before + rank] = j;
            values[base + before + rank] = value;
        }
        base += total;
        __syncthreads();
    }
}

// expands CSR rows into the padded dense layout (exact zeros elsewhere)
__global__ void csr_expand_kernel(const uint64_t *offsets, const uint32_t *columns,
                                  const double *values, uint32_t np, double *J)
{
    const uint32_t i = blockIdx.x;
    for (uint64_t k = offsets[i] + threadIdx.x; k < offsets[i + 1]; k += blockDim.x)
        J[size_t(i) * np + columns[k]] = values[k];
}

// Geometry::validate (physics.cpp:34-56)
int validate_geometry(const double *pos, const double *axes, uint64_t n, const double box[3],
                      int periodic)
{
    if (n == 0) return set_error(WLX_INVALID_ARGUMENT, "Geometry contains no sites");
    if (periodic) {
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(box[k]) || !(box[k] >= 0.0))
                return set_error(WLX_INVALID_ARGUMENT,
                                 "Periodic box lengths must be finite and non-negative");
        if (box[0] == 0.0 && box[1] == 0.0 && box[2] == 0.0)
            return set_error(WLX_INVALID_ARGUMENT,
                             "At least one periodic box length must be positive");
    }
    for (uint64_t i = 0; i < n; ++i) {
        const double *p = pos + 3 * i, *a = axes + 3 * i;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]))
            return set_error(WLX_INVALID_ARGUMENT, "Site coordinates must be finite");
        const double length = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        if (std::abs(length - 1.0) > 1e-10)
            return set_error(WLX_INVALID_ARGUMENT, "Ising axes must be normalized");
    }
    return WLX_OK;
}

struct DeviceBuffer {
    void *p = nullptr;
    ~DeviceBuffer()
    {
        if (p) (void)hipFree(p);
    }
};

} // namespace rewl

using namespace rewl;

// ---- scratch of the no-allocation fields entry point -------------------------------------
// Staging for wlx_system_fields_noalloc: spins (np bytes), fields (np doubles) and the energy
// of each configuration. Each GPU holds its own copy until the process ends; its device
// address is taken once per device (hipGetSymbolAddress) and only that address reaches kernels
// and copies. Calls that overlap on one GPU share it and race.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_guard;
unsigned char *g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up once per device
int scratch_address(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(WLX_DEVICE_ERROR, "Device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_guard);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        REWL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch[device];
    return WLX_OK;
}

size_t round_up_256(size_t bytes) { return (bytes + 255) / 256 * 256; }
} // namespace

extern "C" {

int wlx_system_fields_noalloc(const wlx_system *s, const int8_t *spins, uint64_t count,
                              double *fields, double *energies)
{
    if (!s || !spins || !energies) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (count == 0) return WLX_OK;
    int device = 0;
    REWL_HIP(hipGetDevice(&device));
    if (device != s->device)
        return set_error(WLX_INVALID_ARGUMENT, "The system lives on another device");
    const size_t spin_bytes = round_up_256(count * s->np);
    const size_t field_bytes = round_up_256(count * s->np * sizeof(double));
    const size_t energy_bytes = count * sizeof(double);
    if (count > kNoAllocScratchBytes || spin_bytes + field_bytes + energy_bytes > kNoAllocScratchBytes)
        return set_error(WLX_INVALID_ARGUMENT,
                         "Configurations exceed the 4 MiB scratch of wlx_system_fields_noalloc "
                         "(count * padded spins * 9 bytes + 8 bytes per configuration)");
    for (uint64_t k = 0; k < count * s->n; ++k)
        if (spins[k] != 1 && spins[k] != -1)
            return set_error(WLX_INVALID_ARGUMENT, "Spins must be +/-1");
    unsigned char *base = nullptr;
    if (int status = scratch_address(device, &base)) return status;
    int8_t *d_spins = reinterpret_cast<int8_t *>(base);
    double *d_fields = reinterpret_cast<double *>(base + spin_bytes);
    double *d_energy = reinterpret_cast<double *>(base + spin_bytes + field_bytes);
    // spins into rows of np (the padding sites are never read: couplings stop at N)
    REWL_HIP(hipMemcpy2DAsync(d_spins, s->np, spins, s->n, s->n, count, hipMemcpyHostToDevice,
                              s->stream));
    auto *ms = const_cast<wlx_system *>(s);
    if (int status = system_fields_device(ms, d_spins, count, d_fields, d_energy, s->stream))
        return status;
    if (fields)
        REWL_HIP(hipMemcpy2DAsync(fields, s->n * sizeof(double), d_fields,
                                  s->np * sizeof(double), s->n * sizeof(double), count,
                                  hipMemcpyDeviceToHost, s->stream));
    REWL_HIP(hipMemcpyAsync(energies, d_energy, count * sizeof(double), hipMemcpyDeviceToHost,
                            s->stream));
    REWL_HIP(hipStreamSynchronize(s->stream));
    return WLX_OK;
}

size_t wlx_last_error(char *buffer, size_t capacity)
{
    if (buffer && capacity) {
        const size_t n = std::min(g_error.size(), capacity - 1);
        std::memcpy(buffer, g_error.data(), n);
        buffer[n] = 0;
    }
    return g_error.size();
}

int wlx_system_create(const double *positions, const double *axes, uint64_t n,
                      const double box[3], int periodic, double coupling_scale, double cutoff,
                      int precision, wlx_system **out)
{
    if (!out || (n && (!positions || !axes)) || !box)
        return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    *out = nullptr;
    if (precision != WLX_COUPLINGS_FP64 && precision != WLX_COUPLINGS_FP32)
        return set_error(WLX_INVALID_ARGUMENT, "Unknown coupling precision");
    if (int status = validate_geometry(positions, axes, n, box, periodic)) return status;
    const bool csr = cutoff > 0.0;
    if (n > uint64_t(kWave) * kMaxSitesPerLane)
        return set_error(WLX_INVALID_ARGUMENT,
                         "System too large for the register-resident walker (N <= 4096)");
    auto s = std::make_unique<wlx_system>();
    REWL_HIP(hipGetDevice(&s->device));
    s->n = n;
    s->npl = sites_per_lane(n);
    s->np = uint64_t(s->npl) * kWave;
    s->csr = csr;
    s->precision = precision;
    REWL_HIP(hipStreamCreate(&s->stream));

    // SoA site arrays
    std::vector<double> soa(6 * n);
    for (uint64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            soa[k * n + i] = positions[3 * i + k];
            soa[(3 + k) * n + i] = axes[3 * i + k];
        }
    DeviceBuffer d_sites, d_flag;
    REWL_HIP(hipMalloc(&d_sites.p, soa.size() * sizeof(double)));
    REWL_HIP(hipMalloc(&d_flag.p, sizeof(int)));
    REWL_HIP(hipMemcpyAsync(d_sites.p, soa.data(), soa.size() * sizeof(double),
                            hipMemcpyHostToDevice, s->stream));
    REWL_HIP(hipMemsetAsync(d_flag.p, 0, sizeof(int), s->stream));
    const double *base = static_cast<const double *>(d_sites.p);
    SiteArrays sites{base, base + n, base + 2 * n, base + 3 * n, base + 4 * n, base + 5 * n};
    BoxArgs b{box[0], box[1], box[2], periodic ? 1 : 0};
    int *coincident = static_cast<int *>(d_flag.p);

    REWL_HIP(hipMalloc(&s->d_J64, n * s->np * sizeof(double)));
    if (!csr) {
        dim3 grid(unsigned((s->np + kBuildBlock - 1) / kBuildBlock), unsigned(n));
        hipLaunchKernelGGL(build_dense_kernel, grid, dim3(kBuildBlock), 0, s->stream, sites, b,
                           uint32_t(n), uint32_t(s->np), coupling_scale, s->d_J64, coincident);
        REWL_HIP(hipGetLastError());
    } else {
        REWL_HIP(hipMalloc(&s->d_offsets, (n + 1) * sizeof(uint64_t)));
        hipLaunchKernelGGL(csr_count_kernel, dim3(unsigned(n)), dim3(kBuildBlock), 0, s->stream,
                           sites, b, uint32_t(n), cutoff, s->d_offsets + 1);
        REWL_HIP(hipGetLastError());
        std::vector<uint64_t> offsets(n + 1, 0);
        REWL_HIP(hipMemcpyAsync(offsets.data() + 1, s->d_offsets + 1, n * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, s->stream));
        REWL_HIP(hipStreamSynchronize(s->stream));
        for (uint64_t i = 1; i <= n; ++i) offsets[i] += offsets[i - 1];
        s->nnz = offsets[n];
        REWL_HIP(hipMemcpyAsync(s->d_offsets, offsets.data(), (n + 1) * sizeof(uint64_t),
                                hipMemcpyHostToDevice, s->stream));
        REWL_HIP(hipMalloc(&s->d_columns, std::max<uint64_t>(1, s->nnz) * sizeof(uint32_t)));
        REWL_HIP(hipMalloc(&s->d_values, std::max<uint64_t>(1, s->nnz) * sizeof(double)));
        hipLaunchKernelGGL(csr_fill_kernel, dim3(unsigned(n)), dim3(kBuildBlock), 0, s->stream,
                           sites, b, uint32_t(n), cutoff, coupling_scale, s->d_offsets,
                           s->d_columns, s->d_values, coincident);
        REWL_HIP(hipGetLastError());
        if (precision == WLX_COUPLINGS_FP32 && s->nnz) {
            REWL_HIP(hipMalloc(&s->d_values32, s->nnz * sizeof(float)));
            hipLaunchKernelGGL(round_couplings_kernel, dim3(1024), dim3(256), 0, s->stream,
                               s->d_values, s->d_values32, size_t(s->nnz));
            REWL_HIP(hipGetLastError());
        }
        REWL_HIP(hipMemsetAsync(s->d_J64, 0, n * s->np * sizeof(double), s->stream));
        hipLaunchKernelGGL(csr_expand_kernel, dim3(unsigned(n)), dim3(kBuildBlock), 0, s->stream,
                           s->d_offsets, s->d_columns, s->d_values, uint32_t(s->np), s->d_J64);
        REWL_HIP(hipGetLastError());
    }
    if (precision == WLX_COUPLINGS_FP32 && !csr) { // the CSR expansion copied rounded values
        hipLaunchKernelGGL(round_couplings_kernel, dim3(1024), dim3(256), 0, s->stream, s->d_J64,
                           static_cast<float *>(nullptr), size_t(n) * s->np);
        REWL_HIP(hipGetLastError());
    }
    int flag = 0;
    REWL_HIP(hipMemcpyAsync(&flag, coincident, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    REWL_HIP(hipStreamSynchronize(s->stream));
    if (flag) {
        wlx_system_destroy(s.release());
        return set_error(WLX_INVALID_ARGUMENT, "Coincident dipole sites");
    }
    *out = s.release();
    return WLX_OK;
}

int wlx_system_destroy(wlx_system *s)
{
    if (!s) return WLX_OK;
    (void)hipFree(s->d_J32);
    (void)hipFree(s->d_J64);
    (void)hipFree(s->d_offsets);
    (void)hipFree(s->d_columns);
    (void)hipFree(s->d_values);
    (void)hipFree(s->d_values32);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return WLX_OK;
}

int wlx_system_info(const wlx_system *s, uint64_t *n, uint64_t *backend, uint64_t *nnz)
{
    if (!s) return set_error(WLX_INVALID_ARGUMENT, "Null system");
    if (n) *n = s->n;
    if (backend) *backend = s->csr ? 1 : 0;
    if (nnz) *nnz = s->csr ? s->nnz : s->n * s->n;
    return WLX_OK;
}

int wlx_system_read_dense(const wlx_system *s, double *out)
{
    if (!s || !out) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    REWL_HIP(hipMemcpy2DAsync(out, s->n * sizeof(double), s->d_J64, s->np * sizeof(double),
                              s->n * sizeof(double), s->n, hipMemcpyDeviceToHost, s->stream));
    REWL_HIP(hipStreamSynchronize(s->stream));
    return WLX_OK;
}

int wlx_system_read_csr(const wlx_system *s, uint64_t *offsets, uint64_t *columns,
                        double *values)
{
    if (!s || !offsets || !columns || !values)
        return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (!s->csr) return set_error(WLX_INVALID_ARGUMENT, "System uses the dense backend");
    std::vector<uint32_t> cols(s->nnz);
    REWL_HIP(hipMemcpyAsync(offsets, s->d_offsets, (s->n + 1) * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s->stream));
    if (s->nnz) {
        REWL_HIP(hipMemcpyAsync(cols.data(), s->d_columns, s->nnz * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s->stream));
        REWL_HIP(hipMemcpyAsync(values, s->d_values, s->nnz * sizeof(double),
                                hipMemcpyDeviceToHost, s->stream));
    }
    REWL_HIP(hipStreamSynchronize(s->stream));
    for (uint64_t k = 0; k < s->nnz; ++k) columns[k] = cols[k];
    return WLX_OK;
}

int wlx_system_fields(const wlx_system *s, const int8_t *spins, uint64_t count, double *fields,
                      double *energies)
{
    if (!s || !spins || !energies) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (count == 0) return WLX_OK;
    for (uint64_t k = 0; k < count * s->n; ++k)
        if (spins[k] != 1 && spins[k] != -1)
            return set_error(WLX_INVALID_ARGUMENT, "Spins must be +/-1");
    std::vector<int8_t> padded(count * s->np, 1);
    for (uint64_t c = 0; c < count; ++c)
        std::memcpy(padded.data() + c * s->np, spins + c * s->n, s->n);
    DeviceBuffer d_spins, d_fields, d_energy;
    REWL_HIP(hipMalloc(&d_spins.p, padded.size()));
    REWL_HIP(hipMalloc(&d_fields.p, count * s->np * sizeof(double)));
    REWL_HIP(hipMalloc(&d_energy.p, count * sizeof(double)));
    REWL_HIP(hipMemcpyAsync(d_spins.p, padded.data(), padded.size(), hipMemcpyHostToDevice,
                            s->stream));
    auto *ms = const_cast<wlx_system *>(s);
    if (int status = system_fields_device(ms, static_cast<int8_t *>(d_spins.p), count,
                                          static_cast<double *>(d_fields.p),
                                          static_cast<double *>(d_energy.p), s->stream))
        return status;
    if (fields)
        REWL_HIP(hipMemcpy2DAsync(fields, s->n * sizeof(double), d_fields.p,
                                  s->np * sizeof(double), s->n * sizeof(double), count,
                                  hipMemcpyDeviceToHost, s->stream));
    REWL_HIP(hipMemcpyAsync(energies, d_energy.p, count * sizeof(double), hipMemcpyDeviceToHost,
                            s->stream));
    REWL_HIP(hipStreamSynchronize(s->stream));
    return WLX_OK;
}

} // extern "C"
