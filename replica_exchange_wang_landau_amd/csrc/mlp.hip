// Warm-start network: a 4-layer, 128-wide MLP (6 features -> 128 -> 128 -> 128 -> 1, SiLU
// hidden activations) trained and evaluated by hand-written kernels. Matrix products take BF16
// operands on the matrix cores and accumulate in FP32: v_mfma_f32_16x16x32_bf16 (gfx950's
// 32-deep BF16 step) where the reduction runs over 128 features, v_mfma_f32_16x16x16_bf16
// where it runs over the 16 padded inputs or the 16 samples of a tile; biases,
// pre-activations, the output layer, the loss and Adam run in FP32 on FP32 master weights.
//
// One wave processes a tile of 16 samples through all layers with its activations in LDS.
// Training steps are deterministic: each tile writes its own gradient partial and one kernel
// sums the partials in tile order before the Adam update.
#include "engine.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace rewl {
namespace {

constexpr int kFeatures = 6;  // network inputs
constexpr int kIn = 16;       // inputs padded to one matrix-core K step
constexpr int kWidth = 128;   // hidden width
constexpr int kTile = 16;     // samples per wave

// parameter layout (FP32 master copy): W0 [128 x 16], b0 [128], W1 [128 x 128], b1 [128],
// W2 [128 x 128], b2 [128], w3 [128], b3 [1]; row j of a W holds the weights of output j
constexpr size_t kOffW0 = 0;
constexpr size_t kOffB0 = kOffW0 + size_t(kWidth) * kIn;
constexpr size_t kOffW1 = kOffB0 + kWidth;
constexpr size_t kOffB1 = kOffW1 + size_t(kWidth) * kWidth;
constexpr size_t kOffW2 = kOffB1 + kWidth;
constexpr size_t kOffB2 = kOffW2 + size_t(kWidth) * kWidth;
constexpr size_t kOffW3 = kOffB2 + kWidth;
constexpr size_t kOffB3 = kOffW3 + kWidth;
constexpr size_t kParams = kOffB3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8s __attribute__((ext_vector_type(8))); // BF16 bit patterns
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite)
__host__ __device__ inline uint16_t to_bf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__host__ __device__ inline float from_bf16(uint16_t h)
{
    const uint32_t u = uint32_t(h) << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

__device__ inline float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// v_mfma_f32_16x16x16_bf16 on gfx950: lane l holds A[l % 16][4 (l / 16) + t] and
// B[4 (l / 16) + t][l % 16] for t < 4, and D[4 (l / 16) + r][l % 16] in register r (measured
// with constant-row, constant-column and paired-k operands)
__device__ inline f32x4 mma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_bf16 on gfx950: lane l holds A[l % 16][8 (l / 16) + t] and
// B[8 (l / 16) + t][l % 16] for t < 8, and D[4 (l / 16) + r][l % 16] in register r (measured
// the same way)
__device__ inline f32x4 mma32(bf16x8s a, bf16x8s b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

struct Weights { // BF16 copies used by the matrix products
    const uint16_t *w0, *w1, *w2, *w3;
    const float *p; // FP32 master (biases, b3)
};

struct TileLds {
    uint16_t x[kTile * kIn];          // inputs
    uint16_t h[3][kTile * kWidth];    // activations of the hidden layers
    float z[3][kTile * kWidth];       // pre-activations
    uint16_t dz[kTile * kWidth];      // gradient of the current layer's pre-activation
    float dzf[kTile * kWidth];        // the same in FP32 (bias gradients)
    float y[kTile], err[kTile], dy[kTile];
};

// Z = act_in W^T + b for the tile, K inputs per sample; H = SiLU(Z) in BF16
template <int K>
__device__ void hidden_layer(const uint16_t *act_in, const uint16_t *w, const float *b, float *z,
                             uint16_t *h)
{
    const uint32_t l = threadIdx.x, i = l % 16, g = l / 16;
#pragma unroll 1
    for (int jt = 0; jt < kWidth / 16; ++jt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (K % 32 == 0) {
#pragma unroll
            for (int k0 = 0; k0 < K; k0 += 32) {
                bf16x8s a, bb;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    a[t] = short(act_in[i * K + k0 + 8 * g + t]);
                    bb[t] = short(w[(jt * 16 + i) * K + k0 + 8 * g + t]);
                }
                acc = mma32(a, bb, acc);
            }
        } else {
#pragma unroll
            for (int k0 = 0; k0 < K; k0 += 16) {
                bf16x4 a, bb;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    a[t] = short(act_in[i * K + k0 + 4 * g + t]);
                    bb[t] = short(w[(jt * 16 + i) * K + k0 + 4 * g + t]);
                }
                acc = mma(a, bb, acc);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t p = 4 * g + r, j = jt * 16 + i;
            const float v = acc[r] + b[j];
            z[p * kWidth + j] = v;
            h[p * kWidth + j] = to_bf16(v * sigmoid(v));
        }
    }
}

// the tile's inputs (zero padded) and forward pass; valid[p] = sample p exists
__device__ void forward_tile(const float *x, uint64_t count, uint64_t first, const uint32_t *index,
                             Weights wt, TileLds &s)
{
    const uint32_t l = threadIdx.x;
    for (uint32_t e = l; e < kTile * kIn; e += kWave) {
        const uint32_t p = e / kIn, k = e % kIn;
        const uint64_t n = first + p;
        float v = 0.0f;
        if (n < count && k < kFeatures) v = x[(index ? index[n] : n) * kFeatures + k];
        s.x[e] = to_bf16(v);
    }
    __syncthreads();
    hidden_layer<kIn>(s.x, wt.w0, wt.p + kOffB0, s.z[0], s.h[0]);
    __syncthreads();
    hidden_layer<kWidth>(s.h[0], wt.w1, wt.p + kOffB1, s.z[1], s.h[1]);
    __syncthreads();
    hidden_layer<kWidth>(s.h[1], wt.w2, wt.p + kOffB2, s.z[2], s.h[2]);
    __syncthreads();
    // output: y = H2 . w3 + b3, four lane groups of 32 features, summed in group order
    const uint32_t p = l % 16, g = l / 16;
    float part = 0.0f;
    for (int j = 32 * g; j < 32 * g + 32; ++j)
        part += from_bf16(s.h[2][p * kWidth + j]) * from_bf16(wt.w3[j]);
    const float p1 = __shfl(part, int(p + 16)), p2 = __shfl(part, int(p + 32)),
                p3 = __shfl(part, int(p + 48));
    if (g == 0) s.y[p] = ((part + p1) + p2) + p3 + wt.p[kOffB3];
    __syncthreads();
}

__global__ __launch_bounds__(kWave) void mlp_predict_kernel(const float *x, uint64_t count,
                                                            Weights wt, float *y)
{
    __shared__ TileLds s;
    const uint64_t first = uint64_t(blockIdx.x) * kTile;
    forward_tile(x, count, first, nullptr, wt, s);
    if (threadIdx.x < kTile && first + threadIdx.x < count) y[first + threadIdx.x] = s.y[threadIdx.x];
}

// dW[j][k] += sum_p dZ[p][j] act[p][k] over the tile (K = 16 samples), written to the tile's
// partial; act has `cols` columns
template <int COLS>
__device__ void weight_gradient(const uint16_t *dz, const uint16_t *act, float *out)
{
    const uint32_t l = threadIdx.x, i = l % 16, g = l / 16;
#pragma unroll 1
    for (int jt = 0; jt < kWidth / 16; ++jt)
#pragma unroll 1
        for (int kt = 0; kt < COLS / 16; ++kt) {
            bf16x4 a, b;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                a[t] = short(dz[(4 * g + t) * kWidth + jt * 16 + i]);  // A[j][p]
                b[t] = short(act[(4 * g + t) * COLS + kt * 16 + i]);   // B[p][k]
            }
            const f32x4 d = mma(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(jt * 16 + 4 * g + r) * COLS + kt * 16 + i] = d[r];
        }
}

// dH[p][k] = sum_j dZ[p][j] W[j][k] (K = 128), then dZ_prev = dH * SiLU'(Z_prev): FP32 into
// s.dzf while s.dz is still being read, then BF16 into s.dz
__device__ void back_through(const uint16_t *w, const float *z_prev, TileLds &s)
{
    const uint32_t l = threadIdx.x, i = l % 16, g = l / 16;
#pragma unroll 1
    for (int kt = 0; kt < kWidth / 16; ++kt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j0 = 0; j0 < kWidth; j0 += 32) {
            bf16x8s a, b;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                a[t] = short(s.dz[i * kWidth + j0 + 8 * g + t]);          // A[p][j]
                b[t] = short(w[(j0 + 8 * g + t) * kWidth + kt * 16 + i]); // B[j][k]
            }
            acc = mma32(a, b, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t p = 4 * g + r, k = kt * 16 + i;
            const float zv = z_prev[p * kWidth + k], sg = sigmoid(zv);
            s.dzf[p * kWidth + k] = acc[r] * (sg * (1.0f + zv * (1.0f - sg)));
        }
    }
    __syncthreads(); // every lane has read s.dz
    for (uint32_t e = l; e < kTile * kWidth; e += kWave) s.dz[e] = to_bf16(s.dzf[e]);
    __syncthreads();
}

// one tile's loss and gradient partial (mean squared error over `count` samples)
__global__ __launch_bounds__(kWave) void mlp_gradient_kernel(const float *x, const float *target,
                                                             uint64_t count, const uint32_t *index,
                                                             Weights wt, float *partial,
                                                             float *loss_partial)
{
    __shared__ TileLds s;
    const uint32_t l = threadIdx.x, tile = blockIdx.x;
    const uint64_t first = uint64_t(tile) * kTile;
    forward_tile(x, count, first, index, wt, s);
    float *out = partial + size_t(tile) * kParams;
    // output layer: dL/dy = 2 (y - t) / count for the tile's valid samples
    if (l < kTile) {
        const uint64_t n = first + l;
        const float err = n < count ? s.y[l] - target[index ? index[n] : n] : 0.0f;
        s.err[l] = err;
        s.dy[l] = 2.0f * err / float(count);
    }
    __syncthreads();
    if (l == 0) {
        float sum = 0.0f, dsum = 0.0f;
        for (int q = 0; q < kTile; ++q) {
            sum += s.err[q] * s.err[q];
            dsum += s.dy[q];
        }
        loss_partial[tile] = sum / float(count);
        out[kOffB3] = dsum;
    }
    for (int j = l; j < kWidth; j += kWave) { // w3 gradient
        float acc = 0.0f;
        for (int q = 0; q < kTile; ++q) acc += s.dy[q] * from_bf16(s.h[2][q * kWidth + j]);
        out[kOffW3 + j] = acc;
    }
    for (uint32_t e = l; e < kTile * kWidth; e += kWave) { // dZ2 = dy w3 SiLU'(Z2)
        const uint32_t q = e / kWidth, j = e % kWidth;
        const float zv = s.z[2][e], sg = sigmoid(zv);
        const float d = s.dy[q] * from_bf16(wt.w3[j]) * (sg * (1.0f + zv * (1.0f - sg)));
        s.dzf[e] = d;
        s.dz[e] = to_bf16(d);
    }
    __syncthreads();
    // layer 2: bias and weight gradients, then back to layer 1
    for (int j = l; j < kWidth; j += kWave) {
        float acc = 0.0f;
        for (int q = 0; q < kTile; ++q) acc += s.dzf[q * kWidth + j];
        out[kOffB2 + j] = acc;
    }
    weight_gradient<kWidth>(s.dz, s.h[1], out + kOffW2);
    back_through(wt.w2, s.z[1], s);
    for (int j = l; j < kWidth; j += kWave) {
        float acc = 0.0f;
        for (int q = 0; q < kTile; ++q) acc += s.dzf[q * kWidth + j];
        out[kOffB1 + j] = acc;
    }
    weight_gradient<kWidth>(s.dz, s.h[0], out + kOffW1);
    back_through(wt.w1, s.z[0], s);
    for (int j = l; j < kWidth; j += kWave) {
        float acc = 0.0f;
        for (int q = 0; q < kTile; ++q) acc += s.dzf[q * kWidth + j];
        out[kOffB0 + j] = acc;
    }
    weight_gradient<kIn>(s.dz, s.x, out + kOffW0);
}

// sums the tiles' partials in tile order; optionally an Adam step on the FP32 master weights
// and fresh BF16 copies
__global__ void mlp_reduce_kernel(const float *partial, uint32_t tiles, const float *loss_partial,
                                  float *grad, float *loss, float *params, float *m, float *v,
                                  uint16_t *bf, float lr, float c1, float c2, uint32_t adam)
{
    const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q == 0) {
        float s = 0.0f;
        for (uint32_t t = 0; t < tiles; ++t) s += loss_partial[t];
        *loss = s;
    }
    if (q >= kParams) return;
    float g = 0.0f;
    for (uint32_t t = 0; t < tiles; ++t) g += partial[size_t(t) * kParams + q];
    if (grad) grad[q] = g;
    if (!adam) return;
    const float mq = 0.9f * m[q] + 0.1f * g;
    const float vq = 0.999f * v[q] + 0.001f * g * g;
    m[q] = mq;
    v[q] = vq;
    const float w = params[q] - lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
    params[q] = w;
    bf[q] = to_bf16(w);
}

__global__ void mlp_copy_bf16_kernel(const float *params, uint16_t *bf)
{
    const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < kParams) bf[q] = to_bf16(params[q]);
}

} // namespace
} // namespace rewl

using namespace rewl;

struct wlx_mlp {
    float *d_params = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *d_bf = nullptr;
    uint64_t steps_taken = 0;
    hipStream_t stream = nullptr;

    Weights weights() const
    {
        return {d_bf + kOffW0, d_bf + kOffW1, d_bf + kOffW2, d_bf + kOffW3, d_params};
    }
};

namespace {

struct DeviceBuffer { // owned device memory of one call
    void *p = nullptr;
    ~DeviceBuffer()
    {
        if (p) (void)hipFree(p);
    }
};

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// uniform in [-1, 1) from 53 random bits
double symmetric_uniform(uint64_t &s) { return double(splitmix(s) >> 11) * 0x1.0p-52 - 1.0; }

int run_gradient(wlx_mlp *m, const float *d_x, const float *d_y, uint64_t count,
                 const uint32_t *d_index, float *d_partial, float *d_loss_partial)
{
    const uint32_t tiles = uint32_t((count + kTile - 1) / kTile);
    hipLaunchKernelGGL(mlp_gradient_kernel, dim3(tiles), dim3(kWave), 0, m->stream, d_x, d_y, count,
                       d_index, m->weights(), d_partial, d_loss_partial);
    REWL_HIP(hipGetLastError());
    return WLX_OK;
}

} // namespace

extern "C" {

int wlx_mlp_params(uint64_t *count)
{
    if (!count) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    *count = kParams;
    return WLX_OK;
}

int wlx_mlp_create(uint64_t seed, wlx_mlp **out)
{
    if (!out) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    *out = nullptr;
    auto m = std::make_unique<wlx_mlp>();
    // uniform initialisation with bound sqrt(6 / fan_in) (He, for SiLU); biases 0
    std::vector<float> p(kParams, 0.0f);
    uint64_t s = seed;
    auto fill = [&](size_t off, size_t rows, size_t cols, size_t used_cols) {
        const double bound = std::sqrt(6.0 / double(used_cols));
        for (size_t r = 0; r < rows; ++r)
            for (size_t c = 0; c < cols; ++c)
                p[off + r * cols + c] = c < used_cols ? float(bound * symmetric_uniform(s)) : 0.0f;
    };
    fill(kOffW0, kWidth, kIn, kFeatures);
    fill(kOffW1, kWidth, kWidth, kWidth);
    fill(kOffW2, kWidth, kWidth, kWidth);
    fill(kOffW3, 1, kWidth, kWidth);
    REWL_HIP(hipStreamCreate(&m->stream));
    REWL_HIP(hipMalloc(&m->d_params, kParams * sizeof(float)));
    REWL_HIP(hipMalloc(&m->d_m, kParams * sizeof(float)));
    REWL_HIP(hipMalloc(&m->d_v, kParams * sizeof(float)));
    REWL_HIP(hipMalloc(&m->d_bf, kParams * sizeof(uint16_t)));
    REWL_HIP(hipMemcpyAsync(m->d_params, p.data(), kParams * sizeof(float), hipMemcpyHostToDevice,
                            m->stream));
    std::vector<float> zero(kParams, 0.0f);
    REWL_HIP(hipMemcpyAsync(m->d_m, zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice,
                            m->stream));
    REWL_HIP(hipMemcpyAsync(m->d_v, zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice,
                            m->stream));
    hipLaunchKernelGGL(mlp_copy_bf16_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream,
                       m->d_params, m->d_bf);
    REWL_HIP(hipGetLastError());
    REWL_HIP(hipStreamSynchronize(m->stream));
    *out = m.release();
    return WLX_OK;
}

int wlx_mlp_destroy(wlx_mlp *m)
{
    if (!m) return WLX_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *p : {static_cast<void *>(m->d_params), static_cast<void *>(m->d_m),
                    static_cast<void *>(m->d_v), static_cast<void *>(m->d_bf)})
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return WLX_OK;
}

int wlx_mlp_get_params(wlx_mlp *m, float *params)
{
    if (!m || !params) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    REWL_HIP(hipMemcpyAsync(params, m->d_params, kParams * sizeof(float), hipMemcpyDeviceToHost,
                            m->stream));
    REWL_HIP(hipStreamSynchronize(m->stream));
    return WLX_OK;
}

int wlx_mlp_set_params(wlx_mlp *m, const float *params)
{
    if (!m || !params) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    for (size_t q = 0; q < kParams; ++q)
        if (!std::isfinite(params[q])) return set_error(WLX_INVALID_ARGUMENT, "Non-finite weight");
    REWL_HIP(hipMemcpyAsync(m->d_params, params, kParams * sizeof(float), hipMemcpyHostToDevice,
                            m->stream));
    hipLaunchKernelGGL(mlp_copy_bf16_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream,
                       m->d_params, m->d_bf);
    REWL_HIP(hipGetLastError());
    REWL_HIP(hipStreamSynchronize(m->stream));
    return WLX_OK;
}

int wlx_mlp_predict(wlx_mlp *m, const float *x, uint64_t count, float *y)
{
    if (!m || !x || !y) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (count == 0) return WLX_OK;
    if (count > (uint64_t(1) << 31)) return set_error(WLX_INVALID_ARGUMENT, "Too many samples");
    DeviceBuffer bx, by;
    REWL_HIP(hipMalloc(&bx.p, count * kFeatures * sizeof(float)));
    REWL_HIP(hipMalloc(&by.p, count * sizeof(float)));
    REWL_HIP(hipMemcpyAsync(bx.p, x, count * kFeatures * sizeof(float), hipMemcpyHostToDevice,
                            m->stream));
    const uint32_t tiles = uint32_t((count + kTile - 1) / kTile);
    hipLaunchKernelGGL(mlp_predict_kernel, dim3(tiles), dim3(kWave), 0, m->stream,
                       static_cast<const float *>(bx.p), count, m->weights(),
                       static_cast<float *>(by.p));
    REWL_HIP(hipGetLastError());
    REWL_HIP(hipMemcpyAsync(y, by.p, count * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    REWL_HIP(hipStreamSynchronize(m->stream));
    return WLX_OK;
}

int wlx_mlp_gradient(wlx_mlp *m, const float *x, const float *y, uint64_t count, float *loss,
                     float *grad)
{
    if (!m || !x || !y || !loss || !grad) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (count == 0 || count > (uint64_t(1) << 20))
        return set_error(WLX_INVALID_ARGUMENT, "Sample count out of range");
    const uint32_t tiles = uint32_t((count + kTile - 1) / kTile);
    DeviceBuffer bx, by, bp, bl, bg, bloss;
    REWL_HIP(hipMalloc(&bx.p, count * kFeatures * sizeof(float)));
    REWL_HIP(hipMalloc(&by.p, count * sizeof(float)));
    REWL_HIP(hipMalloc(&bp.p, size_t(tiles) * kParams * sizeof(float)));
    REWL_HIP(hipMalloc(&bl.p, tiles * sizeof(float)));
    REWL_HIP(hipMalloc(&bg.p, kParams * sizeof(float)));
    REWL_HIP(hipMalloc(&bloss.p, sizeof(float)));
    REWL_HIP(hipMemcpyAsync(bx.p, x, count * kFeatures * sizeof(float), hipMemcpyHostToDevice,
                            m->stream));
    REWL_HIP(hipMemcpyAsync(by.p, y, count * sizeof(float), hipMemcpyHostToDevice, m->stream));
    if (int st = run_gradient(m, static_cast<float *>(bx.p), static_cast<float *>(by.p), count,
                              nullptr, static_cast<float *>(bp.p), static_cast<float *>(bl.p)))
        return st;
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream,
                       static_cast<const float *>(bp.p), tiles, static_cast<const float *>(bl.p),
                       static_cast<float *>(bg.p), static_cast<float *>(bloss.p), nullptr,
                       nullptr, nullptr, nullptr, 0.0f, 1.0f, 1.0f, 0u);
    REWL_HIP(hipGetLastError());
    REWL_HIP(hipMemcpyAsync(grad, bg.p, kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    REWL_HIP(hipMemcpyAsync(loss, bloss.p, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    REWL_HIP(hipStreamSynchronize(m->stream));
    return WLX_OK;
}

int wlx_mlp_train(wlx_mlp *m, const float *x, const float *y, uint64_t count, uint64_t steps,
                  uint64_t batch, float learning_rate, uint64_t seed, float *losses)
{
    if (!m || !x || !y) return set_error(WLX_INVALID_ARGUMENT, "Null argument");
    if (count == 0 || count >= (uint64_t(1) << 32) || batch == 0 || batch > count ||
        batch > (uint64_t(1) << 16) || steps > (uint64_t(1) << 24) ||
        !(learning_rate > 0.0f) || !std::isfinite(learning_rate))
        return set_error(WLX_INVALID_ARGUMENT, "Invalid training arguments");
    if (steps == 0) return WLX_OK;
    // every step's minibatch: consecutive slices of per-epoch Fisher-Yates permutations
    std::vector<uint32_t> order(count), index(steps * batch);
    uint64_t s = seed;
    uint64_t at = count;
    for (uint64_t k = 0; k < steps * batch; ++k) {
        if (at == count) {
            for (uint64_t q = 0; q < count; ++q) order[q] = uint32_t(q);
            for (uint64_t q = count - 1; q > 0; --q) std::swap(order[q], order[splitmix(s) % (q + 1)]);
            at = 0;
        }
        index[k] = order[at++];
    }
    const uint32_t tiles = uint32_t((batch + kTile - 1) / kTile);
    DeviceBuffer bx, by, bi, bp, bl, bloss;
    REWL_HIP(hipMalloc(&bx.p, count * kFeatures * sizeof(float)));
    REWL_HIP(hipMalloc(&by.p, count * sizeof(float)));
    REWL_HIP(hipMalloc(&bi.p, index.size() * sizeof(uint32_t)));
    REWL_HIP(hipMalloc(&bp.p, size_t(tiles) * kParams * sizeof(float)));
    REWL_HIP(hipMalloc(&bl.p, tiles * sizeof(float)));
    REWL_HIP(hipMalloc(&bloss.p, steps * sizeof(float)));
    REWL_HIP(hipMemcpyAsync(bx.p, x, count * kFeatures * sizeof(float), hipMemcpyHostToDevice,
                            m->stream));
    REWL_HIP(hipMemcpyAsync(by.p, y, count * sizeof(float), hipMemcpyHostToDevice, m->stream));
    REWL_HIP(hipMemcpyAsync(bi.p, index.data(), index.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, m->stream));
    for (uint64_t k = 0; k < steps; ++k) {
        const uint32_t *idx = static_cast<const uint32_t *>(bi.p) + k * batch;
        if (int st = run_gradient(m, static_cast<float *>(bx.p), static_cast<float *>(by.p), batch,
                                  idx, static_cast<float *>(bp.p), static_cast<float *>(bl.p)))
            return st;
        ++m->steps_taken;
        const float c1 = 1.0f - std::pow(0.9f, float(m->steps_taken));
        const float c2 = 1.0f - std::pow(0.999f, float(m->steps_taken));
        hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream,
                           static_cast<const float *>(bp.p), tiles,
                           static_cast<const float *>(bl.p), nullptr,
                           static_cast<float *>(bloss.p) + k, m->d_params, m->d_m, m->d_v, m->d_bf,
                           learning_rate, c1, c2, 1u);
        REWL_HIP(hipGetLastError());
    }
    if (losses)
        REWL_HIP(hipMemcpyAsync(losses, bloss.p, steps * sizeof(float), hipMemcpyDeviceToHost,
                                m->stream));
    REWL_HIP(hipStreamSynchronize(m->stream));
    return WLX_OK;
}

} // extern "C"
