// Capacity.h -- learned per-query neighbor capacity (kernels).
//
// A small MLP maps local Morton-cell occupancy features of each primitive to the
// log of its neighbor count, so that neighbor rows can be sized before the single
// fill pass.  Features of the primitive at sorted position i (code c, point p):
//   for the two Morton levels mA, mA+1 (cells of side L/2^m, the coarser >= 2r):
//     log2(1 + own-cell count)      - log2(1 + expected per cell)
//     log2(1 + 27-cell neighborhood) - log2(1 + 27 x expected)
//   distance of p to the nearest face of its level-(mA+1) cell, per axis, scaled to [0, 1]
//   log2(expected neighbors per primitive)  (density x sphere volume)
// Network: 8 -> 32 (ReLU) -> 32 (ReLU) -> 1, FP32.  Parameters are one flat array:
//   W1[32][8] b1[32] W2[32][32] b2[32] W3[32] b3.
// Training minimizes the mean squared error on log2(1 + count) with Adam; gradients are
// reduced per wave with shuffles, per block in LDS, across blocks with float atomics.
#ifndef NEIGHBOR_KERNELS_CAPACITY_H_
#define NEIGHBOR_KERNELS_CAPACITY_H_

#include "../Runtime.h"
#include "LBVHBuild.h"

#include <stdint.h>

namespace neighbor
{
namespace gpu
{
namespace capacity
{
constexpr int kFeatures = 8;
constexpr int kHidden = 32;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kHidden * kFeatures;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kHidden;
constexpr int kParams = kB3 + 1;     // 1377

//! Inverse of spread3: gather every third bit (10 bits out).
__device__ __forceinline__ uint32_t compact3(uint32_t v)
    {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x030000ffu;
    v = (v | (v >> 16)) & 0x000003ffu;
    return v;
    }

//! Per-level feature inputs.
struct Levels
    {
    unsigned int mA;                 //!< coarse level (cells of side L/2^mA)
    const uint32_t* countsA;         //!< 8^mA cell counts
    const uint32_t* countsB;         //!< 8^(mA+1) cell counts
    float expectA;                   //!< expected primitives per coarse cell
    float expectB;
    float3 lo;                       //!< scene lower corner
    float3 cellB;                    //!< side of a fine cell per axis
    float log_expected;              //!< log2(expected neighbors per primitive)
    };

namespace kernel
{

//! Cell counts of sorted codes at one Morton level.  The codes are sorted, so the lanes of a
//! wave hold a few runs of equal cells: the first lane of each run adds the run's length
//! (one atomic per run instead of one per primitive on the same few counters).
__global__ void __launch_bounds__(256) cell_histogram(const uint32_t* __restrict__ codes, uint32_t n, unsigned int level,
                                                      uint32_t* __restrict__ counts)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int lane = threadIdx.x & 63u;
    const bool valid = i < n;
    const uint32_t cell = valid ? (codes[i] >> (30u - 3u * level)) : 0xffffffffu;
    const uint32_t prev = static_cast<uint32_t>(__shfl_up(static_cast<int>(cell), 1, 64));
    const bool head = valid && (lane == 0u || prev != cell);
    const unsigned long long heads = __builtin_amdgcn_ballot_w64(head);
    const unsigned long long validm = __builtin_amdgcn_ballot_w64(valid);
    if (!head) return;
    // run = [lane, next head or the end of the valid lanes)
    const unsigned long long above = (lane == 63u) ? 0ull : (heads & (~0ull << (lane + 1u)));
    const unsigned int end = above ? static_cast<unsigned int>(__builtin_ctzll(above))
                                   : static_cast<unsigned int>(64 - __builtin_clzll(validm));
    atomicAdd(&counts[cell], end - lane);
    }

__device__ __forceinline__ float neighborhood(const uint32_t* __restrict__ counts, uint32_t cell, unsigned int level, uint32_t& own)
    {
    const int side = 1 << level;
    const int cx = static_cast<int>(compact3(cell >> 2)), cy = static_cast<int>(compact3(cell >> 1)), cz = static_cast<int>(compact3(cell));
    own = counts[cell];
    uint32_t sum = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
                {
                const uint32_t x = static_cast<uint32_t>((cx + dx + side) & (side - 1));
                const uint32_t y = static_cast<uint32_t>((cy + dy + side) & (side - 1));
                const uint32_t z = static_cast<uint32_t>((cz + dz + side) & (side - 1));
                sum += counts[(gpu::kernel::spread3(x) << 2) | (gpu::kernel::spread3(y) << 1) | gpu::kernel::spread3(z)];
                }
    return static_cast<float>(sum);
    }

//! Features of the primitive at each sorted position.
__global__ void __launch_bounds__(256) features(const uint32_t* __restrict__ codes, const float4* __restrict__ points,
                                               uint32_t n, const Levels lv, float* __restrict__ out)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = codes[i];
    uint32_t ownA, ownB;
    const float nbA = neighborhood(lv.countsA, c >> (30u - 3u * lv.mA), lv.mA, ownA);
    const float nbB = neighborhood(lv.countsB, c >> (30u - 3u * (lv.mA + 1u)), lv.mA + 1u, ownB);
    const float4 p = points[i];
    float f[kFeatures];
    f[0] = log2f(1.0f + ownA) - log2f(1.0f + lv.expectA);
    f[1] = log2f(1.0f + nbA) - log2f(1.0f + 27.0f * lv.expectA);
    f[2] = log2f(1.0f + ownB) - log2f(1.0f + lv.expectB);
    f[3] = log2f(1.0f + nbB) - log2f(1.0f + 27.0f * lv.expectB);
    const float u[3] = {(p.x - lv.lo.x) / lv.cellB.x, (p.y - lv.lo.y) / lv.cellB.y, (p.z - lv.lo.z) / lv.cellB.z};
    for (int a = 0; a < 3; ++a)
        {
        const float frac = u[a] - floorf(u[a]);
        f[4 + a] = 2.0f * fminf(frac, 1.0f - frac);
        }
    f[7] = lv.log_expected;
    for (int k = 0; k < kFeatures; ++k) out[static_cast<size_t>(i) * kFeatures + k] = f[k];
    }

//! Forward pass of one sample; h1, h2 keep the post-ReLU activations.
__device__ __forceinline__ float mlp_forward(const float* __restrict__ w, const float* x, float* h1, float* h2)
    {
    for (int j = 0; j < kHidden; ++j)
        {
        float s = w[kB1 + j];
        for (int k = 0; k < kFeatures; ++k) s = fmaf(w[kW1 + j * kFeatures + k], x[k], s);
        h1[j] = fmaxf(s, 0.0f);
        }
    for (int j = 0; j < kHidden; ++j)
        {
        float s = w[kB2 + j];
        for (int k = 0; k < kHidden; ++k) s = fmaf(w[kW2 + j * kHidden + k], h1[k], s);
        h2[j] = fmaxf(s, 0.0f);
        }
    float y = w[kB3];
    for (int k = 0; k < kHidden; ++k) y = fmaf(w[kW3 + k], h2[k], y);
    return y;
    }

//! Prediction y = MLP(features) for n samples (sample i at features[i]).
__global__ void __launch_bounds__(256) predict(const float* __restrict__ params, const float* __restrict__ feat, uint32_t n,
                                               float* __restrict__ y)
    {
    __shared__ float w[kParams];
    for (int j = threadIdx.x; j < kParams; j += blockDim.x) w[j] = params[j];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[kFeatures], h1[kHidden], h2[kHidden];
    for (int k = 0; k < kFeatures; ++k) x[k] = feat[static_cast<size_t>(i) * kFeatures + k];
    y[i] = mlp_forward(w, x, h1, h2);
    }

__device__ __forceinline__ float wave_sum(float v)
    {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
    }

//! Accumulate the gradient of mean((y - t)^2) over a batch into grad (zeroed by the caller).
/*!
 * Sample b of the batch is sorted position sample[b] (b itself when sample is null); its
 * target is target[sample[b]] (log2(1 + count)).  Also accumulates the batch loss into
 * grad[kParams].
 */
__global__ void __launch_bounds__(256) grad(const float* __restrict__ params, const float* __restrict__ feat,
                                            const float* __restrict__ target, const uint32_t* __restrict__ sample,
                                            uint32_t batch, float* __restrict__ g)
    {
    __shared__ float w[kParams];
    __shared__ float acc[kParams + 1];
    for (int j = threadIdx.x; j < kParams; j += blockDim.x)
        {
        w[j] = params[j];
        acc[j] = 0.0f;
        }
    if (threadIdx.x == 0) acc[kParams] = 0.0f;
    __syncthreads();

    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = b < batch;
    const uint32_t i = valid ? (sample ? sample[b] : b) : 0u;
    float x[kFeatures], h1[kHidden], h2[kHidden];
    for (int k = 0; k < kFeatures; ++k) x[k] = valid ? feat[static_cast<size_t>(i) * kFeatures + k] : 0.0f;
    const float y = mlp_forward(w, x, h1, h2);
    const float r = valid ? (y - target[i]) : 0.0f;
    const float dy = valid ? 2.0f * r / static_cast<float>(batch) : 0.0f;
    const bool lead = (threadIdx.x & 63u) == 0u;

    float s = wave_sum(valid ? r * r / static_cast<float>(batch) : 0.0f);
    if (lead) atomicAdd(&acc[kParams], s);
    s = wave_sum(dy);
    if (lead) atomicAdd(&acc[kB3], s);
    float dh2[kHidden];
    for (int j = 0; j < kHidden; ++j)
        {
        s = wave_sum(dy * h2[j]);
        if (lead) atomicAdd(&acc[kW3 + j], s);
        dh2[j] = (h2[j] > 0.0f) ? dy * w[kW3 + j] : 0.0f;
        }
    float dh1[kHidden];
    for (int k = 0; k < kHidden; ++k) dh1[k] = 0.0f;
    for (int j = 0; j < kHidden; ++j)
        {
        s = wave_sum(dh2[j]);
        if (lead) atomicAdd(&acc[kB2 + j], s);
        for (int k = 0; k < kHidden; ++k)
            {
            s = wave_sum(dh2[j] * h1[k]);
            if (lead) atomicAdd(&acc[kW2 + j * kHidden + k], s);
            dh1[k] = fmaf(w[kW2 + j * kHidden + k], dh2[j], dh1[k]);
            }
        }
    for (int j = 0; j < kHidden; ++j)
        {
        const float d = (h1[j] > 0.0f) ? dh1[j] : 0.0f;
        s = wave_sum(d);
        if (lead) atomicAdd(&acc[kB1 + j], s);
        for (int k = 0; k < kFeatures; ++k)
            {
            s = wave_sum(d * x[k]);
            if (lead) atomicAdd(&acc[kW1 + j * kFeatures + k], s);
            }
        }
    __syncthreads();
    for (int j = threadIdx.x; j <= kParams; j += blockDim.x) atomicAdd(&g[j], acc[j]);
    }

//! One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at step t >= 1; clears grad for the next step.
__global__ void __launch_bounds__(256) adam(float* __restrict__ params, float* __restrict__ g, float* __restrict__ m,
                                            float* __restrict__ v, float lr, int t)
    {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > kParams) return;
    if (j == kParams)
        {
        g[j] = 0.0f;    // loss slot
        return;
        }
    const float gj = g[j];
    const float mj = 0.9f * m[j] + 0.1f * gj;
    const float vj = 0.999f * v[j] + 0.001f * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const float mh = mj / (1.0f - powf(0.9f, static_cast<float>(t)));
    const float vh = vj / (1.0f - powf(0.999f, static_cast<float>(t)));
    params[j] -= lr * mh / (sqrtf(vh) + 1e-8f);
    g[j] = 0.0f;
    }

//! Target log2(1 + count) in sorted order from per-query counts (query q = primitive q).
__global__ void __launch_bounds__(256) targets(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ primitive,
                                               uint32_t n, float* __restrict__ t)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t[i] = log2f(1.0f + static_cast<float>(counts[primitive[i]]));
    }

//! Capacity of query primitive[i]: ceil(2^(y + margin)) - 1 + pad, at least 1.
__global__ void __launch_bounds__(256) capacity(const float* __restrict__ params, const float* __restrict__ feat,
                                                const uint32_t* __restrict__ primitive, uint32_t n, float margin,
                                                uint32_t pad, uint32_t* __restrict__ cap)
    {
    __shared__ float w[kParams];
    for (int j = threadIdx.x; j < kParams; j += blockDim.x) w[j] = params[j];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[kFeatures], h1[kHidden], h2[kHidden];
    for (int k = 0; k < kFeatures; ++k) x[k] = feat[static_cast<size_t>(i) * kFeatures + k];
    const float y = mlp_forward(w, x, h1, h2);
    const float c = ceilf(exp2f(fminf(y + margin, 30.0f))) - 1.0f + static_cast<float>(pad);
    cap[primitive[i]] = c < 1.0f ? 1u : static_cast<uint32_t>(c);
    }

//! Queries whose count exceeds their capacity, appended in index order of discovery.
__global__ void __launch_bounds__(256) overflow_list(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ cap,
                                                     uint32_t n, uint32_t* __restrict__ list, uint32_t* __restrict__ n_list)
    {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (counts[q] > cap[q]) list[atomicAdd(n_list, 1u)] = q;
    }

//! count of overflowed query list[t], for the relocation scan.
__global__ void __launch_bounds__(256) gather_counts(const uint32_t* __restrict__ list, uint32_t n,
                                                     const uint32_t* __restrict__ counts, uint32_t* __restrict__ out)
    {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = counts[list[t]];
    }

//! Move overflowed rows to [tail + offset[t], ... + count): start and limit of query list[t].
__global__ void __launch_bounds__(256) relocate(const uint32_t* __restrict__ list, uint32_t n,
                                                const unsigned long long* __restrict__ offset,
                                                const uint32_t* __restrict__ counts, unsigned long long tail,
                                                unsigned long long* __restrict__ start, uint32_t* __restrict__ limit)
    {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t q = list[t];
    start[q] = tail + offset[t];
    limit[q] = counts[q];
    }

} // end namespace kernel
} // end namespace capacity
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_CAPACITY_H_
