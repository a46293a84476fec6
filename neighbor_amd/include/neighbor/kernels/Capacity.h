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
    const uint32_t* nbA;             //!< 27-cell neighborhood sums of countsA (periodic)
    const uint32_t* nbB;             //!< 27-cell neighborhood sums of countsB
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

//! Sum of the counts of a cell's 27-cell periodic neighborhood (Morton-numbered cells).
__device__ __forceinline__ uint32_t neighborhood_sum(const uint32_t* __restrict__ counts, uint32_t cell, unsigned int level)
    {
    const int side = 1 << level;
    const int cx = static_cast<int>(compact3(cell >> 2)), cy = static_cast<int>(compact3(cell >> 1)), cz = static_cast<int>(compact3(cell));
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
    return sum;
    }

//! Neighborhood sums of every cell of both feature levels (cells of level A first, then B):
//! each primitive then reads two words per level instead of 27.
__global__ void __launch_bounds__(256) cell_neighborhoods(const uint32_t* __restrict__ counts, const unsigned int mA,
                                                          const uint32_t cellsA, const uint32_t cellsB,
                                                          uint32_t* __restrict__ nb)
    {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cellsA) nb[c] = neighborhood_sum(counts, c, mA);
    else if (c < cellsA + cellsB) nb[c] = neighborhood_sum(counts + cellsA, c - cellsA, mA + 1u);
    }

//! Features of the primitive at sorted position i.
__device__ __forceinline__ void feature_vector(const uint32_t* __restrict__ codes, const float4* __restrict__ points,
                                               const uint32_t i, const Levels& lv, float f[kFeatures])
    {
    const uint32_t c = codes[i];
    const uint32_t cellA = c >> (30u - 3u * lv.mA), cellB = c >> (30u - 3u * (lv.mA + 1u));
    const float ownA = static_cast<float>(lv.countsA[cellA]), ownB = static_cast<float>(lv.countsB[cellB]);
    const float nbA = static_cast<float>(lv.nbA[cellA]), nbB = static_cast<float>(lv.nbB[cellB]);
    const float4 p = points[i];
    f[0] = log2f(1.0f + ownA) - log2f(1.0f + lv.expectA);
    f[1] = log2f(1.0f + nbA) - log2f(1.0f + 27.0f * lv.expectA);
    f[2] = log2f(1.0f + ownB) - log2f(1.0f + lv.expectB);
    f[3] = log2f(1.0f + nbB) - log2f(1.0f + 27.0f * lv.expectB);
    const float u[3] = {(p.x - lv.lo.x) / lv.cellB.x, (p.y - lv.lo.y) / lv.cellB.y, (p.z - lv.lo.z) / lv.cellB.z};
    #pragma unroll
    for (int a = 0; a < 3; ++a)
        {
        const float frac = u[a] - floorf(u[a]);
        f[4 + a] = 2.0f * fminf(frac, 1.0f - frac);
        }
    f[7] = lv.log_expected;
    }

//! Features of the primitive at each sorted position.
__global__ void __launch_bounds__(256) features(const uint32_t* __restrict__ codes, const float4* __restrict__ points,
                                               uint32_t n, const Levels lv, float* __restrict__ out)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float f[kFeatures];
    feature_vector(codes, points, i, lv, f);
    for (int k = 0; k < kFeatures; ++k) out[static_cast<size_t>(i) * kFeatures + k] = f[k];
    }

//! FP32 forward pass of one sample (the training kernel's); h1, h2 keep the post-ReLU
//! activations.  Inference uses mlp_forward_mfma.
__device__ __forceinline__ float mlp_forward(const float* __restrict__ w, const float* x, float* h1, float* h2)
    {
    #pragma unroll
    for (int j = 0; j < kHidden; ++j)
        {
        float s = w[kB1 + j];
        #pragma unroll
        for (int k = 0; k < kFeatures; ++k) s = fmaf(w[kW1 + j * kFeatures + k], x[k], s);
        h1[j] = fmaxf(s, 0.0f);
        }
    #pragma unroll
    for (int j = 0; j < kHidden; ++j)
        {
        float s = w[kB2 + j];
        #pragma unroll
        for (int k = 0; k < kHidden; ++k) s = fmaf(w[kW2 + j * kHidden + k], h1[k], s);
        h2[j] = fmaxf(s, 0.0f);
        }
    float y = w[kB3];
    #pragma unroll
    for (int k = 0; k < kHidden; ++k) y = fmaf(w[kW3 + k], h2[k], y);
    return y;
    }

//! The forward pass of the 64 samples of a wave (one per lane) on the matrix cores.
/*!
 * Layers 1 and 2 are batched products over 32 samples: H1^T = W1 X^T (32 x 8 by 8 x 32, one
 * v_mfma_f32_32x32x8_f16) and H2^T = W2 H1^T (32 x 32 by 32 x 32, two v_mfma_f32_32x32x16_f16,
 * K = 16 each), FP16 operands and FP32 accumulation; the 64 lanes hold two such blocks
 * (samples 0-31 and 32-63).  Operand layouts of the 32 x 32 family: lane (h, j) = (lane / 32,
 * lane % 32) supplies row j of A and column j of B at K indices 4h + v (K = 8) or 8h + v
 * (K = 16), and receives D rows 8 (e / 4) + 4h + e % 4 (e = 0..15) of column j.
 * - Layer 1's B operand for block b at lane (h, j) is features 4h..4h+3 of sample 32b + j:
 *   a lane owns sample 32h + j, so it keeps its own half of the features for block h and
 *   swaps the other half with lane j + 32(1 - h) for block 1 - h.
 * - Layer 1's D rows are not the K order layer 2 reads from the same lane; instead of moving
 *   them across lanes, layer 2 uses W2 with its columns permuted to match: K index
 *   16c + 8h + v of chunk c is hidden unit 16c + 8(v / 4) + 4h + v % 4 (bits 2 and 3 of the
 *   index swapped).
 * - Layer 3 (32 weights) is a per-lane dot product over the lane's 16 D rows plus the partner
 *   lane's 16 (one shuffle).
 * The biases, w3 and b3 are applied in FP32.  All 64 lanes must execute it (no early exit).
 */
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

//! Hidden unit of layer 2's K index k (bits 2 and 3 swapped).
__host__ __device__ constexpr int mfma_k_unit(int k) { return (k & ~12) | ((k & 4) << 1) | ((k & 8) >> 1); }

//! Per-lane A operands of the two matrix layers (depend on the lane only: load once per wave).
struct MlpOperands
    {
    half4_t a1;      //!< W1[j][4h + v]
    half8_t a2[2];   //!< W2[j][unit(16c + 8h + v)]
    };

__device__ __forceinline__ MlpOperands mlp_operands(const float* __restrict__ w)
    {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int h = lane >> 5, j = lane & 31;
    MlpOperands o;
    #pragma unroll
    for (int v = 0; v < 4; ++v) o.a1[v] = static_cast<_Float16>(w[kW1 + j * kFeatures + 4 * h + v]);
    #pragma unroll
    for (int c = 0; c < 2; ++c)
        #pragma unroll
        for (int v = 0; v < 8; ++v) o.a2[c][v] = static_cast<_Float16>(w[kW2 + j * kHidden + mfma_k_unit(16 * c + 8 * h + v)]);
    return o;
    }

//! y of this lane's sample; s_vec: the block's copy of b1, b2, w3 (3 x 32 floats, LDS).
__device__ __forceinline__ float mlp_forward_mfma(const MlpOperands& o, const float* s_vec, const float b3, const float x[kFeatures])
    {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int h = lane >> 5;
    // features: own half for block h, the partner's half (swapped across the two half-waves) for block 1 - h
    half4_t own, send;
    #pragma unroll
    for (int v = 0; v < 4; ++v)
        {
        own[v] = static_cast<_Float16>(h ? x[4 + v] : x[v]);
        send[v] = static_cast<_Float16>(h ? x[v] : x[4 + v]);
        }
    const unsigned long long sent = __builtin_bit_cast(unsigned long long, send);
    const unsigned int lo = static_cast<unsigned int>(__shfl_xor(static_cast<int>(sent & 0xffffffffull), 32, 64));
    const unsigned int hi = static_cast<unsigned int>(__shfl_xor(static_cast<int>(sent >> 32), 32, 64));
    const half4_t recv = __builtin_bit_cast(half4_t, (static_cast<unsigned long long>(hi) << 32) | lo);
    float y = 0.0f;
    #pragma unroll
    for (int b = 0; b < 2; ++b)
        {
        const half4_t xb = (b == h) ? own : recv;
        floatx16 d1 = {};
        d1 = __builtin_amdgcn_mfma_f32_32x32x8f16(o.a1, xb, d1, 0, 0, 0);
        half8_t h1[2];
        #pragma unroll
        for (int e = 0; e < 16; ++e)
            {
            const int r = 8 * (e >> 2) + 4 * h + (e & 3);
            h1[e >> 3][e & 7] = static_cast<_Float16>(fmaxf(d1[e] + s_vec[r], 0.0f));
            }
        floatx16 d2 = {};
        d2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.a2[0], h1[0], d2, 0, 0, 0);
        d2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.a2[1], h1[1], d2, 0, 0, 0);
        float part = 0.0f;
        #pragma unroll
        for (int e = 0; e < 16; ++e)
            {
            const int r = 8 * (e >> 2) + 4 * h + (e & 3);
            part = fmaf(s_vec[2 * kHidden + r], fmaxf(d2[e] + s_vec[kHidden + r], 0.0f), part);
            }
        const float full = part + __shfl_xor(part, 32, 64) + b3;
        if (b == h) y = full;
        }
    return y;
    }

//! The block's LDS copy of b1, b2 and w3 for mlp_forward_mfma.
__device__ __forceinline__ void load_vectors(const float* __restrict__ w, float* s_vec)
    {
    for (int t = static_cast<int>(threadIdx.x); t < 3 * kHidden; t += static_cast<int>(blockDim.x))
        s_vec[t] = w[(t < kHidden ? kB1 : t < 2 * kHidden ? kB2 - kHidden : kW3 - 2 * kHidden) + t];
    __syncthreads();
    }

//! Capacity from a prediction y (log2(1 + count)): ceil(2^(y + margin)) - 1 + pad, at least 1.
__device__ __forceinline__ uint32_t capacity_of(const float y, const float margin, const uint32_t pad)
    {
    const float c = ceilf(exp2f(fminf(y + margin, 30.0f))) - 1.0f + static_cast<float>(pad);
    return c < 1.0f ? 1u : static_cast<uint32_t>(c);
    }

//! Features and capacity in one pass: cap[primitive[i]] for every sorted position i, with the
//! weights read straight from params (scalar loads), no feature array written or read.
//! group_total (optional): the capacities of sorted positions [256g, 256g + 256) (one block)
//! summed into group_total[g], for row starts placed per block of 256 queries (PlacedRowPairOp).
//! Launched with 256-thread blocks.
__global__ void __launch_bounds__(256) capacity_direct(const float* __restrict__ params, const uint32_t* __restrict__ codes,
                                                       const float4* __restrict__ points,
                                                       const uint32_t* __restrict__ primitive, const uint32_t n,
                                                       const Levels lv, const float margin, const uint32_t pad,
                                                       uint32_t* __restrict__ cap, unsigned long long* __restrict__ group_total)
    {
    __shared__ float s_vec[3 * kHidden];
    __shared__ unsigned long long s_wave[4];
    load_vectors(params, s_vec);
    const MlpOperands o = mlp_operands(params);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    float x[kFeatures] = {};
    if (valid) feature_vector(codes, points, i, lv, x);
    const float y = mlp_forward_mfma(o, s_vec, params[kB3], x);     // every lane of the wave takes part
    const uint32_t c = valid ? capacity_of(y, margin, pad) : 0u;
    if (valid) cap[primitive[i]] = c;
    if (group_total)
        {
        unsigned long long s = c;        // capacities of up to 2^30 + pad each: 64-bit sums
        #pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if ((threadIdx.x & 63u) == 0u) s_wave[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) group_total[blockIdx.x] = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        }
    }

//! Prediction y = MLP(features) for n samples (sample i at features[i]).
__global__ void __launch_bounds__(256) predict(const float* __restrict__ params, const float* __restrict__ feat, uint32_t n,
                                               float* __restrict__ y)
    {
    __shared__ float s_vec[3 * kHidden];
    load_vectors(params, s_vec);
    const MlpOperands o = mlp_operands(params);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    float x[kFeatures] = {};
    #pragma unroll
    for (int k = 0; k < kFeatures; ++k) x[k] = valid ? feat[static_cast<size_t>(i) * kFeatures + k] : 0.0f;
    const float yi = mlp_forward_mfma(o, s_vec, params[kB3], x);
    if (valid) y[i] = yi;
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

//! Capacity of query primitive[i] from the stored features.
__global__ void __launch_bounds__(256) capacity(const float* __restrict__ params, const float* __restrict__ feat,
                                                const uint32_t* __restrict__ primitive, uint32_t n, float margin,
                                                uint32_t pad, uint32_t* __restrict__ cap)
    {
    __shared__ float s_vec[3 * kHidden];
    load_vectors(params, s_vec);
    const MlpOperands o = mlp_operands(params);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    float x[kFeatures] = {};
    #pragma unroll
    for (int k = 0; k < kFeatures; ++k) x[k] = valid ? feat[static_cast<size_t>(i) * kFeatures + k] : 0.0f;
    const float y = mlp_forward_mfma(o, s_vec, params[kB3], x);
    if (valid) cap[primitive[i]] = capacity_of(y, margin, pad);
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
