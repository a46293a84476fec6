// Scan.h -- exclusive scan of u32 counts into u64 offsets (CSR row starts), in index order
// or in the order of a permutation (rows laid out in traversal order).
//
// Three launches, no inter-block waiting: per-tile sums, a one-block scan of the
// tile sums, then per-tile scans offset by their tile's start.  Tile = 256 threads x
// 8 counts.  offsets has n+1 entries; offsets[n] is the total.
#ifndef NEIGHBOR_KERNELS_SCAN_H_
#define NEIGHBOR_KERNELS_SCAN_H_

#include "../Runtime.h"

#include <stdint.h>

namespace neighbor
{
namespace gpu
{
namespace scan
{
constexpr unsigned int kBlock = 256;
constexpr unsigned int kItems = 8;
constexpr unsigned int kTile = kBlock * kItems;

inline unsigned int num_tiles(uint32_t n) { return (n + kTile - 1) / kTile; }
inline size_t scratch_bytes(uint32_t n) { return (static_cast<size_t>(num_tiles(n)) + 1) * sizeof(unsigned long long); }

namespace kernel
{
__device__ __forceinline__ unsigned long long block_exclusive(unsigned long long v, unsigned long long* s_tmp, unsigned long long* total)
    {
    const unsigned int lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    unsigned long long incl = v;
    #pragma unroll
    for (unsigned int d = 1; d < 64; d <<= 1)
        {
        const unsigned long long y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
        }
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    unsigned long long base = 0, all = 0;
    for (unsigned int w = 0; w < blockDim.x / 64; ++w)
        {
        if (w < wave) base += s_tmp[w];
        all += s_tmp[w];
        }
    __syncthreads();
    if (total) *total = all;
    return base + incl - v;
    }

//! Count of scan position i: counts[i], or counts[perm[i]] for a scan in perm order.
__device__ __forceinline__ uint32_t count_at(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ perm, size_t i)
    {
    return counts[perm ? perm[i] : i];
    }

__global__ void __launch_bounds__(kBlock) tile_sums(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ perm,
                                                    uint32_t n, unsigned long long* __restrict__ sums)
    {
    __shared__ unsigned long long s_tmp[kBlock / 64];
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    unsigned long long s = 0;
    #pragma unroll
    for (unsigned int i = 0; i < kItems; ++i)
        if (base + i < n) s += count_at(counts, perm, base + i);
    unsigned long long total;
    block_exclusive(s, s_tmp, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
    }

__global__ void __launch_bounds__(1024) scan_sums(unsigned long long* __restrict__ sums, uint32_t m)
    {
    __shared__ unsigned long long s_tmp[16];
    unsigned long long carry = 0;
    for (uint32_t start = 0; start < m; start += blockDim.x)
        {
        const uint32_t i = start + threadIdx.x;
        const unsigned long long v = (i < m) ? sums[i] : 0ull;
        unsigned long long total;
        const unsigned long long ex = block_exclusive(v, s_tmp, &total);
        if (i < m) sums[i] = carry + ex;
        carry += total;
        }
    if (threadIdx.x == 0) sums[m] = carry;
    }

__global__ void __launch_bounds__(kBlock) tile_scan(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ perm,
                                                    uint32_t n, const unsigned long long* __restrict__ sums,
                                                    unsigned long long* __restrict__ offsets)
    {
    __shared__ unsigned long long s_tmp[kBlock / 64];
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    uint32_t c[kItems];
    unsigned long long s = 0;
    #pragma unroll
    for (unsigned int i = 0; i < kItems; ++i)
        {
        c[i] = (base + i < n) ? count_at(counts, perm, base + i) : 0u;
        s += c[i];
        }
    unsigned long long run = sums[blockIdx.x] + block_exclusive(s, s_tmp, nullptr);
    #pragma unroll
    for (unsigned int i = 0; i < kItems; ++i)
        {
        if (base + i < n) offsets[perm ? perm[base + i] : base + i] = run;
        run += c[i];
        }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) offsets[n] = sums[gridDim.x];
    }
//! Exclusive scan of n <= kGroupScanMax u64 values in one block, in rounds of 1024 coalesced
//! values (the loads of all rounds issued first), offset by *carry when carry is given (read
//! before any write, so it may be out[0]); out[n] = the offset total and, when spill is
//! given, spill[0] = that total and spill[1] = 0 (the cursor and moved-row count of a row
//! fill's spill area past all rows).
constexpr unsigned int kGroupScanRounds = 16;
constexpr unsigned int kGroupScanMax = 1024u * kGroupScanRounds;

__global__ void __launch_bounds__(1024) group_scan(const unsigned long long* __restrict__ in, uint32_t n,
                                                   unsigned long long* out, const unsigned long long* carry,
                                                   unsigned long long* __restrict__ spill)
    {
    __shared__ unsigned long long s_tmp[16];
    unsigned long long v[kGroupScanRounds];
    #pragma unroll
    for (unsigned int k = 0; k < kGroupScanRounds; ++k)
        {
        const uint32_t i = k * 1024u + threadIdx.x;
        v[k] = i < n ? in[i] : 0ull;
        }
    unsigned long long run = carry ? *carry : 0ull;
    #pragma unroll
    for (unsigned int k = 0; k < kGroupScanRounds; ++k)
        {
        if (k * 1024u >= n) break;                      // uniform
        unsigned long long total;
        const unsigned long long ex = block_exclusive(v[k], s_tmp, &total);   // barriers: carry read before writes
        const uint32_t i = k * 1024u + threadIdx.x;
        if (i < n) out[i] = run + ex;
        run += total;
        }
    if (threadIdx.x == 0)
        {
        out[n] = run;
        if (spill)
            {
            spill[0] = run;
            spill[1] = 0ull;
            }
        }
    }
} // end namespace kernel

//! out[0..n] = exclusive scan of n u64 values (out[n] the total); spill (optional) = {total, 0}.
//! One block per 16384 values, each carried from the previous block's total.
inline void group_totals_scan(const unsigned long long* in, uint32_t n, unsigned long long* out, unsigned long long* spill,
                              hipStream_t stream)
    {
    uint32_t done = 0;
    do
        {
        const uint32_t m = min(n - done, kernel::kGroupScanMax);
        const bool last = done + m == n;
        hipLaunchKernelGGL(kernel::group_scan, dim3(1), dim3(1024), 0, stream, in + done, m, out + done,
                           done ? static_cast<const unsigned long long*>(out + done) : nullptr, last ? spill : nullptr);
        done += m;
        }
    while (done < n);
    check_launch("group_totals_scan");
    }

//! offsets[0..n] = exclusive scan of counts; scratch holds scratch_bytes(n) bytes.  With a
//! permutation perm (n entries), the scan runs in perm order: offsets[perm[t]] = sum of
//! counts[perm[t']] over t' < t, and offsets[n] is the total.
inline void exclusive_scan(const uint32_t* counts, uint32_t n, unsigned long long* offsets, void* scratch, hipStream_t stream,
                           const uint32_t* perm = nullptr)
    {
    unsigned long long* sums = static_cast<unsigned long long*>(scratch);
    const unsigned int tiles = num_tiles(n);
    if (tiles == 0)
        {
        hipLaunchKernelGGL(kernel::scan_sums, dim3(1), dim3(1024), 0, stream, sums, 0u);
        NEIGHBOR_HIP_CHECK(hipMemcpyAsync(offsets, sums, sizeof(unsigned long long), hipMemcpyDeviceToDevice, stream));
        return;
        }
    hipLaunchKernelGGL(kernel::tile_sums, dim3(tiles), dim3(kBlock), 0, stream, counts, perm, n, sums);
    hipLaunchKernelGGL(kernel::scan_sums, dim3(1), dim3(1024), 0, stream, sums, tiles);
    hipLaunchKernelGGL(kernel::tile_scan, dim3(tiles), dim3(kBlock), 0, stream, counts, perm, n, sums, offsets);
    check_launch("exclusive_scan");
    }

} // end namespace scan
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_SCAN_H_
