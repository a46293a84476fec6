// RadixSort.h -- stable LSD radix sort of (u32 or u64 key, u32 value) pairs for gfx950.
//
// Replaces the reference's hipCUB DeviceRadixSort::SortPairs call
// (kernels/LBVH.cuh:421-443): keys are sorted ascending on bits [0, end_bit) and
// ties keep their input order, so with identity values the result is ordered by
// (key, original index).
//
// Design (one kernel per 8-bit digit pass, "onesweep" style):
//  * A tile is 512 threads x 12 keys = 6144 keys, held in registers in a
//    wave-striped arrangement (key w*768 + i*64 + lane of the tile lives in lane
//    `lane` of wave w, item i), so loads are coalesced and (wave, item, lane)
//    order is input order.
//  * Ranking inside a wave64 uses eight ballots per key to find the lanes holding
//    the same digit; the lowest such lane bumps a per-wave LDS counter.  No LDS
//    atomics, no sorting networks.
//  * Tiles obtain their position in the global order from a virtual tile id
//    (atomic counter) and a decoupled look-back over the per-digit counts of
//    earlier tiles.  A tile only ever waits on tiles with smaller ids, which were
//    scheduled before it, so the wait always ends; the spin is also bounded and
//    reports a flag instead of hanging.  (Eight counters, tile 8c + x for the c-th
//    claim of queue x, made each pass 5-9% slower: the look-backs wait longer.)
//  * The digit histograms are kept in eight copies, added to by the producer's
//    workgroups in turn and summed by the passes (one shared copy serialized the
//    producer's atomics: the codes kernel at 10^6 keys took 18.3 instead of 9.2 us).
//  * Keys and values are scattered into LDS in digit order and written out from
//    there, so each digit run of a tile is stored contiguously.
//  * The global digit histograms of all passes are produced in one read of the
//    keys (by the producer of the keys, see lbvh_codes in LBVHBuild.h, or by
//    radix_histogram below).
#ifndef NEIGHBOR_KERNELS_RADIX_SORT_H_
#define NEIGHBOR_KERNELS_RADIX_SORT_H_

#include "../Runtime.h"

#include <stdint.h>

#include <type_traits>

namespace neighbor
{
namespace gpu
{
namespace radix
{
constexpr unsigned int kBits = 8;
constexpr unsigned int kRadix = 1u << kBits;
constexpr unsigned int kBlock = 512;                 // threads per tile (the first kRadix own a digit each)
constexpr unsigned int kWaves = kBlock / 64;
constexpr unsigned int kItems = 12;                  // keys per thread (512 x 8: -3% at 64M; 256 x 16: 140 VGPRs)
constexpr unsigned int kTile = kBlock * kItems;      // 6144 keys
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;

static_assert(kBlock % kRadix == 0 && kBlock >= kRadix, "one thread per digit in the look-back");

//! Copies of the digit histograms: a producer's workgroup b adds its counts to copy b % 8 (one
//! per XCD under the round-robin dispatch), so each bin takes an eighth of the global atomics,
//! and a pass sums the copies of its digit.
constexpr unsigned int kHistParts = 8;

inline unsigned int num_passes(unsigned int end_bit) { return (end_bit + kBits - 1) / kBits; }
inline unsigned int num_tiles(size_t n) { return static_cast<unsigned int>((n + kTile - 1) / kTile); }

//! Words of device workspace for a sort of n keys over end_bit bits.
/*!
 * Layout: hist[kHistParts][passes][256] | tile_counter[passes] | error[1] | pad |
 * status[passes][tiles][256].
 * hist, tile_counter and error must be zero before the histogram is accumulated;
 * status must be zero before the passes.
 */
struct Workspace
    {
    uint32_t* hist;
    uint32_t* tile_counter;
    uint32_t* error;
    uint32_t* status;
    unsigned int passes;
    unsigned int tiles;

    static size_t header_words(unsigned int passes) { return ((passes * (kHistParts * kRadix + 1) + 1 + 63) / 64) * 64; }
    static size_t words(size_t n, unsigned int end_bit)
        {
        const unsigned int p = num_passes(end_bit);
        return header_words(p) + static_cast<size_t>(p) * num_tiles(n) * kRadix;
        }
    static Workspace carve(uint32_t* base, size_t n, unsigned int end_bit)
        {
        Workspace w;
        w.passes = num_passes(end_bit);
        w.tiles = num_tiles(n);
        w.hist = base;
        w.tile_counter = base + kHistParts * w.passes * kRadix;
        w.error = w.tile_counter + w.passes;
        w.status = base + header_words(w.passes);
        return w;
        }
    };

//! Add one to bin d of an LDS histogram.  When every active lane of the wave holds the same
//! digit (the upper digits of sorted or grouped keys), one lane adds the lane count instead
//! of 64 lanes colliding on one bin.
__device__ __forceinline__ void count_digit(uint32_t* s_hist, uint32_t d)
    {
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const unsigned long long same = __builtin_amdgcn_ballot_w64(d == d0);
    const unsigned long long active = __builtin_amdgcn_ballot_w64(true);
    if (same == active)
        {
        if (__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(active >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(active), 0u)) == 0u)
            atomicAdd(&s_hist[d0], static_cast<uint32_t>(__popcll(active)));
        }
    else
        {
        atomicAdd(&s_hist[d], 1u);
        }
    }

//! The copy of the histograms (passes x kRadix words each) that this workgroup adds to.
__device__ __forceinline__ uint32_t* hist_part(uint32_t* hist, const unsigned int passes)
    {
    return hist + (blockIdx.x % kHistParts) * passes * kRadix;
    }

//! No payload gather (the default of the LBVH hierarchy kernel's gather parameter).
struct NoGather
    {
    typedef int Value;
    __device__ __forceinline__ int load(uint32_t) const { return 0; }
    __device__ __forceinline__ void store(uint32_t, uint32_t, int) const {}
    };

namespace kernel
{

//! Zero n words (a graph-safe replacement for a memset node).
__global__ void __launch_bounds__(256) fill_zero(uint32_t* __restrict__ p, size_t n)
    {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
    }

//! Per-pass digit histograms of arbitrary keys (used when the key producer does not fuse it).
__global__ void __launch_bounds__(256) radix_histogram(const uint32_t* __restrict__ keys, uint32_t n,
                                                       unsigned int passes, uint32_t* __restrict__ hist)
    {
    __shared__ uint32_t s_hist[4 * kRadix];
    for (unsigned int j = threadIdx.x; j < passes * kRadix; j += blockDim.x) s_hist[j] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        {
        const uint32_t k = keys[i];
        for (unsigned int p = 0; p < passes; ++p)
            atomicAdd(&s_hist[p * kRadix + ((k >> (p * kBits)) & (kRadix - 1))], 1u);
        }
    __syncthreads();
    uint32_t* part = hist_part(hist, passes);
    for (unsigned int j = threadIdx.x; j < passes * kRadix; j += blockDim.x)
        if (s_hist[j]) atomicAdd(&part[j], s_hist[j]);
    }

//! Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp /* kWaves words */)
    {
    const unsigned int lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    #pragma unroll
    for (unsigned int d = 1; d < 64; d <<= 1)
        {
        const uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
        }
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    uint32_t wave_base = 0;
    for (unsigned int w = 0; w < wave; ++w) wave_base += s_tmp[w];
    __syncthreads();
    return wave_base + incl - v;
    }

//! One stable 8-bit digit pass.  IDENTITY_VALUES: values are the input positions (not read).
template<bool IDENTITY_VALUES, class KeyT>
__global__ void __launch_bounds__(kBlock) onesweep_pass(const KeyT* __restrict__ keys_in,
                                                        const uint32_t* __restrict__ vals_in,
                                                        KeyT* __restrict__ keys_out,
                                                        uint32_t* __restrict__ vals_out,
                                                        const uint32_t n,
                                                        const unsigned int shift,
                                                        const uint32_t* __restrict__ hist,
                                                        const unsigned int hist_stride,
                                                        uint32_t* __restrict__ tile_counter,
                                                        uint32_t* __restrict__ status,
                                                        uint32_t* __restrict__ error)
    {
    __shared__ KeyT s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t s_wave[kWaves][kRadix];
    __shared__ uint32_t s_tile_start[kRadix];
    __shared__ uint32_t s_dst[kRadix];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_tile;

    const unsigned int tid = threadIdx.x;
    const unsigned int lane = tid & 63u;
    const unsigned int wave = tid >> 6;

    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (unsigned int j = tid; j < kWaves * kRadix; j += kBlock) (&s_wave[0][0])[j] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kTile;

    // load (wave-striped); keys past the end carry digit 255 and sort last in the tile
    KeyT k[kItems];
    uint32_t v[kItems];
    #pragma unroll
    for (unsigned int i = 0; i < kItems; ++i)
        {
        const uint32_t idx = base + wave * (64 * kItems) + i * 64 + lane;
        if (idx < n)
            {
            k[i] = keys_in[idx];
            v[i] = IDENTITY_VALUES ? idx : vals_in[idx];
            }
        else
            {
            k[i] = ~KeyT(0);
            v[i] = 0u;
            }
        }

    // rank inside the wave, item by item, lanes in order
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t r[kItems];
    #pragma unroll
    for (unsigned int i = 0; i < kItems; ++i)
        {
        const uint32_t d = static_cast<uint32_t>(k[i] >> shift) & (kRadix - 1);
        uint64_t same = ~0ull;
        #pragma unroll
        for (unsigned int b = 0; b < kBits; ++b)
            {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            same &= bit ? bal : ~bal;
            }
        const uint32_t before = __popcll(same & lt_mask);
        const uint32_t prev = s_wave[wave][d];
        r[i] = prev + before;
        if (before == 0) s_wave[wave][d] = prev + __popcll(same);
        }
    __syncthreads();

    // per digit (thread == digit): exclusive offsets of the waves, tile total
    const bool owner = tid < kRadix;      // thread tid owns digit tid
    uint32_t tile_count = 0;
    if (owner)
        {
        const unsigned int d = tid;
        #pragma unroll
        for (unsigned int w = 0; w < kWaves; ++w)
            {
            const uint32_t c = s_wave[w][d];
            s_wave[w][d] = tile_count;
            tile_count += c;
            }
        }
    // this tile's digit counts go out before the block scans: later tiles look back at them
    // (the look-back's waits on unpublished counts dominate a pass at 10^6 keys)
    uint32_t* my = status + static_cast<size_t>(tile) * kRadix + (owner ? tid : 0u);
    if (owner)
        __hip_atomic_store(my, (tile == 0 ? kFlagPrefix : kFlagAggregate) | tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t hist_d = 0u;
    if (owner)
        {
        #pragma unroll
        for (unsigned int x = 0; x < kHistParts; ++x) hist_d += hist[x * hist_stride + tid];
        }
    const uint32_t tile_start = block_exclusive_scan(tile_count, s_tmp);
    if (owner) s_tile_start[tid] = tile_start;
    const uint32_t digit_base = block_exclusive_scan(hist_d, s_tmp);

    // decoupled look-back over earlier tiles, one digit per owner thread
    if (owner)
    {
        const unsigned int d = tid;
        uint32_t excl = 0;
        if (tile != 0)
            {
            int32_t j = static_cast<int32_t>(tile) - 1;
            while (j >= 0)
                {
                const uint32_t* w = status + static_cast<size_t>(j) * kRadix + d;
                uint32_t s = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                uint32_t spins = 0;
                while ((s >> 30) == 0u)
                    {
                    if (++spins > kSpinLimit)
                        {
                        atomicOr(error, 1u);
                        break;
                        }
                    __builtin_amdgcn_s_sleep(1);
                    s = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                excl += s & kValueMask;
                if ((s >> 30) != 1u) break;     // inclusive prefix found (or gave up)
                --j;
                }
            __hip_atomic_store(my, kFlagPrefix | ((excl + tile_count) & kValueMask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        s_dst[d] = digit_base + excl - tile_start;
    }
    __syncthreads();

    // scatter into LDS in digit order
    #pragma unroll
    for (unsigned int i = 0; i < kItems; ++i)
        {
        const uint32_t d = static_cast<uint32_t>(k[i] >> shift) & (kRadix - 1);
        const uint32_t pos = s_tile_start[d] + s_wave[wave][d] + r[i];
        s_keys[pos] = k[i];
        s_vals[pos] = v[i];
        }
    __syncthreads();

    // contiguous runs per digit to global memory
    const uint32_t nvalid = (n - base < kTile) ? (n - base) : kTile;
    for (uint32_t j = tid; j < nvalid; j += kBlock)
        {
        const KeyT key = s_keys[j];
        const uint32_t dst = s_dst[static_cast<uint32_t>(key >> shift) & (kRadix - 1)] + j;
        if (dst < n)       // always true unless a look-back gave up (error flag set)
            {
            keys_out[dst] = key;
            vals_out[dst] = s_vals[j];
            }
        }
    }

} // end namespace kernel

//! Run all passes.  Keys/values start in (keys, vals); the sorted result ends in
//! (keys, vals) when the number of passes is even, otherwise in (keys_alt, vals_alt).
//! Returns true when the result is in the alternate buffers.  The histogram must
//! already be in ws.hist and ws.status must be zero.
template<class KeyT>
inline bool sort_passes(KeyT* keys, uint32_t* vals, KeyT* keys_alt, uint32_t* vals_alt, uint32_t n,
                        bool identity_values, const Workspace& ws, hipStream_t stream)
    {
    bool alt = false;
    for (unsigned int p = 0; p < ws.passes; ++p)
        {
        const KeyT* ki = alt ? keys_alt : keys;
        const uint32_t* vi = alt ? vals_alt : vals;
        KeyT* ko = alt ? keys : keys_alt;
        uint32_t* vo = alt ? vals : vals_alt;
        uint32_t* st = ws.status + static_cast<size_t>(p) * ws.tiles * kRadix;
        if (p == 0 && identity_values)
            hipLaunchKernelGGL((kernel::onesweep_pass<true, KeyT>), dim3(ws.tiles), dim3(kBlock), 0, stream,
                               ki, vi, ko, vo, n, p * kBits, ws.hist + p * kRadix, ws.passes * kRadix, ws.tile_counter + p, st,
                               ws.error);
        else
            hipLaunchKernelGGL((kernel::onesweep_pass<false, KeyT>), dim3(ws.tiles), dim3(kBlock), 0, stream,
                               ki, vi, ko, vo, n, p * kBits, ws.hist + p * kRadix, ws.passes * kRadix, ws.tile_counter + p, st,
                               ws.error);
        alt = !alt;
        }
    check_launch("radix sort pass");
    return alt;
    }

} // end namespace radix
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_RADIX_SORT_H_
