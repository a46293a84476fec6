// Batch.h -- kernels of the batched (many independent boxes) build and traversal.
//
// All boxes are built as one forest in one pass of the single-tree pipeline:
//   batch_codes      key = box << 30 | Morton code of the point in its own box (the
//                    single-tree code formula against that box's lo/hi), fused with the
//                    digit histograms of a 64-bit onesweep sort;
//   (sort)           RadixSort.h on the 64-bit keys (stable: ties keep input order);
//   batch_hierarchy  Karras over BoxKeys (LBVHBuild.h): each box is one subtree, shaped
//                    like the tree of that box built alone;
//   (bubble)         the single-tree bubble kernels, unchanged;
//   batch_roots      the subtree root of every box;
//   batch_items      work items of up to 64 consecutive (Morton-ordered) queries of one box.
// Traversal: persistent kernels claim items from eight queues, one per XCD, each a
// contiguous range of items (so one XCD's L2 holds a few boxes at a time) and steal from
// the other queues when their own is empty.  Queries are the points themselves; the walk
// of an item covers its box's subtree only, with the box's own periodic images.
#ifndef NEIGHBOR_KERNELS_BATCH_H_
#define NEIGHBOR_KERNELS_BATCH_H_

#include "LBVHBuild.h"
#include "LBVHTraverse.h"

#include <stdint.h>
#include <type_traits>

namespace neighbor
{
namespace gpu
{
namespace batch
{
namespace walk_kernels = ::neighbor::gpu::kernel;

constexpr unsigned int kCodeBits = 30;
constexpr unsigned int kItemQueries = 64;      //!< queries per work item (one wave)
constexpr unsigned int kQueues = walk_kernels::wave::kXcds;  //!< item queues, one per XCD

//! Box of global point i: the last b with offsets[b] <= i (offsets has nbox + 1 entries).
__device__ __forceinline__ unsigned int box_of(const uint32_t* __restrict__ offsets, unsigned int nbox, uint32_t i)
    {
    unsigned int lo = 0, hi = nbox;     // offsets[lo] <= i < offsets[hi]
    while (hi - lo > 1u)
        {
        const unsigned int mid = (lo + hi) >> 1;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid;
        }
    return lo;
    }

//! The 27 (or 1) image vectors of a periodic box of side L, ix outer, iz inner (self at 13).
struct BoxImages
    {
    typedef double3 type;
    double3 L;
    unsigned int n;      //!< 27 periodic, 1 not

    __device__ __forceinline__ double3 get(const unsigned int k) const
        {
        if (n == 1u) return make_double3(0.0, 0.0, 0.0);
        const int ix = static_cast<int>(k / 9u) - 1;
        const int iy = static_cast<int>((k / 3u) % 3u) - 1;
        const int iz = static_cast<int>(k % 3u) - 1;
        return make_double3(L.x * ix, L.y * iy, L.z * iz);
        }
    NEIGHBOR_HD unsigned int size() const { return n; }
    };

//! A work item: box and first query (Morton rank within the box).
struct Item
    {
    uint32_t box;
    uint32_t first;
    };

namespace kernel
{

//! Forest keys (box << 30 | code) + digit histograms of all passes (+ clears of scratch words).
__global__ void __launch_bounds__(256) batch_codes(const float3* __restrict__ points, const uint32_t n,
                                                   const uint32_t* __restrict__ offsets, const unsigned int nbox,
                                                   const float3* __restrict__ box_lo, const float3* __restrict__ box_hi,
                                                   unsigned long long* __restrict__ keys, uint32_t* __restrict__ hist,
                                                   const unsigned int passes, uint32_t* __restrict__ clear,
                                                   const size_t clear_words)
    {
    __shared__ uint32_t s_hist[8 * radix::kRadix];
    for (unsigned int j = threadIdx.x; j < passes * radix::kRadix; j += blockDim.x) s_hist[j] = 0u;
    __syncthreads();
    // each block takes one contiguous range of points, so a thread finds its box once and
    // then only steps forward
    const size_t per = ((static_cast<size_t>(n) + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
    const size_t begin = static_cast<size_t>(blockIdx.x) * per;
    const size_t end = min(begin + per, static_cast<size_t>(n));
    unsigned int b = begin + threadIdx.x < end ? box_of(offsets, nbox, static_cast<uint32_t>(begin + threadIdx.x)) : 0u;
    for (size_t i = begin + threadIdx.x; i < end; i += blockDim.x)
        {
        while (offsets[b + 1] <= i) ++b;
        const float3 lo = box_lo[b], hi = box_hi[b];
        const float3 L = make_float3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
        const float3 c = points[i];
        const uint32_t code = gpu::kernel::morton3(gpu::kernel::fraction_to_bin((c.x - lo.x) / L.x),
                                                   gpu::kernel::fraction_to_bin((c.y - lo.y) / L.y),
                                                   gpu::kernel::fraction_to_bin((c.z - lo.z) / L.z));
        const unsigned long long key = (static_cast<unsigned long long>(b) << kCodeBits) | code;
        keys[i] = key;
        for (unsigned int p = 0; p < passes; ++p)
            radix::count_digit(s_hist + p * radix::kRadix, static_cast<uint32_t>((key >> (p * radix::kBits)) & (radix::kRadix - 1)));
        }
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < clear_words; i += stride) clear[i] = 0u;
    __syncthreads();
    uint32_t* part = radix::hist_part(hist, passes);
    for (unsigned int j = threadIdx.x; j < passes * radix::kRadix; j += blockDim.x)
        if (s_hist[j]) atomicAdd(&part[j], s_hist[j]);
    }

//! Row capacity of every point from its box's density, for the single-pass pair fill (rows that
//! overflow are moved and walked again inside the fill, so any capacity gives exact rows):
//! cap = ceil(lambda + z * sqrt(lambda)) + pad with lambda = (N_b / V_b) (4/3) pi r^3 the mean
//! count of a uniform box at the point's query radius r (sphere w), at least 1; box b holds
//! points [offsets[b], offsets[b+1]).  Uniform boxes have nearly Poisson counts, so z sets the
//! share of moved rows.
__global__ void __launch_bounds__(256) batch_capacity(const float4* __restrict__ spheres, const uint32_t* __restrict__ offsets,
                                                      const unsigned int nbox, const float3* __restrict__ box_lo,
                                                      const float3* __restrict__ box_hi, const float z, const uint32_t pad,
                                                      uint32_t* __restrict__ cap)
    {
    // one block per box (grid-stride over boxes): the box's density once, then its points
    for (unsigned int b = blockIdx.x; b < nbox; b += gridDim.x)
        {
        const uint32_t s = offsets[b], e = offsets[b + 1];
        const float3 lo = box_lo[b], hi = box_hi[b];
        const double volume = static_cast<double>(hi.x - lo.x) * static_cast<double>(hi.y - lo.y) * static_cast<double>(hi.z - lo.z);
        const double density = static_cast<double>(e - s) / volume;
        for (uint32_t i = s + threadIdx.x; i < e; i += blockDim.x)
            {
            const double r = static_cast<double>(spheres[i].w);
            const double lambda = density * (4.0 / 3.0) * 3.14159265358979 * r * r * r;
            const double c = ceil(lambda + static_cast<double>(z) * sqrt(lambda)) + static_cast<double>(pad);
            cap[i] = c < 1.0 ? 1u : (c > 1073741824.0 ? 1073741824u : static_cast<uint32_t>(c));
            }
        }
    }

//! Karras hierarchy of the forest.
__global__ void __launch_bounds__(1024) batch_hierarchy(const LBVHData tree, const unsigned long long* __restrict__ keys,
                                                        const uint32_t* __restrict__ offsets, const uint32_t n,
                                                        const gpu::kernel::SpanLists sl)
    {
    gpu::kernel::hierarchy_node(tree, gpu::kernel::BoxKeys{keys, offsets, n},
                                static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), sl);
    }

//! Roots of the boxes that have no internal node: the leaf of a one-point box, the
//! sentinel otherwise (overwritten by batch_roots for boxes of two or more points).
__global__ void __launch_bounds__(256) batch_roots_init(const uint32_t* __restrict__ offsets, const uint32_t n,
                                                        const unsigned int nbox, int* __restrict__ roots)
    {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const uint32_t s = offsets[b], e = offsets[b + 1];
    roots[b] = (e - s == 1u) ? static_cast<int>(n - 1u + s) : LBVHSentinel;
    }

//! Subtree root of each box of two or more points: the internal node whose leaf range is
//! exactly the box (thread i checks internal node i).
__global__ void __launch_bounds__(256) batch_roots(const int* __restrict__ range_end, const unsigned long long* __restrict__ keys,
                                                   const uint32_t* __restrict__ offsets, const uint32_t n,
                                                   int* __restrict__ roots)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1u >= n) return;
    const int j = range_end[i];
    const uint32_t first = min(i, static_cast<uint32_t>(j)), last = max(i, static_cast<uint32_t>(j));
    const uint32_t b = static_cast<uint32_t>(keys[first] >> kCodeBits);
    if (static_cast<uint32_t>(keys[last] >> kCodeBits) != b) return;
    if (first == offsets[b] && last + 1u == offsets[b + 1]) roots[b] = static_cast<int>(i);
    }

//! Items per box (for the scan that places them).
__global__ void __launch_bounds__(256) batch_item_counts(const uint32_t* __restrict__ offsets, const unsigned int nbox,
                                                         uint32_t* __restrict__ counts)
    {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbox) counts[b] = (offsets[b + 1] - offsets[b] + kItemQueries - 1) / kItemQueries;
    }

__global__ void __launch_bounds__(256) batch_items(const uint32_t* __restrict__ offsets, const unsigned int nbox,
                                                   const unsigned long long* __restrict__ item_offsets, Item* __restrict__ items)
    {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const uint32_t n = offsets[b + 1] - offsets[b];
    unsigned long long t = item_offsets[b];
    for (uint32_t f = 0; f < n; f += kItemQueries, ++t) items[t] = Item{b, f};
    }

//! Claim the next item for this wave from queue `q` (wave-uniform result).
__device__ __forceinline__ unsigned long long claim(unsigned int* __restrict__ counters, unsigned int q)
    {
    unsigned int v = 0;
    if ((threadIdx.x & 63u) == 0u) v = atomicAdd(&counters[q], 1u);
    return __builtin_amdgcn_readfirstlane(v);
    }

//! Items [begin, end) of queue q out of n items.
__device__ __forceinline__ void queue_range(unsigned long long n, unsigned int q, unsigned long long& begin,
                                            unsigned long long& end)
    {
    begin = n * q / kQueues;
    end = n * (q + 1u) / kQueues;
    }

//! Persistent traversal of the forest.  LANE_WALK: each lane walks its own rope path
//! (bucket records, leaves tested one by one); otherwise wave::walk.
template<bool LANE_WALK, class OutputOpT, class QueryOpT>
__device__ __forceinline__ void batch_traverse_body(const OutputOpT& out,
                                                    const float4* __restrict__ nodes,
                                                    const float4* __restrict__ points,
                                                    const unsigned int* __restrict__ primitive,
                                                    const float3* __restrict__ tree_lo,
                                                    const float3* __restrict__ tree_hi,
                                                    const int* __restrict__ rope,
                                                    const int* __restrict__ roots,
                                                    const uint32_t* __restrict__ offsets,
                                                    const float3* __restrict__ box_lo,
                                                    const float3* __restrict__ box_hi,
                                                    const Item* __restrict__ items,
                                                    const unsigned long long* __restrict__ n_items_ptr,
                                                    unsigned int* __restrict__ counters,
                                                    const QueryOpT& query,
                                                    const unsigned int periodic,
                                                    const unsigned int n_nodes)
    {
    // the lane walk writes its hits directly: no staging storage
    __shared__ walk_kernels::StageStorage<typename std::conditional<LANE_WALK, int, OutputOpT>::type, QueryOpT> lds;
    walk_kernels::wave::Stage stage = lds.lane_stage();
    const unsigned int lane = threadIdx.x & 63u;
    const unsigned int n_internal = (n_nodes - 1u) / 2u;
    const unsigned long long n_items = *n_items_ptr;
    const unsigned int home = blockIdx.x % kQueues;

    for (unsigned int qi = 0; qi < kQueues; ++qi)
        {
        const unsigned int q = (home + qi) % kQueues;
        unsigned long long begin, end;
        queue_range(n_items, q, begin, end);
        for (;;)
            {
            const unsigned long long t = begin + claim(counters, q);
            if (t >= end) break;
            const Item it = items[t];
            const uint32_t s = offsets[it.box];
            const uint32_t nb = offsets[it.box + 1] - s;
            const bool valid = it.first + lane < nb;
            const unsigned int idx = valid ? primitive[s + it.first + lane] : 0u;
            // a box without a root (none is expected for a box with points) yields no hits
            const int root = roots[it.box];
            const bool rooted = root != LBVHSentinel;
            const int stop = rooted ? rope[root] : root;
            const BoundingBox root_box = rooted ? BoundingBox(tree_lo[root], tree_hi[root])
                                                : BoundingBox(make_float3(3e38f, 3e38f, 3e38f), make_float3(3e38f, 3e38f, 3e38f));
            const float3 blo = box_lo[it.box], bhi = box_hi[it.box];
            const BoxImages images{make_double3(static_cast<double>(bhi.x - blo.x), static_cast<double>(bhi.y - blo.y),
                                                static_cast<double>(bhi.z - blo.z)),
                                   periodic ? 27u : 1u};
            if constexpr (LANE_WALK)
                {
                // one rope path per lane over the bucket records
                const typename QueryOpT::ThreadData qd = valid ? query.setup(idx) : typename QueryOpT::ThreadData();
                typename OutputOpT::ThreadData result = out.setup(idx, qd);
                uint32_t image_flags = 0;
                for (unsigned int k = 0; k < images.size(); ++k)
                    if (valid && query.overlap(query.get(qd, images.get(k)), root_box)) image_flags |= 1u << k;
                for (int pass = 0; pass < 2; ++pass)
                    {
                    uint32_t flags = image_flags;
                    while (flags)
                        {
                        const unsigned int k = __builtin_ctz(flags);
                        flags &= flags - 1u;
                        const typename QueryOpT::Volume v = query.get(qd, images.get(k));
                        int node = root;
                        for (int budget = 2 * static_cast<int>(n_nodes); node != stop && budget > 0; --budget)
                            {
                            const float4 a = nodes[2u * node];
                            const float4 b = nodes[2u * node + 1u];
                            int next = __float_as_int(b.w);
                            if (walk_kernels::wave::box_distance2(walk_kernels::wave::center(v), a, b) <= walk_kernels::wave::threshold(v))
                                {
                                const int lf = __float_as_int(a.w);
                                if (lf < 0)
                                    {
                                    const unsigned int e = walk_kernels::bucket_end(next, n_internal);
                                    for (unsigned int leaf = ~static_cast<unsigned int>(lf); leaf < e; ++leaf)
                                        {
                                        const float4 p = points[leaf];
                                        if (QueryOpT::refine_point(v, p))
                                            detail::emit_hit(out, result, static_cast<int>(__float_as_uint(p.w)), k, 0);
                                        }
                                    }
                                else
                                    {
                                    next = lf;
                                    }
                                }
                            node = next;
                            }
                        }
                    if constexpr (!detail::relocates<OutputOpT>::value) break;
                    else if (!(valid && out.relocate(result))) break;
                    }
                if (valid) out.finalize(result);
                }
            else
                {
                walk_kernels::wave::walk(out, nodes, points, primitive, root_box, root, stop, query, images, NullTransformOp(), valid, idx,
                           n_internal, n_nodes, stage, lds.wave_compact());
                }
            }
        }
    }

#define NEIGHBOR_BATCH_TRAVERSE_PARAMS                                                                              \
    const OutputOpT out, const float4* __restrict__ nodes, const float4* __restrict__ points,                      \
        const unsigned int* __restrict__ primitive, const float3* __restrict__ tree_lo,                           \
        const float3* __restrict__ tree_hi, const int* __restrict__ rope, const int* __restrict__ roots,          \
        const uint32_t* __restrict__ offsets, const float3* __restrict__ box_lo, const float3* __restrict__ box_hi, \
        const Item* __restrict__ items, const unsigned long long* __restrict__ n_items_ptr,                       \
        unsigned int* __restrict__ counters, const QueryOpT query, const unsigned int periodic, const unsigned int n_nodes
#define NEIGHBOR_BATCH_TRAVERSE_ARGS                                                                              \
    out, nodes, points, primitive, tree_lo, tree_hi, rope, roots, offsets, box_lo, box_hi, items, n_items_ptr, counters, \
        query, periodic, n_nodes

template<bool LANE_WALK, class OutputOpT, class QueryOpT>
__global__ void __launch_bounds__(256) batch_traverse(NEIGHBOR_BATCH_TRAVERSE_PARAMS)
    {
    batch_traverse_body<LANE_WALK, OutputOpT, QueryOpT>(NEIGHBOR_BATCH_TRAVERSE_ARGS);
    }

//! batch_traverse of the wave walk for the compacted exact pair fills, compiled for eight
//! waves per SIMD as the single-tree pair fill is (lbvh_traverse_wave_pairs).
template<class OutputOpT, class QueryOpT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
batch_traverse_pairs(NEIGHBOR_BATCH_TRAVERSE_PARAMS)
    {
    batch_traverse_body<false, OutputOpT, QueryOpT>(NEIGHBOR_BATCH_TRAVERSE_ARGS);
    }

#undef NEIGHBOR_BATCH_TRAVERSE_PARAMS
#undef NEIGHBOR_BATCH_TRAVERSE_ARGS

} // end namespace kernel
} // end namespace batch
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_BATCH_H_
