// LBVHBuild.h -- device kernels of the LBVH build.
//
// Stages and the reference semantics they keep (include/neighbor/kernels/LBVH.cuh):
//   lbvh_codes      Morton code of each leaf-box center (:127-157, :28-52, :105-108),
//                   fused with the digit histograms of the radix sort and with the
//                   clearing of the bubble flags and sort status words.
//   (sort)          RadixSort.h, replaces hipCUB SortPairs (:421-443).
//   lbvh_hierarchy  Karras 2012 Fig. 4 hierarchy (:73-94, :169-243) plus, per node,
//                   its rope, computed from the sorted codes instead of by
//                   backtracking through parents (kernels/LBVHTraverser.cuh:79-95);
//                   for point leaves also the sorted points (SortedPointGather).
//   lbvh_bubble_*   bottom-up AABB union with the second-arrival rule (:270-327),
//                   in LDS per chunk of sorted leaves, then span passes over 64-fold
//                   wider leaf spans (one workgroup each) for the nodes that cross
//                   chunk boundaries, until one span covers every leaf.
//   lbvh_single     N == 1 tree (:340-354).
//   lbvh_refit_single  its refit (new box under the same one-leaf tree).
#ifndef NEIGHBOR_KERNELS_LBVH_BUILD_H_
#define NEIGHBOR_KERNELS_LBVH_BUILD_H_

#include "../BoundingVolumes.h"
#include "../InsertOps.h"
#include "../LBVHData.h"
#include "RadixSort.h"

#include <type_traits>
#include <vector>

namespace neighbor
{
namespace gpu
{
namespace detail
{
//! Position of primitive idx: the op's point when it has a `points` array, else its box center.
template<class InsertOpT>
__device__ __forceinline__ auto insert_point(const InsertOpT& op, unsigned int idx, int) -> decltype(op.points[idx], float3())
    {
    return op.points[idx];
    }
template<class InsertOpT>
__device__ __forceinline__ float3 insert_point(const InsertOpT& op, unsigned int idx, long)
    {
    return op.get(idx).getCenter();
    }
__device__ __forceinline__ float3 insert_point(const PostypeInsertOp& op, unsigned int idx, int)
    {
    return op.point(idx);
    }

//! Insert ops whose leaf box is a function of the primitive's point (the bubble keeps 3
//! floats per leaf in LDS instead of 6); box_of_point repeats the op's own arithmetic.
template<class InsertOpT>
struct point_leaves : std::false_type {};
template<>
struct point_leaves<PointInsertOp> : std::true_type {};
template<>
struct point_leaves<SphereInsertOp> : std::true_type {};
template<>
struct point_leaves<PostypeInsertOp> : std::true_type {};

__device__ __forceinline__ BoundingBox box_of_point(const PointInsertOp&, const float3 p)
    {
    return BoundingBox(p, p);
    }
__device__ __forceinline__ BoundingBox box_of_point(const SphereInsertOp& op, const float3 p)
    {
    return BoundingBox(make_float3(p.x - op.r, p.y - op.r, p.z - op.r), make_float3(p.x + op.r, p.y + op.r, p.z + op.r));
    }
__device__ __forceinline__ BoundingBox box_of_point(const PostypeInsertOp& op, const float3 p)
    {
    if (op.r < 0.0f) return BoundingBox(p, p);
    return BoundingBox(make_float3(p.x - op.r, p.y - op.r, p.z - op.r), make_float3(p.x + op.r, p.y + op.r, p.z + op.r));
    }
template<class InsertOpT>
__device__ __forceinline__ BoundingBox box_of_point(const InsertOpT&, const float3 p)
    {
    return BoundingBox(p, p);   // not used: leaves of other ops keep their whole box
    }
} // end namespace detail

namespace kernel
{

//! Spread the 10 low bits of v so that two zero bits follow each bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v)
    {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
    }

//! 30-bit Morton code, x the most significant bit of each triple.
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z)
    {
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
    }

//! [0,1] fraction to one of 1024 bins, clamped at both ends.
__device__ __forceinline__ uint32_t fraction_to_bin(float f)
    {
    return static_cast<uint32_t>(fminf(fmaxf(f * 1023.0f, 0.0f), 1023.0f));
    }

//! Longest common prefix of the (code, position) keys at sorted positions i and j; -1 off the ends.
__device__ __forceinline__ int common_prefix(const uint32_t* __restrict__ codes, uint32_t n, int i, uint32_t ci, int j)
    {
    if (j < 0 || j >= static_cast<int>(n)) return -1;
    const uint32_t cj = codes[j];
    return (ci == cj) ? 32 + __clz(static_cast<uint32_t>(i ^ j)) : __clz(ci ^ cj);
    }

//! Morton codes + radix digit histograms (+ clears of the per-build scratch words).
template<class InsertOpT>
__global__ void __launch_bounds__(256) lbvh_codes(const InsertOpT insert,
                                                  const float3 lo,
                                                  const float3 hi,
                                                  const uint32_t n,
                                                  uint32_t* __restrict__ codes,
                                                  uint32_t* __restrict__ hist,
                                                  const unsigned int passes,
                                                  uint32_t* __restrict__ clear,
                                                  const size_t clear_words)
    {
    __shared__ uint32_t s_hist[4 * radix::kRadix];
    for (unsigned int j = threadIdx.x; j < passes * radix::kRadix; j += blockDim.x) s_hist[j] = 0u;
    __syncthreads();

    const float3 L = make_float3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t first = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // kUnroll independent primitive loads in flight per thread (few blocks keep the
    // histogram atomics few, so each thread streams many primitives)
    constexpr unsigned int kUnroll = 4;
    for (size_t i0 = first; i0 < n; i0 += kUnroll * stride)
        {
        float3 c[kUnroll];
        #pragma unroll
        for (unsigned int u = 0; u < kUnroll; ++u)
            {
            const size_t i = i0 + u * stride;
            c[u] = i < n ? insert.get(static_cast<uint32_t>(i)).getCenter() : make_float3(0.f, 0.f, 0.f);
            }
        #pragma unroll
        for (unsigned int u = 0; u < kUnroll; ++u)
            {
            const size_t i = i0 + u * stride;
            if (i >= n) break;
            const uint32_t code = morton3(fraction_to_bin((c[u].x - lo.x) / L.x),
                                          fraction_to_bin((c[u].y - lo.y) / L.y),
                                          fraction_to_bin((c[u].z - lo.z) / L.z));
            codes[i] = code;
            for (unsigned int p = 0; p < passes; ++p)
                radix::count_digit(s_hist + p * radix::kRadix, (code >> (p * radix::kBits)) & (radix::kRadix - 1));
            }
        }
    for (size_t i = first; i < clear_words; i += stride) clear[i] = 0u;
    __syncthreads();
    uint32_t* part = radix::hist_part(hist, passes);
    for (unsigned int j = threadIdx.x; j < passes * radix::kRadix; j += blockDim.x)
        if (s_hist[j]) atomicAdd(&part[j], s_hist[j]);
    }

namespace bubble
{
constexpr unsigned int kChunk = 2048;   //!< sorted leaves per workgroup in the chunk pass (up to kLargeN leaves)
constexpr unsigned int kBlock = 1024;   //!< its threads
constexpr unsigned int kChunkLarge = 1024;  //!< above kLargeN leaves: half the LDS, two workgroups per CU
constexpr unsigned int kBlockLarge = 512;
constexpr size_t kLargeN = 6000000;
constexpr unsigned int kFanShift = 6;
constexpr unsigned int kFan = 1u << kFanShift;  //!< sub-spans per workgroup in each span pass
constexpr unsigned int kMaxLevels = 6;  //!< span passes over at most 2^31 leaves in 1024-leaf chunks: 4
//! A span's list is kept as kFan sub-lists, one per sub-span, each node in the sub-list of
//! the sub-span holding its Karras index (one end of its range; appends to a sub-list come
//! from about a dozen nodes instead of the whole span's, so they do not serialize on one
//! counter).  A node of the span's level whose index lies in a sub-span extends past one of
//! that sub-span's two boundaries, so it lies on that boundary's chain of ancestors: at most
//! one node per tree level on each.  Prefix lengths grow strictly along a root-to-leaf path;
//! single-tree keys have 61 of them (30 for distinct 30-bit codes, 31 for equal codes split by
//! position), so a sub-list holds at most 2 x 61 = 122 < kSubCap entries.  Forest keys (box
//! bits above the code) allow up to 83 levels in principle; a sub-list that overflows sets
//! bit 1 of the build's error word instead of losing nodes silently.
constexpr unsigned int kSubCap = 128;
constexpr unsigned int kListCap = kFan * kSubCap;   //!< entries of one span's list
constexpr unsigned int kSpanCap = 1024;     //!< list entries a span workgroup climbs in LDS
constexpr unsigned int kSpanBlock = 1024;   //!< its threads (one entry each)

//! Leaves per workgroup of the chunk pass for n leaves: small trees are latency bound and
//! gain from fewer chunk boundaries; large ones from two resident workgroups per CU.
inline unsigned int chunk_for(size_t n) { return n > kLargeN ? kChunkLarge : kChunk; }
inline unsigned int chunk_shift(unsigned int chunk) { return chunk == kChunkLarge ? 10u : 11u; }

//! Leaves covered by one workgroup of span pass `level` (level 0: the chunk pass).
inline unsigned long long span_leaves(unsigned int level, unsigned int chunk)
    {
    unsigned long long s = chunk;
    for (unsigned int l = 0; l < level; ++l) s *= kFan;
    return s;
    }

//! Number of span passes after the chunk pass (0 when one chunk holds every leaf).
inline unsigned int span_levels(size_t n, unsigned int chunk)
    {
    unsigned int levels = 0;
    while (span_leaves(levels, chunk) < n) ++levels;
    return levels;
    }

inline size_t groups(size_t n, unsigned int level, unsigned int chunk)
    {
    const unsigned long long s = span_leaves(level, chunk);
    return static_cast<size_t>((n + s - 1) / s);
    }

//! Spans of all span passes together (one list each).
inline size_t total_groups(size_t n)
    {
    const unsigned int chunk = chunk_for(n);
    size_t t = 0;
    for (unsigned int l = 1; l <= span_levels(n, chunk); ++l) t += groups(n, l, chunk);
    return t;
    }

//! Words of the span lists' counters, one per sub-list (zero before each build's hierarchy kernel).
inline size_t count_words(size_t n) { return total_groups(n) * kFan; }

//! Words of bubble scratch after the per-node arrival flags: the counters, then (16-byte
//! aligned) kListCap int4 entries per span.
inline size_t list_words(size_t n)
    {
    const size_t t = total_groups(n);
    return t == 0 ? 0 : t * kFan + 3 + t * kListCap * 4;
    }
} // end namespace bubble

//! The span passes' lists of one build, filled by its hierarchy kernel.
/*!
 * Span pass l (1 <= l <= levels) has one workgroup per span of 2^(chunk_shift + 6 l)
 * sorted leaves.  An internal node belongs to level l when its leaf range lies inside one
 * such span but not inside one span of level l - 1 (level 0: inside one chunk, the chunk
 * pass's).  Its entry in its span's list is (node, left, right, flags): flags bit 0 (1) when
 * the left (right) child is of the same level, i.e. completed by the same workgroup; the
 * other children are complete when the pass starts.  A span's list is kFan sub-lists of
 * kSubCap entries with a counter each (sub-list k of span g at index g kFan + k, level l's
 * spans from group0[l]).  The topology never changes under a refit, so a refit reuses the
 * lists of the build.
 */
struct SpanLists
    {
    uint32_t* counts;                           //!< per sub-list
    int4* entries;                              //!< bubble::kSubCap per sub-list
    uint32_t* error;                            //!< bit 1: a list overflowed
    uint32_t group0[bubble::kMaxLevels + 1];
    uint32_t levels;
    uint32_t chunk_shift;

    //! Level of a node with the sorted leaf range [a, b].
    __device__ __forceinline__ unsigned int level(const int a, const int b) const
        {
        const uint32_t x = static_cast<uint32_t>(a ^ b) >> chunk_shift;
        return x == 0u ? 0u : 1u + (31u - static_cast<unsigned int>(__clz(x))) / bubble::kFanShift;
        }
    };

//! The span lists carved from `words` (bubble::list_words(n) words, counters first).
inline SpanLists span_lists(uint32_t* words, const size_t n, uint32_t* error)
    {
    SpanLists sl{};
    const unsigned int chunk = bubble::chunk_for(n);
    sl.levels = bubble::span_levels(n, chunk);
    sl.chunk_shift = bubble::chunk_shift(chunk);
    sl.error = error;
    sl.counts = words;
    size_t g = 0;
    for (unsigned int l = 1; l <= sl.levels && l <= bubble::kMaxLevels; ++l)
        {
        sl.group0[l] = static_cast<uint32_t>(g);
        g += bubble::groups(n, l, chunk);
        }
    const uintptr_t e = reinterpret_cast<uintptr_t>(words + g * bubble::kFan);
    sl.entries = g ? reinterpret_cast<int4*>((e + 15u) & ~static_cast<uintptr_t>(15u)) : nullptr;
    return sl;
    }

//! Sort keys of the single-tree build: (code, position).
struct CodeKeys
    {
    const uint32_t* __restrict__ codes;
    uint32_t n;
    typedef uint32_t Key;

    __device__ __forceinline__ Key key(int i) const { return codes[i]; }
    //! Longest common prefix of the keys at i (key ki) and j; -1 off the ends.
    __device__ __forceinline__ int prefix(int i, Key ki, int j) const { return common_prefix(codes, n, i, ki, j); }
    };

//! Sort keys of the batched (forest) build: (box << 30 | code, position in the box).
/*!
 * Inside one box the prefixes are those of CodeKeys plus 32 (the 64-bit keys share the
 * box bits and the tie-break uses the position relative to the box start), and across
 * boxes they are at most 33, below every in-box prefix (>= 34).  Karras's decisions
 * inside a box are therefore those of the box's own tree: each box is one subtree of
 * the forest with the shape of the tree built for that box alone.
 */
struct BoxKeys
    {
    const unsigned long long* __restrict__ keys;
    const uint32_t* __restrict__ offsets;     //!< first sorted position of each box
    uint32_t n;
    typedef unsigned long long Key;

    __device__ __forceinline__ Key key(int i) const { return keys[i]; }
    __device__ __forceinline__ int prefix(int i, Key ki, int j) const
        {
        if (j < 0 || j >= static_cast<int>(n)) return -1;
        const unsigned long long kj = keys[j];
        if (ki == kj)
            {
            const int s = static_cast<int>(offsets[ki >> 30]);
            return 64 + __clz(static_cast<uint32_t>((i - s) ^ (j - s)));
            }
        return __clzll(static_cast<long long>(ki ^ kj));
        }
    };

//! Rope of a node whose leaf range ends at r (sorted position).
/*!
 * The rope is the right sibling of the first ancestor-or-self that is a left
 * child.  That ancestor ends at r and is the left child of the node split at r,
 * whose right child starts at r+1: the internal node r+1 when its range starts
 * there (its common prefix with r+2 exceeds the one with r), else the leaf r+1.
 */
template<class KeysT>
__device__ __forceinline__ int rope_after(const KeysT& K, int r)
    {
    const int N = static_cast<int>(K.n);
    if (r >= N - 1) return LBVHSentinel;
    const int s = r + 1;
    const typename KeysT::Key cs = K.key(s);
    return (K.prefix(s, cs, s + 1) > K.prefix(s, cs, r)) ? s : (N + r);
    }

//! Karras hierarchy for thread i: internal node i (i < N-1) and leaf i; a node of span level
//! l >= 1 also joins its span's list (SpanLists).
template<class KeysT>
__device__ __forceinline__ void hierarchy_node(const LBVHData& tree, const KeysT& K, const int i, const SpanLists& sl)
    {
    const int N = static_cast<int>(K.n);
    if (i >= N) return;

    // leaf i: rope only (its parent is written by the internal node that owns it)
    tree.rope[N - 1 + i] = rope_after(K, i);
    if (i >= N - 1) return;

    const typename KeysT::Key ci = K.key(i);
    const int up = K.prefix(i, ci, i + 1);
    const int down = K.prefix(i, ci, i - 1);
    const int d = (up >= down) ? 1 : -1;
    const int floor_prefix = (d > 0) ? down : up;

    // exponential then binary search for the far end of the range
    int reach = 2;
    while (K.prefix(i, ci, i + reach * d) > floor_prefix) reach <<= 1;
    int len = 0;
    for (int step = reach >> 1; step > 0; step >>= 1)
        if (K.prefix(i, ci, i + (len + step) * d) > floor_prefix) len += step;
    const int j = i + len * d;
    const int node_prefix = K.prefix(i, ci, j);

    // binary search for the split: last position sharing more than node_prefix with i
    int s = 0;
    int step = len;
    do
        {
        step = (step + 1) >> 1;
        if (s + step < len && K.prefix(i, ci, i + (s + step) * d) > node_prefix) s += step;
        } while (step > 1);
    const int split = i + s * d + (d < 0 ? -1 : 0);

    const int first = (d > 0) ? i : j;
    const int last = (d > 0) ? j : i;
    const int lc = (first == split) ? split + (N - 1) : split;
    const int rc = (last == split + 1) ? split + N : split + 1;
    tree.left[i] = lc;
    tree.right[i] = rc;
    tree.parent[lc] = i;
    tree.parent[rc] = i;
    tree.rope[i] = rope_after(K, last);
    tree.range_end[i] = j;
    if (i == 0) tree.parent[0] = LBVHSentinel;

    const unsigned int level = sl.levels ? sl.level(first, last) : 0u;
    if (level > 0u)
        {
        // the sub-list of the sub-span holding i (span index * kFan + sub-span in the span)
        const unsigned int shift = sl.chunk_shift + (level - 1u) * bubble::kFanShift;
        const uint32_t sub = sl.group0[level] * bubble::kFan + static_cast<uint32_t>(static_cast<unsigned long long>(i) >> shift);
        const int flags = (sl.level(first, split) == level ? 1 : 0) | (sl.level(split + 1, last) == level ? 2 : 0);
        const uint32_t slot = atomicAdd(&sl.counts[sub], 1u);
        if (slot < bubble::kSubCap) sl.entries[static_cast<size_t>(sub) * bubble::kSubCap + slot] = make_int4(i, lc, rc, flags);
        else atomicOr(sl.error, 2u);
        }
    }

//! Sort keys of the single-tree build read through an LDS window of sorted codes.
/*!
 * The searches of node i probe positions i +- 2^k and then bisect; for all but the
 * few nodes with ranges longer than the halo they stay inside [i0 - halo, i0 + block +
 * halo) of the workgroup, whose codes are read once, coalesced, into LDS.  Positions
 * outside the window read the global array: the keys, and so the tree, are those of
 * CodeKeys.
 */
struct WindowKeys
    {
    const uint32_t* __restrict__ codes;
    const uint32_t* s_codes;
    int w0;
    int w1;
    uint32_t n;
    typedef uint32_t Key;

    __device__ __forceinline__ Key key(int i) const { return (i >= w0 && i < w1) ? s_codes[i - w0] : codes[i]; }
    __device__ __forceinline__ int prefix(int i, Key ki, int j) const
        {
        if (j < 0 || j >= static_cast<int>(n)) return -1;
        const uint32_t cj = key(j);
        return (ki == cj) ? 32 + __clz(static_cast<uint32_t>(i ^ j)) : __clz(ki ^ cj);
        }
    };

//! Karras hierarchy of the single-tree build (block size <= 1024).
//! GatherT (not radix::NoGather): thread i also gathers leaf i's point (gather.load of primitive
//! i, issued before the node's searches so that its latency hides behind them) and stores it
//! into the sorted points (gather.store).  clear: the sort's histograms and tile counters
//! (clear_words words), no longer read once the sort is done, zeroed for the next build.
template<class GatherT = radix::NoGather>
__global__ void __launch_bounds__(1024) lbvh_hierarchy(const LBVHData tree, const uint32_t* __restrict__ codes, const uint32_t n,
                                                       const SpanLists sl, uint32_t* __restrict__ clear,
                                                       const uint32_t clear_words, const GatherT gather = GatherT())
    {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < clear_words; j += gridDim.x * blockDim.x) clear[j] = 0u;
    constexpr bool kGather = !std::is_same<GatherT, radix::NoGather>::value;
    constexpr int kHalo = 256;
    __shared__ uint32_t s_codes[1024 + 2 * kHalo];
    const int i0 = static_cast<int>(blockIdx.x * blockDim.x);
    const int i = i0 + static_cast<int>(threadIdx.x);
    uint32_t pid = 0u;
    typename GatherT::Value pt{};
    if constexpr (kGather)
        if (i < static_cast<int>(n))
            {
            pid = tree.primitive[i];
            pt = gather.load(pid);
            }
    const int w0 = max(i0 - kHalo, 0);
    const int w1 = min(i0 + static_cast<int>(blockDim.x) + kHalo, static_cast<int>(n));
    for (int j = w0 + static_cast<int>(threadIdx.x); j < w1; j += static_cast<int>(blockDim.x)) s_codes[j - w0] = codes[j];
    __syncthreads();
    hierarchy_node(tree, WindowKeys{codes, s_codes, w0, w1, n}, i, sl);
    if constexpr (kGather)
        if (i < static_cast<int>(n)) gather.store(static_cast<uint32_t>(i), pid, pt);
    }

//! Union of a (left, right) pair of boxes with strict comparisons: left wins ties.
__device__ __forceinline__ void merge_boxes(float3& lo, float3& hi, const float3& llo, const float3& lhi,
                                            const float3& rlo, const float3& rhi)
    {
    lo = make_float3(rlo.x < llo.x ? rlo.x : llo.x, rlo.y < llo.y ? rlo.y : llo.y, rlo.z < llo.z ? rlo.z : llo.z);
    hi = make_float3(rhi.x > lhi.x ? rhi.x : lhi.x, rhi.y > lhi.y ? rhi.y : lhi.y, rhi.z > lhi.z ? rhi.z : lhi.z);
    }

//! Point gather of the build's hierarchy kernel for point leaves: each primitive's point, in
//! Morton order, as (x, y, z, bits(primitive)) -- the sorted points the chunk pass and the
//! traversals read.  The random reads (a whole line per 12-byte point for unordered input)
//! overlap the hierarchy's searches there instead of heading the chunk pass.
template<class InsertOpT>
struct SortedPointGather
    {
    typedef float3 Value;
    InsertOpT insert;
    float4* out;

    __device__ __forceinline__ float3 load(const uint32_t pid) const { return detail::insert_point(insert, pid, 0); }
    __device__ __forceinline__ void store(const uint32_t dst, const uint32_t pid, const float3& p) const
        {
        out[dst] = make_float4(p.x, p.y, p.z, __uint_as_float(pid));
        }
    };


//! Bubble pass 1: all nodes whose leaf range lies inside one chunk of sorted leaves.
/*!
 * A workgroup owns the leaves [cs, cs+CHUNK).  An internal node inside the chunk
 * has its Karras index (one end of its range) in the chunk too, so its arrival
 * counter, its box, its children and its parent live in LDS at slot index-cs (read
 * once, coalesced, before any climb).  The leaves' boxes are computed once, written
 * out coalesced and kept in LDS; the second arrival at a node merges the sibling box
 * (from LDS) in (left, right) order, keeps the result in LDS and climbs on.  A climb
 * stops at a parent spanning the chunk boundary (a node of the span passes' lists).  The
 * internal boxes leave LDS in one coalesced pass at the end (the span passes read those
 * of the nodes whose parents cross after the kernel boundary).  Only workgroup-scope
 * ordering is used; no global flag.
 */
template<class InsertOpT, unsigned int CHUNK, unsigned int BLOCK>
__global__ void __launch_bounds__(BLOCK) lbvh_bubble_chunk(const LBVHData tree,
                                                                    const InsertOpT insert,
                                                                    const int* __restrict__ range_end,
                                                                    const uint32_t n,
                                                                    const bool sorted_ready)
    {
    constexpr int kInside = static_cast<int>(0x80000000u);
    constexpr bool kPointLeaves = detail::point_leaves<InsertOpT>::value;
    constexpr unsigned int kPer = CHUNK / BLOCK;                     // leaves per thread
    // s_up[j]: the parent of internal node cs + j, chunk-relative, in the low 16 bits (kUpOutside:
    // outside the chunk, kUpRoot: none), and the arrivals at node cs + j above them
    constexpr uint32_t kUpOutside = 0xffffu;
    constexpr uint32_t kUpRoot = 0xfffeu;
    constexpr uint32_t kArrival = 1u << 16;
    static_assert(CHUNK % BLOCK == 0, "whole leaves per thread");
    static_assert(CHUNK < kUpRoot, "chunk-relative parents in 16 bits");
    __shared__ float s_box[6][CHUNK];   // internal node cs + j
    // point leaves keep no LDS copy: a thread holds its own leaves' points, and a sibling leaf's
    // point is read back from the sorted points this workgroup wrote (72 KB of LDS for 2048
    // leaves: two workgroups per CU); other leaves keep their boxes here
    __shared__ float s_leaf[kPointLeaves ? 1 : 6][kPointLeaves ? 1 : CHUNK];
    __shared__ uint32_t s_up[CHUNK];
    __shared__ int s_left[CHUNK];      // left child; sign bit: the node's range lies in the chunk
    __shared__ int s_right[CHUNK];
    int lparent[kPer];                 // parents of this thread's leaves (same leaves in both loops)
    float3 lpoint[kPer];               // their points (point leaves)

    const int cs = static_cast<int>(blockIdx.x * CHUNK);
    const int ce = min(cs + static_cast<int>(CHUNK), static_cast<int>(n));
    const int Ni = static_cast<int>(n) - 1;
    const int tid = static_cast<int>(threadIdx.x);
    const auto leaf_box = [&](const int j) {
        if constexpr (kPointLeaves)
            {
            const float4 sp = tree.sorted_points[cs + j];
            return detail::box_of_point(insert, make_float3(sp.x, sp.y, sp.z));
            }
        else
            {
            return BoundingBox(make_float3(s_leaf[0][j], s_leaf[1][j], s_leaf[2][j]),
                               make_float3(s_leaf[3 % 6][j], s_leaf[4 % 6][j], s_leaf[5 % 6][j]));
            }
    };

    // the chunk's slice of the hierarchy and its leaf boxes: every climb below runs in LDS.  All
    // loads of the thread's leaves are issued before any store, so that they travel together
    // (the stores through the tree's pointers would otherwise keep each load behind them)
    int r_par[kPer], r_end[kPer], r_left[kPer], r_right[kPer], r_up[kPer];
    float4 r_point[kPer];
    unsigned int r_pid[kPer];
    #pragma unroll
    for (unsigned int k = 0; k < kPer; ++k)
        {
        const int i = cs + tid + static_cast<int>(k * BLOCK);
        r_par[k] = LBVHSentinel;
        r_end[k] = r_left[k] = r_right[k] = 0;
        r_up[k] = LBVHSentinel;
        r_point[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        r_pid[k] = 0u;
        if (i >= ce) continue;
        r_par[k] = tree.parent[Ni + i];
        if (i < Ni)
            {
            r_end[k] = range_end[i];
            r_left[k] = tree.left[i];
            r_right[k] = tree.right[i];
            r_up[k] = tree.parent[i];
            }
        if (kPointLeaves && sorted_ready) r_point[k] = tree.sorted_points[i];
        else r_pid[k] = tree.primitive[i];
        }
    #pragma unroll
    for (unsigned int k = 0; k < kPer; ++k)
        {
        const int j = tid + static_cast<int>(k * BLOCK);
        s_up[j] = kUpOutside;
        lparent[k] = LBVHSentinel;
        lpoint[k] = make_float3(0.0f, 0.0f, 0.0f);
        const int i = cs + j;
        if (i >= ce) continue;
        lparent[k] = r_par[k];
        if (i < Ni)
            {
            const int q = r_end[k];
            const bool inside = min(i, q) >= cs && max(i, q) < ce;
            s_left[j] = r_left[k] | (inside ? kInside : 0);
            s_right[j] = r_right[k];
            const int pp = r_up[k];
            s_up[j] = pp == LBVHSentinel ? kUpRoot : (pp >= cs && pp < ce ? static_cast<uint32_t>(pp - cs) : kUpOutside);
            }
        if constexpr (kPointLeaves)
            {
            float3 pt;
            if (sorted_ready)       // gathered by the hierarchy kernel (SortedPointGather)
                {
                pt = make_float3(r_point[k].x, r_point[k].y, r_point[k].z);
                }
            else
                {
                const unsigned int pid = r_pid[k];
                pt = detail::insert_point(insert, pid, 0);
                tree.sorted_points[i] = make_float4(pt.x, pt.y, pt.z, __uint_as_float(pid));
                }
            const BoundingBox b = detail::box_of_point(insert, pt);
            lpoint[k] = pt;
            tree.lo[Ni + i] = b.lo;
            tree.hi[Ni + i] = b.hi;
            }
        else
            {
            const unsigned int pid = r_pid[k];
            const BoundingBox b = insert.get(pid);
            s_leaf[0][j] = b.lo.x; s_leaf[1][j] = b.lo.y; s_leaf[2][j] = b.lo.z;
            s_leaf[3 % 6][j] = b.hi.x; s_leaf[4 % 6][j] = b.hi.y; s_leaf[5 % 6][j] = b.hi.z;
            tree.lo[Ni + i] = b.lo;
            tree.hi[Ni + i] = b.hi;
            const float3 pt = detail::insert_point(insert, pid, 0);
            tree.sorted_points[i] = make_float4(pt.x, pt.y, pt.z, __uint_as_float(pid));
            }
        }
    __syncthreads();    // also orders the sorted points written above before the reads below

    #pragma unroll
    for (unsigned int k = 0; k < kPer; ++k)
        {
        const int i = cs + tid + static_cast<int>(k * BLOCK);
        if (i >= ce) continue;
        const int j0 = i - cs;
        BoundingBox lb;
        if constexpr (kPointLeaves) lb = detail::box_of_point(insert, lpoint[k]);
        else lb = leaf_box(j0);
        float3 lo = lb.lo;
        float3 hi = lb.hi;
        int node = Ni + i;
        int p = lparent[k];
        // keys have at most 96 significant bits (forest keys with position tie-break): bounded depth
        for (int level = 0; level < 128 && p != LBVHSentinel; ++level)
            {
            if (p < cs || p >= ce || s_left[p - cs] >= 0) break;   // the parent crosses the chunk boundary
            const int ps = p - cs;
            const uint32_t up = __hip_atomic_fetch_add(&s_up[ps], kArrival, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if ((up >> 16) == 0u) break;

            const int l = s_left[ps] & ~kInside;
            const bool mine_left = (l == node);
            const int sib = mine_left ? s_right[ps] : l;
            float3 slo, shi;
            if (sib >= Ni)
                {
                const BoundingBox sb = leaf_box(sib - Ni - cs);
                slo = sb.lo;
                shi = sb.hi;
                }
            else
                {
                const int sj = sib - cs;
                slo = make_float3(s_box[0][sj], s_box[1][sj], s_box[2][sj]);
                shi = make_float3(s_box[3][sj], s_box[4][sj], s_box[5][sj]);
                }
            if (mine_left) merge_boxes(lo, hi, lo, hi, slo, shi);
            else merge_boxes(lo, hi, slo, shi, lo, hi);
            // the node's box in LDS before the arrival at its parent (released by that atomic)
            s_box[0][ps] = lo.x; s_box[1][ps] = lo.y; s_box[2][ps] = lo.z;
            s_box[3][ps] = hi.x; s_box[4][ps] = hi.y; s_box[5][ps] = hi.z;
            node = p;
            const uint32_t rel = up & 0xffffu;
            // a parent outside the chunk becomes ce (any index past the chunk ends the climb)
            p = rel == kUpRoot ? LBVHSentinel : (rel == kUpOutside ? ce : cs + static_cast<int>(rel));
            }
        }
    __syncthreads();

    // every internal node inside the chunk is complete: its box leaves LDS coalesced
    for (int j = static_cast<int>(threadIdx.x); j < static_cast<int>(CHUNK); j += static_cast<int>(blockDim.x))
        {
        const int i = cs + j;
        if (i >= ce || i >= Ni || s_left[j] >= 0) continue;
        tree.lo[i] = make_float3(s_box[0][j], s_box[1][j], s_box[2][j]);
        tree.hi[i] = make_float3(s_box[3][j], s_box[4][j], s_box[5][j]);
        }
    }

namespace bubble
{
//! Climb of the span pass's global-memory path: from `node` (complete, its box in tree.lo/hi)
//! through the ancestors whose leaf range lies in [s0, s1), with arrival counters in `flags`
//! (workgroup scope: all climbs into a node of the span come from its workgroup).  Each level
//! loads the next parent's record together with the sibling's box (two round trips per level).
__device__ __forceinline__ void climb_global(const LBVHData& tree, const int* __restrict__ range_end,
                                             uint32_t* __restrict__ flags, int node, const long long s0,
                                             const long long s1)
    {
    float3 lo = tree.lo[node], hi = tree.hi[node];
    int p = tree.parent[node];
    int q = 0, l = 0, r = 0, pp = LBVHSentinel;
    if (p != LBVHSentinel)
        {
        q = range_end[p];
        l = tree.left[p];
        r = tree.right[p];
        pp = tree.parent[p];
        }
    for (int level = 0; level < 128 && p != LBVHSentinel; ++level)
        {
        if (static_cast<long long>(min(p, q)) < s0 || static_cast<long long>(max(p, q)) >= s1) break;
        const uint32_t arrived = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (arrived == 0u) break;
        const bool mine_left = (l == node);
        const int sib = mine_left ? r : l;
        // the sibling's box only after the counter: it was written before the sibling's arrival
        const float3 slo = tree.lo[sib];
        const float3 shi = tree.hi[sib];
        int q2 = 0, l2 = 0, r2 = 0, pp2 = LBVHSentinel;
        if (pp != LBVHSentinel)
            {
            q2 = range_end[pp];
            l2 = tree.left[pp];
            r2 = tree.right[pp];
            pp2 = tree.parent[pp];
            }
        if (mine_left) merge_boxes(lo, hi, lo, hi, slo, shi);
        else merge_boxes(lo, hi, slo, shi, lo, hi);
        tree.lo[p] = lo;
        tree.hi[p] = hi;
        node = p;
        p = pp;
        q = q2;
        l = l2;
        r = r2;
        pp = pp2;
        }
    }

//! Slot of an LDS open-addressing table of 2 kSpanCap keys (node indices).
__device__ __forceinline__ uint32_t span_hash(const int key)
    {
    return (static_cast<uint32_t>(key) * 0x9e3779b1u) >> (32u - 11u);
    }
static_assert(2 * kSpanCap == (1u << 11), "span_hash covers 2 kSpanCap slots");
} // end namespace bubble

//! Bubble span pass: the nodes of one span's list (SpanLists), one workgroup per span.
/*!
 * The list holds every node of the span's level whose range lies in the span, with its
 * children; the children of lower levels are complete.  Up to kSpanCap entries the climb
 * runs in LDS: the entries are numbered by their list position, an LDS hash table from
 * node index to entry links each same-level child to its parent entry, and each entry
 * keeps the boxes of its two children and an arrival counter.  The boxes of the
 * lower-level children are read first, all at once; an entry with two of them starts a
 * climb, and a climb arriving second at a parent merges the sibling's box (left, right
 * order) and goes on.  No global memory is touched during the climbs; the entries' boxes
 * leave LDS at the end.  A longer list (never met with up to 743 entries per span on
 * uniform and clustered points) climbs from the lower-level children with arrival counters
 * in global memory instead (flags, zero on entry), as does every list longer than lds_limit
 * (a testing knob: 0 sends every span through the global-memory path).
 */
__global__ void __launch_bounds__(bubble::kSpanBlock) lbvh_bubble_span(const LBVHData tree,
                                                                       const int* __restrict__ range_end,
                                                                       uint32_t* __restrict__ flags,
                                                                       const uint32_t* __restrict__ counts,
                                                                       const int4* __restrict__ entries,
                                                                       uint32_t* __restrict__ error,
                                                                       const unsigned long long span,
                                                                       const uint32_t lds_limit)
    {
    constexpr unsigned int C = bubble::kSpanCap;
    constexpr unsigned int H = 2 * C;                // hash slots
    constexpr uint32_t kEmpty = 0xffffffffu;
    constexpr uint32_t kTop = 0xffffu;               // s_up: the parent is of a higher level
    // the two children's boxes of entry e: s_box[side * 6 + component][e]; during set-up the
    // first four rows hold the hash table instead (keys, entries)
    __shared__ float s_box[12][C];
    __shared__ uint32_t s_up[C];                     // parent entry | side << 16 (side 1: right child)
    __shared__ uint32_t s_cnt[C];                    // children complete or arrived
    __shared__ uint32_t s_off[bubble::kFan + 1];     // first entry of each sub-list
    uint32_t* s_key = reinterpret_cast<uint32_t*>(&s_box[0][0]);
    uint32_t* s_val = reinterpret_cast<uint32_t*>(&s_box[2][0]);
    static_assert(C <= kTop, "entry indices in 16 bits");
    static_assert(bubble::kFan == 64, "one wave scans the sub-list counts");

    const uint32_t g = blockIdx.x;
    const unsigned int tid = threadIdx.x;
    const int4* list = entries + static_cast<size_t>(g) * bubble::kListCap;
    if (tid < bubble::kFan)
        {
        // the entries are numbered sub-list after sub-list
        const uint32_t c = min(counts[g * bubble::kFan + tid], bubble::kSubCap);
        uint32_t incl = c;
        #pragma unroll
        for (unsigned int d = 1; d < bubble::kFan; d <<= 1)
            {
            const uint32_t o = __shfl_up(incl, d, 64);
            if (tid >= d) incl += o;
            }
        s_off[tid + 1] = incl;
        if (tid == 0) s_off[0] = 0u;
        }
    __syncthreads();
    const uint32_t E = s_off[bubble::kFan];
    // entry e: its sub-list is the last one starting at or before e
    const auto entry = [&](const uint32_t e) {
        unsigned int k = 0;
        #pragma unroll
        for (unsigned int step = bubble::kFan / 2; step > 0; step >>= 1)
            if (s_off[k + step] <= e) k += step;
        return list[k * bubble::kSubCap + (e - s_off[k])];
    };
    if (E > min(lds_limit, C))
        {
        const long long s0 = static_cast<long long>(g) * static_cast<long long>(span);
        const long long s1 = s0 + static_cast<long long>(span);
        for (uint32_t e = tid; e < E; e += blockDim.x)
            {
            const int4 en = entry(e);
            if (!(en.w & 1)) bubble::climb_global(tree, range_end, flags, en.y, s0, s1);
            if (!(en.w & 2)) bubble::climb_global(tree, range_end, flags, en.z, s0, s1);
            }
        return;
        }

    const bool has = tid < E;
    const int4 en = has ? entry(tid) : make_int4(0, 0, 0, 3);
    // boxes of the lower-level children, in flight during the set-up
    float3 clo[2], chi[2];
    #pragma unroll
    for (int side = 0; side < 2; ++side)
        {
        clo[side] = chi[side] = make_float3(0.0f, 0.0f, 0.0f);
        if (!((en.w >> side) & 1))
            {
            const int c = side ? en.z : en.y;
            clo[side] = tree.lo[c];
            chi[side] = tree.hi[c];
            }
        }
    for (unsigned int h = tid; h < H; h += blockDim.x) s_key[h] = kEmpty;
    for (unsigned int e = tid; e < C; e += blockDim.x) s_up[e] = kTop;
    __syncthreads();
    if (has)
        {
        uint32_t h = bubble::span_hash(en.x);
        for (unsigned int probe = 0; probe < H; ++probe, h = (h + 1u) & (H - 1u))
            if (atomicCAS(&s_key[h], kEmpty, static_cast<uint32_t>(en.x)) == kEmpty)
                {
                s_val[h] = tid;
                break;
                }
        }
    __syncthreads();
    // each entry links its same-level children to itself
    if (has)
        {
        #pragma unroll
        for (int side = 0; side < 2; ++side)
            {
            if (!((en.w >> side) & 1)) continue;
            const uint32_t key = static_cast<uint32_t>(side ? en.z : en.y);
            uint32_t h = bubble::span_hash(static_cast<int>(key));
            uint32_t found = kEmpty;
            for (unsigned int probe = 0; probe < H; ++probe, h = (h + 1u) & (H - 1u))
                {
                const uint32_t k = s_key[h];
                if (k == key)
                    {
                    found = s_val[h];
                    break;
                    }
                if (k == kEmpty) break;
                }
            if (found < C) s_up[found] = tid | (static_cast<uint32_t>(side) << 16);
            else atomicOr(error, 4u);    // a child missing from its list (after an overflow)
            }
        }
    __syncthreads();    // the hash table is no longer read: its rows take boxes from here on
    if (has)
        {
        uint32_t c = 0u;
        #pragma unroll
        for (int side = 0; side < 2; ++side)
            {
            if ((en.w >> side) & 1) continue;
            s_box[side * 6 + 0][tid] = clo[side].x; s_box[side * 6 + 1][tid] = clo[side].y; s_box[side * 6 + 2][tid] = clo[side].z;
            s_box[side * 6 + 3][tid] = chi[side].x; s_box[side * 6 + 4][tid] = chi[side].y; s_box[side * 6 + 5][tid] = chi[side].z;
            ++c;
            }
        s_cnt[tid] = c;
        }
    __syncthreads();

    if (has && en.w == 0)
        {
        // both children complete: climb from this entry
        float3 lo, hi;
        merge_boxes(lo, hi, clo[0], chi[0], clo[1], chi[1]);
        uint32_t e = tid;
        for (int level = 0; level < 128; ++level)
            {
            // the entry's own box (read back at the end), then its arrival at the parent entry
            s_box[0][e] = lo.x; s_box[1][e] = lo.y; s_box[2][e] = lo.z;
            s_box[3][e] = hi.x; s_box[4][e] = hi.y; s_box[5][e] = hi.z;
            const uint32_t up = s_up[e];
            const uint32_t q = up & 0xffffu;
            if (q == kTop) break;
            const uint32_t side = up >> 16;
            const uint32_t row = side * 6u;
            s_box[row + 0][q] = lo.x; s_box[row + 1][q] = lo.y; s_box[row + 2][q] = lo.z;
            s_box[row + 3][q] = hi.x; s_box[row + 4][q] = hi.y; s_box[row + 5][q] = hi.z;
            const uint32_t arrived = __hip_atomic_fetch_add(&s_cnt[q], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (arrived == 0u) break;
            const uint32_t srow = 6u - row;     // the sibling's side
            const float3 slo = make_float3(s_box[srow + 0][q], s_box[srow + 1][q], s_box[srow + 2][q]);
            const float3 shi = make_float3(s_box[srow + 3][q], s_box[srow + 4][q], s_box[srow + 5][q]);
            if (side == 0u) merge_boxes(lo, hi, lo, hi, slo, shi);
            else merge_boxes(lo, hi, slo, shi, lo, hi);
            e = q;
            }
        }
    __syncthreads();
    if (has)
        {
        tree.lo[en.x] = make_float3(s_box[0][tid], s_box[1][tid], s_box[2][tid]);
        tree.hi[en.x] = make_float3(s_box[3][tid], s_box[4][tid], s_box[5][tid]);
        }
    }

//! Bottom-up boxes of a built hierarchy: the chunk pass, then span passes of kFan-fold
//! wider spans until one span covers every leaf, over the span lists of the hierarchy
//! kernel (sl).  flags: n-1 words, zero on entry (the span passes' global-memory path).
//! sorted_ready: tree.sorted_points already holds the point leaves' points in Morton order
//! (the hierarchy kernel's SortedPointGather); otherwise the chunk pass gathers and writes them.
//! lds_limit: span lists longer than this climb through global memory (bubble::kSpanCap at most).
template<class InsertOpT>
inline void launch_bubble(const LBVHData& tree, const InsertOpT& insert, const int* range_end, uint32_t* flags,
                          const SpanLists& sl, uint32_t n, hipStream_t stream, bool sorted_ready = false,
                          uint32_t lds_limit = bubble::kSpanCap)
    {
    const unsigned int chunk = bubble::chunk_for(n);
    if (chunk == bubble::kChunk)
        hipLaunchKernelGGL((lbvh_bubble_chunk<InsertOpT, bubble::kChunk, bubble::kBlock>), dim3(div_up(n, chunk)),
                           dim3(bubble::kBlock), 0, stream, tree, insert, range_end, n, sorted_ready);
    else
        hipLaunchKernelGGL((lbvh_bubble_chunk<InsertOpT, bubble::kChunkLarge, bubble::kBlockLarge>), dim3(div_up(n, chunk)),
                           dim3(bubble::kBlockLarge), 0, stream, tree, insert, range_end, n, sorted_ready);
    for (unsigned int l = 1; l <= sl.levels; ++l)
        hipLaunchKernelGGL(lbvh_bubble_span, dim3(static_cast<unsigned int>(bubble::groups(n, l, chunk))), dim3(bubble::kSpanBlock),
                           0, stream, tree, range_end, flags, sl.counts + static_cast<size_t>(sl.group0[l]) * bubble::kFan,
                           sl.entries + static_cast<size_t>(sl.group0[l]) * bubble::kListCap, sl.error,
                           bubble::span_leaves(l, chunk), lds_limit);
    check_launch("lbvh_bubble");
    }

//! One-primitive tree: the leaf is the root (node 0).
template<class InsertOpT>
__global__ void lbvh_single(const LBVHData tree, const InsertOpT insert, uint32_t* __restrict__ codes,
                            const float3 lo, const float3 hi)
    {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const BoundingBox b = insert.get(0);
    const float3 c = b.getCenter();
    codes[0] = morton3(fraction_to_bin((c.x - lo.x) / (hi.x - lo.x)),
                       fraction_to_bin((c.y - lo.y) / (hi.y - lo.y)),
                       fraction_to_bin((c.z - lo.z) / (hi.z - lo.z)));
    tree.parent[0] = LBVHSentinel;
    tree.primitive[0] = 0u;
    tree.rope[0] = LBVHSentinel;
    tree.lo[0] = b.lo;
    tree.hi[0] = b.hi;
    if (tree.sorted_points)
        {
        const float3 pt = detail::insert_point(insert, 0u, 0);
        tree.sorted_points[0] = make_float4(pt.x, pt.y, pt.z, __uint_as_float(0u));
        }
    }

//! Refit of the one-primitive tree: the root leaf's box and sorted point.
template<class InsertOpT>
__global__ void lbvh_refit_single(const LBVHData tree, const InsertOpT insert)
    {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const BoundingBox b = insert.get(0);
    tree.lo[0] = b.lo;
    tree.hi[0] = b.hi;
    if (tree.sorted_points)
        {
        const float3 pt = detail::insert_point(insert, 0u, 0);
        tree.sorted_points[0] = make_float4(pt.x, pt.y, pt.z, __uint_as_float(0u));
        }
    }

} // end namespace kernel
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_LBVH_BUILD_H_
