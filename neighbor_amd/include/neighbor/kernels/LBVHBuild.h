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
    for (unsigned int j = threadIdx.x; j < passes * radix::kRadix; j += blockDim.x)
        if (s_hist[j]) atomicAdd(&hist[j], s_hist[j]);
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

//! Karras hierarchy for thread i: internal node i (i < N-1) and leaf i.
template<class KeysT>
__device__ __forceinline__ void hierarchy_node(const LBVHData& tree, const KeysT& K, const int i)
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
//! into the sorted points (gather.store).
template<class GatherT = radix::NoGather>
__global__ void __launch_bounds__(1024) lbvh_hierarchy(const LBVHData tree, const uint32_t* __restrict__ codes, const uint32_t n,
                                                       const GatherT gather = GatherT())
    {
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
    hierarchy_node(tree, WindowKeys{codes, s_codes, w0, w1, n}, i);
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

namespace bubble
{
constexpr unsigned int kChunk = 2048;   //!< sorted leaves per workgroup in the chunk pass (up to kLargeN leaves)
constexpr unsigned int kBlock = 1024;   //!< its threads
constexpr unsigned int kChunkLarge = 1024;  //!< above kLargeN leaves: half the LDS, two workgroups per CU
constexpr unsigned int kBlockLarge = 512;
constexpr size_t kLargeN = 6000000;
constexpr unsigned int kTopBlock = 1024;
constexpr unsigned int kFan = 64;       //!< sub-spans per workgroup in each span pass
//! Entries a sub-span can hand up: the maximal subtrees inside it whose parent is not,
//! at most two per tree level, with the level count bounded by 64 in practice.
constexpr unsigned int kPerSub = 128;
constexpr unsigned int kListCap = kFan * kPerSub;

//! Leaves per workgroup of the chunk pass for n leaves: small trees are latency bound and
//! gain from fewer chunk boundaries; large ones from two resident workgroups per CU.
inline unsigned int chunk_for(size_t n) { return n > kLargeN ? kChunkLarge : kChunk; }

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

//! Words of bubble scratch after the per-node arrival flags: a counter and a list per
//! workgroup of every span pass.
inline size_t list_words(size_t n)
    {
    const unsigned int chunk = chunk_for(n);
    size_t w = 0;
    for (unsigned int l = 1; l <= span_levels(n, chunk); ++l) w += groups(n, l, chunk) * (1 + kListCap);
    return w;
    }
} // end namespace bubble

//! Bubble pass 1: all nodes whose leaf range lies inside one chunk of sorted leaves.
/*!
 * A workgroup owns the leaves [cs, cs+CHUNK).  An internal node inside the chunk
 * has its Karras index (one end of its range) in the chunk too, so its arrival
 * counter, its box, its children and its parent live in LDS at slot index-cs (read
 * once, coalesced, before any climb).  The leaves' boxes are computed once, written
 * out coalesced and kept in LDS; the second arrival at a node merges the sibling box
 * (from LDS) in (left, right) order, keeps the result in LDS and climbs on.  A climb
 * that reaches a parent spanning the chunk boundary hands its node over to the span
 * pass.  The internal boxes leave LDS in one coalesced pass at the end (handed-up
 * nodes included: the span pass reads them after the kernel boundary).  Only
 * workgroup-scope ordering is used; no global flag.
 */
template<class InsertOpT, unsigned int CHUNK, unsigned int BLOCK>
__global__ void __launch_bounds__(BLOCK) lbvh_bubble_chunk(const LBVHData tree,
                                                                    const InsertOpT insert,
                                                                    const int* __restrict__ range_end,
                                                                    uint32_t* __restrict__ out_count,
                                                                    int* __restrict__ out_list,
                                                                    uint32_t* __restrict__ error,
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
    // nodes handed to the span pass, staged here and appended to the span's list with one
    // global atomic per workgroup (one per node serialized on the span's counter)
    __shared__ int s_hand[bubble::kPerSub];
    __shared__ uint32_t s_nhand;
    __shared__ uint32_t s_hand_base;
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

    // the chunk's slice of the hierarchy and its leaf boxes: every climb below runs in LDS
    if (tid == 0) s_nhand = 0u;
    #pragma unroll
    for (unsigned int k = 0; k < kPer; ++k)
        {
        const int j = tid + static_cast<int>(k * BLOCK);
        s_up[j] = kUpOutside;
        lparent[k] = LBVHSentinel;
        lpoint[k] = make_float3(0.0f, 0.0f, 0.0f);
        const int i = cs + j;
        if (i >= ce) continue;
        lparent[k] = tree.parent[Ni + i];
        if (i < Ni)
            {
            const int q = range_end[i];
            const bool inside = min(i, q) >= cs && max(i, q) < ce;
            s_left[j] = tree.left[i] | (inside ? kInside : 0);
            s_right[j] = tree.right[i];
            const int pp = tree.parent[i];
            s_up[j] = pp == LBVHSentinel ? kUpRoot : (pp >= cs && pp < ce ? static_cast<uint32_t>(pp - cs) : kUpOutside);
            }
        if constexpr (kPointLeaves)
            {
            float3 pt;
            if (sorted_ready)       // gathered by the hierarchy kernel (SortedPointGather)
                {
                const float4 sp = tree.sorted_points[i];
                pt = make_float3(sp.x, sp.y, sp.z);
                }
            else
                {
                const unsigned int pid = tree.primitive[i];
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
            const unsigned int pid = tree.primitive[i];
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
            if (p < cs || p >= ce || s_left[p - cs] >= 0)
                {
                // hand the node to the span pass of the next level (workgroup = kFan chunks)
                const uint32_t slot = atomicAdd(&s_nhand, 1u);
                if (slot < bubble::kPerSub) s_hand[slot] = node;
                else atomicOr(error, 2u);
                break;
                }
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
            // a parent outside the chunk becomes ce (any index past the chunk hands the node up)
            p = rel == kUpRoot ? LBVHSentinel : (rel == kUpOutside ? ce : cs + static_cast<int>(rel));
            }
        }
    __syncthreads();

    // the handed-up nodes join the span's list: one reservation for the workgroup
    const uint32_t nhand = min(s_nhand, bubble::kPerSub);
    const uint32_t g = blockIdx.x / bubble::kFan;
    if (tid == 0 && nhand > 0u)
        {
        if (out_count) s_hand_base = atomicAdd(&out_count[g], nhand);
        else atomicOr(error, 4u);    // one chunk holds every leaf: unreachable
        }
    // every internal node inside the chunk is complete: its box leaves LDS coalesced
    for (int j = static_cast<int>(threadIdx.x); j < static_cast<int>(CHUNK); j += static_cast<int>(blockDim.x))
        {
        const int i = cs + j;
        if (i >= ce || i >= Ni || s_left[j] >= 0) continue;
        tree.lo[i] = make_float3(s_box[0][j], s_box[1][j], s_box[2][j]);
        tree.hi[i] = make_float3(s_box[3][j], s_box[4][j], s_box[5][j]);
        }
    __syncthreads();
    if (out_count && static_cast<uint32_t>(tid) < nhand)
        {
        const uint32_t slot = s_hand_base + static_cast<uint32_t>(tid);
        if (slot < bubble::kListCap) out_list[static_cast<size_t>(g) * bubble::kListCap + slot] = s_hand[tid];
        else atomicOr(error, 2u);
        }
    }

//! Bubble span pass: the nodes whose leaf range lies in one span of kFan sub-spans.
/*!
 * Workgroup g owns the leaves [g*span, (g+1)*span) and starts from the nodes handed up
 * by its sub-spans (their boxes are final and visible after the kernel boundary).  It
 * climbs with arrival counters in global memory: all climbs into a node come from the
 * same workgroup, so workgroup-scope ordering suffices.  A node whose parent leaves the
 * span is handed to the next span pass (out_list, or nowhere at the last level, whose
 * span covers every leaf).  flags[N-1] must be zero on entry.
 */
__global__ void __launch_bounds__(bubble::kTopBlock) lbvh_bubble_span(const LBVHData tree,
                                                                      const int* __restrict__ range_end,
                                                                      uint32_t* __restrict__ flags,
                                                                      const uint32_t* __restrict__ in_count,
                                                                      const int* __restrict__ in_list,
                                                                      uint32_t* __restrict__ out_count,
                                                                      int* __restrict__ out_list,
                                                                      uint32_t* __restrict__ error,
                                                                      const unsigned long long span)
    {
    const uint32_t g = blockIdx.x;
    const long long s0 = static_cast<long long>(g) * static_cast<long long>(span);
    const long long s1 = s0 + static_cast<long long>(span);
    const uint32_t count = min(in_count[g], bubble::kListCap);
    const int* list = in_list + static_cast<size_t>(g) * bubble::kListCap;
    for (uint32_t e = threadIdx.x; e < count; e += blockDim.x)
        {
        int node = list[e];
        float3 lo = tree.lo[node], hi = tree.hi[node];
        int p = tree.parent[node];
        // the parent's record (range end, children, parent); each level loads the next one
        // together with the sibling's box, so a level costs two round trips: its arrival
        // counter, then both loads
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
            if (static_cast<long long>(min(p, q)) < s0 || static_cast<long long>(max(p, q)) >= s1)
                {
                if (out_list)
                    {
                    const uint32_t og = g / bubble::kFan;
                    const uint32_t slot = atomicAdd(&out_count[og], 1u);
                    if (slot < bubble::kListCap) out_list[static_cast<size_t>(og) * bubble::kListCap + slot] = node;
                    else atomicOr(error, 2u);
                    }
                else
                    {
                    atomicOr(error, 4u);    // the last span covers every leaf: unreachable
                    }
                break;
                }
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
    }

//! Bottom-up boxes of a built hierarchy: the chunk pass, then span passes of kFan-fold
//! wider spans until one span covers every leaf.  work: bubble::list_words(n) words, and
//! flags: n-1 words, both zero on entry; error: a device word the passes OR into.
//! sorted_ready: tree.sorted_points already holds the point leaves' points in Morton order
//! (the hierarchy kernel's SortedPointGather); otherwise the chunk pass gathers and writes them.
template<class InsertOpT>
inline void launch_bubble(const LBVHData& tree, const InsertOpT& insert, const int* range_end, uint32_t* flags,
                          uint32_t* work, uint32_t* error, uint32_t n, hipStream_t stream, bool sorted_ready = false)
    {
    const unsigned int chunk = bubble::chunk_for(n);
    const unsigned int levels = bubble::span_levels(n, chunk);
    // per level: counters[groups] then lists[groups][kListCap]
    std::vector<uint32_t*> counts(levels + 1, nullptr);
    std::vector<int*> lists(levels + 1, nullptr);
    uint32_t* w = work;
    for (unsigned int l = 1; l <= levels; ++l)
        {
        const size_t g = bubble::groups(n, l, chunk);
        counts[l] = w;
        lists[l] = reinterpret_cast<int*>(w + g);
        w += g * (1 + bubble::kListCap);
        }
    uint32_t* out_count = counts.size() > 1 ? counts[1] : nullptr;
    int* out_list = lists.size() > 1 ? lists[1] : nullptr;
    if (chunk == bubble::kChunk)
        hipLaunchKernelGGL((lbvh_bubble_chunk<InsertOpT, bubble::kChunk, bubble::kBlock>), dim3(div_up(n, chunk)),
                           dim3(bubble::kBlock), 0, stream, tree, insert, range_end, out_count, out_list, error, n,
                           sorted_ready);
    else
        hipLaunchKernelGGL((lbvh_bubble_chunk<InsertOpT, bubble::kChunkLarge, bubble::kBlockLarge>), dim3(div_up(n, chunk)),
                           dim3(bubble::kBlockLarge), 0, stream, tree, insert, range_end, out_count, out_list, error, n,
                           sorted_ready);
    for (unsigned int l = 1; l <= levels; ++l)
        {
        const bool last = (l == levels);
        hipLaunchKernelGGL(lbvh_bubble_span, dim3(static_cast<unsigned int>(bubble::groups(n, l, chunk))), dim3(bubble::kTopBlock), 0,
                           stream, tree, range_end, flags, counts[l], lists[l], last ? nullptr : counts[l + 1],
                           last ? nullptr : lists[l + 1], error, bubble::span_leaves(l, chunk));
        }
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
