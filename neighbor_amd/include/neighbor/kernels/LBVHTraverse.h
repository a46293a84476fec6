// LBVHTraverse.h -- compression and stackless rope traversal kernels.
//
// Reference semantics kept (include/neighbor/kernels/LBVHTraverser.cuh):
//   lbvh_compress  node boxes quantized to 10 bits per axis against the root box,
//                  lower bins rounded down from the root's lower corner, upper bins
//                  rounded down from the root's upper corner, with the HIP-path
//                  directed rounding (:45-125).  The rope comes from the build.
//   lbvh_traverse  per query: image bits from the unquantized root box, then for
//                  each image in ascending order a rope walk from the root that
//                  descends left on overlap and follows the rope otherwise; a leaf
//                  overlap goes through refine and the output op (:155-252).
// Additions: a float rope record (no decode work in the hot loop), an optional
// query order (thread t serves query order[t]; with the LBVH's primitive order a
// self-query walks neighboring paths in neighboring lanes), and the image index
// handed to output ops that accept it.
#ifndef NEIGHBOR_KERNELS_LBVH_TRAVERSE_H_
#define NEIGHBOR_KERNELS_LBVH_TRAVERSE_H_

#include "../BoundingVolumes.h"
#include "../LBVHData.h"
#include "../LBVHTraverserData.h"

#include <cstddef>
#include <type_traits>
#include <utility>

namespace neighbor
{
namespace gpu
{
namespace detail
{
// prefer refine(ThreadData, Volume, prim) when the query op has it
template<class Q>
__device__ __forceinline__ auto refine_hit(const Q& q, const typename Q::ThreadData& t, const typename Q::Volume& v, int p, int)
    -> decltype(q.refine(t, v, p))
    {
    return q.refine(t, v, p);
    }
template<class Q>
__device__ __forceinline__ bool refine_hit(const Q& q, const typename Q::ThreadData& t, const typename Q::Volume&, int p, long)
    {
    return q.refine(t, p);
    }
// prefer process(ThreadData&, prim, image) when the output op has it
template<class O>
__device__ __forceinline__ auto emit_hit(const O& o, typename O::ThreadData& t, int p, unsigned int k, int)
    -> decltype(o.process(t, p, k))
    {
    o.process(t, p, k);
    }
template<class O>
__device__ __forceinline__ void emit_hit(const O& o, typename O::ThreadData& t, int p, unsigned int, long)
    {
    o.process(t, p);
    }

// output ops whose rows can overflow and be moved provide bool relocate(ThreadData&): true when
// the query's row was moved (the query is then walked again into its new row)
template<class O, class = void>
struct relocates : std::false_type {};
template<class O>
struct relocates<O, decltype(void(std::declval<const O&>().relocate(std::declval<typename O::ThreadData&>())))>
    : std::true_type {};

//! Root quantization scales (kernels/LBVHTraverser.cuh:64-74).
__device__ __forceinline__ float bin_scale(float lo, float hi)
    {
    float L = approx::fsub_ru(hi, lo);
    if (L <= 0.0f) L = 1.0f;
    return approx::fdiv_rd(1023.0f, L);
    }

__device__ __forceinline__ uint32_t quantize(float offset, float scale)
    {
    return approx::sat_f2u(floorf(approx::fmul_rd(offset, scale)));
    }

//! Decompressed (outward) box of a quantized record (kernels/LBVHTraverser.cuh:218-226).
__device__ __forceinline__ BoundingBox decode_box(uint32_t lo3, uint32_t hi3, const float3& rlo, const float3& rhi, const float3& bins)
    {
    BoundingBox b;
    b.lo.x = approx::fadd_rd(rlo.x, approx::fmul_rd(static_cast<float>((lo3 >> 20) & 0x3ffu), bins.x));
    b.lo.y = approx::fadd_rd(rlo.y, approx::fmul_rd(static_cast<float>((lo3 >> 10) & 0x3ffu), bins.y));
    b.lo.z = approx::fadd_rd(rlo.z, approx::fmul_rd(static_cast<float>(lo3 & 0x3ffu), bins.z));
    b.hi.x = approx::fsub_ru(rhi.x, approx::fmul_rd(static_cast<float>((hi3 >> 20) & 0x3ffu), bins.x));
    b.hi.y = approx::fsub_ru(rhi.y, approx::fmul_rd(static_cast<float>((hi3 >> 10) & 0x3ffu), bins.y));
    b.hi.z = approx::fsub_ru(rhi.z, approx::fmul_rd(static_cast<float>(hi3 & 0x3ffu), bins.z));
    return b;
    }
} // end namespace detail

namespace kernel
{

//! Quantized records for all nodes; node 0's thread also writes the root box and bin sizes.
template<class TransformOpT>
__global__ void __launch_bounds__(1024) lbvh_compress(const LBVHCompressedData ctree,
                                                      const TransformOpT transform,
                                                      const ConstLBVHData tree,
                                                      const uint32_t n_internal,
                                                      const uint32_t n_nodes)
    {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_nodes) return;
    const float3 rlo = tree.lo[tree.root];
    const float3 rhi = tree.hi[tree.root];
    const float3 scale = make_float3(detail::bin_scale(rlo.x, rhi.x), detail::bin_scale(rlo.y, rhi.y), detail::bin_scale(rlo.z, rhi.z));

    const float3 lo = tree.lo[idx];
    const float3 hi = tree.hi[idx];
    const uint32_t lo3 = (detail::quantize(approx::fsub_rd(lo.x, rlo.x), scale.x) << 20)
                       + (detail::quantize(approx::fsub_rd(lo.y, rlo.y), scale.y) << 10)
                       + detail::quantize(approx::fsub_rd(lo.z, rlo.z), scale.z);
    const uint32_t hi3 = (detail::quantize(approx::fsub_rd(rhi.x, hi.x), scale.x) << 20)
                       + (detail::quantize(approx::fsub_rd(rhi.y, hi.y), scale.y) << 10)
                       + detail::quantize(approx::fsub_rd(rhi.z, hi.z), scale.z);
    const int lf = (idx < n_internal) ? tree.left[idx] : static_cast<int>(~transform(tree.primitive[idx - n_internal]));
    ctree.data[idx] = make_int4(static_cast<int>(lo3), static_cast<int>(hi3), lf, tree.rope[idx]);

    if (idx == 0)
        {
        *ctree.lo = rlo;
        *ctree.hi = rhi;
        *ctree.bins = make_float3(approx::frcp_rd(scale.x), approx::frcp_rd(scale.y), approx::frcp_rd(scale.z));
        }
    }

//! Float rope records.  DECODED: store the decompressed quantized boxes (results
//! identical to the quantized walk); otherwise store the exact node boxes.
template<bool DECODED>
__global__ void __launch_bounds__(1024) lbvh_float_ropes(float4* __restrict__ out,
                                                         const LBVHCompressedData ctree,
                                                         const ConstLBVHData tree,
                                                         const uint32_t n_nodes)
    {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_nodes) return;
    const int4 r = ctree.data[idx];
    BoundingBox b;
    if (DECODED)
        b = detail::decode_box(static_cast<uint32_t>(r.x), static_cast<uint32_t>(r.y), tree.lo[tree.root], tree.hi[tree.root], *ctree.bins);
    else
        b = BoundingBox(tree.lo[idx], tree.hi[idx]);
    out[2 * idx] = make_float4(b.lo.x, b.lo.y, b.lo.z, __int_as_float(r.z));
    out[2 * idx + 1] = make_float4(b.hi.x, b.hi.y, b.hi.z, __int_as_float(r.w));
    }

//! Quantized rope traversal, one query per thread.
template<class OutputOpT, class QueryOpT, class TranslateOpT>
__global__ void __launch_bounds__(1024) lbvh_traverse_ropes(const OutputOpT out,
                                                            const LBVHCompressedData lbvh,
                                                            const QueryOpT query,
                                                            const TranslateOpT images,
                                                            const unsigned int* __restrict__ order,
                                                            const unsigned int query_budget)
    {
    const unsigned int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= query.size()) return;
    const unsigned int idx = order ? order[t] : t;

    const BoundingBox root_box(*lbvh.lo, *lbvh.hi);
    const float3 rlo = root_box.lo, rhi = root_box.hi;
    const float3 bins = *lbvh.bins;

    const typename QueryOpT::ThreadData qdata = query.setup(idx);
    typename OutputOpT::ThreadData result = out.setup(idx, qdata);

    uint32_t image_flags = 0;
    const unsigned int nimg = images.size() < 32u ? images.size() : 32u;
    for (unsigned int k = 0; k < nimg; ++k)
        if (query.overlap(query.get(qdata, images.get(k)), root_box)) image_flags |= 1u << k;

    for (int pass = 0; pass < 2; ++pass)
        {
        uint32_t flags = image_flags;
        while (flags)
            {
            const unsigned int k = __builtin_ctz(flags);
            flags &= flags - 1u;
            const typename QueryOpT::Volume v = query.get(qdata, images.get(k));
            int node = lbvh.root;
            // ropes and left children only move forward in preorder: a walk visits each node at most once
            for (int budget = 2 * static_cast<int>(query_budget); node != LBVHSentinel && budget > 0; --budget)
                {
                const int4 rec = lbvh.data[node];
                const BoundingBox box = detail::decode_box(static_cast<uint32_t>(rec.x), static_cast<uint32_t>(rec.y), rlo, rhi, bins);
                node = rec.w;
                if (query.overlap(v, box))
                    {
                    if (rec.z < 0)
                        {
                        const int prim = ~rec.z;
                        if (detail::refine_hit(query, qdata, v, prim, 0)) detail::emit_hit(out, result, prim, k, 0);
                        }
                    else
                        {
                        node = rec.z;
                        }
                    }
                }
            }
        if constexpr (!detail::relocates<OutputOpT>::value) break;
        else if (!out.relocate(result)) break;
        }
    out.finalize(result);
    }

//! Float rope traversal (no decode work per node), one query per thread.
template<class OutputOpT, class QueryOpT, class TranslateOpT>
__global__ void __launch_bounds__(1024) lbvh_traverse_float_ropes(const OutputOpT out,
                                                                  const LBVHFloatRopeData lbvh,
                                                                  const float3* __restrict__ root_lo,
                                                                  const float3* __restrict__ root_hi,
                                                                  const QueryOpT query,
                                                                  const TranslateOpT images,
                                                                  const unsigned int* __restrict__ order,
                                                                  const unsigned int query_budget)
    {
    const unsigned int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= query.size()) return;
    const unsigned int idx = order ? order[t] : t;
    const BoundingBox root_box(*root_lo, *root_hi);

    const typename QueryOpT::ThreadData qdata = query.setup(idx);
    typename OutputOpT::ThreadData result = out.setup(idx, qdata);

    uint32_t image_flags = 0;
    const unsigned int nimg = images.size() < 32u ? images.size() : 32u;
    for (unsigned int k = 0; k < nimg; ++k)
        if (query.overlap(query.get(qdata, images.get(k)), root_box)) image_flags |= 1u << k;

    for (int pass = 0; pass < 2; ++pass)
        {
        uint32_t flags = image_flags;
        while (flags)
            {
            const unsigned int k = __builtin_ctz(flags);
            flags &= flags - 1u;
            const typename QueryOpT::Volume v = query.get(qdata, images.get(k));
            int node = lbvh.root;
            for (int budget = 2 * static_cast<int>(query_budget); node != LBVHSentinel && budget > 0; --budget)
                {
                const float4 a = lbvh.data[2 * node];
                const float4 b = lbvh.data[2 * node + 1];
                const int lf = __float_as_int(a.w);
                node = __float_as_int(b.w);
                if (query.overlap(v, BoundingBox(make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z))))
                    {
                    if (lf < 0)
                        {
                        const int prim = ~lf;
                        if (detail::refine_hit(query, qdata, v, prim, 0)) detail::emit_hit(out, result, prim, k, 0);
                        }
                    else
                        {
                        node = lf;
                        }
                    }
                }
            }
        if constexpr (!detail::relocates<OutputOpT>::value) break;
        else if (!out.relocate(result)) break;
        }
    out.finalize(result);
    }

//! One past the last sorted leaf of a subtree, read off its rope.
/*!
 * A subtree ending at sorted leaf r has as rope the internal node r+1, the leaf
 * N-1+(r+1), or the sentinel when r = N-1 (see rope_after in LBVHBuild.h).
 */
__device__ __forceinline__ unsigned int bucket_end(int rope, unsigned int n_internal)
    {
    if (rope == LBVHSentinel) return n_internal + 1u;
    const unsigned int v = static_cast<unsigned int>(rope);
    return (v < n_internal) ? v : v - n_internal;
    }

//! Bucket rope records for the wave walk: float4(lo, bits(a)), float4(hi, bits(rope)).
/*!
 * a >= 0: internal node, descend to left child a.  a < 0: the node's subtree holds at
 * most `bucket` leaves and is handled as one bucket of the consecutive sorted leaves
 * [~a, bucket_end(rope)).  Nodes below a bucket are never reached (a bucket's rope
 * skips its subtree).  DECODED selects the decoded quantized boxes instead of the
 * exact ones.  TOP_ONLY (walks that start at the root only): the records of nodes below
 * a bucket are not written.
 */
template<bool DECODED, bool TOP_ONLY = false>
__global__ void __launch_bounds__(1024) lbvh_bucket_records(float4* __restrict__ out,
                                                            const LBVHCompressedData ctree,
                                                            const ConstLBVHData tree,
                                                            const uint32_t n_internal,
                                                            const uint32_t n_nodes,
                                                            const uint32_t bucket)
    {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_nodes) return;
    int first, count;
    if (idx < n_internal)
        {
        const int j = tree.range_end[idx];
        first = min(static_cast<int>(idx), j);
        count = abs(j - static_cast<int>(idx)) + 1;
        }
    else
        {
        first = static_cast<int>(idx - n_internal);
        count = 1;
        }
    if constexpr (TOP_ONLY)
        {
        const int p = tree.parent[idx];
        if (p != LBVHSentinel && abs(tree.range_end[p] - p) + 1 <= static_cast<int>(bucket)) return;   // below a bucket
        }
    const int a = (count <= static_cast<int>(bucket)) ? ~first : tree.left[idx];
    BoundingBox b;
    if (DECODED)
        {
        const int4 r = ctree.data[idx];
        b = detail::decode_box(static_cast<uint32_t>(r.x), static_cast<uint32_t>(r.y), tree.lo[tree.root], tree.hi[tree.root], *ctree.bins);
        }
    else
        {
        b = BoundingBox(tree.lo[idx], tree.hi[idx]);
        }
    out[2 * idx] = make_float4(b.lo.x, b.lo.y, b.lo.z, __int_as_float(a));
    out[2 * idx + 1] = make_float4(b.hi.x, b.hi.y, b.hi.z, __int_as_float(tree.rope[idx]));
    if (!DECODED && idx == 0 && ctree.lo && ctree.hi)
        {
        // the exact root box (the quantized compression, which otherwise writes it, is skipped)
        *ctree.lo = tree.lo[tree.root];
        *ctree.hi = tree.hi[tree.root];
        }
    }

//! Read-only data addressed through the constant address space: a wave-uniform
//! index then compiles to scalar loads (SMEM) instead of 64 vector lanes.
typedef float native_float4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) native_float4 uniform_native_float4;

__device__ __forceinline__ float4 uniform_load(const float4* base, unsigned int i)
    {
    const native_float4 v = ((uniform_native_float4*)(base))[i];
    return make_float4(v.x, v.y, v.z, v.w);
    }

namespace wave
{
//! Lanes of the wave (64 bits) whose predicate holds: the compare mask itself, with no
//! detour through a VGPR.
__device__ __forceinline__ unsigned long long vote(bool p)
    {
    return __builtin_amdgcn_ballot_w64(p);
    }

//! Number of XCDs (accelerator complex dies, each with its own L2) on gfx950.
constexpr unsigned int kXcds = 8;

//! Logical block index under the round-robin dispatch of workgroups over the XCDs.
/*!
 * Workgroup b runs on XCD b % 8.  Renumbering so that XCD x owns the contiguous
 * logical range [x*q + min(x, r), ...) (q = G/8, r = G%8) gives each XCD's L2 one
 * contiguous stretch of the (Morton-ordered) queries instead of every 8th block.
 */
__device__ __forceinline__ unsigned int xcd_block(unsigned int b, unsigned int G)
    {
    const unsigned int x = b % kXcds;
    const unsigned int j = b / kXcds;
    const unsigned int q = G / kXcds;
    const unsigned int r = G % kXcds;
    return x * q + (x < r ? x : r) + j;
    }

//! Queries of one 256-thread block of the compacted walks (tail-split launch map).
struct BlockSpan
    {
    unsigned int first;      //!< first query position of the block
    unsigned int lead;       //!< positions of the block's group of 256 before `first`
    unsigned int per_wave;   //!< queries per wave (lanes below it serve one each)
    bool live;               //!< false: a block past its XCD's share (exits at once)
    };

//! Map of workgroup b of a tail-split launch over G groups of 256 query positions.
/*!
 * XCD x owns the contiguous groups [x*q + min(x, r), ...) as in xcd_block.  Its first groups
 * run as one block each (64 queries per wave); its last T groups are each split into S
 * blocks of 256/S positions (64/S queries per wave, the other lanes idle).  A launch of
 * about two waves per wave slot ends with waves that started late and run with few others
 * beside them, at the latency of their own dependent loads: split, those last waves are
 * shorter.  The grid is 8 * (blocks of XCD 0); blocks past an XCD's own count are not live.
 * T = 0 is the plain map of xcd_block (grid G, every block live).
 */
__host__ __device__ __forceinline__ BlockSpan tail_block(const unsigned int b, const unsigned int G,
                                                         const unsigned int T, const unsigned int S)
    {
    const unsigned int x = b % kXcds;
    const unsigned int j = b / kXcds;
    const unsigned int q = G / kXcds;
    const unsigned int r = G % kXcds;
    const unsigned int gx = q + (x < r ? 1u : 0u);
    const unsigned int g0 = x * q + (x < r ? x : r);
    const unsigned int tb = gx < T ? gx : T;
    const unsigned int head = gx - tb;
    if (j < head) return BlockSpan{256u * (g0 + j), 0u, 64u, true};
    const unsigned int k = j - head;
    if (k >= S * tb) return BlockSpan{0u, 0u, 0u, false};
    const unsigned int lead = (k % S) * (256u / S);
    return BlockSpan{256u * (g0 + head + k / S) + lead, lead, 64u / S, true};
    }

//! Node test of the wave walk: float distance from the RN center to the box against a
//! threshold that is never tighter than the query's own test (so no hit is pruned).
__device__ __forceinline__ float box_distance2(const float3& c, const float4& lo, const float4& hi)
    {
    const float dx = fmaxf(fmaxf(lo.x - c.x, c.x - hi.x), 0.0f);
    const float dy = fmaxf(fmaxf(lo.y - c.y, c.y - hi.y), 0.0f);
    const float dz = fmaxf(fmaxf(lo.z - c.z, c.z - hi.z), 0.0f);
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    }

//! Exact query: ExactSphere's own box test is already this padded float test.
__device__ __forceinline__ float3 center(const ExactSphere& v) { return v.c; }
__device__ __forceinline__ float threshold(const ExactSphere& v) { return v.T; }

//! Reference query: the directed-rounding test satisfies dr2 >= d2 (1 - 15u) (u = 2^-24)
//! for the exact distance d2 from the float origin, and the RN float distance is at most
//! d2 (1 + 5u); Rsq (1 + 2^-18) therefore admits every box the reference test admits.
__device__ __forceinline__ float3 center(const BoundingSphere& v) { return v.origin; }
__device__ __forceinline__ float threshold(const BoundingSphere& v) { return approx::fmul_ru(v.Rsq, 1.0f + 0x1p-18f); }

//! 64 bytes (four float4) at a wave-uniform address: one scalar load.
typedef float native_float16 __attribute__((ext_vector_type(16), aligned(16)));
typedef const __attribute__((address_space(4))) native_float16 uniform_native_float16;

__device__ __forceinline__ void uniform_load4(const float4* base, unsigned int i, float4 out[4])
    {
    const native_float16 v = *((uniform_native_float16*)(base + i));
    out[0] = make_float4(v[0], v[1], v[2], v[3]);
    out[1] = make_float4(v[4], v[5], v[6], v[7]);
    out[2] = make_float4(v[8], v[9], v[10], v[11]);
    out[3] = make_float4(v[12], v[13], v[14], v[15]);
    }

__device__ __forceinline__ unsigned int uniform_load_u32(const unsigned int* base, unsigned int i)
    {
    return *((const __attribute__((address_space(4))) unsigned int*)(base + i));
    }

//! LDS staging of hits for pair output ops (kStaged): each lane keeps up to kStage
//! (neighbor, image) entries in LDS; at wave-uniform points the lanes holding more than a
//! threshold are written out one lane at a time by the whole wave, so each run of entries
//! of a row becomes one coalesced store instead of one scattered store per hit.
constexpr unsigned int kStage = 16;
constexpr unsigned int kStageStride = kStage + 1;   // odd word stride: conflict-free pushes

template<class O, class = void>
struct staged : std::false_type {};
template<class O>
struct staged<O, typename std::enable_if<O::kStaged>::type> : std::true_type {};

//! Output ops whose rows the wave kernels place before the walk (kPlacedRows: PlacedRowPairOp;
//! every thread of the 256-thread block takes part).
template<class O, class = void>
struct places_rows : std::false_type {};
template<class O>
struct places_rows<O, typename std::enable_if<O::kPlacedRows>::type> : std::true_type {};

template<class OutputOpT>
__device__ __forceinline__ void place_rows(const OutputOpT& out, const unsigned int t, const unsigned int idx, const bool valid,
                                           const unsigned int group = 0u, const unsigned int lead = 0u,
                                           const unsigned int* order = nullptr, const unsigned int n = 0u)
    {
    if constexpr (places_rows<OutputOpT>::value) out.place(t, idx, valid, group, lead, order, n);
    }

struct Stage
    {
    unsigned int* prim;      //!< this wave's 64 x kStageStride words
    unsigned char* img;      //!< this wave's 64 x kStage bytes
    unsigned int lane;
    unsigned int n;          //!< entries staged by this lane
    };

__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, unsigned int l)
    {
    const unsigned int lo = __builtin_amdgcn_readlane(static_cast<int>(v & 0xffffffffull), static_cast<int>(l));
    const unsigned int hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), static_cast<int>(l));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
    }

//! Write out the staged entries of the lanes holding more than `keep` (call from wave-uniform code).
/*!
 * Staged entries of a lane belong at [t.pos, t.pos + n) of its row; positions at or past
 * t.end are dropped (their hits stay counted in t.num_neigh).
 */
template<class OutputOpT>
__device__ __forceinline__ void flush(const OutputOpT& out, typename OutputOpT::ThreadData& t, Stage& s, unsigned int keep)
    {
    const bool mine = s.n > keep;
    unsigned long long m = vote(mine);
    if (!m) return;
    __builtin_amdgcn_wave_barrier();
    while (m)
        {
        const unsigned int l = static_cast<unsigned int>(__builtin_ctzll(m));
        m &= m - 1ull;
        const unsigned int nl = __builtin_amdgcn_readlane(static_cast<int>(s.n), static_cast<int>(l));
        const unsigned long long pl = readlane64(t.pos, l);
        const unsigned long long el = readlane64(t.end, l);
        const unsigned long long p = pl + s.lane;
        if (s.lane < nl && p < el)
            {
            out.neigh[p] = s.prim[l * kStageStride + s.lane];
            if (out.image) out.image[p] = s.img[l * kStage + s.lane];
            }
        }
    if (mine)
        {
        t.pos += s.n;
        s.n = 0;
        }
    __builtin_amdgcn_wave_barrier();
    }

template<class OutputOpT, class TransformOpT>
struct BucketContext
    {
    const OutputOpT& out;
    typename OutputOpT::ThreadData& result;
    const TransformOpT& transform;
    unsigned int image;
    unsigned long long hitm;     //!< lanes whose box test passed at this bucket (wave-uniform)
    Stage& stage;

    __device__ __forceinline__ bool hit() const { return __builtin_amdgcn_inverse_ballot_w64(hitm); }
    };

//! A hit of this lane: staged in LDS for pair ops, handed to the output op otherwise.
template<class OutputOpT, class TransformOpT>
__device__ __forceinline__ void emit(BucketContext<OutputOpT, TransformOpT>& ctx, unsigned int value)
    {
    if constexpr (staged<OutputOpT>::value)
        {
        Stage& s = ctx.stage;
        s.prim[s.lane * kStageStride + s.n] = value;
        s.img[s.lane * kStage + s.n] = static_cast<unsigned char>(ctx.image);
        ++s.n;
        ++ctx.result.num_neigh;
        }
    else
        {
        detail::emit_hit(ctx.out, ctx.result, static_cast<int>(value), ctx.image, 0);
        }
    }

//! Make room for `batch` more hits per lane before a batch of leaf tests.
template<class OutputOpT, class TransformOpT>
__device__ __forceinline__ void reserve(BucketContext<OutputOpT, TransformOpT>& ctx, unsigned int batch)
    {
    if constexpr (staged<OutputOpT>::value) flush(ctx.out, ctx.result, ctx.stage, kStage - batch);
    }

//! Leaf tests of a bucket, four leaves per scalar load.
/*!
 * Exact query (ExactSphereQueryOp): the FP64 definition, decided by the banded float
 * pre-test of ExactSphereQueryOp::refine_point.  Reference query (SphereQueryOp): its
 * own overlap test against the leaf's decoded quantized box (refine() is always true),
 * decided by the RN float distance outside a band: the directed-rounding distance lies
 * within d2 (1 +- 15u) of the exact distance d2 from the origin and the RN one within
 * d2 (1 +- 5u), so RN <= Rsq (1 - 2^-19) implies a reference hit and RN > Rsq (1 + 2^-19)
 * a miss (2^-19 = 32u).  In both cases the slow test is a divergent branch that the
 * wave skips when none of its lanes lands in the band.
 */
//! Hits of this lane among the lanes set in the wave-uniform mask m.
__device__ __forceinline__ bool lane_in(unsigned long long m)
    {
    return __builtin_amdgcn_inverse_ballot_w64(m);
    }

//! Exact leaf tests of M consecutive leaves, decided as lane masks: one float test per leaf,
//! one wave-uniform branch for the band refinements of the whole group, then the hits.
template<unsigned int M, class OutputOpT, class TransformOpT>
__device__ __forceinline__ void leaves_exact(BucketContext<OutputOpT, TransformOpT>& ctx, const ExactSphere& v, const float4* p,
                                             const unsigned long long hitmask)
    {
    unsigned long long in_m[M], band_m[M];
    unsigned long long any = 0ull;
    #pragma unroll
    for (unsigned int j = 0; j < M; ++j)
        {
        const float ex = v.c.x - p[j].x;
        const float ey = v.c.y - p[j].y;
        const float ez = v.c.z - p[j].z;
        const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
        in_m[j] = vote(d2 <= v.t_lo);
        band_m[j] = vote(d2 <= v.t_hi) & ~in_m[j] & hitmask;
        any |= band_m[j];
        }
    if (any)
        {
        #pragma unroll
        for (unsigned int j = 0; j < M; ++j)
            if (band_m[j])
                {
                bool r = false;
                if (lane_in(band_m[j])) r = ExactSphereQueryOp::refine_fp64(v, p[j]);
                in_m[j] |= vote(r);
                }
        }
    #pragma unroll
    for (unsigned int j = 0; j < M; ++j)
        {
        const unsigned long long h = in_m[j] & hitmask;
        if constexpr (std::is_same<OutputOpT, CountNeighborsOp>::value)
            {
            ctx.result.num_neigh += lane_in(h) ? 1u : 0u;
            }
        else
            {
            if (h && lane_in(h)) emit(ctx, ctx.transform(__float_as_uint(p[j].w)));
            }
        }
    }

template<class OutputOpT, class TransformOpT>
__device__ __forceinline__ void bucket(BucketContext<OutputOpT, TransformOpT>& ctx, const ExactSphereQueryOp&, const ExactSphere& v,
                                       const float4* points, const float4*, const unsigned int*, unsigned int, unsigned int first,
                                       unsigned int end)
    {
    const unsigned long long hitmask = ctx.hitm;
    unsigned int leaf = first;
    for (; leaf + 4u <= end; leaf += 4u)
        {
        reserve(ctx, 4u);
        float4 p[4];
        uniform_load4(points, leaf, p);
        leaves_exact<4>(ctx, v, p, hitmask);
        }
    if (leaf < end) reserve(ctx, 3u);
    for (; leaf < end; ++leaf)
        {
        const float4 p = uniform_load(points, leaf);
        leaves_exact<1>(ctx, v, &p, hitmask);
        }
    }

//! Compacted buckets for the exact query with the count or the pair output ops.
/*!
 * At a bucket only the lanes whose box test passed (a of 64) have work, so the wave
 * enumerates the a x n (query, leaf) pairs instead of n leaves for 64 lanes, two leaves
 * per lane: slot t is query rank r = t / n2 (r-th active lane, n2 = ceil(n / 2)) and
 * leaves 2 (t % n2) and the next, 128 pairs per round.  A query's slots sit in consecutive
 * lanes in leaf order, so its hits of a round are numbered by two lane-mask counts and
 * written as one contiguous run of its row; ranks, leaves and running counts are exchanged
 * through LDS, within the wave only.  Emission order per query is unchanged: bucket order,
 * then leaf order.
 */
//! Used for the pair ops (their hits leave as per-query runs with no LDS staging) and the
//! count op.  With two leaves per lane the compacted count takes 0.45 ms at N=10^6 (48 leaves
//! per bucket) against 0.59 ms for the dense bucket (every lane against every leaf); with one
//! leaf per lane it had taken 0.63-0.69 ms against 0.61.
template<class OutputOpT, class QueryOpT>
struct compactable
    : std::integral_constant<bool, std::is_same<QueryOpT, ExactSphereQueryOp>::value
                                       && (staged<OutputOpT>::value || std::is_same<OutputOpT, CountNeighborsOp>::value)>
    {};

//! Image bytes of the run [a, b) of one row: the image of one walk is the same for all of a
//! query's pairs found in it, so the compacted walks write the image array once per image
//! walk and row, as a run (16-byte stores in the aligned middle), not one byte per pair: the
//! bytes of a wave's rows are then written within a short time and leave L2 as whole lines.
__device__ __forceinline__ void image_run(unsigned char* __restrict__ img, unsigned long long a, const unsigned long long b,
                                          const unsigned int k)
    {
    const unsigned char kb = static_cast<unsigned char>(k);
    const unsigned int k4 = (k & 0xffu) * 0x01010101u;
    const uint4 k16 = make_uint4(k4, k4, k4, k4);
    // bytes up to a 4-byte boundary, dwords up to a 16-byte boundary, 16-byte stores, then the tail
    for (; a < b && (reinterpret_cast<uintptr_t>(img + a) & 3u); ++a) img[a] = kb;
    for (; a + 4u <= b && (reinterpret_cast<uintptr_t>(img + a) & 15u); a += 4u) *reinterpret_cast<unsigned int*>(img + a) = k4;
    for (; a + 16u <= b; a += 16u) *reinterpret_cast<uint4*>(img + a) = k16;
    for (; a + 4u <= b; a += 4u) *reinterpret_cast<unsigned int*>(img + a) = k4;
    for (; a < b; ++a) img[a] = kb;
    }

//! Per-wave LDS of the compacted buckets.  The per-rank records share one 16-byte stride, so
//! a round reaches all of a rank's data from one 32-bit offset (immediate offsets in the LDS
//! instructions, no per-array address arithmetic).
struct CompactWave
    {
    float4 leaf[64];                 //!< bucket leaves (x, y, z, bits(primitive))
    float4 qc[64];                   //!< per rank: float center, t_lo
    uint4 qx[64];                    //!< per rank: bits(t_hi), lane, unused, unused
    ulonglong2 row[64];              //!< per rank: (next row position, row end); the count op uses .x only
    };

//! Element of an array member of the wave's CompactWave at byte offset `off` (16-byte records).
template<class T>
__device__ __forceinline__ T& compact_at(CompactWave& cw, const unsigned int member, const unsigned int off)
    {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(&cw) + member + off);
    }

//! Leaves per lane in a round of the compacted buckets (a divisor of 64).  Measured on the
//! N=10^6 exact fill: 1 leaf 0.688 ms, 2 leaves 0.659 ms, 4 leaves 0.725 ms (fewer instructions
//! per candidate, but more tests of padding and of the last partial round of each bucket).
constexpr unsigned int kCompactLeaves = 2;

//! Set bits of the wave-uniform mask m in the lanes below this one.
__device__ __forceinline__ unsigned int count_below(unsigned long long m)
    {
    return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned int>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<unsigned int>(m), 0u));
    }

template<class OutputOpT, class TransformOpT>
__device__ __forceinline__ void bucket_compact(BucketContext<OutputOpT, TransformOpT>& ctx, CompactWave& cw,
                                               const ExactSphere& v, const float4* __restrict__ points,
                                               const unsigned int first, const unsigned int end)
    {
    constexpr bool kPairs = staged<OutputOpT>::value;
    const unsigned int lane = threadIdx.x & 63u;
    const unsigned long long A = ctx.hitm;
    const bool hit = ctx.hit();
    const unsigned int a = static_cast<unsigned int>(__popcll(A));
    const unsigned int n = end - first;                   // 1..64 leaves
    const unsigned int rank = count_below(A);
    unsigned long long pos0 = 0ull;
    // The last slot of a bucket whose size K does not divide pairs its leaves with padding
    // leaves, which only need x = inf: their squared distance is then inf or NaN (stale y, z),
    // which no comparison accepts.  Leaf j sits at (j % K) 64/K + j / K: the lanes of a round
    // read their k-th leaves from consecutive 16-byte records (no LDS bank conflicts).
    constexpr unsigned int K = kCompactLeaves;
    const unsigned int lpos = (lane % K) * (64u / K) + lane / K;
    if (lane < n) cw.leaf[lpos] = points[first + lane];
    else if (lane < (n + K - 1u) / K * K) cw.leaf[lpos].x = __builtin_huge_valf();
    if (hit)
        {
        cw.qc[rank] = make_float4(v.c.x, v.c.y, v.c.z, v.t_lo);
        cw.qx[rank] = make_uint4(__float_as_uint(v.t_hi), lane, 0u, 0u);
        unsigned long long row_end = 0ull;
        if constexpr (kPairs)
            {
            pos0 = ctx.result.pos;
            row_end = ctx.result.end;
            }
        cw.row[rank] = make_ulonglong2(pos0, row_end);
        }
    __builtin_amdgcn_wave_barrier();

    // slots of K consecutive leaves: slot (r, k) tests query rank r against leaves K k ..
    // K k + K - 1 (far points past n), so one pair index serves K leaf tests
    const unsigned int nK = (n + K - 1u) / K;
    const unsigned int total = a * nK;
    // hardware reciprocal (1 ulp): r = t / nK below needs (t + 1/2) * inv_n within 1/(2 nK) of
    // (t + 1/2) / nK, and the error here is below 64 * 2^-22 for t < 4096, nK <= 64
    const float inv_n = __builtin_amdgcn_rcpf(static_cast<float>(nK));
    for (unsigned int t0 = 0; t0 < total; t0 += 64u)
        {
        const unsigned int left = total - t0;
        const unsigned long long validm = left >= 64u ? ~0ull : ((1ull << left) - 1ull);
        const unsigned int t = t0 + lane;
        // r = t / nK exactly: (t + 1/2) / nK is at least 1/(2 nK) away from an integer and the
        // float product errs by far less for t < 4096; lanes past the end take the last slot
        const unsigned int tc = min(t, total - 1u);
        const unsigned int r = static_cast<unsigned int>((static_cast<float>(tc) + 0.5f) * inv_n);
        const unsigned int rn = __umul24(r, nK);
        const unsigned int k0 = tc - rn;                 // slot in the query's row: leaves K k0 ..
        const unsigned int ro = r << 4;                  // byte offset of rank r's records
        const float4 q = compact_at<float4>(cw, offsetof(CompactWave, qc), ro);
        const uint4 qx = compact_at<uint4>(cw, offsetof(CompactWave, qx), ro);
        const float T_hi = __uint_as_float(qx.x);
        float4 p[K];
        unsigned long long m[K], bm[K];
        unsigned long long band = 0ull;
        #pragma unroll
        for (unsigned int k = 0; k < K; ++k)
            {
            p[k] = compact_at<float4>(cw, offsetof(CompactWave, leaf), (k * (64u / K) + k0) << 4);
            const float ex = q.x - p[k].x;
            const float ey = q.y - p[k].y;
            const float ez = q.z - p[k].z;
            const float d = fmaf(ex, ex, fmaf(ey, ey, ez * ez));     // inf for a far point
            m[k] = vote(d <= q.w) & validm;
            bm[k] = vote(d <= T_hi) & ~m[k] & validm;
            band |= bm[k];
            }
        if (band)
            {
            // the FP64 definition for the pairs in the float band; the query's FP64 center
            // and R^2 come from the lane that owns it (every lane takes part in the shuffles)
            const int src = static_cast<int>(qx.y);
            const double cx = __shfl(v.cd.x, src, 64);
            const double cy = __shfl(v.cd.y, src, 64);
            const double cz = __shfl(v.cd.z, src, 64);
            const double R2 = __shfl(v.R2, src, 64);
            #pragma unroll
            for (unsigned int k = 0; k < K; ++k)
                {
                bool e = false;
                if (lane_in(bm[k]))
                    {
                    const double dx = cx - static_cast<double>(p[k].x);
                    const double dy = cy - static_cast<double>(p[k].y);
                    const double dz = cz - static_cast<double>(p[k].z);
                    e = (dx * dx + dy * dy) + dz * dz <= R2;
                    }
                m[k] |= vote(e);
                }
            }
        // lanes of query r in this round: [lo, hi); its hits below this lane are the hits of
        // the lanes from lo up (per-lane masks of the wave-uniform hit masks), K per lane
        const unsigned int lo = rn > t0 ? rn - t0 : 0u;
        const unsigned int hi = min(min(rn + nK - t0, 64u), left);
        const unsigned long long from_lo = ~0ull << lo;
        unsigned int before = 0u;
        #pragma unroll
        for (unsigned int k = 0; k < K; ++k) before += count_below(m[k] & from_lo);
        // one read of the rank's row record serves the hits' stores and the cursor update
        ulonglong2& rowref = compact_at<ulonglong2>(cw, offsetof(CompactWave, row), ro);
        const ulonglong2 row = rowref;
        unsigned long long at = row.x + before;
        #pragma unroll
        for (unsigned int k = 0; k < K; ++k)
            {
            const bool in = lane_in(m[k]);
            // the image bytes are written per image walk (image_run)
            if constexpr (kPairs)
                if (in && at < row.y) ctx.out.neigh[at] = ctx.transform(__float_as_uint(p[k].w));
            at += in ? 1u : 0u;
            }
        if (lane + 1u == hi) rowref.x = at;
        __builtin_amdgcn_wave_barrier();
        }
    if (hit)
        {
        const unsigned int c = static_cast<unsigned int>(cw.row[rank].x - pos0);
        ctx.result.num_neigh += c;
        if constexpr (kPairs) ctx.result.pos += c;
        }
    __builtin_amdgcn_wave_barrier();
    }

template<class OutputOpT, class TransformOpT>
__device__ __forceinline__ void leaf_reference(BucketContext<OutputOpT, TransformOpT>& ctx, const SphereQueryOp& q, const BoundingSphere& v,
                                               float r_lo, float r_hi, const float4& lo, const float4& hi, unsigned int prim)
    {
    const float d2 = box_distance2(v.origin, lo, hi);
    bool in = d2 <= r_lo;
    const bool hit = ctx.hit();
    const bool maybe = hit && !in && d2 <= r_hi;
    if (maybe) in = q.overlap(v, BoundingBox(make_float3(lo.x, lo.y, lo.z), make_float3(hi.x, hi.y, hi.z)));
    if (hit && in) emit(ctx, ctx.transform(prim));
    }

template<class OutputOpT, class TransformOpT>
__device__ __forceinline__ void bucket(BucketContext<OutputOpT, TransformOpT>& ctx, const SphereQueryOp& q, const BoundingSphere& v,
                                       const float4*, const float4* nodes, const unsigned int* primitive, unsigned int n_internal,
                                       unsigned int first, unsigned int end)
    {
    const float r_lo = v.Rsq * (1.0f - 0x1p-19f);
    const float r_hi = v.Rsq * (1.0f + 0x1p-19f);
    unsigned int leaf = first;
    for (; leaf + 2u <= end; leaf += 2u)
        {
        reserve(ctx, 2u);
        float4 b[4];    // lo, hi of two consecutive leaf records
        uniform_load4(nodes, 2u * (n_internal + leaf), b);
        leaf_reference(ctx, q, v, r_lo, r_hi, b[0], b[1], uniform_load_u32(primitive, leaf));
        leaf_reference(ctx, q, v, r_lo, r_hi, b[2], b[3], uniform_load_u32(primitive, leaf + 1u));
        }
    if (leaf < end) reserve(ctx, 1u);
    if (leaf < end)
        leaf_reference(ctx, q, v, r_lo, r_hi, uniform_load(nodes, 2u * (n_internal + leaf)),
                       uniform_load(nodes, 2u * (n_internal + leaf) + 1u), uniform_load_u32(primitive, leaf));
    }
} // end namespace wave

//! Walk LDS for 256-thread blocks (four waves): compacted-bucket storage for the exact
//! count / pair walks, staged pair writes for the other pair walks, nothing otherwise.
template<class OutputOpT, class QueryOpT = void>
struct StageStorage
    {
    static constexpr bool kCompact = wave::compactable<OutputOpT, QueryOpT>::value;
    static constexpr bool kOn = wave::staged<OutputOpT>::value && !kCompact;
    unsigned int prim[kOn ? 4u * 64u * wave::kStageStride : 1u];
    unsigned char img[kOn ? 4u * 64u * wave::kStage : 4u];
    wave::CompactWave compact[kCompact ? 4 : 1];

    __device__ __forceinline__ wave::Stage lane_stage()
        {
        const unsigned int w = kOn ? (threadIdx.x >> 6) : 0u;
        return wave::Stage{prim + w * 64u * wave::kStageStride, img + w * 64u * wave::kStage, threadIdx.x & 63u, 0u};
        }
    __device__ __forceinline__ wave::CompactWave* wave_compact() { return kCompact ? &compact[threadIdx.x >> 6] : nullptr; }
    };

namespace wave
{
//! The walk of one wave over the subtree [root, stop) for one query per lane.
/*!
 * Lane queries: idx (when valid).  The subtree is walked from root until its rope
 * (stop; the sentinel for a whole tree).  Image bits come from each query's own test
 * against root_box; the wave visits the images any lane needs, in ascending order.
 * At each node every active lane runs the padded float box test and the wave descends
 * when any lane overlaps.  Node indices are wave-uniform, so node records and bucket
 * leaves are scalar loads.  At a bucket every lane whose box test passed applies the
 * query's leaf test to each leaf and emits hits in leaf order, so each query's hits
 * come out image-major, then in Morton order, as in the per-thread walk.  The emitted
 * value is transform(primitive).  Output ops that relocate overflowed rows get a second
 * pass for the moved lanes.
 */
template<class OutputOpT, class QueryOpT, class TranslateOpT, class TransformOpT, bool STATS = false>
__device__ __forceinline__ void walk(const OutputOpT& out, const float4* __restrict__ nodes,
                                     const float4* __restrict__ points, const unsigned int* __restrict__ primitive,
                                     const BoundingBox& root_box, const int root, const int stop, const QueryOpT& query,
                                     const TranslateOpT& images, const TransformOpT& transform, const bool valid,
                                     const unsigned int idx, const unsigned int n_internal, const unsigned int n_nodes,
                                     Stage& stage, CompactWave* compact, unsigned int* visit_slot = nullptr)
    {
    constexpr bool kCompact = compactable<OutputOpT, QueryOpT>::value;
    unsigned int n_node = 0, n_bucket = 0, n_leaf = 0;     // wave-uniform walk statistics (STATS)
    constexpr bool kStagedOut = staged<OutputOpT>::value && !kCompact;
    const typename QueryOpT::ThreadData q = valid ? query.setup(idx) : typename QueryOpT::ThreadData();
    typename OutputOpT::ThreadData result = out.setup(idx, q);

    uint32_t image_flags = 0;
    const unsigned int nimg = images.size() < 32u ? images.size() : 32u;
    for (unsigned int k = 0; k < nimg; ++k)
        if (valid && query.overlap(query.get(q, images.get(k)), root_box)) image_flags |= 1u << k;

    // pass 1 walks every lane; pass 2 (output ops that relocate overflowed rows) only the moved lanes
    uint32_t flags = image_flags;
    for (int pass = 0; pass < 2; ++pass)
        {
        uint32_t wave_flags = 0;
        for (unsigned int k = 0; k < nimg; ++k)
            if (vote((flags >> k) & 1u)) wave_flags |= 1u << k;

        while (wave_flags)
            {
            const unsigned int k = __builtin_ctz(wave_flags);
            wave_flags &= wave_flags - 1u;
            const unsigned long long activem = vote((flags >> k) & 1u);
            const typename QueryOpT::Volume v = query.get(q, images.get(k));
            const float3 c = center(v);
            const float T = threshold(v);
            unsigned long long run0 = 0ull;
            if constexpr (kCompact && staged<OutputOpT>::value) run0 = result.pos;
            int node = root;
            int budget = 2 * static_cast<int>(n_nodes);
            for (; node != stop && budget > 0; --budget)
                {
                node = __builtin_amdgcn_readfirstlane(node);
                const float4 a = uniform_load(nodes, 2u * node);
                const float4 b = uniform_load(nodes, 2u * node + 1u);
                // every lane tests (no divergent branch); inactive lanes are masked out
                const unsigned long long hitm = vote(box_distance2(c, a, b) <= T) & activem;
                int next = __float_as_int(b.w);
                if (hitm)
                    {
                    const int lf = __float_as_int(a.w);
                    if (lf < 0)
                        {
                        if constexpr (STATS)
                            {
                            ++n_bucket;
                            n_leaf += bucket_end(next, n_internal) - ~static_cast<unsigned int>(lf);
                            }
                        BucketContext<OutputOpT, TransformOpT> ctx{out, result, transform, k, hitm, stage};
                        if constexpr (kCompact)
                            bucket_compact(ctx, *compact, v, points, ~static_cast<unsigned int>(lf), bucket_end(next, n_internal));
                        else
                            bucket(ctx, query, v, points, nodes, primitive, n_internal, ~static_cast<unsigned int>(lf),
                                   bucket_end(next, n_internal));
                        }
                    else
                        {
                        next = lf;
                        }
                    }
                node = next;
                }
            if constexpr (STATS) n_node += static_cast<unsigned int>(2 * static_cast<int>(n_nodes) - budget);   // records loaded
            if constexpr (kCompact && staged<OutputOpT>::value)
                if (out.image) image_run(out.image, run0, result.pos < result.end ? result.pos : result.end, k);
            }
        if constexpr (kStagedOut) flush(out, result, stage, 0u);
        if constexpr (!detail::relocates<OutputOpT>::value) break;
        else
            {
            const bool again = valid && out.relocate(result);
            if (!vote(again)) break;
            flags = again ? image_flags : 0u;
            }
        }
    if (STATS && visit_slot && (threadIdx.x & 63u) == 0u)
        {
        // node records loaded, buckets whose leaves were tested, and those leaves (per lane)
        visit_slot[0] = n_node;
        visit_slot[1] = n_bucket;
        visit_slot[2] = n_leaf;
        }
    if (valid) out.finalize(result);
    }
} // end namespace wave

//! One wave walks one rope path for 64 consecutive queries (thread t serves query
//! order[t]; Morton order for self-queries, so the 64 lanes of a wave are spatial
//! neighbors).  See wave::walk.  STATS: each wave writes its node records loaded, buckets
//! tested and leaves in them to visits[3w .. 3w+2] (w: the launched wave).
template<class OutputOpT, class QueryOpT, class TranslateOpT, class TransformOpT, bool STATS = false>
__global__ void __launch_bounds__(256) lbvh_traverse_wave(const OutputOpT out,
                                                          const float4* __restrict__ nodes,
                                                          const float4* __restrict__ points,
                                                          const unsigned int* __restrict__ primitive,
                                                          const float3* __restrict__ root_lo,
                                                          const float3* __restrict__ root_hi,
                                                          const int root,
                                                          const QueryOpT query,
                                                          const TranslateOpT images,
                                                          const TransformOpT transform,
                                                          const unsigned int* __restrict__ order,
                                                          const unsigned int n_nodes,
                                                          unsigned int* __restrict__ visits)
    {
    __shared__ StageStorage<OutputOpT, QueryOpT> lds;
    wave::Stage stage = lds.lane_stage();
    const unsigned int t = wave::xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const unsigned int n_internal = (n_nodes - 1u) / 2u;
    const bool valid = t < query.size();
    const unsigned int idx = valid ? (order ? order[t] : t) : 0u;
    const BoundingBox root_box(*root_lo, *root_hi);
    wave::place_rows(out, t, idx, valid);
    wave::walk<OutputOpT, QueryOpT, TranslateOpT, TransformOpT, STATS>(
        out, nodes, points, primitive, root_box, root, LBVHSentinel, query, images, transform, valid, idx, n_internal,
        n_nodes, stage, lds.wave_compact(), STATS ? visits + 3u * (t >> 6) : nullptr);
    }

//! lbvh_traverse_wave for the compacted exact counts and pair fills (wave::compactable ops),
//! compiled for eight waves per SIMD: 64 VGPRs with a few spilled instead of 72 at seven waves
//! (the fill of the N=10^6 exact step measured 0.74 against 0.78 ms).  The reference-query
//! walks and the walks with a visit log keep the kernel above.  Blocks map to queries by
//! wave::tail_block (tail_groups = 0: the plain XCD map).
template<class OutputOpT, class QueryOpT, class TranslateOpT, class TransformOpT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
lbvh_traverse_wave_pairs(const OutputOpT out,
                         const float4* __restrict__ nodes,
                         const float4* __restrict__ points,
                         const unsigned int* __restrict__ primitive,
                         const float3* __restrict__ root_lo,
                         const float3* __restrict__ root_hi,
                         const int root,
                         const QueryOpT query,
                         const TranslateOpT images,
                         const TransformOpT transform,
                         const unsigned int* __restrict__ order,
                         const unsigned int n_nodes,
                         const unsigned int tail_groups,
                         const unsigned int tail_split)
    {
    __shared__ StageStorage<OutputOpT, QueryOpT> lds;
    const unsigned int n = query.size();
    const wave::BlockSpan span = wave::tail_block(blockIdx.x, (n + 255u) / 256u, tail_groups, tail_split);
    if (!span.live) return;
    wave::Stage stage = lds.lane_stage();
    const unsigned int lane = threadIdx.x & 63u;
    const unsigned int t = span.first + (threadIdx.x >> 6) * span.per_wave + lane;
    const unsigned int n_internal = (n_nodes - 1u) / 2u;
    const bool valid = lane < span.per_wave && t < n;
    const unsigned int idx = valid ? (order ? order[t] : t) : 0u;
    const BoundingBox root_box(*root_lo, *root_hi);
    wave::place_rows(out, t, idx, valid, span.first - span.lead, span.lead, order, n);
    wave::walk<OutputOpT, QueryOpT, TranslateOpT, TransformOpT, false>(
        out, nodes, points, primitive, root_box, root, LBVHSentinel, query, images, transform, valid, idx, n_internal,
        n_nodes, stage, lds.wave_compact(), nullptr);
    }

} // end namespace kernel
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_LBVH_TRAVERSE_H_
