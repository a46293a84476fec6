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

    uint32_t flags = 0;
    const unsigned int nimg = images.size() < 32u ? images.size() : 32u;
    for (unsigned int k = 0; k < nimg; ++k)
        if (query.overlap(query.get(qdata, images.get(k)), root_box)) flags |= 1u << k;

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

    uint32_t flags = 0;
    const unsigned int nimg = images.size() < 32u ? images.size() : 32u;
    for (unsigned int k = 0; k < nimg; ++k)
        if (query.overlap(query.get(qdata, images.get(k)), root_box)) flags |= 1u << k;

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
 * exact ones.
 */
template<bool DECODED>
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

//! Exact queries, one wave walking one rope path for 64 consecutive queries.
/*!
 * Thread t serves query order[t] (Morton order for self-queries, so the 64 lanes of
 * a wave are spatial neighbors).  The wave visits the images any lane needs, in
 * ascending order; at each node every active lane tests its own volume and the
 * wave descends when any lane overlaps.  Node indices are wave-uniform, so node
 * records and bucket points are scalar loads.  At a bucket every lane whose box
 * test passed refines each point with the FP64 test and emits hits in leaf order,
 * so each query's hits come out image-major, then in Morton order, as in the
 * per-thread walk.
 */
template<class OutputOpT, class TranslateOpT, class TransformOpT>
__global__ void __launch_bounds__(256) lbvh_traverse_wave_exact(const OutputOpT out,
                                                                const float4* __restrict__ nodes,
                                                                const float4* __restrict__ points,
                                                                const float3* __restrict__ root_lo,
                                                                const float3* __restrict__ root_hi,
                                                                const int root,
                                                                const ExactSphereQueryOp query,
                                                                const TranslateOpT images,
                                                                const TransformOpT transform,
                                                                const unsigned int* __restrict__ order,
                                                                const unsigned int n_nodes)
    {
    const unsigned int t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int n_internal = (n_nodes - 1u) / 2u;
    const bool valid = t < query.size();
    const unsigned int idx = valid ? (order ? order[t] : t) : 0u;
    const float4 q = valid ? query.setup(idx) : make_float4(0.f, 0.f, 0.f, 0.f);
    typename OutputOpT::ThreadData result = out.setup(idx, q);

    const BoundingBox root_box(*root_lo, *root_hi);
    uint32_t flags = 0;
    const unsigned int nimg = images.size() < 32u ? images.size() : 32u;
    for (unsigned int k = 0; k < nimg; ++k)
        if (valid && query.overlap(query.get(q, images.get(k)), root_box)) flags |= 1u << k;
    uint32_t wave_flags = 0;
    for (unsigned int k = 0; k < nimg; ++k)
        if (__ballot((flags >> k) & 1u)) wave_flags |= 1u << k;

    while (wave_flags)
        {
        const unsigned int k = __builtin_ctz(wave_flags);
        wave_flags &= wave_flags - 1u;
        const bool active = (flags >> k) & 1u;
        const ExactSphere v = query.get(q, images.get(k));
        int node = root;
        for (int budget = 2 * static_cast<int>(n_nodes); node != LBVHSentinel && budget > 0; --budget)
            {
            node = __builtin_amdgcn_readfirstlane(node);
            const float4 a = uniform_load(nodes, 2u * node);
            const float4 b = uniform_load(nodes, 2u * node + 1u);
            const bool hit = active && v.overlap(BoundingBox(make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z)));
            int next = __float_as_int(b.w);
            if (__ballot(hit))
                {
                const int lf = __float_as_int(a.w);
                if (lf < 0)
                    {
                    const unsigned int first = ~static_cast<unsigned int>(lf);
                    const unsigned int end = bucket_end(next, n_internal);
                    for (unsigned int c = first; c < end; ++c)
                        {
                        const float4 p = uniform_load(points, c);
                        if (hit && query.refine_point(v, p))
                            detail::emit_hit(out, result, static_cast<int>(transform(__float_as_uint(p.w))), k, 0);
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
    if (valid) out.finalize(result);
    }

} // end namespace kernel
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_LBVH_TRAVERSE_H_
