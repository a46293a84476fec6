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

} // end namespace kernel
} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_LBVH_TRAVERSE_H_
