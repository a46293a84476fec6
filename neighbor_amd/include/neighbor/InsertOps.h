// InsertOps.h -- turn primitives into leaf boxes for the LBVH build.
//
// Concept (reference include/neighbor/InsertOps.h:18-94): an insert op is a
// trivially copyable struct passed by value to kernels with
//   __device__ BoundingBox get(unsigned int idx) const;   // leaf box of primitive idx
//   __host__ __device__ unsigned int size() const;        // number of primitives
#ifndef NEIGHBOR_INSERT_OPS_H_
#define NEIGHBOR_INSERT_OPS_H_

#include "BoundingVolumes.h"

namespace neighbor
{

//! Points: the leaf box of a point is the point itself (InsertOps.h:35-41).
struct PointInsertOp
    {
    PointInsertOp(const float3* points_, unsigned int N_) : points(points_), N(N_) {}

    __device__ __forceinline__ BoundingBox get(const unsigned int idx) const
        {
        const float3 p = points[idx];
        return BoundingBox(p, p);
        }

    NEIGHBOR_HD unsigned int size() const { return N; }

    const float3* points;
    const unsigned int N;
    };

//! Spheres of one constant radius: leaf box (p-r, p+r) in float (InsertOps.h:75-82).
struct SphereInsertOp
    {
    SphereInsertOp(const float3* points_, const float r_, unsigned int N_) : points(points_), r(r_), N(N_) {}

    __device__ __forceinline__ BoundingBox get(const unsigned int idx) const
        {
        const float3 p = points[idx];
        return BoundingBox(make_float3(p.x - r, p.y - r, p.z - r), make_float3(p.x + r, p.y + r, p.z + r));
        }

    NEIGHBOR_HD unsigned int size() const { return N; }

    const float3* points;
    const float r;
    const unsigned int N;
    };

//! Points or constant-radius spheres read in place from 16-byte particle records (x, y, z, w),
//! HOOMD's postype layout (the reference benchmark's ParticleInsertOp reads the same records,
//! benchmark/lbvh_benchmark.cuh:25-60), optionally through an index list: primitive i is record
//! index[i] (index null: record i).  r < 0: point leaves as PointInsertOp; r >= 0: the leaf box
//! (p - r, p + r) as SphereInsertOp.  A per-type tree reads its type's records where they lie,
//! with no gathered copy.  (No reference counterpart in the library itself.)
struct PostypeInsertOp
    {
    PostypeInsertOp(const float4* postype_, const unsigned int* index_, const float r_, unsigned int N_)
        : postype(postype_), index(index_), r(r_), N(N_) {}

    NEIGHBOR_HD float3 point(const unsigned int idx) const
        {
        const float4 p = postype[index ? index[idx] : idx];
        return make_float3(p.x, p.y, p.z);
        }

    __device__ __forceinline__ BoundingBox get(const unsigned int idx) const
        {
        const float3 p = point(idx);
        if (r < 0.0f) return BoundingBox(p, p);
        return BoundingBox(make_float3(p.x - r, p.y - r, p.z - r), make_float3(p.x + r, p.y + r, p.z + r));
        }

    NEIGHBOR_HD unsigned int size() const { return N; }

    const float4* postype;
    const unsigned int* index;
    const float r;
    const unsigned int N;
    };

} // end namespace neighbor

#endif // NEIGHBOR_INSERT_OPS_H_
