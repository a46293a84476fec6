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

} // end namespace neighbor

#endif // NEIGHBOR_INSERT_OPS_H_
