// BoundingVolumes.h -- axis-aligned box and sphere used by the LBVH.
//
// Semantics follow the reference (include/neighbor/BoundingVolumes.h):
//   BoundingBox    : float bounds; double ctor rounds lo down / hi up (:54-58);
//                    getCenter = 0.5f*(lo+hi) (:64-72); inclusive overlap (:83-88).
//   BoundingSphere : float ctor Rsq = fmul_ru(r,r) (:119-123); double ctor takes
//                    origin = lo rounding of o and pads R by the widest rounding
//                    interval (:135-147); overlap = directed-rounded squared
//                    distance from origin to the clamped point <= Rsq (:163-171).
#ifndef NEIGHBOR_BOUNDING_VOLUMES_H_
#define NEIGHBOR_BOUNDING_VOLUMES_H_

#include "ApproximateMath.h"

namespace neighbor
{

struct BoundingBox
    {
    NEIGHBOR_HD BoundingBox() {}

    NEIGHBOR_HD BoundingBox(const float3& lo_, const float3& hi_) : lo(lo_), hi(hi_) {}

    NEIGHBOR_HD BoundingBox(const double3& lo_, const double3& hi_)
        {
        lo.x = approx::double2float_rd(lo_.x);
        lo.y = approx::double2float_rd(lo_.y);
        lo.z = approx::double2float_rd(lo_.z);
        hi.x = approx::double2float_ru(hi_.x);
        hi.y = approx::double2float_ru(hi_.y);
        hi.z = approx::double2float_ru(hi_.z);
        }

    NEIGHBOR_HD float3 getCenter() const
        {
        return make_float3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
        }

    NEIGHBOR_HD bool overlap(const BoundingBox& b) const
        {
        const bool apart = (hi.x < b.lo.x) | (b.hi.x < lo.x) |
                           (hi.y < b.lo.y) | (b.hi.y < lo.y) |
                           (hi.z < b.lo.z) | (b.hi.z < lo.z);
        return !apart;
        }

    float3 lo;
    float3 hi;
    };

struct BoundingSphere
    {
    NEIGHBOR_HD BoundingSphere() {}

    NEIGHBOR_HD BoundingSphere(const float3& o, const float r)
        : origin(o), Rsq(approx::fmul_ru(r, r))
        {}

    NEIGHBOR_HD BoundingSphere(const double3& o, const double r)
        {
        const float3 down = make_float3(approx::double2float_rd(o.x), approx::double2float_rd(o.y), approx::double2float_rd(o.z));
        const float3 up = make_float3(approx::double2float_ru(o.x), approx::double2float_ru(o.y), approx::double2float_ru(o.z));
        const float wx = approx::fsub_ru(up.x, down.x);
        const float wy = approx::fsub_ru(up.y, down.y);
        const float wz = approx::fsub_ru(up.z, down.z);
        const float pad = fmaxf(fmaxf(wx, wy), wz);
        const float R = approx::fadd_ru(approx::double2float_ru(r), pad);
        origin = down;
        Rsq = approx::fmul_ru(R, R);
        }

    //! Squared distance (rounded as the reference does) from the origin to the box.
    NEIGHBOR_HD float distanceSq(const BoundingBox& b) const
        {
        const float dx = approx::fsub_rd(fminf(fmaxf(origin.x, b.lo.x), b.hi.x), origin.x);
        const float dy = approx::fsub_rd(fminf(fmaxf(origin.y, b.lo.y), b.hi.y), origin.y);
        const float dz = approx::fsub_rd(fminf(fmaxf(origin.z, b.lo.z), b.hi.z), origin.z);
        return approx::fmaf_rd(dx, dx, approx::fmaf_rd(dy, dy, approx::fmul_rd(dz, dz)));
        }

    NEIGHBOR_HD bool overlap(const BoundingBox& b) const
        {
        return distanceSq(b) <= Rsq;
        }

    float3 origin;
    float Rsq;
    };

} // end namespace neighbor

#endif // NEIGHBOR_BOUNDING_VOLUMES_H_
