// QueryOps.h -- query volumes for LBVH traversal.
//
// Concept (reference include/neighbor/QueryOps.h:15-31):
//   typedef ... ThreadData;  typedef ... Volume;
//   __device__ ThreadData setup(unsigned int idx) const;
//   __device__ Volume get(const ThreadData&, const Translate::type& image) const;
//   __device__ bool overlap(const Volume&, const BoundingBox&) const;
//   __device__ bool refine(const ThreadData&, int primitive) const;
//   __host__ __device__ unsigned int size() const;
// Extension used by this library: a query op may instead provide
//   __device__ bool refine(const ThreadData&, const Volume&, int primitive) const;
// which the traversal prefers when present (it needs the translated volume).
#ifndef NEIGHBOR_QUERY_OPS_H_
#define NEIGHBOR_QUERY_OPS_H_

#include "BoundingVolumes.h"

namespace neighbor
{

//! Sphere queries (x,y,z,R) translated in float (reference QueryOps.h:32-120).
struct SphereQueryOp
    {
    SphereQueryOp(const float4* spheres_, unsigned int N_) : spheres(spheres_), N(N_) {}

    typedef float4 ThreadData;
    typedef BoundingSphere Volume;

    __device__ __forceinline__ ThreadData setup(const unsigned int idx) const { return spheres[idx]; }

    __device__ __forceinline__ Volume get(const ThreadData& q, const float3& image) const
        {
        return BoundingSphere(make_float3(q.x + image.x, q.y + image.y, q.z + image.z), q.w);
        }

    __device__ __forceinline__ bool overlap(const Volume& v, const BoundingBox& box) const { return v.overlap(box); }

    //! No narrow phase: every primitive whose leaf box overlaps is reported (QueryOps.h:104-107).
    __device__ __forceinline__ bool refine(const ThreadData&, const int) const { return true; }

    NEIGHBOR_HD unsigned int size() const { return N; }

    const float4* spheres;
    const unsigned int N;
    };

//! Traversal volume of ExactSphereQueryOp.
/*!
 * The tree test uses a float center c and a squared threshold T padded so that
 * every point p with  ((double)q + t - (double)p)^2 <= R^2  (FP64, no contraction)
 * passes the test for any box containing p:
 *   FP64 inclusion => |c64 - p| <= R (1 + 2^-50) + 2^-45 |c64|;
 *   |c - c64| <= 2^-24 |c64| per axis; the RN box distance below overestimates by
 *   at most a factor (1 + 5*2^-24) in the square;
 *   Rpad = R (1 + 2^-20) + 2^-21 max|c64| and T = Rpad^2 (1 + 2^-19), rounded up,
 *   dominate all of these.
 */
struct ExactSphere
    {
    float3 c;       //!< center rounded to float
    float T;        //!< conservative squared threshold for box tests
    float t_lo;     //!< float squared distances <= t_lo are FP64 hits
    float t_hi;     //!< float squared distances > t_hi are FP64 misses
    double3 cd;     //!< FP64 center q + t
    double R2;      //!< R*R in FP64

    NEIGHBOR_HD float distanceSq(const BoundingBox& b) const
        {
        const float dx = fmaxf(fmaxf(b.lo.x - c.x, c.x - b.hi.x), 0.0f);
        const float dy = fmaxf(fmaxf(b.lo.y - c.y, c.y - b.hi.y), 0.0f);
        const float dz = fmaxf(fmaxf(b.lo.z - c.z, c.z - b.hi.z), 0.0f);
        return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        }

    NEIGHBOR_HD bool overlap(const BoundingBox& b) const { return distanceSq(b) <= T; }
    };

//! Exact sphere queries: FP64 translation and FP64 narrow phase.
/*!
 * (i, j, k) is reported iff  d2 <= R_i^2  with  d = ((double)q_i + t_k) - (double)x_j
 * per axis and  d2 = (dx*dx + dy*dy) + dz*dz  in IEEE double.  The tree is only
 * used to find candidates (ExactSphere is conservative); every candidate is
 * refined with the FP64 test.  \a points is indexed by the value cached in the
 * leaf records (the primitive index under NullTransformOp); its records are float3, or
 * `stride` floats each (x, y, z first).
 */
struct ExactSphereQueryOp
    {
    ExactSphereQueryOp(const float4* spheres_, unsigned int N_, const float3* points_)
        : spheres(spheres_), points(reinterpret_cast<const float*>(points_)), stride(3u), N(N_) {}
    //! Points given as records of `stride` floats, x, y, z first (4: HOOMD's postype layout).
    ExactSphereQueryOp(const float4* spheres_, unsigned int N_, const float* points_, unsigned int stride_)
        : spheres(spheres_), points(points_), stride(stride_), N(N_) {}

    typedef float4 ThreadData;
    typedef ExactSphere Volume;

    __device__ __forceinline__ ThreadData setup(const unsigned int idx) const { return spheres[idx]; }

    NEIGHBOR_HD static Volume make(const ThreadData& q, const double3& t)
        {
        Volume v;
        v.cd = make_double3(static_cast<double>(q.x) + t.x, static_cast<double>(q.y) + t.y, static_cast<double>(q.z) + t.z);
        v.c = make_float3(static_cast<float>(v.cd.x), static_cast<float>(v.cd.y), static_cast<float>(v.cd.z));
        const double R = static_cast<double>(q.w);
        v.R2 = R * R;
        const double cmax = fmax(fmax(fabs(v.cd.x), fabs(v.cd.y)), fabs(v.cd.z));
        const double Rpad = R * (1.0 + 0x1p-20) + cmax * 0x1p-21;
        v.T = approx::double2float_ru(Rpad * Rpad * (1.0 + 0x1p-19));
        // band of the float pre-test (see refine_point): 8u(|c| + R)R + 8uR^2, u = 2^-24
        const double cnorm = sqrt(v.cd.x * v.cd.x + v.cd.y * v.cd.y + v.cd.z * v.cd.z);
        const double band = 0x1p-21 * ((cnorm + R) * R + R * R);
        v.t_lo = approx::double2float_rd(v.R2 - band);
        v.t_hi = approx::double2float_ru(v.R2 + band);
        return v;
        }

    __device__ __forceinline__ Volume get(const ThreadData& q, const double3& image) const { return make(q, image); }

    __device__ __forceinline__ Volume get(const ThreadData& q, const float3& image) const
        {
        return make(q, make_double3(image.x, image.y, image.z));
        }

    __device__ __forceinline__ bool overlap(const Volume& v, const BoundingBox& box) const { return v.overlap(box); }

    __device__ __forceinline__ bool refine(const ThreadData&, const Volume& v, const int primitive) const
        {
        const float* p = points + static_cast<size_t>(stride) * static_cast<unsigned int>(primitive);
        return refine_point(v, make_float4(p[0], p[1], p[2], 0.f));
        }

    //! The FP64 test, by definition.
    NEIGHBOR_HD static bool refine_fp64(const Volume& v, const float4& p)
        {
        const double dx = v.cd.x - static_cast<double>(p.x);
        const double dy = v.cd.y - static_cast<double>(p.y);
        const double dz = v.cd.z - static_cast<double>(p.z);
        return (dx * dx + dy * dy) + dz * dz <= v.R2;
        }

    //! The FP64 test against a point given by value (x, y, z; w unused), decided in float
    //! outside a thin band.
    /*!
     * With u = 2^-24, e = fl(c - p) per axis (c the float center) differs from the real
     * c64 - p by at most u(|c64| + |c64 - p|)(1 + u), and the float sum of squares (one
     * product and two fused multiply-adds) adds a relative 3u.  A float value
     * <= R^2 - band then implies an FP64 hit and one > R^2 + band an FP64 miss, for
     * band = 8u(|c64| + R)R + 8uR^2 (the FP64 test's own error, ~5*2^-53 relative, is
     * far inside).  Only values in the band pay for FP64.
     */
    NEIGHBOR_HD static bool refine_point(const Volume& v, const float4& p)
        {
        const float ex = v.c.x - p.x;
        const float ey = v.c.y - p.y;
        const float ez = v.c.z - p.z;
        const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
        if (d2 <= v.t_lo) return true;
        if (d2 > v.t_hi) return false;
        return refine_fp64(v, p);
        }

    NEIGHBOR_HD unsigned int size() const { return N; }

    const float4* spheres;
    const float* points;
    const unsigned int stride;
    const unsigned int N;
    };

} // end namespace neighbor

#endif // NEIGHBOR_QUERY_OPS_H_
