// TranslateOps.h -- image vectors applied to each query volume.
//
// Concept (reference include/neighbor/TranslateOps.h:19-91):
//   typedef ... type;                          // vector type handed to QueryOp::get
//   __device__ type get(unsigned int k) const; // k-th image
//   __host__ __device__ unsigned int size() const;
// At most 32 images are traversed per call (the traversal keeps one bit per image).
#ifndef NEIGHBOR_TRANSLATE_OPS_H_
#define NEIGHBOR_TRANSLATE_OPS_H_

#include "ApproximateMath.h"

namespace neighbor
{

//! Only the self image (zero vector).
struct SelfOp
    {
    typedef float3 type;
    SelfOp() {}
    __device__ __forceinline__ float3 get(const unsigned int) const { return make_float3(0.f, 0.f, 0.f); }
    NEIGHBOR_HD unsigned int size() const { return 1; }
    };

//! Images read from an array, in array order.
template<typename Real3>
struct ImageListOp
    {
    typedef Real3 type;
    ImageListOp() : images(nullptr), N(0) {}
    ImageListOp(const Real3* images_, unsigned int N_) : images(images_), N(N_) {}
    __device__ __forceinline__ Real3 get(const unsigned int k) const { return images[k]; }
    NEIGHBOR_HD unsigned int size() const { return N; }
    const Real3* images;
    const unsigned int N;
    };

} // end namespace neighbor

#endif // NEIGHBOR_TRANSLATE_OPS_H_
