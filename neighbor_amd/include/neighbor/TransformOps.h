// TransformOps.h -- value cached in a leaf record in place of the primitive index.
//
// Concept (reference include/neighbor/TransformOps.h:25-64):
//   __host__ __device__ unsigned int operator()(unsigned int primitive) const;
#ifndef NEIGHBOR_TRANSFORM_OPS_H_
#define NEIGHBOR_TRANSFORM_OPS_H_

#include "ApproximateMath.h"

namespace neighbor
{

//! Cache map[primitive].
struct MapTransformOp
    {
    MapTransformOp(const unsigned int* map_) : map(map_) {}
    NEIGHBOR_HD unsigned int operator()(unsigned int primitive) const { return map[primitive]; }
    const unsigned int* map;
    };

//! Cache the primitive index itself.
struct NullTransformOp
    {
    NEIGHBOR_HD unsigned int operator()(unsigned int primitive) const { return primitive; }
    };

} // end namespace neighbor

#endif // NEIGHBOR_TRANSFORM_OPS_H_
