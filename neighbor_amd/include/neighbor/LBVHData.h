// LBVHData.h -- raw views of an LBVH for kernels.
//
// Layout (reference include/neighbor/LBVHData.h:14-43, LBVH.h:41-42): internal
// nodes occupy [0, N-1), leaves [N-1, 2N-1); the root is node 0 and its parent is
// LBVHSentinel.  Leaf N-1+i holds the primitive primitive[i] (i-th in Morton order).
// This library additionally keeps the rope of every node (the node visited next
// when a traversal skips the node's subtree; LBVHSentinel ends the walk) and, for
// internal node i, the far end j of its leaf range (the range is [min(i,j), max(i,j)]).
#ifndef NEIGHBOR_LBVH_DATA_H_
#define NEIGHBOR_LBVH_DATA_H_

#include <hip/hip_runtime.h>

namespace neighbor
{

const int LBVHSentinel = 0x7fffffff;

struct LBVHData
    {
    int* parent;
    int* left;
    int* right;
    unsigned int* primitive;
    float3* lo;
    float3* hi;
    int root;
    int* rope;
    int* range_end;
    float4* sorted_points;   //!< (x, y, z, bits(primitive)) of leaf i, in Morton order
    };

struct ConstLBVHData
    {
    const int* parent;
    const int* left;
    const int* right;
    const unsigned int* primitive;
    const float3* lo;
    const float3* hi;
    int root;
    const int* rope;
    const int* range_end;
    const float4* sorted_points;
    };

} // end namespace neighbor

#endif // NEIGHBOR_LBVH_DATA_H_
