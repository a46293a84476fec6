// LBVHTraverserData.h -- raw view of a compressed LBVH.
//
// Quantized rope record (reference LBVHTraverserData.h:14-21, layout
// LBVHTraverser.h:459-480), 16 B per node:
//   x = lo bins  (x<<20)+(y<<10)+z, rounded outward from the root box lower corner
//   y = hi bins  measured down from the root box upper corner
//   z = left child (>= 0) or ~cached primitive (< 0) for a leaf
//   w = rope (next node when the subtree is skipped; LBVHSentinel ends the walk)
// Float rope record (this library), 32 B per node, two 16-B loads:
//   {lo.x, lo.y, lo.z, bits(z)}, {hi.x, hi.y, hi.z, bits(w)} with z, w as above.
#ifndef NEIGHBOR_LBVH_TRAVERSER_DATA_H_
#define NEIGHBOR_LBVH_TRAVERSER_DATA_H_

#include <hip/hip_runtime.h>

namespace neighbor
{

struct LBVHCompressedData
    {
    int root;
    int4* data;
    float3* lo;
    float3* hi;
    float3* bins;
    };

struct LBVHFloatRopeData
    {
    int root;
    float4* data;   // 2 float4 per node
    };

} // end namespace neighbor

#endif // NEIGHBOR_LBVH_TRAVERSER_DATA_H_
