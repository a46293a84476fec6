// neighbor.h -- umbrella header of the MI355X-native LBVH neighbor search
// (same entry point as the reference's include/neighbor/neighbor.h:8-18).
#ifndef NEIGHBOR_NEIGHBOR_H_
#define NEIGHBOR_NEIGHBOR_H_

#include "BoundingVolumes.h"
#include "InsertOps.h"
#include "OutputOps.h"
#include "QueryOps.h"
#include "TransformOps.h"
#include "TranslateOps.h"

#include "LBVH.h"
#include "LBVHTraverser.h"

#endif // NEIGHBOR_NEIGHBOR_H_
