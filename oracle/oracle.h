/*
 * oracle.h -- CPU restatement of the reference LBVH neighbor-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links or calls
 * it, and the product has no CPU fallback.
 *
 * Every function restates the semantics of the reference (mphowardlab/neighbor,
 * HIP branch of include/neighbor/ApproximateMath.h) and cites the file:line it
 * follows.  Pinning (see tests/test_oracle_*.py):
 *   - directed rounding   : glibc nextafterf over special and random values,
 *                           and the 28 known answers of test/approx_math_test.cu:24-131
 *   - Morton codes        : brute-force bit interleave; test/lbvh_test.cu:94-98
 *   - sort                : numpy lexsort on (code, index)
 *   - Karras hierarchy    : recursive top-down radix-tree split (independent);
 *                           test/lbvh_test.cu:76-92, 534-539
 *   - AABB bubble         : recursive post-order union (independent);
 *                           test/lbvh_test.cu:105-117, 550-552
 *   - compression/ropes   : preorder skip pointers (independent);
 *                           test/lbvh_test.cu:133-190, 571-576
 *   - rope traversal      : flat "root flag and leaf overlap, leaf order" scan
 *                           (independent); test/lbvh_test.cu:193-310, 357-386, 493-502, 578-580
 *   - exact FP64 pairs    : brute force vs cell list; hand-made tiny cases
 *   - BoundingSphere      : closed form of the double ctor (origin = largest float <= o,
 *                           enclosure of the double sphere checked in exact rational
 *                           arithmetic) and of the overlap test (exact distance)
 */
#ifndef NEIGHBOR_ORACLE_H_
#define NEIGHBOR_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SENTINEL 0x7fffffff

/* ---- directed rounding, HIP path (ApproximateMath.h:19-21, 55-345) ---- */
float orc_next_down(float x);               /* nextafterf(x, -FLT_MAX) */
float orc_next_up(float x);                 /* nextafterf(x, +FLT_MAX) */
float orc_fadd_rd(float x, float y);
float orc_fadd_ru(float x, float y);
float orc_fsub_rd(float x, float y);
float orc_fsub_ru(float x, float y);
float orc_fmul_rd(float x, float y);
float orc_fmul_ru(float x, float y);
float orc_fdiv_rd(float x, float y);
float orc_fdiv_ru(float x, float y);
float orc_frcp_rd(float x);
float orc_frcp_ru(float x);
float orc_fmaf_rd(float x, float y, float z);
float orc_fmaf_ru(float x, float y, float z);
float orc_double2float_rd(double x);
float orc_double2float_ru(double x);
uint32_t orc_sat_f2u(float x);              /* saturating float -> u32 conversion */

/* ---- build (kernels/LBVH.cuh) ---- */
/* sphere_r < 0 selects PointInsertOp, otherwise SphereInsertOp with radius sphere_r. */
void orc_leaf_box(const float* xyz, uint32_t idx, float sphere_r, float lo[3], float hi[3]);
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z);
void orc_morton_codes(const float* xyz, uint32_t N, float sphere_r,
                      const float lo[3], const float hi[3], uint32_t* codes);
void orc_sort_codes(uint32_t N, const uint32_t* codes, uint32_t* sorted_codes, uint32_t* primitive);
void orc_karras(uint32_t N, const uint32_t* sorted_codes, int32_t* parent, int32_t* left, int32_t* right);
void orc_bubble(uint32_t N, const int32_t* parent, const int32_t* left, const int32_t* right,
                const uint32_t* primitive, const float* xyz, float sphere_r, float* lo, float* hi);
/* Full build: codes -> sort -> hierarchy -> bubble.  Arrays sized as in LBVH::allocate. */
void orc_build(const float* xyz, uint32_t N, float sphere_r, const float scene_lo[3], const float scene_hi[3],
               uint32_t* sorted_codes, uint32_t* primitive, int32_t* parent, int32_t* left, int32_t* right,
               float* lo, float* hi);

/* ---- compression (kernels/LBVHTraverser.cuh:45-125) ---- */
void orc_compress(uint32_t N, const int32_t* parent, const int32_t* left, const int32_t* right,
                  const uint32_t* primitive, const float* lo, const float* hi, const uint32_t* map,
                  int32_t* records, float root_lo[3], float root_hi[3], float bins[3]);

/* ---- rope traversal (kernels/LBVHTraverser.cuh:155-252) ----
 * spheres: Nq x float4 (x,y,z,R); images: Nimg x float3.
 * counts[Nq] always written.  list (nullable) is the NeighborListOp layout,
 * row stride max_neigh.  triplet outputs (nullable) receive, in emission order,
 * the image index of every hit; trip_offsets[q] gives the start of query q's hits
 * (exclusive scan of counts, caller supplied).  Returns the total number of hits;
 * *node_visits receives the number of node records loaded.
 */
uint64_t orc_traverse(uint32_t N, const int32_t* records, const float root_lo[3], const float root_hi[3],
                      const float bins[3], const float* spheres, uint32_t Nq, const float* images, uint32_t Nimg,
                      uint32_t* counts, uint32_t* list, uint32_t max_neigh,
                      const uint64_t* trip_offsets, uint32_t* trip_prim, uint8_t* trip_image,
                      uint64_t* node_visits);

/* ---- BoundingSphere (BoundingVolumes.h:102-175) ----
 * orc_sphere_double: double ctor (:135-147); orc_sphere_float: float ctor (:119-123);
 * orc_sphere_overlap: overlap (:163-171), 1 when the box [lo, hi] overlaps. */
void orc_sphere_double(const double o[3], double r, float origin[3], float* Rsq);
void orc_sphere_float(const float o[3], float r, float origin[3], float* Rsq);
int orc_sphere_overlap(const float origin[3], float Rsq, const float lo[3], const float hi[3]);

/* ---- exact FP64 pair definition ----
 * (i, j, k) is a pair iff  d2 <= R_i*R_i  with  d = ((double)q_i + t_k) - (double)x_j  per axis and
 * d2 = (dx*dx + dy*dy) + dz*dz  evaluated in IEEE double without contraction.
 * queries: Nq x float4 (x,y,z,R); images: Nimg x double3.
 */
/* OpenMP threads of the exact routines (n <= 0 leaves the current setting). */
void orc_set_threads(int n);
uint64_t orc_exact_bruteforce(const float* xyz, uint32_t N, const float* queries, uint32_t Nq,
                              const double* images, uint32_t Nimg, uint32_t* counts,
                              const uint64_t* offsets, uint32_t* out_j, uint8_t* out_k);
/* Cell-list version for an orthorhombic periodic box [0,L) with the 27 images of
 * make_images() order (ix,iy,iz in -1..1, iz fastest).  Same set, same per-query ordering
 * (image ascending, then j ascending).  nthreads <= 0 uses OpenMP's default. */
uint64_t orc_exact_celllist(const float* xyz, uint32_t N, const float* queries, uint32_t Nq,
                            const double box[3], double rmax, uint32_t* counts,
                            const uint64_t* offsets, uint32_t* out_j, uint8_t* out_k, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* NEIGHBOR_ORACLE_H_ */
