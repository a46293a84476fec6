/*
 * neighbor_amd.h -- C ABI of the MI355X-native LBVH neighbor search.
 *
 * The reference (mphowardlab/neighbor) is a header-only C++ template library; its
 * callers instantiate LBVH::build and LBVHTraverser::traverse with concrete ops.
 * This ABI exports the concrete instantiations a foreign-function caller (ctypes,
 * cgo, a C host code) would bind, with plain pointers and sizes.  The C++
 * template API itself (same class and op names as the reference) lives in
 * neighbor_amd/include/neighbor/neighbor.h.
 *
 * Conventions
 *   - Every function returns an int status: NB_OK (0) or a negative NB_ERR_*;
 *     nb_last_error_string() describes the last failure of the calling thread.
 *   - Array arguments are DEVICE pointers (HBM) unless marked [host].
 *   - All work is enqueued on `stream` (a hipStream_t; NULL = default stream);
 *     nothing synchronizes the host unless stated.
 *   - Handles are not thread-safe; use one handle per host thread.
 */
#ifndef NEIGHBOR_AMD_H_
#define NEIGHBOR_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NB_OK 0
#define NB_ERR_INVALID (-1)   /* bad handle, size or pointer */
#define NB_ERR_TUNING (-2)    /* block size outside the tunable set: Tunable.h:110-118 (std::runtime_error there) */
#define NB_ERR_IMAGES (-3)    /* more than 32 images: LBVHTraverser.h:429-432 (std::runtime_error there) */
#define NB_ERR_HIP (-4)       /* HIP runtime or launch failure */
#define NB_ERR_ALLOC (-5)     /* device allocation failure: Memory.h:185 (std::runtime_error there) */
#define NB_ERR_INTERNAL (-6)

#define NB_SENTINEL 0x7fffffff /* LBVHData.h:14 LBVHSentinel */

/* traversal variants */
#define NB_VARIANT_QUANTIZED 0        /* 16-B quantized rope records (the reference's traverser) */
#define NB_VARIANT_FLOAT_ROPE 1       /* 32-B records of decoded quantized boxes: same results, no decode work */
#define NB_VARIANT_FLOAT_ROPE_EXACT 2 /* 32-B records of exact node boxes: for the exact queries */
#define NB_VARIANT_WAVE 3             /* exact queries only: one rope path per wave of 64 queries, bucket leaves */
#define NB_VARIANT_WAVE_QUANTIZED 4   /* the wave walk over decoded quantized boxes: reference results, any query */

typedef struct nb_lbvh nb_lbvh;
typedef struct nb_traverser nb_traverser;

/* Message of the calling thread's last failed call ("" when none). */
const char* nb_last_error_string(void);
/* ABI version (major*10000 + minor*100 + patch). */
int nb_version(void);

/* ------------------------------------------------------------------ LBVH
 * Replaces: LBVH::LBVH() and ~LBVH (include/neighbor/LBVH.h:224-227). */
int nb_lbvh_create(nb_lbvh** out);
int nb_lbvh_destroy(nb_lbvh* lbvh);

/* Replaces: LBVH::build(LaunchParameters(block, stream), PointInsertOp(points, N), lo, hi)
 * (LBVH.h:242-318, InsertOps.h:18-52).  points: N float3; lo, hi [host]: scene bounds. */
int nb_lbvh_build_points(nb_lbvh* lbvh, const float* points, uint32_t N, const float* lo, const float* hi,
                         uint32_t block, void* stream);

/* Replaces: LBVH::build(..., SphereInsertOp(points, radius, N), lo, hi) (InsertOps.h:57-94). */
int nb_lbvh_build_spheres(nb_lbvh* lbvh, const float* points, uint32_t N, float radius, const float* lo,
                          const float* hi, uint32_t block, void* stream);

/* Build over 16-byte particle records (x, y, z, w; HOOMD's postype layout, w not read), read in
 * place: primitive i is record index[i] (index NULL: record i), N primitives.  radius < 0: point
 * leaves (as nb_lbvh_build_points); radius >= 0: sphere leaves (as nb_lbvh_build_spheres).  The
 * tree is the one the float3 calls build over the gathered points.  (The reference's benchmark
 * reads the same records through its own insert op, benchmark/lbvh_benchmark.cuh:25-60.) */
int nb_lbvh_build_postype(nb_lbvh* lbvh, const float* postype, const uint32_t* index, uint32_t N, float radius,
                          const float* lo, const float* hi, uint32_t block, void* stream);
int nb_lbvh_refit_postype(nb_lbvh* lbvh, const float* postype, const uint32_t* index, uint32_t N, float radius,
                          uint32_t block, void* stream);

/* Refit (no reference counterpart; builds on the grow-only reuse of LBVH::allocate,
 * LBVH.h:336-402): new primitive boxes under the hierarchy of the last build -- the same
 * sorted order, parents, children and ropes, leaf boxes from the new points, internal boxes
 * bubbled up.  N must equal the last build's.  Every node box bounds its leaves for any motion,
 * so traversal stays exact; the boxes only loosen as points leave their Morton order.  A
 * traverser set up on the tree must be set up again (nb_traverser_setup). */
int nb_lbvh_refit_points(nb_lbvh* lbvh, const float* points, uint32_t N, uint32_t block, void* stream);
int nb_lbvh_refit_spheres(nb_lbvh* lbvh, const float* points, uint32_t N, float radius, uint32_t block, void* stream);

/* max over i of the squared displacement of points[i] from ref[i], minimum image in a periodic
 * box of lengths box[0..2] ([host]; NULL: not periodic), in FP64 and rounded up by a relative
 * 2^-40, written to *out (device, one double).  The skin test of a neighbor list built at
 * cutoff + skin: it is still complete while 4 * max <= skin^2.  wrapped (device u32, may be
 * NULL) receives the number of points whose minimum image is not the plain difference (moved
 * across the box; a refit keeps them at their old place in the hierarchy, stretching their
 * ancestors' boxes across the box: rebuild instead).  (No reference counterpart.) */
int nb_max_displacement2(const float* points, const float* ref, uint32_t N, const float* box, double* out,
                         uint32_t* wrapped, void* stream);

/* Replaces: LBVH::getN / getNInternal / getNNodes / getRoot (LBVH.h:111-133).  Outputs [host]. */
int nb_lbvh_info(const nb_lbvh* lbvh, uint32_t* N, uint32_t* n_internal, uint32_t* n_nodes, int32_t* root);

/* Replaces: LBVH::getParents / getLeftChildren / getRightChildren / getPrimitives /
 * getLowerBounds / getUpperBounds / data() (LBVH.h:135-183).  Returns device pointers
 * owned by the handle (valid until the next build or destroy); any output may be NULL.
 * lo, hi: n_nodes float3.  rope and codes are this library's additions (ropes per node,
 * sorted Morton codes). */
int nb_lbvh_arrays(const nb_lbvh* lbvh, const int32_t** parent, const int32_t** left, const int32_t** right,
                   const uint32_t** primitive, const float** lo, const float** hi, const int32_t** rope,
                   const uint32_t** codes);

/* Valid block sizes (Tunable::getTunableParameters, Tunable.h:60-63).  out [host] may be NULL;
 * *count receives the number of values. */
int nb_lbvh_tunable_parameters(const nb_lbvh* lbvh, uint32_t* out, uint32_t capacity, uint32_t* count);

/* ------------------------------------------------------------------ traverser
 * Replaces: LBVHTraverser::LBVHTraverser() (LBVHTraverser.h:357-361). */
int nb_traverser_create(nb_traverser** out);
int nb_traverser_destroy(nb_traverser* trav);

/* Select NB_VARIANT_*; a change discards a replayed setup. (No reference counterpart.) */
int nb_traverser_set_variant(nb_traverser* trav, int variant);

/* Leaves per bucket of NB_VARIANT_WAVE (1..64, default 16); a change discards a replayed
 * setup.  (No reference counterpart.) */
int nb_traverser_set_bucket_size(nb_traverser* trav, uint32_t bucket);

/* Launch shape of the exact wave walks (nb_traverse_exact_count / _pairs / _rows /
 * _rows_placed): per XCD, the last `groups` groups of 256 queries run as `split` blocks (1, 2
 * or 4) of 256/split queries each, so the waves that start last in the launch are shorter.
 * Default 32 groups, split 2 (the last wave of each SIMD of an XCD); groups = 0: one block per
 * group.  Results do not depend on it.  (No reference counterpart.) */
int nb_traverser_set_tail_split(nb_traverser* trav, uint32_t groups, uint32_t split);

/* Floats per record of the `points` argument of the exact calls (nb_traverse_exact_*): 3
 * (float3, the default) or 4 (16-byte postype records, x, y, z first).  (No reference
 * counterpart.) */
int nb_traverser_set_point_stride(nb_traverser* trav, uint32_t stride);

/* No counterpart in the reference (walk statistics of the wave variants, for measurement):
 * while set, every wave-variant count traversal through trav writes, per launched wave w, the
 * node records it loaded, the buckets whose leaves it tested and the leaves in them to
 * log[3w .. 3w+2] (device memory of `words` u32; nb_visit_log_words(n) for n queries).  A
 * traversal whose waves need more words, or a wave traversal with another output op, fails
 * with NB_ERR_INVALID while the log is set.  log NULL switches it off. */
int nb_traverser_set_visit_log(nb_traverser* trav, uint32_t* log, uint64_t words);
size_t nb_visit_log_words(uint32_t n_queries);

/* Replaces: LBVHTraverser::setup(LaunchParameters(block, stream), lbvh[, MapTransformOp(map)])
 * (LBVHTraverser.h:378-390).  map: NULL for NullTransformOp. */
int nb_traverser_setup(nb_traverser* trav, const nb_lbvh* lbvh, const uint32_t* map, uint32_t block, void* stream);

/* Replaces: LBVHTraverser::reset() (LBVHTraverser.h:135-138). */
int nb_traverser_reset(nb_traverser* trav);

/* Replaces: LBVHTraverser::getData() (LBVHTraverser.h:326-329) plus the compression's
 * root box and bin sizes (LBVHTraverserData.h:14-21).  Device pointers owned by trav:
 * records = n_nodes int4; root_lo, root_hi, bins = float3; float_records = 2*n_nodes float4
 * (NULL unless a float or wave variant compressed; the Wave variant leaves the records of
 * nodes below a bucket unwritten, since its walk never reaches them). */
int nb_traverser_data(const nb_traverser* trav, const int32_t** records, const float** root_lo,
                      const float** root_hi, const float** bins, const float** float_records);

/* Replaces: LBVHTraverser::traverse(LaunchParameters(block, stream), lbvh,
 *   SphereQueryOp(spheres, Nq), CountNeighborsOp(nneigh), ImageListOp<float3>(images, Nimg)
 *   [or SelfOp() when images == NULL], MapTransformOp(map) [or NullTransformOp when NULL])
 * (LBVHTraverser.h:414-450, QueryOps.h:32-120, OutputOps.h:28-103, TranslateOps.h:19-91).
 * spheres: Nq float4 (x, y, z, R); images: Nimg float3.  order (NULL = identity, no
 * reference counterpart): thread t serves query order[t]. */
int nb_traverse_count(nb_traverser* trav, const nb_lbvh* lbvh, const float* spheres, uint32_t Nq,
                      const float* images, uint32_t Nimg, const uint32_t* map, const uint32_t* order,
                      uint32_t* nneigh, uint32_t block, void* stream);

/* Replaces: the same traverse with NeighborListOp(neigh_list, nneigh, max_neigh)
 * (OutputOps.h:113-176).  neigh_list: Nq*max_neigh u32 (row offsets in 64 bits). */
int nb_traverse_list(nb_traverser* trav, const nb_lbvh* lbvh, const float* spheres, uint32_t Nq,
                     const float* images, uint32_t Nimg, const uint32_t* map, const uint32_t* order,
                     uint32_t* neigh_list, uint32_t* nneigh, uint32_t max_neigh, uint32_t block, void* stream);

/* Same traversal as nb_traverse_count with the hits written as (neighbor, image) pairs:
 * query q's hits go to [offsets[q], offsets[q+1]) in emission order (image ascending, then
 * leaf order).  offsets: Nq+1 u64 (usually the exclusive scan of nb_traverse_count);
 * capacity: entries of neigh (and image); hits at or past it are dropped but counted.
 * image may be NULL; nneigh may be NULL.  (No reference counterpart: the reference never
 * reports the image.) */
int nb_traverse_pairs(nb_traverser* trav, const nb_lbvh* lbvh, const float* spheres, uint32_t Nq,
                      const float* images, uint32_t Nimg, const uint32_t* map, const uint32_t* order,
                      const uint64_t* offsets, uint32_t* neigh, uint8_t* image, uint32_t* nneigh,
                      uint64_t capacity, uint32_t block, void* stream);

/* Exact FP64 queries (no reference counterpart; the reference's hits are a conservative
 * superset).  (q, j, k) is reported iff d2 <= R_q^2 with d = ((double)q + t_k) - (double)x_j
 * per axis and d2 = (dx*dx + dy*dy) + dz*dz in IEEE double.  points: the float3 positions
 * indexed by the value cached in the leaves (the primitive index when map == NULL; with a
 * map, points must be indexed by map[primitive]).  images: Nimg double3, or NULL for the
 * self image. */
int nb_traverse_exact_count(nb_traverser* trav, const nb_lbvh* lbvh, const float* spheres, uint32_t Nq,
                            const float* points, const double* images, uint32_t Nimg, const uint32_t* map,
                            const uint32_t* order, uint32_t* nneigh, uint32_t block, void* stream);
int nb_traverse_exact_pairs(nb_traverser* trav, const nb_lbvh* lbvh, const float* spheres, uint32_t Nq,
                            const float* points, const double* images, uint32_t Nimg, const uint32_t* map,
                            const uint32_t* order, const uint64_t* offsets, uint32_t* neigh, uint8_t* image,
                            uint32_t* nneigh, uint64_t capacity, uint32_t block, void* stream);

/* Exact queries written into rows given by a start and a length per query (e.g. rows sized by
 * nb_capacity_predict): query q's hits go to [start[q], start[q] + limit[q]) in emission order,
 * and nneigh[q] (required) receives q's exact count.
 * spill == NULL: hits past the row (or past capacity) are dropped but counted
 *   (nneigh[q] > limit[q] marks an overflowed row, for nb_overflow_rows).
 * spill != NULL (device, 2 x u64; on entry spill[0] = first free position after all rows,
 *   spill[1] = 0): an overflowed query gets a new row of exactly nneigh[q] entries at
 *   spill[0] (start[q] and limit[q] are rewritten) and is walked again into it within the
 *   same launch.  On exit spill[0] = end of the used positions (all rows are complete iff
 *   spill[0] <= capacity) and spill[1] = number of moved rows.  (No reference counterpart.) */
int nb_traverse_exact_rows(nb_traverser* trav, const nb_lbvh* lbvh, const float* spheres, uint32_t Nq,
                           const float* points, const double* images, uint32_t Nimg, const uint32_t* map,
                           const uint32_t* order, uint64_t* start, uint32_t* limit, uint32_t* neigh,
                           uint8_t* image, uint32_t* nneigh, uint64_t capacity, uint64_t* spill, uint32_t block,
                           void* stream);
/* nb_traverse_exact_rows for the tree's own primitives (Nq == N, query q = primitive q) in the
 * tree's primitive (Morton) order, with row starts placed by the walk instead of a scan over
 * the queries: the query at sorted position t starts at base[t / 256] plus the limits of the
 * sorted positions before it in its group of 256, and that start is written to start[q].  base
 * (ceil(N / 256) + 1 u64) comes from nb_capacity_predict_rows with the same limits; spill as in
 * nb_traverse_exact_rows (that call also initializes it).  Wave variants only (NB_ERR_INVALID
 * otherwise); their kernels run 256-thread blocks.  (No reference counterpart.) */
int nb_traverse_exact_rows_placed(nb_traverser* trav, const nb_lbvh* lbvh, const float* spheres, uint32_t Nq,
                                const float* points, const double* images, uint32_t Nimg, const uint32_t* map,
                                const uint64_t* base, uint64_t* start, uint32_t* limit, uint32_t* neigh, uint8_t* image,
                                uint32_t* nneigh, uint64_t capacity, uint64_t* spill, void* stream);

/* ------------------------------------------------------------------ learned capacity
 * (No reference counterpart.)  A 8-32-32-1 FP32 MLP on Morton-cell occupancy features predicts
 * each query's neighbor count so that the exact pairs can be written in one pass; overflowed
 * rows are redone by the count-then-fill path (nb_overflow_rows, nb_relocate_rows). */
typedef struct nb_capacity nb_capacity;
int nb_capacity_create(nb_capacity** out, uint64_t seed);
int nb_capacity_destroy(nb_capacity* pred);
/* Copy the 1377 parameters out ([host] params_out) and/or in ([host] params_in); either may be NULL. */
int nb_capacity_parameters(nb_capacity* pred, float* params_out, const float* params_in, uint32_t* count);
/* Features of the primitives of lbvh (lo, hi [host] = the build's scene box, r = query radius);
 * *features receives the device array (N x 8, sorted order) owned by pred. */
int nb_capacity_features(nb_capacity* pred, const nb_lbvh* lbvh, const float* lo, const float* hi, float r,
                         const float** features, void* stream);
/* Adam training on exact counts (indexed by query = primitive); *loss [host] = last batch loss.
 * Synchronizes the stream. */
int nb_capacity_train(nb_capacity* pred, const nb_lbvh* lbvh, const uint32_t* counts, uint32_t steps, uint32_t batch,
                      float lr, uint64_t seed, float* loss, void* stream);
/* Margin (log2 units) so that a fraction `quantile` of the training counts fit; synchronizes. */
int nb_capacity_margin(nb_capacity* pred, const nb_lbvh* lbvh, float quantile, float* margin, void* stream);
/* cap[q] = ceil(2^(prediction + margin)) - 1 + pad for every query q = primitive. */
int nb_capacity_predict(nb_capacity* pred, const nb_lbvh* lbvh, float margin, uint32_t pad, uint32_t* cap, void* stream);
/* The same capacities with the features computed on the fly from the tree (lo, hi [host] = the
 * build's scene box, r = query radius) in one kernel: nb_capacity_features then
 * nb_capacity_predict, without storing the features. */
int nb_capacity_predict_direct(nb_capacity* pred, const nb_lbvh* lbvh, const float* lo, const float* hi, float r,
                               float margin, uint32_t pad, uint32_t* cap, void* stream);
/* nb_capacity_predict_direct plus the row bases of nb_traverse_exact_rows_placed: the kernel
 * also sums the capacities of each group of 256 sorted positions, and one scan of those sums
 * gives base[g] (g < ceil(N / 256)) and base[ceil(N / 256)] = the sum of all capacities.  spill
 * (2 u64, may be NULL) receives {that sum, 0}: the fill's spill cursor and moved-row count. */
int nb_capacity_predict_rows(nb_capacity* pred, const nb_lbvh* lbvh, const float* lo, const float* hi, float r,
                             float margin, uint32_t pad, uint32_t* cap, uint64_t* base, uint64_t* spill, void* stream);
/* y[i] = MLP(features[i]) for n samples (features: n x 8, all device pointers) with the inference
 * forward pass: the two matrix layers on the matrix cores (FP16 operands, FP32 accumulation),
 * biases and the output layer in FP32.  The capacities of nb_capacity_predict(_direct) and the
 * margin of nb_capacity_margin come from this forward pass. */
int nb_capacity_forward(const float* params, const float* features, uint32_t n, float* y, void* stream);
/* Gradient of mean((y - t)^2) over `batch` samples (features: batch x 8, targets: batch) at params;
 * grad receives 1377 gradients and, at index 1377, the loss.  All device pointers. */
int nb_capacity_gradient(const float* params, const float* features, const float* targets, uint32_t batch, float* grad,
                         void* stream);
/* Queries with counts[q] > cap[q] appended to list; *n_list (device u32) must be zero on entry. */
int nb_overflow_rows(const uint32_t* counts, const uint32_t* cap, uint32_t n, uint32_t* list, uint32_t* n_list,
                     void* stream);
/* Give the n_list overflowed queries new rows after `tail`: start[list[t]] = tail + (exclusive
 * scan of their counts)[t], limit[list[t]] = counts[list[t]].  Device scratch of
 * nb_relocate_scratch_bytes(n_list) bytes. */
size_t nb_relocate_scratch_bytes(uint32_t n_list);
int nb_relocate_rows(const uint32_t* list, uint32_t n_list, const uint32_t* counts, uint64_t tail, uint64_t* start,
                     uint32_t* limit, void* scratch, void* stream);

/* Device address of the build's error word (zero after a good build; read it after a
 * synchronization): bit 0 a radix-sort look-back gave up its bounded wait, bits 1-2 a
 * bubble span list overflowed.  A refit keeps the bits of its build.  (No reference
 * counterpart.) */
int nb_lbvh_error_word(const nb_lbvh* lbvh, const uint32_t** word);
/* Synchronizes `stream` and returns NB_ERR_INTERNAL (message in nb_last_error_string) when the
 * last build or refit gave up a bounded wait or list: its tree is incomplete and must not be
 * traversed.  NB_OK otherwise.  (No reference counterpart: the reference has no bounded waits.) */
int nb_lbvh_check(const nb_lbvh* lbvh, void* stream);
/* (No reference counterpart.)  on != 0: every later build or refit of lbvh records an event of
 * the library's once the sorted codes, primitive order and sorted points are complete (for point
 * leaves before the bubble passes); nb_lbvh_wait_sorted makes `stream` wait for it, so work that
 * needs only those arrays (nb_capacity_predict_rows) can overlap the rest of the build on
 * another stream.  NB_ERR_INTERNAL from the wait when the event was never enabled. */
int nb_lbvh_enable_sorted_event(nb_lbvh* lbvh, int on);
int nb_lbvh_wait_sorted(const nb_lbvh* lbvh, void* stream);
/* (No reference counterpart; a testing knob.)  The bubble's span passes climb each span's list of
 * chunk-crossing nodes in LDS up to `limit` entries (default and maximum 1024) and through global
 * memory beyond it; 0 takes the global-memory path for every span.  Applies to later builds and
 * refits of lbvh; the tree is the same either way. */
int nb_lbvh_set_span_lds_limit(nb_lbvh* lbvh, uint32_t limit);

/* ------------------------------------------------------------------ batched boxes
 * (No reference counterpart.)  Many independent boxes built as one forest with one sort and
 * traversed by one persistent launch: points of all boxes concatenated (box b holds
 * [offsets[b], offsets[b+1])), each box with its own lower/upper corner (its periodic cell).
 * Queries are the points themselves: spheres (device, N x float4 x, y, z, r) in the same
 * order, each queried against its own box only (27 images of the box when periodic != 0).
 * Outputs are indexed by the global point index and neighbors are global indices.  The
 * pair semantics are those of nb_traverse_exact_pairs. */
typedef struct nb_batch nb_batch;
#define NB_BATCH_WAVE 0   /* one rope path per wave of 64 Morton-consecutive queries */
#define NB_BATCH_LANE 1   /* one rope path per lane */
int nb_batch_create(nb_batch** out);
int nb_batch_destroy(nb_batch* batch);
/* points: N x 3 float; offsets: nbox + 1 u32 (offsets[0] = 0, offsets[nbox] = N); box_lo/box_hi:
 * nbox x 3 float.  All device arrays, read again by later traversals: keep them alive. */
int nb_batch_build(nb_batch* batch, const float* points, uint32_t N, const uint32_t* offsets, const float* box_lo,
                   const float* box_hi, uint32_t nbox, void* stream);
/* Tunables: block (hierarchy kernel, multiple of 32 in [32, 1024]), bucket (leaves per bucket,
 * 1..64), variant (NB_BATCH_*), grid (persistent blocks, 0 = 8 per CU).  Bucket changes take
 * effect at nb_batch_setup. */
int nb_batch_set_params(nb_batch* batch, uint32_t block, uint32_t bucket, int variant, uint32_t grid);
int nb_batch_get_params(const nb_batch* batch, uint32_t* block, uint32_t* bucket, int* variant, uint32_t* grid);
/* Bucket records of the built forest (after every build). */
int nb_batch_setup(nb_batch* batch, void* stream);
/* Device address of the forest build's error word (as nb_lbvh_error_word). */
int nb_batch_error_word(const nb_batch* batch, const uint32_t** word);
/* As nb_lbvh_check, for the forest's last build. */
int nb_batch_check(const nb_batch* batch, void* stream);
int nb_batch_count(nb_batch* batch, const float* spheres, int periodic, uint32_t* nneigh, void* stream);
int nb_batch_pairs(nb_batch* batch, const float* spheres, int periodic, const uint64_t* offsets, uint32_t* neigh,
                   uint8_t* image, uint32_t* nneigh, uint64_t capacity, void* stream);
/* Pairs into rows [start[q], start[q] + limit[q]) (as nb_traverse_exact_rows, spill may be NULL);
 * with start from nb_exclusive_scan_permuted_u32_u64 over the forest's primitive order and
 * limit = the counts, rows are laid out in Morton order per box (the fastest writes). */
int nb_batch_rows(nb_batch* batch, const float* spheres, int periodic, uint64_t* start, uint32_t* limit, uint32_t* neigh,
                  uint8_t* image, uint32_t* nneigh, uint64_t capacity, uint64_t* spill, void* stream);
/* Row capacities for a single-pass nb_batch_rows without a count pass (no reference
 * counterpart): cap[i] = ceil(lambda + z * sqrt(lambda)) + pad, at least 1, with lambda =
 * (N_b / V_b) * (4/3) pi r_i^3 the mean neighbor count of point i's box b at its query radius
 * (spheres w).  With start from nb_exclusive_scan_permuted_u32_u64 of cap over the forest's
 * primitive order and a spill counter {start[N], 0}, nb_batch_rows(limit = cap) writes every
 * row exactly; rows that overflow are moved past all rows inside the launch. */
int nb_batch_capacity(nb_batch* batch, const float* spheres, float z, uint32_t pad, uint32_t* cap, void* stream);
/* Device arrays of the forest (any output may be NULL): box roots (nbox), primitive (N, sorted
 * order -> point index), keys (N u64: box << 30 | Morton code, sorted), parent, left, right,
 * rope (nodes), lower/upper node bounds (nodes x 3 float); counts on the host. */
int nb_batch_arrays(const nb_batch* batch, const int** roots, const uint32_t** primitive, const uint64_t** keys,
                    const int** parent, const int** left, const int** right, const int** rope, const float** lo,
                    const float** hi, uint32_t* N, uint32_t* nbox, uint32_t* n_nodes);
/* Autotuner: times the build per block size and the count pass per (variant, bucket) on this
 * problem, keeps the fastest on the batch and reports them (host outputs; NULL allowed).
 * nneigh: device scratch of N u32.  Leaves the batch built and set up.  Synchronizes. */
int nb_batch_tune(nb_batch* batch, const float* points, uint32_t N, const uint32_t* offsets, const float* box_lo,
                  const float* box_hi, uint32_t nbox, const float* spheres, int periodic, uint32_t* nneigh,
                  uint32_t* block, uint32_t* bucket, int* variant, float* build_ms, float* count_ms, void* stream);
/* One graph = build + setup + count + scan + pair fill with the batch's current tunables.
 * Buffers: nneigh (N), pair_offsets (N + 1 u64), scan scratch (nb_scan_scratch_bytes(N)),
 * neigh/image (capacity; image may be NULL).  Rows are laid out in the forest's Morton order:
 * point q's pairs are [pair_offsets[q], pair_offsets[q] + nneigh[q]), pair_offsets[N] = total.
 * Creation runs the step once eagerly (and synchronizes); each launch replays it on `stream`.
 * The graph refers to the batch's device buffers and to every array passed here: keep the
 * arrays alive while the plan is used.  nb_batch_plan_launch returns NB_ERR_INVALID and
 * launches nothing once the batch has reallocated a buffer (nb_batch_build with more points or
 * boxes than at capture, or a larger bucket setup) or has been destroyed; create a new plan
 * then.  A launch rebuilds the batch's forest over the plan's own problem. */
typedef struct nb_batch_plan nb_batch_plan;
int nb_batch_plan_create(nb_batch_plan** out, nb_batch* batch, const float* points, uint32_t N, const uint32_t* offsets,
                         const float* box_lo, const float* box_hi, uint32_t nbox, const float* spheres, int periodic,
                         uint32_t* nneigh, uint64_t* pair_offsets, void* scan_scratch, uint32_t* neigh, uint8_t* image,
                         uint64_t capacity);
int nb_batch_plan_launch(nb_batch_plan* plan, void* stream);
int nb_batch_plan_destroy(nb_batch_plan* plan);

/* ------------------------------------------------------------------ utilities
 * Exclusive scan of n u32 counts into n+1 u64 offsets (offsets[n] = total).  Device
 * scratch: `scratch` of nb_scan_scratch_bytes(n) bytes.  (No reference counterpart.) */
size_t nb_scan_scratch_bytes(uint32_t n);
int nb_exclusive_scan_u32_u64(const uint32_t* counts, uint32_t n, uint64_t* offsets, void* scratch, void* stream);
/* The same scan in the order of a permutation perm (n u32): offsets[perm[t]] = sum of
 * counts[perm[t']] over t' < t, offsets[n] = total.  With perm = the traversal order, the
 * rows [offsets[q], offsets[q] + counts[q]) of the queries of one wave are adjacent, which is
 * what nb_traverse_exact_rows (limit = counts) writes fastest. */
int nb_exclusive_scan_permuted_u32_u64(const uint32_t* counts, const uint32_t* perm, uint32_t n, uint64_t* offsets,
                                       void* scratch, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the exclusive scan of
 * nb_exclusive_scan_u32_u64 (perm == NULL) or nb_exclusive_scan_permuted_u32_u64 (perm != NULL)
 * with its tile sums in a static device array of the library instead of a workspace argument.
 * Allocates nothing; n is limited to 2^28 counts (NB_ERR_INVALID above that).  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (different streams, or host threads)
 * race and give wrong offsets; serialize them. */
int nb_exclusive_scan_u32_u64_noalloc(const uint32_t* counts, const uint32_t* perm, uint32_t n, uint64_t* offsets,
                                      void* stream);

/* In-place filter of neighbor rows (no reference counterpart: the half lists and the
 * diameter-shifted per-pair cutoffs of a HOOMD-style per-type neighbor list).  Row q holds
 * the pairs (i, j, k) of particle i = query_index[q] (query_index NULL: i = q) at
 * neigh/image[start[q], start[q] + counts[q]); the kept pairs move to the front of their row,
 * in order, and counts_out[q] (may alias counts) receives their number.
 * half != 0: keep j > i, and for j == i the images k with 2k + 1 < n_images (images listed in
 *   mirrored pairs, images[n_images - 1 - k] = -images[k]; the zero-image self pair is dropped);
 *   image may be NULL only with n_images = 1.
 * points != NULL: also keep only r >= 0 and d2 <= r^2, with d = ((double)x_i + images[k]) - (double)x_j
 *   per axis, d2 = (dx*dx + dy*dy) + dz*dz and r = ((double)radius[t_i * ntypes + t_j]
 *   + shift_i) + shift_j in IEEE double (images: n_images x 3 double; radius: ntypes x ntypes
 *   float; types, shift: per particle, NULL = type 0 / shift 0). */
int nb_rows_filter(const uint64_t* start, const uint32_t* counts, uint32_t Nq, const uint32_t* query_index,
                   uint32_t* neigh, uint8_t* image, uint32_t* counts_out, int half, uint32_t n_images,
                   const float* points, const double* images, const uint32_t* types, const float* shift,
                   const float* radius, uint32_t ntypes, void* stream);
/* Rows [start[q], start[q] + counts[q]) copied to [offsets[q], offsets[q] + counts[q]) (offsets
 * from nb_exclusive_scan_*: gap-free CSR rows, e.g. after nb_rows_filter), with first_out[e] =
 * particle i of the row of pair e (query_index NULL: i = q): first_out, neigh_out, image_out
 * are the (i, j, image) pair stream.  Outputs may be NULL and must not overlap the inputs.
 * (No reference counterpart.) */
int nb_rows_pack(const uint64_t* start, const uint32_t* counts, uint32_t Nq, const uint32_t* query_index,
                 const uint64_t* offsets, const uint32_t* neigh, const uint8_t* image, uint32_t* neigh_out,
                 uint8_t* image_out, uint32_t* first_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NEIGHBOR_AMD_H_ */
