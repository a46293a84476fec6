// capi.hip -- C ABI over the header-only template library (include/neighbor_amd.h).
//
// Each entry point instantiates the template API with the concrete ops named in
// its declaration and translates C++ exceptions into NB_ERR_* codes.
#include "neighbor_amd.h"

#include "neighbor/neighbor.h"
#include "neighbor/BatchPlan.h"
#include "neighbor/CapacityPredictor.h"
#include "neighbor/LBVHBatch.h"
#include "neighbor/kernels/Displacement.h"
#include "neighbor/kernels/RowFilter.h"
#include "neighbor/kernels/Scan.h"

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <memory>
#include <new>
#include <string>
#include <vector>

// Tile sums of nb_exclusive_scan_u32_u64_noalloc: one u64 per scan tile (2048 counts) plus the
// total, for at most 2^17 tiles (n <= 2^28 counts; 1 MiB per GPU, held until the process ends).
// Host code never passes `scratch` itself: each device's copy is reached through the address
// hipGetSymbolAddress returns with that device current (static_scan_scratch below).
static __device__ __align__(256) unsigned long long scratch[(1u << 17) + 1];

struct nb_lbvh
    {
    neighbor::LBVH lbvh;
    };

struct nb_capacity
    {
    neighbor::CapacityPredictor pred;
    };

struct nb_traverser
    {
    neighbor::LBVHTraverser trav;
    unsigned int point_stride = 3;     //!< floats per record of the exact calls' points
    };

struct nb_batch
    {
    neighbor::LBVHBatch batch;
    neighbor::BatchVariant variant = neighbor::BatchVariant::Wave;
    };

struct nb_batch_plan
    {
    std::unique_ptr<neighbor::BatchPlan> plan;
    };

namespace
{
thread_local std::string g_error;

int fail(int code, const std::string& msg)
    {
    g_error = msg;
    return code;
    }

template<class F>
int guarded(F&& f)
    {
    try
        {
        f();
        return NB_OK;
        }
    catch (const neighbor::tuning_error& e) { return fail(NB_ERR_TUNING, e.what()); }
    catch (const neighbor::image_limit_error& e) { return fail(NB_ERR_IMAGES, e.what()); }
    catch (const neighbor::hip_error& e)
        {
        return fail(e.code() == hipErrorOutOfMemory ? NB_ERR_ALLOC : NB_ERR_HIP, e.what());
        }
    catch (const std::bad_alloc& e) { return fail(NB_ERR_ALLOC, e.what()); }
    catch (const std::invalid_argument& e) { return fail(NB_ERR_INVALID, e.what()); }
    catch (const std::exception& e) { return fail(NB_ERR_INTERNAL, e.what()); }
    catch (...) { return fail(NB_ERR_INTERNAL, "unknown exception"); }
    }

typedef neighbor::LBVHTraverser::LaunchParameters TravParams;
typedef neighbor::LBVH::LaunchParameters BuildParams;

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

//! Dispatch the translate and transform ops of a traversal call.
template<class Real3, class QueryOpT, class OutputOpT>
void run_traverse(nb_traverser* t, const nb_lbvh* h, const QueryOpT& query, const OutputOpT& out,
                  const Real3* images, uint32_t nimg, const uint32_t* map, const uint32_t* order,
                  uint32_t block, void* stream)
    {
    neighbor::LBVHTraverser& tr = t->trav;
    tr.setQueryOrder(order, query.size());
    const TravParams p(block, as_stream(stream));
    if (images)
        {
        const neighbor::ImageListOp<Real3> img(images, nimg);
        if (map) tr.traverse(p, h->lbvh, query, out, img, neighbor::MapTransformOp(map));
        else tr.traverse(p, h->lbvh, query, out, img, neighbor::NullTransformOp());
        }
    else
        {
        if (map) tr.traverse(p, h->lbvh, query, out, neighbor::SelfOp(), neighbor::MapTransformOp(map));
        else tr.traverse(p, h->lbvh, query, out, neighbor::SelfOp(), neighbor::NullTransformOp());
        }
    tr.setQueryOrder(nullptr, 0);
    }

int check_traverse_args(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const void* images, uint32_t Nimg)
    {
    if (!t || !h) return fail(NB_ERR_INVALID, "null handle");
    if (Nq > 0 && !spheres) return fail(NB_ERR_INVALID, "null spheres");
    if (images && Nimg > 32) return fail(NB_ERR_IMAGES, "A maximum of 32 image vectors are supported by LBVH traversers.");
    return NB_OK;
    }

constexpr uint32_t kStaticScanTiles = 1u << 17;
static_assert(sizeof(scratch) == (kStaticScanTiles + 1) * sizeof(unsigned long long), "scratch size");
constexpr uint64_t kStaticScanMaxN = static_cast<uint64_t>(kStaticScanTiles) * neighbor::gpu::scan::kTile;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

//! Device address of the current device's `scratch`, looked up on the first call per device.
int static_scan_scratch(void** out)
    {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(NB_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(NB_ERR_INVALID, "device ordinal out of range for the static scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev])
        {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
        }
    *out = g_scratch_ptr[dev];
    return NB_OK;
    }
} // end anonymous namespace

extern "C" {

const char* nb_last_error_string(void) { return g_error.c_str(); }

int nb_version(void) { return 100; }

int nb_lbvh_create(nb_lbvh** out)
    {
    if (!out) return fail(NB_ERR_INVALID, "null output");
    return guarded([&] { *out = new nb_lbvh(); });
    }

int nb_lbvh_destroy(nb_lbvh* lbvh)
    {
    return guarded([&] { delete lbvh; });
    }

int nb_lbvh_build_points(nb_lbvh* h, const float* points, uint32_t N, const float* lo, const float* hi,
                         uint32_t block, void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    if (N > 0 && (!points || !lo || !hi)) return fail(NB_ERR_INVALID, "null points or bounds");
    if (N >= 0x40000000u) return fail(NB_ERR_INVALID, "N must be below 2^30");
    return guarded([&] {
        const neighbor::PointInsertOp insert(reinterpret_cast<const float3*>(points), N);
        const float3 l = lo ? make_float3(lo[0], lo[1], lo[2]) : make_float3(0.f, 0.f, 0.f);
        const float3 u = hi ? make_float3(hi[0], hi[1], hi[2]) : make_float3(0.f, 0.f, 0.f);
        h->lbvh.build(BuildParams(block, as_stream(stream)), insert, l, u);
    });
    }

int nb_lbvh_build_spheres(nb_lbvh* h, const float* points, uint32_t N, float radius, const float* lo,
                          const float* hi, uint32_t block, void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    if (N > 0 && (!points || !lo || !hi)) return fail(NB_ERR_INVALID, "null points or bounds");
    if (N >= 0x40000000u) return fail(NB_ERR_INVALID, "N must be below 2^30");
    return guarded([&] {
        const neighbor::SphereInsertOp insert(reinterpret_cast<const float3*>(points), radius, N);
        const float3 l = lo ? make_float3(lo[0], lo[1], lo[2]) : make_float3(0.f, 0.f, 0.f);
        const float3 u = hi ? make_float3(hi[0], hi[1], hi[2]) : make_float3(0.f, 0.f, 0.f);
        h->lbvh.build(BuildParams(block, as_stream(stream)), insert, l, u);
    });
    }

int nb_lbvh_build_postype(nb_lbvh* h, const float* postype, const uint32_t* index, uint32_t N, float radius,
                          const float* lo, const float* hi, uint32_t block, void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    if (N > 0 && (!postype || !lo || !hi)) return fail(NB_ERR_INVALID, "null records or bounds");
    if (N >= 0x40000000u) return fail(NB_ERR_INVALID, "N must be below 2^30");
    return guarded([&] {
        const neighbor::PostypeInsertOp insert(reinterpret_cast<const float4*>(postype), index, radius < 0.0f ? -1.0f : radius, N);
        const float3 l = lo ? make_float3(lo[0], lo[1], lo[2]) : make_float3(0.f, 0.f, 0.f);
        const float3 u = hi ? make_float3(hi[0], hi[1], hi[2]) : make_float3(0.f, 0.f, 0.f);
        h->lbvh.build(BuildParams(block, as_stream(stream)), insert, l, u);
    });
    }

int nb_lbvh_refit_postype(nb_lbvh* h, const float* postype, const uint32_t* index, uint32_t N, float radius, uint32_t block,
                          void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    if (N > 0 && !postype) return fail(NB_ERR_INVALID, "null records");
    if (N != h->lbvh.getN()) return fail(NB_ERR_INVALID, "refit needs the primitive count of the last build");
    return guarded([&] {
        const neighbor::PostypeInsertOp insert(reinterpret_cast<const float4*>(postype), index, radius < 0.0f ? -1.0f : radius, N);
        h->lbvh.refit(BuildParams(block, as_stream(stream)), insert);
    });
    }

int nb_lbvh_refit_points(nb_lbvh* h, const float* points, uint32_t N, uint32_t block, void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    if (N > 0 && !points) return fail(NB_ERR_INVALID, "null points");
    if (N != h->lbvh.getN()) return fail(NB_ERR_INVALID, "refit needs the primitive count of the last build");
    return guarded([&] {
        const neighbor::PointInsertOp insert(reinterpret_cast<const float3*>(points), N);
        h->lbvh.refit(BuildParams(block, as_stream(stream)), insert);
    });
    }

int nb_lbvh_refit_spheres(nb_lbvh* h, const float* points, uint32_t N, float radius, uint32_t block, void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    if (N > 0 && !points) return fail(NB_ERR_INVALID, "null points");
    if (N != h->lbvh.getN()) return fail(NB_ERR_INVALID, "refit needs the primitive count of the last build");
    return guarded([&] {
        const neighbor::SphereInsertOp insert(reinterpret_cast<const float3*>(points), radius, N);
        h->lbvh.refit(BuildParams(block, as_stream(stream)), insert);
    });
    }

int nb_max_displacement2(const float* points, const float* ref, uint32_t N, const float* box, double* out,
                         uint32_t* wrapped, void* stream)
    {
    if (!out || (N > 0 && (!points || !ref))) return fail(NB_ERR_INVALID, "null pointer");
    return guarded([&] {
        const double3 L = box ? make_double3(box[0], box[1], box[2]) : make_double3(0.0, 0.0, 0.0);
        neighbor::gpu::max_displacement2(reinterpret_cast<const float3*>(points), reinterpret_cast<const float3*>(ref), N,
                                         box != nullptr, L, out, wrapped, as_stream(stream));
    });
    }

int nb_lbvh_info(const nb_lbvh* h, uint32_t* N, uint32_t* n_internal, uint32_t* n_nodes, int32_t* root)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    if (N) *N = h->lbvh.getN();
    if (n_internal) *n_internal = h->lbvh.getNInternal();
    if (n_nodes) *n_nodes = h->lbvh.getNNodes();
    if (root) *root = h->lbvh.getRoot();
    return NB_OK;
    }

int nb_lbvh_arrays(const nb_lbvh* h, const int32_t** parent, const int32_t** left, const int32_t** right,
                   const uint32_t** primitive, const float** lo, const float** hi, const int32_t** rope,
                   const uint32_t** codes)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    const neighbor::LBVH& b = h->lbvh;
    if (parent) *parent = b.getParents().get();
    if (left) *left = b.getLeftChildren().get();
    if (right) *right = b.getRightChildren().get();
    if (primitive) *primitive = b.getPrimitives().get();
    if (lo) *lo = reinterpret_cast<const float*>(b.getLowerBounds().get());
    if (hi) *hi = reinterpret_cast<const float*>(b.getUpperBounds().get());
    if (rope) *rope = b.getRopes().get();
    if (codes) *codes = b.getCodes().get();
    return NB_OK;
    }

int nb_lbvh_tunable_parameters(const nb_lbvh* h, uint32_t* out, uint32_t capacity, uint32_t* count)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    const std::vector<unsigned int> v = h->lbvh.getTunableParameters();
    if (count) *count = static_cast<uint32_t>(v.size());
    for (uint32_t i = 0; out && i < capacity && i < v.size(); ++i) out[i] = v[i];
    return NB_OK;
    }

int nb_traverser_create(nb_traverser** out)
    {
    if (!out) return fail(NB_ERR_INVALID, "null output");
    return guarded([&] { *out = new nb_traverser(); });
    }

int nb_traverser_destroy(nb_traverser* t)
    {
    return guarded([&] { delete t; });
    }

int nb_traverser_set_variant(nb_traverser* t, int variant)
    {
    if (!t) return fail(NB_ERR_INVALID, "null handle");
    if (variant < NB_VARIANT_QUANTIZED || variant > NB_VARIANT_WAVE_QUANTIZED) return fail(NB_ERR_INVALID, "unknown variant");
    t->trav.setVariant(static_cast<neighbor::TraversalVariant>(variant));
    return NB_OK;
    }

int nb_traverser_set_bucket_size(nb_traverser* t, uint32_t bucket)
    {
    if (!t) return fail(NB_ERR_INVALID, "null handle");
    if (bucket < 1 || bucket > 64) return fail(NB_ERR_INVALID, "bucket size must be in [1, 64]");
    t->trav.setBucketSize(bucket);
    return NB_OK;
    }

int nb_traverser_set_tail_split(nb_traverser* t, uint32_t groups, uint32_t split)
    {
    if (!t) return fail(NB_ERR_INVALID, "null handle");
    if (split != 1 && split != 2 && split != 4) return fail(NB_ERR_INVALID, "the tail split must be 1, 2 or 4");
    t->trav.setTailSplit(groups, split);
    return NB_OK;
    }

int nb_traverser_set_point_stride(nb_traverser* t, uint32_t stride)
    {
    if (!t) return fail(NB_ERR_INVALID, "null handle");
    if (stride != 3 && stride != 4) return fail(NB_ERR_INVALID, "point records must hold 3 or 4 floats");
    t->point_stride = stride;
    return NB_OK;
    }

int nb_traverser_set_visit_log(nb_traverser* t, uint32_t* log, uint64_t words)
    {
    if (!t) return fail(NB_ERR_INVALID, "null handle");
    t->trav.setVisitLog(log, log ? static_cast<size_t>(words) : 0);
    return NB_OK;
    }

size_t nb_visit_log_words(uint32_t n_queries) { return neighbor::LBVHTraverser::visitLogWords(n_queries); }

int nb_traverser_setup(nb_traverser* t, const nb_lbvh* h, const uint32_t* map, uint32_t block, void* stream)
    {
    if (!t || !h) return fail(NB_ERR_INVALID, "null handle");
    return guarded([&] {
        const TravParams p(block, as_stream(stream));
        if (map) t->trav.setup(p, h->lbvh, neighbor::MapTransformOp(map));
        else t->trav.setup(p, h->lbvh, neighbor::NullTransformOp());
    });
    }

int nb_traverser_reset(nb_traverser* t)
    {
    if (!t) return fail(NB_ERR_INVALID, "null handle");
    t->trav.reset();
    return NB_OK;
    }

int nb_traverser_data(const nb_traverser* t, const int32_t** records, const float** root_lo, const float** root_hi,
                      const float** bins, const float** float_records)
    {
    if (!t) return fail(NB_ERR_INVALID, "null handle");
    const neighbor::LBVHTraverser& tr = t->trav;
    if (records) *records = reinterpret_cast<const int32_t*>(tr.getData().get());
    if (root_lo) *root_lo = reinterpret_cast<const float*>(tr.getRootLower().get());
    if (root_hi) *root_hi = reinterpret_cast<const float*>(tr.getRootUpper().get());
    if (bins) *bins = reinterpret_cast<const float*>(tr.getBins().get());
    if (float_records) *float_records = reinterpret_cast<const float*>(tr.getFloatRopeData().get());
    return NB_OK;
    }

int nb_traverse_count(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const float* images,
                      uint32_t Nimg, const uint32_t* map, const uint32_t* order, uint32_t* nneigh, uint32_t block,
                      void* stream)
    {
    int rc = check_traverse_args(t, h, spheres, Nq, images, Nimg);
    if (rc) return rc;
    if (Nq > 0 && !nneigh) return fail(NB_ERR_INVALID, "null nneigh");
    return guarded([&] {
        const neighbor::SphereQueryOp q(reinterpret_cast<const float4*>(spheres), Nq);
        const neighbor::CountNeighborsOp out(nneigh);
        run_traverse(t, h, q, out, reinterpret_cast<const float3*>(images), Nimg, map, order, block, stream);
    });
    }

int nb_traverse_list(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const float* images,
                     uint32_t Nimg, const uint32_t* map, const uint32_t* order, uint32_t* neigh_list, uint32_t* nneigh,
                     uint32_t max_neigh, uint32_t block, void* stream)
    {
    int rc = check_traverse_args(t, h, spheres, Nq, images, Nimg);
    if (rc) return rc;
    if (Nq > 0 && (!nneigh || (max_neigh > 0 && !neigh_list))) return fail(NB_ERR_INVALID, "null output");
    return guarded([&] {
        const neighbor::SphereQueryOp q(reinterpret_cast<const float4*>(spheres), Nq);
        const neighbor::NeighborListOp out(neigh_list, nneigh, max_neigh);
        run_traverse(t, h, q, out, reinterpret_cast<const float3*>(images), Nimg, map, order, block, stream);
    });
    }

int nb_traverse_pairs(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const float* images,
                      uint32_t Nimg, const uint32_t* map, const uint32_t* order, const uint64_t* offsets,
                      uint32_t* neigh, uint8_t* image, uint32_t* nneigh, uint64_t capacity, uint32_t block, void* stream)
    {
    int rc = check_traverse_args(t, h, spheres, Nq, images, Nimg);
    if (rc) return rc;
    if (Nq > 0 && (!offsets || !neigh)) return fail(NB_ERR_INVALID, "null output");
    return guarded([&] {
        const neighbor::SphereQueryOp q(reinterpret_cast<const float4*>(spheres), Nq);
        const neighbor::PairListOp out(reinterpret_cast<const unsigned long long*>(offsets), neigh, image, nneigh, capacity);
        run_traverse(t, h, q, out, reinterpret_cast<const float3*>(images), Nimg, map, order, block, stream);
    });
    }

int nb_traverse_exact_count(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const float* points,
                            const double* images, uint32_t Nimg, const uint32_t* map, const uint32_t* order,
                            uint32_t* nneigh, uint32_t block, void* stream)
    {
    int rc = check_traverse_args(t, h, spheres, Nq, images, Nimg);
    if (rc) return rc;
    if (Nq > 0 && (!nneigh || !points)) return fail(NB_ERR_INVALID, "null points or nneigh");
    return guarded([&] {
        const neighbor::ExactSphereQueryOp q(reinterpret_cast<const float4*>(spheres), Nq, points, t->point_stride);
        const neighbor::CountNeighborsOp out(nneigh);
        run_traverse(t, h, q, out, reinterpret_cast<const double3*>(images), Nimg, map, order, block, stream);
    });
    }

int nb_traverse_exact_pairs(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const float* points,
                            const double* images, uint32_t Nimg, const uint32_t* map, const uint32_t* order,
                            const uint64_t* offsets, uint32_t* neigh, uint8_t* image, uint32_t* nneigh,
                            uint64_t capacity, uint32_t block, void* stream)
    {
    int rc = check_traverse_args(t, h, spheres, Nq, images, Nimg);
    if (rc) return rc;
    if (Nq > 0 && (!offsets || !neigh || !points)) return fail(NB_ERR_INVALID, "null output or points");
    return guarded([&] {
        const neighbor::ExactSphereQueryOp q(reinterpret_cast<const float4*>(spheres), Nq, points, t->point_stride);
        const neighbor::PairListOp out(reinterpret_cast<const unsigned long long*>(offsets), neigh, image, nneigh, capacity);
        run_traverse(t, h, q, out, reinterpret_cast<const double3*>(images), Nimg, map, order, block, stream);
    });
    }

int nb_traverse_exact_rows(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const float* points,
                           const double* images, uint32_t Nimg, const uint32_t* map, const uint32_t* order,
                           uint64_t* start, uint32_t* limit, uint32_t* neigh, uint8_t* image, uint32_t* nneigh,
                           uint64_t capacity, uint64_t* spill, uint32_t block, void* stream)
    {
    int rc = check_traverse_args(t, h, spheres, Nq, images, Nimg);
    if (rc) return rc;
    if (Nq > 0 && (!start || !limit || !neigh || !nneigh || !points)) return fail(NB_ERR_INVALID, "null rows, output or points");
    return guarded([&] {
        const neighbor::ExactSphereQueryOp q(reinterpret_cast<const float4*>(spheres), Nq, points, t->point_stride);
        const neighbor::RowPairOp out(reinterpret_cast<unsigned long long*>(start), limit, neigh, image, nneigh, capacity,
                                      reinterpret_cast<unsigned long long*>(spill));
        run_traverse(t, h, q, out, reinterpret_cast<const double3*>(images), Nimg, map, order, block, stream);
    });
    }

int nb_traverse_exact_rows_placed(nb_traverser* t, const nb_lbvh* h, const float* spheres, uint32_t Nq, const float* points,
                                const double* images, uint32_t Nimg, const uint32_t* map, const uint64_t* base,
                                uint64_t* start, uint32_t* limit, uint32_t* neigh, uint8_t* image, uint32_t* nneigh,
                                uint64_t capacity, uint64_t* spill, void* stream)
    {
    int rc = check_traverse_args(t, h, spheres, Nq, images, Nimg);
    if (rc) return rc;
    if (Nq != h->lbvh.getN()) return fail(NB_ERR_INVALID, "wave-placed rows serve the tree's own primitives (Nq == N)");
    if (Nq > 0 && (!base || !start || !limit || !neigh || !nneigh || !points))
        return fail(NB_ERR_INVALID, "null rows, output or points");
    if (t->trav.getVariant() != neighbor::TraversalVariant::Wave && t->trav.getVariant() != neighbor::TraversalVariant::WaveQuantized)
        return fail(NB_ERR_INVALID, "wave-placed rows need a wave traversal variant");
    return guarded([&] {
        const neighbor::ExactSphereQueryOp q(reinterpret_cast<const float4*>(spheres), Nq, points, t->point_stride);
        const neighbor::PlacedRowPairOp out(reinterpret_cast<const unsigned long long*>(base),
                                          reinterpret_cast<unsigned long long*>(start), limit, neigh, image, nneigh, capacity,
                                          reinterpret_cast<unsigned long long*>(spill));
        run_traverse(t, h, q, out, reinterpret_cast<const double3*>(images), Nimg, map, h->lbvh.getPrimitives().get(), 256u,
                     stream);
    });
    }

int nb_capacity_create(nb_capacity** out, uint64_t seed)
    {
    if (!out) return fail(NB_ERR_INVALID, "null output");
    return guarded([&] { *out = new nb_capacity{neighbor::CapacityPredictor(seed)}; });
    }

int nb_capacity_destroy(nb_capacity* p)
    {
    return guarded([&] { delete p; });
    }

int nb_capacity_parameters(nb_capacity* p, float* params_out, const float* params_in, uint32_t* count)
    {
    if (!p) return fail(NB_ERR_INVALID, "null handle");
    if (count) *count = neighbor::CapacityPredictor::kParams;
    return guarded([&] {
        if (params_out)
            {
            const std::vector<float> v = p->pred.getParameters();
            std::copy(v.begin(), v.end(), params_out);
            }
        if (params_in) p->pred.setParameters(std::vector<float>(params_in, params_in + neighbor::CapacityPredictor::kParams));
    });
    }

int nb_capacity_features(nb_capacity* p, const nb_lbvh* h, const float* lo, const float* hi, float r, const float** features,
                         void* stream)
    {
    if (!p || !h || !lo || !hi) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        p->pred.computeFeatures(h->lbvh, make_float3(lo[0], lo[1], lo[2]), make_float3(hi[0], hi[1], hi[2]), r, as_stream(stream));
        if (features) *features = p->pred.features();
    });
    }

int nb_capacity_train(nb_capacity* p, const nb_lbvh* h, const uint32_t* counts, uint32_t steps, uint32_t batch, float lr,
                      uint64_t seed, float* loss, void* stream)
    {
    if (!p || !h || !counts) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        const float l = p->pred.train(h->lbvh, counts, steps, batch, lr, seed, as_stream(stream));
        if (loss) *loss = l;
    });
    }

int nb_capacity_margin(nb_capacity* p, const nb_lbvh* h, float quantile, float* margin, void* stream)
    {
    if (!p || !h || !margin) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] { *margin = p->pred.calibrateMargin(h->lbvh, quantile, as_stream(stream)); });
    }

int nb_capacity_predict(nb_capacity* p, const nb_lbvh* h, float margin, uint32_t pad, uint32_t* cap, void* stream)
    {
    if (!p || !h || !cap) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] { p->pred.predictCapacity(h->lbvh, margin, pad, cap, as_stream(stream)); });
    }

int nb_capacity_predict_direct(nb_capacity* p, const nb_lbvh* h, const float* lo, const float* hi, float r, float margin,
                               uint32_t pad, uint32_t* cap, void* stream)
    {
    if (!p || !h || !lo || !hi || !cap) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        p->pred.predictCapacityDirect(h->lbvh, make_float3(lo[0], lo[1], lo[2]), make_float3(hi[0], hi[1], hi[2]), r, margin,
                                      pad, cap, as_stream(stream));
    });
    }

int nb_capacity_predict_rows(nb_capacity* p, const nb_lbvh* h, const float* lo, const float* hi, float r, float margin,
                             uint32_t pad, uint32_t* cap, uint64_t* base, uint64_t* spill, void* stream)
    {
    if (!p || !h || !lo || !hi || !cap || !base) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        p->pred.predictRows(h->lbvh, make_float3(lo[0], lo[1], lo[2]), make_float3(hi[0], hi[1], hi[2]), r, margin, pad, cap,
                            reinterpret_cast<unsigned long long*>(base), reinterpret_cast<unsigned long long*>(spill),
                            as_stream(stream));
    });
    }

int nb_capacity_forward(const float* params, const float* features, uint32_t n, float* y, void* stream)
    {
    if (!params || (n > 0 && (!features || !y))) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        if (n > 0)
            hipLaunchKernelGGL(neighbor::gpu::capacity::kernel::predict, dim3(neighbor::div_up(n, 256)), dim3(256), 0,
                               as_stream(stream), params, features, n, y);
        neighbor::check_launch("capacity forward");
    });
    }

int nb_capacity_gradient(const float* params, const float* features, const float* targets, uint32_t batch, float* grad,
                         void* stream)
    {
    if (!params || !features || !targets || !grad) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        namespace cap = neighbor::gpu::capacity;
        hipLaunchKernelGGL(neighbor::gpu::radix::kernel::fill_zero, dim3(1), dim3(256), 0, as_stream(stream),
                           reinterpret_cast<uint32_t*>(grad), static_cast<size_t>(cap::kParams + 1));
        if (batch > 0)
            hipLaunchKernelGGL(cap::kernel::grad, dim3(neighbor::div_up(batch, 256)), dim3(256), 0, as_stream(stream), params,
                               features, targets, nullptr, batch, grad);
        neighbor::check_launch("capacity gradient");
    });
    }

int nb_overflow_rows(const uint32_t* counts, const uint32_t* cap, uint32_t n, uint32_t* list, uint32_t* n_list, void* stream)
    {
    if ((n > 0 && (!counts || !cap || !list)) || !n_list) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        if (n > 0)
            hipLaunchKernelGGL(neighbor::gpu::capacity::kernel::overflow_list, dim3(neighbor::div_up(n, 256)), dim3(256), 0,
                               as_stream(stream), counts, cap, n, list, n_list);
        neighbor::check_launch("overflow rows");
    });
    }

size_t nb_relocate_scratch_bytes(uint32_t n_list)
    {
    const size_t a = ((static_cast<size_t>(n_list) * 4 + 255) / 256) * 256;
    const size_t b = ((static_cast<size_t>(n_list + 1) * 8 + 255) / 256) * 256;
    return a + b + neighbor::gpu::scan::scratch_bytes(n_list);
    }

int nb_relocate_rows(const uint32_t* list, uint32_t n_list, const uint32_t* counts, uint64_t tail, uint64_t* start,
                     uint32_t* limit, void* scratch, void* stream)
    {
    if (n_list == 0) return NB_OK;
    if (!list || !counts || !start || !limit || !scratch) return fail(NB_ERR_INVALID, "null argument");
    return guarded([&] {
        namespace cap = neighbor::gpu::capacity;
        char* s = static_cast<char*>(scratch);
        uint32_t* gathered = reinterpret_cast<uint32_t*>(s);
        unsigned long long* offs = reinterpret_cast<unsigned long long*>(s + ((static_cast<size_t>(n_list) * 4 + 255) / 256) * 256);
        void* scan_scratch = s + ((static_cast<size_t>(n_list) * 4 + 255) / 256) * 256 + ((static_cast<size_t>(n_list + 1) * 8 + 255) / 256) * 256;
        const hipStream_t st = as_stream(stream);
        hipLaunchKernelGGL(cap::kernel::gather_counts, dim3(neighbor::div_up(n_list, 256)), dim3(256), 0, st, list, n_list, counts, gathered);
        neighbor::gpu::scan::exclusive_scan(gathered, n_list, offs, scan_scratch, st);
        hipLaunchKernelGGL(cap::kernel::relocate, dim3(neighbor::div_up(n_list, 256)), dim3(256), 0, st, list, n_list, offs, counts,
                           static_cast<unsigned long long>(tail), reinterpret_cast<unsigned long long*>(start), limit);
        neighbor::check_launch("relocate rows");
    });
    }

namespace
{
neighbor::BatchProblem batch_problem(const float* points, uint32_t N, const uint32_t* offsets, const float* box_lo,
                                     const float* box_hi, uint32_t nbox, const float* spheres, int periodic)
    {
    return neighbor::BatchProblem{reinterpret_cast<const float3*>(points), N, offsets, reinterpret_cast<const float3*>(box_lo),
                                  reinterpret_cast<const float3*>(box_hi), nbox, reinterpret_cast<const float4*>(spheres),
                                  periodic != 0};
    }

int check_batch_inputs(const float* points, uint32_t N, const uint32_t* offsets, const float* box_lo, const float* box_hi,
                       uint32_t nbox)
    {
    if (nbox == 0) return fail(NB_ERR_INVALID, "a batch needs at least one box");
    if (!offsets || !box_lo || !box_hi || (N > 0 && !points)) return fail(NB_ERR_INVALID, "null batch input");
    if (N >= 0x40000000u) return fail(NB_ERR_INVALID, "N must be below 2^30");
    if (nbox >= (1u << 22)) return fail(NB_ERR_INVALID, "at most 2^22 boxes per batch");
    return NB_OK;
    }
} // end anonymous namespace

int nb_lbvh_error_word(const nb_lbvh* h, const uint32_t** word)
    {
    if (!h || !word) return fail(NB_ERR_INVALID, "null argument");
    *word = h->lbvh.getSortErrorFlag();
    return NB_OK;
    }

int nb_lbvh_enable_sorted_event(nb_lbvh* h, int on)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    return guarded([&] { h->lbvh.enableSortedEvent(on != 0); });
    }

int nb_lbvh_set_span_lds_limit(nb_lbvh* h, uint32_t limit)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    return guarded([&] { h->lbvh.setSpanLdsLimit(limit); });
    }

int nb_lbvh_wait_sorted(const nb_lbvh* h, void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    return guarded([&] { h->lbvh.waitSorted(as_stream(stream)); });
    }

int nb_lbvh_check(const nb_lbvh* h, void* stream)
    {
    if (!h) return fail(NB_ERR_INVALID, "null handle");
    try
        {
        h->lbvh.checkBuild(as_stream(stream));
        }
    catch (const neighbor::hip_error& e) { return fail(NB_ERR_HIP, e.what()); }
    catch (const std::exception& e) { return fail(NB_ERR_INTERNAL, e.what()); }
    return NB_OK;
    }

int nb_batch_check(const nb_batch* b, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    try
        {
        b->batch.checkBuild(as_stream(stream));
        }
    catch (const neighbor::hip_error& e) { return fail(NB_ERR_HIP, e.what()); }
    catch (const std::exception& e) { return fail(NB_ERR_INTERNAL, e.what()); }
    return NB_OK;
    }

int nb_batch_error_word(const nb_batch* b, const uint32_t** word)
    {
    if (!b || !word) return fail(NB_ERR_INVALID, "null argument");
    *word = b->batch.getSortErrorFlag();
    return NB_OK;
    }

int nb_batch_create(nb_batch** out)
    {
    if (!out) return fail(NB_ERR_INVALID, "null output");
    return guarded([&] { *out = new nb_batch(); });
    }

int nb_batch_destroy(nb_batch* b)
    {
    return guarded([&] { delete b; });
    }

int nb_batch_build(nb_batch* b, const float* points, uint32_t N, const uint32_t* offsets, const float* box_lo,
                   const float* box_hi, uint32_t nbox, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    const int rc = check_batch_inputs(points, N, offsets, box_lo, box_hi, nbox);
    if (rc) return rc;
    return guarded([&] {
        b->batch.build(reinterpret_cast<const float3*>(points), N, offsets, reinterpret_cast<const float3*>(box_lo),
                       reinterpret_cast<const float3*>(box_hi), nbox, as_stream(stream));
    });
    }

int nb_batch_set_params(nb_batch* b, uint32_t block, uint32_t bucket, int variant, uint32_t grid)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    if (block < 32 || block > 1024 || block % 32) return fail(NB_ERR_TUNING, "Invalid block size.");
    if (bucket < 1 || bucket > 64) return fail(NB_ERR_INVALID, "bucket size must be in [1, 64]");
    if (variant != NB_BATCH_WAVE && variant != NB_BATCH_LANE) return fail(NB_ERR_INVALID, "unknown batch variant");
    return guarded([&] {
        b->batch.setBlock(block);
        b->batch.setBucketOnly(bucket);
        b->variant = variant == NB_BATCH_LANE ? neighbor::BatchVariant::Lane : neighbor::BatchVariant::Wave;
        b->batch.setGrid(grid);
    });
    }

int nb_batch_get_params(const nb_batch* b, uint32_t* block, uint32_t* bucket, int* variant, uint32_t* grid)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    return guarded([&] {
        if (block) *block = b->batch.getBlock();
        if (bucket) *bucket = b->batch.getBucket();
        if (variant) *variant = b->variant == neighbor::BatchVariant::Lane ? NB_BATCH_LANE : NB_BATCH_WAVE;
        if (grid) *grid = b->batch.gridSize();
    });
    }

int nb_batch_setup(nb_batch* b, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    return guarded([&] { b->batch.setupTraversal(b->batch.getBucket(), as_stream(stream)); });
    }

int nb_batch_count(nb_batch* b, const float* spheres, int periodic, uint32_t* nneigh, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    if (b->batch.getN() > 0 && (!spheres || !nneigh)) return fail(NB_ERR_INVALID, "null spheres or output");
    return guarded([&] {
        b->batch.traverse(neighbor::CountNeighborsOp(nneigh), reinterpret_cast<const float4*>(spheres), periodic != 0,
                          b->variant, as_stream(stream));
    });
    }

int nb_batch_pairs(nb_batch* b, const float* spheres, int periodic, const uint64_t* offsets, uint32_t* neigh,
                   uint8_t* image, uint32_t* nneigh, uint64_t capacity, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    if (b->batch.getN() > 0 && (!spheres || !offsets || !neigh)) return fail(NB_ERR_INVALID, "null spheres, offsets or output");
    return guarded([&] {
        b->batch.traverse(neighbor::PairListOp(reinterpret_cast<const unsigned long long*>(offsets), neigh, image, nneigh, capacity),
                          reinterpret_cast<const float4*>(spheres), periodic != 0, b->variant, as_stream(stream));
    });
    }

int nb_batch_rows(nb_batch* b, const float* spheres, int periodic, uint64_t* start, uint32_t* limit, uint32_t* neigh,
                  uint8_t* image, uint32_t* nneigh, uint64_t capacity, uint64_t* spill, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    if (b->batch.getN() > 0 && (!spheres || !start || !limit || !neigh || !nneigh))
        return fail(NB_ERR_INVALID, "null spheres, rows or output");
    return guarded([&] {
        b->batch.traverse(neighbor::RowPairOp(reinterpret_cast<unsigned long long*>(start), limit, neigh, image, nneigh, capacity,
                                              reinterpret_cast<unsigned long long*>(spill)),
                          reinterpret_cast<const float4*>(spheres), periodic != 0, b->variant, as_stream(stream));
    });
    }

int nb_batch_capacity(nb_batch* b, const float* spheres, float z, uint32_t pad, uint32_t* cap, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    if (b->batch.getN() > 0 && (!spheres || !cap)) return fail(NB_ERR_INVALID, "null spheres or output");
    return guarded([&] { b->batch.capacity(reinterpret_cast<const float4*>(spheres), z, pad, cap, as_stream(stream)); });
    }

int nb_batch_arrays(const nb_batch* b, const int** roots, const uint32_t** primitive, const uint64_t** keys,
                    const int** parent, const int** left, const int** right, const int** rope, const float** lo,
                    const float** hi, uint32_t* N, uint32_t* nbox, uint32_t* n_nodes)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    const neighbor::LBVHBatch& t = b->batch;
    if (roots) *roots = t.getRoots().get();
    if (primitive) *primitive = t.getPrimitives().get();
    if (keys) *keys = reinterpret_cast<const uint64_t*>(t.getKeys().get());
    if (parent) *parent = t.getParents().get();
    if (left) *left = t.getLeftChildren().get();
    if (right) *right = t.getRightChildren().get();
    if (rope) *rope = t.getRopes().get();
    if (lo) *lo = reinterpret_cast<const float*>(t.getLowerBounds().get());
    if (hi) *hi = reinterpret_cast<const float*>(t.getUpperBounds().get());
    if (N) *N = t.getN();
    if (nbox) *nbox = t.getNBoxes();
    if (n_nodes) *n_nodes = t.getNNodes();
    return NB_OK;
    }

int nb_batch_tune(nb_batch* b, const float* points, uint32_t N, const uint32_t* offsets, const float* box_lo,
                  const float* box_hi, uint32_t nbox, const float* spheres, int periodic, uint32_t* nneigh,
                  uint32_t* block, uint32_t* bucket, int* variant, float* build_ms, float* count_ms, void* stream)
    {
    if (!b) return fail(NB_ERR_INVALID, "null handle");
    const int rc = check_batch_inputs(points, N, offsets, box_lo, box_hi, nbox);
    if (rc) return rc;
    if (N > 0 && (!spheres || !nneigh)) return fail(NB_ERR_INVALID, "null spheres or scratch");
    return guarded([&] {
        neighbor::BatchTuner tuner;
        const neighbor::BatchChoice c =
            tuner.tune(b->batch, batch_problem(points, N, offsets, box_lo, box_hi, nbox, spheres, periodic), nneigh, as_stream(stream));
        b->variant = c.variant;
        if (block) *block = c.block;
        if (bucket) *bucket = c.bucket;
        if (variant) *variant = c.variant == neighbor::BatchVariant::Lane ? NB_BATCH_LANE : NB_BATCH_WAVE;
        if (build_ms) *build_ms = c.build_ms;
        if (count_ms) *count_ms = c.count_ms;
    });
    }

int nb_batch_plan_create(nb_batch_plan** out, nb_batch* b, const float* points, uint32_t N, const uint32_t* offsets,
                         const float* box_lo, const float* box_hi, uint32_t nbox, const float* spheres, int periodic,
                         uint32_t* nneigh, uint64_t* pair_offsets, void* scan_scratch, uint32_t* neigh, uint8_t* image,
                         uint64_t capacity)
    {
    if (!out || !b) return fail(NB_ERR_INVALID, "null handle");
    const int rc = check_batch_inputs(points, N, offsets, box_lo, box_hi, nbox);
    if (rc) return rc;
    if (N > 0 && (!spheres || !nneigh || !pair_offsets || !scan_scratch || !neigh))
        return fail(NB_ERR_INVALID, "null plan buffer");
    return guarded([&] {
        std::unique_ptr<nb_batch_plan> p(new nb_batch_plan());
        p->plan.reset(new neighbor::BatchPlan(b->batch, batch_problem(points, N, offsets, box_lo, box_hi, nbox, spheres, periodic),
                                              b->variant, nneigh, reinterpret_cast<unsigned long long*>(pair_offsets),
                                              scan_scratch, neigh, image, capacity));
        *out = p.release();
    });
    }

int nb_batch_plan_launch(nb_batch_plan* p, void* stream)
    {
    if (!p || !p->plan) return fail(NB_ERR_INVALID, "null plan");
    return guarded([&] { p->plan->launch(as_stream(stream)); });
    }

int nb_batch_plan_destroy(nb_batch_plan* p)
    {
    return guarded([&] { delete p; });
    }

size_t nb_scan_scratch_bytes(uint32_t n) { return neighbor::gpu::scan::scratch_bytes(n); }

int nb_exclusive_scan_permuted_u32_u64(const uint32_t* counts, const uint32_t* perm, uint32_t n, uint64_t* offsets,
                                       void* scratch, void* stream)
    {
    if (!offsets || !scratch || (n > 0 && (!counts || !perm))) return fail(NB_ERR_INVALID, "null pointer");
    return guarded([&] {
        neighbor::gpu::scan::exclusive_scan(counts, n, reinterpret_cast<unsigned long long*>(offsets), scratch, as_stream(stream),
                                            perm);
    });
    }

int nb_rows_filter(const uint64_t* start, const uint32_t* counts, uint32_t Nq, const uint32_t* query_index,
                   uint32_t* neigh, uint8_t* image, uint32_t* counts_out, int half, uint32_t n_images,
                   const float* points, const double* images, const uint32_t* types, const float* shift,
                   const float* radius, uint32_t ntypes, void* stream)
    {
    if (Nq > 0 && (!start || !counts || !neigh || !counts_out)) return fail(NB_ERR_INVALID, "null rows");
    if (n_images == 0 || n_images > 32) return fail(NB_ERR_IMAGES, "n_images must be in [1, 32]");
    if (half && !image && n_images != 1) return fail(NB_ERR_INVALID, "half lists over several images need the image array");
    if (points && (!images || !radius || ntypes == 0)) return fail(NB_ERR_INVALID, "the radius filter needs images and radii");
    return guarded([&] {
        neighbor::gpu::kernel::RowFilterArgs a;
        a.start = reinterpret_cast<const unsigned long long*>(start);
        a.counts = counts;
        a.n = Nq;
        a.query_index = query_index;
        a.neigh = neigh;
        a.image = image;
        a.counts_out = counts_out;
        a.half = half;
        a.n_images = n_images;
        a.points = reinterpret_cast<const float3*>(points);
        a.images = reinterpret_cast<const double3*>(images);
        a.types = types;
        a.shift = shift;
        a.radius = radius;
        a.ntypes = ntypes;
        neighbor::gpu::rows_filter(a, as_stream(stream));
    });
    }

int nb_rows_pack(const uint64_t* start, const uint32_t* counts, uint32_t Nq, const uint32_t* query_index,
                 const uint64_t* offsets, const uint32_t* neigh, const uint8_t* image, uint32_t* neigh_out,
                 uint8_t* image_out, uint32_t* first_out, void* stream)
    {
    if (Nq > 0 && (!start || !counts || !offsets)) return fail(NB_ERR_INVALID, "null rows");
    if ((neigh_out && !neigh) || (image_out && !image)) return fail(NB_ERR_INVALID, "null input for a requested output");
    return guarded([&] {
        neighbor::gpu::rows_pack(reinterpret_cast<const unsigned long long*>(start), counts, Nq, query_index,
                                 reinterpret_cast<const unsigned long long*>(offsets), neigh, image, neigh_out, image_out,
                                 first_out, as_stream(stream));
    });
    }


int nb_exclusive_scan_u32_u64_noalloc(const uint32_t* counts, const uint32_t* perm, uint32_t n, uint64_t* offsets,
                                      void* stream)
    {
    if (!offsets || (n > 0 && !counts)) return fail(NB_ERR_INVALID, "null pointer");
    if (static_cast<uint64_t>(n) > kStaticScanMaxN)
        return fail(NB_ERR_INVALID, "nb_exclusive_scan_u32_u64_noalloc: n exceeds 2^28 counts (the static scratch holds "
                                    "131072 tile sums); use nb_exclusive_scan_u32_u64 with a workspace");
    void* ptr = nullptr;
    const int rc = static_scan_scratch(&ptr);
    if (rc) return rc;
    return guarded([&] {
        neighbor::gpu::scan::exclusive_scan(counts, n, reinterpret_cast<unsigned long long*>(offsets), ptr, as_stream(stream),
                                            perm);
    });
    }

int nb_exclusive_scan_u32_u64(const uint32_t* counts, uint32_t n, uint64_t* offsets, void* scratch, void* stream)
    {
    if (!offsets || !scratch || (n > 0 && !counts)) return fail(NB_ERR_INVALID, "null pointer");
    return guarded([&] {
        neighbor::gpu::scan::exclusive_scan(counts, n, reinterpret_cast<unsigned long long*>(offsets), scratch, as_stream(stream));
    });
    }

} // extern "C"
