// RowFilter.h -- in-place filters of neighbor rows (half lists, per-pair shifted radii) and
// the packing of rows into gap-free CSR rows and an (i, j, image) pair stream.
//
// Row q holds the pairs (i, j, k) of particle i = query_index[q] (or q) at
// neigh/image[start[q], start[q] + counts[q]).  The kept pairs move to the front of the row in
// their order and counts_out[q] receives how many there are.  Filters:
//  - half: keep j > i, and for j == i (a particle and its own periodic image) the image k with
//    2k + 1 < n_images, the images being listed in mirrored pairs (images[n_images-1-k] ==
//    -images[k], as the 27-image list x-outer / z-inner is).  Of a pair (i, j, k) and its mirror
//    (j, i, n_images-1-k) exactly one is kept; the self pair at the zero image is dropped.
//  - radii: keep d2 <= r^2 with d = ((double)x_i + images[k]) - (double)x_j per axis,
//    d2 = (dx*dx + dy*dy) + dz*dz and r = ((double)radius[t_i * ntypes + t_j] + shift_i) + shift_j,
//    all in IEEE double, and r >= 0 (the exact pair test of the sphere query with a per-pair
//    radius: HOOMD-style diameter shifting with shift = (diameter - 1) / 2).
// One thread per row; rows laid out in traversal order put the rows of a wave side by side.
// (No counterpart in the reference.)
#ifndef NEIGHBOR_KERNELS_ROW_FILTER_H_
#define NEIGHBOR_KERNELS_ROW_FILTER_H_

#include "../Runtime.h"

#include <stdint.h>

namespace neighbor
{
namespace gpu
{
namespace kernel
{

struct RowFilterArgs
    {
    const unsigned long long* start;
    const uint32_t* counts;
    uint32_t n;
    const uint32_t* query_index;    //!< row -> particle (null: identity)
    uint32_t* neigh;
    uint8_t* image;                 //!< null: every pair is at image 0
    uint32_t* counts_out;           //!< may alias counts
    int half;
    uint32_t n_images;
    const float3* points;           //!< null: no radius filter
    const double3* images;
    const uint32_t* types;          //!< null: all type 0
    const float* shift;             //!< null: all 0
    const float* radius;            //!< ntypes x ntypes
    uint32_t ntypes;
    };

__global__ void __launch_bounds__(256) rows_filter(const RowFilterArgs a)
    {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    const uint32_t i = a.query_index ? a.query_index[q] : q;
    const unsigned long long s = a.start[q];
    const uint32_t c = a.counts[q];

    double xi = 0.0, yi = 0.0, zi = 0.0, si = 0.0;
    const float* rrow = nullptr;
    if (a.points)
        {
        const float3 p = a.points[i];
        xi = p.x;
        yi = p.y;
        zi = p.z;
        si = a.shift ? static_cast<double>(a.shift[i]) : 0.0;
        rrow = a.radius + (a.types ? a.types[i] : 0u) * a.ntypes;
        }

    uint32_t kept = 0;
    for (uint32_t e = 0; e < c; ++e)
        {
        const uint32_t j = a.neigh[s + e];
        const uint32_t k = a.image ? a.image[s + e] : 0u;
        bool keep = true;
        if (a.half) keep = (j > i) || (j == i && 2u * k + 1u < a.n_images);
        if (keep && a.points)
            {
            const double3 t = a.images[k];
            const float3 pj = a.points[j];
            const double dx = (xi + t.x) - static_cast<double>(pj.x);
            const double dy = (yi + t.y) - static_cast<double>(pj.y);
            const double dz = (zi + t.z) - static_cast<double>(pj.z);
            const double d2 = (dx * dx + dy * dy) + dz * dz;
            const double sj = a.shift ? static_cast<double>(a.shift[j]) : 0.0;
            const double r = (static_cast<double>(rrow[a.types ? a.types[j] : 0u]) + si) + sj;
            keep = r >= 0.0 && d2 <= r * r;
            }
        if (keep)
            {
            a.neigh[s + kept] = j;
            if (a.image) a.image[s + kept] = static_cast<uint8_t>(k);
            ++kept;
            }
        }
    a.counts_out[q] = kept;
    }

//! Row q = [start[q], start[q] + counts[q]) copied to [offsets[q], offsets[q] + counts[q]) of the
//! output arrays (CSR rows without gaps, offsets from a scan of counts), and first[e] = i for each
//! pair e of the row (with neigh and image: the (i, j, image) pair stream).  Any output may be null.
__global__ void __launch_bounds__(256) rows_pack(const unsigned long long* __restrict__ start, const uint32_t* __restrict__ counts,
                                                 const uint32_t n, const uint32_t* __restrict__ query_index,
                                                 const unsigned long long* __restrict__ offsets, const uint32_t* __restrict__ neigh,
                                                 const uint8_t* __restrict__ image, uint32_t* __restrict__ neigh_out,
                                                 uint8_t* __restrict__ image_out, uint32_t* __restrict__ first_out)
    {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t i = query_index ? query_index[q] : q;
    const unsigned long long s = start[q], o = offsets[q];
    const uint32_t c = counts[q];
    for (uint32_t e = 0; e < c; ++e)
        {
        if (neigh_out) neigh_out[o + e] = neigh[s + e];
        if (image_out) image_out[o + e] = image[s + e];
        if (first_out) first_out[o + e] = i;
        }
    }

} // end namespace kernel

inline void rows_pack(const unsigned long long* start, const uint32_t* counts, uint32_t n, const uint32_t* query_index,
                      const unsigned long long* offsets, const uint32_t* neigh, const uint8_t* image, uint32_t* neigh_out,
                      uint8_t* image_out, uint32_t* first_out, hipStream_t stream)
    {
    if (n == 0) return;
    hipLaunchKernelGGL(kernel::rows_pack, dim3((n + 255u) / 256u), dim3(256), 0, stream, start, counts, n, query_index,
                       offsets, neigh, image, neigh_out, image_out, first_out);
    check_launch("rows_pack");
    }

inline void rows_filter(const kernel::RowFilterArgs& a, hipStream_t stream)
    {
    if (a.n == 0) return;
    hipLaunchKernelGGL(kernel::rows_filter, dim3((a.n + 255u) / 256u), dim3(256), 0, stream, a);
    check_launch("rows_filter");
    }

} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_ROW_FILTER_H_
