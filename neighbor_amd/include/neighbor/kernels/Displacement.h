// Displacement.h -- largest squared displacement of points since a reference configuration
// (the skin test that lets a molecular-dynamics loop keep a neighbor list built at
// cutoff + skin while no particle has moved more than skin / 2).
//
// d = (double)x - (double)x_ref per axis, wrapped to the minimum image (d -= L * rint(d / L))
// in a periodic box; d2 = (dx*dx + dy*dy) + dz*dz, rounded up by a relative 2^-40; the maximum
// over all points is written to *out.  Non-negative doubles order like their bit patterns, so
// the blocks combine with one 64-bit integer atomicMax each.  Optionally counts the points whose
// minimum image is not the plain difference (wrapped across the box since the reference): a
// refit tree keeps such a leaf at its old place in the hierarchy, so that its ancestors' boxes
// stretch across the box.  (No counterpart in the reference.)
#ifndef NEIGHBOR_KERNELS_DISPLACEMENT_H_
#define NEIGHBOR_KERNELS_DISPLACEMENT_H_

#include "../Runtime.h"

#include <stdint.h>

#include <algorithm>

namespace neighbor
{
namespace gpu
{
namespace kernel
{

__global__ void zero_outputs(unsigned long long* __restrict__ p, uint32_t* __restrict__ wrapped)
    {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        {
        *p = 0ull;
        if (wrapped) *wrapped = 0u;
        }
    }

__global__ void __launch_bounds__(256) max_displacement2(const float3* __restrict__ points, const float3* __restrict__ ref,
                                                         const uint32_t n, const int periodic, const double3 L,
                                                         unsigned long long* __restrict__ out, uint32_t* __restrict__ wrapped)
    {
    double m = 0.0;
    uint32_t w = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        {
        const float3 p = points[i];
        const float3 r = ref[i];
        double dx = static_cast<double>(p.x) - static_cast<double>(r.x);
        double dy = static_cast<double>(p.y) - static_cast<double>(r.y);
        double dz = static_cast<double>(p.z) - static_cast<double>(r.z);
        if (periodic)
            {
            const double nx = rint(dx / L.x), ny = rint(dy / L.y), nz = rint(dz / L.z);
            dx -= L.x * nx;
            dy -= L.y * ny;
            dz -= L.z * nz;
            w += (nx != 0.0 || ny != 0.0 || nz != 0.0) ? 1u : 0u;
            }
        const double d2 = ((dx * dx + dy * dy) + dz * dz) * (1.0 + 0x1p-40);
        m = fmax(m, d2);
        }
    // wave, then workgroup maximum
    for (int off = 32; off > 0; off >>= 1)
        {
        m = fmax(m, __shfl_xor(m, off, 64));
        w += __shfl_xor(w, off, 64);
        }
    __shared__ double s_m[4];
    __shared__ uint32_t s_w[4];
    const unsigned int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0)
        {
        s_m[wave] = m;
        s_w[wave] = w;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        {
        double b = s_m[0];
        uint32_t c = s_w[0];
        for (unsigned int v = 1; v < (blockDim.x + 63u) / 64u; ++v)
            {
            b = fmax(b, s_m[v]);
            c += s_w[v];
            }
        atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(b)));
        if (wrapped && c) atomicAdd(wrapped, c);
        }
    }

} // end namespace kernel

inline void max_displacement2(const float3* points, const float3* ref, uint32_t n, bool periodic, double3 L, double* out,
                              uint32_t* wrapped, hipStream_t stream)
    {
    unsigned long long* o = reinterpret_cast<unsigned long long*>(out);
    hipLaunchKernelGGL(kernel::zero_outputs, dim3(1), dim3(64), 0, stream, o, wrapped);
    if (n > 0)
        {
        const unsigned int blocks = std::min<unsigned int>(div_up(n, 256), 1024u);
        hipLaunchKernelGGL(kernel::max_displacement2, dim3(blocks), dim3(256), 0, stream, points, ref, n,
                           periodic ? 1 : 0, L, o, wrapped);
        }
    check_launch("max_displacement2");
    }

} // end namespace gpu
} // end namespace neighbor

#endif // NEIGHBOR_KERNELS_DISPLACEMENT_H_
