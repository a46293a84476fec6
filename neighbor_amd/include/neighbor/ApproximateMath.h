// ApproximateMath.h -- directed-rounding float helpers.
//
// Semantics follow the HIP branch of the reference (include/neighbor/
// ApproximateMath.h:19-21, 55-345): every *_rd / *_ru operation is the IEEE
// round-to-nearest result stepped one ulp outward with C99 nextafterf toward
// -FLT_MAX / +FLT_MAX, even when the RN result is exact; double2float_rd/ru step
// only when the RN conversion lies on the wrong side.  The step is written here
// from the IEEE-754 binary32 layout (sign-magnitude integers are ordered like
// the floats they encode), branch-free so that a wave64 executes it without
// divergence.  Everything is __host__ __device__ so the same code is checked on
// the CPU.  Kernels that include this header must be compiled with
// -ffp-contract=off: a contracted x*y+z would change the RN result that is stepped.
#ifndef NEIGHBOR_APPROXIMATE_MATH_H_
#define NEIGHBOR_APPROXIMATE_MATH_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NEIGHBOR_HD __host__ __device__ __forceinline__

namespace neighbor
{
namespace approx
{

NEIGHBOR_HD uint32_t as_bits(float x) { return __builtin_bit_cast(uint32_t, x); }
NEIGHBOR_HD float from_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

//! nextafterf(x, -FLT_MAX)
NEIGHBOR_HD float next_down(float x)
    {
    uint32_t u = as_bits(x);
    if ((u << 1) == 0u) u = 0x80000000u;                  // +0 and -0 both step to -denorm_min
    uint32_t s = (u & 0x80000000u) ? u + 1u : u - 1u;     // negative: grow magnitude; positive: shrink
    if (u >= 0xff7fffffu) s = (u == 0xff800000u) ? 0xff7fffffu : u;   // -FLT_MAX stays, -inf -> -FLT_MAX, -NaN stays
    else if (u > 0x7f800000u && u < 0x80000000u) s = u;               // +NaN stays
    return from_bits(s);
    }

//! nextafterf(x, +FLT_MAX)
NEIGHBOR_HD float next_up(float x)
    {
    uint32_t u = as_bits(x);
    if ((u << 1) == 0u) u = 0u;                           // +0 and -0 both step to +denorm_min
    uint32_t s = (u & 0x80000000u) ? u - 1u : u + 1u;
    if (u >= 0x7f7fffffu && u < 0x80000000u) s = (u == 0x7f800000u) ? 0x7f7fffffu : u;  // FLT_MAX stays, +inf -> FLT_MAX, +NaN stays
    else if (u > 0xff800000u) s = u;                                                      // -NaN stays
    return from_bits(s);
    }

NEIGHBOR_HD float fadd_rd(float x, float y) { return next_down(x + y); }
NEIGHBOR_HD float fadd_ru(float x, float y) { return next_up(x + y); }
NEIGHBOR_HD float fsub_rd(float x, float y) { return next_down(x - y); }
NEIGHBOR_HD float fsub_ru(float x, float y) { return next_up(x - y); }
NEIGHBOR_HD float fmul_rd(float x, float y) { return next_down(x * y); }
NEIGHBOR_HD float fmul_ru(float x, float y) { return next_up(x * y); }
NEIGHBOR_HD float fdiv_rd(float x, float y) { return next_down(x / y); }
NEIGHBOR_HD float fdiv_ru(float x, float y) { return next_up(x / y); }
NEIGHBOR_HD float frcp_rd(float x) { return next_down(1.0f / x); }
NEIGHBOR_HD float frcp_ru(float x) { return next_up(1.0f / x); }
NEIGHBOR_HD float fmaf_rd(float x, float y, float z) { return next_down(fmaf(x, y, z)); }
NEIGHBOR_HD float fmaf_ru(float x, float y, float z) { return next_up(fmaf(x, y, z)); }

NEIGHBOR_HD float double2float_rd(double x)
    {
    const float y = static_cast<float>(x);
    return (static_cast<double>(y) <= x) ? y : next_down(y);
    }

NEIGHBOR_HD float double2float_ru(double x)
    {
    const float y = static_cast<float>(x);
    return (static_cast<double>(y) >= x) ? y : next_up(y);
    }

//! Saturating float -> uint32 conversion (NaN and x <= 0 give 0, x >= 2^32 gives 0xffffffff).
/*!
 * The reference casts floorf() results to unsigned int
 * (kernels/LBVHTraverser.cuh:100-110); on a degenerate axis that value is -1.0f
 * and gfx9 hardware saturates it to 0.  The saturation is spelled out here so the
 * result does not depend on undefined behaviour.
 */
NEIGHBOR_HD uint32_t sat_f2u(float x)
    {
    return (x > 0.0f) ? ((x < 4294967296.0f) ? static_cast<uint32_t>(x) : 0xffffffffu) : 0u;
    }

} // end namespace approx
} // end namespace neighbor

#endif // NEIGHBOR_APPROXIMATE_MATH_H_
