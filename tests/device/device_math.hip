// device_math.hip -- runs the product's directed-rounding helpers (ApproximateMath.h) and
// bounding-volume constructors (BoundingVolumes.h) inside a gfx950 kernel, one thread per
// case, so that a test can compare their device results bit for bit with the oracle.
//
// Input  (binary, n records): float a, b, c; float pad; double ox, oy, oz, r;
//                             float lo[3], hi[3]
// Output (binary, n records of 22 u32 words):
//   0 next_down(a)   1 next_up(a)    2 fadd_rd(a,b)  3 fadd_ru(a,b)  4 fsub_rd(a,b)
//   5 fsub_ru(a,b)   6 fmul_rd(a,b)  7 fmul_ru(a,b)  8 fdiv_rd(a,b)  9 fdiv_ru(a,b)
//  10 frcp_rd(b)    11 frcp_ru(b)   12 fmaf_rd(a,b,c) 13 fmaf_ru(a,b,c)
//  14 double2float_rd(ox)  15 double2float_ru(ox)  16 sat_f2u(a)
//  17 Rsq of BoundingSphere(double3(ox, oy, oz), r)   18-20 its origin x, y, z
//  21 bit 0: that sphere overlaps BoundingBox(lo, hi);
//     bit 1: BoundingSphere(float3(a, b, c), (float)r) overlaps it
//
// Build: hipcc -std=c++17 -O3 --offload-arch=gfx950 -ffp-contract=off
//        -Xclang -target-feature -Xclang -packed-fp32-ops -I neighbor_amd/include
//        tests/device/device_math.hip -o tests/device/device_math.bin
// Run:   tests/device/device_math.bin in.bin out.bin
#include "neighbor/BoundingVolumes.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

struct Case
    {
    float a, b, c, pad;
    double ox, oy, oz, r;
    float lo[3], hi[3];
    };
static_assert(sizeof(Case) == 72, "record layout");

constexpr unsigned int kWords = 22;

__global__ void probe(const Case* __restrict__ in, uint32_t n, uint32_t* __restrict__ out)
    {
    using namespace neighbor;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Case k = in[i];
    uint32_t* o = out + static_cast<size_t>(i) * kWords;
    const float v[16] = {approx::next_down(k.a), approx::next_up(k.a), approx::fadd_rd(k.a, k.b), approx::fadd_ru(k.a, k.b),
                         approx::fsub_rd(k.a, k.b), approx::fsub_ru(k.a, k.b), approx::fmul_rd(k.a, k.b),
                         approx::fmul_ru(k.a, k.b), approx::fdiv_rd(k.a, k.b), approx::fdiv_ru(k.a, k.b),
                         approx::frcp_rd(k.b), approx::frcp_ru(k.b), approx::fmaf_rd(k.a, k.b, k.c),
                         approx::fmaf_ru(k.a, k.b, k.c), approx::double2float_rd(k.ox), approx::double2float_ru(k.ox)};
    for (int w = 0; w < 16; ++w) o[w] = approx::as_bits(v[w]);
    o[16] = approx::sat_f2u(k.a);
    const BoundingSphere s(make_double3(k.ox, k.oy, k.oz), k.r);
    o[17] = approx::as_bits(s.Rsq);
    o[18] = approx::as_bits(s.origin.x);
    o[19] = approx::as_bits(s.origin.y);
    o[20] = approx::as_bits(s.origin.z);
    const BoundingBox box(make_float3(k.lo[0], k.lo[1], k.lo[2]), make_float3(k.hi[0], k.hi[1], k.hi[2]));
    const BoundingSphere f(make_float3(k.a, k.b, k.c), static_cast<float>(k.r));
    o[21] = (s.overlap(box) ? 1u : 0u) | (f.overlap(box) ? 2u : 0u);
    }

#define CHECK(x)                                                                        \
    do                                                                                  \
        {                                                                               \
        const hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess)                                                           \
            {                                                                           \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));        \
            return 2;                                                                   \
            }                                                                           \
        } while (0)

int main(int argc, char** argv)
    {
    if (argc != 3)
        {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 1;
        }
    FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) return 1;
    std::vector<Case> cases;
    Case k;
    while (std::fread(&k, sizeof(Case), 1, fi) == 1) cases.push_back(k);
    std::fclose(fi);
    const uint32_t n = static_cast<uint32_t>(cases.size());
    std::vector<uint32_t> res(static_cast<size_t>(n) * kWords, 0u);
    if (n > 0)
        {
        Case* d_in = nullptr;
        uint32_t* d_out = nullptr;
        CHECK(hipMalloc(&d_in, sizeof(Case) * n));
        CHECK(hipMalloc(&d_out, sizeof(uint32_t) * kWords * n));
        CHECK(hipMemcpy(d_in, cases.data(), sizeof(Case) * n, hipMemcpyHostToDevice));
        const unsigned int block = 256;
        hipLaunchKernelGGL(probe, dim3((n + block - 1) / block), dim3(block), 0, 0, d_in, n, d_out);
        CHECK(hipGetLastError());
        CHECK(hipMemcpy(res.data(), d_out, sizeof(uint32_t) * kWords * n, hipMemcpyDeviceToHost));
        CHECK(hipFree(d_in));
        CHECK(hipFree(d_out));
        }
    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 1;
    std::fwrite(res.data(), sizeof(uint32_t), res.size(), fo);
    std::fclose(fo);
    std::printf("OK n=%u\n", n);
    return 0;
    }
