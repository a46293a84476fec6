// periodic_neighbors.hip -- the header-only template API used the way a caller of
// mphowardlab/neighbor uses it (reference test/lbvh_test.cu, benchmark/lbvh_benchmark.cu):
// build an LBVH over points with PointInsertOp, then traverse it with SphereQueryOp over the
// 27 periodic images (ImageListOp<float3>), counting with CountNeighborsOp and listing with
// NeighborListOp.  The counts are checked against a host brute force over the same images
// (pairs inside r must be reported; the reference's query also reports pairs whose quantized
// leaf box overlaps the sphere, up to sqrt(3) quantization bins beyond r, never farther), and
// the list pass against the count pass.
//
// Build: hipcc -std=c++17 -O2 --offload-arch=gfx950 -ffp-contract=off
//        -Xclang -target-feature -Xclang -packed-fp32-ops -I neighbor_amd/include
//        neighbor_amd/examples/periodic_neighbors.hip -o neighbor_amd/examples/periodic_neighbors.bin
// Run:   neighbor_amd/examples/periodic_neighbors.bin [N] [r]   (prints "OK" and exits 0 on success)

#include "neighbor/neighbor.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

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
    const unsigned int N = argc > 1 ? static_cast<unsigned int>(std::atoi(argv[1])) : 4000u;
    const float r = argc > 2 ? static_cast<float>(std::atof(argv[2])) : 2.0f;
    const unsigned int max_neigh = 256;
    const float L = std::cbrt(static_cast<float>(N) / 0.85f);
    if (N == 0 || !(r > 0.0f) || 2.0f * r >= L)
        {
        std::fprintf(stderr, "need N > 0 and 0 < r < L/2 (L = %g)\n", L);
        return 2;
        }

    // points uniform in [0, L)^3; one query sphere of radius r per point
    std::mt19937 gen(7);
    std::uniform_real_distribution<float> u(0.0f, L);
    std::vector<float3> h_points(N);
    std::vector<float4> h_spheres(N);
    for (unsigned int i = 0; i < N; ++i)
        {
        h_points[i] = make_float3(u(gen), u(gen), u(gen));
        h_spheres[i] = make_float4(h_points[i].x, h_points[i].y, h_points[i].z, r);
        }
    std::vector<float3> h_images;
    for (int ix = -1; ix <= 1; ++ix)
        for (int iy = -1; iy <= 1; ++iy)
            for (int iz = -1; iz <= 1; ++iz)
                h_images.push_back(make_float3(L * ix, L * iy, L * iz));

    float3* d_points;
    float4* d_spheres;
    float3* d_images;
    unsigned int* d_counts;
    unsigned int* d_list_counts;
    unsigned int* d_list;
    CHECK(hipMalloc(&d_points, N * sizeof(float3)));
    CHECK(hipMalloc(&d_spheres, N * sizeof(float4)));
    CHECK(hipMalloc(&d_images, h_images.size() * sizeof(float3)));
    CHECK(hipMalloc(&d_counts, N * sizeof(unsigned int)));
    CHECK(hipMalloc(&d_list_counts, N * sizeof(unsigned int)));
    CHECK(hipMalloc(&d_list, static_cast<size_t>(N) * max_neigh * sizeof(unsigned int)));
    CHECK(hipMemcpy(d_points, h_points.data(), N * sizeof(float3), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(d_spheres, h_spheres.data(), N * sizeof(float4), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(d_images, h_images.data(), h_images.size() * sizeof(float3), hipMemcpyHostToDevice));

    // the reference's calls, unchanged
    neighbor::LBVH lbvh;
    lbvh.build(neighbor::PointInsertOp(d_points, N), make_float3(0.0f, 0.0f, 0.0f), make_float3(L, L, L));
    neighbor::LBVHTraverser traverser;
    traverser.traverse(lbvh, neighbor::SphereQueryOp(d_spheres, N), neighbor::CountNeighborsOp(d_counts),
                       neighbor::ImageListOp<float3>(d_images, static_cast<unsigned int>(h_images.size())));
    traverser.traverse(lbvh, neighbor::SphereQueryOp(d_spheres, N),
                       neighbor::NeighborListOp(d_list, d_list_counts, max_neigh),
                       neighbor::ImageListOp<float3>(d_images, static_cast<unsigned int>(h_images.size())));
    CHECK(hipDeviceSynchronize());

    std::vector<unsigned int> counts(N), list_counts(N), list(static_cast<size_t>(N) * max_neigh);
    CHECK(hipMemcpy(counts.data(), d_counts, N * sizeof(unsigned int), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(list_counts.data(), d_list_counts, N * sizeof(unsigned int), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(list.data(), d_list, list.size() * sizeof(unsigned int), hipMemcpyDeviceToHost));

    // host brute force over the 27 images: pairs clearly within r (FP64) must be counted.  The
    // reference's query reports every leaf whose box overlaps the sphere (its refine accepts all)
    // and the default traversal tests the leaves' 10-bit quantized boxes (bins of L/1023,
    // rounded outward), so pairs up to sqrt(3) bins beyond r may be reported too, never farther
    const double bin = static_cast<double>(L) / 1023.0;
    const double r_in = static_cast<double>(r) * (1.0 - 1e-4), r_out = static_cast<double>(r) + 1.75 * bin;
    const double r2 = r_in * r_in;
    const double slack2 = r_out * r_out;
    unsigned long long total = 0;
    for (unsigned int i = 0; i < N; ++i)
        {
        unsigned int inside = 0, near = 0;
        for (unsigned int j = 0; j < N; ++j)
            for (const float3& t : h_images)
                {
                const double dx = static_cast<double>(h_points[i].x) + t.x - h_points[j].x;
                const double dy = static_cast<double>(h_points[i].y) + t.y - h_points[j].y;
                const double dz = static_cast<double>(h_points[i].z) + t.z - h_points[j].z;
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 <= r2) ++inside;
                else if (d2 <= slack2) ++near;
                }
        if (counts[i] < inside || counts[i] > inside + near)
            {
            std::fprintf(stderr, "query %u: %u neighbors, brute force %u (+%u within the quantization band)\n", i,
                         counts[i], inside, near);
            return 1;
            }
        if (list_counts[i] != counts[i])
            {
            std::fprintf(stderr, "query %u: list pass counted %u, count pass %u\n", i, list_counts[i], counts[i]);
            return 1;
            }
        const unsigned int m = counts[i] < max_neigh ? counts[i] : max_neigh;
        for (unsigned int k = 0; k < m; ++k)
            if (list[static_cast<size_t>(i) * max_neigh + k] >= N)
                {
                std::fprintf(stderr, "query %u: neighbor %u out of range\n", i, list[static_cast<size_t>(i) * max_neigh + k]);
                return 1;
                }
        total += counts[i];
        }

    hipFree(d_points);
    hipFree(d_spheres);
    hipFree(d_images);
    hipFree(d_counts);
    hipFree(d_list_counts);
    hipFree(d_list);
    std::printf("OK N=%u r=%g pairs=%llu\n", N, static_cast<double>(r), total);
    return 0;
    }
