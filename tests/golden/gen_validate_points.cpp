// Writes the lbvh_validate ideal-gas positions as raw float32 (x,y,z) triples.
// Box 20 x 15 x 25, N = 7500, std::mt19937(42) and
// std::uniform_real_distribution<float>(-0.5, 0.5), draws taken x, y, z in order
// (the reference test forms all three inside one make_float3 call, whose argument
// evaluation order C++ leaves unspecified; any order gives a valid ideal gas).
#include <cstdio>
#include <random>

int main(int argc, char** argv)
    {
    const float Lx = 20.f, Ly = 15.f, Lz = 25.f;
    const unsigned int N = static_cast<unsigned int>(1.0 * Lx * Ly * Lz);
    std::mt19937 mt(42);
    std::uniform_real_distribution<float> U(-0.5, 0.5);
    FILE* f = std::fopen(argc > 1 ? argv[1] : "validate_points.f32", "wb");
    if (!f) return 1;
    for (unsigned int i = 0; i < N; ++i)
        {
        const float x = Lx * U(mt);
        const float y = Ly * U(mt);
        const float z = Lz * U(mt);
        const float p[3] = {x, y, z};
        std::fwrite(p, sizeof(float), 3, f);
        }
    std::fclose(f);
    return 0;
    }
