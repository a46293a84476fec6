// Runtime.h -- thin HIP runtime helpers for the neighbor library (gfx950).
//
// The reference reaches the GPU runtime through the hipper interop layer
// (absent submodule extern/hipper); this library calls HIP directly.
#ifndef NEIGHBOR_RUNTIME_H_
#define NEIGHBOR_RUNTIME_H_

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace neighbor
{
//! Stream handle (the reference's hipper::stream_t).
typedef hipStream_t stream_t;

//! Error raised for a failed HIP runtime call.
class hip_error : public std::runtime_error
    {
    public:
        hip_error(hipError_t code, const std::string& what)
            : std::runtime_error(what + ": " + hipGetErrorString(code)), m_code(code)
            {}
        hipError_t code() const { return m_code; }
    private:
        hipError_t m_code;
    };

inline void check_hip(hipError_t code, const char* what)
    {
    if (code != hipSuccess) throw hip_error(code, what);
    }

#define NEIGHBOR_HIP_CHECK(call) ::neighbor::check_hip((call), #call)

//! Check the launch status of the most recent kernel.
inline void check_launch(const char* what)
    {
    check_hip(hipGetLastError(), what);
    }

//! Number of blocks needed to cover n items with blocks of size b.
inline unsigned int div_up(size_t n, unsigned int b)
    {
    return static_cast<unsigned int>((n + b - 1) / b);
    }

} // end namespace neighbor

#endif // NEIGHBOR_RUNTIME_H_
