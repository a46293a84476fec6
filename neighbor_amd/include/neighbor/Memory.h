// Memory.h -- owning device arrays with shared ownership.
//
// API follows the reference's shared_array / buffered_array
// (include/neighbor/Memory.h:29-296): shared ownership, get(), size(), swap(),
// host operator[] on managed allocations.  Two allocation kinds exist:
//   Managed : hipMallocManaged, host-indexable after a stream sync (the
//             reference's only kind; the default for drop-in use).
//   Device  : hipMalloc, HBM-resident; operator[] must not be used on the host,
//             use copyToHost() instead.  The LBVH and traverser keep their tree
//             arrays in Device memory.
#ifndef NEIGHBOR_MEMORY_H_
#define NEIGHBOR_MEMORY_H_

#include "Runtime.h"

#include <memory>
#include <utility>
#include <vector>

namespace neighbor
{

enum class MemoryKind { Managed, Device };

template<typename T>
class shared_array
    {
    public:
        shared_array() : m_size(0), m_kind(MemoryKind::Managed) {}

        explicit shared_array(size_t size, MemoryKind kind = MemoryKind::Managed)
            : m_size(0), m_kind(kind)
            {
            allocate(size);
            }

        void swap(shared_array& other)
            {
            std::swap(m_data, other.m_data);
            std::swap(m_size, other.m_size);
            std::swap(m_kind, other.m_kind);
            }

        T& operator[](size_t i) { return m_data.get()[i]; }
        const T& operator[](size_t i) const { return m_data.get()[i]; }

        T* get() { return m_data.get(); }
        const T* get() const { return m_data.get(); }

        size_t size() const { return m_size; }
        MemoryKind kind() const { return m_kind; }

        //! Copy the first n elements (all when n==0) to host memory, synchronously.
        std::vector<T> copyToHost(size_t n = 0) const
            {
            if (n == 0 || n > m_size) n = m_size;
            std::vector<T> out(n);
            if (n) NEIGHBOR_HIP_CHECK(hipMemcpy(out.data(), m_data.get(), n * sizeof(T), hipMemcpyDefault));
            return out;
            }

    private:
        std::shared_ptr<T> m_data;
        size_t m_size;
        MemoryKind m_kind;

        struct deleter
            {
            void operator()(T* p) const
                {
                if (p) (void)hipFree(p);
                }
            };

        void allocate(size_t size)
            {
            m_data.reset();
            m_size = 0;
            if (size == 0) return;
            T* p = nullptr;
            hipError_t code = (m_kind == MemoryKind::Managed)
                            ? hipMallocManaged(reinterpret_cast<void**>(&p), size * sizeof(T))
                            : hipMalloc(reinterpret_cast<void**>(&p), size * sizeof(T));
            if (code != hipSuccess)
                {
                (void)hipGetLastError();
                throw hip_error(code, "neighbor::shared_array allocation");
                }
            m_data = std::shared_ptr<T>(p, deleter());
            m_size = size;
            }
    };

//! Two arrays with a selector bit, for ping-pong passes (reference Memory.h:212-296).
template<typename T>
class buffered_array
    {
    public:
        buffered_array() : m_sel(0) {}
        explicit buffered_array(size_t size, MemoryKind kind = MemoryKind::Managed)
            : m_a(size, kind), m_b(size, kind), m_sel(0) {}

        void swap(buffered_array& other)
            {
            m_a.swap(other.m_a);
            m_b.swap(other.m_b);
            std::swap(m_sel, other.m_sel);
            }

        shared_array<T>& current() { return m_sel ? m_b : m_a; }
        const shared_array<T>& current() const { return m_sel ? m_b : m_a; }
        shared_array<T>& alternate() { return m_sel ? m_a : m_b; }
        const shared_array<T>& alternate() const { return m_sel ? m_a : m_b; }
        void flip() { m_sel ^= 1; }
        size_t size() const { return m_a.size(); }

    private:
        shared_array<T> m_a;
        shared_array<T> m_b;
        int m_sel;
    };

} // end namespace neighbor

#endif // NEIGHBOR_MEMORY_H_
