// LBVH.h -- linear bounding volume hierarchy, built on the GPU.
//
// Public interface of the reference (include/neighbor/LBVH.h:47-219): setup(),
// build() with LaunchParameters / stream / default overloads, getRoot, getN,
// getNInternal, getNNodes, getParents, getLeftChildren, getRightChildren,
// getLowerBounds, getUpperBounds, getPrimitives, data(); Tunable block size.
// Node layout: internal nodes [0, N-1), leaves [N-1, 2N-1), root 0.
//
// Pipeline on one stream, no host synchronization:
//   lbvh_codes (Morton + digit histograms + clears)
//   -> 4 onesweep radix passes -> lbvh_hierarchy (Karras + ropes + span lists; point leaves:
//   also each primitive's point gathered into Morton order; clears the sort's histograms for
//   the next build)
//   -> lbvh_bubble_chunk (LDS, per 2048 sorted leaves; 1024 above 6M leaves) -> lbvh_bubble_span passes
//   (one workgroup per 64-fold wider span of leaves, until one span covers all).
// refit() reruns only the two bubble stages over new primitive boxes.
// Arrays live in HBM (MemoryKind::Device) unless another kind is requested.
#ifndef NEIGHBOR_LBVH_H_
#define NEIGHBOR_LBVH_H_

#include "InsertOps.h"
#include "LBVHData.h"
#include "Memory.h"
#include "Tunable.h"
#include "kernels/LBVHBuild.h"
#include "kernels/RadixSort.h"

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>

namespace neighbor
{

//! Description of a nonzero build error word (bit 0: radix-sort look-back gave up its bounded
//! wait; bit 1: a bubble span list overflowed; bit 2: a span list entry missing its child).
inline std::string build_error_message(unsigned int word)
    {
    std::string m = "LBVH build incomplete (error word " + std::to_string(word) + "):";
    if (word & 1u) m += " a radix-sort look-back gave up its bounded wait;";
    if (word & 6u) m += " a bubble span list overflowed;";
    return m + " the tree must not be used";
    }

class LBVH : public Tunable<unsigned int>
    {
    public:
        //! Bits of the Morton codes (the sort key width).
        static constexpr unsigned int kCodeBits = 30;
        //! Block size used by the overloads without LaunchParameters.
        static constexpr unsigned int kDefaultBlock = 256;

        explicit LBVH(MemoryKind kind = MemoryKind::Device)
            : Tunable<unsigned int>(32, 1024, 32), m_kind(kind),
              m_root(LBVHSentinel), m_N(0), m_N_internal(0), m_N_nodes(0)
            {}

        //! Record an event in every later build (and refit) once the sorted codes, primitive order
        //! and sorted points are complete -- before the bubble passes for point leaves -- so that
        //! work that needs only those can run on another stream meanwhile (waitSorted).
        void enableSortedEvent(bool on)
            {
            if (on && !m_sorted_event)
                {
                hipEvent_t e;
                NEIGHBOR_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                m_sorted_event = std::shared_ptr<void>(static_cast<void*>(e),
                                                       [](void* p) { (void)hipEventDestroy(static_cast<hipEvent_t>(p)); });
                }
            m_record_sorted = on;
            }

        //! Span lists longer than `limit` entries climb through global memory instead of LDS in
        //! the bubble's span passes (this library's addition; default and maximum
        //! gpu::kernel::bubble::kSpanCap; 0 takes the global-memory path for every span, which
        //! real inputs never need: it exists to test that path).
        void setSpanLdsLimit(unsigned int limit) { m_span_lds_limit = std::min(limit, gpu::kernel::bubble::kSpanCap); }

        //! Make `stream` wait until the sorted data of the last build are complete.
        void waitSorted(hipStream_t stream) const
            {
            if (!m_sorted_event) throw std::runtime_error("LBVH::waitSorted needs enableSortedEvent(true) first");
            NEIGHBOR_HIP_CHECK(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(m_sorted_event.get()), 0));
            }

        template<class InsertOpT>
        void setup(const LaunchParameters&, const InsertOpT& insert) { allocate(insert.size()); }
        template<class InsertOpT>
        void setup(stream_t stream, const InsertOpT& insert) { setup(LaunchParameters(kDefaultBlock, stream), insert); }
        template<class InsertOpT>
        void setup(const InsertOpT& insert) { setup(stream_t(0), insert); }

        template<class InsertOpT>
        void build(const LaunchParameters& params, const InsertOpT& insert, const float3& lo, const float3& hi);
        template<class InsertOpT>
        void build(stream_t stream, const InsertOpT& insert, const float3& lo, const float3& hi)
            {
            build(LaunchParameters(kDefaultBlock, stream), insert, lo, hi);
            }
        //! Refit: new primitive boxes under the hierarchy of the last build (same sorted order,
        //! parents, children and ropes; leaf boxes from \a insert, internal boxes bubbled up).
        //! \a insert must hold the same number of primitives as the last build.  The tree stays
        //! valid for any motion (every node box bounds its leaves); it only loosens as the
        //! primitives leave their Morton order.  (No counterpart in the reference.)
        template<class InsertOpT>
        void refit(const LaunchParameters& params, const InsertOpT& insert);
        template<class InsertOpT>
        void refit(stream_t stream, const InsertOpT& insert) { refit(LaunchParameters(kDefaultBlock, stream), insert); }

        template<class InsertOpT>
        void build(const InsertOpT& insert, const float3& lo, const float3& hi)
            {
            build(stream_t(0), insert, lo, hi);
            }

        int getRoot() const { return m_root; }
        unsigned int getN() const { return m_N; }
        unsigned int getNInternal() const { return m_N_internal; }
        unsigned int getNNodes() const { return m_N_nodes; }
        const shared_array<int>& getParents() const { return m_parent; }
        const shared_array<int>& getLeftChildren() const { return m_left; }
        const shared_array<int>& getRightChildren() const { return m_right; }
        const shared_array<float3>& getLowerBounds() const { return m_lo; }
        const shared_array<float3>& getUpperBounds() const { return m_hi; }
        const shared_array<unsigned int>& getPrimitives() const { return m_indexes.current(); }
        //! Ropes of all nodes (this library's addition).
        const shared_array<int>& getRopes() const { return m_rope; }
        //! Leaf positions in Morton order, (x, y, z, bits(primitive)) (this library's addition).
        const shared_array<float4>& getSortedPoints() const { return m_sorted; }
        //! Sorted Morton codes (this library's addition).
        const shared_array<unsigned int>& getCodes() const { return m_codes.current(); }
        //! Nonzero if the build gave up a bounded wait or list (device word; read after a sync):
        //! bit 0 a radix-sort look-back, bits 1-2 a bubble span list.
        const unsigned int* getSortErrorFlag() const { return m_error_ptr; }

        //! Synchronize \a stream and throw std::runtime_error if the last build or refit gave up
        //! a bounded wait or list (its tree is then incomplete).  (No counterpart in the reference.)
        void checkBuild(stream_t stream = stream_t(0)) const
            {
            if (!m_error_ptr) return;
            unsigned int word = 0;
            NEIGHBOR_HIP_CHECK(hipMemcpyAsync(&word, m_error_ptr, sizeof(word), hipMemcpyDeviceToHost, stream));
            NEIGHBOR_HIP_CHECK(hipStreamSynchronize(stream));
            if (word) throw std::runtime_error(build_error_message(word));
            }

        const ConstLBVHData data() const
            {
            ConstLBVHData t;
            t.parent = m_parent.get();
            t.left = m_left.get();
            t.right = m_right.get();
            t.primitive = m_indexes.current().get();
            t.lo = m_lo.get();
            t.hi = m_hi.get();
            t.root = m_root;
            t.rope = m_rope.get();
            t.range_end = m_range_end.get();
            t.sorted_points = m_sorted.get();
            return t;
            }

    private:
        MemoryKind m_kind;
        int m_root;
        unsigned int m_N;
        unsigned int m_N_internal;
        unsigned int m_N_nodes;

        shared_array<int> m_parent;
        shared_array<int> m_left;
        shared_array<int> m_right;
        shared_array<float3> m_lo;
        shared_array<float3> m_hi;
        shared_array<int> m_rope;
        shared_array<int> m_range_end;
        shared_array<float4> m_sorted;
        buffered_array<unsigned int> m_codes;
        buffered_array<unsigned int> m_indexes;
        shared_array<unsigned int> m_work;     //!< sort header | sort status | bubble flags | bubble span lists
        const unsigned int* m_error_ptr = nullptr;
        std::shared_ptr<void> m_sorted_event;   //!< hipEvent_t of enableSortedEvent (shared by copies)
        bool m_record_sorted = false;
        bool m_header_dirty = true;     //!< the sort header of m_work is not known to be zero
        unsigned int m_span_lds_limit = gpu::kernel::bubble::kSpanCap;

        void recordSorted(hipStream_t stream)
            {
            if (m_record_sorted && m_sorted_event)
                NEIGHBOR_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(m_sorted_event.get()), stream));
            }

        LBVHData mutable_data()
            {
            LBVHData t;
            t.parent = m_parent.get();
            t.left = m_left.get();
            t.right = m_right.get();
            t.primitive = m_indexes.current().get();
            t.lo = m_lo.get();
            t.hi = m_hi.get();
            t.root = m_root;
            t.rope = m_rope.get();
            t.range_end = m_range_end.get();
            t.sorted_points = m_sorted.get();
            return t;
            }

        //! Grow-only allocation (reference LBVH.h:336-402); no-op when N is unchanged.
        void allocate(unsigned int N)
            {
            // the radix sort keeps tile counts in 30 bits and node ids are int (2N - 1 nodes)
            if (N >= (1u << 30)) throw std::invalid_argument("LBVH: N must be below 2^30 primitives");
            if (N == m_N && m_parent.size() > 0) return;
            m_root = 0;
            m_N = N;
            m_N_internal = (N > 0) ? N - 1 : 0;
            m_N_nodes = N + m_N_internal;
            const MemoryKind dev = MemoryKind::Device;
            if (m_N_nodes > m_parent.size())
                {
                shared_array<int>(m_N_nodes, m_kind).swap(m_parent);
                shared_array<float3>(m_N_nodes, m_kind).swap(m_lo);
                shared_array<float3>(m_N_nodes, m_kind).swap(m_hi);
                shared_array<int>(m_N_nodes, m_kind).swap(m_rope);
                shared_array<float4>(N, dev).swap(m_sorted);
                }
            if (m_N_internal > m_left.size())
                {
                shared_array<int>(m_N_internal, m_kind).swap(m_left);
                shared_array<int>(m_N_internal, m_kind).swap(m_right);
                shared_array<int>(m_N_internal, dev).swap(m_range_end);
                }
            if (N > m_codes.size())
                {
                buffered_array<unsigned int>(N, m_kind).swap(m_codes);
                buffered_array<unsigned int>(N, m_kind).swap(m_indexes);
                }
            const size_t words = gpu::radix::Workspace::words(N, kCodeBits) + m_N_internal + gpu::kernel::bubble::list_words(N) + 64;
            if (words > m_work.size())
                {
                shared_array<unsigned int>(words, dev).swap(m_work);
                m_header_dirty = true;      // the next build clears the sort header first
                }
            m_error_ptr = m_work.get() ? gpu::radix::Workspace::carve(m_work.get(), N, kCodeBits).error : nullptr;
            }
    };

template<class InsertOpT>
void LBVH::build(const LaunchParameters& params, const InsertOpT& insert, const float3& lo, const float3& hi)
    {
    setup(params, insert);
    const hipStream_t stream = params.stream;
    if (m_N == 0)
        {
        recordSorted(stream);
        return;
        }
    if (m_N == 1)
        {
        hipLaunchKernelGGL(gpu::kernel::lbvh_single<InsertOpT>, dim3(1), dim3(64), 0, stream,
                           mutable_data(), insert, m_codes.current().get(), lo, hi);
        hipLaunchKernelGGL(gpu::radix::kernel::fill_zero, dim3(1), dim3(64), 0, stream, m_work.get(),
                           gpu::radix::Workspace::header_words(gpu::radix::num_passes(kCodeBits)));
        check_launch("lbvh_single");
        recordSorted(stream);
        return;
        }
    const unsigned int block = checkParameter(params);

    // workspace: [sort header (histograms, tile counters, error word) | sort status | bubble flags |
    // span list counters | span lists]; the histograms and tile counters are zero on entry (zeroed
    // at allocation and by the previous build's hierarchy kernel, after its sort), the rest from
    // the error word on is cleared by the codes kernel
    const gpu::radix::Workspace ws = gpu::radix::Workspace::carve(m_work.get(), m_N, kCodeBits);
    uint32_t* clear_begin = ws.error;
    const size_t status_words = static_cast<size_t>(ws.passes) * ws.tiles * gpu::radix::kRadix;
    uint32_t* flags = ws.status + status_words;
    const gpu::kernel::SpanLists lists = gpu::kernel::span_lists(flags + m_N_internal, m_N, ws.error);
    const size_t clear_words = static_cast<size_t>(ws.status - ws.error) + status_words + m_N_internal
                               + gpu::kernel::bubble::count_words(m_N);
    const uint32_t sort_header_words = static_cast<uint32_t>(ws.error - ws.hist);   // histograms + tile counters
    if (m_header_dirty)
        {
        // a new workspace (or one a single-primitive build left): clear the header on this stream
        hipLaunchKernelGGL(gpu::radix::kernel::fill_zero, dim3(1), dim3(256), 0, stream, ws.hist,
                           gpu::radix::Workspace::header_words(ws.passes));
        m_header_dirty = false;
        }

    // one block per 1024 primitives up to 2048 blocks: each folds its histogram into one of
    // the eight copies (radix::kHistParts) with 256 atomics per pass (at 10^6 primitives 10.9 us
    // at 2048 per block, 9.2 at 1024, 16 at 4096; one shared copy: 18.3 us, slower with more
    // blocks)
    const unsigned int code_blocks = std::min<unsigned int>(div_up(m_N, 1024), 2048u);
    hipLaunchKernelGGL(gpu::kernel::lbvh_codes<InsertOpT>, dim3(code_blocks), dim3(256), 0, stream,
                       insert, lo, hi, m_N, m_codes.current().get(), ws.hist, ws.passes, clear_begin, clear_words);
    check_launch("lbvh_codes");

    const bool alt = gpu::radix::sort_passes(m_codes.current().get(), m_indexes.current().get(),
                                             m_codes.alternate().get(), m_indexes.alternate().get(),
                                             m_N, true, ws, stream);
    if (alt)
        {
        m_codes.flip();
        m_indexes.flip();
        }

    // point leaves: the hierarchy kernel also gathers each primitive's point into Morton order
    constexpr bool kGather = gpu::detail::point_leaves<InsertOpT>::value;
    LBVHData tree = mutable_data();
    if constexpr (kGather)
        hipLaunchKernelGGL(gpu::kernel::lbvh_hierarchy<gpu::kernel::SortedPointGather<InsertOpT>>, dim3(div_up(m_N, block)),
                           dim3(block), 0, stream, tree, m_codes.current().get(), m_N, lists, ws.hist, sort_header_words,
                           gpu::kernel::SortedPointGather<InsertOpT>{insert, m_sorted.get()});
    else
        hipLaunchKernelGGL(gpu::kernel::lbvh_hierarchy<>, dim3(div_up(m_N, block)), dim3(block), 0, stream, tree,
                           m_codes.current().get(), m_N, lists, ws.hist, sort_header_words, gpu::radix::NoGather());
    check_launch("lbvh_hierarchy");
    if (kGather) recordSorted(stream);      // sorted points written by the hierarchy kernel
    gpu::kernel::launch_bubble(tree, insert, m_range_end.get(), flags, lists, m_N, stream, kGather, m_span_lds_limit);
    if (!kGather) recordSorted(stream);     // ... or by the chunk pass
    }

template<class InsertOpT>
void LBVH::refit(const LaunchParameters& params, const InsertOpT& insert)
    {
    if (insert.size() != m_N)
        throw std::runtime_error("refit needs a built tree over the same number of primitives");
    const hipStream_t stream = params.stream;
    if (m_N == 0)
        {
        recordSorted(stream);
        return;
        }
    if (m_N == 1)
        {
        // the one leaf is the root: its box and sorted point
        hipLaunchKernelGGL(gpu::kernel::lbvh_refit_single<InsertOpT>, dim3(1), dim3(64), 0, stream, mutable_data(), insert);
        hipLaunchKernelGGL(gpu::radix::kernel::fill_zero, dim3(1), dim3(64), 0, stream, m_work.get(),
                           gpu::radix::Workspace::header_words(gpu::radix::num_passes(kCodeBits)));
        check_launch("lbvh_refit_single");
        recordSorted(stream);
        return;
        }
    (void)checkParameter(params);
    const gpu::radix::Workspace ws = gpu::radix::Workspace::carve(m_work.get(), m_N, kCodeBits);
    const size_t status_words = static_cast<size_t>(ws.passes) * ws.tiles * gpu::radix::kRadix;
    uint32_t* flags = ws.status + status_words;
    // the span lists of the build describe the same topology: only the flags are cleared
    const gpu::kernel::SpanLists lists = gpu::kernel::span_lists(flags + m_N_internal, m_N, ws.error);
    const size_t clear_words = m_N_internal;
    hipLaunchKernelGGL(gpu::radix::kernel::fill_zero, dim3(std::min<unsigned int>(div_up(clear_words, 256), 1024u)),
                       dim3(256), 0, stream, flags, clear_words);
    // the error word keeps the build's bits: a refit of an incomplete tree is incomplete too
    check_launch("lbvh_refit clear");
    gpu::kernel::launch_bubble(mutable_data(), insert, m_range_end.get(), flags, lists, m_N, stream, false, m_span_lds_limit);
    recordSorted(stream);       // the refit's chunk pass rewrites the sorted points
    }

} // end namespace neighbor

#endif // NEIGHBOR_LBVH_H_
