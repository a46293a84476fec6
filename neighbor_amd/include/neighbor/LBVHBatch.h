// LBVHBatch.h -- many independent periodic boxes built and traversed as one forest.
//
// Points of all boxes are concatenated, box b holding [offsets[b], offsets[b+1]); each
// box has its own lower/upper corner (its periodic cell when periodic).  build() runs
// the single-tree pipeline once over all boxes with 64-bit (box, Morton code) keys
// (kernels/Batch.h), so every box is a subtree shaped like its own tree, and finds the
// box roots and the 64-query work items.  setupTraversal() writes the bucket records;
// traverse() runs one persistent launch over all items with any output op (count, pair
// list, rows).  Queries are the points themselves (one sphere per point), each against
// its own box, with the 27 periodic images of that box.  Every launch is on one stream
// with no host synchronization, so the whole sequence can be captured in a graph.
// (No counterpart in the reference, which builds one tree per call.)
#ifndef NEIGHBOR_LBVH_BATCH_H_
#define NEIGHBOR_LBVH_BATCH_H_

#include "InsertOps.h"
#include "LBVH.h"
#include "LBVHData.h"
#include "Memory.h"
#include "OutputOps.h"
#include "QueryOps.h"
#include "kernels/Batch.h"
#include "kernels/LBVHBuild.h"
#include "kernels/LBVHTraverse.h"
#include "kernels/RadixSort.h"
#include "kernels/Scan.h"

#include <algorithm>
#include <memory>
#include <stdexcept>

namespace neighbor
{

//! Traversal variant of the batched walk.
enum class BatchVariant
    {
    Wave = 0,   //!< one rope path per wave of 64 Morton-consecutive queries (wave::walk)
    Lane = 1    //!< one rope path per lane
    };

class LBVHBatch
    {
    public:
        static constexpr unsigned int kCodeBits = gpu::batch::kCodeBits;

        LBVHBatch() : m_N(0), m_nbox(0), m_bucket(16), m_block(256), m_grid(0),
                      m_generation(std::make_shared<unsigned long long>(0ull)) {}
        ~LBVHBatch() { *m_generation = kDestroyed; }
        LBVHBatch(const LBVHBatch&) = delete;
        LBVHBatch& operator=(const LBVHBatch&) = delete;

        //! Generation of the device buffers: changes whenever build() or setupTraversal()
        //! reallocates one (more points, boxes or nodes than before) and becomes kDestroyed when
        //! the batch is destroyed.  A captured graph (BatchPlan) is valid only while it is
        //! unchanged.  The counter outlives the batch for whoever holds the pointer.
        static constexpr unsigned long long kDestroyed = ~0ull;
        std::shared_ptr<const unsigned long long> generation() const { return m_generation; }

        //! Synchronize \a stream and throw std::runtime_error if the last build gave up a
        //! bounded wait or list (see LBVH::checkBuild).
        void checkBuild(hipStream_t stream) const
            {
            const unsigned int* word_ptr = getSortErrorFlag();
            if (!word_ptr || m_N == 0) return;
            unsigned int word = 0;
            NEIGHBOR_HIP_CHECK(hipMemcpyAsync(&word, word_ptr, sizeof(word), hipMemcpyDeviceToHost, stream));
            NEIGHBOR_HIP_CHECK(hipStreamSynchronize(stream));
            if (word) throw std::runtime_error(build_error_message(word));
            }

        //! Build the forest.  points: N float3; offsets: nbox + 1 u32 (offsets[0] = 0,
        //! offsets[nbox] = N); box_lo/box_hi: nbox float3.  All device pointers; the
        //! arrays must stay valid for later traversals (the tree refers to points).
        void build(const float3* points, uint32_t N, const uint32_t* offsets, const float3* box_lo,
                   const float3* box_hi, uint32_t nbox, hipStream_t stream)
            {
            if (nbox == 0) throw std::runtime_error("a batch needs at least one box");
            if (nbox >= (1u << 22)) throw std::runtime_error("at most 2^22 boxes per batch");
            if (N >= 0x40000000u) throw std::runtime_error("N must be below 2^30");
            allocate(N, nbox);
            m_points = points;
            m_offsets = offsets;
            m_box_lo = box_lo;
            m_box_hi = box_hi;
            if (N == 0) return;

            const gpu::radix::Workspace ws = gpu::radix::Workspace::carve(m_work.get(), N, key_bits());
            const size_t header = gpu::radix::Workspace::header_words(ws.passes);
            const size_t status_words = static_cast<size_t>(ws.passes) * ws.tiles * gpu::radix::kRadix;
            uint32_t* flags = ws.status + status_words;
            const gpu::kernel::SpanLists lists = gpu::kernel::span_lists(flags + m_Ni, N, ws.error);
            const size_t clear_words = status_words + m_Ni + gpu::kernel::bubble::count_words(N);
            const unsigned int boxes_blocks = div_up(m_nbox, 256);

            hipLaunchKernelGGL(gpu::radix::kernel::fill_zero, dim3(1), dim3(256), 0, stream, ws.hist, header);
            hipLaunchKernelGGL(gpu::batch::kernel::batch_codes, dim3(std::min<unsigned int>(div_up(N, 256), 2048u)), dim3(256), 0,
                               stream, points, N, offsets, m_nbox, box_lo, box_hi, m_keys.current().get(), ws.hist, ws.passes,
                               ws.status, clear_words);
            check_launch("batch_codes");
            if (gpu::radix::sort_passes(m_keys.current().get(), m_indexes.current().get(), m_keys.alternate().get(),
                                        m_indexes.alternate().get(), N, true, ws, stream))
                {
                m_keys.flip();
                m_indexes.flip();
                }
            const PointInsertOp insert(points, N);
            LBVHData tree = data();
            if (N == 1)
                {
                // one point: the leaf is the root (its code is not used afterwards)
                hipLaunchKernelGGL(gpu::kernel::lbvh_single<PointInsertOp>, dim3(1), dim3(64), 0, stream, tree, insert,
                                   m_code_one.get(), make_float3(0.0f, 0.0f, 0.0f), make_float3(1.0f, 1.0f, 1.0f));
                }
            else
                {
                hipLaunchKernelGGL(gpu::batch::kernel::batch_hierarchy, dim3(div_up(N, m_block)), dim3(m_block), 0, stream, tree,
                                   m_keys.current().get(), offsets, N, lists);
                check_launch("batch_hierarchy");
                gpu::kernel::launch_bubble(tree, insert, m_range_end.get(), flags, lists, N, stream);
                }
            hipLaunchKernelGGL(gpu::batch::kernel::batch_roots_init, dim3(boxes_blocks), dim3(256), 0, stream, offsets, N, m_nbox,
                               m_roots.get());
            if (N > 1)
                hipLaunchKernelGGL(gpu::batch::kernel::batch_roots, dim3(div_up(N, 256)), dim3(256), 0, stream, m_range_end.get(),
                                   m_keys.current().get(), offsets, N, m_roots.get());
            hipLaunchKernelGGL(gpu::batch::kernel::batch_item_counts, dim3(boxes_blocks), dim3(256), 0, stream, offsets, m_nbox,
                               m_item_counts.get());
            gpu::scan::exclusive_scan(m_item_counts.get(), m_nbox, m_item_offsets.get(), m_scan_scratch.get(), stream);
            hipLaunchKernelGGL(gpu::batch::kernel::batch_items, dim3(boxes_blocks), dim3(256), 0, stream, offsets, m_nbox,
                               m_item_offsets.get(), m_items.get());
            check_launch("batch roots/items");
            }

        //! Bucket records of the forest for the traversal (leaves per bucket 1..64).
        void setupTraversal(unsigned int bucket, hipStream_t stream)
            {
            if (bucket < 1 || bucket > 64) throw std::runtime_error("bucket size must be in [1, 64]");
            m_bucket = bucket;
            if (m_N == 0) return;
            const unsigned int nodes = m_N + m_Ni;
            if (2 * static_cast<size_t>(nodes) > m_records.size())
                {
                shared_array<float4>(2 * static_cast<size_t>(nodes), MemoryKind::Device).swap(m_records);
                ++*m_generation;
                }
            LBVHCompressedData none{};
            hipLaunchKernelGGL(gpu::kernel::lbvh_bucket_records<false>, dim3(div_up(nodes, 256)), dim3(256), 0, stream,
                               m_records.get(), none, const_data(), m_Ni, nodes, m_bucket);
            check_launch("batch records");
            }

        //! One persistent launch over all items.  spheres: one float4 (x, y, z, r) per point
        //! (original order); the output op is indexed by the original point index.
        template<class OutputOpT>
        void traverse(const OutputOpT& out, const float4* spheres, bool periodic, BatchVariant variant, hipStream_t stream)
            {
            if (m_N == 0) return;
            if (m_records.size() == 0) throw std::runtime_error("setupTraversal() first");
            const ExactSphereQueryOp query(spheres, m_N, m_points);
            hipLaunchKernelGGL(gpu::radix::kernel::fill_zero, dim3(1), dim3(64), 0, stream, m_counters.get(),
                               static_cast<size_t>(gpu::batch::kQueues));
            const unsigned long long* n_items = m_item_offsets.get() + m_nbox;
            const unsigned int grid = gridSize();
            if (variant == BatchVariant::Lane)
                hipLaunchKernelGGL((gpu::batch::kernel::batch_traverse<true, OutputOpT, ExactSphereQueryOp>), dim3(grid), dim3(256), 0,
                                   stream, out, m_records.get(), m_sorted.get(), m_indexes.current().get(), m_lo.get(), m_hi.get(),
                                   m_rope.get(), m_roots.get(), m_offsets, m_box_lo, m_box_hi, m_items.get(), n_items,
                                   m_counters.get(), query, periodic ? 1u : 0u, m_N + m_Ni);
            else if constexpr (gpu::kernel::wave::compactable<OutputOpT, ExactSphereQueryOp>::value)
                hipLaunchKernelGGL((gpu::batch::kernel::batch_traverse_pairs<OutputOpT, ExactSphereQueryOp>), dim3(grid), dim3(256),
                                   0, stream, out, m_records.get(), m_sorted.get(), m_indexes.current().get(), m_lo.get(),
                                   m_hi.get(), m_rope.get(), m_roots.get(), m_offsets, m_box_lo, m_box_hi, m_items.get(), n_items,
                                   m_counters.get(), query, periodic ? 1u : 0u, m_N + m_Ni);
            else
                hipLaunchKernelGGL((gpu::batch::kernel::batch_traverse<false, OutputOpT, ExactSphereQueryOp>), dim3(grid), dim3(256), 0,
                                   stream, out, m_records.get(), m_sorted.get(), m_indexes.current().get(), m_lo.get(), m_hi.get(),
                                   m_rope.get(), m_roots.get(), m_offsets, m_box_lo, m_box_hi, m_items.get(), n_items,
                                   m_counters.get(), query, periodic ? 1u : 0u, m_N + m_Ni);
            check_launch("batch_traverse");
            }

        //! Row capacities of the single-pass pair fill from each box's density (batch_capacity):
        //! cap[i] for every point of the last build, spheres as in traverse().
        void capacity(const float4* spheres, float z, unsigned int pad, uint32_t* cap, hipStream_t stream) const
            {
            if (m_N == 0) return;
            hipLaunchKernelGGL(gpu::batch::kernel::batch_capacity, dim3(std::min<unsigned int>(m_nbox, 65535u)), dim3(256), 0,
                               stream, spheres, m_offsets, m_nbox, m_box_lo, m_box_hi, z, pad, cap);
            check_launch("batch capacity");
            }

        //! Block size of the hierarchy kernel (tunable, multiple of 32 in [32, 1024]).
        void setBlock(unsigned int b)
            {
            if (b < 32 || b > 1024 || b % 32) throw std::runtime_error("invalid block size");
            m_block = b;
            }
        unsigned int getBlock() const { return m_block; }
        //! Set the bucket size used by the next setupTraversal() without writing records.
        void setBucketOnly(unsigned int bucket)
            {
            if (bucket < 1 || bucket > 64) throw std::runtime_error("bucket size must be in [1, 64]");
            m_bucket = bucket;
            }
        unsigned int getBucket() const { return m_bucket; }
        //! Blocks of the persistent traversal (0: from the device's CU count).
        void setGrid(unsigned int g) { m_grid = g; }

        unsigned int getN() const { return m_N; }
        unsigned int getNBoxes() const { return m_nbox; }
        unsigned int getNNodes() const { return m_N + m_Ni; }
        const shared_array<int>& getRoots() const { return m_roots; }
        const shared_array<int>& getParents() const { return m_parent; }
        const shared_array<int>& getLeftChildren() const { return m_left; }
        const shared_array<int>& getRightChildren() const { return m_right; }
        const shared_array<int>& getRopes() const { return m_rope; }
        const shared_array<float3>& getLowerBounds() const { return m_lo; }
        const shared_array<float3>& getUpperBounds() const { return m_hi; }
        const shared_array<unsigned int>& getPrimitives() const { return m_indexes.current(); }
        const shared_array<unsigned long long>& getKeys() const { return m_keys.current(); }
        const unsigned long long* getItemOffsets() const { return m_item_offsets.get(); }
        //! Nonzero if the build gave up a bounded wait or list (see LBVH::getSortErrorFlag).
        const unsigned int* getSortErrorFlag() const
            {
            return m_work.get() ? gpu::radix::Workspace::carve(const_cast<unsigned int*>(m_work.get()), m_N, key_bits()).error : nullptr;
            }

    private:
        unsigned int m_N;
        unsigned int m_Ni = 0;
        unsigned int m_nbox;
        unsigned int m_bucket;
        unsigned int m_block;
        unsigned int m_grid;
        const float3* m_points = nullptr;
        const uint32_t* m_offsets = nullptr;
        const float3* m_box_lo = nullptr;
        const float3* m_box_hi = nullptr;

        shared_array<int> m_parent, m_left, m_right, m_rope, m_range_end, m_roots;
        shared_array<float3> m_lo, m_hi;
        shared_array<float4> m_sorted, m_records;
        buffered_array<unsigned long long> m_keys;
        buffered_array<unsigned int> m_indexes;
        shared_array<unsigned int> m_work, m_item_counts, m_counters, m_code_one;
        shared_array<unsigned long long> m_item_offsets;
        shared_array<gpu::batch::Item> m_items;
        shared_array<unsigned char> m_scan_scratch;
        std::shared_ptr<unsigned long long> m_generation;

        unsigned int key_bits() const
            {
            unsigned int bits = 1;
            while ((1u << bits) < m_nbox) ++bits;
            return kCodeBits + bits;
            }

    public:
        //! Blocks of the persistent traversal: the set value, else 8 per CU.
        unsigned int gridSize() const
            {
            if (m_grid) return m_grid;
            int dev = 0, cus = 0;
            NEIGHBOR_HIP_CHECK(hipGetDevice(&dev));
            NEIGHBOR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            // enough 256-thread blocks to fill every CU; waves leave when the items run out
            return static_cast<unsigned int>(std::max(cus, 1)) * 8u;
            }

    private:

        LBVHData data()
            {
            LBVHData t;
            t.parent = m_parent.get();
            t.left = m_left.get();
            t.right = m_right.get();
            t.primitive = m_indexes.current().get();
            t.lo = m_lo.get();
            t.hi = m_hi.get();
            t.root = 0;
            t.rope = m_rope.get();
            t.range_end = m_range_end.get();
            t.sorted_points = m_sorted.get();
            return t;
            }

        ConstLBVHData const_data() const
            {
            ConstLBVHData t;
            t.parent = m_parent.get();
            t.left = m_left.get();
            t.right = m_right.get();
            t.primitive = m_indexes.current().get();
            t.lo = m_lo.get();
            t.hi = m_hi.get();
            t.root = 0;
            t.rope = m_rope.get();
            t.range_end = m_range_end.get();
            t.sorted_points = m_sorted.get();
            return t;
            }

        //! Grow-only allocation.
        void allocate(uint32_t N, uint32_t nbox)
            {
            const MemoryKind dev = MemoryKind::Device;
            bool grew = false;
            m_N = N;
            m_Ni = N > 0 ? N - 1 : 0;
            m_nbox = nbox;
            const size_t nodes = static_cast<size_t>(N) + m_Ni;
            if (nodes > m_parent.size())
                {
                grew = true;
                shared_array<int>(nodes, dev).swap(m_parent);
                shared_array<int>(nodes, dev).swap(m_rope);
                shared_array<float3>(nodes, dev).swap(m_lo);
                shared_array<float3>(nodes, dev).swap(m_hi);
                }
            if (std::max<size_t>(m_Ni, 1) > m_left.size())
                {
                grew = true;
                shared_array<int>(std::max<size_t>(m_Ni, 1), dev).swap(m_left);
                shared_array<int>(std::max<size_t>(m_Ni, 1), dev).swap(m_right);
                shared_array<int>(std::max<size_t>(m_Ni, 1), dev).swap(m_range_end);
                }
            if (N > m_keys.size())
                {
                grew = true;
                buffered_array<unsigned long long>(N, dev).swap(m_keys);
                buffered_array<unsigned int>(N, dev).swap(m_indexes);
                shared_array<float4>(N, dev).swap(m_sorted);
                }
            const size_t words = gpu::radix::Workspace::words(std::max<uint32_t>(N, 1), key_bits()) + m_Ni
                                 + gpu::kernel::bubble::list_words(N) + 64;
            if (words > m_work.size())
                {
                grew = true;
                shared_array<unsigned int>(words, dev).swap(m_work);
                }
            if (nbox > m_roots.size())
                {
                grew = true;
                shared_array<int>(nbox, dev).swap(m_roots);
                shared_array<unsigned int>(nbox, dev).swap(m_item_counts);
                shared_array<unsigned long long>(static_cast<size_t>(nbox) + 1, dev).swap(m_item_offsets);
                shared_array<unsigned char>(gpu::scan::scratch_bytes(nbox) + 256, dev).swap(m_scan_scratch);
                }
            const size_t max_items = N / gpu::batch::kItemQueries + nbox + 1;
            if (max_items > m_items.size())
                {
                grew = true;
                shared_array<gpu::batch::Item>(max_items, dev).swap(m_items);
                }
            if (m_counters.size() == 0)
                {
                grew = true;
                shared_array<unsigned int>(gpu::batch::kQueues, dev).swap(m_counters);
                shared_array<unsigned int>(1, dev).swap(m_code_one);
                }
            if (grew) ++*m_generation;
            }
    };

} // end namespace neighbor

#endif // NEIGHBOR_LBVH_BATCH_H_
