// LBVHTraverser.h -- stackless rope traversal of an LBVH.
//
// Public interface of the reference (include/neighbor/LBVHTraverser.h:60-355):
// setup() compresses once and arms replay, reset() disarms it, traverse() with
// LaunchParameters / stream / default overloads and optional translate and
// transform ops, getData() for the compressed records; Tunable block size.
// traverse() silently returns for an empty tree, no queries or no images and
// throws std::runtime_error for more than 32 images (LBVHTraverser.h:423-432).
//
// Variants (this library): Quantized walks the reference's 16-B records;
// FloatRope walks 32-B records holding the decoded quantized boxes (same
// results, no decode work per visit); FloatRopeExact holds the exact node boxes
// and is meant for ExactSphereQueryOp (with the float SphereQueryOp it reports
// fewer, tighter hits than the reference); Wave (ExactSphereQueryOp only) walks
// exact-box records with bucket leaves, one rope path per wave of 64 queries;
// WaveQuantized is the same walk over decoded quantized boxes, for SphereQueryOp
// (results identical to Quantized) and ExactSphereQueryOp.
#ifndef NEIGHBOR_LBVH_TRAVERSER_H_
#define NEIGHBOR_LBVH_TRAVERSER_H_

#include "LBVH.h"
#include "LBVHTraverserData.h"
#include "Memory.h"
#include "OutputOps.h"
#include "QueryOps.h"
#include "TransformOps.h"
#include "TranslateOps.h"
#include "Tunable.h"
#include "kernels/LBVHTraverse.h"

#include <stdexcept>
#include <type_traits>

namespace neighbor
{

//! Raised by traverse() for more than 32 images.
class image_limit_error : public std::runtime_error
    {
    public:
        image_limit_error() : std::runtime_error("A maximum of 32 image vectors are supported by LBVH traversers.") {}
    };

enum class TraversalVariant { Quantized = 0, FloatRope = 1, FloatRopeExact = 2, Wave = 3, WaveQuantized = 4 };

class LBVHTraverser : public Tunable<unsigned int>
    {
    public:
        static constexpr unsigned int kDefaultBlock = 256;

        LBVHTraverser()
            : Tunable<unsigned int>(32, 1024, 32),
              m_root(LBVHSentinel),
              m_lbvh_lo(1, MemoryKind::Device), m_lbvh_hi(1, MemoryKind::Device), m_bins(1, MemoryKind::Device),
              m_replay(false), m_variant(TraversalVariant::Quantized), m_order(nullptr), m_order_n(0)
            {}

        template<class TransformOpT>
        void setup(const LaunchParameters& params, const LBVH& lbvh, const TransformOpT& transform)
            {
            reset();
            if (lbvh.getN() != 0)
                {
                compress(params, lbvh, transform);
                m_replay = true;
                }
            }
        void setup(const LaunchParameters& params, const LBVH& lbvh) { setup(params, lbvh, NullTransformOp()); }
        template<class TransformOpT>
        void setup(stream_t stream, const LBVH& lbvh, const TransformOpT& transform)
            {
            setup(LaunchParameters(kDefaultBlock, stream), lbvh, transform);
            }
        void setup(stream_t stream, const LBVH& lbvh) { setup(stream, lbvh, NullTransformOp()); }
        template<class TransformOpT>
        void setup(const LBVH& lbvh, const TransformOpT& transform) { setup(stream_t(0), lbvh, transform); }
        void setup(const LBVH& lbvh) { setup(stream_t(0), lbvh, NullTransformOp()); }

        void reset() { m_replay = false; }

        template<class QueryOpT, class OutputOpT, class TranslateOpT, class TransformOpT>
        void traverse(const LaunchParameters& params, const LBVH& lbvh, const QueryOpT& query,
                      const OutputOpT& out, const TranslateOpT& images, const TransformOpT& transform);

        template<class QueryOpT, class OutputOpT, class TranslateOpT>
        void traverse(const LaunchParameters& params, const LBVH& lbvh, const QueryOpT& query,
                      const OutputOpT& out, const TranslateOpT& images)
            {
            traverse(params, lbvh, query, out, images, NullTransformOp());
            }
        template<class QueryOpT, class OutputOpT>
        void traverse(const LaunchParameters& params, const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out)
            {
            traverse(params, lbvh, query, out, SelfOp(), NullTransformOp());
            }
        template<class QueryOpT, class OutputOpT, class TranslateOpT, class TransformOpT>
        void traverse(stream_t stream, const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out,
                      const TranslateOpT& images, const TransformOpT& transform)
            {
            traverse(LaunchParameters(kDefaultBlock, stream), lbvh, query, out, images, transform);
            }
        template<class QueryOpT, class OutputOpT, class TranslateOpT>
        void traverse(stream_t stream, const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out,
                      const TranslateOpT& images)
            {
            traverse(stream, lbvh, query, out, images, NullTransformOp());
            }
        template<class QueryOpT, class OutputOpT>
        void traverse(stream_t stream, const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out)
            {
            traverse(stream, lbvh, query, out, SelfOp(), NullTransformOp());
            }
        template<class QueryOpT, class OutputOpT, class TranslateOpT, class TransformOpT>
        void traverse(const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out, const TranslateOpT& images,
                      const TransformOpT& transform)
            {
            traverse(stream_t(0), lbvh, query, out, images, transform);
            }
        template<class QueryOpT, class OutputOpT, class TranslateOpT>
        void traverse(const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out, const TranslateOpT& images)
            {
            traverse(stream_t(0), lbvh, query, out, images, NullTransformOp());
            }
        template<class QueryOpT, class OutputOpT>
        void traverse(const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out)
            {
            traverse(stream_t(0), lbvh, query, out, SelfOp(), NullTransformOp());
            }

        //! Quantized rope records (reference layout); not produced by the Wave variant.
        const shared_array<int4>& getData() const { return m_data; }
        //! Float rope records (FloatRope variants) or bucket records (wave variants); the Wave
        //! variant does not write the records of nodes below a bucket (never visited).
        const shared_array<float4>& getFloatRopeData() const { return m_float; }
        const shared_array<float3>& getRootLower() const { return m_lbvh_lo; }
        const shared_array<float3>& getRootUpper() const { return m_lbvh_hi; }
        const shared_array<float3>& getBins() const { return m_bins; }

        //! Select the node record walked by traverse(); takes effect at the next compression.
        void setVariant(TraversalVariant v)
            {
            if (v != m_variant) m_replay = false;
            m_variant = v;
            }
        TraversalVariant getVariant() const { return m_variant; }

        //! Thread t serves query order[t] (order must hold a permutation of [0, n)); null clears.
        void setQueryOrder(const unsigned int* order, unsigned int n)
            {
            m_order = order;
            m_order_n = n;
            }

        //! Wave-walk statistics (wave variants, count traversals): each launched wave w writes
        //! its node records loaded, buckets tested and leaves in those buckets to log[3w ..
        //! 3w+2].  words: the log's size; a traversal launching more waves than it holds, or
        //! another output op while the log is set, throws.  null clears.
        void setVisitLog(unsigned int* log, size_t words)
            {
            m_visits = log;
            m_visit_words = log ? words : 0;
            }
        //! Words of visit log a wave traversal of n queries writes.
        static size_t visitLogWords(unsigned int n) { return 3 * static_cast<size_t>(div_up(n, 256)) * 4; }

    private:
        int m_root;
        shared_array<int4> m_data;
        shared_array<float4> m_float;
        shared_array<float3> m_lbvh_lo;
        shared_array<float3> m_lbvh_hi;
        shared_array<float3> m_bins;
        bool m_replay;
        TraversalVariant m_variant;
        const unsigned int* m_order;
        unsigned int m_order_n;
        unsigned int* m_visits = nullptr;
        size_t m_visit_words = 0;

        LBVHCompressedData data()
            {
            LBVHCompressedData c;
            c.root = m_root;
            c.data = m_data.get();
            c.lo = m_lbvh_lo.get();
            c.hi = m_lbvh_hi.get();
            c.bins = m_bins.get();
            return c;
            }

        template<class TransformOpT>
        void compress(const LaunchParameters& params, const LBVH& lbvh, const TransformOpT& transform);

        //! Wave walk: defined for ExactSphereQueryOp only.
        template<class QueryOpT, class OutputOpT, class TranslateOpT, class TransformOpT>
        void traverse_wave(const LaunchParameters& params, const LBVH& lbvh, const QueryOpT& query, const OutputOpT& out,
                           const TranslateOpT& images, const TransformOpT& transform, const unsigned int* order)
            {
            constexpr bool exact = std::is_same<QueryOpT, ExactSphereQueryOp>::value;
            constexpr bool reference = std::is_same<QueryOpT, SphereQueryOp>::value;
            if constexpr (exact || reference)
                {
                // exact node boxes are only conservative for the exact query's FP64 test
                if (!exact && m_variant == TraversalVariant::Wave)
                    throw std::runtime_error("the Wave variant needs ExactSphereQueryOp; use WaveQuantized");
                // whole waves: 256-thread blocks regardless of the tunable block size
                const dim3 grid(div_up(query.size(), 256));
                if (m_visits)
                    {
                    if constexpr (std::is_same<OutputOpT, CountNeighborsOp>::value)
                        {
                        if (m_visit_words < visitLogWords(query.size()))
                            throw std::invalid_argument("the visit log holds fewer words than the traversal's waves write");
                        hipLaunchKernelGGL((gpu::kernel::lbvh_traverse_wave<OutputOpT, QueryOpT, TranslateOpT, TransformOpT, true>),
                                           grid, dim3(256), 0, params.stream, out, m_float.get(), lbvh.getSortedPoints().get(),
                                           lbvh.getPrimitives().get(), m_lbvh_lo.get(), m_lbvh_hi.get(), m_root, query, images,
                                           transform, order, lbvh.getNNodes(), m_visits);
                        return;
                        }
                    else
                        {
                        throw std::invalid_argument("the visit log records count traversals only");
                        }
                    }
                if constexpr (gpu::kernel::wave::compactable<OutputOpT, QueryOpT>::value)
                    {
                    // tail-split map (wave::tail_block): XCD 0 holds the most groups; the other
                    // XCDs' blocks past their own count exit at once
                    const unsigned int G = div_up(query.size(), 256);
                    const unsigned int x0 = G / 8u + (G % 8u ? 1u : 0u);
                    const dim3 tgrid = m_tail_groups ? dim3(8u * (x0 + (m_tail_split - 1u) * std::min(x0, m_tail_groups)))
                                                     : grid;
                    hipLaunchKernelGGL((gpu::kernel::lbvh_traverse_wave_pairs<OutputOpT, QueryOpT, TranslateOpT, TransformOpT>),
                                       tgrid, dim3(256), 0, params.stream, out, m_float.get(), lbvh.getSortedPoints().get(),
                                       lbvh.getPrimitives().get(), m_lbvh_lo.get(), m_lbvh_hi.get(), m_root, query, images,
                                       transform, order, lbvh.getNNodes(), m_tail_groups, m_tail_split);
                    }
                else
                    hipLaunchKernelGGL((gpu::kernel::lbvh_traverse_wave<OutputOpT, QueryOpT, TranslateOpT, TransformOpT>),
                                       grid, dim3(256), 0, params.stream, out, m_float.get(), lbvh.getSortedPoints().get(),
                                       lbvh.getPrimitives().get(), m_lbvh_lo.get(), m_lbvh_hi.get(), m_root, query, images,
                                       transform, order, lbvh.getNNodes(), nullptr);
                }
            else
                {
                throw std::runtime_error("the wave traversal variants need SphereQueryOp or ExactSphereQueryOp");
                }
            }

        unsigned int m_bucket = 16;
        //! Tail split of the compacted walks: per XCD, the last m_tail_groups groups of 256 queries
        //! run as m_tail_split blocks each.  32 groups of four waves are one wave for each SIMD of
        //! an XCD (32 CUs x 4 SIMDs), the last wave each SIMD receives.  Measured at N=10^6
        //! (uniform, 27 images, 48 leaves per bucket): exact count 0.452 -> 0.438 ms, placed fill
        //! 0.632 -> 0.623 ms; 64 or 128 split groups, or blocks of 64 queries, were slower.
        unsigned int m_tail_groups = 32;
        unsigned int m_tail_split = 2;

    public:
        //! Leaves per bucket of the wave variant (1..64); takes effect at the next compression.
        void setBucketSize(unsigned int b)
            {
            if (b < 1 || b > 64) throw std::runtime_error("bucket size must be in [1, 64]");
            if (b != m_bucket) m_replay = false;
            m_bucket = b;
            }
        unsigned int getBucketSize() const { return m_bucket; }

        //! Tail split of the compacted exact walks (counts and pair fills): per XCD, the last
        //! `groups` groups of 256 queries run as `split` blocks of 256/split queries each, so the
        //! waves that start last are shorter (groups = 0: one block per group).
        void setTailSplit(unsigned int groups, unsigned int split)
            {
            if (split != 1 && split != 2 && split != 4) throw std::invalid_argument("the tail split must be 1, 2 or 4");
            m_tail_groups = groups;
            m_tail_split = split;
            }
        unsigned int getTailGroups() const { return m_tail_groups; }
        unsigned int getTailSplit() const { return m_tail_split; }
    };

template<class TransformOpT>
void LBVHTraverser::compress(const LaunchParameters& params, const LBVH& lbvh, const TransformOpT& transform)
    {
    const unsigned int block = checkParameter(params);
    const unsigned int nodes = lbvh.getNNodes();
    if (nodes > m_data.size()) shared_array<int4>(nodes, MemoryKind::Device).swap(m_data);
    m_root = lbvh.getRoot();
    const ConstLBVHData tree = lbvh.data();
    const LBVHCompressedData c = data();
    // the Wave variant walks its own exact bucket records (and the root box they also write):
    // the quantized records are made for the other variants only
    if (m_variant != TraversalVariant::Wave)
        {
        hipLaunchKernelGGL(gpu::kernel::lbvh_compress<TransformOpT>, dim3(div_up(nodes, block)), dim3(block), 0, params.stream,
                           c, transform, tree, lbvh.getNInternal(), nodes);
        check_launch("lbvh_compress");
        }
    if (m_variant != TraversalVariant::Quantized)
        {
        if (2 * static_cast<size_t>(nodes) > m_float.size()) shared_array<float4>(2 * static_cast<size_t>(nodes), MemoryKind::Device).swap(m_float);
        if (m_variant == TraversalVariant::Wave)
            hipLaunchKernelGGL((gpu::kernel::lbvh_bucket_records<false, true>), dim3(div_up(nodes, block)), dim3(block), 0,
                               params.stream, m_float.get(), c, tree, lbvh.getNInternal(), nodes, m_bucket);
        else if (m_variant == TraversalVariant::WaveQuantized)
            hipLaunchKernelGGL(gpu::kernel::lbvh_bucket_records<true>, dim3(div_up(nodes, block)), dim3(block), 0, params.stream,
                               m_float.get(), c, tree, lbvh.getNInternal(), nodes, m_bucket);
        else if (m_variant == TraversalVariant::FloatRope)
            hipLaunchKernelGGL(gpu::kernel::lbvh_float_ropes<true>, dim3(div_up(nodes, block)), dim3(block), 0, params.stream,
                               m_float.get(), c, tree, nodes);
        else
            hipLaunchKernelGGL(gpu::kernel::lbvh_float_ropes<false>, dim3(div_up(nodes, block)), dim3(block), 0, params.stream,
                               m_float.get(), c, tree, nodes);
        check_launch("lbvh_float_ropes");
        }
    }

template<class QueryOpT, class OutputOpT, class TranslateOpT, class TransformOpT>
void LBVHTraverser::traverse(const LaunchParameters& params, const LBVH& lbvh, const QueryOpT& query,
                             const OutputOpT& out, const TranslateOpT& images, const TransformOpT& transform)
    {
    if constexpr (gpu::kernel::wave::places_rows<OutputOpT>::value)
        if (m_variant != TraversalVariant::Wave && m_variant != TraversalVariant::WaveQuantized)
            throw std::invalid_argument("PlacedRowPairOp: rows are placed by the wave traversal variants only");
    if (lbvh.getN() == 0) return;
    if (query.size() == 0 || images.size() == 0) return;
    if (images.size() > 32) throw image_limit_error();
    const unsigned int block = checkParameter(params);
    if (!m_replay) compress(params, lbvh, transform);

    const unsigned int* order = (m_order && m_order_n == query.size()) ? m_order : nullptr;
    const dim3 grid(div_up(query.size(), block));
    if (m_variant == TraversalVariant::Quantized)
        {
        hipLaunchKernelGGL((gpu::kernel::lbvh_traverse_ropes<OutputOpT, QueryOpT, TranslateOpT>), grid, dim3(block), 0,
                           params.stream, out, data(), query, images, order, lbvh.getNNodes());
        }
    else if (m_variant == TraversalVariant::Wave || m_variant == TraversalVariant::WaveQuantized)
        {
        traverse_wave(params, lbvh, query, out, images, transform, order);
        }
    else
        {
        LBVHFloatRopeData f;
        f.root = m_root;
        f.data = m_float.get();
        // the image bits use the unquantized root box stored by the compression
        hipLaunchKernelGGL((gpu::kernel::lbvh_traverse_float_ropes<OutputOpT, QueryOpT, TranslateOpT>), grid, dim3(block), 0,
                           params.stream, out, f, m_lbvh_lo.get(), m_lbvh_hi.get(), query, images, order,
                           lbvh.getNNodes());
        }
    check_launch("lbvh_traverse");
    }

} // end namespace neighbor

#endif // NEIGHBOR_LBVH_TRAVERSER_H_
