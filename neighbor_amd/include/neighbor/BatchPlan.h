// BatchPlan.h -- autotuner and graph-captured plan of the batched neighbor search.
//
// BatchTuner times the batched build for each hierarchy block size and the batched
// count pass for each (variant, leaves per bucket), on the caller's batch itself, and
// keeps the fastest of each on the LBVHBatch.  BatchPlan captures one whole step
// (forest build, bucket records, exact count, scan, exact pair fill) into a HIP graph on
// a stream of its own and replays it with one launch.  Every buffer is allocated before
// the capture, and no node of the graph is a memset (counters are cleared by kernels).
// (No counterpart in the reference.)
#ifndef NEIGHBOR_BATCH_PLAN_H_
#define NEIGHBOR_BATCH_PLAN_H_

#include "LBVHBatch.h"
#include "OutputOps.h"
#include "kernels/Scan.h"

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <vector>

namespace neighbor
{

//! The batched step's inputs and outputs (device pointers).
struct BatchProblem
    {
    const float3* points;
    uint32_t N;
    const uint32_t* offsets;
    const float3* box_lo;
    const float3* box_hi;
    uint32_t nbox;
    const float4* spheres;
    bool periodic;
    };

struct BatchChoice
    {
    unsigned int block;
    unsigned int bucket;
    BatchVariant variant;
    float build_ms;
    float count_ms;
    };

class BatchTuner
    {
    public:
        //! Candidate values (tried in this order; ties keep the earlier one).
        std::vector<unsigned int> blocks{128, 256, 512, 1024};
        std::vector<unsigned int> wave_buckets{8, 16, 32, 64};
        std::vector<unsigned int> lane_buckets{1, 4, 8};
        unsigned int reps = 3;

        //! Tune on the problem; nneigh is a scratch array of N counts.  Leaves the batch built
        //! with the chosen block and its records set up for the chosen bucket.  Synchronizes.
        BatchChoice tune(LBVHBatch& batch, const BatchProblem& pb, uint32_t* nneigh, hipStream_t stream)
            {
            hipEvent_t e0, e1;
            NEIGHBOR_HIP_CHECK(hipEventCreate(&e0));
            NEIGHBOR_HIP_CHECK(hipEventCreate(&e1));
            auto timed = [&](auto&& f) {
                f();    // warm-up
                std::vector<float> t;
                for (unsigned int r = 0; r < reps; ++r)
                    {
                    NEIGHBOR_HIP_CHECK(hipEventRecord(e0, stream));
                    f();
                    NEIGHBOR_HIP_CHECK(hipEventRecord(e1, stream));
                    NEIGHBOR_HIP_CHECK(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    NEIGHBOR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    t.push_back(ms);
                    }
                std::sort(t.begin(), t.end());
                return t[t.size() / 2];
            };
            BatchChoice best{batch.getBlock(), batch.getBucket(), BatchVariant::Wave, 1e30f, 1e30f};
            for (unsigned int b : blocks)
                {
                batch.setBlock(b);
                const float ms = timed([&] { batch.build(pb.points, pb.N, pb.offsets, pb.box_lo, pb.box_hi, pb.nbox, stream); });
                if (ms < best.build_ms)
                    {
                    best.build_ms = ms;
                    best.block = b;
                    }
                }
            batch.setBlock(best.block);
            batch.build(pb.points, pb.N, pb.offsets, pb.box_lo, pb.box_hi, pb.nbox, stream);
            const CountNeighborsOp count(nneigh);
            for (int v = 0; v < 2; ++v)
                {
                const BatchVariant variant = v ? BatchVariant::Lane : BatchVariant::Wave;
                for (unsigned int bucket : (v ? lane_buckets : wave_buckets))
                    {
                    batch.setupTraversal(bucket, stream);
                    const float ms = timed([&] { batch.traverse(count, pb.spheres, pb.periodic, variant, stream); });
                    if (ms < best.count_ms)
                        {
                        best.count_ms = ms;
                        best.bucket = bucket;
                        best.variant = variant;
                        }
                    }
                }
            batch.setupTraversal(best.bucket, stream);
            NEIGHBOR_HIP_CHECK(hipStreamSynchronize(stream));
            NEIGHBOR_HIP_CHECK(hipEventDestroy(e0));
            NEIGHBOR_HIP_CHECK(hipEventDestroy(e1));
            return best;
            }
    };

//! One graph launch = build + records + count + scan + pair fill of the whole batch; point q's
//! pairs are [pair_offsets[q], pair_offsets[q] + nneigh[q]), rows in the forest's Morton order.
//! The graph holds the device addresses of the batch's buffers and of the problem's arrays: it
//! is valid while the batch keeps its buffers (launch() checks) and the caller keeps the
//! problem's arrays and the plan's buffers alive.  A replay rebuilds the batch's forest over
//! the plan's problem.
class BatchPlan
    {
    public:
        //! pair_offsets: N + 1 u64; scan_scratch: gpu::scan::scratch_bytes(N) bytes; neigh/image:
        //! capacity entries (image may be null).  Pairs past capacity are dropped but counted.
        BatchPlan(LBVHBatch& batch, const BatchProblem& pb, BatchVariant variant, uint32_t* nneigh,
                  unsigned long long* pair_offsets, void* scan_scratch, uint32_t* neigh, uint8_t* image,
                  unsigned long long capacity)
            : m_graph(nullptr), m_exec(nullptr), m_stream(nullptr)
            {
            NEIGHBOR_HIP_CHECK(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking));
            try
                {
                const unsigned int bucket = batch.getBucket();
                // rows in the forest's Morton order: the queries of one wave write adjacent rows
                auto step = [&] {
                    batch.build(pb.points, pb.N, pb.offsets, pb.box_lo, pb.box_hi, pb.nbox, m_stream);
                    batch.setupTraversal(bucket, m_stream);
                    batch.traverse(CountNeighborsOp(nneigh), pb.spheres, pb.periodic, variant, m_stream);
                    gpu::scan::exclusive_scan(nneigh, pb.N, pair_offsets, scan_scratch, m_stream, batch.getPrimitives().get());
                    batch.traverse(RowPairOp(pair_offsets, nneigh, neigh, image, nullptr, capacity), pb.spheres, pb.periodic, variant,
                                   m_stream);
                };
                // eager run first: every allocation happens here, outside the capture
                step();
                NEIGHBOR_HIP_CHECK(hipStreamSynchronize(m_stream));
                batch.setGrid(batch.gridSize());
                NEIGHBOR_HIP_CHECK(hipStreamBeginCapture(m_stream, hipStreamCaptureModeRelaxed));
                step();
                NEIGHBOR_HIP_CHECK(hipStreamEndCapture(m_stream, &m_graph));
                NEIGHBOR_HIP_CHECK(hipGraphInstantiate(&m_exec, m_graph, nullptr, nullptr, 0));
                m_generation = batch.generation();
                m_captured = *m_generation;
                }
            catch (...)
                {
                release();
                throw;
                }
            }

        ~BatchPlan() { release(); }
        BatchPlan(const BatchPlan&) = delete;
        BatchPlan& operator=(const BatchPlan&) = delete;

        //! Replay the step.  Throws std::invalid_argument, launching nothing, when the batch has
        //! reallocated a buffer the graph refers to (rebuilt with more points or boxes than at
        //! capture) or has been destroyed: create a new plan then.
        void launch(hipStream_t stream)
            {
            if (!m_generation) throw std::invalid_argument("BatchPlan: not captured");
            if (*m_generation != m_captured)
                throw std::invalid_argument(*m_generation == LBVHBatch::kDestroyed
                                                ? "BatchPlan: its batch was destroyed"
                                                : "BatchPlan: the batch reallocated its buffers since the plan was "
                                                  "captured (rebuilt with more points or boxes); create a new plan");
            NEIGHBOR_HIP_CHECK(hipGraphLaunch(m_exec, stream));
            }

    private:
        hipGraph_t m_graph;
        hipGraphExec_t m_exec;
        hipStream_t m_stream;
        std::shared_ptr<const unsigned long long> m_generation;
        unsigned long long m_captured = 0;

        void release()
            {
            if (m_exec) (void)hipGraphExecDestroy(m_exec);
            if (m_graph) (void)hipGraphDestroy(m_graph);
            if (m_stream)
                {
                (void)hipStreamSynchronize(m_stream);
                (void)hipStreamDestroy(m_stream);
                }
            m_exec = nullptr;
            m_graph = nullptr;
            m_stream = nullptr;
            }
    };

} // end namespace neighbor

#endif // NEIGHBOR_BATCH_PLAN_H_
