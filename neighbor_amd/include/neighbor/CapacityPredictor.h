// CapacityPredictor.h -- learned neighbor capacity for single-pass list fill.
//
// Owns the MLP parameters (device), its Adam state and the per-primitive features of
// the last LBVH it saw (kernels/Capacity.h).  Typical use:
//   computeFeatures(lbvh, lo, hi, r)      -- Morton-cell occupancy features
//   train(lbvh, counts, steps, batch, lr) -- counts from an exact count pass
//   predictCapacity(lbvh, margin, pad, cap)
//   predictCapacityDirect(lbvh, lo, hi, r, margin, pad, cap) -- features and prediction in one
//                                           kernel, for a tree the network was not trained on
// The capacity only sizes the rows of the single fill pass; rows that overflow it are
// found afterwards and redone by the exact count-then-fill path, so the neighbor sets
// never depend on the network.  (No counterpart in the reference.)
#ifndef NEIGHBOR_CAPACITY_PREDICTOR_H_
#define NEIGHBOR_CAPACITY_PREDICTOR_H_

#include "LBVH.h"
#include "Memory.h"
#include "OutputOps.h"
#include "kernels/Capacity.h"
#include "kernels/Scan.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <vector>

namespace neighbor
{

class CapacityPredictor
    {
    public:
        static constexpr int kParams = gpu::capacity::kParams;
        static constexpr int kFeatures = gpu::capacity::kFeatures;

        explicit CapacityPredictor(unsigned long long seed = 1)
            : m_params(kParams, MemoryKind::Device), m_grad(kParams + 1, MemoryKind::Device),
              m_m(kParams, MemoryKind::Device), m_v(kParams, MemoryKind::Device), m_step(0), m_n(0)
            {
            // He-uniform initialization of the weights, zero biases, output bias at log2(1 + 64)
            std::mt19937_64 rng(seed);
            std::vector<float> p(kParams, 0.0f);
            auto fill = [&](int off, int fan_in, int count) {
                const float a = std::sqrt(6.0f / static_cast<float>(fan_in));
                std::uniform_real_distribution<float> U(-a, a);
                for (int j = 0; j < count; ++j) p[off + j] = U(rng);
            };
            fill(gpu::capacity::kW1, kFeatures, gpu::capacity::kHidden * kFeatures);
            fill(gpu::capacity::kW2, gpu::capacity::kHidden, gpu::capacity::kHidden * gpu::capacity::kHidden);
            fill(gpu::capacity::kW3, gpu::capacity::kHidden, gpu::capacity::kHidden);
            for (int j = 0; j < gpu::capacity::kHidden; ++j) p[gpu::capacity::kW3 + j] *= 0.1f;
            p[gpu::capacity::kB3] = 6.0f;
            setParameters(p);
            }

        void setParameters(const std::vector<float>& p)
            {
            if (p.size() != static_cast<size_t>(kParams)) throw std::runtime_error("wrong parameter count");
            NEIGHBOR_HIP_CHECK(hipMemcpy(m_params.get(), p.data(), kParams * sizeof(float), hipMemcpyHostToDevice));
            std::vector<float> zero(kParams + 1, 0.0f);
            NEIGHBOR_HIP_CHECK(hipMemcpy(m_m.get(), zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice));
            NEIGHBOR_HIP_CHECK(hipMemcpy(m_v.get(), zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice));
            NEIGHBOR_HIP_CHECK(hipMemcpy(m_grad.get(), zero.data(), (kParams + 1) * sizeof(float), hipMemcpyHostToDevice));
            m_step = 0;
            }

        std::vector<float> getParameters() const { return m_params.copyToHost(kParams); }
        float* parameters() { return m_params.get(); }
        const float* features() const { return m_features.get(); }
        unsigned int numFeatureRows() const { return m_n; }

        //! Features of every primitive of lbvh (sorted order); lo/hi is the scene box of the build.
        void computeFeatures(const LBVH& lbvh, const float3& lo, const float3& hi, float r, stream_t stream)
            {
            const unsigned int N = lbvh.getN();
            m_n = N;
            if (N == 0) return;
            if (static_cast<size_t>(N) * kFeatures > m_features.size())
                shared_array<float>(static_cast<size_t>(N) * kFeatures, MemoryKind::Device).swap(m_features);
            const gpu::capacity::Levels lv = levels(lbvh, lo, hi, r, stream);
            hipLaunchKernelGGL(gpu::capacity::kernel::features, dim3(div_up(N, 256)), dim3(256), 0, stream, lbvh.getCodes().get(),
                               lbvh.getSortedPoints().get(), N, lv, m_features.get());
            check_launch("capacity features");
            }

        //! Capacity per query straight from the tree (features computed on the fly, none stored):
        //! cap[q] = ceil(2^(prediction + margin)) - 1 + pad, query q = primitive q.  Same values as
        //! computeFeatures() followed by predictCapacity().
        void predictCapacityDirect(const LBVH& lbvh, const float3& lo, const float3& hi, float r, float margin,
                                   unsigned int pad, uint32_t* cap, stream_t stream)
            {
            const unsigned int N = lbvh.getN();
            if (N == 0) return;
            const gpu::capacity::Levels lv = levels(lbvh, lo, hi, r, stream);
            hipLaunchKernelGGL(gpu::capacity::kernel::capacity_direct, dim3(div_up(N, 256)), dim3(256), 0, stream, m_params.get(),
                               lbvh.getCodes().get(), lbvh.getSortedPoints().get(), lbvh.getPrimitives().get(), N, lv, margin,
                               pad, cap, static_cast<unsigned long long*>(nullptr));
            check_launch("capacity direct");
            }

        //! predictCapacityDirect, then the row bases of the placed fill (PlacedRowPairOp): the
        //! kernel also sums the capacities of each group of 256 sorted positions, and one scan of
        //! those sums gives base[g] = the capacities of the sorted positions before 256 g
        //! (g < ceil(N / 256)), base[ceil(N / 256)] = all of them, and spill (optional) = {that
        //! total, 0}: the cursor of the overflow area past all rows and its moved-row count, as
        //! the fill expects them.  base holds ceil(N / 256) + 1 values.
        void predictRows(const LBVH& lbvh, const float3& lo, const float3& hi, float r, float margin, unsigned int pad,
                         uint32_t* cap, unsigned long long* base, unsigned long long* spill, stream_t stream)
            {
            const unsigned int N = lbvh.getN();
            const unsigned int G = div_up(N, PlacedRowPairOp::kGroup);
            if (N > 0)
                {
                const gpu::capacity::Levels lv = levels(lbvh, lo, hi, r, stream);
                if (G > m_group_total.size()) shared_array<unsigned long long>(G, MemoryKind::Device).swap(m_group_total);
                hipLaunchKernelGGL(gpu::capacity::kernel::capacity_direct, dim3(G), dim3(PlacedRowPairOp::kGroup), 0, stream,
                                   m_params.get(), lbvh.getCodes().get(), lbvh.getSortedPoints().get(),
                                   lbvh.getPrimitives().get(), N, lv, margin, pad, cap, m_group_total.get());
                check_launch("capacity direct");
                }
            gpu::scan::group_totals_scan(m_group_total.get(), G, base, spill, stream);
            }

        //! Train on exact counts (indexed by query = primitive) for `steps` Adam steps of `batch` samples.
        //! Returns the loss of the last step (synchronizes the stream).
        float train(const LBVH& lbvh, const uint32_t* counts, unsigned int steps, unsigned int batch, float lr,
                    unsigned long long seed, stream_t stream)
            {
            const unsigned int N = lbvh.getN();
            if (N == 0 || m_n != N) throw std::runtime_error("computeFeatures() first");
            batch = std::max(1u, std::min(batch, N));
            if (N > m_targets.size()) shared_array<float>(N, MemoryKind::Device).swap(m_targets);
            hipLaunchKernelGGL(gpu::capacity::kernel::targets, dim3(div_up(N, 256)), dim3(256), 0, stream, counts,
                               lbvh.getPrimitives().get(), N, m_targets.get());
            // a random sample order, consumed batch by batch
            const unsigned int pool = std::min<unsigned int>(N, std::max<unsigned int>(batch * 8, 1u << 18));
            std::vector<uint32_t> order(N);
            for (unsigned int i = 0; i < N; ++i) order[i] = i;
            std::mt19937_64 rng(seed);
            for (unsigned int i = 0; i < pool; ++i)
                {
                std::uniform_int_distribution<unsigned int> U(i, N - 1);
                std::swap(order[i], order[U(rng)]);
                }
            if (pool > m_sample.size()) shared_array<uint32_t>(pool, MemoryKind::Device).swap(m_sample);
            NEIGHBOR_HIP_CHECK(hipMemcpyAsync(m_sample.get(), order.data(), pool * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
            unsigned int at = 0;
            for (unsigned int s = 0; s < steps; ++s)
                {
                if (at + batch > pool) at = 0;
                hipLaunchKernelGGL(gpu::capacity::kernel::grad, dim3(div_up(batch, 256)), dim3(256), 0, stream, m_params.get(),
                                   m_features.get(), m_targets.get(), m_sample.get() + at, batch, m_grad.get());
                float loss = 0.0f;
                if (s + 1 == steps)
                    {
                    NEIGHBOR_HIP_CHECK(hipMemcpyAsync(&loss, m_grad.get() + kParams, sizeof(float), hipMemcpyDeviceToHost, stream));
                    NEIGHBOR_HIP_CHECK(hipStreamSynchronize(stream));
                    m_last_loss = loss;
                    }
                ++m_step;
                hipLaunchKernelGGL(gpu::capacity::kernel::adam, dim3(div_up(kParams + 1, 256)), dim3(256), 0, stream, m_params.get(),
                                   m_grad.get(), m_m.get(), m_v.get(), lr, m_step);
                at += batch;
                }
            check_launch("capacity train");
            NEIGHBOR_HIP_CHECK(hipStreamSynchronize(stream));
            return m_last_loss;
            }

        //! Margin (in log2 units) so that a fraction q of the training targets fit under the prediction.
        float calibrateMargin(const LBVH& lbvh, float q, stream_t stream)
            {
            const unsigned int N = lbvh.getN();
            if (N == 0 || m_n != N || m_targets.size() < N) throw std::runtime_error("train() first");
            if (N > m_pred.size()) shared_array<float>(N, MemoryKind::Device).swap(m_pred);
            hipLaunchKernelGGL(gpu::capacity::kernel::predict, dim3(div_up(N, 256)), dim3(256), 0, stream, m_params.get(),
                               m_features.get(), N, m_pred.get());
            check_launch("capacity predict");
            std::vector<float> y(N), t(N);
            NEIGHBOR_HIP_CHECK(hipMemcpyAsync(y.data(), m_pred.get(), N * sizeof(float), hipMemcpyDeviceToHost, stream));
            NEIGHBOR_HIP_CHECK(hipMemcpyAsync(t.data(), m_targets.get(), N * sizeof(float), hipMemcpyDeviceToHost, stream));
            NEIGHBOR_HIP_CHECK(hipStreamSynchronize(stream));
            std::vector<float> r(N);
            for (unsigned int i = 0; i < N; ++i) r[i] = t[i] - y[i];
            const size_t k = std::min<size_t>(N - 1, static_cast<size_t>(std::ceil(q * (N - 1))));
            std::nth_element(r.begin(), r.begin() + k, r.end());
            return std::max(r[k], 0.0f);
            }

        //! Capacity per query (query = primitive index): ceil(2^(y + margin)) - 1 + pad.
        void predictCapacity(const LBVH& lbvh, float margin, unsigned int pad, uint32_t* cap, stream_t stream)
            {
            const unsigned int N = lbvh.getN();
            if (N == 0) return;
            if (m_n != N) throw std::runtime_error("computeFeatures() first");
            hipLaunchKernelGGL(gpu::capacity::kernel::capacity, dim3(div_up(N, 256)), dim3(256), 0, stream, m_params.get(),
                               m_features.get(), lbvh.getPrimitives().get(), N, margin, pad, cap);
            check_launch("capacity");
            }

        float lastLoss() const { return m_last_loss; }

    private:
        //! Cell counts of the sorted codes at the two feature levels and the per-build constants.
        gpu::capacity::Levels levels(const LBVH& lbvh, const float3& lo, const float3& hi, float r, stream_t stream)
            {
            const unsigned int N = lbvh.getN();
            const float Lx = hi.x - lo.x, Ly = hi.y - lo.y, Lz = hi.z - lo.z;
            const float Lmin = std::min(Lx, std::min(Ly, Lz));
            int mA = static_cast<int>(std::floor(std::log2(std::max(Lmin / (2.0f * std::max(r, 1e-30f)), 1.0f))));
            mA = std::max(1, std::min(mA, 6));
            const size_t cellsA = size_t(1) << (3 * mA), cellsB = size_t(1) << (3 * (mA + 1));
            // counts of both levels, then their 27-cell neighborhood sums
            if (2 * (cellsA + cellsB) > m_counts.size())
                shared_array<uint32_t>(2 * (cellsA + cellsB), MemoryKind::Device).swap(m_counts);
            hipLaunchKernelGGL(gpu::radix::kernel::fill_zero, dim3(std::min<size_t>(1024, div_up(cellsA + cellsB, 256))), dim3(256), 0,
                               stream, m_counts.get(), cellsA + cellsB);
            const uint32_t* codes = lbvh.getCodes().get();
            hipLaunchKernelGGL(gpu::capacity::kernel::cell_histogram, dim3(div_up(N, 256)), dim3(256), 0, stream, codes, N,
                               static_cast<unsigned int>(mA), m_counts.get());
            hipLaunchKernelGGL(gpu::capacity::kernel::cell_histogram, dim3(div_up(N, 256)), dim3(256), 0, stream, codes, N,
                               static_cast<unsigned int>(mA + 1), m_counts.get() + cellsA);
            uint32_t* nb = m_counts.get() + cellsA + cellsB;
            hipLaunchKernelGGL(gpu::capacity::kernel::cell_neighborhoods, dim3(div_up(cellsA + cellsB, 256)), dim3(256), 0, stream,
                               m_counts.get(), static_cast<unsigned int>(mA), static_cast<uint32_t>(cellsA),
                               static_cast<uint32_t>(cellsB), nb);
            gpu::capacity::Levels lv;
            lv.mA = static_cast<unsigned int>(mA);
            lv.countsA = m_counts.get();
            lv.countsB = m_counts.get() + cellsA;
            lv.nbA = nb;
            lv.nbB = nb + cellsA;
            lv.expectA = static_cast<float>(N) / static_cast<float>(cellsA);
            lv.expectB = static_cast<float>(N) / static_cast<float>(cellsB);
            lv.lo = lo;
            const float sB = static_cast<float>(1 << (mA + 1));
            lv.cellB = make_float3(Lx / sB, Ly / sB, Lz / sB);
            const double volume = static_cast<double>(Lx) * Ly * Lz;
            const double expected = static_cast<double>(N) / volume * 4.0 / 3.0 * 3.14159265358979 * r * r * r;
            lv.log_expected = static_cast<float>(std::log2(std::max(expected, 1e-6)));
            return lv;
            }

        shared_array<float> m_params;
        shared_array<unsigned long long> m_group_total;   //!< per-group capacity sums of predictRows
        shared_array<float> m_grad;
        shared_array<float> m_m;
        shared_array<float> m_v;
        shared_array<float> m_features;
        shared_array<float> m_targets;
        shared_array<float> m_pred;
        shared_array<uint32_t> m_counts;
        shared_array<uint32_t> m_sample;
        int m_step;
        unsigned int m_n;
        float m_last_loss = 0.0f;
    };

} // end namespace neighbor

#endif // NEIGHBOR_CAPACITY_PREDICTOR_H_
