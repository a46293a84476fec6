// OutputOps.h -- per-query output of a traversal.
//
// Concept (reference include/neighbor/OutputOps.h:13-27):
//   struct ThreadData;
//   template<class QueryDataT> __device__ ThreadData setup(unsigned int idx, const QueryDataT&) const;
//   __device__ void process(ThreadData&, int primitive) const;
//   __device__ void finalize(const ThreadData&) const;
// Extension used by this library: an output op may instead provide
//   __device__ void process(ThreadData&, int primitive, unsigned int image) const;
// which the traversal prefers when present, handing over the image index.
#ifndef NEIGHBOR_OUTPUT_OPS_H_
#define NEIGHBOR_OUTPUT_OPS_H_

#include "ApproximateMath.h"

namespace neighbor
{

//! Count the hits of each query (OutputOps.h:28-103).
struct CountNeighborsOp
    {
    CountNeighborsOp(unsigned int* nneigh_) : nneigh(nneigh_) {}

    struct ThreadData
        {
        __device__ ThreadData(unsigned int idx_) : idx(idx_), num_neigh(0) {}
        unsigned int idx;
        unsigned int num_neigh;
        };

    template<class QueryDataT>
    __device__ __forceinline__ ThreadData setup(const unsigned int idx, const QueryDataT&) const { return ThreadData(idx); }

    __device__ __forceinline__ void process(ThreadData& t, const int) const { ++t.num_neigh; }

    __device__ __forceinline__ void finalize(const ThreadData& t) const { nneigh[t.idx] = t.num_neigh; }

    unsigned int* nneigh;
    };

//! Fixed-stride neighbor list (OutputOps.h:113-176): row idx starts at max_neigh*idx,
//! entries beyond max_neigh are dropped but still counted.  The row offset is
//! computed in 64 bits (the reference's 32-bit product wraps once N*max_neigh >= 2^32).
struct NeighborListOp
    {
    NeighborListOp(unsigned int* neigh_list_, unsigned int* nneigh_, unsigned int max_neigh_)
        : neigh_list(neigh_list_), nneigh(nneigh_), max_neigh(max_neigh_) {}

    struct ThreadData
        {
        __device__ ThreadData(unsigned int idx_, unsigned int* row_) : idx(idx_), row(row_), num_neigh(0) {}
        unsigned int idx;
        unsigned int* row;
        unsigned int num_neigh;
        };

    template<class QueryDataT>
    __device__ __forceinline__ ThreadData setup(const unsigned int idx, const QueryDataT&) const
        {
        return ThreadData(idx, neigh_list + static_cast<size_t>(max_neigh) * idx);
        }

    __device__ __forceinline__ void process(ThreadData& t, const int primitive) const
        {
        if (t.num_neigh < max_neigh) t.row[t.num_neigh] = static_cast<unsigned int>(primitive);
        ++t.num_neigh;
        }

    __device__ __forceinline__ void finalize(const ThreadData& t) const { nneigh[t.idx] = t.num_neigh; }

    unsigned int* neigh_list;
    unsigned int* nneigh;
    const unsigned int max_neigh;
    };

//! Compact pair list: hits of query idx go to [offsets[idx], offsets[idx+1]) as
//! (neighbor, image) pairs, in emission order.  offsets is normally the exclusive
//! scan of a CountNeighborsOp pass of the same traversal; writes past offsets[idx+1]
//! or past the buffer capacity are dropped and still counted (nneigh may be null).
struct PairListOp
    {
    //! The wave walk may stage hits in LDS and write them to neigh/image[pos, end) in runs.
    static constexpr bool kStaged = true;

    PairListOp(const unsigned long long* offsets_, unsigned int* neigh_, unsigned char* image_, unsigned int* nneigh_,
               unsigned long long capacity_)
        : offsets(offsets_), neigh(neigh_), image(image_), nneigh(nneigh_), capacity(capacity_) {}

    struct ThreadData
        {
        __device__ ThreadData(unsigned int idx_, unsigned long long first_, unsigned long long end_)
            : idx(idx_), pos(first_), end(end_), num_neigh(0) {}
        unsigned int idx;
        unsigned long long pos;
        unsigned long long end;
        unsigned int num_neigh;
        };

    template<class QueryDataT>
    __device__ __forceinline__ ThreadData setup(const unsigned int idx, const QueryDataT&) const
        {
        const unsigned long long end = offsets[idx + 1];
        return ThreadData(idx, offsets[idx], end < capacity ? end : capacity);
        }

    __device__ __forceinline__ void process(ThreadData& t, const int primitive, const unsigned int k) const
        {
        if (t.pos < t.end)
            {
            neigh[t.pos] = static_cast<unsigned int>(primitive);
            if (image) image[t.pos] = static_cast<unsigned char>(k);
            }
        ++t.pos;
        ++t.num_neigh;
        }

    __device__ __forceinline__ void finalize(const ThreadData& t) const
        {
        if (nneigh) nneigh[t.idx] = t.num_neigh;
        }

    const unsigned long long* offsets;
    unsigned int* neigh;
    unsigned char* image;
    unsigned int* nneigh;
    unsigned long long capacity;
    };

//! Pair rows given by a start and a length per query: hits of query idx go to
//! [start[idx], start[idx] + limit[idx]) (and below capacity) as (neighbor, image) pairs.
//! Used for rows sized by a predicted capacity.
//!
//! Without a spill counter, hits past the row are dropped but counted.  With one
//! (spill[0] = first free position past all rows, spill[1] = 0 on entry), a query whose
//! count exceeds its row takes a new row of exactly its count at spill[0] (atomically),
//! start[idx] and limit[idx] are rewritten, and the traversal walks that query again
//! into the new row: its exact count, then its fill, for overflowed rows only.  On exit
//! spill[0] is the end of the used positions and spill[1] the number of moved rows.
struct RowPairOp
    {
    static constexpr bool kStaged = true;

    RowPairOp(unsigned long long* start_, unsigned int* limit_, unsigned int* neigh_, unsigned char* image_,
              unsigned int* nneigh_, unsigned long long capacity_, unsigned long long* spill_ = nullptr)
        : start(start_), limit(limit_), neigh(neigh_), image(image_), nneigh(nneigh_), capacity(capacity_), spill(spill_) {}

    typedef PairListOp::ThreadData ThreadData;

    template<class QueryDataT>
    __device__ __forceinline__ ThreadData setup(const unsigned int idx, const QueryDataT&) const
        {
        const unsigned long long s = start[idx];
        const unsigned long long e = s + limit[idx];
        return ThreadData(idx, s, e < capacity ? e : capacity);
        }

    __device__ __forceinline__ void process(ThreadData& t, const int primitive, const unsigned int k) const
        {
        if (t.pos < t.end)
            {
            neigh[t.pos] = static_cast<unsigned int>(primitive);
            if (image) image[t.pos] = static_cast<unsigned char>(k);
            }
        ++t.pos;
        ++t.num_neigh;
        }

    //! Move an overflowed row to the spill area; true when the query must be walked again.
    __device__ __forceinline__ bool relocate(ThreadData& t) const
        {
        if (!spill || t.num_neigh <= limit[t.idx]) return false;
        const unsigned long long n = t.num_neigh;
        const unsigned long long s = atomicAdd(&spill[0], n);
        atomicAdd(&spill[1], 1ull);
        start[t.idx] = s;
        limit[t.idx] = t.num_neigh;
        const unsigned long long e = s + n;
        t = ThreadData(t.idx, s, e < capacity ? e : capacity);
        return true;
        }

    __device__ __forceinline__ void finalize(const ThreadData& t) const
        {
        if (nneigh) nneigh[t.idx] = t.num_neigh;
        }

    unsigned long long* start;
    unsigned int* limit;
    unsigned int* neigh;
    unsigned char* image;
    unsigned int* nneigh;
    unsigned long long capacity;
    unsigned long long* spill;
    };

//! RowPairOp whose row starts the traversal places itself, with no scan over the queries:
//! query position t of the traversal order (the tree's primitive order) belongs to group
//! t / kGroup (one 256-thread block of the wave kernels) and starts at base[t / kGroup] plus
//! the limits of the positions before it in its group; that start is written to start[idx]
//! before the walk.  base[g] is the sum of the limits of the positions before kGroup * g
//! (CapacityPredictor::predictRows writes it with the limits).  Only the wave variants, whose
//! kernels run kGroup-thread blocks, place rows; the other traversals refuse this op.
struct PlacedRowPairOp : public RowPairOp
    {
    static constexpr bool kPlacedRows = true;
    static constexpr unsigned int kGroup = 256;

    PlacedRowPairOp(const unsigned long long* base_, unsigned long long* start_, unsigned int* limit_, unsigned int* neigh_,
                    unsigned char* image_, unsigned int* nneigh_, unsigned long long capacity_,
                    unsigned long long* spill_ = nullptr)
        : RowPairOp(start_, limit_, neigh_, image_, nneigh_, capacity_, spill_), base(base_) {}

    //! Called by every thread of a kGroup-thread block (position t, query idx) before the walk.
    //! A block that serves the later part of a group of kGroup positions (lead > 0: positions
    //! group .. group + lead - 1 belong to other blocks; order, n: the launch's query order and
    //! count) places its rows after the capacities of those lead positions.
    __device__ __forceinline__ void place(const unsigned int t, const unsigned int idx, const bool valid,
                                          const unsigned int group = 0u, const unsigned int lead = 0u,
                                          const unsigned int* order = nullptr, const unsigned int n = 0u) const
        {
        __shared__ unsigned long long s_wave[kGroup / 64];
        __shared__ unsigned long long s_lead[kGroup / 64];
        const unsigned long long c = valid ? limit[idx] : 0u;
        const unsigned int lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        unsigned long long incl = c;
        #pragma unroll
        for (unsigned int d = 1; d < 64; d <<= 1)
            {
            const unsigned long long y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
            }
        if (lane == 63u) s_wave[wave] = incl;
        if (lead)
            {
            const unsigned int p = group + threadIdx.x;
            unsigned long long l = (threadIdx.x < lead && p < n) ? limit[order ? order[p] : p] : 0ull;
            #pragma unroll
            for (unsigned int d = 32; d; d >>= 1) l += __shfl_xor(l, d, 64);
            if (lane == 0u) s_lead[wave] = l;
            }
        __syncthreads();
        unsigned long long before = 0;
        for (unsigned int w = 0; w < wave; ++w) before += s_wave[w];
        if (lead)
            for (unsigned int w = 0; w < kGroup / 64; ++w) before += s_lead[w];
        if (valid) start[idx] = base[t / kGroup] + before + (incl - c);
        }

    const unsigned long long* base;
    };

} // end namespace neighbor

#endif // NEIGHBOR_OUTPUT_OPS_H_
