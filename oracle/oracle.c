/*
 * oracle.c -- CPU restatement of the reference LBVH neighbor-search path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with gcc, -O2, no fast-math and
 * -ffp-contract=off so that every float operation is the single IEEE-754
 * round-to-nearest operation the reference's HIP path performs; denormals are
 * preserved (no FTZ/DAZ on x86-64 by default).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* bit helpers                                                               */
/* ------------------------------------------------------------------------- */
static inline uint32_t f2u_bits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f_bits(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* ------------------------------------------------------------------------- */
/* directed rounding: the HIP branch of ApproximateMath.h (lines 19-21) routes
 * every *_rd / *_ru through nextafterf(RN(op), -/+FLT_MAX) (lines 101, 123, 145,
 * 167, 189, 211, 233, 255, 276, 297, 320, 343), i.e. one ulp outward even when
 * the RN result is exact.  double2float_rd/ru (lines 60-61, 78-79) step only
 * when the RN conversion lies on the wrong side.  nextafterf is restated from
 * the IEEE-754 bit layout (C99 7.12.11.3 semantics).                         */
/* ------------------------------------------------------------------------- */
float orc_next_down(float x)
{
    if (x != x) return x;                       /* NaN propagates */
    if (x == -FLT_MAX) return x;                /* already at the target */
    if (x < -FLT_MAX) return -FLT_MAX;          /* -inf steps toward -FLT_MAX */
    if (x == 0.0f) return u2f_bits(0x80000001u); /* +-0 -> -denorm_min */
    uint32_t u = f2u_bits(x);
    u = (x > 0.0f) ? u - 1u : u + 1u;
    return u2f_bits(u);
}

float orc_next_up(float x)
{
    if (x != x) return x;
    if (x == FLT_MAX) return x;
    if (x > FLT_MAX) return FLT_MAX;            /* +inf steps toward FLT_MAX */
    if (x == 0.0f) return u2f_bits(0x00000001u);
    uint32_t u = f2u_bits(x);
    u = (x > 0.0f) ? u + 1u : u - 1u;
    return u2f_bits(u);
}

float orc_fadd_rd(float x, float y) { return orc_next_down(x + y); }
float orc_fadd_ru(float x, float y) { return orc_next_up(x + y); }
float orc_fsub_rd(float x, float y) { return orc_next_down(x - y); }
float orc_fsub_ru(float x, float y) { return orc_next_up(x - y); }
float orc_fmul_rd(float x, float y) { return orc_next_down(x * y); }
float orc_fmul_ru(float x, float y) { return orc_next_up(x * y); }
float orc_fdiv_rd(float x, float y) { return orc_next_down(x / y); }
float orc_fdiv_ru(float x, float y) { return orc_next_up(x / y); }
float orc_frcp_rd(float x) { return orc_next_down(1.0f / x); }
float orc_frcp_ru(float x) { return orc_next_up(1.0f / x); }
float orc_fmaf_rd(float x, float y, float z) { return orc_next_down(fmaf(x, y, z)); }
float orc_fmaf_ru(float x, float y, float z) { return orc_next_up(fmaf(x, y, z)); }

float orc_double2float_rd(double x)
{
    float y = (float)x;
    return ((double)y <= x) ? y : orc_next_down(y);
}

float orc_double2float_ru(double x)
{
    float y = (float)x;
    return ((double)y >= x) ? y : orc_next_up(y);
}

/* The reference converts floorf(...) results to unsigned int
 * (kernels/LBVHTraverser.cuh:100-110).  On a degenerate axis that value is -1.0f,
 * which gfx9 hardware (v_cvt_u32_f32) saturates to 0; restated here explicitly:
 * NaN and values <= 0 map to 0, values >= 2^32 map to 0xffffffff, the rest truncate. */
uint32_t orc_sat_f2u(float x)
{
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}

/* ------------------------------------------------------------------------- */
/* insert ops (InsertOps.h:35-41 PointInsertOp::get, :75-82 SphereInsertOp::get)
 * and BoundingBox::getCenter (BoundingVolumes.h:64-72).                       */
/* ------------------------------------------------------------------------- */
void orc_leaf_box(const float* xyz, uint32_t idx, float sphere_r, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a) {
        const float p = xyz[3 * (size_t)idx + a];
        if (sphere_r < 0.0f) { lo[a] = p; hi[a] = p; }
        else { lo[a] = p - sphere_r; hi[a] = p + sphere_r; }
    }
}

/* Spread the low 10 bits of v so that two zero bits follow each bit
 * (classic "part1by2" shift-and-mask formulation). */
static uint32_t spread3(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

/* kernels/LBVH.cuh:49-52: x is the most significant of each bit triple. */
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}

/* kernels/LBVH.cuh:105-108 fractionToBin */
static uint32_t fraction_to_bin(float f)
{
    const float s = fminf(fmaxf(f * 1023.0f, 0.0f), 1023.0f);
    return (uint32_t)s;
}

/* kernels/LBVH.cuh:127-157 lbvh_gen_codes: center of the leaf box, fraction by
 * IEEE division, 10-bit bins, Morton interleave. */
void orc_morton_codes(const float* xyz, uint32_t N, float sphere_r,
                      const float lo[3], const float hi[3], uint32_t* codes)
{
    for (uint32_t i = 0; i < N; ++i) {
        float blo[3], bhi[3];
        orc_leaf_box(xyz, i, sphere_r, blo, bhi);
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            const float c = 0.5f * (blo[a] + bhi[a]);
            const float f = (c - lo[a]) / (hi[a] - lo[a]);
            q[a] = fraction_to_bin(f);
        }
        codes[i] = orc_morton3(q[0], q[1], q[2]);
    }
}

/* ------------------------------------------------------------------------- */
/* sort (kernels/LBVH.cuh:421-443): stable ascending sort of 30-bit keys with
 * values = original index, i.e. order by (code, index).                     */
/* ------------------------------------------------------------------------- */
static int cmp_u64(const void* a, const void* b)
{
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

void orc_sort_codes(uint32_t N, const uint32_t* codes, uint32_t* sorted_codes, uint32_t* primitive)
{
    uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * (N ? N : 1));
    for (uint32_t i = 0; i < N; ++i) k[i] = ((uint64_t)(codes[i] & 0x3fffffffu) << 32) | i;
    qsort(k, N, sizeof(uint64_t), cmp_u64);
    for (uint32_t i = 0; i < N; ++i) {
        sorted_codes[i] = (uint32_t)(k[i] >> 32);
        primitive[i] = (uint32_t)(k[i] & 0xffffffffu);
    }
    free(k);
}

/* ------------------------------------------------------------------------- */
/* Karras 2012, Fig. 4 (kernels/LBVH.cuh:73-94 delta, :169-243 lbvh_gen_tree).  */
/* ------------------------------------------------------------------------- */
static int karras_delta(const uint32_t* c, uint32_t N, int i, uint32_t ci, int j)
{
    if (j < 0 || j >= (int)N) return -1;
    const uint32_t cj = c[j];
    if (ci == cj) return 32 + clz32((uint32_t)(i ^ j));
    return clz32(ci ^ cj);
}

void orc_karras(uint32_t N, const uint32_t* c, int32_t* parent, int32_t* left, int32_t* right)
{
    if (N < 2) { if (N == 1) parent[0] = ORC_SENTINEL; return; }
    const int n = (int)N;
    for (int i = 0; i < n - 1; ++i) {
        const uint32_t ci = c[i];
        const int fwd = karras_delta(c, N, i, ci, i + 1);
        const int bwd = karras_delta(c, N, i, ci, i - 1);
        const int d = (fwd >= bwd) ? 1 : -1;
        const int dmin = karras_delta(c, N, i, ci, i - d);

        int lmax = 2;
        while (karras_delta(c, N, i, ci, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (karras_delta(c, N, i, ci, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = karras_delta(c, N, i, ci, j);

        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (s + t < l && karras_delta(c, N, i, ci, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int split = i + s * d + (d < 0 ? d : 0);

        const int lo_end = (i < j) ? i : j, hi_end = (i < j) ? j : i;
        const int lc = (lo_end == split) ? split + (n - 1) : split;
        const int rc = (hi_end == split + 1) ? split + n : split + 1;
        left[i] = lc;
        right[i] = rc;
        parent[lc] = i;
        parent[rc] = i;
    }
    parent[0] = ORC_SENTINEL;
}

/* ------------------------------------------------------------------------- */
/* AABB bubble (kernels/LBVH.cuh:270-327): leaf box stored at N-1+idx, internal
 * boxes are the union of the children, merged with strict < / > comparisons.
 * The reference merges (own, sibling) in arrival order; this restatement merges
 * in the fixed order (left, right), one of the orders the reference may take
 * (the two orders differ only in the sign of a zero bound).                    */
/* ------------------------------------------------------------------------- */
void orc_bubble(uint32_t N, const int32_t* parent, const int32_t* left, const int32_t* right,
                const uint32_t* primitive, const float* xyz, float sphere_r, float* lo, float* hi)
{
    (void)parent;
    if (N == 0) return;
    const uint32_t Ni = N - 1;
    for (uint32_t i = 0; i < N; ++i) orc_leaf_box(xyz, primitive[i], sphere_r, &lo[3 * (size_t)(Ni + i)], &hi[3 * (size_t)(Ni + i)]);
    if (N == 1) return;

    /* iterative post-order from the root */
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)N);
    uint8_t* seen = (uint8_t*)calloc(Ni, 1);
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int32_t nd = stack[sp - 1];
        if (!seen[nd]) {
            seen[nd] = 1;
            if (right[nd] < (int32_t)Ni) stack[sp++] = right[nd];
            if (left[nd] < (int32_t)Ni) stack[sp++] = left[nd];
            continue;
        }
        --sp;
        const float* l_lo = &lo[3 * (size_t)left[nd]];
        const float* l_hi = &hi[3 * (size_t)left[nd]];
        const float* r_lo = &lo[3 * (size_t)right[nd]];
        const float* r_hi = &hi[3 * (size_t)right[nd]];
        for (int a = 0; a < 3; ++a) {
            float b = l_lo[a]; if (r_lo[a] < b) b = r_lo[a]; lo[3 * (size_t)nd + a] = b;
            float e = l_hi[a]; if (r_hi[a] > e) e = r_hi[a]; hi[3 * (size_t)nd + a] = e;
        }
    }
    free(seen);
    free(stack);
}

void orc_build(const float* xyz, uint32_t N, float sphere_r, const float scene_lo[3], const float scene_hi[3],
               uint32_t* sorted_codes, uint32_t* primitive, int32_t* parent, int32_t* left, int32_t* right,
               float* lo, float* hi)
{
    if (N == 0) return;
    if (N == 1) {
        /* LBVH.h:252-257 + kernels/LBVH.cuh:340-354: one leaf at node 0.  The
         * reference never writes primitive[0] here; its freshly allocated managed
         * buffer reads 0, which test/lbvh_test.cu:541-543 asserts. */
        uint32_t code;
        orc_morton_codes(xyz, 1, sphere_r, scene_lo, scene_hi, &code);
        sorted_codes[0] = code;
        primitive[0] = 0;
        parent[0] = ORC_SENTINEL;
        orc_leaf_box(xyz, 0, sphere_r, lo, hi);
        return;
    }
    uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * N);
    orc_morton_codes(xyz, N, sphere_r, scene_lo, scene_hi, codes);
    orc_sort_codes(N, codes, sorted_codes, primitive);
    free(codes);
    orc_karras(N, sorted_codes, parent, left, right);
    orc_bubble(N, parent, left, right, primitive, xyz, sphere_r, lo, hi);
}

/* ------------------------------------------------------------------------- */
/* compression (kernels/LBVHTraverser.cuh:45-125; layout LBVHTraverser.h:459-480) */
/* ------------------------------------------------------------------------- */
void orc_compress(uint32_t N, const int32_t* parent, const int32_t* left, const int32_t* right,
                  const uint32_t* primitive, const float* lo, const float* hi, const uint32_t* map,
                  int32_t* records, float root_lo[3], float root_hi[3], float bins[3])
{
    if (N == 0) return;
    const uint32_t Ni = N - 1, Nn = 2 * N - 1;
    float bininv[3];
    for (int a = 0; a < 3; ++a) {
        root_lo[a] = lo[a];
        root_hi[a] = hi[a];
        float L = orc_fsub_ru(root_hi[a], root_lo[a]);
        if (L <= 0.0f) L = 1.0f;
        bininv[a] = orc_fdiv_rd(1023.0f, L);
    }
    for (uint32_t idx = 0; idx < Nn; ++idx) {
        /* rope: right sibling of the first ancestor-or-self that is a left child */
        int32_t rope = ORC_SENTINEL;
        int32_t cur = (int32_t)idx;
        while (cur != 0 && rope == ORC_SENTINEL) {
            const int32_t p = parent[cur];
            if (left[p] == cur) rope = right[p];
            else cur = p;
        }
        uint32_t lb[3], hb[3];
        for (int a = 0; a < 3; ++a) {
            lb[a] = orc_sat_f2u(floorf(orc_fmul_rd(orc_fsub_rd(lo[3 * (size_t)idx + a], root_lo[a]), bininv[a])));
            hb[a] = orc_sat_f2u(floorf(orc_fmul_rd(orc_fsub_rd(root_hi[a], hi[3 * (size_t)idx + a]), bininv[a])));
        }
        const uint32_t lo3 = (lb[0] << 20) + (lb[1] << 10) + lb[2];
        const uint32_t hi3 = (hb[0] << 20) + (hb[1] << 10) + hb[2];
        int32_t lf;
        if (idx < Ni) lf = left[idx];
        else {
            const uint32_t p = primitive[idx - Ni];
            lf = (int32_t)~(map ? map[p] : p);
        }
        records[4 * (size_t)idx + 0] = (int32_t)lo3;
        records[4 * (size_t)idx + 1] = (int32_t)hi3;
        records[4 * (size_t)idx + 2] = lf;
        records[4 * (size_t)idx + 3] = rope;
    }
    for (int a = 0; a < 3; ++a) bins[a] = orc_frcp_rd(bininv[a]);
}

/* ------------------------------------------------------------------------- */
/* BoundingSphere float ctor (BoundingVolumes.h:119-123) and overlap
 * (BoundingVolumes.h:163-171); BoundingBox::overlap is not used on this path. */
/* ------------------------------------------------------------------------- */
static int sphere_box_overlap(const float o[3], float Rsq, const float blo[3], const float bhi[3])
{
    float dr[3];
    for (int a = 0; a < 3; ++a) dr[a] = orc_fsub_rd(fminf(fmaxf(o[a], blo[a]), bhi[a]), o[a]);
    const float dr2 = orc_fmaf_rd(dr[0], dr[0], orc_fmaf_rd(dr[1], dr[1], orc_fmul_rd(dr[2], dr[2])));
    return dr2 <= Rsq;
}

/* BoundingSphere double ctor (BoundingVolumes.h:135-147): the origin is the per-axis
 * round-down of o; the radius is the round-up of r padded (rounded up) by the widest
 * per-axis interval [round-down, round-up] of o, so the float sphere encloses the double
 * one.  Rsq = fmul_ru(R, R). */
void orc_sphere_double(const double o[3], double r, float origin[3], float* Rsq)
{
    float w = 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float down = orc_double2float_rd(o[a]);
        const float up = orc_double2float_ru(o[a]);
        const float wa = orc_fsub_ru(up, down);
        w = (a == 0) ? wa : fmaxf(w, wa);
        origin[a] = down;
    }
    const float R = orc_fadd_ru(orc_double2float_ru(r), w);
    *Rsq = orc_fmul_ru(R, R);
}

/* BoundingSphere float ctor (BoundingVolumes.h:119-123): origin o, Rsq = fmul_ru(r, r). */
void orc_sphere_float(const float o[3], float r, float origin[3], float* Rsq)
{
    for (int a = 0; a < 3; ++a) origin[a] = o[a];
    *Rsq = orc_fmul_ru(r, r);
}

int orc_sphere_overlap(const float origin[3], float Rsq, const float lo[3], const float hi[3])
{
    return sphere_box_overlap(origin, Rsq, lo, hi);
}

/* rope traversal (kernels/LBVHTraverser.cuh:155-252) with SphereQueryOp
 * (QueryOps.h:43-116), ImageListOp (TranslateOps.h:54-91), CountNeighborsOp /
 * NeighborListOp (OutputOps.h:28-176). */
uint64_t orc_traverse(uint32_t N, const int32_t* records, const float root_lo[3], const float root_hi[3],
                      const float bins[3], const float* spheres, uint32_t Nq, const float* images, uint32_t Nimg,
                      uint32_t* counts, uint32_t* list, uint32_t max_neigh,
                      const uint64_t* trip_offsets, uint32_t* trip_prim, uint8_t* trip_image,
                      uint64_t* node_visits)
{
    uint64_t total = 0, visits = 0;
    if (N == 0 || Nq == 0 || Nimg == 0) { if (node_visits) *node_visits = 0; return 0; }
    const uint32_t nbits = Nimg <= 32u ? Nimg : 32u;
    for (uint32_t q = 0; q < Nq; ++q) {
        const float* s = &spheres[4 * (size_t)q];
        const float Rsq = orc_fmul_ru(s[3], s[3]);
        uint32_t flags = 0;
        for (uint32_t k = 0; k < nbits; ++k) {
            const float o[3] = { s[0] + images[3 * k + 0], s[1] + images[3 * k + 1], s[2] + images[3 * k + 2] };
            if (sphere_box_overlap(o, Rsq, root_lo, root_hi)) flags |= 1u << k;
        }
        uint32_t count = 0;
        uint64_t tpos = trip_offsets ? trip_offsets[q] : 0;
        while (flags) {
            const uint32_t k = (uint32_t)__builtin_ctz(flags);
            flags &= ~(1u << k);
            const float o[3] = { s[0] + images[3 * k + 0], s[1] + images[3 * k + 1], s[2] + images[3 * k + 2] };
            int32_t node = 0;
            while (node != ORC_SENTINEL) {
                const int32_t* r = &records[4 * (size_t)node];
                ++visits;
                const uint32_t lo3 = (uint32_t)r[0], hi3 = (uint32_t)r[1];
                const uint32_t lbin[3] = { (lo3 >> 20) & 0x3ffu, (lo3 >> 10) & 0x3ffu, lo3 & 0x3ffu };
                const uint32_t hbin[3] = { (hi3 >> 20) & 0x3ffu, (hi3 >> 10) & 0x3ffu, hi3 & 0x3ffu };
                float blo[3], bhi[3];
                for (int a = 0; a < 3; ++a) {
                    blo[a] = orc_fadd_rd(root_lo[a], orc_fmul_rd((float)lbin[a], bins[a]));
                    bhi[a] = orc_fsub_ru(root_hi[a], orc_fmul_rd((float)hbin[a], bins[a]));
                }
                const int32_t lf = r[2];
                node = r[3];
                if (sphere_box_overlap(o, Rsq, blo, bhi)) {
                    if (lf < 0) {
                        const uint32_t prim = (uint32_t)~lf;
                        if (list && count < max_neigh) list[(size_t)max_neigh * q + count] = prim;
                        if (trip_prim) trip_prim[tpos] = prim;
                        if (trip_image) trip_image[tpos] = (uint8_t)k;
                        ++tpos;
                        ++count;
                    } else {
                        node = lf;
                    }
                }
            }
        }
        counts[q] = count;
        total += count;
    }
    if (node_visits) *node_visits = visits;
    return total;
}

/* ------------------------------------------------------------------------- */
/* exact FP64 pairs                                                          */
/* ------------------------------------------------------------------------- */
static inline int exact_hit(const float* q, const double* t, const float* x, double R2)
{
    const double dx = ((double)q[0] + t[0]) - (double)x[0];
    const double dy = ((double)q[1] + t[1]) - (double)x[1];
    const double dz = ((double)q[2] + t[2]) - (double)x[2];
    const double d2 = (dx * dx + dy * dy) + dz * dz;
    return d2 <= R2;
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* The reference's only CPU path is the O(N^2) loop of test/lbvh_test.cu:459-490; this is its
 * FP64 restatement over explicit images.  Queries are independent: with OpenMP they are split
 * over the threads of orc_set_threads (each writes its own row at offsets[q]). */
uint64_t orc_exact_bruteforce(const float* xyz, uint32_t N, const float* queries, uint32_t Nq,
                              const double* images, uint32_t Nimg, uint32_t* counts,
                              const uint64_t* offsets, uint32_t* out_j, uint8_t* out_k)
{
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t qq = 0; qq < (int64_t)Nq; ++qq) {
        const uint32_t q = (uint32_t)qq;
        const float* s = &queries[4 * (size_t)q];
        const double R2 = (double)s[3] * (double)s[3];
        uint64_t pos = offsets ? offsets[q] : 0;
        uint32_t c = 0;
        for (uint32_t k = 0; k < Nimg; ++k)
            for (uint32_t j = 0; j < N; ++j)
                if (exact_hit(s, &images[3 * k], &xyz[3 * (size_t)j], R2)) {
                    if (out_j) out_j[pos] = j;
                    if (out_k) out_k[pos] = (uint8_t)k;
                    ++pos;
                    ++c;
                }
        counts[q] = c;
        total += c;
    }
    return total;
}

typedef struct { uint32_t k, j; } kj_t;
static int cmp_kj(const void* a, const void* b)
{
    const kj_t* x = (const kj_t*)a; const kj_t* y = (const kj_t*)b;
    if (x->k != y->k) return (x->k > y->k) - (x->k < y->k);
    return (x->j > y->j) - (x->j < y->j);
}

static int cell_of(double x, double cs, int n)
{
    int c = (int)floor(x / cs);
    if (c < 0) c = 0;
    if (c >= n) c = n - 1;
    return c;
}

uint64_t orc_exact_celllist(const float* xyz, uint32_t N, const float* queries, uint32_t Nq,
                            const double box[3], double rmax, uint32_t* counts,
                            const uint64_t* offsets, uint32_t* out_j, uint8_t* out_k, int nthreads)
{
    int n[3];
    double cs[3];
    for (int a = 0; a < 3; ++a) {
        n[a] = (int)floor(box[a] / (rmax * (1.0 + 1e-9)));
        if (n[a] < 3) n[a] = 0;
        cs[a] = n[a] ? box[a] / n[a] : box[a];
    }
    double images[27 * 3];
    {
        int idx = 0;
        for (int ix = -1; ix <= 1; ++ix)
            for (int iy = -1; iy <= 1; ++iy)
                for (int iz = -1; iz <= 1; ++iz) {
                    images[3 * idx + 0] = ix * box[0];
                    images[3 * idx + 1] = iy * box[1];
                    images[3 * idx + 2] = iz * box[2];
                    ++idx;
                }
    }
    if (!n[0] || !n[1] || !n[2])
        return orc_exact_bruteforce(xyz, N, queries, Nq, images, 27, counts, offsets, out_j, out_k);

    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    uint32_t* head = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
    uint32_t* cellof = (uint32_t*)malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t* members = (uint32_t*)malloc(sizeof(uint32_t) * (N ? N : 1));
    for (uint32_t j = 0; j < N; ++j) {
        const size_t c = ((size_t)cell_of(xyz[3 * (size_t)j], cs[0], n[0]) * n[1] + cell_of(xyz[3 * (size_t)j + 1], cs[1], n[1])) * n[2]
                         + cell_of(xyz[3 * (size_t)j + 2], cs[2], n[2]);
        cellof[j] = (uint32_t)c;
        head[c + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) head[c + 1] += head[c];
    {
        uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * ncell);
        memcpy(fill, head, sizeof(uint32_t) * ncell);
        for (uint32_t j = 0; j < N; ++j) members[fill[cellof[j]]++] = j;
        free(fill);
    }
    uint64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(+ : total)
    {
        size_t cap = 4096;
        kj_t* buf = (kj_t*)malloc(sizeof(kj_t) * cap);
#pragma omp for schedule(dynamic, 256)
        for (int64_t qq = 0; qq < (int64_t)Nq; ++qq) {
            const uint32_t q = (uint32_t)qq;
            const float* s = &queries[4 * (size_t)q];
            const double R2 = (double)s[3] * (double)s[3];
            const int cq[3] = { cell_of(s[0], cs[0], n[0]), cell_of(s[1], cs[1], n[1]), cell_of(s[2], cs[2], n[2]) };
            size_t m = 0;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        int c[3] = { cq[0] + dx, cq[1] + dy, cq[2] + dz };
                        int img[3];
                        for (int a = 0; a < 3; ++a) {
                            img[a] = 0;
                            if (c[a] < 0) { c[a] += n[a]; img[a] = 1; }     /* neighbor sits at x - L: image +L */
                            else if (c[a] >= n[a]) { c[a] -= n[a]; img[a] = -1; }
                        }
                        const uint32_t k = (uint32_t)((img[0] + 1) * 9 + (img[1] + 1) * 3 + (img[2] + 1));
                        const size_t cell = ((size_t)c[0] * n[1] + c[1]) * n[2] + c[2];
                        for (uint32_t e = head[cell]; e < head[cell + 1]; ++e) {
                            const uint32_t j = members[e];
                            if (exact_hit(s, &images[3 * k], &xyz[3 * (size_t)j], R2)) {
                                if (m == cap) { cap *= 2; buf = (kj_t*)realloc(buf, sizeof(kj_t) * cap); }
                                buf[m].k = k; buf[m].j = j; ++m;
                            }
                        }
                    }
            counts[q] = (uint32_t)m;
            total += m;
            if (offsets && (out_j || out_k)) {
                qsort(buf, m, sizeof(kj_t), cmp_kj);
                uint64_t pos = offsets[q];
                for (size_t e = 0; e < m; ++e, ++pos) {
                    if (out_j) out_j[pos] = buf[e].j;
                    if (out_k) out_k[pos] = (uint8_t)buf[e].k;
                }
            }
        }
        free(buf);
    }
    free(members);
    free(cellof);
    free(head);
    return total;
}
