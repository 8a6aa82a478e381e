// Extend-add of one target tile of a large front (see k_extadd_tiles),
// shared by the extend-add launch and the cholesky launch that also
// factors a level's small fronts.
#pragma once
#include "common.h"

// Deterministic extend-add for large fronts.  One workgroup owns the target
// tile rows [r0, r0+256) x columns [c0, c0+16) of front k (lower part).  The
// tile is read into LDS once (its F22 entries start at zero when zero22 is
// set), every child's update is added into it there, in reverse child order
// (the summation order of cholesky.c:41-52), and the tile is written back
// once: each target cache line is read and written once per launch instead
// of once per child, and no atomics are needed.  (llt.c:36-46 adds into an
// F22 already holding the product, so zeroing is optional.)  Relative
// indices increase, so a child's rows and columns landing in the tile are
// contiguous index ranges (their bounds come with the tile list); thread i
// adds child row i's up to 16 columns.
// Target (row, col) with col < nn lives in the blkval block (ld nj); with
// col >= nn in k's update slot (ld na).
#define XT_LD 17
#define XT_TPB 256
// LDS: XT_TPB * XT_LD doubles at S; blockDim.x == XT_TPB.
__device__ inline void extadd_tile(const int4 t, const int4* __restrict__ tiles,
                                   const FrontInfo* __restrict__ fi, double* __restrict__ blkval,
                                   double* __restrict__ upd, const int* __restrict__ relidx,
                                   int zero22, double* S) {
    const int k = t.x, c0 = t.y, r0 = t.z;
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int c1 = min(c0 + 16, nj), r1 = min(r0 + XT_TPB, nj);
    const int tid = threadIdx.x;
    double* blk = blkval + f.blkptr;
    double* U = upd + f.updptr;
    // the tile's current values, one row per thread (coalesced columns)
    const int tr = r0 + tid;
    {
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int tc = c0 + q;
            const bool in = tr < r1 && tc < c1 && tc <= tr;
            const bool mine = tc < nn;  // (uniform: tc does not depend on the thread)
            v[q] = 0.0;
            if (in && (mine || !zero22))
                v[q] = mine ? blk[tr + (i64)tc * nj] : U[(tr - nn) + (i64)(tc - nn) * na];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) S[tid * XT_LD + q] = v[q];
    }
    // The child index ranges landing in the tile (first child index with
    // r >= c0, c1, r0, r1) and the child's na, relptr and updptr, found on
    // the host when the plan was built: three int4 per child in reverse
    // child order, at tiles + t.w.
    const int4* bnd = tiles + t.w;
    const int nch = f.chend - f.chbeg;
    __syncthreads();
    for (int cb = 0; cb < nch; ++cb) {
        const int4 bc = bnd[3 * cb], g1 = bnd[3 * cb + 1], g2 = bnd[3 * cb + 2];
        const int N = g1.x;
        const int* r = relidx + (i64)(((unsigned long long)(unsigned)g1.w << 32) | (unsigned)g1.z);
        const double* Uc = upd + (i64)(((unsigned long long)(unsigned)g2.y << 32) | (unsigned)g2.x);
        const int ja = bc.x, jb = bc.y, i = bc.z + tid;
        if (i < bc.w) {
            // at most 16 child columns land in the 16 target columns
            // (relative indices increase): the loads first, then the adds
            const int ri = r[i] - r0;
            const int je = min(jb, i + 1);  // lower part: column j <= row i
            double u[16];
            int tc[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int j = ja + q;
                tc[q] = -1;
                u[q] = 0.0;
                if (j < je) {
                    tc[q] = r[j] - c0;
                    u[q] = Uc[i + (i64)j * N];
                }
            }
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (tc[q] >= 0) S[ri * XT_LD + tc[q]] += u[q];
        }
        __syncthreads();  // the next child may add into the same entries
    }
    {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int tc = c0 + q;
            if (tr < r1 && tc < c1 && tc <= tr) {
                const double v = S[tid * XT_LD + q];
                if (tc < nn) blk[tr + (i64)tc * nj] = v;
                else U[(tr - nn) + (i64)(tc - nn) * na] = v;
            }
        }
    }
}
