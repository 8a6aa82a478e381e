// Chains of small fronts: runs of consecutive elimination-tree levels that
// hold one front each (a path of the tree, as banded and arrow patterns
// have), walked by one workgroup of CHAIN_TPB threads per chain.  Two consecutive
// fronts of a chain are child and parent, and the matrix one of them hands
// to the other -- the child's update matrix in cholesky (cholesky.c:41-52,
// 74-78) and llt (llt.c:36-50), the child's V_c extracted from the parent in
// the projected inverse (projected_inverse.c:60-70) -- stays in LDS instead
// of a round trip through the update arena.  A front's descriptor
// (ChainFront) carries what the step needs, and the next front's columns
// and relative indices are loaded while the current front is processed, so
// the walk does not wait on global memory between fronts.
//
// Fronts here have order nj <= 64.  A front is held in LDS as a 64 x 65
// row-major array (row i, column j at i * CLD + j); lane l of every wave
// owns row l and wave w the columns w, w + CW, w + 2 CW, ...  The per-front
// steps are chains of dependent LDS and integer operations, so several waves
// per SIMD (CW = 16 waves) keep the SIMDs busy where one wave would wait out
// each latency.
#include "common.h"
#include "kernels.h"

#define CT CHAIN_TPB
#define CW (CT / 64)
#define CLD 65
#define CQ (64 / CW)  // columns per thread
#define NARROW 4      // fronts with nn below this split their dot products over the waves

namespace {

__device__ inline int rdl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Workgroup barrier that orders LDS only.  Nothing a chain kernel stores to
// global memory is read back by the same launch, and a full __syncthreads()
// would also wait for the next front's loads in flight.
__device__ inline void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A ChainFront read with vector loads, lane i holding its 32-bit word i
// (scalar loads would share the LDS wait counter), and unpacked by readlane.
constexpr int CF_WORDS = (int)(sizeof(ChainFront) / 4);
static_assert(sizeof(ChainFront) == 56, "ChainFront layout");

__device__ inline int desc_load(const ChainFront* __restrict__ cf, int q, int l) {
    return reinterpret_cast<const int*>(cf + q)[min(l, CF_WORDS - 1)];
}

__device__ inline i64 rdl64(int v, int lane) {
    return (i64)(((unsigned long long)(unsigned)rdl(v, lane + 1) << 32) | (unsigned)rdl(v, lane));
}

struct Desc {
    i64 blkptr, updptr, nbr_relptr;
    int front, nn, na, chbeg, chend, nbr, nbr_na, keep;
};

__device__ inline Desc desc_get(int v) {
    Desc d;
    d.blkptr = rdl64(v, 0);
    d.updptr = rdl64(v, 2);
    d.nbr_relptr = rdl64(v, 4);
    d.front = rdl(v, 6);
    d.nn = rdl(v, 7);
    d.na = rdl(v, 8);
    d.chbeg = rdl(v, 9);
    d.chend = rdl(v, 10);
    d.nbr = rdl(v, 11);
    d.nbr_na = rdl(v, 12);
    d.keep = rdl(v, 13);
    return d;
}

__device__ inline double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Row l of the front's first nn columns (lower part), the CQ columns of
// wave w, loaded for the next front while the current one is processed.
// Entries outside the lower part read a valid address and are masked out
// where they are used (bit m of the returned mask), so nothing here waits
// for the loads.
__device__ inline unsigned load_cols(const double* __restrict__ blk, int nn, int nj, int w, int l,
                                     double (&v)[CQ]) {
    unsigned okm = 0;
#pragma unroll
    for (int m = 0; m < CQ; ++m) {
        const int j = w + CW * m;
        v[m] = 0.0;
        if (j < nn) {  // (uniform in the wave)
            const bool ok = l < nj && l >= j;
            v[m] = blk[ok ? l + (i64)j * nj : 0];
            okm |= (unsigned)ok << m;
        }
    }
    return okm;
}

// Wait for this wave's loads in flight (the next front's data) before the
// front's global stores are issued: the vector memory counter also counts
// stores, so a wait for the prefetched data placed after them would also
// wait for the stores to complete.
__device__ inline void wait_prefetch() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

__device__ inline int load_rel(const int* __restrict__ relidx, i64 relptr, int N, int l) {
    return (N > 0 && l < N) ? relidx[relptr + l] : 0;
}

// Extend-add of the children of a front into F (cholesky.c:41-52,
// llt.c:36-46), in reverse child order, each followed by a barrier.  The
// chain child `nbr` takes its update matrix from G, at rows and columns
// gnn.. (its own columns come first), with its relative indices already in
// rl_nbr (lane l: r[l]); the others from their update slots.
__device__ inline void chain_extend_add(int nbr, int nbr_na, int rl_nbr, int chbeg, int chend,
                                        double* F, const double* G, int gnn,
                                        const FrontInfo* __restrict__ fi,
                                        const double* __restrict__ upd,
                                        const int* __restrict__ relidx,
                                        const int* __restrict__ chidx, int w, int l) {
    for (int ch = chend - 1; ch >= chbeg; --ch) {
        // a lone chain child needs no index loads
        const int cid = (nbr >= 0 && chend - chbeg == 1) ? nbr : chidx[ch];
        if (cid == nbr) {
            // every read issued before the first write (the targets are
            // distinct, so the adds do not depend on each other)
            const int N = nbr_na;
            double g[CQ], f[CQ];
            int ix[CQ];
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                const bool ok = j < N && l < N && l >= j;
                ix[m] = ok ? rl_nbr * CLD + rdl(rl_nbr, j) : -1;
                g[m] = G[ok ? (gnn + l) * CLD + gnn + j : 0];
                f[m] = F[ok ? ix[m] : 0];
            }
#pragma unroll
            for (int m = 0; m < CQ; ++m)
                if (ix[m] >= 0) F[ix[m]] = f[m] + g[m];
        } else {
            const FrontInfo g = fi[cid];
            const int N = g.na;
            const int rl = load_rel(relidx, g.relptr, N, l);
            const double* Ug = upd + g.updptr;
            double u[CQ];
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                const bool ok = j < N && l < N && l >= j;
                u[m] = Ug[ok ? l + (i64)j * N : 0];
            }
            double f[CQ];
            int ix[CQ];
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                const bool ok = j < N && l < N && l >= j;
                ix[m] = ok ? rl * CLD + rdl(rl, j) : -1;
                f[m] = F[ok ? ix[m] : 0];
            }
#pragma unroll
            for (int m = 0; m < CQ; ++m)
                if (ix[m] >= 0) F[ix[m]] = f[m] + u[m];
        }
        lds_sync();  // the next child may add into the same entries
    }
}

// The front's update matrix F22 (rows and columns nn..) into its slot
// (lower part, ld na).
__device__ inline void store_update(const double* F, int nn, int na, double* U, int w, int l) {
#pragma unroll
    for (int m = 0; m < CQ; ++m) {
        const int j = w + CW * m;
        if (j < na && l < na && l >= j) U[l + (i64)j * na] = F[(nn + l) * CLD + nn + j];
    }
}

}  // namespace

// Cholesky along a chain (cholesky.c:26-82 per front): assemble the front's
// columns and a zero F22, extend-add the children, factor the nn columns
// (right-looking, F22 updated as it goes), B21 = L21 L11^{-1}, store the
// block; F22 is the update matrix, kept in LDS when the parent is the next
// front of the chain.  Column p's step reads the column unscaled and scales
// it in registers, so the step takes one barrier; the scaled column is
// written after it (the next step reads only column p + 1 and right).
__global__ __launch_bounds__(CT) void k_chol_chain(const ChainFront* __restrict__ cf, int count,
                                                   const FrontInfo* __restrict__ fi,
                                                   double* __restrict__ blkval,
                                                   double* __restrict__ upd,
                                                   const int* __restrict__ relidx,
                                                   const int* __restrict__ chidx,
                                                   unsigned long long* status) {
    __shared__ double S0[64 * CLD], S1[64 * CLD];
    double* F = S0;  // the current front
    double* G = S1;  // the previous front (its update matrix at rows / columns gnn..)
    int gnn = 0;
    const int l = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Desc c = desc_get(desc_load(cf, 0, l));
    int nxw = desc_load(cf, count > 1 ? 1 : 0, l);
    double pv[CQ];
    unsigned pok = load_cols(blkval + c.blkptr, c.nn, c.nn + c.na, w, l, pv);
    int prl = load_rel(relidx, c.nbr_relptr, c.nbr >= 0 ? c.nbr_na : 0, l);
    wait_prefetch();  // (so that the loop head has no loads of its own to wait for)
    for (int q = 0; q < count; ++q) {
        const i64 blkptr = c.blkptr, updptr = c.updptr;
        const int nn = c.nn, na = c.na, nj = nn + na, front = c.front;
        const int nbr = c.nbr, nbr_na = c.nbr_na, chbeg = c.chbeg, chend = c.chend;
        const int keep = c.keep, rl = prl;
#pragma unroll
        for (int m = 0; m < CQ; ++m) F[l * CLD + w + CW * m] = ((pok >> m) & 1) ? pv[m] : 0.0;
        if (q + 1 < count) {  // the next front's data in flight during this one
            c = desc_get(nxw);
            pok = load_cols(blkval + c.blkptr, c.nn, c.nn + c.na, w, l, pv);
            prl = load_rel(relidx, c.nbr_relptr, c.nbr >= 0 ? c.nbr_na : 0, l);
            if (q + 2 < count) nxw = desc_load(cf, q + 2, l);
        }
        lds_sync();
        chain_extend_add(nbr, nbr_na, rl, chbeg, chend, F, G, gnn, fi, upd, relidx, chidx, w, l);
        double s = 1.0, xl = 0.0;
        for (int p = 0; p < nn; ++p) {
            const double d = F[p * CLD + p];
            if (!(d > 0.0)) {
                if (threadIdx.x == 0) chp_record_failure(status, front, p + 1);
                return;
            }
            s = sqrt(d);
            const double rs = 1.0 / s;
            xl = (l > p && l < nj) ? F[l * CLD + p] * rs : 0.0;  // L(l, p)
            double a[CQ], b[CQ];
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                a[m] = F[l * CLD + j];
                b[m] = F[j * CLD + p];
            }
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                if (j > p && j < nj && l >= j && l < nj) F[l * CLD + j] = a[m] - xl * (b[m] * rs);
            }
            lds_sync();
            if (w == (p & (CW - 1)) && l >= p && l < nj) F[l * CLD + p] = (l == p) ? s : xl;
        }
        double* blk = blkval + blkptr;
        wait_prefetch();
        if (nn == 1) {
            // B21 = L21 / L11 from wave 0's registers
            if (w == 0 && l < nj) blk[l] = (l == 0) ? s : xl / s;
        } else {
            lds_sync();  // the last column written
            // B21 := L21 L11^{-1}, row by row (backward substitution)
            if (w == 0 && l >= nn && l < nj) {
                for (int jj = nn - 1; jj >= 0; --jj) {
                    double x = F[l * CLD + jj];
                    for (int m2 = jj + 1; m2 < nn; ++m2) x -= F[l * CLD + m2] * F[m2 * CLD + jj];
                    F[l * CLD + jj] = x / F[jj * CLD + jj];
                }
            }
            lds_sync();
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                if (j < nn && l < nj && l >= j) blk[l + (i64)j * nj] = F[l * CLD + j];
            }
        }
        if (keep) {
            // the next front fills the other buffer; this one's F22 is read
            // after the next front's first barrier
            double* x = F;
            F = G;
            G = x;
            gnn = nn;
        } else {
            if (na > 0) store_update(F, nn, na, upd + updptr, w, l);
            lds_sync();  // F is read before the next front overwrites it
        }
    }
}

// llt along a chain (llt.c:26-57 per front): B21 := B21 L11, F = A A^T with
// A = [L11; B21 L11] (A in its own buffer, so F is written directly),
// extend-add the children, store the block; F22 is the update matrix, kept
// in LDS when the parent is the next front.
__global__ __launch_bounds__(CT) void k_llt_chain(const ChainFront* __restrict__ cf, int count,
                                                  const FrontInfo* __restrict__ fi,
                                                  double* __restrict__ blkval,
                                                  double* __restrict__ upd,
                                                  const int* __restrict__ relidx,
                                                  const int* __restrict__ chidx) {
    __shared__ double S0[64 * CLD], S1[64 * CLD], SA[64 * CLD];
    double* F = S0;
    double* G = S1;
    double* A = SA;
    int gnn = 0;
    const int l = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Desc c = desc_get(desc_load(cf, 0, l));
    int nxw = desc_load(cf, count > 1 ? 1 : 0, l);
    double pv[CQ];
    unsigned pok = load_cols(blkval + c.blkptr, c.nn, c.nn + c.na, w, l, pv);
    int prl = load_rel(relidx, c.nbr_relptr, c.nbr >= 0 ? c.nbr_na : 0, l);
    wait_prefetch();  // (so that the loop head has no loads of its own to wait for)
    for (int q = 0; q < count; ++q) {
        const i64 blkptr = c.blkptr, updptr = c.updptr;
        const int nn = c.nn, na = c.na, nj = nn + na;
        const int nbr = c.nbr, nbr_na = c.nbr_na, chbeg = c.chbeg, chend = c.chend;
        const int keep = c.keep, rl = prl;
#pragma unroll
        for (int m = 0; m < CQ; ++m) A[l * CLD + w + CW * m] = ((pok >> m) & 1) ? pv[m] : 0.0;
        if (q + 1 < count) {
            c = desc_get(nxw);
            pok = load_cols(blkval + c.blkptr, c.nn, c.nn + c.na, w, l, pv);
            prl = load_rel(relidx, c.nbr_relptr, c.nbr >= 0 ? c.nbr_na : 0, l);
            if (q + 2 < count) nxw = desc_load(cf, q + 2, l);
        }
        lds_sync();
        if (nn > 1) {
            // B21 := B21 L11 (dtrmm), row l: column j uses the row's entries
            // right of it, not yet overwritten
            if (w == 0 && l >= nn && l < nj) {
                for (int j = 0; j < nn; ++j) {
                    double s = 0.0;
                    for (int m2 = j; m2 < nn; ++m2) s += A[l * CLD + m2] * A[m2 * CLD + j];
                    A[l * CLD + j] = s;
                }
            }
            lds_sync();
        }
        // F = A A^T (lower).  With nn = 1, A(l, 0) = B21(l) L11 is formed
        // here (L11 itself for l = 0).
        {
            const double l00 = A[0];
            double v[CQ];
#pragma unroll
            for (int m = 0; m < CQ; ++m) v[m] = 0.0;
            for (int p = 0; p < nn; ++p) {
                double a = A[l * CLD + p];
                if (nn == 1 && l >= 1) a *= l00;
#pragma unroll
                for (int m = 0; m < CQ; ++m) {
                    const int j = w + CW * m;
                    double b = A[j * CLD + p];
                    if (nn == 1 && j >= 1) b *= l00;
                    v[m] += a * b;  // A(j, p) = 0 for p > j (L11 lower), and rows >= nj
                }
            }
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                if (l >= j) F[l * CLD + j] = v[m];
            }
        }
        lds_sync();
        chain_extend_add(nbr, nbr_na, rl, chbeg, chend, F, G, gnn, fi, upd, relidx, chidx, w, l);
        double* blk = blkval + blkptr;
        wait_prefetch();
#pragma unroll
        for (int m = 0; m < CQ; ++m) {
            const int j = w + CW * m;
            if (j < nn && l < nj && l >= j) blk[l + (i64)j * nj] = F[l * CLD + j];
        }
        if (keep) {
            double* x = F;
            F = G;
            G = x;
            gnn = nn;
        } else {
            if (na > 0) store_update(F, nn, na, upd + updptr, w, l);
            lds_sync();
        }
    }
}

// Projected inverse along a chain, parent first (projected_inverse.c:29-72
// per front).  The front's V_k is in LDS when its parent came just before it
// (placed there by the parent's extraction), else in its update slot (both
// triangles).  Steps: X = L11^{-1} (dtrtri, into G), S11 = X^T X,
// T = V_k B21, S11 += T^T B21, S21 = -T; store the block; then V_c of every
// child from the finished front: into G for the chain's next front, into the
// update slot (both triangles) for the others.  For fronts with fewer than
// NARROW columns the dot products over the na separator rows are split
// between the waves (T) or summed across the lanes (T^T B21).
__global__ __launch_bounds__(CT) void k_projinv_chain(const ChainFront* __restrict__ cf, int count,
                                                      const FrontInfo* __restrict__ fi,
                                                      double* __restrict__ blkval,
                                                      double* __restrict__ upd,
                                                      const int* __restrict__ relidx,
                                                      const int* __restrict__ chidx,
                                                      unsigned long long* status) {
    __shared__ double S0[64 * CLD], S1[64 * CLD];
    __shared__ double TP[CW][64][NARROW];    // partial sums of T (narrow fronts)
    __shared__ double PS[NARROW][NARROW];    // T^T B21 (narrow fronts)
    double* F = S0;  // the current front; V_k at rows / columns nn..
    double* G = S1;  // X = L11^{-1}, then the next front's V
    const int l = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Desc c = desc_get(desc_load(cf, 0, l));
    int nxw = desc_load(cf, count > 1 ? 1 : 0, l);
    double pv[CQ];
    unsigned pok = load_cols(blkval + c.blkptr, c.nn, c.nn + c.na, w, l, pv);
    int prl = load_rel(relidx, c.nbr_relptr, c.keep ? c.nbr_na : 0, l);
    wait_prefetch();
    for (int q = 0; q < count; ++q) {
        const i64 blkptr = c.blkptr, updptr = c.updptr;
        const int nn = c.nn, na = c.na, nj = nn + na, front = c.front;
        const int nbr = c.nbr, nbr_na = c.nbr_na, chbeg = c.chbeg, chend = c.chend;
        const int keep = c.keep, rl_next = prl;
        // the next front (its V goes to G at rows / columns nnx..)
        const int nnx = rdl(nxw, 7), next_front = rdl(nxw, 6);
        if (nbr < 0 && na > 0) {  // V_k from the update slot
            const double* V = upd + updptr;
            double u[CQ];
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                const bool ok = j < na && l < na;
                u[m] = V[ok ? l + (i64)j * na : 0];
            }
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                if (j < na && l < na) F[(nn + l) * CLD + nn + j] = u[m];
            }
        }
#pragma unroll
        for (int m = 0; m < CQ; ++m) {  // L11 (lower) and B21
            const int j = w + CW * m;
            if (j < nn) F[l * CLD + j] = ((pok >> m) & 1) ? pv[m] : 0.0;
        }
        if (q + 1 < count) {
            c = desc_get(nxw);
            pok = load_cols(blkval + c.blkptr, c.nn, c.nn + c.na, w, l, pv);
            prl = load_rel(relidx, c.nbr_relptr, c.keep ? c.nbr_na : 0, l);
            if (q + 2 < count) nxw = desc_load(cf, q + 2, l);
        }
        lds_sync();
        // dtrtri: the first zero on the diagonal is an error
        {
            const unsigned long long z = __ballot(l < nn && F[l * CLD + l] == 0.0);
            if (z) {
                if (threadIdx.x == 0) chp_record_failure(status, front, __builtin_ctzll(z) + 1);
                return;
            }
        }
        const bool narrow = nn < NARROW;
        // X = L11^{-1}, lane l forming column l (rows i > l from the rows
        // above them: X(i, l) = -(sum_{l <= m < i} L(i, m) X(m, l)) / L(i, i))
        if (w == 0 && l < nn) {
            G[l * CLD + l] = 1.0 / F[l * CLD + l];
            for (int i = l + 1; i < nn; ++i) {
                double s = 0.0;
                for (int m2 = l; m2 < i; ++m2) s += F[i * CLD + m2] * G[m2 * CLD + l];
                G[i * CLD + l] = -s / F[i * CLD + i];
            }
        }
        if (narrow && l < na) {
            // partial sums of T(l, j) = sum_m V(l, m) B21(m, j) over wave w's quarter of m
            const int q4 = (na + CW - 1) / CW, m0 = w * q4, m1 = min(m0 + q4, na);
            for (int j = 0; j < nn; ++j) {
                double s0 = 0.0, s1 = 0.0;
                int m2 = m0;
                for (; m2 + 1 < m1; m2 += 2) {
                    s0 += F[(nn + l) * CLD + nn + m2] * F[(nn + m2) * CLD + j];
                    s1 += F[(nn + l) * CLD + nn + m2 + 1] * F[(nn + m2 + 1) * CLD + j];
                }
                if (m2 < m1) s0 += F[(nn + l) * CLD + nn + m2] * F[(nn + m2) * CLD + j];
                TP[w][l][j] = s0 + s1;
            }
        }
        lds_sync();
        // S11 = X^T X (lower) and T = V_k B21, in registers
        double cs[CQ], ts[CQ];
#pragma unroll
        for (int m = 0; m < CQ; ++m) {
            const int j = w + CW * m;
            double s = 0.0;
            if (l < nn && j <= l)
                for (int m2 = l; m2 < nn; ++m2) s += G[m2 * CLD + l] * G[m2 * CLD + j];
            cs[m] = s;
            double t = 0.0;
            if (l < na && j < nn) {
                if (narrow) {
#pragma unroll
                    for (int u = 0; u < CW; ++u) t += TP[u][l][j];
                } else {
                    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
                    int m2 = 0;
                    for (; m2 + 3 < na; m2 += 4) {
                        t0 += F[(nn + l) * CLD + nn + m2] * F[(nn + m2) * CLD + j];
                        t1 += F[(nn + l) * CLD + nn + m2 + 1] * F[(nn + m2 + 1) * CLD + j];
                        t2 += F[(nn + l) * CLD + nn + m2 + 2] * F[(nn + m2 + 2) * CLD + j];
                        t3 += F[(nn + l) * CLD + nn + m2 + 3] * F[(nn + m2 + 3) * CLD + j];
                    }
                    for (; m2 < na; ++m2) t0 += F[(nn + l) * CLD + nn + m2] * F[(nn + m2) * CLD + j];
                    t = (t0 + t1) + (t2 + t3);
                }
            }
            ts[m] = t;
        }
        if (narrow) {
            // (T^T B21)(i, j) for the pairs i >= j, one pair per wave at a
            // time, the rows across the lanes
            int pr = 0;
            for (int i = 0; i < nn; ++i)
                for (int j = 0; j <= i; ++j, ++pr) {
                    if ((pr & (CW - 1)) != w) continue;
                    double t = 0.0;
                    if (l < na) {
#pragma unroll
                        for (int u = 0; u < CW; ++u) t += TP[u][l][i];
                    }
                    const double b = l < na ? F[(nn + l) * CLD + j] : 0.0;
                    const double s = wsum(t * b);
                    if (l == 0) PS[i][j] = s;
                }
        } else {
            // T beside V_k's rows in G (rows nn.. of G are free)
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                if (l < na && j < nn) G[(nn + l) * CLD + j] = ts[m];
            }
        }
        lds_sync();
#pragma unroll
        for (int m = 0; m < CQ; ++m) {
            const int j = w + CW * m;
            if (l < nn && j <= l && na > 0) {
                if (narrow) {
                    cs[m] += PS[l][j];
                } else {
                    double s = 0.0;
                    for (int m2 = 0; m2 < na; ++m2) s += G[(nn + m2) * CLD + l] * F[(nn + m2) * CLD + j];
                    cs[m] += s;
                }
            }
        }
        if (!narrow) lds_sync();  // B21 is read before the front's columns are replaced
#pragma unroll
        for (int m = 0; m < CQ; ++m) {
            const int j = w + CW * m;
            if (j < nn) {
                if (l < nn) F[l * CLD + j] = j <= l ? cs[m] : 0.0;
                if (l < na) F[(nn + l) * CLD + j] = -ts[m];  // row l of T is front row nn + l
            }
        }
        lds_sync();
        double* blk = blkval + blkptr;
        wait_prefetch();
#pragma unroll
        for (int m = 0; m < CQ; ++m) {
            const int j = w + CW * m;
            if (j < nn && l < nj) blk[l + (i64)j * nj] = F[l * CLD + j];
        }
        // V_c = F[r, r] of each child (lower triangle of F read)
        for (int ch = chbeg; ch < chend; ++ch) {
            const int cid = (keep && chend - chbeg == 1) ? next_front : chidx[ch];
            const bool inlds = keep && cid == next_front;
            int N, rl;
            double* Vg = upd;
            if (inlds) {
                N = nbr_na;
                rl = rl_next;
            } else {
                const FrontInfo g = fi[cid];
                N = g.na;
                rl = load_rel(relidx, g.relptr, N, l);
                Vg = upd + g.updptr;
            }
            double v[CQ];
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                const int rj = rdl(rl, j);
                v[m] = F[(j < N && l < N) ? max(rl, rj) * CLD + min(rl, rj) : 0];
            }
#pragma unroll
            for (int m = 0; m < CQ; ++m) {
                const int j = w + CW * m;
                if (j < N && l < N) {
                    if (inlds) G[(nnx + l) * CLD + nnx + j] = v[m];
                    else Vg[l + (i64)j * N] = v[m];
                }
            }
        }
        if (keep) {
            // the next front writes its columns into G (away from its V) and
            // reads only after its first barrier
            double* x = F;
            F = G;
            G = x;
        } else {
            lds_sync();
        }
    }
}
