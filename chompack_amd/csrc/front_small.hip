// Small-front kernels: one workgroup owns one frontal matrix, held entirely
// in LDS, for fronts whose working set fits (the many small fronts near the
// leaves of the clique tree).  All arithmetic is FP64.
//
// Frontal layout in LDS: column-major, leading dimension nj, matching the
// block layout of blkval so block loads/stores are straight copies.
#include "common.h"
#include "kernels.h"

#define TPB 256

__global__ void k_status_init(unsigned long long* status) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        status[0] = CHP_STATUS_OK;
        status[1] = 0;
    }
}

// Extend-add of all children of front k (reverse child order, one child at a
// time) into the LDS front F (ld = nj).  F[r[i] + r[j]*nj] += U[i + j*N],
// i >= j.  Follows cholesky.c:41-52 / cbase.c:11-34.
__device__ inline void lds_extend_add(double* F, int nj, const FrontInfo& f,
                                      const FrontInfo* __restrict__ fi,
                                      const int* __restrict__ chidx,
                                      const int* __restrict__ relidx,
                                      const double* __restrict__ upd) {
    for (int l = f.chend - 1; l >= f.chbeg; --l) {
        const FrontInfo g = fi[chidx[l]];
        const int N = g.na;
        const int* r = relidx + g.relptr;
        const double* U = upd + g.updptr;
        // one pass over the child's N x N array (lower part used): a thread's
        // loads are independent of each other, so they are all in flight at
        // once instead of one global round trip per column
        for (int e = threadIdx.x; e < N * N; e += TPB) {
            const int i = e % N, j = e / N;
            if (i >= j) F[r[i] + r[j] * nj] += U[e];
        }
        __syncthreads();
    }
}

// Supernodal Cholesky of one small front (cholesky.c:30-81):
// assemble, factor F11, L21 = F21 L11^-T, F22 -= L21 L21^T,
// B21 = L21 L11^-1 (stored convention), write block and update matrix.
__device__ inline void chol_small_front(
    const int k, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ upd,
    const int* __restrict__ relidx, const int* __restrict__ chidx,
    unsigned long long* status) {
    extern __shared__ double F[];
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int tid = threadIdx.x;
    double* blk = blkval + f.blkptr;

    for (int e = tid; e < nj * nj; e += TPB) {
        int i = e % nj, j = e / nj;
        F[e] = (j < nn && i >= j) ? blk[e] : 0.0;
    }
    __syncthreads();
    lds_extend_add(F, nj, f, fi, chidx, relidx, upd);

    // right-looking unblocked factorization of the first nn columns,
    // updating the trailing part (including F22) as it goes
    for (int c = 0; c < nn; ++c) {
        const double d = F[c + c * nj];
        if (!(d > 0.0)) {
            if (tid == 0) chp_record_failure(status, k, c + 1);
            return;
        }
        const double s = sqrt(d);
        const double rs = 1.0 / s;
        __syncthreads();
        if (tid == 0) F[c + c * nj] = s;
        for (int i = c + 1 + tid; i < nj; i += TPB) F[i + c * nj] *= rs;
        __syncthreads();
        const int m = nj - c - 1;
        for (int e = tid; e < m * m; e += TPB) {
            int i = c + 1 + e % m, j = c + 1 + e / m;
            if (i >= j) F[i + j * nj] -= F[i + c * nj] * F[j + c * nj];
        }
        __syncthreads();
    }
    // B21 := L21 L11^{-1}, one row per thread (backward substitution)
    for (int i = nn + tid; i < nj; i += TPB) {
        for (int j = nn - 1; j >= 0; --j) {
            double x = F[i + j * nj];
            for (int m = j + 1; m < nn; ++m) x -= F[i + m * nj] * F[m + j * nj];
            F[i + j * nj] = x / F[j + j * nj];
        }
    }
    __syncthreads();
    for (int e = tid; e < nj * nn; e += TPB) {
        int i = e % nj, j = e / nj;
        if (i >= j) blk[e] = F[e];
    }
    if (na > 0) {
        double* U = upd + f.updptr;
        for (int e = tid; e < na * na; e += TPB) {
            int i = e % na, j = e / na;
            if (i >= j) U[e] = F[(nn + i) + (nn + j) * nj];
        }
    }
}

// A persistent launch over the level's small fronts (largest first): each
// workgroup takes fronts blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ __launch_bounds__(TPB) void k_chol_small(
    const int* __restrict__ fronts, int count, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ upd,
    const int* __restrict__ relidx, const int* __restrict__ chidx,
    unsigned long long* status) {
    for (int q = blockIdx.x; q < count; q += gridDim.x) {
        chol_small_front(fronts[q], fi, blkval, upd, relidx, chidx, status);
        __syncthreads();  // the front's LDS is free before the next one loads
    }
}


// Gather V_k = F_p[r, r] from the parent's finished front into k's update
// slot, both triangles (projected_inverse.c:60-70 / cbase.c:39-59, done by
// the child).  The parent's front is its blkval block (columns < nn_p) and
// its own V_p (columns >= nn_p, lower triangle read).  One workgroup per
// (front, 16-column tile, 256-row block); the block goes through LDS so that
// the mirrored upper-triangle writes are coalesced too.
__global__ __launch_bounds__(TPB) void k_extract(
    const int4* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ upd,
    const int* __restrict__ relidx, i64 bstride, i64 ustride,
    const double* __restrict__ scratch, const i64* __restrict__ tptr) {
    __shared__ double S[256][17];
    // grid y: the instance (a Hessian direction), as in k_extadd_tiles
    blkval += (i64)blockIdx.y * bstride;
    upd += (i64)blockIdx.y * ustride;
    const int4 t = tiles[blockIdx.x];
    const int k = t.x, j0 = t.y, ib = t.z;
    if (t.w == 2) {
        // F21 = -T for columns j0 .. j0 + 15 of front k (T at scratch +
        // tptr[k], ld na): the projected inverse's last step for the front,
        // in the launch that extracts its children's V from T directly
        const FrontInfo f = fi[k];
        const int nn = f.nn, na = f.na, nj = nn + na;
        double* blk = blkval + f.blkptr + nn;
        const double* T = scratch + tptr[k];
        const int nc = min(j0 + 16, nn) - j0;
        for (int c = 0; c < nc; c += 4)
            for (int i = threadIdx.x; i < na; i += TPB) {
                double v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = (c + u < nc) ? T[i + (i64)(j0 + c + u) * na] : 0.0;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (c + u < nc) blk[i + (i64)(j0 + c + u) * nj] = -v[u];
            }
        return;
    }
    const FrontInfo f = fi[k];
    const int na = f.na;
    const FrontInfo p = fi[f.parent];
    const int nnp = p.nn, njp = p.nn + p.na, nap = p.na;
    const int* r = relidx + f.relptr;
    const double* pblk = blkval + p.blkptr;
    const double* pV = upd + p.updptr;
    // t.w == 1: the parent's F21 is still -T in scratch (ld nap)
    const double* pT = t.w == 1 ? scratch + tptr[f.parent] : nullptr;
    double* V = upd + f.updptr;
    const int j1 = min(j0 + 16, na);
    const int tid = threadIdx.x;
    {
        const int i0 = max(ib, j0), i1 = min(ib + 256, na);
        const int i = i0 + tid;
        const int ri = (i < i1) ? r[i] : 0;
        // all 16 columns' loads issued before any store
        double v[16];
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int j = j0 + jj;
            v[jj] = 0.0;
            if (j < j1 && i < i1 && i >= j) {
                const int rj = r[j];
                v[jj] = (rj >= nnp) ? pV[(ri - nnp) + (i64)(rj - nnp) * nap]
                        : (pT && ri >= nnp) ? -pT[(ri - nnp) + (i64)rj * nap]
                                            : pblk[ri + (i64)rj * njp];
            }
        }
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int j = j0 + jj;
            if (j < j1 && i < i1 && i >= j) {
                V[i + (i64)j * na] = v[jj];
                S[tid][jj] = v[jj];
            }
        }
        __syncthreads();
        // mirrored part V[j, i] = V[i, j] for i > j: 16 consecutive j per row i
        for (int e = tid; e < (i1 - i0) * 16; e += TPB) {
            const int jj = e & 15, ii = e >> 4;
            const int i = i0 + ii, j = j0 + jj;
            if (j < j1 && i > j) V[j + (i64)i * na] = S[ii][jj];
        }
    }
}

// Projected inverse of one small front (projected_inverse.c:29-72), with
// V_k already gathered into k's update slot.
// LDS: Li (nn x nn), C (nn x nn), B (na x nn), T (na x nn), tmp (nn),
// V (na x na, both triangles: the extraction stores both).
__device__ inline void projinv_small_front(
    const int k, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, const double* __restrict__ upd,
    unsigned long long* status) {
    extern __shared__ double S[];
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int tid = threadIdx.x;
    double* blk = blkval + f.blkptr;
    double* Li = S;
    double* C = Li + nn * nn;
    double* B = C + nn * nn;
    double* T = B + na * nn;
    double* tmp = T + na * nn;
    double* Vs = tmp + nn;

    for (int e = tid; e < nn * nn; e += TPB) {
        int i = e % nn, j = e / nn;
        Li[e] = (i >= j) ? blk[i + j * nj] : 0.0;
    }
    if (na > 0) {  // one flat copy: every thread's loads in flight at once
        const double* V = upd + f.updptr;
        for (int e = tid; e < na * na; e += TPB) Vs[e] = V[e];
    }
    for (int e = tid; e < na * nn; e += TPB) {
        int i = e % na, j = e / na;
        B[e] = blk[nn + i + j * nj];
    }
    __syncthreads();
    // dtrtri: first zero on the diagonal is an error
    for (int j = 0; j < nn; ++j) {
        if (Li[j + j * nn] == 0.0) {
            if (tid == 0) chp_record_failure(status, k, j + 1);
            return;
        }
    }
    for (int j = nn - 1; j >= 0; --j) {
        if (tid == 0) Li[j + j * nn] = 1.0 / Li[j + j * nn];
        __syncthreads();
        const double ajj = -Li[j + j * nn];
        for (int i = j + 1 + tid; i < nn; i += TPB) {
            double y = 0.0;
            for (int m = j + 1; m <= i; ++m) y += Li[i + m * nn] * Li[m + j * nn];
            tmp[i] = y;
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < nn; i += TPB) Li[i + j * nn] = tmp[i] * ajj;
        __syncthreads();
    }
    // C = Li^T Li (lower)
    for (int e = tid; e < nn * nn; e += TPB) {
        int i = e % nn, j = e / nn;
        if (i < j) continue;
        double s = 0.0;
        for (int m = i; m < nn; ++m) s += Li[m + i * nn] * Li[m + j * nn];
        C[e] = s;
    }
    if (na > 0) {
        // T = V B21 (V in LDS, symmetric with both triangles stored)
        for (int e = tid; e < na * nn; e += TPB) {
            int i = e % na, j = e / na;
            double s0 = 0.0, s1 = 0.0;
            int m = 0;
            for (; m + 1 < na; m += 2) {
                s0 += Vs[i + m * na] * B[m + j * na];
                s1 += Vs[i + (m + 1) * na] * B[m + 1 + j * na];
            }
            if (m < na) s0 += Vs[i + m * na] * B[m + j * na];
            T[e] = s0 + s1;
        }
        __syncthreads();
        // C += T^T B21
        for (int e = tid; e < nn * nn; e += TPB) {
            int i = e % nn, j = e / nn;
            if (i < j) continue;
            double s = 0.0;
            for (int m = 0; m < na; ++m) s += T[m + i * na] * B[m + j * na];
            C[e] += s;
        }
    }
    __syncthreads();
    for (int e = tid; e < nn * nn; e += TPB) {
        int i = e % nn, j = e / nn;
        blk[i + j * nj] = (i >= j) ? C[e] : 0.0;
    }
    for (int e = tid; e < na * nn; e += TPB) {
        int i = e % na, j = e / na;
        blk[nn + i + j * nj] = -T[e];
    }
}

// Persistent over the level's small fronts, as k_chol_small.
__global__ __launch_bounds__(TPB) void k_projinv_small(
    const int* __restrict__ fronts, int count, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, const double* __restrict__ upd,
    unsigned long long* status) {
    for (int q = blockIdx.x; q < count; q += gridDim.x) {
        projinv_small_front(fronts[q], fi, blkval, upd, status);
        __syncthreads();
    }
}


// Chordal product X = L L^T for one small front (llt.c:30-57).
// LDS: A (nj x nn, the block as stored) and F (nj x nj).
__device__ inline void llt_small_front(
    const int k, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ upd,
    const int* __restrict__ relidx, const int* __restrict__ chidx) {
    extern __shared__ double S[];
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int tid = threadIdx.x;
    double* blk = blkval + f.blkptr;
    double* A = S;
    double* F = A + nj * nn;
    for (int e = tid; e < nj * nn; e += TPB) A[e] = blk[e];
    __syncthreads();
    // B21 := B21 * L11 (L11 lower triangular), one row per thread
    for (int i = nn + tid; i < nj; i += TPB) {
        for (int j = 0; j < nn; ++j) {
            double s = 0.0;
            for (int m = j; m < nn; ++m) s += A[i + m * nj] * A[m + j * nj];
            A[i + j * nj] = s;
        }
    }
    __syncthreads();
    // F = A A^T (lower)
    for (int e = tid; e < nj * nj; e += TPB) {
        int i = e % nj, j = e / nj;
        double s = 0.0;
        if (i >= j)
            for (int m = 0; m < nn; ++m) s += A[i + m * nj] * A[j + m * nj];
        F[e] = s;
    }
    __syncthreads();
    lds_extend_add(F, nj, f, fi, chidx, relidx, upd);
    for (int e = tid; e < nj * nn; e += TPB) {
        int i = e % nj, j = e / nj;
        if (i >= j) blk[e] = F[e];
    }
    if (na > 0) {
        double* U = upd + f.updptr;
        for (int e = tid; e < na * na; e += TPB) {
            int i = e % na, j = e / na;
            if (i >= j) U[e] = F[(nn + i) + (nn + j) * nj];
        }
    }
}

// Persistent over the level's small fronts, as k_chol_small.
__global__ __launch_bounds__(TPB) void k_llt_small(
    const int* __restrict__ fronts, int count, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ upd,
    const int* __restrict__ relidx, const int* __restrict__ chidx) {
    for (int q = blockIdx.x; q < count; q += gridDim.x) {
        llt_small_front(fronts[q], fi, blkval, upd, relidx, chidx);
        __syncthreads();
    }
}

