// Large-front kernels.  Fronts too big for one workgroup's LDS are processed
// by many workgroups at once through per-level work lists: a column-tiled
// extend-add, a blocked right-looking factorization (64-wide panels), and a
// batched FP64 MFMA GEMM that carries the Schur-complement updates and the
// projected-inverse products.  Work lists are built once per plan on the host.
#include "common.h"
#include "kernels.h"

#define TPB 256

// ---------------------------------------------------------------------------
// Deterministic extend-add for large fronts.  One workgroup owns 16 target
// columns [c0, c0+16) of front k: it zeroes the F22 part of those columns,
// then adds every child's update in reverse child order (the summation order
// of cholesky.c:41-52), so each target entry is updated by exactly one
// workgroup, in the reference's order, without atomics.
// Target (row, col) with col < nn lives in the blkval block (ld nj); with
// col >= nn in k's update slot (ld na).
__global__ __launch_bounds__(TPB) void k_extadd_tiles(
    const int2* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ upd,
    const int* __restrict__ relidx, const int* __restrict__ chidx) {
    const int2 t = tiles[blockIdx.x];
    const int k = t.x, c0 = t.y;
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int c1 = min(c0 + 16, nj);
    const int tid = threadIdx.x;
    double* blk = blkval + f.blkptr;
    double* U = upd + f.updptr;
    for (int tc = max(c0, nn); tc < c1; ++tc)
        for (int tr = tc + tid; tr < nj; tr += TPB)
            U[(tr - nn) + (i64)(tc - nn) * na] = 0.0;
    __syncthreads();
    __shared__ int s_ja, s_jb;
    for (int l = f.chend - 1; l >= f.chbeg; --l) {
        const FrontInfo g = fi[chidx[l]];
        const int N = g.na;
        const int* r = relidx + g.relptr;
        const double* Uc = upd + g.updptr;
        if (tid == 0) {
            int lo = 0, hi = N;
            while (lo < hi) { int m = (lo + hi) >> 1; if (r[m] < c0) lo = m + 1; else hi = m; }
            s_ja = lo;
            hi = N;
            while (lo < hi) { int m = (lo + hi) >> 1; if (r[m] < c1) lo = m + 1; else hi = m; }
            s_jb = lo;
        }
        __syncthreads();
        const int ja = s_ja, jb = s_jb;
        for (int j = ja; j < jb; ++j) {
            const int tc = r[j];
            const double* col = Uc + (i64)j * N;
            if (tc < nn) {
                double* dst = blk + (i64)tc * nj;
                for (int i = j + tid; i < N; i += TPB) dst[r[i]] += col[i];
            } else {
                double* dst = U + (i64)(tc - nn) * na - nn;
                for (int i = j + tid; i < N; i += TPB) dst[r[i]] += col[i];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Diagonal blocks.  One 256-thread workgroup per 64 x 64 diagonal block,
// padded with the identity beyond nb.  Thread (w = tid/64, l = tid%64) owns
// the 16 elements (w + 4q, l), q = 0..15, in registers.  Both the Cholesky
// factorization and the inverse of the factor are right-looking with two
// barriers per column; the column (or row) being eliminated is broadcast
// through LDS.  Inverses are stored row-major: Dinv(r, j) at dinv[j + r*NB].

#define DQ 16  // elements per thread

// L (lower, padded identity) in registers -> factor in place.  Returns false
// (uniformly) if a pivot is not positive; *bad receives its column.
__device__ inline bool block_potrf(double (&a)[DQ], double* colc, int w, int l, int* bad) {
    __shared__ double sd;
    __shared__ int sfail;
    if (w == 0 && l == 0) sfail = -1;
#pragma unroll 1
    for (int c = 0; c < NB; ++c) {
        const int qc = c >> 2, wc = c & 3;
        // owner of (c, c) publishes the pivot
        if (w == wc && l == c) {
            double d = 0.0;
#pragma unroll
            for (int q = 0; q < DQ; ++q) if (q == qc) d = a[q];
            if (!(d > 0.0)) sfail = c;
            sd = sqrt(d);
        }
        __syncthreads();
        if (sfail >= 0) { *bad = sfail; return false; }
        const double s = sd, rs = 1.0 / s;
        // column c: L(i, c) = a(i, c) / s for i > c, L(c, c) = s
        if (l == c) {
#pragma unroll
            for (int q = 0; q < DQ; ++q) {
                const int i = w + 4 * q;
                if (i > c) { a[q] *= rs; colc[i] = a[q]; }
                else if (i == c) { a[q] = s; colc[i] = s; }
            }
        }
        __syncthreads();
        // trailing update a(i, j) -= L(i, c) L(j, c) for i >= j > c
        if (l > c) {
            const double lj = colc[l];
#pragma unroll
            for (int q = 0; q < DQ; ++q) {
                const int i = w + 4 * q;
                if (i >= l) a[q] -= colc[i] * lj;
            }
        }
        __syncthreads();
    }
    return true;
}

// x = L^{-1} for the lower triangular L held (full, padded) in LDS Ls.
__device__ inline void block_trtri(const double (*Ls)[NB + 1], double (&x)[DQ], double* rowc,
                                   int w, int l) {
#pragma unroll
    for (int q = 0; q < DQ; ++q) x[q] = (w + 4 * q == l) ? 1.0 : 0.0;
#pragma unroll 1
    for (int c = 0; c < NB; ++c) {
        const int qc = c >> 2, wc = c & 3;
        // row c of X is final once divided by L(c, c)
        if (w == wc) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < DQ; ++q) if (q == qc) { x[q] = x[q] / Ls[c][c]; v = x[q]; }
            rowc[l] = v;
        }
        __syncthreads();
        // X(i, j) -= L(i, c) X(c, j) for i > c, j <= c
        if (l <= c) {
            const double xc = rowc[l];
#pragma unroll
            for (int q = 0; q < DQ; ++q) {
                const int i = w + 4 * q;
                if (i > c) x[q] -= Ls[i][c] * xc;
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(TPB) void k_potrf_diag(
    const int* __restrict__ fronts, int c0, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ dinv,
    const i64* __restrict__ dinvptr, unsigned long long* status) {
    __shared__ double Ls[NB][NB + 1];
    __shared__ double vec[NB];
    const int k = fronts[blockIdx.x];
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double* blk = blkval + f.blkptr + c0 + (i64)c0 * nj;
    double a[DQ];
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int i = w + 4 * q;
        a[q] = (i < nb && l < nb && i >= l) ? blk[i + (i64)l * nj] : (i == l ? 1.0 : 0.0);
    }
    int bad = 0;
    if (!block_potrf(a, vec, w, l, &bad)) {
        if (threadIdx.x == 0) chp_record_failure(status, k, c0 + bad + 1);
        return;
    }
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int i = w + 4 * q;
        Ls[i][l] = (i >= l) ? a[q] : 0.0;
        if (i < nb && l < nb && i >= l) blk[i + (i64)l * nj] = a[q];
    }
    __syncthreads();
    double x[DQ];
    block_trtri(Ls, x, vec, w, l);
    double* inv = dinv + dinvptr[k] + (i64)(c0 / NB) * NB * NB;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int i = w + 4 * q;
        inv[l + i * NB] = (i < nb && l < nb) ? x[q] : 0.0;
    }
}

// Inverse of every diagonal block of L11 (projected inverse, the diagonal
// part of dtrtri), also written into the diagonal blocks of X = L11^{-1}
// (column-major, ld nn).  blocks[i] = (front, block index).
__global__ __launch_bounds__(TPB) void k_trtri_diag(
    const int2* __restrict__ blocks, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, double* __restrict__ dinv,
    const i64* __restrict__ dinvptr, double* __restrict__ X,
    const i64* __restrict__ xptr, unsigned long long* status) {
    __shared__ double Ls[NB][NB + 1];
    __shared__ double vec[NB];
    const int2 b = blocks[blockIdx.x];
    const int k = b.x, c0 = b.y * NB;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const double* blk = blkval + f.blkptr + c0 + (i64)c0 * nj;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int i = w + 4 * q;
        Ls[i][l] = (i < nb && l < nb && i >= l) ? blk[i + (i64)l * nj] : (i == l ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (Ls[j][j] == 0.0) {
            if (threadIdx.x == 0) chp_record_failure(status, k, c0 + j + 1);
            return;
        }
    }
    double x[DQ];
    block_trtri(Ls, x, vec, w, l);
    double* inv = dinv + dinvptr[k] + (i64)b.y * NB * NB;
    double* xd = X + xptr[k] + c0 + (i64)c0 * nn;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int i = w + 4 * q;
        inv[l + i * NB] = (i < nb && l < nb) ? x[q] : 0.0;
        if (i < nb && l < nb) xd[i + (i64)l * nn] = x[q];
    }
}

// ---------------------------------------------------------------------------
// Batched FP64 GEMM on MFMA (v_mfma_f64_16x16x4f64).  One workgroup computes
// one 64 x 64 tile of one task; 4 waves in a 2 x 2 arrangement, each wave a
// 32 x 32 sub-tile made of 2 x 2 MFMA 16 x 16 accumulators.  K is consumed in
// chunks of 32 staged through LDS, the next chunk prefetched into registers
// while the current one is multiplied.  Accumulator layout (16 x 16 f64): lane l,
// element v holds row (l / 16) + 4 v, column l % 16.
typedef double dbl4 __attribute__((ext_vector_type(4)));

#define BK 32

__device__ inline const double* bufptr(int b, double* b0, double* b1, double* b2, double* b3) {
    return b == 0 ? b0 : (b == 1 ? b1 : (b == 2 ? b2 : b3));
}

// Element (gi, gk) of opA and (gk, gj) of opB with the task's masks; 0 outside.
__device__ inline double ldA(const GemmTask& T, const double* A, int gi, int gk) {
    bool ok = gi < T.M && gk < T.K;
    if (T.triA == 1) ok = ok && gk <= gi;
    else if (T.triA == 2) ok = ok && gk >= gi;
    if (!ok) return 0.0;
    if (T.opA == 0) return A[gi + (i64)gk * T.lda];
    if (T.opA == 1) return A[gk + (i64)gi * T.lda];
    return (gi >= gk) ? A[gi + (i64)gk * T.lda] : A[gk + (i64)gi * T.lda];
}
__device__ inline double ldB(const GemmTask& T, const double* B, int gk, int gj) {
    bool ok = gj < T.N && gk < T.K;
    if (T.pad == 1) ok = ok && gk >= gj;
    if (!ok) return 0.0;
    return (T.opB == 0) ? B[gk + (i64)gj * T.ldb] : B[gj + (i64)gk * T.ldb];
}

__global__ __launch_bounds__(TPB) void k_gemm_tiles(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks,
    double* b0, double* b1, double* b2, double* b3) {
    __shared__ double As[BK][TILE + 2];
    __shared__ double Bs[BK][TILE + 2];
    const GemmTile tl = tiles[blockIdx.x];
    const GemmTask T = tasks[tl.task];
    const double* A = bufptr(T.bufA, b0, b1, b2, b3) + T.offA;
    const double* B = bufptr(T.bufB, b0, b1, b2, b3) + T.offB;
    double* C = const_cast<double*>(bufptr(T.bufC, b0, b1, b2, b3)) + T.offC;
    const int row0 = tl.tm * TILE, col0 = tl.tn * TILE;
    const int K = T.K;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;

    if (T.tri == 2 && row0 + TILE - 1 < col0) {
        for (int e = tid; e < TILE * TILE; e += TPB) {
            int i = row0 + e % TILE, j = col0 + e / TILE;
            if (i < T.M && j < T.N) C[i + (i64)j * T.ldc] = 0.0;
        }
        return;
    }
    // 8 elements of A and 8 of B per thread and chunk, mapped so that
    // consecutive lanes read consecutive addresses of the stored operand
    const bool a_kfast = (T.opA != 0);
    const bool b_kfast = (T.opB == 0);
    double ra[8], rb[8];
    auto fetch = [&](int kb) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * TPB;
            int i, kk;
            if (a_kfast) { kk = e & 31; i = e >> 5; } else { i = e & 63; kk = e >> 6; }
            ra[q] = ldA(T, A, row0 + i, kb + kk);
            int j, k2;
            if (b_kfast) { k2 = e & 31; j = e >> 5; } else { j = e & 63; k2 = e >> 6; }
            rb[q] = ldB(T, B, kb + k2, col0 + j);
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * TPB;
            if (a_kfast) As[e & 31][e >> 5] = ra[q]; else As[e >> 6][e & 63] = ra[q];
            if (b_kfast) Bs[e & 31][e >> 5] = rb[q]; else Bs[e >> 6][e & 63] = rb[q];
        }
    };
    dbl4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = (dbl4){0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int kb = 0; kb < K; kb += BK) {
        stash();
        __syncthreads();
        if (kb + BK < K) fetch(kb + BK);
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + (lane >> 4);
            const double a0 = As[kr][wr * 32 + (lane & 15)];
            const double a1 = As[kr][wr * 32 + 16 + (lane & 15)];
            const double q0 = Bs[kr][wc * 32 + (lane & 15)];
            const double q1 = Bs[kr][wc * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, q0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, q1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, q0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, q1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const int i = row0 + wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int j = col0 + wc * 32 + b * 16 + (lane & 15);
                if (i >= T.M || j >= T.N) continue;
                double* c = C + i + (i64)j * T.ldc;
                if (T.tri == 0 || i >= j) {
                    double r = T.alpha * acc[a][b][v];
                    if (T.beta != 0.0) r += T.beta * (*c);
                    *c = r;
                } else if (T.tri == 2) {
                    *c = 0.0;
                }
            }
}

// blk[nn + i, j] = -S[i + j*na] for the listed (front, 16-column) tiles:
// F21 = -V_k B21 written back after F11 has consumed B21
// (projected_inverse.c:51-56).
__global__ __launch_bounds__(TPB) void k_neg_copy(
    const int2* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, const double* __restrict__ scratch,
    const i64* __restrict__ sptr) {
    const int2 t = tiles[blockIdx.x];
    const int k = t.x, c0 = t.y;
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    double* blk = blkval + f.blkptr + nn;
    const double* S = scratch + sptr[k];
    const int c1 = min(c0 + 16, nn);
    for (int j = c0; j < c1; ++j)
        for (int i = threadIdx.x; i < na; i += TPB) blk[i + (i64)j * nj] = -S[i + (i64)j * na];
}
