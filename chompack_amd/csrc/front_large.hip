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
// Diagonal-block factorization for panel [c0, c0+nb) of each listed front:
// unblocked right-looking Cholesky of the nb x nb block in LDS (one thread
// per row), then its triangular inverse (one row per thread, no barriers),
// stored for the panel solve and the stored-convention solve.
__device__ inline void lds_lower_inverse(double (*D)[NB + 1], double (*E)[NB + 1], int nb, int t) {
    if (t < nb) {
        E[t][t] = 1.0 / D[t][t];
        for (int j = t - 1; j >= 0; --j) {
            double s = 0.0;
            for (int m = j + 1; m <= t; ++m) s += E[t][m] * D[m][j];
            E[t][j] = -s / D[j][j];
        }
    }
}

__global__ __launch_bounds__(64) void k_potrf_diag(
    const int* __restrict__ fronts, int c0, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ dinv,
    const i64* __restrict__ dinvptr, unsigned long long* status) {
    __shared__ double D[NB][NB + 1];
    __shared__ double E[NB][NB + 1];
    const int k = fronts[blockIdx.x];
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int t = threadIdx.x;
    double* blk = blkval + f.blkptr + c0 + (i64)c0 * nj;
    for (int j = 0; j < nb; ++j)
        if (t < nb && t >= j) D[t][j] = blk[t + (i64)j * nj];
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
        const double d = D[c][c];
        if (!(d > 0.0)) {
            if (t == 0) chp_record_failure(status, k, c0 + c + 1);
            return;
        }
        const double s = sqrt(d);
        const double rs = 1.0 / s;
        __syncthreads();
        if (t == c) D[c][c] = s;
        if (t > c && t < nb) D[t][c] *= rs;
        __syncthreads();
        if (t > c && t < nb) {
            const double lt = D[t][c];
            for (int j = c + 1; j <= t; ++j) D[t][j] -= lt * D[j][c];
        }
        __syncthreads();
    }
    lds_lower_inverse(D, E, nb, t);
    __syncthreads();
    double* inv = dinv + dinvptr[k] + (i64)(c0 / NB) * NB * NB;
    for (int j = 0; j < NB; ++j) {
        if (j < nb && t < nb && t >= j) blk[t + (i64)j * nj] = D[t][j];
        inv[t + j * NB] = (j < nb && t < nb && t >= j) ? E[t][j] : 0.0;
    }
}

// Inverse of every diagonal block of L11 (projected inverse, dtrtri's
// diagonal part).  blocks[i] = (front, block index).
__global__ __launch_bounds__(64) void k_trtri_diag(
    const int2* __restrict__ blocks, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, double* __restrict__ dinv,
    const i64* __restrict__ dinvptr, double* __restrict__ X,
    const i64* __restrict__ xptr, unsigned long long* status) {
    __shared__ double D[NB][NB + 1];
    __shared__ double E[NB][NB + 1];
    const int2 b = blocks[blockIdx.x];
    const int k = b.x, c0 = b.y * NB;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int t = threadIdx.x;
    const double* blk = blkval + f.blkptr + c0 + (i64)c0 * nj;
    for (int j = 0; j < nb; ++j)
        if (t < nb && t >= j) D[t][j] = blk[t + (i64)j * nj];
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (D[j][j] == 0.0) {
            if (t == 0) chp_record_failure(status, k, c0 + j + 1);
            return;
        }
    }
    lds_lower_inverse(D, E, nb, t);
    __syncthreads();
    double* inv = dinv + dinvptr[k] + (i64)b.y * NB * NB;
    for (int j = 0; j < NB; ++j)
        inv[t + j * NB] = (t < nb && j < nb && t >= j) ? E[t][j] : 0.0;
    // the diagonal block of X = L11^{-1} (ld nn)
    double* x = X + xptr[k] + c0 + (i64)c0 * nn;
    for (int j = 0; j < nb; ++j)
        if (t < nb) x[t + (i64)j * nn] = (t >= j) ? E[t][j] : 0.0;
}

// In-place right multiplication of a 64-row tile of a panel by a diagonal
// block inverse: tiles[i] = (front, row0, c0, mode).
//   mode 0: X = F[rows, c0:c0+nb] * Dinv^T   (L21 = F21 L_dd^-T, panel solve)
//   mode 1: X = F[rows, c0:c0+nb] * Dinv     (B21 = L21 L11^-1, block step)
__global__ __launch_bounds__(TPB) void k_panel_mul(
    const int4* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, const double* __restrict__ dinv,
    const i64* __restrict__ dinvptr) {
    __shared__ double Fs[64][NB + 1];
    __shared__ double Ms[NB][NB + 1];
    const int4 t = tiles[blockIdx.x];
    const int k = t.x, row0 = t.y, c0 = t.z, mode = t.w;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int rows = min(64, nj - row0);
    const int tid = threadIdx.x;
    double* blk = blkval + f.blkptr + row0 + (i64)c0 * nj;
    const double* inv = dinv + dinvptr[k] + (i64)(c0 / NB) * NB * NB;
    for (int e = tid; e < 64 * NB; e += TPB) {
        int i = e % 64, m = e / 64;
        Fs[i][m] = (i < rows && m < nb) ? blk[i + (i64)m * nj] : 0.0;
        int a = e % NB, b = e / NB;  // Ms[a][b]
        Ms[a][b] = mode == 0 ? inv[b + a * NB] : inv[a + b * NB];
    }
    __syncthreads();
    const int i = tid % 64;
    for (int j = tid / 64; j < nb; j += 4) {
        double s = 0.0;
        for (int m = 0; m < nb; ++m) s += Fs[i][m] * Ms[m][j];
        if (i < rows) blk[i + (i64)j * nj] = s;
    }
}

// ---------------------------------------------------------------------------
// Batched FP64 GEMM on MFMA (v_mfma_f64_16x16x4f64).  One workgroup computes
// one 64 x 64 tile of one task; 4 waves in a 2 x 2 arrangement, each wave a
// 32 x 32 sub-tile made of 2 x 2 MFMA 16 x 16 accumulators.  K is consumed in
// chunks of 16 staged through LDS.  Accumulator layout (16 x 16 f64): lane l,
// element v holds row (l / 16) + 4 v, column l % 16.
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ inline const double* bufptr(int b, double* b0, double* b1, double* b2, double* b3) {
    return b == 0 ? b0 : (b == 1 ? b1 : (b == 2 ? b2 : b3));
}

__global__ __launch_bounds__(TPB) void k_gemm_tiles(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks,
    double* b0, double* b1, double* b2, double* b3) {
    __shared__ double As[16][TILE + 2];
    __shared__ double Bs[16][TILE + 2];
    const GemmTile tl = tiles[blockIdx.x];
    const GemmTask T = tasks[tl.task];
    const double* A = bufptr(T.bufA, b0, b1, b2, b3) + T.offA;
    const double* B = bufptr(T.bufB, b0, b1, b2, b3) + T.offB;
    double* C = const_cast<double*>(bufptr(T.bufC, b0, b1, b2, b3)) + T.offC;
    const int row0 = tl.tm * TILE, col0 = tl.tn * TILE;
    const int M = T.M, N = T.N, K = T.K;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    const int triB = T.pad;

    if (T.tri == 2 && row0 + TILE - 1 < col0) {
        for (int e = tid; e < TILE * TILE; e += TPB) {
            int i = row0 + e % TILE, j = col0 + e / TILE;
            if (i < M && j < N) C[i + (i64)j * T.ldc] = 0.0;
        }
        return;
    }
    dbl4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = (dbl4){0.0, 0.0, 0.0, 0.0};

    for (int kb = 0; kb < K; kb += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + q * TPB;
            int i, kk;
            if (T.opA == 1) { kk = e & 15; i = e >> 4; }
            else { i = e & 63; kk = e >> 6; }
            const int gi = row0 + i, gk = kb + kk;
            double v = 0.0;
            bool ok = gi < M && gk < K;
            if (T.triA == 1) ok = ok && gk <= gi;
            else if (T.triA == 2) ok = ok && gk >= gi;
            if (ok) {
                if (T.opA == 0) v = A[gi + (i64)gk * T.lda];
                else if (T.opA == 1) v = A[gk + (i64)gi * T.lda];
                else v = (gi >= gk) ? A[gi + (i64)gk * T.lda] : A[gk + (i64)gi * T.lda];
            }
            As[kk][i] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + q * TPB;
            int j, kk;
            if (T.opB == 0) { kk = e & 15; j = e >> 4; }
            else { j = e & 63; kk = e >> 6; }
            const int gj = col0 + j, gk = kb + kk;
            double v = 0.0;
            bool ok = gj < N && gk < K;
            if (triB == 1) ok = ok && gk >= gj;
            if (ok) v = (T.opB == 0) ? B[gk + (i64)gj * T.ldb] : B[gj + (i64)gk * T.ldb];
            Bs[kk][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int kr = ks * 4 + (lane >> 4);
            double a0 = As[kr][wr * 32 + (lane & 15)];
            double a1 = As[kr][wr * 32 + 16 + (lane & 15)];
            double bb0 = Bs[kr][wc * 32 + (lane & 15)];
            double bb1 = Bs[kr][wc * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bb0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bb1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bb0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bb1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const int i = row0 + wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int j = col0 + wc * 32 + b * 16 + (lane & 15);
                if (i >= M || j >= N) continue;
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
