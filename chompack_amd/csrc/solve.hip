// Multi-RHS triangular solve with the chordal factor (trsm.c:32-136) and
// small utility kernels (trace inner product, frontal add/extract).
//
// Each front k owns a column-major nj x nrhs slab of the work buffer
// (offset wptr[k] * nrhs, ld nj).  In the forward sweep a parent reads its
// children's update vectors straight out of the children's slabs (rows
// nn_c..nj_c); in the backward sweep a child reads its separator values out
// of its parent's slab.  One workgroup per front, one launch per level.
#include "common.h"
#include "kernels.h"

#define TPB 256

// B := alpha L^{-1} B, leaves to root (trsm.c:36-81).
__global__ __launch_bounds__(TPB) void k_solve_fwd(
    const int* __restrict__ fronts, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, const i64* __restrict__ snptr,
    const int* __restrict__ relidx, const int* __restrict__ chidx,
    const i64* __restrict__ perm, double* __restrict__ B, int ldb, int nrhs,
    double alpha, double* __restrict__ work, const i64* __restrict__ wptr) {
    const int k = fronts[blockIdx.x];
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int tid = threadIdx.x;
    const double* L = blkval + f.blkptr;
    double* W = work + wptr[k] * nrhs;
    const i64* pk = perm + snptr[k];  // original rows of this supernode's columns
    for (int e = tid; e < nj * nrhs; e += TPB) {
        int i = e % nj, j = e / nj;
        W[i + (i64)j * nj] = (i < nn) ? alpha * B[pk[i] + (i64)j * ldb] : 0.0;
    }
    __syncthreads();
    for (int l = f.chend - 1; l >= f.chbeg; --l) {
        const int c = chidx[l];
        const FrontInfo g = fi[c];
        const int N = g.na, njc = g.nn + g.na;
        const int* r = relidx + g.relptr;
        const double* Wc = work + wptr[c] * nrhs + g.nn;
        for (int e = tid; e < N * nrhs; e += TPB) {
            int i = e % N, j = e / N;
            W[r[i] + (i64)j * nj] += Wc[i + (i64)j * njc];
        }
        __syncthreads();
    }
    // x_A -= B21 x_N  (uses x_N before the solve: B21 = L21 L11^-1)
    for (int e = tid; e < na * nrhs; e += TPB) {
        int i = e % na, j = e / na;
        double s = W[nn + i + (i64)j * nj];
        for (int m = 0; m < nn; ++m) s -= L[nn + i + (i64)m * nj] * W[m + (i64)j * nj];
        W[nn + i + (i64)j * nj] = s;
    }
    // x_N := L11^{-1} x_N, column oriented
    for (int c = 0; c < nn; ++c) {
        __syncthreads();
        const double d = L[c + (i64)c * nj];
        for (int j = tid; j < nrhs; j += TPB) W[c + (i64)j * nj] /= d;
        __syncthreads();
        const int m = nn - c - 1;
        for (int e = tid; e < m * nrhs; e += TPB) {
            int i = c + 1 + e % m, j = e / m;
            W[i + (i64)j * nj] -= L[i + (i64)c * nj] * W[c + (i64)j * nj];
        }
    }
    __syncthreads();
    for (int e = tid; e < nn * nrhs; e += TPB) {
        int i = e % nn, j = e / nn;
        B[pk[i] + (i64)j * ldb] = W[i + (i64)j * nj];
    }
}

// B := alpha L^{-T} B, root to leaves (trsm.c:84-135).
__global__ __launch_bounds__(TPB) void k_solve_bwd(
    const int* __restrict__ fronts, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, const i64* __restrict__ snptr,
    const int* __restrict__ relidx, const i64* __restrict__ perm,
    double* __restrict__ B, int ldb, int nrhs, double alpha,
    double* __restrict__ work, const i64* __restrict__ wptr) {
    const int k = fronts[blockIdx.x];
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int tid = threadIdx.x;
    const double* L = blkval + f.blkptr;
    double* W = work + wptr[k] * nrhs;
    const i64* pk = perm + snptr[k];
    for (int e = tid; e < nn * nrhs; e += TPB) {
        int i = e % nn, j = e / nn;
        W[i + (i64)j * nj] = alpha * B[pk[i] + (i64)j * ldb];
    }
    // x_N := L11^{-T} x_N (backward substitution, axpy form)
    for (int c = nn - 1; c >= 0; --c) {
        __syncthreads();
        const double d = L[c + (i64)c * nj];
        for (int j = tid; j < nrhs; j += TPB) W[c + (i64)j * nj] /= d;
        __syncthreads();
        for (int e = tid; e < c * nrhs; e += TPB) {
            int i = e % c, j = e / c;
            W[i + (i64)j * nj] -= L[c + (i64)i * nj] * W[c + (i64)j * nj];
        }
    }
    if (na > 0) {
        const FrontInfo p = fi[f.parent];
        const int njp = p.nn + p.na;
        const double* Wp = work + wptr[f.parent] * nrhs;
        const int* r = relidx + f.relptr;
        for (int e = tid; e < na * nrhs; e += TPB) {
            int i = e % na, j = e / na;
            W[nn + i + (i64)j * nj] = Wp[r[i] + (i64)j * njp];
        }
        __syncthreads();
        // x_N -= B21^T x_A: one wave per output (m, j), lanes over the
        // separator rows (coalesced column reads of B21 and x_A)
        const int lane = tid & 63, wv = tid >> 6;
        for (int o = wv; o < nn * nrhs; o += TPB / 64) {
            const int m = o % nn, j = o / nn;
            const double* lc = L + nn + (i64)m * nj;
            const double* xa = W + nn + (i64)j * nj;
            double acc = 0.0;
            for (int i = lane; i < na; i += 64) acc += lc[i] * xa[i];
#pragma unroll
            for (int sh = 32; sh > 0; sh >>= 1) acc += __shfl_xor(acc, sh, 64);
            if (lane == 0) W[m + (i64)j * nj] -= acc;
        }
    }
    __syncthreads();
    for (int e = tid; e < nn * nrhs; e += TPB) {
        int i = e % nn, j = e / nn;
        B[pk[i] + (i64)j * ldb] = W[i + (i64)j * nj];
    }
}

// Large-front solve steps (the dense products run as GEMM tasks between
// them).  Each workgroup owns rows [r0, r0 + 256) of one front's slab,
// tiles[i] = (front, r0).
// Forward gather: W[:nn] = alpha B[perm], W[nn:] = 0, then the children's
// update rows added in reverse child order (trsm.c:42-60); a child's rows that
// fall in the tile are found by binary search (relative indices increase).
__global__ __launch_bounds__(TPB) void k_sgather_fwd(
    const int2* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const i64* __restrict__ snptr, const int* __restrict__ relidx,
    const int* __restrict__ chidx, const i64* __restrict__ perm, const double* __restrict__ B,
    int ldb, int nrhs, double alpha, double* __restrict__ work, const i64* __restrict__ wptr) {
    const int2 t = tiles[blockIdx.x];
    const int k = t.x, r0 = t.y;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int r1 = min(r0 + 256, nj), rows = r1 - r0;
    const int tid = threadIdx.x;
    double* W = work + wptr[k] * nrhs;
    const i64* pk = perm + snptr[k];
    for (int e = tid; e < rows * nrhs; e += TPB) {
        const int i = r0 + e % rows, j = e / rows;
        W[i + (i64)j * nj] = (i < nn) ? alpha * B[pk[i] + (i64)j * ldb] : 0.0;
    }
    __shared__ int s_ja, s_jb;
    for (int l = f.chend - 1; l >= f.chbeg; --l) {
        const int c = chidx[l];
        const FrontInfo g = fi[c];
        const int N = g.na, njc = g.nn + g.na;
        const int* r = relidx + g.relptr;
        __syncthreads();
        if (tid == 0) {
            int lo = 0, hi = N;
            while (lo < hi) { int m = (lo + hi) >> 1; if (r[m] < r0) lo = m + 1; else hi = m; }
            s_ja = lo;
            hi = N;
            while (lo < hi) { int m = (lo + hi) >> 1; if (r[m] < r1) lo = m + 1; else hi = m; }
            s_jb = lo;
        }
        __syncthreads();
        const int ja = s_ja, cnt = s_jb - s_ja;
        const double* Wc = work + wptr[c] * nrhs + g.nn;
        for (int e = tid; e < cnt * nrhs; e += TPB) {
            const int i = ja + e % cnt, j = e / cnt;
            W[r[i] + (i64)j * nj] += Wc[i + (i64)j * njc];
        }
    }
}

// Backward gather: W[:nn] = alpha B[perm], W[nn:] = the parent's values of
// the separator rows (trsm.c:97-115).
__global__ __launch_bounds__(TPB) void k_sgather_bwd(
    const int2* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const i64* __restrict__ snptr, const int* __restrict__ relidx,
    const i64* __restrict__ perm, const double* __restrict__ B, int ldb, int nrhs,
    double alpha, double* __restrict__ work, const i64* __restrict__ wptr) {
    const int2 t = tiles[blockIdx.x];
    const int k = t.x, r0 = t.y;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int r1 = min(r0 + 256, nj), rows = r1 - r0;
    double* W = work + wptr[k] * nrhs;
    const i64* pk = perm + snptr[k];
    const FrontInfo p = fi[f.parent];
    const int njp = p.nn + p.na;
    const double* Wp = work + wptr[f.parent] * nrhs;
    const int* r = relidx + f.relptr;
    for (int e = threadIdx.x; e < rows * nrhs; e += TPB) {
        const int i = r0 + e % rows, j = e / rows;
        W[i + (i64)j * nj] = (i < nn) ? alpha * B[pk[i] + (i64)j * ldb]
                                      : Wp[r[i - nn] + (i64)j * njp];
    }
}

// B[perm] = T[:nn] for the listed tiles (rows r0 .. r0 + 255 below nn), T the
// front's solved block (column-major, ld nn, at tmp + t.tofs); with copyback
// also W[:nn] = T (the backward sweep's children read it from the slab).
__global__ __launch_bounds__(TPB) void k_sscatter(
    const STile* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const i64* __restrict__ snptr, const i64* __restrict__ perm, double* __restrict__ B,
    int ldb, int nrhs, double* __restrict__ work, const i64* __restrict__ wptr,
    const double* __restrict__ tmp, int copyback) {
    const STile t = tiles[blockIdx.x];
    const int k = t.k, r0 = t.r0;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int r1 = min(r0 + 256, nn), rows = r1 - r0;
    double* W = work + wptr[k] * nrhs;
    const double* T = tmp + t.tofs;
    const i64* pk = perm + snptr[k];
    for (int e = threadIdx.x; e < rows * nrhs; e += TPB) {
        const int i = r0 + e % rows, j = e / rows;
        const double v = T[i + (i64)j * nn];
        B[pk[i] + (i64)j * ldb] = v;
        if (copyback) W[i + (i64)j * nj] = v;
    }
}

// Partial trace inner products, one per supernode column j of the listed
// range: 2 * sum_{i >= j} X[i,j] Y[i,j] - X[j,j] Y[j,j] (dot.c:9-20).
// colfront[j] = supernode of column j; partial[j] receives the value.
__global__ __launch_bounds__(TPB) void k_dot_partial(
    const int* __restrict__ cols, int ncols, const FrontInfo* __restrict__ fi,
    const int* __restrict__ colfront, const double* __restrict__ x,
    const double* __restrict__ y, double* __restrict__ partial) {
    __shared__ double red[TPB];
    const int j = blockIdx.x;
    if (j >= ncols) return;
    const int k = colfront[j];
    const FrontInfo f = fi[k];
    const int nj = f.nn + f.na;
    const int jl = cols[j];  // column index inside the supernode
    const i64 off = f.blkptr + jl + (i64)jl * nj;
    double s = 0.0;
    for (int i = threadIdx.x; i < nj - jl; i += TPB) s += x[off + i] * y[off + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[j] = 2.0 * red[0] - x[off] * y[off];
}

// partial[j] = log L(j, j) for every column j of a factor (the diagonal of
// the stored factor is the diagonal of L, symbolic.py:1113-1128).
__global__ __launch_bounds__(TPB) void k_logdiag(const int* __restrict__ cols, int ncols,
                                                 const FrontInfo* __restrict__ fi,
                                                 const int* __restrict__ colfront,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ partial) {
    const int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= ncols) return;
    const FrontInfo f = fi[colfront[j]];
    const int jl = cols[j];
    partial[j] = log(x[f.blkptr + jl + (i64)jl * (f.nn + f.na)]);
}

// R = B - A X for a CSR matrix A (n x n, full symmetric pattern) and
// column-major n x nrhs X, B, R; one thread per (row, column), rows fastest.
__global__ __launch_bounds__(TPB) void k_spmv_residual(
    i64 n, const i64* __restrict__ rowptr, const int* __restrict__ colind,
    const double* __restrict__ vals, const double* __restrict__ X, int ldx,
    const double* __restrict__ B, int ldb, double* __restrict__ R, int ldr, int nrhs) {
    const i64 total = n * (i64)nrhs;
    for (i64 e = (i64)blockIdx.x * TPB + threadIdx.x; e < total; e += (i64)gridDim.x * TPB) {
        const i64 i = e % n;
        const int j = (int)(e / n);
        const double* x = X + (i64)j * ldx;
        double s0 = 0.0, s1 = 0.0;
        i64 p = rowptr[i];
        const i64 p1 = rowptr[i + 1];
        for (; p + 1 < p1; p += 2) {
            s0 += vals[p] * x[colind[p]];
            s1 += vals[p + 1] * x[colind[p + 1]];
        }
        if (p < p1) s0 += vals[p] * x[colind[p]];
        R[i + (i64)j * ldr] = B[i + (i64)j * ldb] - (s0 + s1);
    }
}

// dst[0:n] = src[0:n], 16-byte vector accesses, grid-stride.
__global__ __launch_bounds__(TPB) void k_copy(double* __restrict__ dst,
                                              const double* __restrict__ src, i64 n) {
    const i64 n2 = n / 2;
    const i64 stride = (i64)gridDim.x * TPB;
    const double2* s2 = reinterpret_cast<const double2*>(src);
    double2* d2 = reinterpret_cast<double2*>(dst);
    for (i64 e = (i64)blockIdx.x * TPB + threadIdx.x; e < n2; e += stride) d2[e] = s2[e];
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) dst[n - 1] = src[n - 1];
}

// F[r[i], r[j]] += alpha U[i, j] (i >= j), general strides (cbase.c:11-34).
__global__ void k_frontal_add(double* F, i64 N, const i64* r, const double* U,
                              i64 fr, i64 fc, i64 ur, i64 uc, double alpha) {
    const i64 j = blockIdx.x;
    for (i64 i = j + threadIdx.x; i < N; i += blockDim.x)
        F[r[i] * fr + r[j] * fc] += alpha * U[i * ur + j * uc];
}

// U[i, j] = F[r[i], r[j]] (i >= j) (cbase.c:39-59).
__global__ void k_frontal_get(const double* F, i64 N, const i64* r, double* U,
                              i64 fr, i64 fc, i64 ur, i64 uc) {
    const i64 j = blockIdx.x;
    for (i64 i = j + threadIdx.x; i < N; i += blockDim.x)
        U[i * ur + j * uc] = F[r[i] * fr + r[j] * fc];
}
