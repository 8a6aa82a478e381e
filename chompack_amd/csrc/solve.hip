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

// Small fronts (nn <= 16) of the solve: one wave per (front, chunk of up to
// 16 right-hand sides), items[i] = (front, first rhs).  L11 and the chunk of
// x_N live in LDS; each of the first rc lanes solves its right-hand side's
// triangular system in registers.  Rows of the slab are spread over the
// lanes with all right-hand sides of a row in registers, so that a lane's
// loads (B, gather map, children's slabs, a row of B21) are independent.
#define SRC 16
typedef double sdbl4 __attribute__((ext_vector_type(4)));

// B := alpha L^{-1} B, leaves to root (trsm.c:36-81).  Row i of the front's
// slab is alpha B[p[i]] (i < nn) or 0, plus the children's update rows
// (gather map, reverse child order).  x_A -= B21 x_N uses x_N before its
// solve (B21 = L21 L11^{-1}); then x_N := L11^{-1} x_N.  The slab keeps x_A
// for the parent; x_N goes to B.
__global__ __launch_bounds__(64) void k_solve_fwd(
    const int2* __restrict__ items, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, const i64* __restrict__ snptr,
    const i64* __restrict__ perm, double* __restrict__ B, int ldb, int nrhs,
    double alpha, double* __restrict__ work, const i64* __restrict__ wptr,
    const int* __restrict__ gptr, const GEnt* __restrict__ gent) {
    __shared__ double Ls[16][17];
    __shared__ double xs[16][17];
    const int2 it = items[blockIdx.x];
    const int k = it.x, j0 = it.y;
    const int rc = min(SRC, nrhs - j0);
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int lane = threadIdx.x;
    const double* L = blkval + f.blkptr;
    const i64 wk = wptr[k];
    double* W = work + wk * nrhs;
    const i64* pk = perm + snptr[k];
#pragma unroll
    for (int e = lane; e < 256; e += 64) {
        const int i = e & 15, c = e >> 4;
        if (i < nn && c <= i) Ls[i][c] = L[i + (i64)c * nj];
    }
    if (lane < nn) {  // gathered right-hand side rows of x_N
        const int i = lane;
        const i64 bi = pk[i];
        double g[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) g[j] = (j < rc) ? alpha * B[bi + (i64)(j0 + j) * ldb] : 0.0;
        const int e1 = gptr[wk + i + 1];
        for (int q = gptr[wk + i]; q < e1; ++q) {
            const GEnt t = gent[q];
            const double* src = work + t.wu * nrhs + t.row + (i64)j0 * t.ld;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < rc) g[j] += src[(i64)j * t.ld];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) xs[i][j] = g[j];
    }
    __syncthreads();
    // x_A = gathered rows - B21 x_N, one row per lane
    for (int i = nn + lane; i < nj; i += 64) {
        double g[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) g[j] = 0.0;
        const int e1 = gptr[wk + i + 1];
        for (int q = gptr[wk + i]; q < e1; ++q) {
            const GEnt t = gent[q];
            const double* src = work + t.wu * nrhs + t.row + (i64)j0 * t.ld;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < rc) g[j] += src[(i64)j * t.ld];
        }
#pragma unroll 1
        for (int m = 0; m < nn; ++m) {
            const double lv = L[i + (i64)m * nj];
#pragma unroll
            for (int j = 0; j < 16; ++j) g[j] -= lv * xs[m][j];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < rc) W[i + (i64)(j0 + j) * nj] = g[j];
    }
    if (lane < rc) {
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = (i < nn) ? xs[i][lane] : 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            if (c < nn) {
                x[c] /= Ls[c][c];
#pragma unroll
                for (int i = c + 1; i < 16; ++i)
                    if (i < nn) x[i] -= Ls[i][c] * x[c];
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < nn) B[pk[i] + (i64)(j0 + lane) * ldb] = x[i];
    }
}

// B := alpha L^{-T} B, root to leaves (trsm.c:84-135).  With the stored
// B21 = L21 L11^{-1}: x_N = L11^{-T} (alpha B[p]) - B21^T x_A, x_A from the
// parent's slab (kept in this front's slab for the children); x_N goes to B
// and to the slab.  B21^T x_A runs on v_mfma_f64_16x16x4f64 over 16-row
// chunks of x_A and B21 staged in LDS (operands: lane l gives A(l % 16, k)
// and B(k, l % 16) with k = l / 16; accumulator element v of lane l is row
// l / 16 + 4 v, column l % 16).
__global__ __launch_bounds__(64) void k_solve_bwd(
    const int2* __restrict__ items, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, const i64* __restrict__ snptr,
    const int* __restrict__ relidx, const i64* __restrict__ perm,
    double* __restrict__ B, int ldb, int nrhs, double alpha,
    double* __restrict__ work, const i64* __restrict__ wptr) {
    __shared__ double Ls[16][17];
    __shared__ double xs[16][17];
    __shared__ double xa[16][17];
    __shared__ double lb[16][17];
    const int2 it = items[blockIdx.x];
    const int k = it.x, j0 = it.y;
    const int rc = min(SRC, nrhs - j0);
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int lane = threadIdx.x;
    const double* L = blkval + f.blkptr;
    double* W = work + wptr[k] * nrhs;
    const i64* pk = perm + snptr[k];
#pragma unroll
    for (int e = lane; e < 256; e += 64) {
        const int i = e & 15, c = e >> 4;
        if (i < nn && c <= i) Ls[i][c] = L[i + (i64)c * nj];
        if (i < nn && c < rc) xs[i][c] = alpha * B[pk[i] + (i64)(j0 + c) * ldb];
    }
    __syncthreads();
    // x_N := L11^{-T} x_N, one right-hand side per lane, in registers
    if (lane < rc) {
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = (i < nn) ? xs[i][lane] : 0.0;
#pragma unroll
        for (int c = 15; c >= 0; --c) {
            if (c < nn) {
                x[c] /= Ls[c][c];
#pragma unroll
                for (int i = 0; i < c; ++i) x[i] -= Ls[c][i] * x[c];
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < nn) xs[i][lane] = x[i];
    }
    const int col = lane & 15, kq = lane >> 4;
    if (na > 0) {
        const FrontInfo pf = fi[f.parent];
        const int njp = pf.nn + pf.na;
        const double* Wp = work + wptr[f.parent] * nrhs;
        const int* r = relidx + f.relptr;
        sdbl4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int i0 = 0; i0 < na; i0 += 16) {
            // stage rows i0 .. i0 + 15 of x_A (also kept in the slab) and B21
            const int ii = lane & 15;
            const int i = i0 + ii;
            const bool ok = i < na;
            const int ri = ok ? r[i] : 0;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int c = (lane >> 4) + 4 * h;  // column 0..15
                double v = 0.0, lv = 0.0;
                if (ok && c < rc) v = Wp[ri + (i64)(j0 + c) * njp];
                if (ok && c < nn) lv = L[nn + i + (i64)c * nj];
                xa[ii][c] = v;
                lb[ii][c] = lv;
                if (ok && c < rc) W[nn + i + (i64)(j0 + c) * nj] = v;
            }
            __syncthreads();
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int kk = 4 * s4 + kq;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(lb[kk][col], xa[kk][col], acc, 0, 0, 0);
            }
            __syncthreads();
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = kq + 4 * v;
            if (m < nn && col < rc) xs[m][col] -= acc[v];
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = lane; e < 256; e += 64) {
        const int i = e & 15, c = e >> 4;
        if (i < nn && c < rc) {
            const double v = xs[i][c];
            B[pk[i] + (i64)(j0 + c) * ldb] = v;
            W[i + (i64)(j0 + c) * nj] = v;
        }
    }
}

// *flag = 0 (a kernel rather than a memset, so that a captured graph replays it).
__global__ void k_flag_clear(int* flag) {
    if (threadIdx.x == 0) *flag = 0;
}

// Compares the lower triangle of L11 of the listed pieces (columns c0 ..
// c0 + 15, rows r0 .. r1 - 1) with the copy the solve's inverses were formed
// from (at tofs, ld nn); refreshes the copy and sets *changed where they
// differ.
__global__ __launch_bounds__(TPB) void k_l11_check(
    const CheckTile* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, double* __restrict__ copy, int* __restrict__ changed) {
    const CheckTile t = tiles[blockIdx.x];
    const FrontInfo f = fi[t.k];
    const int nn = f.nn, nj = nn + f.na;
    const int c0 = t.c0, ncol = min(c0 + 16, nn) - c0, r0 = t.r0, nr = t.r1 - t.r0;
    const int total = ncol * nr;  // rows r0 .. r1 - 1 of each column (those above the diagonal skipped)
    const double* src = blkval + f.blkptr + (i64)c0 * nj + r0;
    double* dst = copy + t.tofs + (i64)c0 * nn + r0;
    bool diff = false;
    for (int e0 = 0; e0 < total; e0 += 8 * TPB) {  // eight element pairs in flight per thread
        double v[8], o[8];
        int at[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * TPB + threadIdx.x;
            const int c = e / nr, i = e - c * nr;
            const bool ok = e < total && r0 + i >= c0 + c;
            const int cc = ok ? c : 0, ii = ok ? i : 0;
            at[u] = ok ? ii + cc * nn : -1;
            v[u] = src[ii + (i64)cc * nj];
            o[u] = dst[ii + (i64)cc * nn];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (at[u] < 0 || __double_as_longlong(v[u]) == __double_as_longlong(o[u])) continue;
            dst[at[u]] = v[u];
            diff = true;
        }
    }
    if (diff) *changed = 1;
}

// Large-front solve steps (the dense products run as GEMM tasks between
// them).  Each workgroup owns rows [r0, r0 + 64) of one front's slab,
// tiles[i] = (front, r0).
// Forward gather: row i of the slab = alpha B[p[i]] (i < nn) or 0, plus the
// children's update rows that land on it (gather map, reverse child order as
// in trsm.c:46-58); with bonly, only rows i < nn and only alpha B[p[i]] (the
// product with L adds the children's rows after the front's own product).
// tiles[i] = (front, r0): rows r0 .. r0 + 63, thread (row, g) handles
// right-hand sides j0 + g, j0 + g + 4, ... of the chunk j0 = blockIdx.y *
// SOLVE_RHS_CHUNK (many right-hand sides spread over workgroups).
__global__ __launch_bounds__(TPB) void k_sgather_fwd(
    const int2* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const i64* __restrict__ snptr, const i64* __restrict__ perm, const double* __restrict__ B,
    int ldb, int nrhs, double alpha, double* __restrict__ work, const i64* __restrict__ wptr,
    const int* __restrict__ gptr, const GEnt* __restrict__ gent, int bonly) {
    const int2 t = tiles[blockIdx.x];
    const int k = t.x;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int i = t.y + (int)(threadIdx.x & 63);
    if (i >= (bonly ? nn : nj)) return;
    const i64 wk = wptr[k];
    double* W = work + wk * nrhs;
    const int e0 = gptr[wk + i], e1 = bonly ? e0 : gptr[wk + i + 1];
    const i64 bi = (i < nn) ? perm[snptr[k] + i] : 0;
    const int j0 = blockIdx.y * SOLVE_RHS_CHUNK, j1 = min(nrhs, j0 + SOLVE_RHS_CHUNK);
    for (int j = j0 + (int)(threadIdx.x >> 6); j < j1; j += TPB / 64) {
        double g = (i < nn) ? alpha * B[bi + (i64)j * ldb] : 0.0;
        for (int e = e0; e < e1; ++e) {
            const GEnt q = gent[e];
            g += work[q.wu * nrhs + q.row + (i64)j * q.ld];
        }
        W[i + (i64)j * nj] = g;
    }
}

// Backward gather: W[:nn] = alpha B[perm], W[nn:] = the parent's values of
// the separator rows (trsm.c:97-115).  tiles[i] = (front, r0): rows r0 ..
// r0 + 63, thread (row, g) handles right-hand sides j0 + g, j0 + g + 4, ...
// of the chunk blockIdx.y.
__global__ __launch_bounds__(TPB) void k_sgather_bwd(
    const int2* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const i64* __restrict__ snptr, const int* __restrict__ relidx,
    const i64* __restrict__ perm, const double* __restrict__ B, int ldb, int nrhs,
    double alpha, double* __restrict__ work, const i64* __restrict__ wptr) {
    const int2 t = tiles[blockIdx.x];
    const int k = t.x;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int i = t.y + (int)(threadIdx.x & 63);
    if (i >= nj) return;
    double* W = work + wptr[k] * nrhs;
    const double* src;
    i64 ld;
    double scale;
    if (i < nn) {
        src = B + perm[snptr[k] + i];
        ld = ldb;
        scale = alpha;
    } else {
        const FrontInfo p = fi[f.parent];
        src = work + wptr[f.parent] * nrhs + relidx[f.relptr + i - nn];
        ld = p.nn + p.na;
        scale = 1.0;
    }
    const int j0 = blockIdx.y * SOLVE_RHS_CHUNK, j1 = min(nrhs, j0 + SOLVE_RHS_CHUNK);
    for (int j = j0 + (int)(threadIdx.x >> 6); j < j1; j += TPB / 64)
        W[i + (i64)j * nj] = scale * src[(i64)j * ld];
}

// B[perm] = T[:nn] for the listed tiles (rows r0 .. r0 + 63 below nn), T the
// front's solved block (column-major, ld nn, at tmp + t.tofs); with tmp2 (the
// backward sweep) the value is T - T2, T2 at tmp2 + t.tofs, for fronts with a
// separator; with copyback also W[:nn] = the value (the backward sweep's
// children read it from the slab).  Thread (row, g) handles right-hand sides
// j0 + g, j0 + g + 4, ... of the chunk blockIdx.y.
__global__ __launch_bounds__(TPB) void k_sscatter(
    const STile* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const i64* __restrict__ snptr, const i64* __restrict__ perm, double* __restrict__ B,
    int ldb, int nrhs, double* __restrict__ work, const i64* __restrict__ wptr,
    const double* __restrict__ tmp, const double* __restrict__ tmp2, int copyback) {
    const STile t = tiles[blockIdx.x];
    const int k = t.k;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int i = t.r0 + (int)(threadIdx.x & 63);
    if (i >= nn) return;
    double* W = work + wptr[k] * nrhs;
    const double* T = tmp + t.tofs;
    const double* T2 = (tmp2 && f.na > 0) ? tmp2 + t.tofs : nullptr;
    const i64 bi = perm[snptr[k] + i];
    const int j0 = blockIdx.y * SOLVE_RHS_CHUNK, j1 = min(nrhs, j0 + SOLVE_RHS_CHUNK);
    for (int j = j0 + (int)(threadIdx.x >> 6); j < j1; j += TPB / 64) {
        double v = T[i + (i64)j * nn];
        if (T2) v -= T2[i + (i64)j * nn];
        B[bi + (i64)j * ldb] = v;
        if (copyback) W[i + (i64)j * nj] = v;
    }
}

// ---------------------------------------------------------------------------
// Products with the chordal factor, B := alpha L B or alpha L^T B
// (pybase/trmm.py:44-114), on the solve's level schedule, slabs and gather
// map.  With the stored B21 = L21 L11^{-1}:
//   'N', leaves to root: y_N = L11 (alpha b_N), y_A = B21 y_N (= L21 alpha
//        b_N); then the children's update rows are added (reverse child
//        order); x_N goes to B, the slab keeps y_A + children's rows for the
//        parent.
//   'T', root to leaves: u = [alpha b_N; the parent's u at the separator
//        rows] stays in the slab for the children (u before any product);
//        x_N = L11^T (u_N + B21^T u_A) goes to B.

// Small fronts, 'N': one wave per (front, chunk of up to 16 right-hand sides).
__global__ __launch_bounds__(64) void k_trmm_fwd(
    const int2* __restrict__ items, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, const i64* __restrict__ snptr,
    const i64* __restrict__ perm, double* __restrict__ B, int ldb, int nrhs,
    double alpha, double* __restrict__ work, const i64* __restrict__ wptr,
    const int* __restrict__ gptr, const GEnt* __restrict__ gent) {
    __shared__ double Ls[16][17];
    __shared__ double ys[16][17];
    const int2 it = items[blockIdx.x];
    const int k = it.x, j0 = it.y;
    const int rc = min(SRC, nrhs - j0);
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int lane = threadIdx.x;
    const double* L = blkval + f.blkptr;
    const i64 wk = wptr[k];
    double* W = work + wk * nrhs;
    const i64* pk = perm + snptr[k];
#pragma unroll
    for (int e = lane; e < 256; e += 64) {
        const int i = e & 15, c = e >> 4;
        if (i < nn && c <= i) Ls[i][c] = L[i + (i64)c * nj];
    }
    __syncthreads();
    if (lane < rc) {  // y_N = L11 (alpha b_N), one right-hand side per lane
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = (i < nn) ? alpha * B[pk[i] + (i64)(j0 + lane) * ldb] : 0.0;
#pragma unroll
        for (int i = 15; i >= 0; --i) {
            if (i < nn) {
                double y = 0.0;
#pragma unroll
                for (int c = 0; c <= i; ++c) y += Ls[i][c] * x[c];
                x[i] = y;
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) ys[i][lane] = (i < nn) ? x[i] : 0.0;
    }
    __syncthreads();
    // every row: own product, then the children's rows in map order
    for (int i = lane; i < nj; i += 64) {
        double g[16];
        if (i < nn) {
#pragma unroll
            for (int j = 0; j < 16; ++j) g[j] = ys[i][j];
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) g[j] = 0.0;
#pragma unroll 1
            for (int m = 0; m < nn; ++m) {
                const double lv = L[i + (i64)m * nj];
#pragma unroll
                for (int j = 0; j < 16; ++j) g[j] += lv * ys[m][j];
            }
        }
        const int e1 = gptr[wk + i + 1];
        for (int q = gptr[wk + i]; q < e1; ++q) {
            const GEnt t = gent[q];
            const double* src = work + t.wu * nrhs + t.row + (i64)j0 * t.ld;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < rc) g[j] += src[(i64)j * t.ld];
        }
        if (i < nn) {
            const i64 bi = pk[i];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < rc) B[bi + (i64)(j0 + j) * ldb] = g[j];
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < rc) W[i + (i64)(j0 + j) * nj] = g[j];
        }
    }
}

// Small fronts, 'T'.  B21^T u_A on v_mfma_f64_16x16x4f64 over 16-row chunks
// staged in LDS (operand and accumulator layout as in k_solve_bwd).
__global__ __launch_bounds__(64) void k_trmm_bwd(
    const int2* __restrict__ items, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, const i64* __restrict__ snptr,
    const int* __restrict__ relidx, const i64* __restrict__ perm,
    double* __restrict__ B, int ldb, int nrhs, double alpha,
    double* __restrict__ work, const i64* __restrict__ wptr) {
    __shared__ double Ls[16][17];
    __shared__ double zs[16][17];
    __shared__ double ua[16][17];
    __shared__ double lb[16][17];
    const int2 it = items[blockIdx.x];
    const int k = it.x, j0 = it.y;
    const int rc = min(SRC, nrhs - j0);
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int lane = threadIdx.x;
    const double* L = blkval + f.blkptr;
    double* W = work + wptr[k] * nrhs;
    const i64* pk = perm + snptr[k];
#pragma unroll
    for (int e = lane; e < 256; e += 64) {
        const int i = e & 15, c = e >> 4;
        if (i < nn && c <= i) Ls[i][c] = L[i + (i64)c * nj];
        double u = 0.0;
        if (i < nn && c < rc) {
            u = alpha * B[pk[i] + (i64)(j0 + c) * ldb];
            W[i + (i64)(j0 + c) * nj] = u;  // the children's view of this front
        }
        zs[i][c] = u;
    }
    const int col = lane & 15, kq = lane >> 4;
    if (na > 0) {
        const FrontInfo pf = fi[f.parent];
        const int njp = pf.nn + pf.na;
        const double* Wp = work + wptr[f.parent] * nrhs;
        const int* r = relidx + f.relptr;
        sdbl4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int i0 = 0; i0 < na; i0 += 16) {
            const int ii = lane & 15;
            const int i = i0 + ii;
            const bool ok = i < na;
            const int ri = ok ? r[i] : 0;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int c = (lane >> 4) + 4 * h;
                double v = 0.0, lv = 0.0;
                if (ok && c < rc) v = Wp[ri + (i64)(j0 + c) * njp];
                if (ok && c < nn) lv = L[nn + i + (i64)c * nj];
                ua[ii][c] = v;
                lb[ii][c] = lv;
                if (ok && c < rc) W[nn + i + (i64)(j0 + c) * nj] = v;
            }
            __syncthreads();
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int kk = 4 * s4 + kq;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(lb[kk][col], ua[kk][col], acc, 0, 0, 0);
            }
            __syncthreads();
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = kq + 4 * v;
            if (m < nn && col < rc) zs[m][col] += acc[v];
        }
    }
    __syncthreads();
    if (lane < rc) {  // x_N = L11^T z
        double z[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) z[i] = (i < nn) ? zs[i][lane] : 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < nn) {
                double y = 0.0;
#pragma unroll
                for (int c = i; c < 16; ++c)
                    if (c < nn) y += Ls[c][i] * z[c];
                B[pk[i] + (i64)(j0 + lane) * ldb] = y;
            }
        }
    }
}

// Large fronts, 'N', after the level's products (T = L11 (alpha b_N) at
// tmp + t.tofs, ld nn; slab rows >= nn = B21 T): the children's rows are
// added in map order; x_N = T + children's rows goes to B, the slab's rows
// >= nn keep their sums for the parent.  tiles[i] = (front, r0, T offset):
// rows r0 .. r0 + 63 of the slab.
__global__ __launch_bounds__(TPB) void k_trmm_finish(
    const STile* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const i64* __restrict__ snptr, const i64* __restrict__ perm, double* __restrict__ B,
    int ldb, int nrhs, double* __restrict__ work, const i64* __restrict__ wptr,
    const int* __restrict__ gptr, const GEnt* __restrict__ gent, const double* __restrict__ tmp) {
    const STile t = tiles[blockIdx.x];
    const int k = t.k;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int i = t.r0 + (int)(threadIdx.x & 63);
    if (i >= nj) return;
    const i64 wk = wptr[k];
    double* W = work + wk * nrhs;
    const double* T = tmp + t.tofs;
    const int e0 = gptr[wk + i], e1 = gptr[wk + i + 1];
    const i64 bi = (i < nn) ? perm[snptr[k] + i] : 0;
    for (int j = (int)(threadIdx.x >> 6); j < nrhs; j += TPB / 64) {
        double g = (i < nn) ? T[i + (i64)j * nn] : W[i + (i64)j * nj];
        for (int e = e0; e < e1; ++e) {
            const GEnt q = gent[e];
            g += work[q.wu * nrhs + q.row + (i64)j * q.ld];
        }
        if (i < nn) B[bi + (i64)j * ldb] = g;
        else W[i + (i64)j * nj] = g;
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

// Xt = X as a row-major n x nrhs array (Xt[i * nrhs + j] = X[i + j * ldx]):
// 64 x 64 tiles through LDS, grid (ceil(n / 64), ceil(nrhs / 64)).
__global__ __launch_bounds__(TPB) void k_to_rowmajor(i64 n, int nrhs,
                                                     const double* __restrict__ X, int ldx,
                                                     double* __restrict__ Xt) {
    __shared__ double t[64][65];
    const i64 r0 = (i64)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = w; c < 64; c += 4) {
        const i64 i = r0 + l;
        t[l][c] = (i < n && c0 + c < nrhs) ? X[i + (i64)(c0 + c) * ldx] : 0.0;
    }
    __syncthreads();
    for (int r = w; r < 64; r += 4) {
        const i64 i = r0 + r;
        if (i < n && c0 + l < nrhs) Xt[i * nrhs + c0 + l] = t[r][l];
    }
}

// R = B - A X for a CSR matrix A (full symmetric pattern) with X given
// row-major (Xt, n x nrhs).  A workgroup owns 64 rows and up to 64
// right-hand sides (blockIdx.y); B and R pass through an LDS tile so that
// their column-major accesses are coalesced; in the product, lpr lanes share
// a row (one right-hand side each), so each nonzero reads a contiguous piece
// of a row of Xt.  Each row sums its nonzeros in CSR order.
__global__ __launch_bounds__(TPB) void k_spmv_residual_rm(
    i64 n, const i64* __restrict__ rowptr, const int* __restrict__ colind,
    const double* __restrict__ vals, const double* __restrict__ Xt, int nrhs,
    const double* __restrict__ B, int ldb, double* __restrict__ R, int ldr) {
    __shared__ double t[64][65];
    const i64 r0 = (i64)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int cw = min(64, nrhs - c0);
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = w; c < 64; c += 4) {
        const i64 i = r0 + l;
        t[l][c] = (i < n && c < cw) ? B[i + (i64)(c0 + c) * ldb] : 0.0;
    }
    __syncthreads();
    int lpr = 1;
    while (lpr < cw) lpr <<= 1;
    const int rpw = 64 / lpr;            // rows per wave and pass
    const int sub = l / lpr, col = l % lpr;
    for (int r = w * rpw + sub; r < 64; r += 4 * rpw) {
        const i64 i = r0 + r;
        if (i >= n || col >= cw) continue;
        double s = 0.0;
        const i64 p1 = rowptr[i + 1];
        for (i64 p = rowptr[i]; p < p1; ++p)
            s += vals[p] * Xt[(i64)colind[p] * nrhs + c0 + col];
        t[r][col] -= s;
    }
    __syncthreads();
    for (int c = w; c < 64; c += 4) {
        const i64 i = r0 + l;
        if (i < n && c < cw) R[i + (i64)(c0 + c) * ldr] = t[l][c];
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
