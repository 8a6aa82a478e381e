// Large-front kernels.  Fronts too big for one workgroup's LDS are processed
// by many workgroups at once through per-level work lists: a column-tiled
// extend-add, a blocked right-looking factorization (64-wide panels), and a
// batched FP64 MFMA GEMM that carries the Schur-complement updates and the
// projected-inverse products.  Work lists are built once per plan on the host.
#include "common.h"
#include "kernels.h"

#define TPB 256

typedef double dbl4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
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
__global__ __launch_bounds__(TPB) void k_extadd_tiles(
    const int4* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ upd,
    const int* __restrict__ relidx, const int* __restrict__ chidx, int zero22,
    i64 bstride, i64 ustride) {
    __shared__ double S[TPB * XT_LD];  // S[(row - r0) * XT_LD + (col - c0)]
    // grid y: the instance (a Hessian direction) whose blkval and update
    // arena lie bstride and ustride doubles after the previous one's
    blkval += (i64)blockIdx.y * bstride;
    upd += (i64)blockIdx.y * ustride;
    const int4 t = tiles[blockIdx.x];
    const int k = t.x, c0 = t.y, r0 = t.z;
    const FrontInfo f = fi[k];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int c1 = min(c0 + 16, nj), r1 = min(r0 + TPB, nj);
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

// ---------------------------------------------------------------------------
// Diagonal blocks.  One 256-thread workgroup per 64 x 64 diagonal block,
// padded with the identity beyond nb.  Serial steps, not flops, bound these
// kernels (a barrier step costs about a microsecond once the dependent FP64
// pivot chain is included), so:
//   * the factorization eliminates 4 columns per step (16 steps): every
//     thread factors the 4 x 4 pivot block redundantly from LDS, computes its
//     own row of the 4-column panel, and applies the rank-4 update to the 16
//     elements it owns, (l, w + 4q) for thread (w = tid/64, l = tid%64);
//   * the inverse uses recursive doubling: closed-form inverses of the 16
//     diagonal 4 x 4 blocks, then 4 levels of X21 = -X22 (L21 X11), 8 barriers.
// Inverses are stored row-major: Dinv(r, j) at dinv[j + r*NB].

#define DQ 16  // elements per thread

// 1/sqrt(d) from the hardware estimate refined by one Newton step
// (relative error ~ ulp); the pivot chain is latency bound.
__device__ inline double prsqrt(double d) {
    const double r = __builtin_amdgcn_rsq(d);
    const double h = d * r;
    return fma(fma(-h, r, 1.0), 0.5 * r, r);
}

// Factor the padded block held in a[] in place.  Returns false (uniformly)
// on a non-positive pivot, whose column goes to *bad.  rdiag[c] = 1/L(c,c).
__device__ inline bool block_potrf4(double (&a)[DQ], double (*P)[5], double (*Q)[5],
                                    double* rdiag, int w, int l, int* bad) {
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) {
        const int r0 = 4 * s;
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < DQ; ++q) if (q == s) v = a[q];
        P[l][w] = v;  // column r0 + w
        __syncthreads();
        // 4 x 4 pivot block, redundantly in every thread
        const double d00 = P[r0][0], d10 = P[r0 + 1][0], d20 = P[r0 + 2][0], d30 = P[r0 + 3][0];
        const double d11 = P[r0 + 1][1], d21 = P[r0 + 2][1], d31 = P[r0 + 3][1];
        const double d22 = P[r0 + 2][2], d32 = P[r0 + 3][2], d33 = P[r0 + 3][3];
        const double i0 = prsqrt(d00);
        const double L10 = d10 * i0, L20 = d20 * i0, L30 = d30 * i0;
        const double e11 = d11 - L10 * L10;
        const double i1 = prsqrt(e11);
        const double L21 = (d21 - L20 * L10) * i1, L31 = (d31 - L30 * L10) * i1;
        const double e22 = d22 - L20 * L20 - L21 * L21;
        const double i2 = prsqrt(e22);
        const double L32 = (d32 - L30 * L20 - L31 * L21) * i2;
        const double e33 = d33 - L30 * L30 - L31 * L31 - L32 * L32;
        const double i3 = prsqrt(e33);
        const int fail = !(d00 > 0.0) ? 0 : (!(e11 > 0.0) ? 1 : (!(e22 > 0.0) ? 2 : (!(e33 > 0.0) ? 3 : -1)));
        if (fail >= 0) { *bad = r0 + fail; return false; }
        // this thread's row of the panel: x = P[l, 0:4] L_pp^{-T}
        double x0, x1, x2, x3;
        if (l >= r0 + 4) {
            const double p0 = P[l][0], p1 = P[l][1], p2 = P[l][2], p3 = P[l][3];
            x0 = p0 * i0;
            x1 = (p1 - x0 * L10) * i1;
            x2 = (p2 - x0 * L20 - x1 * L21) * i2;
            x3 = (p3 - x0 * L30 - x1 * L31 - x2 * L32) * i3;
        } else {
            const int r = l - r0;  // rows of the pivot block, zero above it
            x0 = r == 0 ? d00 * i0 : (r == 1 ? L10 : (r == 2 ? L20 : (r == 3 ? L30 : 0.0)));
            x1 = r == 1 ? e11 * i1 : (r == 2 ? L21 : (r == 3 ? L31 : 0.0));
            x2 = r == 2 ? e22 * i2 : (r == 3 ? L32 : 0.0);
            x3 = r == 3 ? e33 * i3 : 0.0;
        }
        const double xw = w == 0 ? x0 : (w == 1 ? x1 : (w == 2 ? x2 : x3));
#pragma unroll
        for (int q = 0; q < DQ; ++q) if (q == s) a[q] = xw;
        if (w == 0) { Q[l][0] = x0; Q[l][1] = x1; Q[l][2] = x2; Q[l][3] = x3; }
        if (threadIdx.x == 0) { rdiag[r0] = i0; rdiag[r0 + 1] = i1; rdiag[r0 + 2] = i2; rdiag[r0 + 3] = i3; }
        __syncthreads();
        // rank-4 update of owned a(l, j), j = w + 4q > r0 + 3, l >= j
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int j = w + 4 * q;
            const double q0 = Q[j][0], q1 = Q[j][1], q2 = Q[j][2], q3 = Q[j][3];
            const double m = (j >= r0 + 4 && l >= j) ? 1.0 : 0.0;
            a[q] -= m * (x0 * q0 + x1 * q1 + x2 * q2 + x3 * q3);
        }
    }
    __syncthreads();
    return true;
}

// X = L^{-1} (64 x 64, in LDS) from L (lower, in LDS) and rdiag = 1/diag(L),
// by recursive doubling.  W is a 64 x 65 LDS temporary.
__device__ inline void block_trtri_rd(const double (*Ls)[NB + 1], const double* rdiag,
                                      double (*X)[NB + 1], double (*W)[NB + 1]) {
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += TPB) X[e / NB][e % NB] = 0.0;
    __syncthreads();
    if (tid < NB / 4) {
        const int b = 4 * tid;
        // closed-form inverse of the 4 x 4 lower block at (b, b)
        const double y0 = rdiag[b], y1 = rdiag[b + 1], y2 = rdiag[b + 2], y3 = rdiag[b + 3];
        const double l10 = Ls[b + 1][b], l20 = Ls[b + 2][b], l30 = Ls[b + 3][b];
        const double l21 = Ls[b + 2][b + 1], l31 = Ls[b + 3][b + 1], l32 = Ls[b + 3][b + 2];
        const double x10 = -y1 * l10 * y0;
        const double x21 = -y2 * l21 * y1;
        const double x20 = -y2 * (l20 * y0 + l21 * x10);
        const double x32 = -y3 * l32 * y2;
        const double x31 = -y3 * (l31 * y1 + l32 * x21);
        const double x30 = -y3 * (l30 * y0 + l31 * x10 + l32 * x20);
        X[b][b] = y0; X[b + 1][b + 1] = y1; X[b + 2][b + 2] = y2; X[b + 3][b + 3] = y3;
        X[b + 1][b] = x10; X[b + 2][b] = x20; X[b + 3][b] = x30;
        X[b + 2][b + 1] = x21; X[b + 3][b + 1] = x31; X[b + 3][b + 2] = x32;
    }
    __syncthreads();
#pragma unroll 1
    for (int bs = 4; bs < NB; bs *= 2) {
        const int nout = (NB / (2 * bs)) * bs * bs;
        // W = L21 X11 (X11 lower: k >= j)
        for (int e = tid; e < nout; e += TPB) {
            const int pr = e / (bs * bs), i = (e / bs) % bs, j = e % bs;
            const int a0 = pr * 2 * bs;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int k = j;
            for (; k + 3 < bs; k += 4) {
                s0 += Ls[a0 + bs + i][a0 + k] * X[a0 + k][a0 + j];
                s1 += Ls[a0 + bs + i][a0 + k + 1] * X[a0 + k + 1][a0 + j];
                s2 += Ls[a0 + bs + i][a0 + k + 2] * X[a0 + k + 2][a0 + j];
                s3 += Ls[a0 + bs + i][a0 + k + 3] * X[a0 + k + 3][a0 + j];
            }
            for (; k < bs; ++k) s0 += Ls[a0 + bs + i][a0 + k] * X[a0 + k][a0 + j];
            W[a0 + bs + i][a0 + j] = (s0 + s1) + (s2 + s3);
        }
        __syncthreads();
        // X21 = -X22 W (X22 lower: k <= i)
        for (int e = tid; e < nout; e += TPB) {
            const int pr = e / (bs * bs), i = (e / bs) % bs, j = e % bs;
            const int a0 = pr * 2 * bs;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int k = 0;
            for (; k + 3 <= i; k += 4) {
                s0 += X[a0 + bs + i][a0 + bs + k] * W[a0 + bs + k][a0 + j];
                s1 += X[a0 + bs + i][a0 + bs + k + 1] * W[a0 + bs + k + 1][a0 + j];
                s2 += X[a0 + bs + i][a0 + bs + k + 2] * W[a0 + bs + k + 2][a0 + j];
                s3 += X[a0 + bs + i][a0 + bs + k + 3] * W[a0 + bs + k + 3][a0 + j];
            }
            for (; k <= i; ++k) s0 += X[a0 + bs + i][a0 + bs + k] * W[a0 + bs + k][a0 + j];
            X[a0 + bs + i][a0 + j] = -((s0 + s1) + (s2 + s3));
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(TPB) void k_potrf_diag(
    const int* __restrict__ fronts, int c0, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ dinv,
    const i64* __restrict__ dinvptr, unsigned long long* status) {
    __shared__ double Ls[NB][NB + 1];
    __shared__ double X[NB][NB + 1];
    __shared__ double W[NB][NB + 1];
    __shared__ double P[NB][5], Q[NB][5];
    __shared__ double rdiag[NB];
    const int k = fronts[blockIdx.x];
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double* blk = blkval + f.blkptr + c0 + (i64)c0 * nj;
    double a[DQ];
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int j = w + 4 * q;
        a[q] = (l < nb && j < nb && l >= j) ? blk[l + (i64)j * nj] : (l == j ? 1.0 : 0.0);
    }
    int bad = 0;
    if (!block_potrf4(a, P, Q, rdiag, w, l, &bad)) {
        if (threadIdx.x == 0) chp_record_failure(status, k, c0 + bad + 1);
        return;
    }
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int j = w + 4 * q;
        Ls[l][j] = (l >= j) ? a[q] : 0.0;
        if (l < nb && j < nb && l >= j) blk[l + (i64)j * nj] = a[q];
    }
    __syncthreads();
    block_trtri_rd(Ls, rdiag, X, W);
    double* inv = dinv + dinvptr[k] + (i64)(c0 / NB) * NB * NB;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int r = w + 4 * q;  // row r, column l of Dinv
        inv[l + r * NB] = (r < nb && l < nb) ? X[r][l] : 0.0;
    }
}

// One inner step of the large-front Cholesky: tasks[i] = (front, c0, r0, kp).
// Every workgroup factors the 64 x 64 diagonal block at (c0, c0) from the
// front's current values (redundantly: the chain of dependent pivots is the
// same length for all of them) together with the 64 rows r0 .. r0 + 63 of the
// panel below it, and writes those rows of the factor.  Columns kp .. c0 - 1
// (the earlier inner panels of the outer panel, already factored) are applied
// first, left-looking: D -= Lc Lc^T and R -= Lr Lc^T with Lc, Lr their rows
// c0 .. c0 + 63 and r0 .. r0 + 63, on v_mfma_f64_16x16x4f64 from LDS, 64
// columns at a time (each workgroup forms the diagonal block's update
// itself, as it factors the block itself).  The workgroup with
// r0 < 0 reports a failed pivot and stages the factored diagonal block in the
// front's inverse area (dinv + dinvptr[k] + (c0 / 64) * 4096, ld 64): writing
// it into blkval here would race with workgroups of the same launch that
// have not read the block yet.  A later launch moves it into blkval.
// PANEL_WAVES (8) waves: thread (w, l) owns diagonal-block element
// (l, w + 8 q) in a[q] and panel element (r0 + l, w + 8 q) in b[q].  Each of
// the 16 steps eliminates 4 columns: the owners of those columns publish
// them (P), every thread factors the 4 x 4 pivot block redundantly and forms
// its row of the 4-column panel, and the rank-4 update of the columns right
// of the step is spread over all 8 waves.  The step is latency bound (a
// dependent FP64 operation costs tens of cycles): two waves per SIMD hide
// part of that, one wave per SIMD (4 waves) measured about 20% slower.
#define PANEL_WAVES 8
__global__ __launch_bounds__(64 * PANEL_WAVES) void k_panel_factor(
    const int4* __restrict__ tasks, const FrontInfo* __restrict__ fi,
    double* __restrict__ blkval, double* __restrict__ dinv, const i64* __restrict__ dinvptr,
    unsigned long long* status) {
    constexpr int W = PANEL_WAVES, C = 4, DQW = NB / W;
    __shared__ double P[2 * NB][C + 1];
    __shared__ __attribute__((aligned(16))) double Q[NB][C];
    __shared__ double S[2][NB][NB + 2];  // left-looking update: [0] Lc, [1] Lr as [k][row]; then the updates as [col][row]
    const int4 t = tasks[blockIdx.x];
    const int k = t.x, c0 = t.y, r0 = t.z, kp = t.w;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double* blk = blkval + f.blkptr;
    double* dg = blk + c0 + (i64)c0 * nj;
    const bool prow = r0 >= 0 && r0 + l < nj;
    double a[DQW], b[DQW];
#pragma unroll
    for (int q = 0; q < DQW; ++q) {
        const int j = w + W * q;
        a[q] = (l < nb && j < nb && l >= j) ? dg[l + (i64)j * nj] : (l == j ? 1.0 : 0.0);
        b[q] = (prow && j < nb) ? blk[(r0 + l) + (i64)(c0 + j) * nj] : 0.0;
    }
    if (kp < c0) {
        // output blocks (I, J) of 16 x 16: slots 0 .. 9 the lower blocks of
        // D (J <= I), 10 .. 25 the blocks of R; wave w takes slots w + 8 s
        const int lane = l;
        const int nslot = r0 >= 0 ? 26 : 10;
        dbl4 acc[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) acc[s4] = (dbl4){0.0, 0.0, 0.0, 0.0};
        double vc[DQW], vr[DQW];
        auto load_chunk = [&](int kc) {
#pragma unroll
            for (int m = 0; m < DQW; ++m) {  // thread (w, l): row l, chunk column w + 8 m
                const i64 col = (i64)(kc + w + W * m) * nj;
                vc[m] = blk[(l < nb ? c0 + l : c0) + col];
                vr[m] = blk[(prow ? r0 + l : c0) + col];
            }
        };
        load_chunk(kp);
        for (int kc = kp; kc < c0; kc += NB) {
            if (kc > kp) __syncthreads();  // the previous chunk's reads are done
#pragma unroll
            for (int m = 0; m < DQW; ++m) {
                S[0][w + W * m][l] = l < nb ? vc[m] : 0.0;
                S[1][w + W * m][l] = prow ? vr[m] : 0.0;
            }
            __syncthreads();
            if (kc + NB < c0) load_chunk(kc + NB);  // in flight during this chunk's products
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int slot = w + W * s4;
                if (slot >= nslot) break;
                int I, J, src;
                if (slot < 10) {
                    I = slot >= 6 ? 3 : slot >= 3 ? 2 : slot >= 1 ? 1 : 0;
                    J = slot - I * (I + 1) / 2;
                    src = 0;
                } else {
                    I = (slot - 10) >> 2;
                    J = (slot - 10) & 3;
                    src = 1;
                }
#pragma unroll 4
                for (int k4 = 0; k4 < NB; k4 += 4) {
                    const int kr = k4 + (lane >> 4);
                    const double av = S[src][kr][I * 16 + (lane & 15)];
                    const double bv = S[0][kr][J * 16 + (lane & 15)];
                    acc[s4] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[s4], 0, 0, 0);
                }
            }
        }
        __syncthreads();  // the last chunk's reads are done: S takes the updates
        // accumulator lane l, element v: row 16 I + l / 16 + 4 v, column 16 J + l % 16
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int slot = w + W * s4;
            if (slot >= nslot) break;
            int I, J, src;
            if (slot < 10) {
                I = slot >= 6 ? 3 : slot >= 3 ? 2 : slot >= 1 ? 1 : 0;
                J = slot - I * (I + 1) / 2;
                src = 0;
            } else {
                I = (slot - 10) >> 2;
                J = (slot - 10) & 3;
                src = 1;
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
                S[src][J * 16 + (lane & 15)][I * 16 + (lane >> 4) + 4 * v] = acc[s4][v];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DQW; ++q) {
            const int j = w + W * q;
            if (l < nb && j < nb && l >= j) a[q] -= S[0][j][l];
            if (prow && j < nb) b[q] -= S[1][j][l];
        }
    }
    int bad = -1;
#pragma unroll
    for (int s = 0; s < NB / C; ++s) {
        const int rs = C * s, qs = rs / W, lo = rs % W;
        const bool own = w >= lo && w < lo + C;  // this thread holds a column of the step
        if (own) { P[l][w - lo] = a[qs]; P[NB + l][w - lo] = b[qs]; }
        __syncthreads();
        // the pivot block's factor Lb (lower) and 1/diag
        double Lb[C][C], inv[C], x[C], y[C];
        int fail = -1;
#pragma unroll
        for (int tt = 0; tt < C; ++tt) {
            double d = P[rs + tt][tt];
#pragma unroll
            for (int m = 0; m < tt; ++m) d -= Lb[tt][m] * Lb[tt][m];
            if (fail < 0 && !(d > 0.0)) fail = tt;
            inv[tt] = prsqrt(d);
#pragma unroll
            for (int i2 = tt + 1; i2 < C; ++i2) {
                double v = P[rs + i2][tt];
#pragma unroll
                for (int m = 0; m < tt; ++m) v -= Lb[i2][m] * Lb[tt][m];
                Lb[i2][tt] = v * inv[tt];
            }
        }
        if (fail >= 0) { bad = rs + fail; break; }  // (the same in every thread)
        // this thread's block row l and panel row r0 + l in the step's columns
        // (the block's own rows come out as its factor's entries; zero above)
#pragma unroll
        for (int tt = 0; tt < C; ++tt) {
            double v = P[l][tt], u = P[NB + l][tt];
#pragma unroll
            for (int m = 0; m < tt; ++m) { v -= x[m] * Lb[tt][m]; u -= y[m] * Lb[tt][m]; }
            x[tt] = v * inv[tt];
            y[tt] = u * inv[tt];
        }
        const int r = l - rs;
#pragma unroll
        for (int tt = 0; tt < C; ++tt)
            if (r < tt) x[tt] = 0.0;
        if (own) {
            double xw = x[0], yw = y[0];
#pragma unroll
            for (int tt = 1; tt < C; ++tt)
                if (tt == w - lo) { xw = x[tt]; yw = y[tt]; }
            a[qs] = xw;
            b[qs] = yw;
        }
        if (w == 0) {
#pragma unroll
            for (int tt = 0; tt < C; ++tt) Q[l][tt] = x[tt];
        }
        __syncthreads();
        // rank-C update of the owned columns right of the step
#pragma unroll
        for (int q = qs; q < DQW; ++q) {
            const int j = w + W * q;
            if (j < rs + C) continue;
            double ta = 0.0, tb = 0.0;
#pragma unroll
            for (int tt = 0; tt < C; ++tt) {
                const double c = Q[j][tt];
                ta = fma(x[tt], c, ta);
                tb = fma(y[tt], c, tb);
            }
            const double m = (l >= j) ? 1.0 : 0.0;
            a[q] -= m * ta;
            b[q] -= tb;
        }
    }
    if (bad >= 0) {
        if (r0 < 0 && threadIdx.x == 0) chp_record_failure(status, k, c0 + bad + 1);
        return;
    }
#pragma unroll
    for (int q = 0; q < DQW; ++q) {
        const int j = w + W * q;
        if (j >= nb) continue;
        if (r0 < 0) {
            if (l < nb) dinv[dinvptr[k] + (i64)(c0 / NB) * NB * NB + l + (i64)j * NB] = a[q];
        } else if (prow) {
            blk[(r0 + l) + (i64)(c0 + j) * nj] = b[q];
        }
    }
}

// Inverse of every diagonal block of L11 (projected inverse, the diagonal
// part of dtrtri), also written into the diagonal blocks of X = L11^{-1}
// (column-major, ld nn).  blocks[i] = (front, block index).
__global__ __launch_bounds__(TPB) void k_trtri_diag(
    const int2* __restrict__ blocks, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, double* __restrict__ dinv,
    const i64* __restrict__ dinvptr, double* __restrict__ Xg,
    const i64* __restrict__ xptr, unsigned long long* status) {
    __shared__ double Ls[NB][NB + 1];
    __shared__ double X[NB][NB + 1];
    __shared__ double W[NB][NB + 1];
    __shared__ double rdiag[NB];
    __shared__ int zero_at;
    const int2 b = blocks[blockIdx.x];
    const int k = b.x, c0 = b.y * NB;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const int nb = min(NB, nn - c0);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const double* blk = blkval + f.blkptr + c0 + (i64)c0 * nj;
    if (threadIdx.x == 0) zero_at = NB;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int j = w + 4 * q;
        Ls[l][j] = (l < nb && j < nb && l >= j) ? blk[l + (i64)j * nj] : (l == j ? 1.0 : 0.0);
    }
    __syncthreads();
    if (w == 0) {
        if (l < nb && Ls[l][l] == 0.0) atomicMin(&zero_at, l);
        rdiag[l] = 1.0 / Ls[l][l];
    }
    __syncthreads();
    if (zero_at < NB) {
        if (threadIdx.x == 0) chp_record_failure(status, k, c0 + zero_at + 1);
        return;
    }
    block_trtri_rd(Ls, rdiag, X, W);
    double* inv = dinv + dinvptr[k] + (i64)b.y * NB * NB;
    double* xd = Xg + xptr[k] + c0 + (i64)c0 * nn;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int j = w + 4 * q;  // element (l, j) of the block of X
        if (l < nb && j < nb) xd[l + (i64)j * nn] = X[l][j];
        const int r = w + 4 * q;  // row r, column l of Dinv
        inv[l + r * NB] = (r < nb && l < nb) ? X[r][l] : 0.0;
    }
}

// Descriptor-addressed diagonal blocks (dense operations of the completion
// and Hessian sweeps): mode 0 factors the block in place and writes the
// inverse of the factor row-major to the output; mode 1 writes the inverse
// of the lower-triangular block column-major (zero above the diagonal).
// Launches whose BufTab gate reads 0 do nothing.
__global__ __launch_bounds__(TPB) void k_diag_tasks(const DiagTask* __restrict__ tasks,
                                                    BufTab bt, unsigned long long* status) {
    __shared__ double Ls[NB][NB + 1];
    __shared__ double X[NB][NB + 1];
    __shared__ double W[NB][NB + 1];
    __shared__ double P[NB][5], Q[NB][5];
    __shared__ double rdiag[NB];
    __shared__ int zero_at;
    if (gated_off(bt)) return;
    const DiagTask t = tasks[blockIdx.x];
    const int nb = t.nb, ld = t.ld;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double* blk = buf_sel(bt, t.buf) + t.off;
    double* out = buf_sel(bt, t.obuf) + t.outoff;
    if (t.mode == 0) {
        double a[DQ];
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int j = w + 4 * q;
            a[q] = (l < nb && j < nb && l >= j) ? blk[l + (i64)j * ld] : (l == j ? 1.0 : 0.0);
        }
        int bad = 0;
        if (!block_potrf4(a, P, Q, rdiag, w, l, &bad)) {
            if (threadIdx.x == 0) chp_record_failure(status, t.key, t.col + bad + 1);
            return;
        }
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int j = w + 4 * q;
            Ls[l][j] = (l >= j) ? a[q] : 0.0;
            if (l < nb && j < nb && l >= j) blk[l + (i64)j * ld] = a[q];
        }
        __syncthreads();
        block_trtri_rd(Ls, rdiag, X, W);
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int r = w + 4 * q;
            out[l + r * NB] = (r < nb && l < nb) ? X[r][l] : 0.0;
        }
        return;
    }
    if (threadIdx.x == 0) zero_at = NB;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int j = w + 4 * q;
        Ls[l][j] = (l < nb && j < nb && l >= j) ? blk[l + (i64)j * ld] : (l == j ? 1.0 : 0.0);
    }
    __syncthreads();
    if (t.mode == 2) {
        double* cp2 = buf_sel(bt, t.buf2) + t.off2;
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int j = w + 4 * q;
            if (l < nb && j < nb && l >= j) cp2[l + (i64)j * t.ld2] = Ls[l][j];
        }
    }
    if (w == 0) {
        if (l < nb && Ls[l][l] == 0.0) atomicMin(&zero_at, l);
        rdiag[l] = 1.0 / Ls[l][l];
    }
    __syncthreads();
    // a zero pivot is recorded; the inverse is still formed (non-finite, as
    // a triangular solve with the block would be)
    if (zero_at < NB && threadIdx.x == 0) chp_record_failure(status, t.key, t.col + zero_at + 1);
    block_trtri_rd(Ls, rdiag, X, W);
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int j = w + 4 * q;
        if (l < nb && j < nb) out[l + (i64)j * t.ldo] = X[l][j];
    }
}

// Element-wise tile operations (see MatOpKind).  One workgroup per 64 x 64
// tile; consecutive threads walk rows.
__global__ __launch_bounds__(TPB) void k_matop(const MatOp* __restrict__ ops, BufTab bt) {
    if (gated_off(bt)) return;
    const MatOp o = ops[blockIdx.x];
    const double* S = buf_sel(bt, o.sbuf) + o.src;
    double* D = buf_sel(bt, o.dbuf) + o.dst;
    const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
    const int i = o.i0 + il;
    if (i >= o.m) return;
    for (int jl = jl0; jl < 64; jl += 4) {
        const int j = o.j0 + jl;
        if (j >= o.n) break;
        double* d = D + i + (i64)j * o.ldd;
        switch (o.kind) {
        case MO_COPY: *d = o.alpha * S[i + (i64)j * o.lds]; break;
        case MO_COPY_LOWER: if (i >= j) *d = o.alpha * S[i + (i64)j * o.lds]; break;
        case MO_COPY_LOWER_Z: *d = (i >= j) ? o.alpha * S[i + (i64)j * o.lds] : 0.0; break;
        case MO_RT: *d = (i >= j) ? S[(o.n - 1 - j) + (i64)(o.n - 1 - i) * o.lds] : 0.0; break;
        case MO_SYMM: if (i < j) *d = D[j + (i64)i * o.ldd]; break;
        case MO_ZERO_UPPER: if (i < j) *d = 0.0; break;
        case MO_PHI: if (i < j) *d = 0.0; else if (i == j) *d *= 0.5; break;
        case MO_SYMAVG:
            *d = (i >= j) ? o.alpha * (S[i + (i64)j * o.lds] + S[j + (i64)i * o.lds]) : 0.0;
            break;
        default: *d = 0.0; break;
        }
    }
}

// ---------------------------------------------------------------------------
// Batched FP64 GEMM on MFMA (v_mfma_f64_16x16x4f64).  One workgroup computes
// one 64 x 64 tile of one task; 4 waves in a 2 x 2 arrangement, each wave a
// 32 x 32 sub-tile made of 2 x 2 MFMA 16 x 16 accumulators.  K is consumed in
// chunks of 32 staged through LDS, the next chunk prefetched into registers
// while the current one is multiplied.  Accumulator layout (16 x 16 f64): lane l,
// element v holds row (l / 16) + 4 v, column l % 16.

#define BK 32


// Global loads and stores of the GEMM tiles.  COH = true: agent-scope
// atomic (cache-bypassing, coherent across the XCDs' L2 caches) accesses,
// for data that other workgroups of the same launch write (k_chol_dag).
template <bool COH>
__device__ inline double gld(const double* p) {
    if (COH)
        return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    return *p;
}
template <bool COH>
__device__ inline void gst(double* p, double v) {
    if (COH)
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __builtin_bit_cast(unsigned long long, v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

// Element (gi, gk) of opA and (gk, gj) of opB with the task's bounds and
// triangular masks; 0 outside.
template <int OPA, bool COH = false>
__device__ inline double ldA_chk(const GemmTask& T, const double* A, int gi, int gk) {
    bool ok = gi < T.M && gk < T.K;
    if (T.triA == 1) ok = ok && gk <= gi;
    else if (T.triA == 2) ok = ok && gk >= gi;
    if (!ok) return 0.0;
    if (OPA == 0) return gld<COH>(A + gi + (i64)gk * T.lda);
    if (OPA == 1) return gld<COH>(A + gk + (i64)gi * T.lda);
    return (gi >= gk) ? gld<COH>(A + gi + (i64)gk * T.lda) : gld<COH>(A + gk + (i64)gi * T.lda);
}
template <int OPB, bool COH = false>
__device__ inline double ldB_chk(const GemmTask& T, const double* B, int gk, int gj) {
    bool ok = gj < T.N && gk < T.K;
    if (T.pad == 1) ok = ok && gk >= gj;
    else if (T.pad == 2) ok = ok && gk <= gj;
    if (!ok) return 0.0;
    if (OPB == 0) return gld<COH>(B + gk + (i64)gj * T.ldb);
    if (OPB == 1) return gld<COH>(B + gj + (i64)gk * T.ldb);
    return (gk >= gj) ? gld<COH>(B + gk + (i64)gj * T.ldb) : gld<COH>(B + gj + (i64)gk * T.ldb);
}

// One 64 x 64 tile.  OPA: 0 A(i,k) = A[i + k lda], 1 A[k + i lda],
// 2 symmetric lower.  OPB: 0 B(k,j) = B[k + j ldb], 1 B[j + k ldb],
// 2 symmetric lower.
// CHK: per-element bounds and masks (edge tiles, triangular operands);
// otherwise a check-free interior path.  Threads are mapped so consecutive
// lanes read consecutive addresses of the stored operand.
template <int OPA, int OPB, bool CHK, bool COH = false>
__device__ inline void gemm_tile(const GemmTask& T, const double* __restrict__ A,
                                 const double* __restrict__ B, double* __restrict__ C,
                                 int row0, int col0, int kbase, int kend, double* Pp,
                                 double (*As)[TILE + 2], double (*Bs)[TILE + 2]) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    constexpr bool AK = (OPA == 1);   // A read along k
    constexpr bool BK_ = (OPB != 1);  // B read along k
    // per-thread coordinates of the 8 A and 8 B elements of a chunk
    const int ai = AK ? (tid >> 5) : (tid & 63);
    const int ak = AK ? (tid & 31) : (tid >> 6);
    const int bj = BK_ ? (tid >> 5) : (tid & 63);
    const int bk = BK_ ? (tid & 31) : (tid >> 6);
    // q-th element: AK: (ai + 8q, ak); else (ai, ak + 4q).  Same for B.
    double ra[8], rb[8];
    // chunks that need per-element checks: the whole task if CHK, else none;
    // with CHK, a chunk away from the triangles' diagonals and inside all
    // bounds still takes the plain path
    const bool tile_in = row0 + TILE <= T.M && col0 + TILE <= T.N;
    auto chunk_plain = [&](int kb) -> bool {
        if (!CHK) return true;
        if (!tile_in || kb + BK > T.K) return false;
        if (T.triA == 1 && !(kb + BK - 1 <= row0)) return false;
        if (T.triA == 2 && !(kb >= row0 + TILE - 1)) return false;
        if (T.pad == 1 && !(kb >= col0 + TILE - 1)) return false;
        if (T.pad == 2 && !(kb + BK - 1 <= col0)) return false;
        // a symmetric B chunk below the diagonal of B is a plain N read
        return OPA != 2 && (OPB != 2 || kb >= col0 + TILE - 1);
    };
    auto fetch = [&](int kb) {
        if (chunk_plain(kb)) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int gi = row0 + ai + (AK ? 8 * q : 0), gk = kb + ak + (AK ? 0 : 4 * q);
                ra[q] = (OPA == 1) ? gld<COH>(A + gk + (i64)gi * T.lda) : gld<COH>(A + gi + (i64)gk * T.lda);
                const int gj = col0 + bj + (BK_ ? 8 * q : 0), gk2 = kb + bk + (BK_ ? 0 : 4 * q);
                rb[q] = (OPB == 1) ? gld<COH>(B + gj + (i64)gk2 * T.ldb) : gld<COH>(B + gk2 + (i64)gj * T.ldb);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int gi = row0 + ai + (AK ? 8 * q : 0), gk = kb + ak + (AK ? 0 : 4 * q);
                ra[q] = ldA_chk<OPA, COH>(T, A, gi, gk);
                const int gj = col0 + bj + (BK_ ? 8 * q : 0), gk2 = kb + bk + (BK_ ? 0 : 4 * q);
                rb[q] = ldB_chk<OPB, COH>(T, B, gk2, gj);
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            As[ak + (AK ? 0 : 4 * q)][ai + (AK ? 8 * q : 0)] = ra[q];
            Bs[bk + (BK_ ? 0 : 4 * q)][bj + (BK_ ? 8 * q : 0)] = rb[q];
        }
    };
    dbl4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (dbl4){0.0, 0.0, 0.0, 0.0};
    // C is read in the epilogue (registers for an early read would cost
    // occupancy); split-K slices write unscaled-by-beta partials instead
    const double alpha = T.alpha, beta = Pp ? 0.0 : T.beta;
    if (kbase < kend) fetch(kbase);
    for (int kb = kbase; kb < kend; kb += BK) {
        stash();
        __syncthreads();
        if (kb + BK < kend) fetch(kb + BK);
#pragma unroll 2
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
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = row0 + wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int j = col0 + wc * 32 + b * 16 + (lane & 15);
                if (CHK && (i >= T.M || j >= T.N)) continue;
                if (Pp) { Pp[i + (i64)j * T.M] = alpha * acc[a][b][v]; continue; }
                double* c = C + i + (i64)j * T.ldc;
                if (T.tri == 0 || i >= j)
                    gst<COH>(c, alpha * acc[a][b][v] + (beta != 0.0 ? beta * gld<COH>(c) : 0.0));
                else if (T.tri == 2)
                    gst<COH>(c, 0.0);
            }
}

// The same tile for launches too small to fill the device (the host picks it
// for at most LAT_TILES tiles, split-K slices included): 8 waves, two groups of four
// that take alternate K chunks with a staging area each (the K loop runs in
// half the chunk rounds), partial sums added through LDS; C is read at the
// start, while the first chunks load, instead of after the K loop.  Groups:
// g = tid / 256; inside a group the layout of gemm_tile.  A split-K slice
// (Pp not null) writes alpha times its partial sum to Pp (ld M), as
// gemm_tile does.
template <int OPA, int OPB, bool CHK>
__device__ inline void gemm_tile_lat(const GemmTask& T, const double* __restrict__ A,
                                     const double* __restrict__ B, double* __restrict__ C,
                                     int row0, int col0, int kbase, int kend, double* Pp,
                                     double (*As)[BK][TILE + 2], double (*Bs)[BK][TILE + 2]) {
    const int g = threadIdx.x >> 8, tid = threadIdx.x & 255, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    constexpr bool AK = (OPA == 1);
    constexpr bool BK_ = (OPB != 1);
    const int ai = AK ? (tid >> 5) : (tid & 63);
    const int ak = AK ? (tid & 31) : (tid >> 6);
    const int bj = BK_ ? (tid >> 5) : (tid & 63);
    const int bk = BK_ ? (tid & 31) : (tid >> 6);
    double ra[8], rb[8];
    const bool tile_in = row0 + TILE <= T.M && col0 + TILE <= T.N;
    auto chunk_plain = [&](int kb) -> bool {
        if (!CHK) return true;
        if (!tile_in || kb + BK > T.K) return false;
        if (T.triA == 1 && !(kb + BK - 1 <= row0)) return false;
        if (T.triA == 2 && !(kb >= row0 + TILE - 1)) return false;
        if (T.pad == 1 && !(kb >= col0 + TILE - 1)) return false;
        if (T.pad == 2 && !(kb + BK - 1 <= col0)) return false;
        return OPA != 2 && (OPB != 2 || kb >= col0 + TILE - 1);
    };
    auto fetch = [&](int kb) {
        if (chunk_plain(kb)) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int gi = row0 + ai + (AK ? 8 * q : 0), gk = kb + ak + (AK ? 0 : 4 * q);
                ra[q] = (OPA == 1) ? A[gk + (i64)gi * T.lda] : A[gi + (i64)gk * T.lda];
                const int gj = col0 + bj + (BK_ ? 8 * q : 0), gk2 = kb + bk + (BK_ ? 0 : 4 * q);
                rb[q] = (OPB == 1) ? B[gj + (i64)gk2 * T.ldb] : B[gk2 + (i64)gj * T.ldb];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int gi = row0 + ai + (AK ? 8 * q : 0), gk = kb + ak + (AK ? 0 : 4 * q);
                ra[q] = ldA_chk<OPA>(T, A, gi, gk);
                const int gj = col0 + bj + (BK_ ? 8 * q : 0), gk2 = kb + bk + (BK_ ? 0 : 4 * q);
                rb[q] = ldB_chk<OPB>(T, B, gk2, gj);
            }
        }
    };
    double (*as)[TILE + 2] = As[g];
    double (*bs)[TILE + 2] = Bs[g];
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            as[ak + (AK ? 0 : 4 * q)][ai + (AK ? 8 * q : 0)] = ra[q];
            bs[bk + (BK_ ? 0 : 4 * q)][bj + (BK_ ? 8 * q : 0)] = rb[q];
        }
    };
    const double alpha = T.alpha, beta = Pp ? 0.0 : T.beta;
    // C of this thread's accumulator entries (group 0), loaded up front
    double cpre[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = row0 + wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int j = col0 + wc * 32 + b * 16 + (lane & 15);
                const bool need = g == 0 && beta != 0.0 && (!CHK || (i < T.M && j < T.N)) &&
                                  (T.tri == 0 || i >= j);
                cpre[a][b][v] = need ? C[i + (i64)j * T.ldc] : 0.0;
            }
    dbl4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (dbl4){0.0, 0.0, 0.0, 0.0};
    const int nch = (kend - kbase + BK - 1) / BK;
    // group g takes chunks g, g + 2, ...; both groups run the same rounds
    if (g < nch) fetch(kbase + g * BK);
    for (int c0 = 0; c0 < nch; c0 += 2) {
        const int c = c0 + g;
        if (c < nch) stash();
        __syncthreads();
        if (c + 2 < nch) fetch(kbase + (c + 2) * BK);
        if (c < nch) {
#pragma unroll 2
            for (int ks = 0; ks < BK / 4; ++ks) {
                const int kr = ks * 4 + (lane >> 4);
                const double a0 = as[kr][wr * 32 + (lane & 15)];
                const double a1 = as[kr][wr * 32 + 16 + (lane & 15)];
                const double q0 = bs[kr][wc * 32 + (lane & 15)];
                const double q1 = bs[kr][wc * 32 + 16 + (lane & 15)];
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, q0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, q1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, q0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, q1, acc[1][1], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // group 1's partial sums through LDS (the staging areas are free now)
    double (*red)[TILE + 2] = reinterpret_cast<double (*)[TILE + 2]>(&As[0][0][0]);
    if (g == 1) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    red[wr * 32 + a * 16 + (lane >> 4) + 4 * v][wc * 32 + b * 16 + (lane & 15)] = acc[a][b][v];
    }
    __syncthreads();
    if (g == 1) return;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int il = wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int jl = wc * 32 + b * 16 + (lane & 15);
                const int i = row0 + il, j = col0 + jl;
                if (CHK && (i >= T.M || j >= T.N)) continue;
                const double sum = acc[a][b][v] + red[il][jl];
                if (Pp) { Pp[i + (i64)j * T.M] = alpha * sum; continue; }
                double* c = C + i + (i64)j * T.ldc;
                if (T.tri == 0 || i >= j) *c = alpha * sum + (beta != 0.0 ? beta * cpre[a][b][v] : 0.0);
                else if (T.tri == 2) *c = 0.0;
            }
}

// Tiles of launches too small to fill the device (no split-K): 512 threads,
// gemm_tile_lat.  Same tile list and task table as k_gemm_tiles.
__global__ __launch_bounds__(2 * TPB) void k_gemm_tiles_lat(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks, BufTab bt) {
    __shared__ double As[2][BK][TILE + 2];
    __shared__ double Bs[2][BK][TILE + 2];
    if (gated_off(bt)) return;
    const GemmTile tl = tiles[blockIdx.x];
    const GemmTask T = tasks[tl.task];
    const double* A = bt.p[T.bufA] + T.offA;
    const double* B = bt.p[T.bufB] + T.offB;
    double* C = bt.p[T.bufC] + T.offC;
    const int row0 = tl.tm * TILE, col0 = tl.tn * TILE;
    int kbase = 0, kend = T.K;
    double* Pp = nullptr;
    if (T.ksplit > 1) {
        kbase = tl.ks * T.kchunk;
        kend = min(T.K, kbase + T.kchunk);
        Pp = bt.p[4] + T.offP + (i64)tl.ks * T.M * T.N;
        if (T.tri != 0 && row0 + TILE - 1 < col0) return;  // the reduction handles it
    }
    if (T.triA == 1) kend = min(kend, row0 + TILE);
    if (T.triA == 2) kbase = max(kbase, row0 & ~(BK - 1));
    if (T.pad == 1) kbase = max(kbase, col0 & ~(BK - 1));
    if (T.pad == 2) kend = min(kend, col0 + TILE);
    if (T.tri == 2 && row0 + TILE - 1 < col0) {
        for (int e = threadIdx.x; e < TILE * TILE; e += 2 * TPB) {
            int i = row0 + e % TILE, j = col0 + e / TILE;
            if (i < T.M && j < T.N) C[i + (i64)j * T.ldc] = 0.0;
        }
        return;
    }
    int opa = T.opA;
    const bool full = row0 + TILE <= T.M && col0 + TILE <= T.N && ((kend - kbase) % BK) == 0 &&
                      T.triA == 0 && T.pad == 0 && T.opB != 2;
    bool chk = !full;
    if (opa == 2 && full) {
        if (row0 >= T.K) opa = 0;
        else chk = true;
    }
    const int key = opa * 3 + T.opB;
    if (!chk) {
        switch (key) {
        case 0: gemm_tile_lat<0, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 1: gemm_tile_lat<0, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 3: gemm_tile_lat<1, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        default: gemm_tile_lat<1, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        }
    } else {
        switch (key) {
        case 0: gemm_tile_lat<0, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 1: gemm_tile_lat<0, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 2: gemm_tile_lat<0, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 3: gemm_tile_lat<1, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 4: gemm_tile_lat<1, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 5: gemm_tile_lat<1, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 6: gemm_tile_lat<2, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 7: gemm_tile_lat<2, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        default: gemm_tile_lat<2, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        }
    }
}

// GEMM tasks far smaller than a tile (the host sends single-tile tasks with
// K <= SMALL_TASK_K = 32 whose C does not overlap A or B here, e.g. the
// per-direction products of the Hessian's many tiny fronts): one task per
// 256-thread workgroup, no MFMA tile padding.  op(A) (M x K) and op(B)
// (K x N) are staged in LDS with all of a thread's loads in flight (bounds
// and triangular / symmetric forms through ldA_chk / ldB_chk), then each
// thread forms up to 16 outputs (rows across consecutive threads, so C
// writes coalesce) as K-term FMA sums from LDS.
template <int OPA, int OPB>
__device__ inline void small_task(const GemmTask& T, const double* __restrict__ A,
                                  const double* __restrict__ B, double* __restrict__ C,
                                  double (*As)[TILE + 1], double (*Bs)[TILE + 1]) {
    const int M = T.M, N = T.N, K = T.K, tid = threadIdx.x;
    double va[8], vb[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {  // element e = tid + 256 q of the K x 64 staging
        const int e = tid + TPB * q, k = e >> 6, x = e & 63;
        va[q] = (k < K && x < M) ? ldA_chk<OPA>(T, A, x, k) : 0.0;
        vb[q] = (k < K && x < N) ? ldB_chk<OPB>(T, B, k, x) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = tid + TPB * q, k = e >> 6, x = e & 63;
        if (k < SMALL_TASK_K) { As[k][x] = va[q]; Bs[k][x] = vb[q]; }
    }
    __syncthreads();
    for (int e = tid; e < M * N; e += TPB) {
        const int i = e % M, j = e / M;
        double* c = C + i + (i64)j * T.ldc;
        if (T.tri != 0 && i < j) {
            if (T.tri == 2) *c = 0.0;
            continue;
        }
        double s0 = 0.0, s1 = 0.0;
        int k = 0;
        for (; k + 1 < K; k += 2) {
            s0 = fma(As[k][i], Bs[k][j], s0);
            s1 = fma(As[k + 1][i], Bs[k + 1][j], s1);
        }
        if (k < K) s0 = fma(As[k][i], Bs[k][j], s0);
        *c = T.alpha * (s0 + s1) + (T.beta != 0.0 ? T.beta * *c : 0.0);
    }
}

__global__ __launch_bounds__(TPB) void k_gemm_small(const int* __restrict__ ids, int count,
                                                    const GemmTask* __restrict__ tasks, BufTab bt) {
    __shared__ double As[SMALL_TASK_K][TILE + 1];
    __shared__ double Bs[SMALL_TASK_K][TILE + 1];
    if (gated_off(bt)) return;
    const GemmTask T = tasks[ids[blockIdx.x]];
    const double* A = bt.p[T.bufA] + T.offA;
    const double* B = bt.p[T.bufB] + T.offB;
    double* C = bt.p[T.bufC] + T.offC;
    switch (T.opA * 3 + T.opB) {
    case 0: small_task<0, 0>(T, A, B, C, As, Bs); break;
    case 1: small_task<0, 1>(T, A, B, C, As, Bs); break;
    case 2: small_task<0, 2>(T, A, B, C, As, Bs); break;
    case 3: small_task<1, 0>(T, A, B, C, As, Bs); break;
    case 4: small_task<1, 1>(T, A, B, C, As, Bs); break;
    case 5: small_task<1, 2>(T, A, B, C, As, Bs); break;
    case 6: small_task<2, 0>(T, A, B, C, As, Bs); break;
    case 7: small_task<2, 1>(T, A, B, C, As, Bs); break;
    default: small_task<2, 2>(T, A, B, C, As, Bs); break;
    }
}

// Four waves per SIMD (the LDS of four workgroups fits a CU): the register
// budget is capped at 128 so that the K loop's load and LDS latency overlap
// with the other workgroups' MFMA work.
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_gemm_tiles(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks, BufTab bt) {
    __shared__ double As[BK][TILE + 2];
    __shared__ double Bs[BK][TILE + 2];
    if (gated_off(bt)) return;
    const GemmTile tl = tiles[blockIdx.x];
    const GemmTask T = tasks[tl.task];
    const double* A = bt.p[T.bufA] + T.offA;
    const double* B = bt.p[T.bufB] + T.offB;
    double* C = bt.p[T.bufC] + T.offC;
    const int row0 = tl.tm * TILE, col0 = tl.tn * TILE;
    int kbase = 0, kend = T.K;
    double* Pp = nullptr;
    if (T.ksplit > 1) {
        kbase = tl.ks * T.kchunk;
        kend = min(T.K, kbase + T.kchunk);
        Pp = bt.p[4] + T.offP + (i64)tl.ks * T.M * T.N;
        if (T.tri != 0 && row0 + TILE - 1 < col0) return;  // the reduction handles it
    }

    // triangular operands: chunks where opA or opB is identically zero are
    // skipped (A lower: k <= i < row0 + TILE; A upper: k >= i >= row0;
    // B lower: k >= j >= col0; B upper: k <= j < col0 + TILE)
    if (T.triA == 1) kend = min(kend, row0 + TILE);
    if (T.triA == 2) kbase = max(kbase, row0 & ~(BK - 1));
    if (T.pad == 1) kbase = max(kbase, col0 & ~(BK - 1));
    if (T.pad == 2) kend = min(kend, col0 + TILE);

    if (T.tri == 2 && row0 + TILE - 1 < col0) {
        for (int e = threadIdx.x; e < TILE * TILE; e += TPB) {
            int i = row0 + e % TILE, j = col0 + e / TILE;
            if (i < T.M && j < T.N) C[i + (i64)j * T.ldc] = 0.0;
        }
        return;
    }
    // interior: full tile, K a multiple of the chunk, no operand masks; the
    // symmetric operand qualifies when the whole tile sits on one side of the
    // diagonal of A (then it is a plain N or T read)
    int opa = T.opA;
    const bool full = row0 + TILE <= T.M && col0 + TILE <= T.N && ((kend - kbase) % BK) == 0 &&
                      T.triA == 0 && T.pad == 0 && T.opB != 2;
    bool chk = !full;
    if (opa == 2 && full) {
        if (row0 >= T.K) opa = 0;            // all gi >= gk: lower part, N read
        else chk = true;
    }
    const int key = opa * 3 + T.opB;
    if (!chk) {
        switch (key) {
        case 0: gemm_tile<0, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 1: gemm_tile<0, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 3: gemm_tile<1, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        default: gemm_tile<1, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        }
    } else {
        switch (key) {
        case 0: gemm_tile<0, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 1: gemm_tile<0, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 2: gemm_tile<0, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 3: gemm_tile<1, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 4: gemm_tile<1, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 5: gemm_tile<1, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 6: gemm_tile<2, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 7: gemm_tile<2, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        default: gemm_tile<2, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        }
    }
}

// ---------------------------------------------------------------------------
// The same tiles with FP32 arithmetic (mixed-precision Schur updates): FP64
// operands are rounded to FP32 as they are staged in LDS, multiplied on
// v_mfma_f32_16x16x4f32 with FP32 accumulation, and C = alpha acc + beta C is
// formed in FP64.  Accumulator layout (16 x 16 f32): lane l, element v holds
// row 4 (l / 16) + v, column l % 16.
typedef float flt4 __attribute__((ext_vector_type(4)));

template <int OPA, int OPB, bool CHK>
__device__ inline void gemm_tile32(const GemmTask& T, const double* __restrict__ A,
                                   const double* __restrict__ B, double* __restrict__ C,
                                   int row0, int col0, int kbase, int kend, double* Pp,
                                   float (*As)[TILE + 4], float (*Bs)[TILE + 4]) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    constexpr bool AK = (OPA == 1);
    constexpr bool BK_ = (OPB != 1);
    const int ai = AK ? (tid >> 5) : (tid & 63);
    const int ak = AK ? (tid & 31) : (tid >> 6);
    const int bj = BK_ ? (tid >> 5) : (tid & 63);
    const int bk = BK_ ? (tid & 31) : (tid >> 6);
    double ra[8], rb[8];
    const bool tile_in = row0 + TILE <= T.M && col0 + TILE <= T.N;
    auto chunk_plain = [&](int kb) -> bool {
        if (!CHK) return true;
        if (!tile_in || kb + BK > T.K) return false;
        if (T.triA == 1 && !(kb + BK - 1 <= row0)) return false;
        if (T.triA == 2 && !(kb >= row0 + TILE - 1)) return false;
        if (T.pad == 1 && !(kb >= col0 + TILE - 1)) return false;
        if (T.pad == 2 && !(kb + BK - 1 <= col0)) return false;
        return OPA != 2 && (OPB != 2 || kb >= col0 + TILE - 1);
    };
    auto fetch = [&](int kb) {
        const bool plain = chunk_plain(kb);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int gi = row0 + ai + (AK ? 8 * q : 0), gk = kb + ak + (AK ? 0 : 4 * q);
            const int gj = col0 + bj + (BK_ ? 8 * q : 0), gk2 = kb + bk + (BK_ ? 0 : 4 * q);
            if (plain) {
                ra[q] = (OPA == 1) ? A[gk + (i64)gi * T.lda] : A[gi + (i64)gk * T.lda];
                rb[q] = (OPB == 1) ? B[gj + (i64)gk2 * T.ldb] : B[gk2 + (i64)gj * T.ldb];
            } else {
                ra[q] = ldA_chk<OPA>(T, A, gi, gk);
                rb[q] = ldB_chk<OPB>(T, B, gk2, gj);
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            As[ak + (AK ? 0 : 4 * q)][ai + (AK ? 8 * q : 0)] = (float)ra[q];
            Bs[bk + (BK_ ? 0 : 4 * q)][bj + (BK_ ? 8 * q : 0)] = (float)rb[q];
        }
    };
    flt4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (flt4){0.f, 0.f, 0.f, 0.f};
    if (kbase < kend) fetch(kbase);
    for (int kb = kbase; kb < kend; kb += BK) {
        stash();
        __syncthreads();
        if (kb + BK < kend) fetch(kb + BK);
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + (lane >> 4);
            const float a0 = As[kr][wr * 32 + (lane & 15)];
            const float a1 = As[kr][wr * 32 + 16 + (lane & 15)];
            const float q0 = Bs[kr][wc * 32 + (lane & 15)];
            const float q1 = Bs[kr][wc * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, q0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, q1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, q0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, q1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    const double alpha = T.alpha, beta = Pp ? 0.0 : T.beta;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = row0 + wr * 32 + a * 16 + 4 * (lane >> 4) + v;
                const int j = col0 + wc * 32 + b * 16 + (lane & 15);
                if (CHK && (i >= T.M || j >= T.N)) continue;
                const double r = (double)acc[a][b][v];
                if (Pp) { Pp[i + (i64)j * T.M] = alpha * r; continue; }
                double* c = C + i + (i64)j * T.ldc;
                if (T.tri == 0 || i >= j) *c = alpha * r + (beta != 0.0 ? beta * (*c) : 0.0);
                else if (T.tri == 2) *c = 0.0;
            }
}

__global__ __launch_bounds__(TPB) void k_gemm_tiles32(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks, BufTab bt) {
    __shared__ float As[BK][TILE + 4];
    __shared__ float Bs[BK][TILE + 4];
    const GemmTile tl = tiles[blockIdx.x];
    const GemmTask T = tasks[tl.task];
    const double* A = bt.p[T.bufA] + T.offA;
    const double* B = bt.p[T.bufB] + T.offB;
    double* C = bt.p[T.bufC] + T.offC;
    const int row0 = tl.tm * TILE, col0 = tl.tn * TILE;
    int kbase = 0, kend = T.K;
    double* Pp = nullptr;
    if (T.ksplit > 1) {
        kbase = tl.ks * T.kchunk;
        kend = min(T.K, kbase + T.kchunk);
        Pp = bt.p[4] + T.offP + (i64)tl.ks * T.M * T.N;
        if (T.tri != 0 && row0 + TILE - 1 < col0) return;  // the reduction handles it
    }
    if (T.triA == 1) kend = min(kend, row0 + TILE);
    if (T.triA == 2) kbase = max(kbase, row0 & ~(BK - 1));
    if (T.pad == 1) kbase = max(kbase, col0 & ~(BK - 1));
    if (T.pad == 2) kend = min(kend, col0 + TILE);
    if (T.tri == 2 && row0 + TILE - 1 < col0) {
        for (int e = threadIdx.x; e < TILE * TILE; e += TPB) {
            const int i = row0 + e % TILE, j = col0 + e / TILE;
            if (i < T.M && j < T.N) C[i + (i64)j * T.ldc] = 0.0;
        }
        return;
    }
    const bool full = row0 + TILE <= T.M && col0 + TILE <= T.N && ((kend - kbase) % BK) == 0 &&
                      T.triA == 0 && T.pad == 0 && T.opA < 2 && T.opB < 2;
    const int key = T.opA * 3 + T.opB;
    if (full) {
        switch (key) {
        case 0: gemm_tile32<0, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 1: gemm_tile32<0, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 3: gemm_tile32<1, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        default: gemm_tile32<1, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        }
    } else {
        switch (key) {
        case 0: gemm_tile32<0, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 1: gemm_tile32<0, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 2: gemm_tile32<0, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 3: gemm_tile32<1, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 4: gemm_tile32<1, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 5: gemm_tile32<1, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 6: gemm_tile32<2, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        case 7: gemm_tile32<2, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        default: gemm_tile32<2, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, As, Bs); break;
        }
    }
}

// ---------------------------------------------------------------------------
// The same tiles in BF16x3 split precision (mixed-precision Schur updates):
// each FP64 operand is rounded to FP32, f, and split as it is staged in LDS
// into hi = bf16(f) and lo = bf16(f - hi) (round to nearest even; hi + lo
// carries about 16 bits of f).  Each 16 x 16 x 16 block product takes three
// v_mfma_f32_16x16x16bf16_1k (lo*hi, hi*lo, hi*hi; lo*lo is dropped) into
// one FP32 accumulator; C = beta C + alpha acc is formed in FP64.  Relative
// error about 2^-16 per product, removed by FP64 iterative refinement.
// Operand layout of the MFMA: lane l holds row (A) / column (B) l % 16 at
// k = 4 (l / 16) .. +3, so LDS keeps each row's k values contiguous and a
// lane reads its four bf16 values in one 8-byte load.
typedef short bf4 __attribute__((ext_vector_type(4)));
#define BKP (BK + 4)   // LDS row length (bf16 elements) of the split tiles

__device__ inline void split_bf16(double x, unsigned short& hi, unsigned short& lo) {
    const float f = (float)x;
    const unsigned u = __float_as_uint(f);
    const unsigned h = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    const float d = f - __uint_as_float(h << 16);  // exact in FP32
    const unsigned v = __float_as_uint(d);
    hi = (unsigned short)h;
    lo = (unsigned short)((v + 0x7fffu + ((v >> 16) & 1u)) >> 16);
}

template <int OPA, int OPB, bool CHK>
__device__ inline void gemm_tile_bf3(const GemmTask& T, const double* __restrict__ A,
                                     const double* __restrict__ B, double* __restrict__ C,
                                     int row0, int col0, int kbase, int kend, double* Pp,
                                     unsigned short (*Ah)[BKP], unsigned short (*Al)[BKP],
                                     unsigned short (*Bh)[BKP], unsigned short (*Bl)[BKP]) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    constexpr bool AK = (OPA == 1);
    constexpr bool BK_ = (OPB != 1);
    const int ai = AK ? (tid >> 5) : (tid & 63);
    const int ak = AK ? (tid & 31) : (tid >> 6);
    const int bj = BK_ ? (tid >> 5) : (tid & 63);
    const int bk = BK_ ? (tid & 31) : (tid >> 6);
    double ra[8], rb[8];
    const bool tile_in = row0 + TILE <= T.M && col0 + TILE <= T.N;
    auto chunk_plain = [&](int kb) -> bool {
        if (!CHK) return true;
        if (!tile_in || kb + BK > T.K) return false;
        if (T.triA == 1 && !(kb + BK - 1 <= row0)) return false;
        if (T.triA == 2 && !(kb >= row0 + TILE - 1)) return false;
        if (T.pad == 1 && !(kb >= col0 + TILE - 1)) return false;
        if (T.pad == 2 && !(kb + BK - 1 <= col0)) return false;
        return OPA != 2 && (OPB != 2 || kb >= col0 + TILE - 1);
    };
    auto fetch = [&](int kb) {
        const bool plain = chunk_plain(kb);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int gi = row0 + ai + (AK ? 8 * q : 0), gk = kb + ak + (AK ? 0 : 4 * q);
            const int gj = col0 + bj + (BK_ ? 8 * q : 0), gk2 = kb + bk + (BK_ ? 0 : 4 * q);
            if (plain) {
                ra[q] = (OPA == 1) ? A[gk + (i64)gi * T.lda] : A[gi + (i64)gk * T.lda];
                rb[q] = (OPB == 1) ? B[gj + (i64)gk2 * T.ldb] : B[gk2 + (i64)gj * T.ldb];
            } else {
                ra[q] = ldA_chk<OPA>(T, A, gi, gk);
                rb[q] = ldB_chk<OPB>(T, B, gk2, gj);
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = ai + (AK ? 8 * q : 0), k = ak + (AK ? 0 : 4 * q);
            split_bf16(ra[q], Ah[i][k], Al[i][k]);
            const int j = bj + (BK_ ? 8 * q : 0), k2 = bk + (BK_ ? 0 : 4 * q);
            split_bf16(rb[q], Bh[j][k2], Bl[j][k2]);
        }
    };
    flt4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (flt4){0.f, 0.f, 0.f, 0.f};
    if (kbase < kend) fetch(kbase);
    for (int kb = kbase; kb < kend; kb += BK) {
        stash();
        __syncthreads();
        if (kb + BK < kend) fetch(kb + BK);
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            const int kq = ks * 16 + 4 * (lane >> 4);
            bf4 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int r = wr * 32 + a * 16 + (lane & 15);
                ah[a] = *reinterpret_cast<const bf4*>(&Ah[r][kq]);
                al[a] = *reinterpret_cast<const bf4*>(&Al[r][kq]);
                const int c = wc * 32 + a * 16 + (lane & 15);
                bh[a] = *reinterpret_cast<const bf4*>(&Bh[c][kq]);
                bl[a] = *reinterpret_cast<const bf4*>(&Bl[c][kq]);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[a], bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[a], bl[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[a], bh[b], acc[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    const double alpha = T.alpha, beta = Pp ? 0.0 : T.beta;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = row0 + wr * 32 + a * 16 + 4 * (lane >> 4) + v;
                const int j = col0 + wc * 32 + b * 16 + (lane & 15);
                if (CHK && (i >= T.M || j >= T.N)) continue;
                const double r = (double)acc[a][b][v];
                if (Pp) { Pp[i + (i64)j * T.M] = alpha * r; continue; }
                double* c = C + i + (i64)j * T.ldc;
                if (T.tri == 0 || i >= j) *c = alpha * r + (beta != 0.0 ? beta * (*c) : 0.0);
                else if (T.tri == 2) *c = 0.0;
            }
}

__global__ __launch_bounds__(TPB) void k_gemm_tiles_bf3(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks, BufTab bt) {
    __shared__ __align__(16) unsigned short Ah[TILE][BKP];
    __shared__ __align__(16) unsigned short Al[TILE][BKP];
    __shared__ __align__(16) unsigned short Bh[TILE][BKP];
    __shared__ __align__(16) unsigned short Bl[TILE][BKP];
    const GemmTile tl = tiles[blockIdx.x];
    const GemmTask T = tasks[tl.task];
    const double* A = bt.p[T.bufA] + T.offA;
    const double* B = bt.p[T.bufB] + T.offB;
    double* C = bt.p[T.bufC] + T.offC;
    const int row0 = tl.tm * TILE, col0 = tl.tn * TILE;
    int kbase = 0, kend = T.K;
    double* Pp = nullptr;
    if (T.ksplit > 1) {
        kbase = tl.ks * T.kchunk;
        kend = min(T.K, kbase + T.kchunk);
        Pp = bt.p[4] + T.offP + (i64)tl.ks * T.M * T.N;
        if (T.tri != 0 && row0 + TILE - 1 < col0) return;  // the reduction handles it
    }
    if (T.triA == 1) kend = min(kend, row0 + TILE);
    if (T.triA == 2) kbase = max(kbase, row0 & ~(BK - 1));
    if (T.pad == 1) kbase = max(kbase, col0 & ~(BK - 1));
    if (T.pad == 2) kend = min(kend, col0 + TILE);
    if (T.tri == 2 && row0 + TILE - 1 < col0) {
        for (int e = threadIdx.x; e < TILE * TILE; e += TPB) {
            const int i = row0 + e % TILE, j = col0 + e / TILE;
            if (i < T.M && j < T.N) C[i + (i64)j * T.ldc] = 0.0;
        }
        return;
    }
    const bool full = row0 + TILE <= T.M && col0 + TILE <= T.N && ((kend - kbase) % BK) == 0 &&
                      T.triA == 0 && T.pad == 0 && T.opA < 2 && T.opB < 2;
    const int key = T.opA * 3 + T.opB;
    if (full) {
        switch (key) {
        case 0: gemm_tile_bf3<0, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 1: gemm_tile_bf3<0, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 3: gemm_tile_bf3<1, 0, false>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        default: gemm_tile_bf3<1, 1, false>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        }
    } else {
        switch (key) {
        case 0: gemm_tile_bf3<0, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 1: gemm_tile_bf3<0, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 2: gemm_tile_bf3<0, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 3: gemm_tile_bf3<1, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 4: gemm_tile_bf3<1, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 5: gemm_tile_bf3<1, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 6: gemm_tile_bf3<2, 0, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        case 7: gemm_tile_bf3<2, 1, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        default: gemm_tile_bf3<2, 2, true>(T, A, B, C, row0, col0, kbase, kend, Pp, Ah, Al, Bh, Bl); break;
        }
    }
}


// scratch[sptr[k] + i + j*nj] = blk[i + j*nj] for the rows i < nn of the
// listed (front, 16-column) tiles: the diagonal block as stored, kept for
// the product [L11; B21 L11][L11; B21 L11]^T of llt.c:32-33.
__global__ __launch_bounds__(TPB) void k_copy_top(
    const int2* __restrict__ tiles, const FrontInfo* __restrict__ fi,
    const double* __restrict__ blkval, double* __restrict__ scratch,
    const i64* __restrict__ sptr) {
    const int2 t = tiles[blockIdx.x];
    const int k = t.x, c0 = t.y;
    const FrontInfo f = fi[k];
    const int nn = f.nn, nj = nn + f.na;
    const double* blk = blkval + f.blkptr;
    double* S = scratch + sptr[k];
    const int c1 = min(c0 + 16, nn);
    for (int j = c0; j < c1; ++j)
        for (int i = threadIdx.x; i < nn; i += TPB) S[i + (i64)j * nj] = blk[i + (i64)j * nj];
}

// ---------------------------------------------------------------------------
// GEMM tasks with at most NARROW_N columns (the solves' products with a few
// right-hand sides): one workgroup per 64-row tile (and split-K slice), the
// four waves taking four consecutive parts of the slice's K range, summed in
// LDS in wave order.  A wave steps through K sixteen at a time with
// v_mfma_f64_16x16x4f64 on four 16-row blocks; in the step's MFMA j, lane
// group g = lane / 16 supplies k = k0 + 4 g + j, so a lane reads four
// consecutive k of one row of op(A) (or column of op(B)), which is what makes
// the transposed operand's reads contiguous.  The next step's operands are
// loaded before the current step's MFMAs.
template <int OPA, int OPB>
__device__ inline void narrow_load(const GemmTask& T, const double* __restrict__ A,
                                   const double* __restrict__ B, const int* row, int c, int k0,
                                   int klim, bool mask, double (&a)[4][4], double (&b)[4]) {
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int k = k0 + 4 * g + j;
        if (mask) k = k < klim ? k : k0;  // a valid address; zeroed at use
        const int cc = mask ? (c < T.N ? c : 0) : c;
        b[j] = (OPB == 0) ? B[k + (i64)cc * T.ldb] : B[cc + (i64)k * T.ldb];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const int r = mask ? (row[rb] < T.M ? row[rb] : 0) : row[rb];
            a[rb][j] = (OPA == 0) ? A[r + (i64)k * T.lda] : A[k + (i64)r * T.lda];
        }
    }
}

template <int OPA, int OPB, int TRIA>
__device__ inline void gemm_narrow(const GemmTask& T, const double* __restrict__ A,
                                   const double* __restrict__ B, double* __restrict__ C,
                                   int row0, int kbase, int kend, double* Pp,
                                   double (*red)[TILE][NARROW_N + 1]) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, c = lane & 15;
    // a triangular op(A) (TRIA 1: zero for k > i; 2: zero for k < i) limits K
    if (TRIA == 1) kend = min(kend, row0 + TILE);
    if (TRIA == 2) kbase = max(kbase, row0);
    const int q = ((max(kend - kbase, 0) + 63) / 64) * 16;  // each wave's share, a multiple of 16
    const int k0w = min(kend, kbase + w * q), k1w = max(k0w, min(kend, k0w + q));
    int row[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) row[rb] = row0 + rb * 16 + c;
    // steps [kc0, kc1) need no masks: full rows and columns, whole steps, and
    // clear of the diagonal of a triangular op(A)
    int kc0 = k0w, kc1 = k0w + ((k1w - k0w) & ~15);
    if (TRIA == 2 && row0 + TILE - 1 > k0w) kc0 = k0w + ((row0 + TILE - 1 - k0w + 15) & ~15);
    if (TRIA == 1) kc1 = min(kc1, row0 + 1 > k0w ? k0w + ((row0 + 1 - k0w) & ~15) : k0w);
    if (!(row0 + TILE <= T.M && T.N == NARROW_N)) kc0 = kc1 = k0w;
    kc0 = min(kc0, k1w);
    kc1 = max(kc1, kc0);
    dbl4 acc[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) acc[rb] = (dbl4){0.0, 0.0, 0.0, 0.0};
    double a[4][4], b[4], na[4][4], nb[4];
    auto masked_step = [&](int k0) {
        narrow_load<OPA, OPB>(T, A, B, row, c, k0, k1w, true, a, b);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * g + j;
            const bool kok = k < k1w;
            const double bj = (kok && c < T.N) ? b[j] : 0.0;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                const bool on = kok && row[rb] < T.M && (TRIA != 1 || k <= row[rb]) &&
                                (TRIA != 2 || k >= row[rb]);
                acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(on ? a[rb][j] : 0.0, bj, acc[rb], 0, 0, 0);
            }
        }
    };
    for (int k0 = k0w; k0 < kc0; k0 += 16) masked_step(k0);
    if (kc0 < kc1) narrow_load<OPA, OPB>(T, A, B, row, c, kc0, k1w, false, a, b);
    for (int k0 = kc0; k0 < kc1; k0 += 16) {
        const bool more = k0 + 16 < kc1;
        if (more) narrow_load<OPA, OPB>(T, A, B, row, c, k0 + 16, k1w, false, na, nb);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
                acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb][j], b[j], acc[rb], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                b[j] = nb[j];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) a[rb][j] = na[rb][j];
            }
        }
    }
    for (int k0 = kc1; k0 < k1w; k0 += 16) masked_step(k0);
    // accumulator lane l, element v: row 16 rb + l / 16 + 4 v, column l % 16
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[w][rb * 16 + g + 4 * v][c] = acc[rb][v];
    __syncthreads();
    const double alpha = T.alpha, beta = Pp ? 0.0 : T.beta;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = tid + u * TPB, i = e & 63, j = e >> 6;
        const int gi = row0 + i;
        if (gi >= T.M || j >= T.N) continue;
        const double s = (red[0][i][j] + red[1][i][j]) + (red[2][i][j] + red[3][i][j]);
        if (Pp) { Pp[gi + (i64)j * T.M] = alpha * s; continue; }
        double* cp = C + gi + (i64)j * T.ldc;
        *cp = alpha * s + (beta != 0.0 ? beta * (*cp) : 0.0);
    }
}

__global__ __launch_bounds__(TPB) void k_gemm_narrow(const GemmTile* __restrict__ tiles,
                                                     const GemmTask* __restrict__ tasks, BufTab bt) {
    __shared__ double red[4][TILE][NARROW_N + 1];
    if (gated_off(bt)) return;
    const GemmTile tl = tiles[blockIdx.x];
    const GemmTask T = tasks[tl.task];
    const double* A = buf_sel(bt, T.bufA) + T.offA;
    const double* B = buf_sel(bt, T.bufB) + T.offB;
    double* C = buf_sel(bt, T.bufC) + T.offC;
    const int row0 = tl.tm * TILE;
    int kbase = 0, kend = T.K;
    double* Pp = nullptr;
    if (T.ksplit > 1) {
        kbase = tl.ks * T.kchunk;
        kend = min(T.K, kbase + T.kchunk);
        Pp = bt.p[4] + T.offP + (i64)tl.ks * T.M * T.N;
    }
    switch (T.triA * 4 + T.opA * 2 + T.opB) {
    case 0: gemm_narrow<0, 0, 0>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    case 1: gemm_narrow<0, 1, 0>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    case 2: gemm_narrow<1, 0, 0>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    case 3: gemm_narrow<1, 1, 0>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    case 4: gemm_narrow<0, 0, 1>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    case 6: gemm_narrow<1, 0, 1>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    case 8: gemm_narrow<0, 0, 2>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    case 10: gemm_narrow<1, 0, 2>(T, A, B, C, row0, kbase, kend, Pp, red); break;
    default: break;  // not a narrow task (the planner sends none)
    }
}

// C = beta*C + sum_ks partial_ks over one 64 x 64 tile of a split-K task, in
// slice order (deterministic), honouring the task's triangular output mode.
__global__ __launch_bounds__(TPB) void k_splitk_reduce(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks, BufTab bt) {
    // tiles[i / 8] is the output tile; workgroup i % 8 reduces its 8 columns
    // i % 8 * 8 .. +7 (slices summed in slice order)
    if (gated_off(bt)) return;
    const GemmTile tl = tiles[blockIdx.x >> 3];
    const int q = blockIdx.x & 7;
    const GemmTask T = tasks[tl.task];
    double* C = bt.p[T.bufC] + T.offC;
    const double* Pp = bt.p[4] + T.offP;
    const i64 mn = (i64)T.M * T.N;
    for (int e = threadIdx.x; e < TILE * 8; e += TPB) {
        const int i = tl.tm * TILE + (e & 63), j = tl.tn * TILE + q * 8 + (e >> 6);
        if (i >= T.M || j >= T.N) continue;
        double* c = C + i + (i64)j * T.ldc;
        if (T.tri != 0 && i < j) {
            if (T.tri == 2) *c = 0.0;
            continue;
        }
        double s = 0.0;
        for (int ks = 0; ks < T.ksplit; ++ks) s += Pp[ks * mn + i + (i64)j * T.M];
        *c = (T.beta != 0.0 ? T.beta * (*c) : 0.0) + s;
    }
}

// ---------------------------------------------------------------------------
// Tile-DAG factorization of one level's large fronts (the dense part of
// cholesky.c:55-72 for every large front of the level, in one persistent
// launch).  Each front is cut into 64 x 64 tiles (DagFront); the tasks are
//   POTRF(k):      factor diagonal tile (k, k) in place and write the inverse
//                  of its factor (column-major, 64 x 64) to the front's
//                  inverse area, block k;
//   TRSM(i, k):    L(i, k) = A(i, k) L(k, k)^{-T}, a product with that inverse;
//   UPDATE(i,j,k): A(i, j) -= L(i, k) L(j, k)^T (lower part of diagonal tiles;
//                  tiles right of the supernode columns are the update matrix).
// A tile's counter counts the operations finished on it: UPDATE k = 0, 1, ...
// in order, then its POTRF or TRSM.  A task waits until its tile's counter
// equals its k (the updates before it are done; the accumulation order is
// fixed, so results do not depend on timing) and its operand tiles' counters
// equal k + 1.  Workgroups take tasks from a queue in a topological order
// (the host lists them in the start order of a simulated list schedule that
// favours the critical path), so every task waited for has been taken by a
// running workgroup: the waits end without any assumption on how many
// workgroups are resident.  Each wait is also bounded (spin_limit polls); a
// wait that runs out records an internal error in status[1] and goes on.
// Tiles cross workgroups (and the eight XCDs' L2 caches, which are not
// coherent with each other) as agent-scope atomic loads and stores of their
// elements, which bypass those caches; the counter updates follow the
// stores, and the tile loads the polls, by waits on the memory counters.

// X = L^{-1} (64 x 64) from L in LDS (lower, zero above) and rdiag = 1/diag:
// recursive doubling as in block_trtri_rd, with the temporary W = L21 X11 of
// each level kept in the X21 block it is about to be replaced by, so only
// two 64 x 65 arrays are needed.
__device__ inline void trtri_lean(const double (*Ls)[NB + 1], const double* rdiag,
                                  double (*X)[NB + 1]) {
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += TPB) X[e / NB][e % NB] = 0.0;
    __syncthreads();
    if (tid < NB / 4) {
        const int b = 4 * tid;
        const double y0 = rdiag[b], y1 = rdiag[b + 1], y2 = rdiag[b + 2], y3 = rdiag[b + 3];
        const double l10 = Ls[b + 1][b], l20 = Ls[b + 2][b], l30 = Ls[b + 3][b];
        const double l21 = Ls[b + 2][b + 1], l31 = Ls[b + 3][b + 1], l32 = Ls[b + 3][b + 2];
        const double x10 = -y1 * l10 * y0;
        const double x21 = -y2 * l21 * y1;
        const double x20 = -y2 * (l20 * y0 + l21 * x10);
        const double x32 = -y3 * l32 * y2;
        const double x31 = -y3 * (l31 * y1 + l32 * x21);
        const double x30 = -y3 * (l30 * y0 + l31 * x10 + l32 * x20);
        X[b][b] = y0; X[b + 1][b + 1] = y1; X[b + 2][b + 2] = y2; X[b + 3][b + 3] = y3;
        X[b + 1][b] = x10; X[b + 2][b] = x20; X[b + 3][b] = x30;
        X[b + 2][b + 1] = x21; X[b + 3][b + 1] = x31; X[b + 3][b + 2] = x32;
    }
    __syncthreads();
#pragma unroll 1
    for (int bs = 4; bs < NB; bs *= 2) {
        const int nout = (NB / (2 * bs)) * bs * bs;  // <= 4 * TPB
        // W = L21 X11 into the X21 block (read by no thread of this phase)
        for (int e = tid; e < nout; e += TPB) {
            const int pr = e / (bs * bs), i = (e / bs) % bs, j = e % bs;
            const int a0 = pr * 2 * bs;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int k = j;
            for (; k + 3 < bs; k += 4) {
                s0 += Ls[a0 + bs + i][a0 + k] * X[a0 + k][a0 + j];
                s1 += Ls[a0 + bs + i][a0 + k + 1] * X[a0 + k + 1][a0 + j];
                s2 += Ls[a0 + bs + i][a0 + k + 2] * X[a0 + k + 2][a0 + j];
                s3 += Ls[a0 + bs + i][a0 + k + 3] * X[a0 + k + 3][a0 + j];
            }
            for (; k < bs; ++k) s0 += Ls[a0 + bs + i][a0 + k] * X[a0 + k][a0 + j];
            X[a0 + bs + i][a0 + j] = (s0 + s1) + (s2 + s3);
        }
        __syncthreads();
        // X21 = -X22 W, formed in registers and stored once every thread has read W
        double r[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int e = tid + m * TPB;
            r[m] = 0.0;
            if (e >= nout) continue;
            const int pr = e / (bs * bs), i = (e / bs) % bs, j = e % bs;
            const int a0 = pr * 2 * bs;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int k = 0;
            for (; k + 3 <= i; k += 4) {
                s0 += X[a0 + bs + i][a0 + bs + k] * X[a0 + bs + k][a0 + j];
                s1 += X[a0 + bs + i][a0 + bs + k + 1] * X[a0 + bs + k + 1][a0 + j];
                s2 += X[a0 + bs + i][a0 + bs + k + 2] * X[a0 + bs + k + 2][a0 + j];
                s3 += X[a0 + bs + i][a0 + bs + k + 3] * X[a0 + bs + k + 3][a0 + j];
            }
            for (; k <= i; ++k) s0 += X[a0 + bs + i][a0 + bs + k] * X[a0 + bs + k][a0 + j];
            r[m] = -((s0 + s1) + (s2 + s3));
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int e = tid + m * TPB;
            if (e >= nout) continue;
            const int pr = e / (bs * bs), i = (e / bs) % bs, j = e % bs;
            const int a0 = pr * 2 * bs;
            X[a0 + bs + i][a0 + j] = r[m];
        }
        __syncthreads();
    }
}

__device__ inline int dag_row0(const DagFront& F, int i) {
    return i < F.P ? NB * i : F.nn + NB * (i - F.P);
}
__device__ inline int dag_rows(const DagFront& F, int i) {
    const int r0 = dag_row0(F, i);
    return min(NB, (i < F.P ? F.nn : F.nn + F.na) - r0);
}

// Ordering without cache maintenance (the tiles themselves are read and
// written with coherent accesses, gld<true> / gst<true>): wait until this
// thread's memory operations have completed, so that a counter update after
// the next barrier follows the stores, and the tile loads after a poll
// follow it.  Agent-scope fences would write back and invalidate the L2
// caches on every task.
__device__ inline void dag_order_after_stores() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
}
__device__ inline void dag_order_after_loads() {
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Polls counter c until it reaches target, at most `limit` times; gives up at
// once when another wait has already run out (err != 0), so that a launch
// with a failed wait drains quickly.
__device__ inline bool dag_wait(int* c, int target, int limit, unsigned long long* err) {
    int polls = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++polls > limit) return false;
        if ((polls & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

// POTRF task of k_chol_dag (kept out of line: inlined into the task loop its
// unrolled pivot steps would take the registers of a second workgroup per CU)
// The LDS of k_chol_dag.  POTRF: L and its inverse (64 x 65 each), the pivot
// exchange arrays and 1/diag; TRSM and UPDATE: the GEMM's two 32 x 66
// staging arrays at the start.  (At namespace scope, so that the out-of-line
// POTRF addresses it as LDS.)
__shared__ __attribute__((aligned(16))) double dag_smem[DAG_LDS_DOUBLES];

__device__ __noinline__ void dag_potrf(const DagFront F, int k, double* blkval, double* dinv,
                                       unsigned long long* status) {
    double (*Ls)[NB + 1] = reinterpret_cast<double (*)[NB + 1]>(dag_smem);
    double (*X)[NB + 1] = reinterpret_cast<double (*)[NB + 1]>(dag_smem + NB * (NB + 1));
    double (*Pp)[5] = reinterpret_cast<double (*)[5]>(dag_smem + 2 * NB * (NB + 1));
    double (*Qp)[5] = reinterpret_cast<double (*)[5]>(dag_smem + 2 * NB * (NB + 1) + NB * 5);
    double* rdiag = dag_smem + 2 * NB * (NB + 1) + 2 * NB * 5;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int nj = F.nn + F.na;
    double* blk = blkval + F.blkptr;
    const int c0 = NB * k, kw = min(NB, F.nn - c0);
    double* dg = blk + c0 + (i64)c0 * nj;
    double a[DQ];
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int jj = w + 4 * q;
        a[q] = (l < kw && jj < kw && l >= jj) ? gld<true>(dg + l + (i64)jj * nj) : (l == jj ? 1.0 : 0.0);
    }
    int bad = 0;
    if (!block_potrf4(a, Pp, Qp, rdiag, w, l, &bad)) {
        if (tid == 0) chp_record_failure(status, F.key, c0 + bad + 1);
    } else {
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int jj = w + 4 * q;
            Ls[l][jj] = (l >= jj) ? a[q] : 0.0;
            if (l < kw && jj < kw && l >= jj) gst<true>(dg + l + (i64)jj * nj, a[q]);
        }
        __syncthreads();
        trtri_lean(Ls, rdiag, X);
        double* inv = dinv + F.dinvoff + (i64)k * NB * NB;
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int c = w + 4 * q;  // column c, row l of the inverse
            gst<true>(inv + l + c * NB, (l < kw && c < kw) ? X[l][c] : 0.0);
        }
    }
}

__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_chol_dag(const int4* __restrict__ list,
                                                  double* __restrict__ blkval,
                                                  double* __restrict__ upd,
                                                  double* __restrict__ dinv, int* cnt,
                                                  unsigned long long* status, int spin_limit,
                                                  i64* trace) {
    __shared__ int s_task;
    double (*As)[TILE + 2] = reinterpret_cast<double (*)[TILE + 2]>(dag_smem);
    double (*Bs)[TILE + 2] = reinterpret_cast<double (*)[TILE + 2]>(dag_smem + BK * (TILE + 2));

    const int4 h = list[0];
    const DagFront* fronts = reinterpret_cast<const DagFront*>(list + 1);
    const DagTask* tasks = reinterpret_cast<const DagTask*>(fronts + h.y);
    int* qhead = cnt + h.z;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    // Thread 0 takes the first task here and the next one together with the
    // counter update at the end of each iteration, one thread-0 block per
    // iteration between two barriers: two thread-0 blocks on either side of
    // the back edge would be merged by the compiler into a path on which lane
    // 0 leaves the loop while the other lanes of its wave run the loop body
    // again with a stale task index.  The index itself is read as a
    // wave-uniform value, so the loop exit is uniform.
    if (tid == 0) s_task = atomicAdd(qhead, 1);
    __syncthreads();
    i64 t_grab = 0, t_ready = 0;
    for (;;) {
        const int t = __builtin_amdgcn_readfirstlane(s_task);
        if (t >= h.x) break;
        if (trace && tid == 0) t_grab = wall_clock64();
        const DagTask T = tasks[t];
        const DagFront F = fronts[T.f];
        const int k = T.kt & 0xffff, type = T.kt >> 16, i = T.i, j = T.j;
        int* ctile = cnt + F.cbase + i * (i + 1) / 2 + j;
        if (tid == 0) {
            unsigned long long* err = status + 1;
            bool ok = dag_wait(ctile, k, spin_limit, err);
            if (type == DAG_TRSM) {
                ok = ok && dag_wait(cnt + F.cbase + k * (k + 1) / 2 + k, k + 1, spin_limit, err);
            } else if (type == DAG_UPDATE) {
                ok = ok && dag_wait(cnt + F.cbase + i * (i + 1) / 2 + k, k + 1, spin_limit, err);
                if (j != i)
                    ok = ok && dag_wait(cnt + F.cbase + j * (j + 1) / 2 + k, k + 1, spin_limit, err);
            }
            if (!ok) atomicOr(status + 1, 1ULL);
            dag_order_after_loads();
            if (trace) t_ready = wall_clock64();
        }
        __syncthreads();
        const int nj = F.nn + F.na;
        double* blk = blkval + F.blkptr;
        const int c0 = NB * k, kw = min(NB, F.nn - c0);
        if (type == DAG_POTRF) {
            dag_potrf(F, k, blkval, dinv, status);
        } else {
            const int ri = dag_row0(F, i);
            GemmTask G;
            G.offA = G.offB = G.offC = G.offP = 0;
            G.ksplit = 1; G.kchunk = 0;
            G.bufA = G.bufB = G.bufC = 0;
            G.opA = 0; G.opB = 1; G.triA = 0; G.pad = 0;
            G.lda = nj;
            const double* Ap = blk + ri + (i64)c0 * nj;
            const double* Bp;
            double* Cp;
            if (type == DAG_TRSM) {  // in place: the tile is read in full before it is written
                Bp = dinv + F.dinvoff + (i64)k * NB * NB;
                G.ldb = NB;
                Cp = blk + ri + (i64)c0 * nj;
                G.ldc = nj;
                G.M = dag_rows(F, i); G.N = kw; G.K = kw;
                G.alpha = 1.0; G.beta = 0.0; G.tri = 0;
            } else {
                const int rj = dag_row0(F, j);
                Bp = blk + rj + (i64)c0 * nj;
                G.ldb = nj;
                if (j < F.P) {
                    Cp = blk + ri + (i64)rj * nj;
                    G.ldc = nj;
                } else {
                    Cp = upd + F.updptr + (ri - F.nn) + (i64)(rj - F.nn) * F.na;
                    G.ldc = F.na;
                }
                G.M = dag_rows(F, i); G.N = dag_rows(F, j); G.K = kw;
                G.alpha = -1.0; G.beta = 1.0; G.tri = (i == j) ? 1 : 0;
            }
            if (G.M == TILE && G.N == TILE && (G.K % BK) == 0)
                gemm_tile<0, 1, false, true>(G, Ap, Bp, Cp, 0, 0, 0, G.K, nullptr, As, Bs);
            else
                gemm_tile<0, 1, true, true>(G, Ap, Bp, Cp, 0, 0, 0, G.K, nullptr, As, Bs);
        }
        dag_order_after_stores();
        __syncthreads();
        if (tid == 0) {
            __hip_atomic_fetch_add(ctile, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (trace) {  // per task: kind, grab, dependencies met, done (100 MHz clock)
                i64* r = trace + 4 * ((i64)h.w + t);
                r[0] = (i64)type | ((i64)k << 8) | ((i64)i << 24) | ((i64)j << 40) | ((i64)blockIdx.x << 52);
                r[1] = t_grab; r[2] = t_ready; r[3] = wall_clock64();
            }
            s_task = atomicAdd(qhead, 1);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(TPB) void k_fill_int(int* p, i64 n, int v) {
    for (i64 e = (i64)blockIdx.x * TPB + threadIdx.x; e < n; e += (i64)gridDim.x * TPB) p[e] = v;
}
