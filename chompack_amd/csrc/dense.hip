// update_factor on the GPU: one 1024-thread workgroup per child update
// (update_factor.c:3-108).  The child's update V = F[r, r] is produced in
// factored form V = Li^T Li from the parent's frontal matrix whose (2,2)
// block holds the parent's factor Lk (V_k = Lk^T Lk):
//   * A = F[r1, r1], B = F[r2, r1] are gathered into Li;
//   * C = Lk[:, r2 - nn] is reduced to lower-triangular form by Householder
//     reflections from its last column (the LAPACK dlarfg convention:
//     beta = -sign(alpha) ||[alpha; x]||), giving Li22.  Pivots can vanish
//     structurally; a pivot below PIVOT_NOISE ||x|| is taken as +0, the
//     exact-arithmetic value, so the signs of Li22's rows do not follow
//     rounding noise.  The reflections go in blocks of UF_NBH: the block's
//     own columns take them one by one, the columns left of it all at once
//     in compact WY form;
//   * Li21 solves Li22^T Li21 = B;
//   * Li11 is the reversed Cholesky factor of A - Li21^T Li21
//     (A = Li11^T Li11, eliminated from the last row up).
// All work is in global memory (workspace C in the scratch buffer); the
// sequential steps are separated by workgroup barriers.  The steps are
// latency-bound chains (a dependent FP64 operation costs tens of cycles on
// one wave), so the workgroup has 16 waves.  The planner hands the solve
// and the (1,1) block of tasks with long separators (and the reversed
// Cholesky of large (1,1) blocks) to device-wide launches (UFTask flags).
#include "common.h"
#include "kernels.h"

#define UF_TPB 1024
#define UF_WAVES (UF_TPB / 64)
#define UF_ROWMAX 4096  // rows staged in LDS (longer rows are read from global memory)
#define PIVOT_NOISE 1e-13
#define UF_NBH 8       // reflections per block
#define UF_VMAX 1024   // rows of a block's V held in LDS

// v of lane `lane` (uniform across the wave)
__device__ inline double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(UF_TPB) void k_update_factor(const UFTask* __restrict__ tasks,
                                                       BufTab bt, const int* __restrict__ relidx,
                                                       unsigned long long* status) {
    const UFTask t = tasks[blockIdx.x];
    const int nn = t.nn, na = t.na, N = t.N, nj = nn + na;
    const int* r = relidx + t.relpos;
    const double* F = buf_sel(bt, t.fbuf) + t.foff;
    const double* Lk = bt.p[1] + t.lkoff;
    double* Li = bt.p[1] + t.outoff;
    double* C = bt.p[2] + t.woff;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ int s_n1;
    __shared__ double red[UF_WAVES];
    __shared__ double s_tau, s_beta, s_scal, s_piv;
    __shared__ double rowbuf[UF_ROWMAX];
    __shared__ double vbuf[UF_NBH][UF_VMAX];
    __shared__ double gram[UF_NBH][UF_NBH], tmat[UF_NBH][UF_NBH], s_btau[UF_NBH];
    __shared__ int s_bjj[UF_NBH];
    if (tid == 0) {
        int c = 0;
        while (c < N && r[c] < nn) ++c;
        s_n1 = c;
    }
    __syncthreads();
    const int N1 = s_n1, N2 = N - N1;
    const bool ext = t.ext_reduce;  // copies and reflections done by k_uf_*
    // element loops: one column per wave, its rows over the lanes
    if (!ext)
    for (int j = wv; j < N; j += UF_WAVES) {
        const i64 fj = (i64)r[j] * nj;
        for (int i = lane; i < N; i += 64)
            Li[i + (i64)j * N] = (j < N1 && i >= j) ? F[r[i] + fj] : 0.0;
    }
    if (!ext)
    for (int j = wv; j < N2; j += UF_WAVES) {
        const int cj = r[N1 + j] - nn;
        for (int i = lane; i < na; i += 64)
            C[i + (i64)j * na] = (i >= cj) ? Lk[i + (i64)cj * na] : 0.0;
    }
    __syncthreads();
    if (N2 > 0) {
        // Blocks of UF_NBH consecutive reflections.  Within a block each
        // reflection is formed and applied to the block's own columns in
        // turn; the columns left of the block then take the whole block at
        // once as H_{nb-1} ... H_0 = I - V T^T V^T (T upper triangular, the
        // LAPACK dlarft forward recurrence), one wave per column.  A block
        // whose bands span more than UF_VMAX rows applies its reflections one
        // by one instead.
        for (int js0 = 0; js0 < (ext ? 0 : N2); js0 += UF_NBH) {
            const int nb = min(UF_NBH, N2 - js0);
            const int colmin = N2 - js0 - nb;  // leftmost column of the block
            for (int b = 0; b < nb; ++b) {
                const int js = js0 + b;
                const int col = N2 - 1 - js;
                const int jj = r[N - 1 - js] - nn;
                if (tid == 0) {
                    s_btau[b] = 0.0;
                    s_bjj[b] = jj;
                }
                if (na - jj - (js + 1) == 0) continue;
                const int piv = na - 1 - js;
                double* cc = C + (i64)col * na;
                double s = 0.0;
                for (int i = jj + tid; i < piv; i += UF_TPB) s += cc[i] * cc[i];
                s = wave_sum(s);
                if (lane == 0) red[wv] = s;
                __syncthreads();
                if (tid == 0) {
                    double ss = 0.0;
                    for (int w = 0; w < UF_WAVES; ++w) ss += red[w];
                    const double xn = sqrt(ss);
                    double alpha = cc[piv];
                    if (fabs(alpha) <= PIVOT_NOISE * xn) alpha = 0.0;
                    if (xn == 0.0) {
                        s_tau = 0.0;
                    } else {
                        const double beta = -copysign(hypot(alpha, xn), alpha);
                        s_tau = (beta - alpha) / beta;
                        s_beta = beta;
                        s_scal = 1.0 / (alpha - beta);
                    }
                    s_btau[b] = s_tau;
                }
                __syncthreads();
                const double tau = s_tau;
                if (tau == 0.0) continue;
                const double scal = s_scal;
                for (int i = jj + tid; i < piv; i += UF_TPB) cc[i] *= scal;
                if (tid == 0) cc[piv] = s_beta;
                __syncthreads();
                // H = I - tau v v^T, v = [cc[jj:piv]; 1], on rows jj..piv of
                // the block's columns colmin..col-1.  A short band (most
                // reflections of a nearly triangular C zero one or two
                // entries) takes one thread per column, a long one one wave
                // per column.
                if (piv - jj < 32) {
                    for (int c = colmin + tid; c < col; c += UF_TPB) {
                        double* cd = C + (i64)c * na;
                        double d = cd[piv];
                        for (int i = jj; i < piv; ++i) d += cc[i] * cd[i];
                        const double f = tau * d;
                        for (int i = jj; i < piv; ++i) cd[i] -= f * cc[i];
                        cd[piv] -= f;
                    }
                } else {
                    for (int c = colmin + wv; c < col; c += UF_WAVES) {
                        double* cd = C + (i64)c * na;
                        double d0 = 0.0, d1 = 0.0;
                        int i = jj + lane;
                        for (; i + 64 < piv; i += 128) {
                            d0 += cc[i] * cd[i];
                            d1 += cc[i + 64] * cd[i + 64];
                        }
                        if (i < piv) d0 += cc[i] * cd[i];
                        const double d = wave_sum(d0 + d1) + cd[piv];
                        const double f = tau * d;
                        for (int i = jj + lane; i < piv; i += 64) cd[i] -= f * cc[i];
                        if (lane == 0) cd[piv] -= f;
                    }
                }
                __syncthreads();
            }
            __syncthreads();
            if (colmin == 0) continue;
            const int hi = na - 1 - js0;  // the block's first pivot row
            int lo = hi;
            for (int b = 0; b < nb; ++b)
                if (s_btau[b] != 0.0) lo = min(lo, s_bjj[b]);
            if (lo == hi) continue;  // no reflection in the block
            const int len = hi - lo + 1;
            if (len <= UF_VMAX) {
                // V over rows lo..hi: v_b zero outside its band, 1 at its pivot
                for (int e = tid; e < UF_NBH * UF_VMAX; e += UF_TPB) {
                    const int b = e / UF_VMAX, i = e % UF_VMAX, row = lo + i;
                    double v = 0.0;
                    if (b < nb && i < len && s_btau[b] != 0.0) {
                        const int pb = na - 1 - (js0 + b);
                        if (row == pb) v = 1.0;
                        else if (row >= s_bjj[b] && row < pb) v = C[row + (i64)(N2 - 1 - js0 - b) * na];
                    }
                    vbuf[b][i] = v;
                }
                __syncthreads();
                for (int q = wv; q < UF_NBH * UF_NBH; q += UF_WAVES) {
                    const int a = q / UF_NBH, b = q % UF_NBH;
                    if (a >= b) continue;
                    double g = 0.0;
                    for (int i = lane; i < len; i += 64) g += vbuf[a][i] * vbuf[b][i];
                    g = wave_sum(g);
                    if (lane == 0) gram[a][b] = g;
                }
                __syncthreads();
                if (tid == 0) {
                    for (int b = 0; b < UF_NBH; ++b) {
                        const double tb = b < nb ? s_btau[b] : 0.0;
                        for (int a = 0; a < b; ++a) {
                            double g = 0.0;
                            for (int k = a; k < b; ++k) g += tmat[a][k] * gram[k][b];
                            tmat[a][b] = -tb * g;
                        }
                        tmat[b][b] = tb;
                        for (int a = b + 1; a < UF_NBH; ++a) tmat[a][b] = 0.0;
                    }
                }
                __syncthreads();
                for (int c = wv; c < colmin; c += 2 * UF_WAVES) {
                    // w = V^T c, y = T^T w, c -= V y for columns c and c2 at
                    // once (twice the loads in flight); row lo + i in lane i % 64
                    const int c2 = c + UF_WAVES;
                    const bool two = c2 < colmin;
                    double* ca = C + (i64)c * na + lo;
                    double* cb = C + (i64)(two ? c2 : c) * na + lo;
                    double wa[UF_NBH], wb[UF_NBH];
#pragma unroll
                    for (int b = 0; b < UF_NBH; ++b) wa[b] = wb[b] = 0.0;
                    for (int i = lane; i < len; i += 64) {
                        const double xa = ca[i], xb = cb[i];
#pragma unroll
                        for (int b = 0; b < UF_NBH; ++b) {
                            const double v = vbuf[b][i];
                            wa[b] += v * xa;
                            wb[b] += v * xb;
                        }
                    }
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
                        for (int b = 0; b < UF_NBH; ++b) {
                            wa[b] += __shfl_xor(wa[b], o, 64);
                            wb[b] += __shfl_xor(wb[b], o, 64);
                        }
                    // y = T^T w: lane b < UF_NBH forms y_b from column b of T
                    double ya = 0.0, yb = 0.0;
#pragma unroll
                    for (int a = 0; a < UF_NBH; ++a) {
                        const double ta = tmat[a][lane & (UF_NBH - 1)];
                        ya += ta * wa[a];
                        yb += ta * wb[a];
                    }
#pragma unroll
                    for (int b = 0; b < UF_NBH; ++b) {
                        wa[b] = readlane_d(ya, b);
                        wb[b] = readlane_d(yb, b);
                    }
                    for (int i = lane; i < len; i += 64) {
                        double xa = ca[i], xb = cb[i];
#pragma unroll
                        for (int b = 0; b < UF_NBH; ++b) {
                            const double v = vbuf[b][i];
                            xa -= v * wa[b];
                            xb -= v * wb[b];
                        }
                        ca[i] = xa;
                        if (two) cb[i] = xb;
                    }
                }
            } else {
                // one reflection after another; row i of a column always in
                // lane i % 64, so no lane reads what another lane wrote
                for (int c = wv; c < colmin; c += UF_WAVES) {
                    double* cd = C + (i64)c * na;
                    for (int b = 0; b < nb; ++b) {
                        const double tau = s_btau[b];
                        if (tau == 0.0) continue;
                        const int jb = s_bjj[b], pb = na - 1 - (js0 + b);
                        const double* cc = C + (i64)(N2 - 1 - js0 - b) * na;
                        double d0 = 0.0;
                        for (int i = (jb & ~63) + lane; i < pb; i += 64)
                            if (i >= jb) d0 += cc[i] * cd[i];
                        if (lane == (pb & 63)) d0 += cd[pb];
                        const double f = tau * wave_sum(d0);
                        for (int i = (jb & ~63) + lane; i < pb; i += 64)
                            if (i >= jb) cd[i] -= f * cc[i];
                        if (lane == (pb & 63)) cd[pb] -= f;
                    }
                }
            }
            __syncthreads();
        }
        for (int j = wv; j < N2; j += UF_WAVES)
            for (int i = j + lane; i < N2; i += 64)
                Li[(N1 + i) + (i64)(N1 + j) * N] = C[(na - N2 + i) + (i64)j * na];
        if (t.defer_solve) return;  // Li21 = Li22^{-T} B and the rest follow as launches
        __syncthreads();
        // Li21 from Li22^T Li21 = B: row i of every column once rows > i are
        // done, then its contribution to rows < i
        const double* L22 = Li + N1 + (i64)N1 * N;
        double* X = Li + N1;  // column c at X + c N
        if (N1 < 32) {
            // few columns: each thread substitutes its own column
            for (int c = tid; c < N1; c += UF_TPB) {
                double* x = X + (i64)c * N;
                for (int i = N2 - 1; i >= 0; --i) {
                    double v = x[i];
                    for (int k = i + 1; k < N2; ++k) v -= L22[k + (i64)i * N] * x[k];
                    x[i] = v / L22[i + (i64)i * N];
                }
            }
            __syncthreads();
        } else
        for (int i = N2 - 1; i >= 0; --i) {
            const double dinv = 1.0 / L22[i + (i64)i * N];
            const bool st = i <= UF_ROWMAX;  // row i of L22 staged in LDS
            for (int c = tid; c < N1; c += UF_TPB) X[i + (i64)c * N] *= dinv;
            if (st)
                for (int k = tid; k < i; k += UF_TPB) rowbuf[k] = L22[i + (i64)k * N];
            __syncthreads();
            for (int c = wv; c < N1; c += UF_WAVES) {
                double* xc = X + (i64)c * N;
                const double xi = xc[i];
                for (int k = lane; k < i; k += 64)
                    xc[k] -= (st ? rowbuf[k] : L22[i + (i64)k * N]) * xi;
            }
            __syncthreads();
        }
        for (int j = wv; j < N1; j += UF_WAVES) {
            for (int i = j + lane; i < N1; i += 64) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                const double* xi = Li + N1 + (i64)i * N;
                const double* xj = Li + N1 + (i64)j * N;
                int k = 0;
                for (; k + 3 < N2; k += 4) {
                    s0 += xi[k] * xj[k];
                    s1 += xi[k + 1] * xj[k + 1];
                    s2 += xi[k + 2] * xj[k + 2];
                    s3 += xi[k + 3] * xj[k + 3];
                }
                for (; k < N2; ++k) s0 += xi[k] * xj[k];
                Li[i + (i64)j * N] -= (s0 + s1) + (s2 + s3);
            }
        }
        __syncthreads();
    }
    if (t.defer_rchol) return;  // a blocked potrf on the reversed block follows
    // reversed Cholesky of the (1,1) block: A = Lr^T Lr, last row first
    for (int p = N1 - 1; p >= 0; --p) {
        if (tid == 0) s_piv = Li[p + (i64)p * N];
        __syncthreads();
        const double d = s_piv;
        if (!(d > 0.0)) {
            if (tid == 0) chp_record_failure(status, t.key, N1 - p);
            return;
        }
        const double sd = sqrt(d), rs = 1.0 / sd;
        const bool st = p <= UF_ROWMAX;  // row p staged in LDS
        for (int j = tid; j < p; j += UF_TPB) {
            const double v = Li[p + (i64)j * N] * rs;
            Li[p + (i64)j * N] = v;
            if (st) rowbuf[j] = v;
        }
        if (tid == 0) Li[p + (i64)p * N] = sd;
        __syncthreads();
        // rank-1 update of the lower triangle of the leading p x p block
        for (int j = wv; j < p; j += UF_WAVES) {
            const double rj = st ? rowbuf[j] : Li[p + (i64)j * N];
            double* cj = Li + (i64)j * N;
            for (int i = j + lane; i < p; i += 64)
                cj[i] -= (st ? rowbuf[i] : Li[p + (i64)i * N]) * rj;
        }
        __syncthreads();
    }
}

// Barrier of the one-wave kernels below that orders LDS only: their global
// stores are read by later launches, never by the same one, and a full
// __syncthreads() would wait for each of them to complete.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Triangular solves of one right-hand side held across the lanes of a wave
// (lane l: row l), with L lower triangular of order n <= 64 at
// L(i, j) = T[r0 + i][r0 + j] in LDS and rdg = 1 / L(l, l) (0 for l >= n).
// Step i subtracts c_i(l) z_i from every lane, with z_i lane i's running
// value (final by then) broadcast by readlane and c_i(l) = L(i, l) / L(i, i)
// (L^T z = b, last row first; zero unless l < i) or L(l, i) / L(i, i)
// (L w = b; zero unless l > i); the solution is z rdg.  The coefficients of
// the next eight steps are read from LDS while the current eight run, so the
// dependent chain is a broadcast and an fma per step.
template <bool TRANS>
__device__ inline double wave_solve(const double (*T)[65], int r0, int n, double z, double rdg) {
    const int l = threadIdx.x & 63;
    const int nb = (n + 7) / 8;
    double cur[8], nxt[8];
    auto fetch = [&](int b, double (&v)[8]) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = 8 * b + u;
            const bool ok = i < n && (TRANS ? l < i : (l > i && l < n));
            v[u] = TRANS ? T[ok ? r0 + i : 0][ok ? r0 + l : 0] : T[ok ? r0 + l : 0][ok ? r0 + i : 0];
        }
    };
    auto scale = [&](int b, double (&v)[8]) {  // times 1 / L(i, i), zero where unused
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = 8 * b + u;
            const bool ok = i < n && (TRANS ? l < i : (l > i && l < n));
            v[u] = v[u] * (ok ? readlane_d(rdg, i & 63) : 0.0);
        }
    };
#pragma unroll
    for (int bb = 0; bb < 8; ++bb) {
        const int b = TRANS ? 7 - bb : bb;  // block order of the sweep
        if (b < nb) {
            const bool first = TRANS ? (b == nb - 1) : (b == 0);
            if (first) fetch(b, cur);
            const int bn = TRANS ? b - 1 : b + 1;
            if (bn >= 0 && bn < nb) fetch(bn, nxt);
            scale(b, cur);
#pragma unroll
            for (int uu = 0; uu < 8; ++uu) {  // (steps i >= n subtract zero)
                const int u = TRANS ? 7 - uu : uu;
                z -= cur[u] * readlane_d(z, 8 * b + u);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) cur[u] = nxt[u];
        }
    }
    return z * rdg;
}

__device__ inline double wave_solve_lt(const double (*T)[65], int r0, int n, double z,
                                       double rdg) {
    return wave_solve<true>(T, r0, n, z, rdg);
}

__device__ inline double wave_solve_l(const double (*T)[65], int r0, int n, double w,
                                      double rdg) {
    return wave_solve<false>(T, r0, n, w, rdg);
}

// The same update for small tasks (N <= 64 and na <= 64): one wave, with C
// and Li held in LDS; lanes walk rows for the reductions and the copies and
// own a column each in the reflections, the solve and the rank-1 updates.
// The reflections and the (1,1) update follow k_update_factor's arithmetic
// (the thread-per-column path, the four partial sums); short bands sum
// their norm in every lane, and the solve runs down the columns with the
// rows across the lanes, multiplying by reciprocal pivots.
// Returns false after recording a failed reversed Cholesky.
__device__ __forceinline__ bool update_factor_wave(const UFTask t, BufTab bt, const int* __restrict__ relidx,
                                   unsigned long long* status, double (*Cs)[65],
                                   double (*Ls)[65]) {
    const int nn = t.nn, na = t.na, N = t.N, nj = nn + na;
    const int l = threadIdx.x;
    const int* r = relidx + t.relpos;
    const double* F = buf_sel(bt, t.fbuf) + t.foff;
    const double* Lk = bt.p[1] + t.lkoff;
    double* Li = bt.p[1] + t.outoff;
    const int rl = l < N ? r[l] : 0;  // lane j's r[j] is broadcast by readlane below
    const int N1 = __popcll(__ballot(l < N && rl < nn)), N2 = N - N1;
    wave_lds_sync();
    // eight columns at a time, loads first (from a valid address also where
    // the entry is zero), so that they are in flight together
    {
    for (int j0 = 0; j0 < N; j0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u;
            const bool ok = j < N1 && l < N && l >= j;
            v[u] = F[ok ? rl + (i64)__builtin_amdgcn_readlane(rl, j & 63) * nj : 0];
            v[u] = ok ? v[u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < N) Ls[l][j0 + u] = v[u];
    }
    for (int j0 = 0; j0 < N2; j0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u;
            const int cj = j < N2 ? __builtin_amdgcn_readlane(rl, (N1 + j) & 63) - nn : 0;
            const bool ok = j < N2 && l < na && l >= cj;
            v[u] = Lk[ok ? l + (i64)cj * na : 0];
            v[u] = ok ? v[u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < N2) Cs[l][j0 + u] = v[u];
    }
    }
    wave_lds_sync();
    // Reflection js has nothing to reduce when its band is empty, i.e. when
    // piv - jj = (na - 1 - js) - (r[N-1-js] - nn) is 0.  That difference
    // never decreases with js (piv drops by one, r by at least one), so the
    // empty reflections are a prefix: lane q tests reflection q, and the loop
    // starts after the leading run.
    const int rq = __shfl(rl, max(N - 1 - l, 0), 64);
    const unsigned long long emp = __ballot(l < N2 && na - (rq - nn) - (l + 1) == 0);
    const int js_first = ~emp ? __builtin_ctzll(~emp) : 64;
    for (int js = js_first; js < N2; ++js) {
        const int col = N2 - 1 - js;
        const int jj = __builtin_amdgcn_readlane(rl, N - 1 - js) - nn;
        if (na - jj - (js + 1) == 0) continue;
        const int piv = na - 1 - js;
        const double x = (l >= jj && l < piv) ? Cs[l][col] : 0.0;
        double ss = 0.0;
        if (piv - jj <= 8) {  // short band: every lane sums it (no reduction)
            for (int i = jj; i < piv; ++i) ss += Cs[i][col] * Cs[i][col];
        } else {
            ss = wave_sum(x * x);
        }
        const double xn = sqrt(ss);
        double alpha = Cs[piv][col];
        if (fabs(alpha) <= PIVOT_NOISE * xn) alpha = 0.0;
        if (xn == 0.0) continue;
        const double beta = -copysign(hypot(alpha, xn), alpha);
        const double tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        wave_lds_sync();  // every lane has read the pivot
        if (l >= jj && l < piv) Cs[l][col] = x * scal;
        if (l == piv) Cs[piv][col] = beta;
        wave_lds_sync();
        if (l < col) {
            double d = Cs[piv][l];
            for (int i = jj; i < piv; ++i) d += Cs[i][col] * Cs[i][l];
            const double f = tau * d;
            for (int i = jj; i < piv; ++i) Cs[i][l] -= f * Cs[i][col];
            Cs[piv][l] -= f;
        }
        wave_lds_sync();
    }
    for (int j0 = 0; j0 < N2; j0 += 8) {  // Li22 into place, reads before writes
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u;
            const bool ok = l >= j && l < N2 && j < N2;
            v[u] = Cs[ok ? na - N2 + l : 0][ok ? j : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u;
            if (l >= j && l < N2 && j < N2) Ls[N1 + l][N1 + j] = v[u];
        }
    }
    wave_lds_sync();
    // Li21 from Li22^T Li21 = B, column by column, lane l holding row l: once
    // x_i is known, rows l < i subtract Li22(i, l) x_i
    // (multiplied by the reciprocal pivots, computed off the dependent chain)
    const double rdg = l < N2 ? 1.0 / Ls[N1 + l][N1 + l] : 0.0;
    for (int c = 0; c < N1; ++c) {
        double z = l < N2 ? Ls[N1 + l][c] : 0.0;
        z = wave_solve_lt(Ls, N1, N2, z, rdg);
        if (l < N2) Ls[N1 + l][c] = z;
    }
    wave_lds_sync();
    // Li11 -= Li21^T Li21 (lower), row l
    if (l < N1 && N2 > 0) {
        for (int j = 0; j <= l; ++j) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int k = 0;
            for (; k + 3 < N2; k += 4) {
                s0 += Ls[N1 + k][l] * Ls[N1 + k][j];
                s1 += Ls[N1 + k + 1][l] * Ls[N1 + k + 1][j];
                s2 += Ls[N1 + k + 2][l] * Ls[N1 + k + 2][j];
                s3 += Ls[N1 + k + 3][l] * Ls[N1 + k + 3][j];
            }
            for (; k < N2; ++k) s0 += Ls[N1 + k][l] * Ls[N1 + k][j];
            Ls[l][j] -= (s0 + s1) + (s2 + s3);
        }
    }
    wave_lds_sync();
    // reversed Cholesky of the (1,1) block, last row first
    for (int p = N1 - 1; p >= 0; --p) {
        const double d = Ls[p][p];
        if (!(d > 0.0)) {
            if (l == 0) chp_record_failure(status, t.key, N1 - p);
            return false;
        }
        const double sd = sqrt(d), rs = 1.0 / sd;
        wave_lds_sync();  // every lane has read the pivot
        if (l < p) Ls[p][l] *= rs;
        if (l == p) Ls[p][p] = sd;
        wave_lds_sync();
        if (l < p) {
            const double ri = Ls[p][l];
            for (int j = 0; j <= l; ++j) Ls[l][j] -= ri * Ls[p][j];
        }
        wave_lds_sync();
    }
#pragma unroll 8
    for (int j = 0; j < N; ++j)
        if (l < N) Li[l + (i64)j * N] = Ls[l][j];
    wave_lds_sync();  // the LDS tiles are free again
    return true;
}

__global__ __launch_bounds__(64) void k_update_factor_small(const UFTask* __restrict__ tasks,
                                                            BufTab bt,
                                                            const int* __restrict__ relidx,
                                                            unsigned long long* status) {
    __shared__ double Cs[64][65];  // C(i, j)
    __shared__ double Ls[64][65];  // Li(i, j)
    const UFTask t = tasks[blockIdx.x];
    if (t.N > 64 || t.na > 64) {  // the planner sends only small tasks here
        if (threadIdx.x == 0) chp_record_failure(status, t.key, 1);
        return;
    }
    update_factor_wave(t, bt, relidx, status, Cs, Ls);
}
