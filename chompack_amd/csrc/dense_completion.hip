// Dense completions of a chordal partial matrix A (a cspmatrix, reordered):
//
//   psdcompletion (pybase/psdcompletion.py:4-70): the maximum-determinant
//     positive semidefinite completion X, dense n x n.  Supernodes are
//     visited in descending order; for supernode k with separator alpha,
//     columns nu and the rows eta (> nu, not in the clique) still unset,
//         X[eta, nu] = X[eta, alpha] X[alpha, alpha]^+ X[alpha, nu]
//     with the Cholesky factor of X[alpha, alpha], or its eigendecomposition
//     pseudo-inverse (eigenvalues <= tol * lambda_max dropped) when the
//     Cholesky factorization fails.
//   mrcompletion (pybase/mrcompletion.py:6-108): the minimum-rank completion
//     Y (n x r) with P(Y Y^T) = A.  Every clique matrix F_k = A[beta, beta]
//     is eigendecomposed (all cliques in one launch: F_k holds only entries
//     of A); then, from the roots down one etree level per launch, the rows
//     of the clique's own columns are set and rotated into the basis its
//     ancestors fixed on the separator rows (two singular value
//     decompositions per clique).
//
// The per-clique dense linear algebra runs in one workgroup: a right-looking
// Cholesky factorization, and cyclic Jacobi methods (two-sided for the
// symmetric eigendecomposition, one-sided for the singular value
// decomposition) in round-robin pair order, so that all rotations of a round
// act on disjoint rows and columns and the workgroup applies them at once.
// Work arrays live in a device arena (L2-resident for the clique sizes these
// routines are meant for); cliques are limited to DC_MAXN rows.
#include <hip/hip_runtime.h>

#include "common.h"
#include "../../include/chompack_amd.h"

#include <algorithm>
#include <vector>

namespace {

constexpr int DC_TPB = 256;
constexpr int DC_WAVES = DC_TPB / 64;
constexpr int DC_MAXN = 2048;            // largest clique / separator order
constexpr int DC_MAXPAIR = DC_MAXN / 2;
constexpr int DC_SWEEPS = 40;
constexpr double DC_JTOL = 1e-15;       // rotation skipped below JTOL * sqrt(|a_pp a_qq|)

// Pair pr of round r of the round-robin (circle) schedule on mp players
// (mp even): player 0 stays, players 1 .. mp-1 sit on a ring of odd length
// mp - 1 that turns by one each round, so every pair meets once in mp - 1
// rounds and the pairs of one round are disjoint.
__device__ inline void rr_pair(int pr, int r, int mp, int& p, int& q) {
    const int L = mp - 1;
    if (pr == 0) {
        p = 0;
        q = 1 + r % L;
    } else {
        p = 1 + (r + pr) % L;
        q = 1 + (r + L - pr) % L;
    }
}

// Jacobi rotation (c, s) that annihilates the off-diagonal entry of the
// symmetric 2 x 2 matrix [app apq; apq aqq] under J^T M J, J = [c s; -s c]
// (Golub and Van Loan, Algorithm 8.4.1).  Returns false (c = 1, s = 0) when
// apq is negligible.
__device__ inline bool jacobi_rot(double app, double aqq, double apq, double floor_, double& c,
                                  double& s) {
    c = 1.0;
    s = 0.0;
    const double a = fabs(apq);
    if (!(a > floor_) || !(a > DC_JTOL * sqrt(fabs(app) * fabs(aqq)))) return false;
    const double tau = (aqq - app) / (2.0 * apq);
    const double t = fabs(tau) > 1e150 ? 0.5 / tau
                                       : (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    c = 1.0 / sqrt(1.0 + t * t);
    s = t * c;
    return s != 0.0;
}

__device__ double wg_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int t = threadIdx.x;
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < DC_WAVES; ++w) s += red[w];
    __syncthreads();
    return s;
}

struct JacobiLds {
    double c[DC_MAXPAIR], s[DC_MAXPAIR];
    int p[DC_MAXPAIR], q[DC_MAXPAIR];
    double red[DC_WAVES];
    int flag;
};

// Two-sided cyclic Jacobi eigendecomposition of the symmetric m x m matrix A
// (column-major, ld m, both triangles): A := J^T A J, V := V J (V = I on
// entry) until every off-diagonal entry is negligible.  On exit the diagonal
// of A holds the eigenvalues and the columns of V the eigenvectors.
__device__ void wg_jacobi_evd(double* A, double* V, int m, JacobiLds& sh) {
    const int t = threadIdx.x;
    const i64 mm = (i64)m * m;
    double nrm = 0.0;
    for (i64 e = t; e < mm; e += DC_TPB) {
        V[e] = (e % m == e / m) ? 1.0 : 0.0;
        nrm += A[e] * A[e];
    }
    // entries below 1e-18 ||A||_F are left alone: they move eigenvalues by
    // less than that, far below the rank and pseudo-inverse thresholds
    const double floor_ = 1e-18 * sqrt(wg_sum(nrm, sh.red));
    if (m < 2) return;
    const int mp = m + (m & 1), np = mp / 2;
    for (int sweep = 0; sweep < DC_SWEEPS; ++sweep) {
        __syncthreads();
        if (t == 0) sh.flag = 0;
        __syncthreads();
        for (int r = 0; r < mp - 1; ++r) {
            for (int pr = t; pr < np; pr += DC_TPB) {
                int p, q;
                rr_pair(pr, r, mp, p, q);
                double c = 1.0, s = 0.0;
                if (p < m && q < m &&
                    jacobi_rot(A[p + (i64)p * m], A[q + (i64)q * m], A[p + (i64)q * m], floor_, c, s))
                    sh.flag = 1;
                sh.c[pr] = c;
                sh.s[pr] = s;
                sh.p[pr] = p;
                sh.q[pr] = q;
            }
            __syncthreads();
            // A := A J, V := V J on columns p, q of every pair
            for (i64 e = t; e < (i64)np * m; e += DC_TPB) {
                const int pr = (int)(e / m), i = (int)(e % m);
                const double s = sh.s[pr];
                if (s == 0.0) continue;
                const double c = sh.c[pr];
                double* ap = A + (i64)sh.p[pr] * m;
                double* aq = A + (i64)sh.q[pr] * m;
                double x = ap[i], y = aq[i];
                ap[i] = c * x - s * y;
                aq[i] = s * x + c * y;
                double* vp = V + (i64)sh.p[pr] * m;
                double* vq = V + (i64)sh.q[pr] * m;
                x = vp[i];
                y = vq[i];
                vp[i] = c * x - s * y;
                vq[i] = s * x + c * y;
            }
            __syncthreads();
            // A := J^T A on rows p, q; the annihilated pair is set to zero
            for (i64 e = t; e < (i64)np * m; e += DC_TPB) {
                const int pr = (int)(e / m), j = (int)(e % m);
                const double s = sh.s[pr];
                if (s == 0.0) continue;
                const double c = sh.c[pr];
                const int p = sh.p[pr], q = sh.q[pr];
                double* cj = A + (i64)j * m;
                const double x = cj[p], y = cj[q];
                cj[p] = j == q ? 0.0 : c * x - s * y;
                cj[q] = j == p ? 0.0 : s * x + c * y;
            }
            __syncthreads();
        }
        if (sh.flag == 0) break;
    }
    __syncthreads();
}

// One-sided (Hestenes) Jacobi on the columns of the rows x cols matrix B
// (column-major, ld rows): B := B W, W := W J (W = I on entry, cols x cols)
// until the columns of B are mutually orthogonal.  Then B = U diag(sigma),
// sigma_c = ||B[:, c]||, and B_in = U diag(sigma) W^T.  Each wave rotates
// the pairs of a round it owns (disjoint columns), with its 64 lanes over
// the rows.
__device__ void wg_jacobi_svd(double* B, int rows, int cols, double* W, JacobiLds& sh) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const i64 cc = (i64)cols * cols;
    double nrm = 0.0;
    for (i64 e = t; e < cc; e += DC_TPB) W[e] = (e % cols == e / cols) ? 1.0 : 0.0;
    for (i64 e = t; e < (i64)rows * cols; e += DC_TPB) nrm += B[e] * B[e];
    const double floor_ = 1e-32 * wg_sum(nrm, sh.red);  // on inner products
    if (cols < 2) return;
    const int mp = cols + (cols & 1), np = mp / 2;
    for (int sweep = 0; sweep < DC_SWEEPS; ++sweep) {
        __syncthreads();
        if (t == 0) sh.flag = 0;
        __syncthreads();
        for (int r = 0; r < mp - 1; ++r) {
            for (int pr = wv; pr < np; pr += DC_WAVES) {
                int p, q;
                rr_pair(pr, r, mp, p, q);
                if (p >= cols || q >= cols) continue;
                double* bp = B + (i64)p * rows;
                double* bq = B + (i64)q * rows;
                double a = 0.0, b = 0.0, g = 0.0;
                for (int i = lane; i < rows; i += 64) {
                    const double x = bp[i], y = bq[i];
                    a += x * x;
                    b += y * y;
                    g += x * y;
                }
                for (int o = 32; o > 0; o >>= 1) {
                    a += __shfl_xor(a, o);
                    b += __shfl_xor(b, o);
                    g += __shfl_xor(g, o);
                }
                double c, s;
                if (!jacobi_rot(a, b, g, floor_, c, s)) continue;
                if (lane == 0) sh.flag = 1;
                for (int i = lane; i < rows; i += 64) {
                    const double x = bp[i], y = bq[i];
                    bp[i] = c * x - s * y;
                    bq[i] = s * x + c * y;
                }
                double* wp = W + (i64)p * cols;
                double* wq = W + (i64)q * cols;
                for (int i = lane; i < cols; i += 64) {
                    const double x = wp[i], y = wq[i];
                    wp[i] = c * x - s * y;
                    wq[i] = s * x + c * y;
                }
            }
            __syncthreads();
        }
        if (sh.flag == 0) break;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// psdcompletion

// X (n x n, ld ldx) := the dense symmetric copy of the cspmatrix blocks
// (psdcompletion.py:19): for supernode k = blockIdx.x, entry (i, j) of its
// lower trapezoid goes to X[row_i, col_j] and X[col_j, row_i].  X is zero on
// entry.
__global__ __launch_bounds__(DC_TPB) void k_blk_to_dense(const i64* __restrict__ snptr,
                                                         const i64* __restrict__ sncolptr,
                                                         const i64* __restrict__ snrowidx,
                                                         const i64* __restrict__ blkptr,
                                                         const double* __restrict__ blkval,
                                                         double* __restrict__ X, i64 ldx) {
    const int k = blockIdx.x;
    const i64 c0 = snptr[k];
    const int nn = (int)(snptr[k + 1] - c0);
    const int nj = (int)(sncolptr[k + 1] - sncolptr[k]);
    const i64* rows = snrowidx + sncolptr[k];
    const double* B = blkval + blkptr[k];
    for (i64 e = threadIdx.x; e < (i64)nj * nn; e += DC_TPB) {
        const int j = (int)(e / nj), i = (int)(e % nj);
        if (i < j) continue;
        const double v = B[e];
        const i64 r = rows[i], c = c0 + j;
        X[r + c * ldx] = v;
        X[c + r * ldx] = v;
    }
}

// Separator step of supernode k (psdcompletion.py:38-62), one workgroup:
//   W := X[alpha, alpha]^{-1} X[alpha, nu]   (Cholesky: potrf + 2 trtrs)
// or, if the factorization fails,
//   W := Z diag(w^+) Z^T X[alpha, nu],  X[alpha, alpha] = Z diag(w) Z^T,
//   w^+_i = 1 / w_i for w_i > tol * max(w), else 0.
// Also marks the separator rows (mark[alpha_a] = k) for k_psd_update.
// Work: S, Z (na x na), T (na x nn).
//
// edm != 0: the separator step of edmcompletion (edmcompletion.py:17-46),
// on the Gram transform relative to the first separator row a0,
//   Y[i, j] = -X[i, j] / 2 - X[a0, a0] / 2 + X[i, a0] / 2 + X[j, a0] / 2:
// W := Yaa^+ Yan, always through the eigendecomposition, and E receives
// the vector c_a = -X[a0, a0] / 2 + X[a0, alpha_a] / 2 (na) of the rows'
// Yea and the diagonal Ynn[j, j] (nn) for k_psd_update.
__global__ __launch_bounds__(DC_TPB) void k_psd_sep(int k, const double* __restrict__ X, i64 ldx,
                                                    const i64* __restrict__ alpha, int na,
                                                    i64 c0, int nn, double tol, int edm,
                                                    double* __restrict__ S,
                                                    double* __restrict__ Z,
                                                    double* __restrict__ W,
                                                    double* __restrict__ T,
                                                    double* __restrict__ E,
                                                    int* __restrict__ mark) {
    __shared__ JacobiLds sh;
    __shared__ int fail;
    const int t = threadIdx.x;
    const i64 aa = (i64)na * na, an = (i64)na * nn;
    if (t == 0) fail = edm;
    for (int a = t; a < na; a += DC_TPB) mark[alpha[a]] = k;
    if (edm) {
        const i64 a0 = alpha[0];
        const double h00 = 0.5 * X[a0 + a0 * ldx];
        for (i64 e = t; e < aa; e += DC_TPB) {
            const i64 ra = alpha[e % na], rb = alpha[e / na];
            S[e] = -0.5 * X[ra + rb * ldx] - h00 + 0.5 * X[ra + a0 * ldx] + 0.5 * X[rb + a0 * ldx];
        }
        for (i64 e = t; e < an; e += DC_TPB) {
            const i64 ra = alpha[e % na], cj = c0 + e / na;
            W[e] = -0.5 * X[ra + cj * ldx] - h00 + 0.5 * X[a0 + cj * ldx] + 0.5 * X[ra + a0 * ldx];
        }
        for (int a = t; a < na; a += DC_TPB) E[a] = -h00 + 0.5 * X[a0 + alpha[a] * ldx];
        for (int j = t; j < nn; j += DC_TPB) {
            const i64 cj = c0 + j;
            E[na + j] = -0.5 * X[cj + cj * ldx] - h00 + X[cj + a0 * ldx];
        }
        __syncthreads();
    } else {
        for (i64 e = t; e < aa; e += DC_TPB) S[e] = X[alpha[e % na] + alpha[e / na] * ldx];
        for (i64 e = t; e < an; e += DC_TPB) W[e] = X[alpha[e % na] + (c0 + e / na) * ldx];
        __syncthreads();
    }
    // right-looking Cholesky, lower triangle (dpotf2's failure test: a
    // pivot <= 0 or NaN)
    for (int j = 0; j < na && !edm; ++j) {
        const double d = S[j + (i64)j * na];
        if (!(d > 0.0)) {
            if (t == 0) fail = 1;
            break;  // uniform: every thread read the same pivot
        }
        const double l = sqrt(d);
        __syncthreads();
        if (t == 0) S[j + (i64)j * na] = l;
        for (int i = j + 1 + t; i < na; i += DC_TPB) S[i + (i64)j * na] /= l;
        __syncthreads();
        const int m = na - j - 1;
        for (i64 e = t; e < (i64)m * m; e += DC_TPB) {
            const int jj = j + 1 + (int)(e / m), ii = j + 1 + (int)(e % m);
            if (ii >= jj) S[ii + (i64)jj * na] -= S[ii + (i64)j * na] * S[jj + (i64)j * na];
        }
        __syncthreads();
    }
    __syncthreads();
    if (!fail) {
        // W := L^{-T} L^{-1} W, one row of the substitution at a time
        for (int a = 0; a < na; ++a) {
            const double l = S[a + (i64)a * na];
            for (int j = t; j < nn; j += DC_TPB) W[a + (i64)j * na] /= l;
            __syncthreads();
            const int m = na - a - 1;
            for (i64 e = t; e < (i64)m * nn; e += DC_TPB) {
                const int b = a + 1 + (int)(e % m), j = (int)(e / m);
                W[b + (i64)j * na] -= S[b + (i64)a * na] * W[a + (i64)j * na];
            }
            __syncthreads();
        }
        for (int a = na - 1; a >= 0; --a) {
            const double l = S[a + (i64)a * na];
            for (int j = t; j < nn; j += DC_TPB) W[a + (i64)j * na] /= l;
            __syncthreads();
            for (i64 e = t; e < (i64)a * nn; e += DC_TPB) {
                const int b = (int)(e % a), j = (int)(e / a);
                W[b + (i64)j * na] -= S[a + (i64)b * na] * W[a + (i64)j * na];
            }
            __syncthreads();
        }
        return;
    }
    // eigendecomposition pseudo-inverse (S still holds Yaa in the edm case)
    if (!edm)
        for (i64 e = t; e < aa; e += DC_TPB) S[e] = X[alpha[e % na] + alpha[e / na] * ldx];
    __syncthreads();
    wg_jacobi_evd(S, Z, na, sh);
    double lmax = -INFINITY;
    for (int i = 0; i < na; ++i) lmax = fmax(lmax, S[i + (i64)i * na]);
    // T = diag(w^+) Z^T W
    for (i64 e = t; e < an; e += DC_TPB) {
        const int i = (int)(e % na), j = (int)(e / na);
        const double wi = S[i + (i64)i * na];
        double acc = 0.0;
        if (wi > lmax * tol) {
            for (int a = 0; a < na; ++a) acc += Z[a + (i64)i * na] * W[a + (i64)j * na];
            acc /= wi;
        }
        T[e] = acc;
    }
    __syncthreads();
    for (i64 e = t; e < an; e += DC_TPB) {
        const int a = (int)(e % na), j = (int)(e / na);
        double acc = 0.0;
        for (int i = 0; i < na; ++i) acc += Z[a + (i64)i * na] * T[i + (i64)j * na];
        W[e] = acc;
    }
}

constexpr int PSD_NB = 8;  // columns of nu per thread

// X[i, nu] = X[i, alpha] W and X[nu, i] = its transpose for every row
// i >= c0 + nn outside the separator (mark[i] != k): one thread per row,
// PSD_NB columns of nu per blockIdx.y.  edm != 0 (edmcompletion.py:32-48):
// X[i, nu_j] = -2 Yea[i, :] W[:, j] + Ynn[j, j] + Yee[i, i] with
// Yea[i, a] = -X[i, alpha_a] / 2 + c_a + X[i, a0] / 2 and
// Yee[i, i] = -X[i, i] / 2 - X[a0, a0] / 2 + X[i, a0].
__global__ __launch_bounds__(DC_TPB) void k_psd_update(int k, double* __restrict__ X, i64 ldx,
                                                       i64 n, const i64* __restrict__ alpha,
                                                       int na, i64 c0, int nn, int edm,
                                                       const double* __restrict__ W,
                                                       const double* __restrict__ E,
                                                       const int* __restrict__ mark) {
    const i64 i = c0 + nn + (i64)blockIdx.x * DC_TPB + threadIdx.x;
    if (i >= n || mark[i] == k) return;
    const int j0 = blockIdx.y * PSD_NB;
    const int nb = min(PSD_NB, nn - j0);
    double acc[PSD_NB];
#pragma unroll
    for (int j = 0; j < PSD_NB; ++j) acc[j] = 0.0;
    const i64 a0 = alpha[0];
    const double xi0 = edm ? 0.5 * X[i + a0 * ldx] : 0.0;
    for (int a = 0; a < na; ++a) {
        double x = X[i + alpha[a] * ldx];
        if (edm) x = -0.5 * x + E[a] + xi0;
        const double* w = W + a + (i64)j0 * na;
#pragma unroll
        for (int j = 0; j < PSD_NB; ++j)
            if (j < nb) acc[j] += x * w[(i64)j * na];
    }
    const double yee = edm ? -0.5 * X[i + i * ldx] - 0.5 * X[a0 + a0 * ldx] + 2.0 * xi0 : 0.0;
#pragma unroll
    for (int j = 0; j < PSD_NB; ++j) {
        if (j >= nb) break;
        const i64 c = c0 + j0 + j;
        const double v = edm ? -2.0 * acc[j] + E[na + j0 + j] + yee : acc[j];
        X[i + c * ldx] = v;
        X[c + i * ldx] = v;
    }
}

// out (rows x cols, ld ldo) = X[ip[a], permute_cols ? ip[b] : b]: the
// inverse permutation of the reordered result (psdcompletion.py:67-70,
// mrcompletion.py:176-179).
__global__ __launch_bounds__(DC_TPB) void k_dense_permute(const double* __restrict__ X, i64 ldx,
                                                          const i64* __restrict__ ip, i64 rows,
                                                          i64 cols, int permute_cols,
                                                          double* __restrict__ out, i64 ldo) {
    const i64 a = (i64)blockIdx.x * DC_TPB + threadIdx.x;
    if (a >= rows) return;
    const i64 ra = ip[a];
    for (i64 b = blockIdx.y; b < cols; b += gridDim.y) {
        const i64 cb = permute_cols ? ip[b] : b;
        out[a + b * ldo] = X[ra + cb * ldx];
    }
}

// ---------------------------------------------------------------------------
// mrcompletion

// Value of A[r, c], r >= c, in the filled pattern: column c belongs to
// supernode s = snode_of[c]; r is found in s's clique rows by binary search.
__device__ inline double pattern_entry(i64 r, i64 c, const int* snode_of, const i64* snptr,
                                       const i64* sncolptr, const i64* snrowidx,
                                       const i64* blkptr, const double* blkval) {
    const int s = snode_of[c];
    const i64* rows = snrowidx + sncolptr[s];
    const i64 nj = sncolptr[s + 1] - sncolptr[s];
    i64 lo = 0, hi = nj;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if (rows[mid] < r) lo = mid + 1; else hi = mid;
    }
    return blkval[blkptr[s] + (c - snptr[s]) * nj + lo];
}

// Clique k = blockIdx.x (mrcompletion.py:135-156): F_k = A[beta, beta],
// its eigendecomposition sorted ascending (as syevr), the numerical rank
// rk = #{w_i > 1e-14 w_max}, and Zs = the eigenvectors with the last rk
// columns scaled by sqrt(w).  Arena of k at zoff[k]: F (nj^2), V (nj^2),
// Zs (nj^2), w (nj), perm (nj, as doubles).
__global__ __launch_bounds__(DC_TPB) void k_mr_clique(const int* __restrict__ snode_of,
                                                      const i64* __restrict__ snptr,
                                                      const i64* __restrict__ sncolptr,
                                                      const i64* __restrict__ snrowidx,
                                                      const i64* __restrict__ blkptr,
                                                      const double* __restrict__ blkval,
                                                      const i64* __restrict__ zoff,
                                                      double* __restrict__ arena,
                                                      int* __restrict__ rank) {
    __shared__ JacobiLds sh;
    const int k = blockIdx.x, t = threadIdx.x;
    const int nj = (int)(sncolptr[k + 1] - sncolptr[k]);
    const i64* beta = snrowidx + sncolptr[k];
    const i64 mm = (i64)nj * nj;
    double* F = arena + zoff[k];
    double* V = F + mm;
    double* Zs = V + mm;
    double* w = Zs + mm;
    double* perm = w + nj;
    for (i64 e = t; e < mm; e += DC_TPB) {
        const i64 ra = beta[e % nj], rb = beta[e / nj];
        F[e] = pattern_entry(max(ra, rb), min(ra, rb), snode_of, snptr, sncolptr, snrowidx,
                             blkptr, blkval);
    }
    __syncthreads();
    wg_jacobi_evd(F, V, nj, sh);
    for (int i = t; i < nj; i += DC_TPB) {
        const double wi = F[i + (i64)i * nj];
        int pos = 0;
        for (int j = 0; j < nj; ++j) {
            const double wj = F[j + (i64)j * nj];
            pos += (wj < wi) || (wj == wi && j < i);
        }
        w[pos] = wi;
        perm[pos] = (double)i;
    }
    __syncthreads();
    const double wmax = w[nj - 1];
    int rk = 0;
    for (int i = 0; i < nj; ++i) rk += w[i] > 1e-14 * wmax;
    for (i64 e = t; e < mm; e += DC_TPB) {
        const int i = (int)(e % nj), j = (int)(e / nj);
        const double z = V[i + (i64)perm[j] * nj];
        Zs[e] = j >= nj - rk ? z * sqrt(w[j]) : z;
    }
    if (t == 0) rank[k] = rk;
}

// Rows of clique k's own columns, from the roots down (mrcompletion.py
// :154-174), for the cliques of one level (list[blockIdx.x]).  R = the
// largest clique rank (the reference's running r ends there; columns
// beyond a clique's rank are zero either way).  With Yn = [Zs[:nn, last rk]
// 0] and, for a clique with a separator,
//   V  = [Zs[nn:, last rk] 0] = P S Q1t,   Ya = Y[alpha, :R] = . S Q2t,
//   Q2t[i, :] := P[:, i]^T Ya / S_i   for i < min(na, rk), S_i > 1e-14 S_0,
//   Y[nu, :R] = Yn Q1t^T Q2t.
// The decompositions are one-sided Jacobi on V and on a copy of Ya, with the
// singular values sorted in descending order as gesvd returns them.
// Slot of the clique at soff[k]: Vb, Ya, Yc (na x R each), W1, W2, Q
// (R x R each), M (nn x R), ord1, ord2 (R each, as doubles).
__global__ __launch_bounds__(DC_TPB) void k_mr_level(const int* __restrict__ list,
                                                     const i64* __restrict__ snptr,
                                                     const i64* __restrict__ sncolptr,
                                                     const i64* __restrict__ snrowidx,
                                                     const i64* __restrict__ zoff,
                                                     const double* __restrict__ arena,
                                                     const int* __restrict__ rank,
                                                     const i64* __restrict__ soff,
                                                     double* __restrict__ slots, int R,
                                                     double* __restrict__ Y, i64 ldy) {
    __shared__ JacobiLds sh;
    const int k = list[blockIdx.x], t = threadIdx.x;
    const i64 c0 = snptr[k];
    const int nn = (int)(snptr[k + 1] - c0);
    const int nj = (int)(sncolptr[k + 1] - sncolptr[k]);
    const int na = nj - nn;
    const int rk = rank[k];
    const double* Zs = arena + zoff[k] + 2 * (i64)nj * nj;
    const int z0 = nj - rk;  // first scaled eigenvector column
    if (na == 0) {
        for (i64 e = t; e < (i64)nn * rk; e += DC_TPB) {
            const int i = (int)(e % nn), c = (int)(e / nn);
            Y[(c0 + i) + c * ldy] = Zs[i + (i64)(z0 + c) * nj];
        }
        return;
    }
    const i64* Ia = snrowidx + sncolptr[k] + nn;
    const i64 aR = (i64)na * R, RR = (i64)R * R;
    double* Vb = slots + soff[k];
    double* Ya = Vb + aR;
    double* Yc = Ya + aR;
    double* W1 = Yc + aR;
    double* W2 = W1 + RR;
    double* Q = W2 + RR;
    double* M = Q + RR;  // 2 R singular values, then the nn x R product
    double* ord1 = M + (2 + (i64)nn) * R;
    double* ord2 = ord1 + R;
    for (i64 e = t; e < aR; e += DC_TPB) {
        const int a = (int)(e % na), c = (int)(e / na);
        Vb[e] = c < rk ? Zs[(nn + a) + (i64)(z0 + c) * nj] : 0.0;
        const double y = Y[Ia[a] + c * ldy];
        Ya[e] = y;
        Yc[e] = y;
    }
    __syncthreads();
    wg_jacobi_svd(Vb, na, R, W1, sh);
    wg_jacobi_svd(Yc, na, R, W2, sh);
    // column norms, then the descending orders (ties by index)
    for (int c = t; c < R; c += DC_TPB) {
        double s1 = 0.0, s2 = 0.0;
        for (int a = 0; a < na; ++a) {
            s1 += Vb[a + (i64)c * na] * Vb[a + (i64)c * na];
            s2 += Yc[a + (i64)c * na] * Yc[a + (i64)c * na];
        }
        M[c] = sqrt(s1);
        M[R + c] = sqrt(s2);
    }
    __syncthreads();
    for (int c = t; c < R; c += DC_TPB) {
        const double s1 = M[c], s2 = M[R + c];
        int p1 = 0, p2 = 0;
        for (int d = 0; d < R; ++d) {
            p1 += (M[d] > s1) || (M[d] == s1 && d < c);
            p2 += (M[R + d] > s2) || (M[R + d] == s2 && d < c);
        }
        ord1[p1] = (double)c;
        ord2[p2] = (double)c;
    }
    __syncthreads();
    // Q (= Q2t): row i is the i-th right singular vector of Ya ...
    for (i64 e = t; e < RR; e += DC_TPB) {
        const int i = (int)(e % R), c = (int)(e / R);
        Q[e] = W2[c + (i64)ord2[i] * R];
    }
    __syncthreads();
    // ... replaced by P[:, i]^T Ya / S_i on the significant leading rows
    // (S = Ya's singular values, the second gesvd's output)
    const double S0 = M[R + (int)ord2[0]];
    const int lim = min(na, rk);
    for (i64 e = t; e < (i64)lim * R; e += DC_TPB) {
        const int i = (int)(e % lim), c = (int)(e / lim);
        const double Si = M[R + (int)ord2[i]];
        if (!(Si > 1e-14 * S0)) continue;
        const int c1 = (int)ord1[i];
        const double s1 = M[c1];
        double acc = 0.0;
        if (s1 > 0.0)
            for (int a = 0; a < na; ++a) acc += Vb[a + (i64)c1 * na] * Ya[a + (i64)c * na];
        Q[i + (i64)c * R] = s1 > 0.0 ? acc / s1 / Si : 0.0;
    }
    __syncthreads();
    // M = Yn Q1t^T: M[i, j] = sum_c Yn[i, c] W1[c, ord1[j]], Yn[i, c] =
    // Zs[i, z0 + c] for c < rk
    for (i64 e = t; e < (i64)nn * R; e += DC_TPB) {
        const int i = (int)(e % nn), j = (int)(e / nn);
        const double* w1 = W1 + (i64)ord1[j] * R;
        double acc = 0.0;
        for (int c = 0; c < rk; ++c) acc += Zs[i + (i64)(z0 + c) * nj] * w1[c];
        M[e + 2 * (i64)R] = acc;
    }
    __syncthreads();
    // Y[nu, :R] = M Q2t
    for (i64 e = t; e < (i64)nn * R; e += DC_TPB) {
        const int i = (int)(e % nn), c = (int)(e / nn);
        const double* m = M + 2 * (i64)R;
        double acc = 0.0;
        for (int j = 0; j < R; ++j) acc += m[i + (i64)j * nn] * Q[j + (i64)c * R];
        Y[(c0 + i) + c * ldy] = acc;
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

template <class T>
int dev_upload(DevBuf& b, const T* h, size_t count) {
    HIP_CHECK(hipMalloc(&b.p, std::max<size_t>(count, 1) * sizeof(T)));
    if (count) HIP_CHECK(hipMemcpy(b.p, h, count * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

int check_structure(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr) {
    if (n <= 0 || nsn <= 0) {
        chp_set_error("n and nsn must be positive");
        return -1;
    }
    for (int64_t k = 0; k < nsn; ++k) {
        const int64_t nj = sncolptr[k + 1] - sncolptr[k], nn = snptr[k + 1] - snptr[k];
        if (nn <= 0 || nj < nn) {
            chp_set_error("inconsistent supernode arrays");
            return -1;
        }
        if (nj > DC_MAXN) {
            chp_set_error("clique larger than " + std::to_string(DC_MAXN) +
                          " rows: not supported by the dense completions");
            return -1;
        }
    }
    return 0;
}

// The sweep of psdcompletion (edm = 0) and edmcompletion (edm = 1).
int dense_sweep(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr,
                const int64_t* snrowidx, const int64_t* blkptr, const double* blkval,
                double* X, int64_t ldx, double tol, int edm, void* stream) {
    if (!snptr || !sncolptr || !snrowidx || !blkptr || !blkval || !X) {
        chp_set_error("null argument");
        return -1;
    }
    if (int rc = check_structure(n, nsn, snptr, sncolptr)) return rc;
    if (ldx < n) {
        chp_set_error("ldx is smaller than n");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    DevBuf d_snptr, d_colptr, d_rowidx, d_blkptr, d_work, d_mark;
    if (int rc = dev_upload(d_snptr, snptr, nsn + 1)) return rc;
    if (int rc = dev_upload(d_colptr, sncolptr, nsn + 1)) return rc;
    if (int rc = dev_upload(d_rowidx, snrowidx, sncolptr[nsn])) return rc;
    if (int rc = dev_upload(d_blkptr, blkptr, nsn + 1)) return rc;
    i64 amax = 0, abmax = 0;
    for (int64_t k = 0; k < nsn; ++k) {
        const i64 nn = snptr[k + 1] - snptr[k], na = sncolptr[k + 1] - sncolptr[k] - nn;
        amax = std::max(amax, na);
        abmax = std::max(abmax, na * nn);
    }
    const i64 aa = std::max<i64>(amax * amax, 1), an = std::max<i64>(abmax, 1);
    const i64 ne = std::max<i64>(amax + (i64)(snptr[nsn] - snptr[0]), 1);  // na + nn <= n
    HIP_CHECK(hipMalloc(&d_work.p, (2 * aa + 2 * an + ne) * sizeof(double)));
    HIP_CHECK(hipMalloc(&d_mark.p, n * sizeof(int)));
    double* S = (double*)d_work.p;
    double* Z = S + aa;
    double* W = Z + aa;
    double* T = W + an;
    double* E = T + an;
    int* mark = (int*)d_mark.p;
    HIP_CHECK(hipMemsetAsync(mark, 0xff, n * sizeof(int), s));  // -1: no supernode
    hipLaunchKernelGGL(k_blk_to_dense, dim3((unsigned)nsn), dim3(DC_TPB), 0, s,
                       (const i64*)d_snptr.p, (const i64*)d_colptr.p, (const i64*)d_rowidx.p,
                       (const i64*)d_blkptr.p, blkval, X, (i64)ldx);
    const i64* rowidx = (const i64*)d_rowidx.p;
    for (int64_t k = nsn - 1; k >= 0; --k) {
        const i64 c0 = snptr[k];
        const int nn = (int)(snptr[k + 1] - c0);
        const int na = (int)(sncolptr[k + 1] - sncolptr[k] - nn);
        if (na == 0) continue;
        const i64* alpha = rowidx + sncolptr[k] + nn;
        hipLaunchKernelGGL(k_psd_sep, dim3(1), dim3(DC_TPB), 0, s, (int)k, (const double*)X,
                           (i64)ldx, alpha, na, c0, nn, tol, edm, S, Z, W, T, E, mark);
        const i64 rows = n - (c0 + nn);
        if (rows > 0)
            hipLaunchKernelGGL(k_psd_update,
                               dim3((unsigned)((rows + DC_TPB - 1) / DC_TPB),
                                    (unsigned)((nn + PSD_NB - 1) / PSD_NB)),
                               dim3(DC_TPB), 0, s, (int)k, X, (i64)ldx, (i64)n, alpha, na, c0,
                               nn, edm, (const double*)W, (const double*)E, (const int*)mark);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));  // the work arrays are freed on return
    return 0;
}

}  // namespace

extern "C" {

int chp_psdcompletion(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr,
                      const int64_t* snrowidx, const int64_t* blkptr, const double* blkval,
                      double* X, int64_t ldx, double tol, void* stream) {
    return dense_sweep(n, nsn, snptr, sncolptr, snrowidx, blkptr, blkval, X, ldx, tol, 0, stream);
}

int chp_edmcompletion(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr,
                      const int64_t* snrowidx, const int64_t* blkptr, const double* blkval,
                      double* X, int64_t ldx, double tol, void* stream) {
    return dense_sweep(n, nsn, snptr, sncolptr, snrowidx, blkptr, blkval, X, ldx, tol, 1, stream);
}

int chp_mrcompletion(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr,
                     const int64_t* snrowidx, const int64_t* blkptr, const int64_t* snpar,
                     const double* blkval, double* Y, int64_t ldy, int maxr, int* rank_out,
                     void* stream) {
    if (!snptr || !sncolptr || !snrowidx || !blkptr || !snpar || !blkval || !Y || !rank_out) {
        chp_set_error("null argument");
        return -1;
    }
    if (int rc = check_structure(n, nsn, snptr, sncolptr)) return rc;
    if (ldy < n || maxr <= 0) {
        chp_set_error("ldy is smaller than n or maxr is not positive");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    // column -> supernode, etree depth of every supernode (parents have
    // larger indices: depth from the top down), cliques per depth
    std::vector<int> snode_of(n), depth(nsn, 0);
    for (int64_t k = 0; k < nsn; ++k)
        for (int64_t c = snptr[k]; c < snptr[k + 1]; ++c) snode_of[c] = (int)k;
    int H = 0;
    for (int64_t k = nsn - 1; k >= 0; --k) {
        const int64_t p = snpar[k];
        if (p < 0 || p > nsn - 1 || (p != k && p < k)) {
            chp_set_error("snpar is not a postordered supernodal tree");
            return -1;
        }
        depth[k] = p == k ? 0 : depth[p] + 1;
        H = std::max(H, depth[k] + 1);
    }
    std::vector<std::vector<int>> bylevel(H);
    for (int64_t k = 0; k < nsn; ++k) bylevel[depth[k]].push_back((int)k);
    std::vector<i64> zoff(nsn);
    i64 ztot = 0;
    for (int64_t k = 0; k < nsn; ++k) {
        const i64 nj = sncolptr[k + 1] - sncolptr[k];
        zoff[k] = ztot;
        ztot += 3 * nj * nj + 2 * nj;
    }
    DevBuf d_snode, d_snptr, d_colptr, d_rowidx, d_blkptr, d_zoff, d_arena, d_rank, d_soff,
        d_slots, d_list;
    if (int rc = dev_upload(d_snode, snode_of.data(), n)) return rc;
    if (int rc = dev_upload(d_snptr, snptr, nsn + 1)) return rc;
    if (int rc = dev_upload(d_colptr, sncolptr, nsn + 1)) return rc;
    if (int rc = dev_upload(d_rowidx, snrowidx, sncolptr[nsn])) return rc;
    if (int rc = dev_upload(d_blkptr, blkptr, nsn + 1)) return rc;
    if (int rc = dev_upload(d_zoff, zoff.data(), nsn)) return rc;
    HIP_CHECK(hipMalloc(&d_arena.p, std::max<i64>(ztot, 1) * sizeof(double)));
    HIP_CHECK(hipMalloc(&d_rank.p, nsn * sizeof(int)));
    hipLaunchKernelGGL(k_mr_clique, dim3((unsigned)nsn), dim3(DC_TPB), 0, s,
                       (const int*)d_snode.p, (const i64*)d_snptr.p, (const i64*)d_colptr.p,
                       (const i64*)d_rowidx.p, (const i64*)d_blkptr.p, blkval,
                       (const i64*)d_zoff.p, (double*)d_arena.p, (int*)d_rank.p);
    HIP_CHECK(hipGetLastError());
    std::vector<int> rank(nsn);
    HIP_CHECK(hipMemcpyAsync(rank.data(), d_rank.p, nsn * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    int R = 0;
    for (int64_t k = 0; k < nsn; ++k) R = std::max(R, rank[k]);
    if (R > maxr) {
        chp_set_error("clique rank exceeds maxr");
        return -1;
    }
    *rank_out = R;
    for (int64_t c = 0; c < maxr; ++c)
        HIP_CHECK(hipMemsetAsync(Y + c * ldy, 0, n * sizeof(double), s));
    if (R == 0) {
        HIP_CHECK(hipStreamSynchronize(s));
        return 0;
    }
    // per-level slots (levels run one after the other and reuse the space)
    std::vector<i64> soff(nsn, 0);
    i64 smax = 0;
    for (int L = 0; L < H; ++L) {
        i64 cur = 0;
        for (int k : bylevel[L]) {
            const i64 nn = snptr[k + 1] - snptr[k], na = sncolptr[k + 1] - sncolptr[k] - nn;
            soff[k] = cur;
            if (na > 0) cur += 3 * na * R + 3 * (i64)R * R + (nn + 4) * R;
        }
        smax = std::max(smax, cur);
    }
    std::vector<int> flat;
    std::vector<size_t> lstart(H + 1, 0);
    for (int L = 0; L < H; ++L) {
        lstart[L] = flat.size();
        flat.insert(flat.end(), bylevel[L].begin(), bylevel[L].end());
    }
    lstart[H] = flat.size();
    if (int rc = dev_upload(d_soff, soff.data(), nsn)) return rc;
    if (int rc = dev_upload(d_list, flat.data(), flat.size())) return rc;
    HIP_CHECK(hipMalloc(&d_slots.p, std::max<i64>(smax, 1) * sizeof(double)));
    for (int L = 0; L < H; ++L) {
        const unsigned cnt = (unsigned)(lstart[L + 1] - lstart[L]);
        hipLaunchKernelGGL(k_mr_level, dim3(cnt), dim3(DC_TPB), 0, s,
                           (const int*)d_list.p + lstart[L], (const i64*)d_snptr.p,
                           (const i64*)d_colptr.p, (const i64*)d_rowidx.p, (const i64*)d_zoff.p,
                           (const double*)d_arena.p, (const int*)d_rank.p, (const i64*)d_soff.p,
                           (double*)d_slots.p, R, Y, (i64)ldy);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

int chp_dense_permute(const double* X, int64_t ldx, const int64_t* ip, int64_t rows,
                      int64_t cols, int permute_cols, double* out, int64_t ldo, void* stream) {
    if (!X || !ip || !out || rows < 0 || cols < 0 || ldx < 1 || ldo < rows) {
        chp_set_error("invalid argument");
        return -1;
    }
    if (rows == 0 || cols == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_dense_permute,
                       dim3((unsigned)((rows + DC_TPB - 1) / DC_TPB), (unsigned)std::min<i64>(cols, 65535)),
                       dim3(DC_TPB), 0, s, X, (i64)ldx, (const i64*)ip, (i64)rows, (i64)cols,
                       permute_cols, out, (i64)ldo);
    HIP_CHECK(hipGetLastError());
    return 0;
}

}  // extern "C"
