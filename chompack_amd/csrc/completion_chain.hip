// Completion with factored updates (completion.c:24-118, update_factor.c:3-108)
// along a chain of small fronts (one front per level, nj <= 64), in one
// 64-thread workgroup: every front's columns B = [X11; X21] and the factor Lk
// of its update matrix (V_k = Lk^T Lk) stay in LDS, and the update factor of
// the chain's next front is built in place in a second factor buffer, which
// becomes that front's Lk.  One wave, so LDS needs no workgroup barrier
// beyond ordering; every dependent chain is kept in registers where it can
// be:
//   * the triangular solves hold the coefficients of the next eight steps in
//     registers (a step is a broadcast and an fma);
//   * a reflection's band, when short, is summed and applied from registers.
// The items run in order: a front (kind 1, carrying the number nch of its
// children) is followed by the update factors of its children (kind 0, and
// kind 2 for the child that is the chain's next front, which comes last).
#include "common.h"
#include "kernels.h"

#define PIVOT_NOISE 1e-13  // as k_update_factor: a pivot below PIVOT_NOISE ||x|| is +0
#define SB 8               // band rows held in registers

namespace {

typedef double Tile[65];

__device__ inline double rdl_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__device__ inline double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Orders LDS accesses of the wave (no workgroup to wait for).
__device__ inline void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One right-hand side across the lanes (lane l: row l) against L lower
// triangular of order n <= 64, L(i, j) = T[rb + i][cb + j], rdg = 1 / L(l, l)
// (0 for l >= n).  Step i subtracts c_i(l) z_i, z_i being lane i's running
// value (final by then), with c_i(l) = L(i, l) / L(i, i) for l < i
// (TRANS: L^T z = b, last row first) or L(l, i) / L(i, i) for l > i (L w = b);
// the solution is z rdg.  The next eight steps' coefficients are read while
// the current eight run.
template <bool TRANS>
__device__ inline double tsolve(const Tile* T, int rb, int cb, int n, double z, double rdg) {
    const int l = threadIdx.x;
    const int nb = (n + 7) / 8;
    double cur[8], nxt[8];
    auto fetch = [&](int b, double (&v)[8]) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = 8 * b + u;
            const bool ok = i < n && (TRANS ? l < i : (l > i && l < n));
            v[u] = TRANS ? T[ok ? rb + i : 0][ok ? cb + l : 0] : T[ok ? rb + l : 0][ok ? cb + i : 0];
        }
    };
#pragma unroll
    for (int u = 0; u < 8; ++u) nxt[u] = 0.0;
#pragma unroll
    for (int bb = 0; bb < 8; ++bb) {
        const int b = TRANS ? 7 - bb : bb;
        if (b < nb) {
            if (TRANS ? (b == nb - 1) : (b == 0)) fetch(b, cur);
            const int bn = TRANS ? b - 1 : b + 1;
            if (bn >= 0 && bn < nb) fetch(bn, nxt);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = 8 * b + u;
                const bool ok = i < n && (TRANS ? l < i : (l > i && l < n));
                cur[u] *= ok ? rdl_d(rdg, i) : 0.0;
            }
#pragma unroll
            for (int uu = 0; uu < 8; ++uu) {  // (steps i >= n subtract zero)
                const int u = TRANS ? 7 - uu : uu;
                z -= cur[u] * rdl_d(z, 8 * b + u);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) cur[u] = nxt[u];
        }
    }
    return z * rdg;
}

// Reversed Cholesky (A = Lr^T Lr, last row first) of the n x n lower block
// at T[rb + i][j], in place.  Returns the failing step n - p (0 on success).
__device__ inline int rev_chol(Tile* T, int rb, int n) {
    const int l = threadIdx.x;
    for (int p = n - 1; p >= 0; --p) {
        const double d = T[rb + p][p];
        if (!(d > 0.0)) return n - p;
        const double sd = sqrt(d), rs = 1.0 / sd;
        wsync();  // every lane has read the pivot
        if (l < p) T[rb + p][l] *= rs;
        if (l == p) T[rb + p][p] = sd;
        wsync();
        if (l < p) {
            const double ri = T[rb + p][l];
            for (int j = 0; j <= l; ++j) T[rb + l][j] -= ri * T[rb + p][j];
        }
        wsync();
    }
    return 0;
}

// Update factor Li of child t (update_factor.c:3-108) from the front's
// columns Bf and its factor Lk = V[voff + i][j], built in T: Li(i, j) at
// T[toff + i][j] for j <= i (the entries above the diagonal are not
// cleared).  C = Lk[:, r2 - nn] is placed at T rows crow + sh, columns
// N1 + j, so that the lower-triangular form the reflections leave in its
// last N2 rows is Li22 in place.  Returns false after recording a failed
// reversed Cholesky.
__device__ inline bool child_factor(const UFTask t, const Tile* V, int voff, const Tile* Bf,
                                    Tile* T, int& toff, const int* __restrict__ relidx,
                                    unsigned long long* status) {
    const int l = threadIdx.x;
    const int nn = t.nn, na = t.na, N = t.N;
    const int rl = l < N ? relidx[t.relpos + l] : 0;  // lane j's r[j], broadcast by readlane
    const int N1 = __popcll(__ballot(l < N && rl < nn)), N2 = N - N1;
    const int sh = max(N - na, 0);
    toff = max(na - N, 0);
    // C, eight columns at a time, reads before writes
    for (int j0 = 0; j0 < N2; j0 += 8) {
        double v[8];
        bool ok[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u;
            const int cj = j < N2 ? __builtin_amdgcn_readlane(rl, (N1 + j) & 63) - nn : 0;
            ok[u] = j < N2 && l < na && l >= cj;
            v[u] = V[ok[u] ? voff + l : 0][ok[u] ? cj : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < N2 && l < na) T[l + sh][N1 + j0 + u] = ok[u] ? v[u] : 0.0;
    }
    // A = F[r1, r1] and B = F[r2, r1] into Li's first N1 columns
    for (int j = 0; j < N1; ++j) {
        const int rj = __builtin_amdgcn_readlane(rl, j);
        const bool ok = l < N && l >= j;
        const double v = Bf[ok ? rl : 0][ok ? rj : 0];
        if (l < N) T[toff + l][j] = ok ? v : 0.0;
    }
    wsync();
    // Reduce C to lower-triangular form by reflections from its last column
    // (dlarfg convention beta = -sign(alpha) ||[alpha; x]||).  Reflection js
    // works on column col = N2 - 1 - js, rows jj .. piv of C; its band is
    // empty for a leading run of js (piv - jj never decreases with js).
    const int rq = __shfl(rl, max(N - 1 - l, 0), 64);
    const unsigned long long emp = __ballot(l < N2 && na - (rq - nn) - (l + 1) == 0);
    const int js_first = ~emp ? __builtin_ctzll(~emp) : 64;
    for (int js = js_first; js < N2; ++js) {
        const int col = N1 + N2 - 1 - js;  // T column
        const int jj = __builtin_amdgcn_readlane(rl, N - 1 - js) - nn;
        const int piv = na - 1 - js;
        const int bl = piv - jj;           // band length
        if (bl == 0) continue;
        if (bl <= SB) {
            // short band: every lane holds it and sums it
            double xb[SB];
            double ss = 0.0;
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                xb[u] = u < bl ? T[min(jj + u, piv) + sh][col] : 0.0;
                ss += xb[u] * xb[u];
            }
            const double xn = sqrt(ss);
            double alpha = T[piv + sh][col];
            if (fabs(alpha) <= PIVOT_NOISE * xn) alpha = 0.0;
            if (xn == 0.0) continue;
            const double beta = -copysign(hypot(alpha, xn), alpha);
            const double tau = (beta - alpha) / beta;
            const double scal = 1.0 / (alpha - beta);
#pragma unroll
            for (int u = 0; u < SB; ++u) xb[u] *= scal;  // the reflector (below its unit pivot)
            // apply to the columns left of col (lane l: column N1 + l)
            const int cl = N1 + l;
            if (cl < col) {
                double rows[SB];
#pragma unroll
                for (int u = 0; u < SB; ++u) rows[u] = u < bl ? T[min(jj + u, piv) + sh][cl] : 0.0;
                const double pv = T[piv + sh][cl];
                double d = pv;
#pragma unroll
                for (int u = 0; u < SB; ++u) d += xb[u] * rows[u];
                const double fct = tau * d;
#pragma unroll
                for (int u = 0; u < SB; ++u)
                    if (u < bl) T[jj + u + sh][cl] = rows[u] - fct * xb[u];
                T[piv + sh][cl] = pv - fct;
            }
            // (the eliminated entries of column col are never read again:
            // they lie above Li22's diagonal)
            if (l == 0) T[piv + sh][col] = beta;
        } else {
            const double x = (l >= jj && l < piv) ? T[l + sh][col] : 0.0;
            const double xn = sqrt(wsum(x * x));
            double alpha = T[piv + sh][col];
            if (fabs(alpha) <= PIVOT_NOISE * xn) alpha = 0.0;
            if (xn == 0.0) continue;
            const double beta = -copysign(hypot(alpha, xn), alpha);
            const double tau = (beta - alpha) / beta;
            const double scal = 1.0 / (alpha - beta);
            wsync();  // every lane has read the pivot
            if (l >= jj && l < piv) T[l + sh][col] = x * scal;
            if (l == piv) T[piv + sh][col] = beta;
            wsync();
            const int cl = N1 + l;
            if (cl < col) {
                double d = T[piv + sh][cl];
                for (int i = jj; i < piv; ++i) d += T[i + sh][col] * T[i + sh][cl];
                const double fct = tau * d;
                for (int i = jj; i < piv; ++i) T[i + sh][cl] -= fct * T[i + sh][col];
                T[piv + sh][cl] -= fct;
            }
        }
        wsync();
    }
    // Li21 from Li22^T Li21 = B
    const double rdg = l < N2 ? 1.0 / T[min(toff + N1 + l, 63)][min(N1 + l, 63)] : 0.0;
    for (int c = 0; c < N1; ++c) {
        double z = l < N2 ? T[toff + N1 + l][c] : 0.0;
        z = tsolve<true>(T, toff + N1, N1, N2, z, rdg);
        if (l < N2) T[toff + N1 + l][c] = z;
    }
    wsync();
    // Li11 = A - Li21^T Li21 (lower), then its reversed Cholesky
    for (int i = 0; i < N1; ++i)
        for (int j = 0; j <= i; ++j) {
            const double s = wsum(l < N2 ? T[toff + N1 + l][i] * T[toff + N1 + l][j] : 0.0);
            if (l == 0) T[toff + i][j] -= s;
        }
    wsync();
    const int bad = rev_chol(T, toff, N1);
    if (bad) {
        if (l == 0) chp_record_failure(status, t.key, bad);
        return false;
    }
    return true;
}

}  // namespace

__global__ __launch_bounds__(64) void k_completion_chain(const ChainItem* __restrict__ items,
                                                         int nitems,
                                                         const FrontInfo* __restrict__ fi,
                                                         BufTab bt, const int* __restrict__ relidx,
                                                         unsigned long long* status) {
    __shared__ Tile Bf[64];          // the front's columns [X11; X21]
    __shared__ Tile Fa[64], Fb[64];  // Lk of the front and the chain child's factor
    __shared__ Tile Ws[64];          // another child's factor; the front's inverse
    Tile* V = Fa;
    Tile* U = Fb;
    int voff = 0;  // Lk(i, j) = V[voff + i][j]
    bool have_v = false;
    const int l = threadIdx.x;
    double* blkval = bt.p[0];
    double* arena = bt.p[1];
    for (int it = 0; it < nitems;) {
        const ChainItem ci = items[it];
        const FrontInfo f = fi[ci.front];
        const int nn = f.nn, na = f.na, nj = nn + na, nch = ci.nch;
        double* blk = blkval + f.blkptr;
        for (int j0 = 0; j0 < nn; j0 += 8) {  // the front's columns, lower part
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool ok = l < nj && j0 + u < nn && l >= j0 + u;
                v[u] = blk[ok ? l + (i64)(j0 + u) * nj : 0];
                v[u] = ok ? v[u] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < nn) Bf[l][j0 + u] = v[u];
        }
        if (!have_v) {  // the chain's first front: Lk from its update slot
            const double* Lk = arena + f.updptr;
            voff = 0;
            for (int j0 = 0; j0 < na; j0 += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool ok = l < na && j0 + u < na && l >= j0 + u;
                    v[u] = Lk[ok ? l + (i64)(j0 + u) * na : 0];
                    v[u] = ok ? v[u] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (j0 + u < na) V[l][j0 + u] = v[u];
            }
        }
        wsync();
        // the children's update factors (completion.c:48-66)
        bool keep = false;
        int voff_next = 0;
        for (int c = 1; c <= nch; ++c) {
            const ChainItem cu = items[it + c];
            const bool chain = cu.kind == 2;
            Tile* T = chain ? U : Ws;
            int toff = 0;
            if (!child_factor(cu.uf, V, voff, Bf, T, toff, relidx, status)) return;
            wsync();
            if (chain) {
                keep = true;
                voff_next = toff;
            } else {  // the factor goes to the child's update slot (zero above the diagonal)
                const int N = cu.uf.N;
                double* Li = arena + cu.uf.outoff;
                for (int j = 0; j < N; ++j)
                    if (l < N) Li[l + (i64)j * N] = l >= j ? T[toff + l][j] : 0.0;
                wsync();
            }
        }
        it += 1 + nch;
        if (na > 0) {
            const int vr = min(voff + l, 63);
            const unsigned long long zb = __ballot(l < na && V[vr][l] == 0.0);
            if (zb) {
                if (l == 0) chp_record_failure(status, ci.front, __builtin_ctzll(zb) + 1);
                return;
            }
            // X21 <- Lk^{-T} X21, X11 -= X21^T X21, X21 <- -Lk^{-1} X21
            const double rdg = l < na ? 1.0 / V[vr][l] : 0.0;
            for (int c = 0; c < nn; ++c) {
                double z = l < na ? Bf[nn + l][c] : 0.0;
                z = tsolve<true>(V, voff, 0, na, z, rdg);
                if (l < na) Bf[nn + l][c] = z;
            }
            wsync();
            if (nn <= 4) {
                for (int i = 0; i < nn; ++i)
                    for (int j = 0; j <= i; ++j) {
                        const double s = wsum(l < na ? Bf[nn + l][i] * Bf[nn + l][j] : 0.0);
                        if (l == 0) Bf[i][j] -= s;
                    }
            } else if (l < nn) {
                for (int j = 0; j <= l; ++j) {
                    double s0 = 0.0, s1 = 0.0;
                    int q = 0;
                    for (; q + 1 < na; q += 2) {
                        s0 += Bf[nn + q][l] * Bf[nn + q][j];
                        s1 += Bf[nn + q + 1][l] * Bf[nn + q + 1][j];
                    }
                    if (q < na) s0 += Bf[nn + q][l] * Bf[nn + q][j];
                    Bf[l][j] -= s0 + s1;
                }
            }
            wsync();
            for (int c = 0; c < nn; ++c) {
                double w = l < na ? Bf[nn + l][c] : 0.0;
                w = tsolve<false>(V, voff, 0, na, w, rdg);
                if (l < na) Bf[nn + l][c] = -w;
            }
            wsync();
        }
        // X11 = Lr^T Lr (reversed Cholesky), L11 = Lr^{-1} into Ws
        const int bad = rev_chol(Bf, 0, nn);
        if (bad) {
            if (l == 0) chp_record_failure(status, ci.front, bad);
            return;
        }
        if (l < nn) {
            Ws[l][l] = 1.0 / Bf[l][l];
            for (int i = l + 1; i < nn; ++i) {
                double v = 0.0;
                for (int q = l; q < i; ++q) v -= Bf[i][q] * Ws[q][l];
                Ws[i][l] = v / Bf[i][i];
            }
        }
        wsync();
        for (int j = 0; j < nn; ++j) {
            if (l < nn && l >= j) blk[l + (i64)j * nj] = Ws[l][j];
            if (l >= nn && l < nj) blk[l + (i64)j * nj] = Bf[l][j];
        }
        wsync();
        if (keep) {  // the chain child's factor is the next front's Lk
            Tile* x = V;
            V = U;
            U = x;
            voff = voff_next;
        }
        have_v = keep;
    }
}
