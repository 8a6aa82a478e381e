// Completion and Hessian sweeps (completion.c, hessian.c, update_factor.c of
// the reference) on the level schedule of a plan.
//
// Every front-local dense operation is expressed on descriptors and batched
// over all fronts of a level:
//   * Cholesky of a set of matrices: 64 x 64 diagonal-block factorizations
//     (k_diag_tasks, mode 0), in-place panel products with the block inverse
//     and trailing updates (k_gemm_tiles);
//   * inverse of a set of lower-triangular matrices: diagonal-block inverses
//     (mode 1) and recursive halving X21 = -X22 (L21 X11) on MFMA;
//   * triangular solves and products: GEMM tasks with a triangular operand
//     (a solve multiplies by the explicit inverse);
//   * the reversed Cholesky factor of the completion (S = Lr^T Lr): reversed
//     transposed copy, Cholesky, reversed transposed copy;
//   * update matrices: k_extract (unfactored) or k_update_factor
//     (Householder-factored updates), issued at the parent's level before
//     the parent's own (2,2) block is factored in place.
// The sweeps run on their own program (work lists, tasks) and scratch; the
// update arena is the plan's.
#include "plan.h"
#include "../../include/chompack_amd.h"

#include <algorithm>
#include <array>
#include <functional>

namespace chpi {

template <class T>
void push_tasks(chp_plan* P, std::vector<Launch>& seq, Kid kid, const std::vector<T>& v,
                int bsel) {
    if (v.empty()) return;
    Launch L{kid, (int)v.size(), 256, 0, P->pg->lists.add(v), 0};
    L.bsel = bsel;
    seq.push_back(L);
}

void add_matop(std::vector<MatOp>& ops, MatOpKind kind, int sbuf, i64 src, int lds, int dbuf,
               i64 dst, int ldd, int m, int n, double alpha) {
    for (int j0 = 0; j0 < n; j0 += 64)
        for (int i0 = 0; i0 < m; i0 += 64) {
            const bool above = i0 + 63 < j0;  // every element has i < j
            const bool below = i0 > j0 + 63;  // every element has i > j
            if (kind == MO_COPY_LOWER && above) continue;
            if ((kind == MO_SYMM || kind == MO_ZERO_UPPER || kind == MO_PHI) && below) continue;
            MatOp o;
            std::memset(&o, 0, sizeof(o));
            o.src = src; o.dst = dst; o.lds = lds; o.ldd = ldd; o.m = m; o.n = n;
            o.i0 = i0; o.j0 = j0;
            o.sbuf = (unsigned char)sbuf; o.dbuf = (unsigned char)dbuf;
            o.kind = (unsigned char)kind; o.alpha = alpha;
            ops.push_back(o);
        }
}

// In-place lower Cholesky of each matrix (right-looking, 64-wide blocks).
void emit_potrf(chp_plan* P, std::vector<Launch>& seq, const std::vector<Mat>& ms,
                const std::vector<int>& keys, Bump& dv) {
    int maxn = 0;
    std::vector<i64> dvo(ms.size());
    for (size_t q = 0; q < ms.size(); ++q) {
        maxn = std::max(maxn, ms[q].n);
        dvo[q] = dv.take((i64)((ms[q].n + NB - 1) / NB) * NB * NB);
    }
    for (int c0 = 0; c0 < maxn; c0 += NB) {
        std::vector<DiagTask> dt;
        std::vector<GemmTile> pt, ut;
        for (size_t q = 0; q < ms.size(); ++q) {
            const Mat& m = ms[q];
            if (m.n <= c0) continue;
            const int c1 = std::min(c0 + NB, m.n);
            DiagTask t;
            std::memset(&t, 0, sizeof(t));
            t.off = m.off + c0 + (i64)c0 * m.ld;
            t.outoff = dvo[q] + (i64)(c0 / NB) * NB * NB;
            t.ld = m.ld; t.ldo = NB; t.nb = c1 - c0; t.key = keys[q]; t.col = c0;
            t.buf = (unsigned char)m.buf; t.obuf = 3; t.mode = 0;
            dt.push_back(t);
            if (c1 < m.n) {
                const i64 pan = m.off + c1 + (i64)c0 * m.ld;
                // L[c1:, c0:c1] = A[c1:, c0:c1] Dinv^T (in place; Dinv row-major)
                add_gemm_tiles(P, pt, mk(m.buf, pan, m.ld, 0, 3, t.outoff, NB, 0, m.buf, pan, m.ld,
                                         m.n - c1, c1 - c0, c1 - c0, 1.0, 0.0, 0), false);
                // A[c1:, c1:] -= L[c1:, c0:c1] L[c1:, c0:c1]^T (lower)
                add_gemm_tiles(P, ut, mk(m.buf, pan, m.ld, 0, m.buf, pan, m.ld, 1, m.buf,
                                         m.off + c1 + (i64)c1 * m.ld, m.ld, m.n - c1, m.n - c1,
                                         c1 - c0, -1.0, 1.0, 1), true);
            }
        }
        push_tasks(P, seq, K_DIAG, dt);
        push_gemm(P, seq, pt);
        push_gemm(P, seq, ut);
    }
}

// X = L^{-1} for lower-triangular L (only its lower triangle is read); X is
// written in full (zero above the diagonal).  W: scratch n x n (ld n) each.
void emit_trtri(chp_plan* P, std::vector<Launch>& seq, const std::vector<Mat>& Ls,
                const std::vector<Mat>& Xs, const std::vector<i64>& Ws,
                const std::vector<int>& keys) {
    std::vector<MatOp> z;
    std::vector<DiagTask> dt;
    for (size_t q = 0; q < Ls.size(); ++q) {
        const Mat &L = Ls[q], &X = Xs[q];
        const int n = L.n;
        for (int j0 = 64; j0 < n; j0 += 64)
            for (int i0 = 0; i0 < j0; i0 += 64) {
                MatOp o;
                std::memset(&o, 0, sizeof(o));
                o.dst = X.off; o.ldd = X.ld; o.m = n; o.n = n; o.i0 = i0; o.j0 = j0;
                o.dbuf = (unsigned char)X.buf; o.sbuf = (unsigned char)X.buf; o.kind = MO_ZERO;
                z.push_back(o);
            }
        for (int b = 0; b * NB < n; ++b) {
            DiagTask t;
            std::memset(&t, 0, sizeof(t));
            t.off = L.off + (i64)b * NB * (1 + (i64)L.ld);
            t.outoff = X.off + (i64)b * NB * (1 + (i64)X.ld);
            t.ld = L.ld; t.ldo = X.ld; t.nb = std::min(NB, n - b * NB);
            t.key = keys[q]; t.col = b * NB;
            t.buf = (unsigned char)L.buf; t.obuf = (unsigned char)X.buf; t.mode = 1;
            dt.push_back(t);
        }
    }
    push_tasks(P, seq, K_MATOP, z);
    push_tasks(P, seq, K_DIAG, dt);
    std::vector<std::vector<std::array<int, 4>>> byh;  // (q, a, m, b) in blocks
    std::function<int(int, int, int)> tree = [&](int q, int a, int b) -> int {
        if (b - a <= 1) return 0;
        int m = (a + b) / 2;
        int h = 1 + std::max(tree(q, a, m), tree(q, m, b));
        if ((int)byh.size() < h) byh.resize(h);
        byh[h - 1].push_back({q, a, m, b});
        return h;
    };
    for (size_t q = 0; q < Ls.size(); ++q) tree((int)q, 0, (Ls[q].n + NB - 1) / NB);
    for (auto& nodes : byh) {
        std::vector<GemmTile> g1, g2;
        for (auto& nd : nodes) {
            const Mat &L = Ls[nd[0]], &X = Xs[nd[0]];
            const int n = L.n;
            const i64 W = Ws[nd[0]];
            const int r1 = nd[1] * NB, r2 = nd[2] * NB, r3 = std::min(nd[3] * NB, n);
            const int m1 = r2 - r1, m2 = r3 - r2;
            // W[r2:r3, r1:r2] = L[r2:r3, r1:r2] X[r1:r2, r1:r2]   (X11 lower)
            add_gemm_tiles(P, g1, mk(L.buf, L.off + r2 + (i64)r1 * L.ld, L.ld, 0,
                                     X.buf, X.off + r1 + (i64)r1 * X.ld, X.ld, 0,
                                     2, W + r2 + (i64)r1 * n, n, m2, m1, m1, 1.0, 0.0, 0, 0, 1),
                           false);
            // X[r2:r3, r1:r2] = -X[r2:r3, r2:r3] W[r2:r3, r1:r2]   (X22 lower)
            add_gemm_tiles(P, g2, mk(X.buf, X.off + r2 + (i64)r2 * X.ld, X.ld, 0,
                                     2, W + r2 + (i64)r1 * n, n, 0,
                                     X.buf, X.off + r2 + (i64)r1 * X.ld, X.ld, m2, m1, m2,
                                     -1.0, 0.0, 0, 1, 0), false);
        }
        push_gemm(P, seq, g1);
        push_gemm(P, seq, g2);
    }
}

template void push_tasks<DiagTask>(chp_plan*, std::vector<Launch>&, Kid, const std::vector<DiagTask>&, int);
template void push_tasks<MatOp>(chp_plan*, std::vector<Launch>&, Kid, const std::vector<MatOp>&, int);

namespace {

// Update-factor task of child c from its parent's front (buffer fbuf) and the
// factor of the parent's update matrix; the workspace offset is left at 0.
UFTask uf_task(const chp_plan* P, int c, int fbuf) {
    const FrontInfo& g = P->fi[c];
    const FrontInfo& f = P->fi[g.parent];
    UFTask t;
    std::memset(&t, 0, sizeof(t));
    t.foff = f.blkptr; t.lkoff = f.updptr; t.outoff = g.updptr;
    t.relpos = g.relptr;
    t.nn = f.nn; t.na = f.na; t.N = g.na; t.key = g.parent;
    t.fbuf = (unsigned char)fbuf;
    return t;
}

// Update matrices of the children of the fronts `fr`, read from the fronts
// [blk (buffer fbuf); V or its factor (update arena)].
// Blocks of consecutive reflections of a task with many of them (k_uf_step):
// up to UF_BMAX reflections whose bands, rows lo .. na - 1 - js0, fit UF_VCAP
// doubles together; runs of reflections with empty bands are left out.  The
// blocks' T go to the two slots at toff in turn.
std::vector<UFBlock> uf_blocks(const chp_plan* P, const UFTask& t, int n1, i64 toff) {
    const int* r = P->h_relidx.data() + t.relpos;
    const int N = t.N, N2 = N - n1, nn = t.nn, na = t.na;
    std::vector<UFBlock> out;
    int js = 0;
    while (js < N2) {
        const int start = js, hi = na - 1 - start;
        int lo = INT32_MAX, cnt = 0;
        while (js < N2 && cnt < UF_BMAX) {
            const int jj = r[N - 1 - js] - nn, piv = na - 1 - js;
            const int nlo = piv - jj > 0 ? std::min(lo, jj) : lo;
            const int nlen = nlo == INT32_MAX ? 1 : hi - nlo + 1;
            if (cnt > 0 && (i64)nlen * (cnt + 1) > UF_VCAP) break;
            lo = nlo;
            ++cnt;
            ++js;
        }
        if (lo == INT32_MAX) continue;
        UFBlock B;
        std::memset(&B, 0, sizeof(B));
        B.woff = t.woff; B.relpos = t.relpos;
        B.toff = toff + (out.size() & 1) * (UF_BMAX * UF_BMAX + 2 * UF_BMAX);
        B.nn = nn; B.na = na; B.N = N; B.N2 = N2;
        B.js0 = start; B.nbk = cnt; B.lo = lo; B.len = hi - lo + 1;
        B.colmin = N2 - start - cnt; B.key = t.key;
        out.push_back(B);
    }
    return out;
}

void emit_child_updates(chp_plan* P, std::vector<Launch>& seq, const std::vector<int>& fr,
                        bool factored, int fbuf, Bump& sc, Bump& dv) {
    const std::vector<int> ch = children_of(P, fr);
    if (ch.empty()) return;
    if (!factored) {
        const std::vector<int4> xt = extract_tiles(P, ch);
        push_tasks(P, seq, K_EXTRACT, xt, fbuf);
        return;
    }
    std::vector<UFTask> ut, us;
    // Large tasks leave the kernel early and finish on the whole device:
    //   * N1 >= UF_SOLVE_N1 and N2 >= UF_SOLVE_N2: X = Li22^{-1} (blocked
    //     triangular inverse), Z = X^T B (= Li21), Li11 = A - Z^T Z;
    //   * those and all N1 >= UF_BLOCKED_N1: the reversed Cholesky of Li11 as
    //     J A J = M M^T by the blocked potrf (J the reversal), Li11 = J M^T J,
    //     through reversed-transpose copies.
    // Smaller tasks are latency chains that stay inside the kernel.
    constexpr int UF_BLOCKED_N1 = 192, UF_SOLVE_N1 = 1, UF_SOLVE_N2 = 64;
    // tasks with at least UF_EXT_N2 reflections reduce C on the whole device
    // (k_uf_setup, k_uf_step) before k_update_factor's tail
    constexpr int UF_EXT_N2 = 64;
    std::vector<UFTask> xt;
    std::vector<int4> xcols;
    std::vector<std::vector<UFBlock>> xblocks;
    std::vector<MatOp> rt1, rt2;
    std::vector<Mat> R, L22s, Xs;
    std::vector<i64> Ws;
    std::vector<int> rkeys, skeys;
    std::vector<GemmTile> gz, gs;
    for (int c : ch) {
        const FrontInfo& g = P->fi[c];
        const FrontInfo& f = P->fi[g.parent];
        int n1 = 0;
        while (n1 < g.na && P->h_relidx[g.relptr + n1] < f.nn) ++n1;
        UFTask t = uf_task(P, c, fbuf);
        t.woff = sc.take((i64)f.na * (g.na - n1));
        const int N = g.na, n2 = N - n1;
        if (n1 >= UF_SOLVE_N1 && n2 >= UF_SOLVE_N2) {
            t.defer_solve = 1;
            const Mat L22{1, g.updptr + n1 + (i64)n1 * N, N, n2};
            const Mat X{2, sc.take((i64)n2 * n2), n2, n2};
            const i64 z = sc.take((i64)n2 * n1);
            L22s.push_back(L22);
            Xs.push_back(X);
            Ws.push_back(sc.take((i64)n2 * n2));
            skeys.push_back(g.parent);
            // Z = X^T B (B = Li21 as gathered), Li11 -= Z^T Z (lower), Li21 = Z
            add_gemm_tiles(P, gz, mk(2, X.off, n2, 1, 1, g.updptr + n1, N, 0, 2, z, n2, n2, n1, n2,
                                     1.0, 0.0, 0, 2), false);
            add_gemm_tiles(P, gs, mk(2, z, n2, 1, 2, z, n2, 0, 1, g.updptr, N, n1, n1, n2, -1.0,
                                     1.0, 1), true);
            add_matop(rt1, MO_COPY, 2, z, n2, 1, g.updptr + n1, N, n2, n1);
        }
        if (n1 >= UF_BLOCKED_N1 || t.defer_solve) {
            t.defer_rchol = 1;
            const i64 rs = sc.take((i64)n1 * n1);
            add_matop(rt1, MO_RT, 1, g.updptr, g.na, 2, rs, n1, n1, n1);
            add_matop(rt2, MO_RT, 2, rs, n1, 1, g.updptr, g.na, n1, n1);
            R.push_back({2, rs, n1, n1});
            rkeys.push_back(g.parent);
        }
        // tasks that fit one wave's LDS tiles run in 64-thread workgroups
        const bool small = N <= 64 && f.na <= 64 && !t.defer_solve && !t.defer_rchol;
        if (!small && n2 >= UF_EXT_N2 && f.na <= UF_VCAP) {
            t.ext_reduce = 1;
            const int ti = (int)xt.size();
            xt.push_back(t);
            for (int j = 0; j < N + n2; ++j) xcols.push_back(make_int4(ti, j, n1, 0));
            xblocks.push_back(uf_blocks(P, t, n1, sc.take(2 * (UF_BMAX * UF_BMAX + 2 * UF_BMAX))));
        }
        (small ? us : ut).push_back(t);
    }
    if (!xt.empty()) {
        std::vector<char> lb(xt.size() * sizeof(UFTask) + xcols.size() * sizeof(int4));
        std::memcpy(lb.data(), xt.data(), xt.size() * sizeof(UFTask));
        std::memcpy(lb.data() + xt.size() * sizeof(UFTask), xcols.data(), xcols.size() * sizeof(int4));
        seq.push_back(Launch{K_UF_SETUP, (int)xcols.size(), 256, 0, P->pg->lists.add(lb),
                             (int)xt.size()});
        // step s: block s of each task (role 1, first in the list: the
        // critical path) and block s - 1 applied to the columns left of it
        // but for block s's own (role 0)
        size_t nbl = 0;
        for (const auto& v : xblocks) nbl = std::max(nbl, v.size());
        for (size_t s = 0; s <= nbl; ++s) {
            std::vector<UFBlock> items, tiles;
            for (const auto& v : xblocks) {
                if (s < v.size()) {
                    UFBlock B = v[s];
                    B.role = 1;
                    if (s > 0) {
                        const UFBlock& Q = v[s - 1];
                        B.pjs0 = Q.js0; B.pnbk = Q.nbk; B.plo = Q.lo; B.plen = Q.len;
                        B.ptoff = Q.toff;
                    }
                    items.push_back(B);
                }
                if (s == 0 || s > v.size()) continue;
                UFBlock T = v[s - 1];
                T.role = 0;
                if (s < v.size()) {
                    T.xlo = v[s].colmin;
                    T.xhi = v[s].colmin + v[s].nbk;
                }
                for (int c0 = 0; c0 < T.colmin; c0 += 16) {
                    if (c0 >= T.xlo && c0 + 16 <= T.xhi) continue;  // all in the excluded range
                    T.c0 = c0;
                    tiles.push_back(T);
                }
            }
            items.insert(items.end(), tiles.begin(), tiles.end());
            push_tasks(P, seq, K_UF_STEP, items, 0);
        }
    }
    push_tasks(P, seq, K_UPDATE_FACTOR, us);
    if (!us.empty()) seq.back().block = 64;
    push_tasks(P, seq, K_UPDATE_FACTOR, ut);
    if (!ut.empty()) seq.back().block = 1024;  // k_update_factor runs 16 waves
    if (!L22s.empty()) {
        emit_trtri(P, seq, L22s, Xs, Ws, skeys);
        push_gemm(P, seq, gz);
        push_gemm(P, seq, gs);
    }
    if (!R.empty()) {
        push_tasks(P, seq, K_MATOP, rt1);
        emit_potrf(P, seq, R, rkeys, dv);
        push_tasks(P, seq, K_MATOP, rt2);
    }
}

// One level of completion (completion.c:24-118), root to leaves.
void build_completion_level(chp_plan* P, std::vector<Launch>& seq, const std::vector<int>& fr,
                            bool fu, Bump& sc, Bump& dv) {
    emit_child_updates(P, seq, fr, fu, 0, sc, dv);
    std::vector<int> fa;
    for (int k : fr) if (P->fi[k].na > 0) fa.push_back(k);
    if (!fa.empty()) {
        std::vector<Mat> V, Xv;
        std::vector<i64> Wv, Zo;
        for (int k : fa) {
            const FrontInfo& f = P->fi[k];
            V.push_back({1, f.updptr, f.na, f.na});
            Xv.push_back({2, sc.take((i64)f.na * f.na), f.na, f.na});
            Wv.push_back(sc.take((i64)f.na * f.na));
            Zo.push_back(sc.take((i64)f.na * f.nn));
        }
        if (!fu) emit_potrf(P, seq, V, fa, dv);  // V_k = Lv Lv^T
        emit_trtri(P, seq, V, Xv, Wv, fa);
        std::vector<GemmTile> gz, gs;
        for (size_t q = 0; q < fa.size(); ++q) {
            const FrontInfo& f = P->fi[fa[q]];
            const int nn = f.nn, na = f.na, nj = nn + na;
            // Z = Lv^{-1} X21 (unfactored) or Lk^{-T} X21 (factored, V_k = Lk^T Lk)
            add_gemm_tiles(P, gz, mk(2, Xv[q].off, na, fu ? 1 : 0, 0, f.blkptr + nn, nj, 0,
                                     2, Zo[q], na, na, nn, na, 1.0, 0.0, 0, fu ? 2 : 1), false);
            // X11 -= Z^T Z (lower) and X21 = -V_k^{-1} X21
            add_gemm_tiles(P, gs, mk(2, Zo[q], na, 1, 2, Zo[q], na, 0, 0, f.blkptr, nj,
                                     nn, nn, na, -1.0, 1.0, 1), true);
            add_gemm_tiles(P, gs, mk(2, Xv[q].off, na, fu ? 0 : 1, 2, Zo[q], na, 0,
                                     0, f.blkptr + nn, nj, na, nn, na, -1.0, 0.0, 0,
                                     fu ? 1 : 2), false);
        }
        push_gemm(P, seq, gz);
        push_gemm(P, seq, gs);
    }
    // L11 = Lr^{-1} with S11 = Lr^T Lr (completion.c:93-116)
    std::vector<MatOp> rt1, rt2, cp;
    std::vector<Mat> T, T2, X11;
    std::vector<i64> W11;
    std::vector<int> keys;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        const int nn = f.nn, nj = nn + f.na;
        const i64 t1 = sc.take((i64)nn * nn), t2 = sc.take((i64)nn * nn);
        const i64 x = sc.take((i64)nn * nn), w = sc.take((i64)nn * nn);
        T.push_back({2, t1, nn, nn});
        T2.push_back({2, t2, nn, nn});
        X11.push_back({2, x, nn, nn});
        W11.push_back(w);
        keys.push_back(k);
        add_matop(rt1, MO_RT, 0, f.blkptr, nj, 2, t1, nn, nn, nn);
        add_matop(rt2, MO_RT, 2, t1, nn, 2, t2, nn, nn, nn);
        add_matop(cp, MO_COPY_LOWER, 2, x, nn, 0, f.blkptr, nj, nn, nn);
    }
    push_tasks(P, seq, K_MATOP, rt1);
    emit_potrf(P, seq, T, keys, dv);
    push_tasks(P, seq, K_MATOP, rt2);
    emit_trtri(P, seq, T2, X11, W11, keys);
    push_tasks(P, seq, K_MATOP, cp);
}

// One level of the scaling map (hessian.c:201-352): shared part (update
// matrices, factors, inverses) and the per-direction part.
void build_scale_level(chp_plan* P, std::vector<Launch>& sh, std::vector<Launch>& sht,
                       std::vector<Launch>& dir, const std::vector<int>& fr, bool fu, bool inv,
                       bool adj, Bump& sc, Bump& dv, int nb, i64 bstride) {
    emit_child_updates(P, sh, fr, fu, 6, sc, dv);
    std::vector<int> fa;
    for (int k : fr) if (P->fi[k].na > 0) fa.push_back(k);
    std::vector<Mat> V, Xv;
    std::vector<i64> Wv;
    for (int k : fa) {
        const FrontInfo& f = P->fi[k];
        V.push_back({1, f.updptr, f.na, f.na});
        if (inv) {
            Xv.push_back({2, sc.take((i64)f.na * f.na), f.na, f.na});
            Wv.push_back(sc.take((i64)f.na * f.na));
        }
    }
    if (!fa.empty() && !fu) emit_potrf(P, sh, V, fa, dv);
    if (!fa.empty() && inv) emit_trtri(P, sht, V, Xv, Wv, fa);
    // M1 = L11^{-1} (trsm, !inv) or L11 itself (trmm, inv)
    std::vector<Mat> L11, X11;
    std::vector<i64> W11;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        L11.push_back({5, f.blkptr, f.nn + f.na, f.nn});
        if (!inv) {
            X11.push_back({2, sc.take((i64)f.nn * f.nn), f.nn, f.nn});
            W11.push_back(sc.take((i64)f.nn * f.nn));
        }
    }
    if (!inv) emit_trtri(P, sht, L11, X11, W11, std::vector<int>(fr.begin(), fr.end()));
    // per direction (nb of them, blkval bstride doubles apart): U11
    // symmetric, Tu = U op1(M1), U11 = op2(M1) Tu11 (upper zeroed),
    // U21 = op3(M3) Tu21
    std::vector<MatOp> sy;
    std::vector<GemmTile> g1, g2;
    const bool t1 = !adj, t2 = adj;                       // op1 / op2 transposed
    const bool t3 = fu ? adj : !adj;                      // op3 transposed
    for (int u = 0; u < nb; ++u) {
        const i64 ub = (i64)u * bstride;
        size_t qa = 0;
        for (size_t q = 0; q < fr.size(); ++q) {
            const int k = fr[q];
            const FrontInfo& f = P->fi[k];
            const int nn = f.nn, na = f.na, nj = nn + na;
            const Mat M1 = inv ? L11[q] : X11[q];
            const i64 tu = sc.take((i64)nj * nn);
            add_matop(sy, MO_SYMM, 0, ub + f.blkptr, nj, 0, ub + f.blkptr, nj, nn, nn);
            add_gemm_tiles(P, g1, mk(0, ub + f.blkptr, nj, 0, M1.buf, M1.off, M1.ld, t1 ? 1 : 0,
                                     2, tu, nj, nj, nn, nn, 1.0, 0.0, 0, 0, t1 ? 2 : 1), false);
            add_gemm_tiles(P, g2, mk(M1.buf, M1.off, M1.ld, t2 ? 1 : 0, 2, tu, nj, 0,
                                     0, ub + f.blkptr, nj, nn, nn, nn, 1.0, 0.0, 2, t2 ? 2 : 1),
                           false);
            if (na > 0) {
                const Mat M3 = inv ? Xv[qa] : V[qa];
                add_gemm_tiles(P, g2, mk(M3.buf, M3.off, M3.ld, t3 ? 1 : 0, 2, tu + nn, nj, 0,
                                         0, ub + f.blkptr + nn, nj, na, nn, na, 1.0, 0.0, 0,
                                         t3 ? 2 : 1),
                               false);
                ++qa;
            }
        }
    }
    push_tasks(P, dir, K_MATOP, sy);
    push_gemm(P, dir, g1);
    push_gemm(P, dir, g2);
}

// Copies of a GEMM task for nb directions: operands on a per-direction
// buffer (str[buf] != 0) move str[buf] doubles per direction.
void add_gemm_dirs(chp_plan* P, std::vector<GemmTile>& tiles, const GemmTask& t,
                   bool skip_upper, int nb, const i64* str) {
    for (int u = 0; u < nb; ++u) {
        GemmTask c = t;
        c.offA += u * str[c.bufA];
        c.offB += u * str[c.bufB];
        c.offC += u * str[c.bufC];
        add_gemm_tiles(P, tiles, c, skip_upper);
    }
}

// The extend-add / extract launches appended to seq since `from` cover nb
// directions (grid y).
void mark_dirs(std::vector<Launch>& seq, size_t from, int nb, const i64* str) {
    for (size_t i = from; i < seq.size(); ++i)
        if (seq[i].kid == K_EXTADD || seq[i].kid == K_EXTRACT) {
            seq[i].gy = nb;
            seq[i].bstride = str[0];
            seq[i].ustride = str[1];
        }
}

// One level of _Y2K (hessian.c:3-101), leaves to root, for nb directions:
// str[0] / str[1] = the distance of consecutive directions' blkval / update
// arena (buffers 0 and 1).
void build_y2k_level(chp_plan* P, std::vector<Launch>& seq, const std::vector<int>& fr, bool inv,
                     int nb, const i64* str) {
    const double alpha = inv ? 1.0 : -1.0;
    int net = 0;
    const std::vector<int4> et = extadd_tiles(P, fr, &net);
    const size_t from = seq.size();
    if (!inv) {
        Launch L{K_EXTADD, net, 256, 0, P->pg->lists.add(et), 1};
        seq.push_back(L);
    } else {
        std::vector<MatOp> z;
        for (int u = 0; u < nb; ++u)
            for (int k : fr) {
                const FrontInfo& f = P->fi[k];
                const i64 o = (i64)u * str[1] + f.updptr;
                if (f.na > 0) add_matop(z, MO_ZERO, 1, o, f.na, 1, o, f.na, f.na, f.na);
            }
        push_tasks(P, seq, K_MATOP, z);
    }
    std::vector<GemmTile> g1, g2, g3;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        const int nn = f.nn, na = f.na, nj = nn + na;
        if (na == 0) continue;
        const i64 b21 = f.blkptr + nn;
        // F22 += alpha L21 F21^T
        add_gemm_dirs(P, g1, mk(5, b21, nj, 0, 0, b21, nj, 1, 1, f.updptr, na,
                                na, na, nn, alpha, 1.0, 1), true, nb, str);
        // F21 += alpha L21 sym(F11)
        add_gemm_dirs(P, g2, mk(5, b21, nj, 0, 0, f.blkptr, nj, 2, 0, b21, nj,
                                na, nn, nn, alpha, 1.0, 0), false, nb, str);
        // F22 += alpha F21 L21^T
        add_gemm_dirs(P, g3, mk(0, b21, nj, 0, 5, b21, nj, 1, 1, f.updptr, na,
                                na, na, nn, alpha, 1.0, 1), true, nb, str);
    }
    push_gemm(P, seq, g1);
    push_gemm(P, seq, g2);
    push_gemm(P, seq, g3);
    if (inv) {
        Launch L{K_EXTADD, net, 256, 0, P->pg->lists.add(et), 0};
        seq.push_back(L);
    }
    mark_dirs(seq, from, nb, str);
}

// One level of _M2T (hessian.c:103-198), root to leaves, for nb directions
// (str as for _Y2K).
void build_m2t_level(chp_plan* P, std::vector<Launch>& seq, const std::vector<int>& fr, bool inv,
                     Bump& sc, int nb, const i64* str) {
    const double alpha = inv ? 1.0 : -1.0;
    const size_t from = seq.size();
    if (inv) {
        Bump dv;  // unused: unfactored updates need no block inverses
        emit_child_updates(P, seq, fr, false, 0, sc, dv);  // before this level's update
    }
    std::vector<GemmTile> g1, g2, g3;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        const int nn = f.nn, na = f.na, nj = nn + na;
        if (na == 0) continue;
        const i64 b21 = f.blkptr + nn;
        // F11 += alpha F21^T L21 (lower)
        add_gemm_dirs(P, g1, mk(0, b21, nj, 1, 5, b21, nj, 0, 0, f.blkptr, nj,
                                nn, nn, na, alpha, 1.0, 1), true, nb, str);
        // F21 += alpha V_k L21 (V_k stored with both triangles)
        add_gemm_dirs(P, g2, mk(1, f.updptr, na, 0, 5, b21, nj, 0, 0, b21, nj,
                                na, nn, na, alpha, 1.0, 0), false, nb, str);
        // F11 += alpha L21^T F21 (lower)
        add_gemm_dirs(P, g3, mk(5, b21, nj, 1, 0, b21, nj, 0, 0, f.blkptr, nj,
                                nn, nn, na, alpha, 1.0, 1), true, nb, str);
    }
    push_gemm(P, seq, g1);
    push_gemm(P, seq, g2);
    push_gemm(P, seq, g3);
    // without inv: the children's V from the finished fronts of this level,
    // here at the parents' level (an update slot is touched only at its
    // front's level and its parent's)
    if (!inv) push_tasks(P, seq, K_EXTRACT, extract_tiles(P, children_of(P, fr)), 0);
    mark_dirs(seq, from, nb, str);
}

// One level of the Cholesky adjoint sweep (the reverse of cholesky.c:26-82,
// root to leaves).  Buffers: 0 the output Fbar (blkval-shaped), 5 the factor
// [L11; B21], 6 the gradient [Gb11; Gb21] of the stored factor, 1 the update
// arena, whose slot k holds Ub_k (both triangles) from the parent's level.
// Per front, X11 = L11^{-1}, L21 = B21 L11, T = Gb21 X11^T:
//   M = tril(Gb11) - B21^T T;  T -= 2 Ub L21;  Fbar21 = T X11 / 2;
//   M -= 2 Fbar21^T L21;  P = Phi(L11^T tril(M));  S = X11^T (P X11);
//   Fbar11 = lower((S + S^T) / 2);
// then the children's Ub_c = Fbar[r_c, r_c] (k_extract on buffer 0 and the
// arena).  The gradient with respect to X's lower-stored entries is Fbar
// with the strictly-lower entries doubled.
void build_chol_adjoint_level(chp_plan* P, std::vector<Launch>& seq, const std::vector<int>& fr,
                              Bump& sc) {
    std::vector<Mat> Ls, Xs;
    std::vector<i64> Ws;
    std::vector<int> keys;
    std::vector<MatOp> m0, phi, sym;
    std::vector<GemmTile> ga, gb, gc, gd, ge, gf, gg, gh;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        const int nn = f.nn, na = f.na, nj = nn + na;
        const i64 bp = f.blkptr;
        const i64 nn2 = (i64)nn * nn;
        const i64 x = sc.take(nn2), w = sc.take(nn2), m = sc.take(nn2), p = sc.take(nn2),
                  w2 = sc.take(nn2), s = sc.take(nn2);
        Ls.push_back({5, bp, nj, nn});
        Xs.push_back({2, x, nn, nn});
        Ws.push_back(w);
        keys.push_back(k);
        add_matop(m0, MO_COPY_LOWER_Z, 6, bp, nj, 2, m, nn, nn, nn);
        if (na > 0) {
            const i64 l21 = sc.take((i64)na * nn), t = sc.take((i64)na * nn);
            // L21 = B21 L11 (L11 lower), T = Gb21 X11^T (X11^T upper)
            add_gemm_tiles(P, ga, mk(5, bp + nn, nj, 0, 5, bp, nj, 0, 2, l21, na,
                                     na, nn, nn, 1.0, 0.0, 0, 0, 1), false);
            add_gemm_tiles(P, ga, mk(6, bp + nn, nj, 0, 2, x, nn, 1, 2, t, na,
                                     na, nn, nn, 1.0, 0.0, 0, 0, 2), false);
            // M -= B21^T T
            add_gemm_tiles(P, gb, mk(5, bp + nn, nj, 1, 2, t, na, 0, 2, m, nn,
                                     nn, nn, na, -1.0, 1.0, 0), false);
            // T -= 2 Ub L21 (Ub symmetric, both triangles stored)
            add_gemm_tiles(P, gc, mk(1, f.updptr, na, 0, 2, l21, na, 0, 2, t, na,
                                     na, nn, na, -2.0, 1.0, 0), false);
            // Fbar21 = T X11 / 2 (X11 lower)
            add_gemm_tiles(P, gd, mk(2, t, na, 0, 2, x, nn, 0, 0, bp + nn, nj,
                                     na, nn, nn, 0.5, 0.0, 0, 0, 1), false);
            // M -= 2 Fbar21^T L21
            add_gemm_tiles(P, ge, mk(0, bp + nn, nj, 1, 2, l21, na, 0, 2, m, nn,
                                     nn, nn, na, -2.0, 1.0, 0), false);
        }
        // P = L11^T tril(M) (lower part), then Phi
        add_gemm_tiles(P, gf, mk(5, bp, nj, 1, 2, m, nn, 0, 2, p, nn,
                                 nn, nn, nn, 1.0, 0.0, 1, 2, 1), true);
        add_matop(phi, MO_PHI, 2, p, nn, 2, p, nn, nn, nn);
        // W2 = P X11 (both lower), S = X11^T W2
        add_gemm_tiles(P, gg, mk(2, p, nn, 0, 2, x, nn, 0, 2, w2, nn,
                                 nn, nn, nn, 1.0, 0.0, 0, 1, 1), false);
        add_gemm_tiles(P, gh, mk(2, x, nn, 1, 2, w2, nn, 0, 2, s, nn,
                                 nn, nn, nn, 1.0, 0.0, 0, 2, 0), false);
        add_matop(sym, MO_SYMAVG, 2, s, nn, 0, bp, nj, nn, nn, 0.5);
    }
    emit_trtri(P, seq, Ls, Xs, Ws, keys);
    push_tasks(P, seq, K_MATOP, m0);
    for (auto* g : {&ga, &gb, &gc, &gd, &ge, &gf}) push_gemm(P, seq, *g);
    push_tasks(P, seq, K_MATOP, phi);
    push_gemm(P, seq, gg);
    push_gemm(P, seq, gh);
    push_tasks(P, seq, K_MATOP, sym);
    // the children's Ub from this level's finished Fbar
    push_tasks(P, seq, K_EXTRACT, extract_tiles(P, children_of(P, fr)), 0);
}

int ensure_dense(chp_plan* P, hipStream_t s) {
    if (sync_program(P->dense)) return -2;
    struct { double** d; i64* have; i64 need; } bufs[] = {
        {&P->d_dscratch, &P->dscratch_cap, P->dscratch_len},
        {&P->d_ddinv, &P->ddinv_cap, P->ddinv_len},
        {&P->d_dsplitp, &P->dsplitp_cap, P->dense.split_doubles()},
    };
    for (auto& b : bufs) {
        if (*b.d && b.need <= *b.have) continue;
        HIP_CHECK(hipStreamSynchronize(s));
        if (*b.d) HIP_CHECK(hipFree(*b.d));
        *b.d = nullptr;
        if (alloc_doubles(b.d, b.need)) return -2;
        *b.have = std::max<i64>(b.need, 1);
        if (b.d == &P->d_dsplitp)  // arrival counters start at zero
            HIP_CHECK(hipMemset(P->d_dsplitp, 0, std::max<i64>(b.need, 1) * sizeof(double)));
    }
    return 0;
}

// Builds the requested sweeps on first use; the needs (scratch, inverses)
// grow the plan's dense buffers.
struct DenseBuild {
    chp_plan* P;
    i64 sc_need = 0, dv_need = 0;
    explicit DenseBuild(chp_plan* p) : P(p) { P->pg = &P->dense; }
    ~DenseBuild() { P->pg = &P->main; }
    void level_done(const Bump& sc, const Bump& dv) {
        sc_need = std::max(sc_need, sc.cur);
        dv_need = std::max(dv_need, dv.cur);
    }
};

void note_needs(chp_plan* P, const DenseBuild& b, i64& sc_req, i64& dv_req) {
    sc_req = std::max(sc_req, b.sc_need);
    dv_req = std::max(dv_req, b.dv_need);
}

}  // namespace

int build_completion(chp_plan* P, int fu, i64& sc_req, i64& dv_req) {
    DenseBuild b(P);
    auto& seq = P->comp[fu].seq;
    // With factored updates, runs of at least CHAIN_MIN levels that each hold
    // one front with nn + na <= 64 go to one k_completion_chain launch.
    constexpr int CHAIN_MIN = 3;
    auto chain_front = [&](int L) {
        const std::vector<int>& fr = P->bylevel[L];
        return fr.size() == 1 && P->fi[fr[0]].nn + P->fi[fr[0]].na <= 64;
    };
    for (int L = P->nlevels - 1; L >= 0;) {
        int L2 = L;
        while (fu && L2 >= 0 && chain_front(L2)) --L2;
        if (L - L2 >= CHAIN_MIN) {
            std::vector<ChainItem> items;
            for (int M = L; M > L2; --M) {
                const int k = P->bylevel[M][0];
                const int next = M - 1 > L2 ? P->bylevel[M - 1][0] : -1;
                ChainItem ci;
                std::memset(&ci, 0, sizeof(ci));
                ci.front = k;
                ci.kind = 1;
                ci.nch = P->fi[k].chend - P->fi[k].chbeg;
                items.push_back(ci);
                for (int pass = 0; pass < 2; ++pass)  // the chain's next front last
                    for (int l = P->fi[k].chbeg; l < P->fi[k].chend; ++l) {
                        const int c = P->h_chidx[l];
                        if ((c == next) != (pass == 1)) continue;
                        std::memset(&ci, 0, sizeof(ci));
                        ci.uf = uf_task(P, c, 0);
                        ci.kind = c == next ? 2 : 0;
                        items.push_back(ci);
                    }
            }
            push_tasks(P, seq, K_COMP_CHAIN, items);
            seq.back().grid = 1;
            seq.back().block = 64;
            seq.back().iarg = (int)items.size();
            L = L2;
            continue;
        }
        Bump sc, dv;
        build_completion_level(P, seq, P->bylevel[L], fu != 0, sc, dv);
        b.level_done(sc, dv);
        --L;
    }
    P->comp[fu].built = true;
    note_needs(P, b, sc_req, dv_req);
    return 0;
}

// Distances of consecutive directions of a chunk: blkval (buffer 0) and
// update arena (buffer 1), rounded to 256-byte multiples; 0 for the other
// buffers.
void dir_strides(const chp_plan* P, i64* str) {
    for (int i = 0; i < 8; ++i) str[i] = 0;
    str[0] = (P->blk_len + 31) / 32 * 32;
    str[1] = (P->upd_len + 31) / 32 * 32;
}

int build_scale(chp_plan* P, int idx, int nb, i64& sc_req, i64& dv_req) {
    DenseBuild b(P);
    const bool fu = idx & 4, inv = idx & 2, adj = idx & 1;
    auto& S = P->scale[idx][nb];
    S.shared.assign(P->nlevels, {});
    S.shared_t.assign(P->nlevels, {});
    S.dir.assign(P->nlevels, {});
    i64 str[8];
    dir_strides(P, str);
    // the scaling sweeps keep every update slot (d_supd): built on the
    // one-slot-per-front table, and their front kernels read that table
    std::swap(P->fi, P->fi_full);
    for (int L = P->nlevels - 1; L >= 0; --L) {
        Bump sc, dv;
        build_scale_level(P, S.shared[L], S.shared_t[L], S.dir[L], P->bylevel[L], fu, inv, adj,
                          sc, dv, nb, str[0]);
        b.level_done(sc, dv);
        for (auto* v : {&S.shared[L], &S.shared_t[L], &S.dir[L]})
            for (Launch& l : *v) l.fisel = 1;
    }
    std::swap(P->fi, P->fi_full);
    S.built = true;
    note_needs(P, b, sc_req, dv_req);
    return 0;
}

int build_chol_adjoint(chp_plan* P, i64& sc_req) {
    DenseBuild b(P);
    for (int L = P->nlevels - 1; L >= 0; --L) {
        Bump sc, dv;
        build_chol_adjoint_level(P, P->chol_adj.seq, P->bylevel[L], sc);
        b.level_done(sc, dv);
    }
    P->chol_adj.built = true;
    sc_req = std::max(sc_req, b.sc_need);
    return 0;
}

int build_y2k(chp_plan* P, int inv, int nb) {
    DenseBuild b(P);
    i64 str[8];
    dir_strides(P, str);
    auto& S = P->y2k[inv][nb];
    for (int L = 0; L < P->nlevels; ++L)
        build_y2k_level(P, S.seq, P->bylevel[L], inv != 0, nb, str);
    S.built = true;
    return 0;
}

int build_m2t(chp_plan* P, int inv, int nb, i64& sc_req) {
    DenseBuild b(P);
    i64 str[8];
    dir_strides(P, str);
    auto& S = P->m2t[inv][nb];
    for (int L = P->nlevels - 1; L >= 0; --L) {
        Bump sc, dv;
        build_m2t_level(P, S.seq, P->bylevel[L], inv != 0, sc, nb, str);
        b.level_done(sc, dv);
    }
    S.built = true;
    sc_req = std::max(sc_req, b.sc_need);
    return 0;
}

BufTab dense_tab(chp_plan* P, double* x, const double* L, const double* Y) {
    BufTab t;
    t.p[0] = x; t.p[1] = P->d_upd; t.p[2] = P->d_dscratch; t.p[3] = P->d_ddinv;
    t.p[4] = P->d_dsplitp; t.p[5] = const_cast<double*>(L); t.p[6] = const_cast<double*>(Y);
    t.p[7] = nullptr;
    t.gate = nullptr;
    return t;
}

}  // namespace chpi

using namespace chpi;

namespace {
// Requested dense buffer sizes (scratch, inverses) before allocation.
int prepare(chp_plan* P, i64 sc_req, i64 dv_req, hipStream_t s) {
    P->dscratch_len = std::max(P->dscratch_len, sc_req);
    P->ddinv_len = std::max(P->ddinv_len, dv_req);
    return ensure_dense(P, s);
}
}  // namespace

extern "C" {

int chp_completion(chp_plan* P, double* blkval, int factored, void* stream, i64* info) {
    if (!P || !blkval) { chp_set_error("null argument"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    const int fu = factored ? 1 : 0;
    i64 sc = 0, dv = 0;
    if (!P->comp[fu].built) build_completion(P, fu, sc, dv);
    if (prepare(P, sc, dv, s)) return -2;
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, s, P->d_status);
    int rc = run_seq(P, P->dense, P->comp[fu].seq, dense_tab(P, blkval, nullptr, nullptr), s);
    if (rc) return rc;
    return read_status(P, s, info);
}

int chp_cholesky_adjoint(chp_plan* P, const double* lblkval, const double* gblkval, double* out,
                         void* stream, i64* info) {
    if (!P || !lblkval || !gblkval || !out) { chp_set_error("null argument"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    i64 sc = 0, dv = 0;
    if (!P->chol_adj.built) build_chol_adjoint(P, sc);
    if (prepare(P, sc, dv, s)) return -2;
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, s, P->d_status);
    BufTab bt = dense_tab(P, out, lblkval, gblkval);
    int rc = run_seq(P, P->dense, P->chol_adj.seq, bt, s);
    if (rc) return rc;
    return read_status(P, s, info);
}

int chp_hessian(chp_plan* P, const double* lblkval, const double* yblkval, double* const* ublkval,
                int nu, int adj, int inv, int factored, void* stream, i64* info) {
    if (!P || !lblkval || !yblkval || (nu > 0 && !ublkval)) {
        chp_set_error("null argument");
        return -1;
    }
    if (nu <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int fu = factored ? 1 : 0, iv = inv ? 1 : 0;
    const int a0 = adj < 0 ? iv : (adj ? 1 : 0);
    const int npass = adj < 0 ? 2 : 1;
    i64 str[8];
    dir_strides(P, str);
    // Directions go in chunks of up to HESS_CHUNK, stacked in one device
    // array (blkval bstride apart, each with its own update arena): every
    // launch of a level covers the whole chunk.  The chunk's device memory is
    // capped at HESS_BYTES.
    constexpr int HESS_CHUNK = 32;
    constexpr double HESS_BYTES = 16e9;
    int nbmax = std::min(nu, HESS_CHUNK);
    while (nbmax > 1 && 8.0 * nbmax * (double)(str[0] + str[1]) > HESS_BYTES) --nbmax;
    i64 sc = 0, dv = 0;
    std::vector<int> sizes = {nbmax};
    if (nu % nbmax) sizes.push_back(nu % nbmax);
    for (int nb : sizes)
        for (int pass = 0; pass < npass; ++pass) {
            const int a = pass == 0 ? a0 : 1 - a0;
            const int idx = 4 * fu + 2 * iv + a;
            if (!P->scale[idx][nb].built) build_scale(P, idx, nb, sc, dv);
            if (a) { if (!P->m2t[iv][nb].built) build_m2t(P, iv, nb, sc); }
            else if (!P->y2k[iv][nb].built) build_y2k(P, iv, nb);
        }
    if (prepare(P, sc, dv, s)) return -2;
    if (!P->d_supd && alloc_doubles(&P->d_supd, P->upd_full_len)) return -2;
    const i64 need_u = (i64)nbmax * str[0], need_a = (i64)nbmax * str[1];
    if (P->ustack_cap < need_u || P->uupd_cap < need_a) {
        HIP_CHECK(hipStreamSynchronize(s));
        if (P->ustack_cap < need_u) {
            if (P->d_ustack) HIP_CHECK(hipFree(P->d_ustack));
            P->d_ustack = nullptr;
            if (alloc_doubles(&P->d_ustack, need_u)) return -2;
            P->ustack_cap = need_u;
        }
        if (P->uupd_cap < need_a) {
            if (P->d_uupd) HIP_CHECK(hipFree(P->d_uupd));
            P->d_uupd = nullptr;
            if (alloc_doubles(&P->d_uupd, need_a)) return -2;
            P->uupd_cap = need_a;
        }
    }
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, s, P->d_status);
    // The scaling sweeps keep their update matrices (factored or not; the
    // same for adj and its adjoint) in d_supd, which the M2 sweeps between
    // them leave alone: after the first scaling sweep of the call the others
    // (later chunks, the second pass of adj=None) reuse them.
    bool have_upd = false;
    for (int u0 = 0; u0 < nu; u0 += nbmax) {
        const int nb = std::min(nbmax, nu - u0);
        for (int u = 0; u < nb; ++u)
            HIP_CHECK(hipMemcpyAsync(P->d_ustack + u * str[0], ublkval[u0 + u],
                                     P->blk_len * sizeof(double), hipMemcpyDeviceToDevice, s));
        auto run_scale = [&](int idx) -> int {
            const auto& S = P->scale[idx][nb];
            for (int L = P->nlevels - 1; L >= 0; --L) {
                BufTab bt = dense_tab(P, P->d_ustack, lblkval, yblkval);
                bt.p[1] = P->d_supd;
                int rc = have_upd ? 0 : run_seq(P, P->dense, S.shared[L], bt, s);
                if (!rc) rc = run_seq(P, P->dense, S.shared_t[L], bt, s);
                if (!rc) rc = run_seq(P, P->dense, S.dir[L], bt, s);
                if (rc) return rc;
            }
            have_upd = true;
            return 0;
        };
        for (int pass = 0; pass < npass; ++pass) {
            const int a = pass == 0 ? a0 : 1 - a0;
            const int idx = 4 * fu + 2 * iv + a;
            int rc = 0;
            if (a != iv && (rc = run_scale(idx))) return rc;
            const auto& seq = a ? P->m2t[iv][nb].seq : P->y2k[iv][nb].seq;
            BufTab bt = dense_tab(P, P->d_ustack, lblkval, yblkval);
            bt.p[1] = P->d_uupd;
            if ((rc = run_seq(P, P->dense, seq, bt, s))) return rc;
            if (a == iv && (rc = run_scale(idx))) return rc;
        }
        for (int u = 0; u < nb; ++u)
            HIP_CHECK(hipMemcpyAsync(ublkval[u0 + u], P->d_ustack + u * str[0],
                                     P->blk_len * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    // a failed update factorization is reported in info but, as in the
    // reference (hessian.c:354-374 returns 0), is not an error
    read_status(P, s, info);
    return 0;
}

}  // extern "C"
