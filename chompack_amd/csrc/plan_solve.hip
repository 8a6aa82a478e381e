// Triangular solve with the chordal factor, B := alpha L^{-1} B or
// alpha L^{-T} B (trsm.c:3-138 as called by cbase.c:965-1077), on the level
// schedule of a plan.
//
// Fronts with a small diagonal block and little separator work are solved by
// one workgroup each (k_solve_fwd / k_solve_bwd).  For the larger fronts the
// inverses X_k = L11^{-1} are formed first, all at once (diagonal-block
// inverses and recursive halving on MFMA, from the factor passed in), so that
// every level is a few batched launches:
//   forward:  gather (alpha B rows + children's update rows); one GEMM launch
//             with x_A -= B21 x_N (B21 = L21 L11^{-1}, so x_N is used before
//             its solve) and T = X_k x_N; scatter T to B;
//   backward: gather (alpha B rows + the parent's separator values);
//             T = X_k^T x_N; T -= B21^T x_A; scatter T to B and back into the
//             slab, where the children read their separator values.
// Solving with the explicit inverse instead of substitution changes the last
// bits only (forward error of the order of cond(L11) eps either way).
// The schedule depends on the number of right-hand sides and is cached per
// count.
#include "plan.h"
#include <hip/hip_ext.h>
#include "../../include/chompack_amd.h"

#include <algorithm>

namespace chpi {
namespace {

// Fronts that one wave solves (k_solve_fwd / k_solve_bwd): at most 16
// columns and 64 rows.  Everything else takes the batched gather / GEMM /
// scatter path, whose launches spread a front's rows over many workgroups.
bool solve_large(const FrontInfo& f, int nrhs) {
    (void)nrhs;
    return f.nn > 16 || f.nn + f.na > 64;
}

// (front, first row) tiles of `step` rows over rows [0, nrow(front)).
template <class F>
std::vector<int2> row_tiles(const std::vector<int>& fr, F nrow, int step = 256) {
    std::vector<int2> t;
    for (int k : fr)
        for (int r0 = 0; r0 < nrow(k); r0 += step) t.push_back(make_int2(k, r0));
    return t;
}

// (front, first right-hand side) items of the small-front solve kernels,
// 16 right-hand sides per wave.
std::vector<int2> rhs_items(const std::vector<int>& fr, int nrhs) {
    std::vector<int2> t;
    for (int k : fr)
        for (int j0 = 0; j0 < nrhs; j0 += 16) t.push_back(make_int2(k, j0));
    return t;
}

template <class T>
void push_list(chp_plan* P, std::vector<Launch>& seq, Kid kid, const std::vector<T>& v,
               int block = 256) {
    if (v.empty()) return;
    seq.push_back({kid, (int)v.size(), block, 0, P->pg->lists.add(v), 0});
}

SolvePlan* build_solve(chp_plan* P, int nrhs) {
    SolvePlan* S = new SolvePlan();
    S->nrhs = nrhs;
    P->pg = &S->pg;
    const int H = P->nlevels;
    // X_k = L11^{-1} of every large front, one batched recursive-halving
    // inverse (buffer 6, ld nn; workspace buffer 2), formed again only when
    // L11 differs from the copy it was formed from (k_l11_check)
    std::vector<i64> xo(P->fi.size(), -1), to(P->fi.size(), -1);
    std::vector<Mat> Ls, Xs;
    std::vector<i64> Ws;
    std::vector<int> keys;
    Bump bx, bw, bt;
    for (size_t k = 0; k < P->fi.size(); ++k) {
        const FrontInfo& f = P->fi[k];
        if (!solve_large(f, nrhs)) continue;
        xo[k] = bx.take((i64)f.nn * f.nn);
        to[k] = bt.take((i64)f.nn * nrhs);
        Ls.push_back({0, f.blkptr, f.nn + f.na, f.nn});
        Xs.push_back({6, xo[k], f.nn, f.nn});
        Ws.push_back(bw.take((i64)f.nn * f.nn));
        keys.push_back((int)k);
        // 16 columns x at most 1024 rows per workgroup
        for (int c0 = 0; c0 < f.nn; c0 += 16)
            for (int r0 = c0; r0 < f.nn; r0 += 1024)
                S->chk.push_back({xo[k], (int)k, c0, r0, std::min(r0 + 1024, f.nn)});
    }
    S->inv_len = bx.cur;
    S->w_len = bw.cur;
    // T = X_k x_N (or X_k^T x_N) at to[k]; the backward sweep's
    // T2 = B21^T x_A at to[k] + t2base, subtracted by the scatter
    S->t2base = bt.cur;
    S->tmp_len = 2 * bt.cur;
    if (!Ls.empty()) emit_trtri(P, S->inv, Ls, Xs, Ws, keys);
    auto wofs = [&](int k) { return P->h_wunit[k] * nrhs; };
    auto stiles = [&](const std::vector<int>& large) {
        std::vector<STile> t;
        for (int k : large)
            for (int r0 = 0; r0 < P->fi[k].nn; r0 += 64) t.push_back({k, r0, to[k]});
        return t;
    };
    for (int L = 0; L < H; ++L) {
        std::vector<int> small, large;
        for (int k : P->bylevel[L]) (solve_large(P->fi[k], nrhs) ? large : small).push_back(k);
        // small and large fronts of a level touch disjoint slabs and rows of
        // B: the small fronts run on side stream 1
        const bool fork = !small.empty() && !large.empty();
        if (fork) S->fwd.push_back({K_FORK, 0, 0, 0, 0, 1});
        push_list(P, S->fwd, K_SOLVE_FWD, rhs_items(small, nrhs), 64);
        if (fork) S->fwd.back().side = 1;
        if (large.empty()) continue;
        auto nj_of = [&](int k) { return P->fi[k].nn + P->fi[k].na; };
        push_list(P, S->fwd, K_SGATHER_F, row_tiles(large, nj_of, 64));
        std::vector<GemmTile> g;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            const int nn = f.nn, na = f.na, nj = nn + na;
            if (na > 0)  // x_A -= B21 x_N  (B21 = L21 L11^{-1}: x_N before its solve)
                add_gemm_tiles(P, g, mk(0, f.blkptr + nn, nj, 0, 7, wofs(k), nj, 0, 7, wofs(k) + nn,
                                        nj, na, nrhs, nn, -1.0, 1.0, 0), false);
            // T = L11^{-1} x_N
            add_gemm_tiles(P, g, mk(6, xo[k], nn, 0, 7, wofs(k), nj, 0, 5, to[k], nn,
                                    nn, nrhs, nn, 1.0, 0.0, 0, 1), false);
        }
        push_gemm(P, S->fwd, g);
        std::vector<STile> st = stiles(large);
        push_list(P, S->fwd, K_SSCATTER, st);
        if (fork) S->fwd.push_back({K_JOIN, 0, 0, 0, 0, 1});
    }
    for (int L = H - 1; L >= 0; --L) {
        std::vector<int> small, large;
        for (int k : P->bylevel[L]) (solve_large(P->fi[k], nrhs) ? large : small).push_back(k);
        const bool fork = !small.empty() && !large.empty();
        if (fork) S->bwd.push_back({K_FORK, 0, 0, 0, 0, 1});
        push_list(P, S->bwd, K_SOLVE_BWD, rhs_items(small, nrhs), 64);
        if (fork) S->bwd.back().side = 1;
        if (large.empty()) continue;
        auto nj_of = [&](int k) { return P->fi[k].nn + P->fi[k].na; };
        push_list(P, S->bwd, K_SGATHER_B, row_tiles(large, nj_of, 64));
        std::vector<GemmTile> g1;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            const int nn = f.nn, na = f.na, nj = nn + na;
            // T = L11^{-T} x_N
            add_gemm_tiles(P, g1, mk(6, xo[k], nn, 1, 7, wofs(k), nj, 0, 5, to[k], nn,
                                     nn, nrhs, nn, 1.0, 0.0, 0, 2), false);
            if (na > 0)  // T2 = B21^T x_A (the scatter writes T - T2)
                add_gemm_tiles(P, g1, mk(0, f.blkptr + nn, nj, 1, 7, wofs(k) + nn, nj, 0, 5,
                                         to[k] + S->t2base, nn, nn, nrhs, na, 1.0, 0.0, 0), false);
        }
        push_gemm(P, S->bwd, g1);
        std::vector<STile> st = stiles(large);
        Launch Lc{K_SSCATTER, (int)st.size(), 256, 0, P->pg->lists.add(st), 1};
        S->bwd.push_back(Lc);
        if (fork) S->bwd.push_back({K_JOIN, 0, 0, 0, 0, 1});
    }
    P->pg = &P->main;
    return S;
}

// Schedules of B := alpha L B ('N', leaves to root) and alpha L^T B ('T',
// root to leaves) on the solve plan's level schedule, slabs and per-front
// blocks T (to[k]) and T2 (to[k] + t2base), pybase/trmm.py:44-114.
void build_trmm(chp_plan* P, SolvePlan* S) {
    const int nrhs = S->nrhs;
    P->pg = &S->pg;
    const int H = P->nlevels;
    std::vector<i64> to(P->fi.size(), -1);
    {
        Bump bt;  // the same offsets build_solve gave the large fronts
        for (size_t k = 0; k < P->fi.size(); ++k)
            if (solve_large(P->fi[k], nrhs)) to[k] = bt.take((i64)P->fi[k].nn * nrhs);
    }
    auto wofs = [&](int k) { return P->h_wunit[k] * nrhs; };
    auto nj_of = [&](int k) { return P->fi[k].nn + P->fi[k].na; };
    for (int L = 0; L < H; ++L) {
        std::vector<int> small, large;
        for (int k : P->bylevel[L]) (solve_large(P->fi[k], nrhs) ? large : small).push_back(k);
        const bool fork = !small.empty() && !large.empty();
        if (fork) S->mfwd.push_back({K_FORK, 0, 0, 0, 0, 1});
        push_list(P, S->mfwd, K_TRMM_FWD, rhs_items(small, nrhs), 64);
        if (fork) S->mfwd.back().side = 1;
        if (large.empty()) continue;
        Launch Lg{K_SGATHER_F, 0, 256, 0, 0, 1};  // iarg 1: alpha B rows only
        {
            std::vector<int2> t = row_tiles(large, [&](int k) { return P->fi[k].nn; }, 64);
            Lg.grid = (int)t.size();
            Lg.list = P->pg->lists.add(t);
        }
        S->mfwd.push_back(Lg);
        std::vector<GemmTile> g1, g2;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            const int nn = f.nn, na = f.na, nj = nn + na;
            // T = tril(L11) (alpha b_N)
            add_gemm_tiles(P, g1, mk(0, f.blkptr, nj, 0, 7, wofs(k), nj, 0, 5, to[k], nn,
                                     nn, nrhs, nn, 1.0, 0.0, 0, 1, 0), false);
            if (na > 0)  // slab rows >= nn = B21 T
                add_gemm_tiles(P, g2, mk(0, f.blkptr + nn, nj, 0, 5, to[k], nn, 0, 7,
                                         wofs(k) + nn, nj, na, nrhs, nn, 1.0, 0.0, 0, 0, 0), false);
        }
        push_gemm(P, S->mfwd, g1);
        push_gemm(P, S->mfwd, g2);
        std::vector<STile> ft;
        for (int k : large)
            for (int r0 = 0; r0 < nj_of(k); r0 += 64) ft.push_back({k, r0, to[k]});
        push_list(P, S->mfwd, K_TRMM_FIN, ft);
        if (fork) S->mfwd.push_back({K_JOIN, 0, 0, 0, 0, 1});
    }
    for (int L = H - 1; L >= 0; --L) {
        std::vector<int> small, large;
        for (int k : P->bylevel[L]) (solve_large(P->fi[k], nrhs) ? large : small).push_back(k);
        const bool fork = !small.empty() && !large.empty();
        if (fork) S->mbwd.push_back({K_FORK, 0, 0, 0, 0, 1});
        push_list(P, S->mbwd, K_TRMM_BWD, rhs_items(small, nrhs), 64);
        if (fork) S->mbwd.back().side = 1;
        if (large.empty()) continue;
        push_list(P, S->mbwd, K_SGATHER_B, row_tiles(large, nj_of, 64));
        std::vector<GemmTile> g1, g2;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            const int nn = f.nn, na = f.na, nj = nn + na;
            // T = tril(L11)^T u_N; T2 = B21^T u_A; then T += tril(L11)^T T2
            add_gemm_tiles(P, g1, mk(0, f.blkptr, nj, 1, 7, wofs(k), nj, 0, 5, to[k], nn,
                                     nn, nrhs, nn, 1.0, 0.0, 0, 2, 0), false);
            if (na > 0) {
                add_gemm_tiles(P, g1, mk(0, f.blkptr + nn, nj, 1, 7, wofs(k) + nn, nj, 0, 5,
                                         to[k] + S->t2base, nn, nn, nrhs, na, 1.0, 0.0, 0, 0, 0),
                               false);
                add_gemm_tiles(P, g2, mk(0, f.blkptr, nj, 1, 5, to[k] + S->t2base, nn, 0, 5, to[k],
                                         nn, nn, nrhs, nn, 1.0, 1.0, 0, 2, 0), false);
            }
        }
        push_gemm(P, S->mbwd, g1);
        push_gemm(P, S->mbwd, g2);
        std::vector<STile> st;
        for (int k : large)
            for (int r0 = 0; r0 < P->fi[k].nn; r0 += 64) st.push_back({k, r0, to[k]});
        push_list(P, S->mbwd, K_SSCATTER, st);  // iarg 0: B only, no T2
        if (fork) S->mbwd.push_back({K_JOIN, 0, 0, 0, 0, 1});
    }
    S->mbuilt = true;
    P->pg = &P->main;
}

// gate: when not null, the dense launches of seq do nothing unless *gate != 0.
int run_solve(chp_plan* P, const SolvePlan* S, const std::vector<Launch>& seq,
              const double* blkval, double* B, int ldb, int nrhs, double alpha, hipStream_t s,
              const int* gate = nullptr) {
    BufTab bt;
    for (int i = 0; i < 8; ++i) bt.p[i] = nullptr;
    bt.gate = gate;
    bt.p[0] = const_cast<double*>(blkval);
    bt.p[2] = S->d_w;
    bt.p[4] = S->d_splitp;
    bt.p[5] = S->d_tmp;
    bt.p[6] = S->d_inv;
    bt.p[7] = P->d_work;
#define CHP_SLAUNCH(K, ...)                                                        \
    do {                                                                       \
        if (P->prof)                                                           \
            hipExtLaunchKernelGGL(K, g, b, 0, s, P->ev0, P->ev1, 0, __VA_ARGS__); \
        else                                                                   \
            hipLaunchKernelGGL(K, g, b, 0, s, __VA_ARGS__);                    \
    } while (0)
    const bool two = !P->prof;  // profiling serializes everything on one stream
    if (two && ensure_side_stream(P)) return -2;
    const hipStream_t s0 = s;
    for (const Launch& L : seq) {
        if (L.kid == K_FORK || L.kid == K_JOIN) {
            if (!two) continue;
            const int i = L.iarg;
            if (L.kid == K_FORK) {
                HIP_CHECK(hipEventRecord(P->sev_fork[i], s0));
                HIP_CHECK(hipStreamWaitEvent(P->sstream[i], P->sev_fork[i], 0));
            } else {
                HIP_CHECK(hipEventRecord(P->sev_join[i], P->sstream[i]));
                HIP_CHECK(hipStreamWaitEvent(s0, P->sev_join[i], 0));
            }
            continue;
        }
        s = (two && L.side) ? P->sstream[L.side] : s0;
        const void* list = S->pg.d_lists + L.list;
        // the slab gather/scatter kernels take the right-hand sides in chunks
        // along the grid's y dimension
        const bool rhs_y = L.kid == K_SGATHER_F || L.kid == K_SGATHER_B || L.kid == K_SSCATTER;
        const dim3 g(L.grid, rhs_y ? (nrhs + SOLVE_RHS_CHUNK - 1) / SOLVE_RHS_CHUNK : 1), b(L.block);
        switch (L.kid) {
        case K_SOLVE_FWD:
            CHP_SLAUNCH(k_solve_fwd, (const int2*)list, P->d_fi, blkval, P->d_snptr,
                        P->d_perm, B, ldb, nrhs, alpha, P->d_work, P->d_wunit, P->d_gptr,
                        P->d_gent);
            break;
        case K_SOLVE_BWD:
            CHP_SLAUNCH(k_solve_bwd, (const int2*)list, P->d_fi, blkval, P->d_snptr,
                               P->d_relidx, P->d_perm, B, ldb, nrhs, alpha, P->d_work, P->d_wunit);
            break;
        case K_SGATHER_F:
            CHP_SLAUNCH(k_sgather_fwd, (const int2*)list, P->d_fi, P->d_snptr, P->d_perm, B,
                        ldb, nrhs, alpha, P->d_work, P->d_wunit, P->d_gptr, P->d_gent, L.iarg);
            break;
        case K_TRMM_FWD:
            CHP_SLAUNCH(k_trmm_fwd, (const int2*)list, P->d_fi, blkval, P->d_snptr, P->d_perm, B,
                        ldb, nrhs, alpha, P->d_work, P->d_wunit, P->d_gptr, P->d_gent);
            break;
        case K_TRMM_BWD:
            CHP_SLAUNCH(k_trmm_bwd, (const int2*)list, P->d_fi, blkval, P->d_snptr, P->d_relidx,
                        P->d_perm, B, ldb, nrhs, alpha, P->d_work, P->d_wunit);
            break;
        case K_TRMM_FIN:
            CHP_SLAUNCH(k_trmm_finish, (const STile*)list, P->d_fi, P->d_snptr, P->d_perm, B, ldb,
                        nrhs, P->d_work, P->d_wunit, P->d_gptr, P->d_gent, S->d_tmp);
            break;
        case K_SGATHER_B:
            CHP_SLAUNCH(k_sgather_bwd, (const int2*)list, P->d_fi, P->d_snptr,
                               P->d_relidx, P->d_perm, B, ldb, nrhs, alpha, P->d_work, P->d_wunit);
            break;
        case K_SSCATTER:
            CHP_SLAUNCH(k_sscatter, (const STile*)list, P->d_fi, P->d_snptr,
                        P->d_perm, B, ldb, nrhs, P->d_work, P->d_wunit, S->d_tmp,
                        L.iarg ? S->d_tmp + S->t2base : nullptr, L.iarg);
            break;
        case K_DIAG:
            CHP_SLAUNCH(k_diag_tasks, (const DiagTask*)list, bt, P->d_status);
            break;
        case K_MATOP:
            CHP_SLAUNCH(k_matop, (const MatOp*)list, bt);
            break;
        case K_GEMM:
            if (L.block == 512)
                CHP_SLAUNCH(k_gemm_tiles_lat, (const GemmTile*)list, S->pg.d_tasks, bt);
            else
                CHP_SLAUNCH(k_gemm_tiles, (const GemmTile*)list, S->pg.d_tasks, bt);
            break;
        case K_SPLITK_REDUCE:
            CHP_SLAUNCH(k_splitk_reduce, (const GemmTile*)list, S->pg.d_tasks, bt);
            break;
        case K_GEMM_NARROW:
            CHP_SLAUNCH(k_gemm_narrow, (const GemmTile*)list, S->pg.d_tasks, bt);
            break;
        case K_GEMM_SMALL:
            CHP_SLAUNCH(k_gemm_small, (const int*)list, L.iarg, S->pg.d_tasks, bt);
            break;
        default:
            chp_set_error("internal: unexpected launch kind in solve");
            return -3;
        }
        if (P->prof) {  // per-kind times, as for the factorization sweeps
            HIP_CHECK(hipEventSynchronize(P->ev1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, P->ev0, P->ev1));
            P->prof_ms[L.kid] += ms;
            P->prof_cnt[L.kid] += 1;
            P->prof_flops[L.kid] += L.flops;
            P->prof_rows.insert(P->prof_rows.end(), {(double)L.kid, (double)L.grid, L.flops, (double)ms});
        }
    }
#undef CHP_SLAUNCH
    HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace
}  // namespace chpi

using namespace chpi;

namespace {

// B := alpha L^{-1} B / alpha L^{-T} B (product = false) or alpha L B /
// alpha L^T B (product = true) on the cached schedule for nrhs.
int triangular_op(chp_plan* P, const double* blkval, double* B, int ldb, int nrhs, double alpha,
                  int trans, bool product, void* stream) {
    if (!P || !blkval || !B) { chp_set_error("null argument"); return -1; }
    if (trans != 'N' && trans != 'T') { chp_set_error("trans must be 'N' or 'T'"); return -1; }
    if (nrhs <= 0) return 0;
    if (ldb < P->n) { chp_set_error("ldB is smaller than n"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    SolvePlan* S = nullptr;
    for (SolvePlan* q : P->solves) if (q->nrhs == nrhs) S = q;
    if (!S) {
        S = build_solve(P, nrhs);
        P->solves.push_back(S);
        if (sync_program(S->pg)) return -2;
        if (alloc_doubles(&S->d_splitp, S->pg.split_doubles())) return -2;
        HIP_CHECK(hipMemset(S->d_splitp, 0, S->pg.split_doubles() * sizeof(double)));
        if (alloc_doubles(&S->d_inv, S->inv_len)) return -2;
        if (alloc_doubles(&S->d_w, S->w_len)) return -2;
        if (alloc_doubles(&S->d_tmp, S->tmp_len)) return -2;
        if (alloc_doubles(&S->d_l11, S->inv_len)) return -2;
        if (upload(&S->d_chk, S->chk)) return -2;
        HIP_CHECK(hipMalloc((void**)&S->d_changed, sizeof(int)));
    }
    if (product && !S->mbuilt) {
        // the product schedules extend the plan's lists and tasks, which may
        // move them and the split-K partials: the plan's graphs go first
        HIP_CHECK(hipDeviceSynchronize());
        for (auto& g : S->graphs) (void)hipGraphExecDestroy(g.exec);
        S->graphs.clear();
        S->last = SolvePlan::Key();
        const i64 before = S->pg.split_doubles();
        build_trmm(P, S);
        if (sync_program(S->pg)) return -2;
        if (S->pg.split_doubles() > before) {
            HIP_CHECK(hipFree(S->d_splitp));
            S->d_splitp = nullptr;
            if (alloc_doubles(&S->d_splitp, S->pg.split_doubles())) return -2;
            HIP_CHECK(hipMemset(S->d_splitp, 0, S->pg.split_doubles() * sizeof(double)));
        }
    }
    const i64 need = P->solve_len * nrhs;
    // captured solves hold the addresses of the work buffers: drop them all
    // before either buffer moves
    auto drop_graphs = [&]() {
        for (SolvePlan* q : P->solves) {
            for (auto& g : q->graphs) (void)hipGraphExecDestroy(g.exec);
            q->graphs.clear();
            q->last = SolvePlan::Key();
        }
    };
    if (need > P->work_len) {
        HIP_CHECK(hipDeviceSynchronize());
        drop_graphs();
    }
    if (need > P->work_len) {
        if (P->d_work) HIP_CHECK(hipFree(P->d_work));
        P->d_work = nullptr;
        HIP_CHECK(hipMalloc((void**)&P->d_work, need * sizeof(double)));
        P->work_len = need;
    }
    // The inverses X_k of the large fronts' L11: formed on first use, then
    // formed again only when L11 is not the one they were formed from.  The
    // exact comparison with a copy (k_l11_check) sets the gate that the
    // re-inversion's launches test on the device, so no call waits for the
    // host.  A zero pivot gives non-finite inverses and results, as a
    // triangular solve with the block would.
    const bool inv = !product && !S->chk.empty();
    auto refresh = [&](hipStream_t st, bool gated) -> int {
        hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(64), 0, st, S->d_changed);
        hipLaunchKernelGGL(k_l11_check, dim3((unsigned)S->chk.size()), dim3(256), 0, st, S->d_chk,
                           P->d_fi, blkval, S->d_l11, S->d_changed);
        return run_solve(P, S, S->inv, blkval, B, ldb, nrhs, alpha, st,
                         gated ? S->d_changed : nullptr);
    };
    bool fresh = false;
    if (inv && !S->inv_valid) {
        int rc = refresh(s, false);
        if (rc) return rc;
        S->inv_valid = true;
        fresh = true;
    }
    const std::vector<Launch>& seq = product ? (trans == 'N' ? S->mfwd : S->mbwd)
                                             : (trans == 'N' ? S->fwd : S->bwd);
    SolvePlan::Key key;
    key.blk = blkval; key.B = B; key.ldb = ldb; key.trans = trans + (product ? 256 : 0);
    key.alpha = alpha;
    SolvePlan::Graph* G = nullptr;
    for (auto& g : S->graphs) if (g.key == key) G = &g;
    if (P->use_graphs && !P->prof && (G || S->last == key)) {
        if (!P->gstream) {
            HIP_CHECK(hipStreamCreateWithFlags(&P->gstream, hipStreamNonBlocking));
            HIP_CHECK(hipEventCreateWithFlags(&P->gev_in, hipEventDisableTiming));
            HIP_CHECK(hipEventCreateWithFlags(&P->gev_out, hipEventDisableTiming));
        }
        if (ensure_side_stream(P)) return -2;  // created outside any capture
        if (!G) {
            if (S->graphs.size() >= 4) {
                HIP_CHECK(hipStreamSynchronize(s));
                (void)hipGraphExecDestroy(S->graphs.front().exec);
                S->graphs.erase(S->graphs.begin());
            }
            hipGraph_t graph = nullptr;
            HIP_CHECK(hipStreamBeginCapture(P->gstream, hipStreamCaptureModeThreadLocal));
            int rc = inv ? refresh(P->gstream, true) : 0;
            if (!rc) rc = run_solve(P, S, seq, blkval, B, ldb, nrhs, alpha, P->gstream);
            hipError_t e = hipStreamEndCapture(P->gstream, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            if (e != hipSuccess) { chp_set_error(std::string("graph capture: ") + hipGetErrorString(e)); return -2; }
            SolvePlan::Graph ng;
            ng.key = key;
            e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) { chp_set_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return -2; }
            S->graphs.push_back(ng);
            G = &S->graphs.back();
        }
        HIP_CHECK(hipEventRecord(P->gev_in, s));
        HIP_CHECK(hipStreamWaitEvent(P->gstream, P->gev_in, 0));
        HIP_CHECK(hipGraphLaunch(G->exec, P->gstream));
        HIP_CHECK(hipEventRecord(P->gev_out, P->gstream));
        HIP_CHECK(hipStreamWaitEvent(s, P->gev_out, 0));
    } else {
        S->last = key;
        int rc = (inv && !fresh) ? refresh(s, true) : 0;
        if (!rc) rc = run_solve(P, S, seq, blkval, B, ldb, nrhs, alpha, s);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace

extern "C" {

int chp_trsm(chp_plan* P, const double* blkval, double* B, int ldb, int nrhs, double alpha,
             int trans, void* stream) {
    return triangular_op(P, blkval, B, ldb, nrhs, alpha, trans, false, stream);
}

int chp_trmm(chp_plan* P, const double* blkval, double* B, int ldb, int nrhs, double alpha,
             int trans, void* stream) {
    return triangular_op(P, blkval, B, ldb, nrhs, alpha, trans, true, stream);
}

}  // extern "C"
