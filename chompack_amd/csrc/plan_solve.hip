// Triangular solve with the chordal factor, B := alpha L^{-1} B or
// alpha L^{-T} B (trsm.c:3-138 as called by cbase.c:965-1077), on the level
// schedule of a plan.
//
// Fronts with a small diagonal block and little separator work are solved by
// one workgroup each (k_solve_fwd / k_solve_bwd).  Larger fronts run as a
// short sequence per level, batched over the level's fronts:
//   forward:  gather (alpha B rows + children's update rows), x_A -= B21 x_N
//             on MFMA (B21 = L21 L11^{-1}, so x_N is used before its solve),
//             then x_N := L11^{-1} x_N by 64-row blocks: x_b := Dinv_b x_b,
//             x_below -= L[below, b] x_b; scatter x_N to B;
//   backward: gather (alpha B rows + the parent's separator values),
//             x_N := L11^{-T} x_N by blocks from the last, x_N -= B21^T x_A,
//             scatter.
// Dinv_b, the inverses of the 64 x 64 diagonal blocks of L11, are formed at
// the start of each solve from the factor passed in (k_diag_tasks, mode 1).
// The schedule depends on the number of right-hand sides and is cached per
// count.
#include "plan.h"
#include "../../include/chompack_amd.h"

#include <algorithm>

namespace chpi {
namespace {

bool solve_large(const FrontInfo& f, int nrhs) {
    return f.nn > 32 || (i64)f.na * f.nn * nrhs > 256 * 1024;
}

template <class T>
void push_list(chp_plan* P, std::vector<Launch>& seq, Kid kid, const std::vector<T>& v) {
    if (v.empty()) return;
    seq.push_back({kid, (int)v.size(), 256, 0, P->pg->lists.add(v), 0});
}

SolvePlan* build_solve(chp_plan* P, int nrhs) {
    SolvePlan* S = new SolvePlan();
    S->nrhs = nrhs;
    P->pg = &S->pg;
    const int H = P->nlevels;
    // diagonal-block inverses of all large fronts (mode 1, column-major, ld 64)
    std::vector<i64> dvo(P->fi.size(), -1);
    std::vector<DiagTask> dt;
    i64 dv = 0;
    for (size_t k = 0; k < P->fi.size(); ++k) {
        const FrontInfo& f = P->fi[k];
        if (!solve_large(f, nrhs)) continue;
        dvo[k] = dv;
        for (int b = 0; b * NB < f.nn; ++b) {
            DiagTask t;
            std::memset(&t, 0, sizeof(t));
            t.off = f.blkptr + (i64)b * NB * (1 + (i64)(f.nn + f.na));
            t.outoff = dv + (i64)b * NB * NB;
            t.ld = f.nn + f.na; t.ldo = NB; t.nb = std::min(NB, f.nn - b * NB);
            t.key = (int)k; t.col = b * NB;
            t.buf = 0; t.obuf = 3; t.mode = 1;
            dt.push_back(t);
        }
        dv += (i64)((f.nn + NB - 1) / NB) * NB * NB;
    }
    P->sdinv_cap = std::max(P->sdinv_cap, dv);  // requested; allocated by the caller
    push_list(P, S->fwd, K_DIAG, dt);
    S->bwd = S->fwd;
    auto wofs = [&](int k) { return P->h_wunit[k] * nrhs; };
    for (int L = 0; L < H; ++L) {
        std::vector<int> small, large;
        for (int k : P->bylevel[L]) (solve_large(P->fi[k], nrhs) ? large : small).push_back(k);
        push_list(P, S->fwd, K_SOLVE_FWD, small);
        if (large.empty()) continue;
        push_list(P, S->fwd, K_SGATHER_F, large);
        std::vector<GemmTile> g;
        int maxnb = 0;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            const int nn = f.nn, na = f.na, nj = nn + na;
            maxnb = std::max(maxnb, (nn + NB - 1) / NB);
            if (na > 0)  // x_A -= B21 x_N
                add_gemm_tiles(P, g, mk(0, f.blkptr + nn, nj, 0, 7, wofs(k), nj, 0, 7, wofs(k) + nn,
                                        nj, na, nrhs, nn, -1.0, 1.0, 0), false);
        }
        push_gemm(P, S->fwd, g);
        for (int b = 0; b < maxnb; ++b) {
            std::vector<GemmTile> gd, gu;
            for (int k : large) {
                const FrontInfo& f = P->fi[k];
                const int nn = f.nn, nj = nn + f.na;
                const int b0 = b * NB;
                if (b0 >= nn) continue;
                const int b1 = std::min(b0 + NB, nn);
                // x_b := Dinv_b x_b (in place: one tile row reads all of x_b first)
                add_gemm_tiles(P, gd, mk(3, dvo[k] + (i64)b * NB * NB, NB, 0, 7, wofs(k) + b0, nj, 0,
                                         7, wofs(k) + b0, nj, b1 - b0, nrhs, b1 - b0, 1.0, 0.0, 0, 1),
                               false);
                if (b1 < nn)  // x[b1:nn] -= L[b1:nn, b0:b1] x_b
                    add_gemm_tiles(P, gu, mk(0, f.blkptr + b1 + (i64)b0 * nj, nj, 0, 7, wofs(k) + b0,
                                             nj, 0, 7, wofs(k) + b1, nj, nn - b1, nrhs, b1 - b0,
                                             -1.0, 1.0, 0), false);
            }
            push_gemm(P, S->fwd, gd);
            push_gemm(P, S->fwd, gu);
        }
        push_list(P, S->fwd, K_SSCATTER, large);
    }
    for (int L = H - 1; L >= 0; --L) {
        std::vector<int> small, large;
        for (int k : P->bylevel[L]) (solve_large(P->fi[k], nrhs) ? large : small).push_back(k);
        push_list(P, S->bwd, K_SOLVE_BWD, small);
        if (large.empty()) continue;
        push_list(P, S->bwd, K_SGATHER_B, large);
        int maxnb = 0;
        for (int k : large) maxnb = std::max(maxnb, (P->fi[k].nn + NB - 1) / NB);
        for (int bb = 0; bb < maxnb; ++bb) {
            std::vector<GemmTile> gd, gu;
            for (int k : large) {
                const FrontInfo& f = P->fi[k];
                const int nn = f.nn, nj = nn + f.na;
                const int nb = (nn + NB - 1) / NB;
                if (bb >= nb) continue;
                const int b = nb - 1 - bb, b0 = b * NB, b1 = std::min(b0 + NB, nn);
                // x_b := Dinv_b^T x_b
                add_gemm_tiles(P, gd, mk(3, dvo[k] + (i64)b * NB * NB, NB, 1, 7, wofs(k) + b0, nj, 0,
                                         7, wofs(k) + b0, nj, b1 - b0, nrhs, b1 - b0, 1.0, 0.0, 0, 2),
                               false);
                if (b0 > 0)  // x[0:b0] -= L[b0:b1, 0:b0]^T x_b
                    add_gemm_tiles(P, gu, mk(0, f.blkptr + b0, nj, 1, 7, wofs(k) + b0, nj, 0,
                                             7, wofs(k), nj, b0, nrhs, b1 - b0, -1.0, 1.0, 0), false);
            }
            push_gemm(P, S->bwd, gd);
            push_gemm(P, S->bwd, gu);
        }
        std::vector<GemmTile> g;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            const int nn = f.nn, na = f.na, nj = nn + na;
            if (na > 0)  // x_N -= B21^T x_A
                add_gemm_tiles(P, g, mk(0, f.blkptr + nn, nj, 1, 7, wofs(k) + nn, nj, 0, 7, wofs(k),
                                        nj, nn, nrhs, na, -1.0, 1.0, 0), false);
        }
        push_gemm(P, S->bwd, g);
        push_list(P, S->bwd, K_SSCATTER, large);
    }
    P->pg = &P->main;
    return S;
}

int run_solve(chp_plan* P, const SolvePlan* S, const std::vector<Launch>& seq,
              const double* blkval, double* B, int ldb, int nrhs, double alpha, hipStream_t s) {
    BufTab bt;
    for (int i = 0; i < 8; ++i) bt.p[i] = nullptr;
    bt.p[0] = const_cast<double*>(blkval);
    bt.p[3] = P->d_sdinv;
    bt.p[4] = S->d_splitp;
    bt.p[7] = P->d_work;
    for (const Launch& L : seq) {
        const void* list = S->pg.d_lists + L.list;
        const dim3 g(L.grid), b(L.block);
        switch (L.kid) {
        case K_SOLVE_FWD:
            hipLaunchKernelGGL(k_solve_fwd, g, b, 0, s, (const int*)list, P->d_fi, blkval, P->d_snptr,
                               P->d_relidx, P->d_chidx, P->d_perm, B, ldb, nrhs, alpha, P->d_work,
                               P->d_wunit);
            break;
        case K_SOLVE_BWD:
            hipLaunchKernelGGL(k_solve_bwd, g, b, 0, s, (const int*)list, P->d_fi, blkval, P->d_snptr,
                               P->d_relidx, P->d_perm, B, ldb, nrhs, alpha, P->d_work, P->d_wunit);
            break;
        case K_SGATHER_F:
            hipLaunchKernelGGL(k_sgather_fwd, g, b, 0, s, (const int*)list, P->d_fi, P->d_snptr,
                               P->d_relidx, P->d_chidx, P->d_perm, B, ldb, nrhs, alpha, P->d_work,
                               P->d_wunit);
            break;
        case K_SGATHER_B:
            hipLaunchKernelGGL(k_sgather_bwd, g, b, 0, s, (const int*)list, P->d_fi, P->d_snptr,
                               P->d_relidx, P->d_perm, B, ldb, nrhs, alpha, P->d_work, P->d_wunit);
            break;
        case K_SSCATTER:
            hipLaunchKernelGGL(k_sscatter, g, b, 0, s, (const int*)list, P->d_fi, P->d_snptr,
                               P->d_perm, B, ldb, nrhs, P->d_work, P->d_wunit);
            break;
        case K_DIAG:
            hipLaunchKernelGGL(k_diag_tasks, g, b, 0, s, (const DiagTask*)list, bt, P->d_status);
            break;
        case K_GEMM:
            hipLaunchKernelGGL(k_gemm_tiles, g, b, 0, s, (const GemmTile*)list, S->pg.d_tasks, bt);
            break;
        case K_SPLITK_REDUCE:
            hipLaunchKernelGGL(k_splitk_reduce, g, b, 0, s, (const GemmTile*)list, S->pg.d_tasks, bt);
            break;
        default:
            chp_set_error("internal: unexpected launch kind in solve");
            return -3;
        }
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace
}  // namespace chpi

using namespace chpi;

extern "C" {

int chp_trsm(chp_plan* P, const double* blkval, double* B, int ldb, int nrhs, double alpha,
             int trans, void* stream) {
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
        if (alloc_doubles(&S->d_splitp, S->pg.splitp_len)) return -2;
    }
    const i64 need = P->solve_len * nrhs;
    if (need > P->work_len) {
        HIP_CHECK(hipStreamSynchronize(s));
        if (P->d_work) HIP_CHECK(hipFree(P->d_work));
        P->d_work = nullptr;
        HIP_CHECK(hipMalloc((void**)&P->d_work, need * sizeof(double)));
        P->work_len = need;
    }
    if (P->sdinv_cap > P->dinv_len_solve) {
        HIP_CHECK(hipStreamSynchronize(s));
        if (P->d_sdinv) HIP_CHECK(hipFree(P->d_sdinv));
        P->d_sdinv = nullptr;
        if (alloc_doubles(&P->d_sdinv, P->sdinv_cap)) return -2;
        P->dinv_len_solve = P->sdinv_cap;
    }
    const std::vector<Launch>& seq = trans == 'N' ? S->fwd : S->bwd;
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, s, P->d_status);
    int rc = run_solve(P, S, seq, blkval, B, ldb, nrhs, alpha, s);
    if (rc) return rc;
    i64 info[2];
    if (read_status(P, s, info) != 0) {
        chp_set_error("trsm: singular diagonal block in supernode " + std::to_string(info[0]));
        return -4;
    }
    return 0;
}

}  // extern "C"
