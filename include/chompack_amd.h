/*
 * chompack_amd C-ABI: MI355X (gfx950) supernodal multifrontal numeric sweeps.
 *
 * This is the boundary a CHOMPACK-style host binds (ctypes from Python, see
 * chompack_amd/_native.py).  All index arrays are int64 host arrays with the
 * meaning of the reference's symbolic object (src/python/symbolic.py:499-890);
 * all value arrays (blkval, B, U, F) are float64 DEVICE pointers in HBM;
 * `stream` is a hipStream_t (NULL = default stream).  No torch types appear.
 *
 * Return codes: 0 success; 1 numerical failure (see info); negative = usage or
 * runtime error, with a message from chp_last_error().
 */
#ifndef CHOMPACK_AMD_H
#define CHOMPACK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chp_plan chp_plan;

/* Last error message of the calling thread. */
const char* chp_last_error(void);

/* ABI version of this library. */
int chp_version(void);

/* ---------------------------------------------------------------------------
 * Host-side symbolic routines (integer, bit-exact with the reference).
 * ------------------------------------------------------------------------- */

/* Elimination tree (symbolic.py:49-72).  cp/ri: CSC of the symmetric pattern. */
void chp_etree(int64_t n, const int64_t* cp, const int64_t* ri, int64_t* parent);
/* Postorder of a forest, roots have parent[j] == j (symbolic.py:26-47). */
void chp_post_order(int64_t n, const int64_t* parent, int64_t* post);
/* Column counts of the Cholesky factor (symbolic.py:95-131). */
void chp_counts(int64_t n, const int64_t* cp, const int64_t* ri, const int64_t* parent,
                const int64_t* post, int64_t* colcount);
/* Pothen-Sun maximal supernodes (symbolic.py:133-244); returns Nsn. */
int64_t chp_supernodes(int64_t n, const int64_t* parent, const int64_t* post,
                       const int64_t* colcount, int64_t* snode, int64_t* snptr,
                       int64_t* snpar);
/* Clique row sets (symbolic.py:328-384); returns 0 or -1-k on a size mismatch. */
int64_t chp_embed(int64_t nsn, const int64_t* cpl, const int64_t* ril, const int64_t* snptr,
                  const int64_t* snpar, const int64_t* sncolptr, int64_t* snrowidx);
/* Sorted merge of left[0..nl) and right[0..nr) into out (room for nl + nr);
 * equal heads are taken once.  Returns the merged length.  Replaces lmerge
 * (cbase.c:64-101; Python twin misc.py:178-204), whose caller copies out
 * back into left. */
int64_t chp_lmerge(const int64_t* left, const int64_t* right, int64_t nl, int64_t nr,
                   int64_t* out);
/* Relative indices (symbolic.py:386-411); returns 0 or -1-k. */
int64_t chp_relative_idx(int64_t nsn, const int64_t* snptr, const int64_t* snpar,
                         const int64_t* sncolptr, const int64_t* snrowidx,
                         const int64_t* relptr, int64_t* relidx);
/* Maximum cardinality search ordering (mcs.py:4-57); cp/ri: the full
 * symmetric pattern (CCS of symmetrize(A), rows ascending per column); ve:
 * the vertex eliminated last.  Writes p[0..n-1]. */
void chp_maxcardsearch(int64_t n, const int64_t* cp, const int64_t* ri, int64_t ve, int64_t* p);
/* Maximal chordal subgraph (maxchord.py:5-83): writes the ordering p and the
 * kept lower-triangular entries (rows[i], cols[i]; src[i] = position of the
 * entry in ri, so the caller copies its value); returns their number
 * (at most cp[n]). */
int64_t chp_maxchord(int64_t n, const int64_t* cp, const int64_t* ri, int64_t ve, int64_t* p,
                     int64_t* rows, int64_t* cols, int64_t* src);
/* Approximate minimum degree ordering; replaces cvxopt.amd.order as called
 * through symbolic(A, p=amd.order) (symbolic.py:540-546). */
int chp_amd_order(int64_t n, const int64_t* cp, const int64_t* ri, int64_t* perm);

/* ---------------------------------------------------------------------------
 * Device plan: the symbolic arrays of one factorization in HBM plus the
 * precomputed level schedule.  Replaces the per-call argument extraction and
 * workspace malloc of cbase.c (e.g. cchol, cbase.c:187-258).
 * small_max: largest clique order handled by the LDS-resident front kernels
 * (<= 0 selects the default, 64).
 * ------------------------------------------------------------------------- */
chp_plan* chp_plan_create(int64_t n, int64_t nsn, const int64_t* snpost,
                          const int64_t* snptr, const int64_t* relptr,
                          const int64_t* relidx, const int64_t* chptr,
                          const int64_t* chidx, const int64_t* blkptr,
                          const int64_t* p, int small_max);
/* The same with tunable parameters (no reference counterpart; chosen per
 * plan by chompack_amd.autotune): params[0] small_max (as above), [1] outer
 * panel width of the large-front factorization (a multiple of 64, default
 * 256), [2] cap of the split-K factor of GEMM tasks (1 = no split-K,
 * default 16), [3] LDS bytes a small-front workgroup may use (default
 * 65536, capped by the device), [4] the large fronts of the FP64 cholesky:
 * 1 one tile-DAG launch per level, 2 panel and GEMM launches (default);
 * entries <= 0 or beyond nparams keep their defaults. */
chp_plan* chp_plan_create_params(int64_t n, int64_t nsn, const int64_t* snpost,
                                 const int64_t* snptr, const int64_t* relptr,
                                 const int64_t* relidx, const int64_t* chptr,
                                 const int64_t* chidx, const int64_t* blkptr,
                                 const int64_t* p, const int* params, int nparams);
/* out[5]: the parameters of a plan (small_max, panel width, split-K cap, LDS
 * cap, large-front cholesky 1 tile DAG / 2 panels). */
int chp_plan_params(chp_plan* plan, int* out);
void chp_plan_destroy(chp_plan* plan);
/* Update-arena layout of a plan (host only, no device call): updptr[k] =
 * offset of supernode k's na_k x na_k update matrix, level_out[k] = its
 * level (height in the supernodal elimination tree; either output may be
 * NULL); returns the arena length in doubles.  Slot k is live over the
 * levels [level(k), level(parent(k))]; slots live at a common level are
 * disjoint, the others share memory, so the arena holds the largest live
 * set instead of sum(na^2).  Replaces the reference's LIFO stack sizing
 * stack_mem (symbolic.py:634-680), which assumes a sequential postorder. */
int64_t chp_update_layout(int64_t nsn, const int64_t* snpost, const int64_t* relptr,
                          const int64_t* chptr, const int64_t* chidx, int64_t* updptr,
                          int64_t* level_out);
/* out[8]: levels, large fronts (cholesky), large fronts (projected inverse),
 * GEMM tiles, update-arena doubles, diagonal-inverse doubles, scratch doubles,
 * cholesky launches. */
int chp_plan_info(chp_plan* plan, int64_t* out);
/* on != 0 (the default): each of the cholesky / projected-inverse / llt
 * sweeps times its first calls launched one by one and replayed from a HIP
 * graph, and keeps the faster way (calls 1-2 eager, 3-4 graph, decided from
 * call 5 on); on == 0: always one by one.  Graphs are captured per blkval
 * pointer and run on a plan-owned stream ordered after and before the
 * caller's stream with events. */
int chp_set_graphs(chp_plan* plan, int on);

/* Measurement support (no counterpart in the reference).  chp_profile(plan, 1)
 * times every launch of the sweeps with HIP events on the caller's stream
 * (synchronizing after each launch, every launch on that stream) and clears
 * the counters; chp_profile_read fills ms/cnt/flops[min(max, 28)] per
 * kernel kind (0 chol_small, 1 extadd, 2 potrf_diag, 3 panel_factor, 4 gemm,
 * 5 extract, 6 projinv_small, 7 trtri_diag, 8 neg_copy, 9 llt_small,
 * 10 solve_fwd, 11 solve_bwd, 12 copy_top, 13 splitk_reduce, 14 diag,
 * 15 matop, 16 update_factor, 17 sgather_fwd, 18 sgather_bwd, 19 sscatter,
 * 20 trmm_fwd, 21 trmm_bwd, 22 trmm_finish, 23 completion_chain,
 * 24 projinv_chain, 25 chol_dag, 26 fill_int, 27 gemm_small) and returns
 * the number of kinds (28).
 * chp_plan_flops: algorithmic GEMM flops of one cholesky (out[0]) and one
 * projected-inverse sweep (out[1]). */
int chp_profile(chp_plan* plan, int on);
int chp_profile_read(chp_plan* plan, double* ms, int64_t* cnt, double* flops, int max);
/* Per-task trace of the tile-DAG launches of the cholesky sweep (measurement
 * support): chp_dag_trace(plan, 1) allocates a device buffer of 4 int64 per
 * task that later sweeps fill (kind | k << 8 | i << 24 | j << 40 |
 * workgroup << 52; the 100 MHz wall clock when the task was taken, when its
 * dependencies were met, when it finished), 0 frees it; both return the
 * number of tasks.  chp_dag_trace_read copies min(max, tasks) rows to out. */
int64_t chp_dag_trace(chp_plan* plan, int on);
int64_t chp_dag_trace_read(chp_plan* plan, int64_t* out, int64_t max);
/* Bound of every dependency wait of the tile-DAG launches, in polls (default
 * 262144, about a quarter of a second); a wait that runs out makes the sweep
 * return -3 ("internal: a dependency wait ... timed out") after the launch
 * has drained.  For tests of that error path. */
int chp_set_dag_wait_limit(chp_plan* plan, int polls);
/* Per-launch rows (kind, workgroups, algorithmic flops, ms) of the profiled
 * sweeps: copies min(rows, max) rows into out[4*max], returns the row count. */
int64_t chp_profile_launches(chp_plan* plan, double* out, int64_t max);
int chp_plan_flops(chp_plan* plan, double* out);
/* out[0..3]: algorithmic bytes of the cholesky / projected-inverse GEMM
 * launches (A, B, C read once, C written once) and their launch counts. */
int chp_plan_gemm_bytes(chp_plan* plan, double* out);

/* ---------------------------------------------------------------------------
 * Numeric sweeps, in place on a device blkval array of length blkptr[nsn].
 * ------------------------------------------------------------------------- */

/* Supernodal multifrontal Cholesky; replaces cholesky() (chompack.h:63-76,
 * cholesky.c:3-84) as called by cbase.c:262-266.  On failure returns 1 with
 * info[0] = supernode, info[1] = LAPACK-style column info within it. */
int chp_cholesky(chp_plan* plan, double* blkval, void* stream, int64_t* info);

/* Chordal product X = L L^T; replaces llt() (llt.c:3-60, cbase.c:364-368). */
int chp_llt(chp_plan* plan, double* blkval, void* stream);

/* Cholesky with the Schur-complement updates (the trailing and inner-panel
 * updates of large fronts) in FP32 arithmetic on MFMA; diagonal blocks,
 * panels, the stored-convention solves and small fronts stay FP64.  The
 * factor is accurate to about FP32 precision; pair it with iterative
 * refinement (chp_spmv_residual + chp_trsm).  No reference counterpart. */
int chp_cholesky_mixed(chp_plan* plan, double* blkval, void* stream, int64_t* info);
/* The same with a choice of Schur-update arithmetic: schur = 1 FP32 (as
 * chp_cholesky_mixed), 2 BF16x3 split precision (each operand rounded to
 * FP32 and split into bf16 hi + lo; lo*hi + hi*lo + hi*hi on
 * v_mfma_f32_16x16x16bf16_1k with FP32 accumulation, about 2^-16 relative
 * per product).  No reference counterpart. */
int chp_cholesky_lowprec(chp_plan* plan, double* blkval, int schur, void* stream,
                         int64_t* info);

/* Reverse-mode sweep of chp_cholesky (cholesky.c:26-82 backwards, root to
 * leaves; no reference counterpart: the backward of a differentiable
 * cholesky).  lblkval: the factor in the stored convention; gblkval: the
 * gradient of a scalar loss with respect to that stored factor (lower
 * trapezoid of each block read); out (blkval length, distinct from both):
 * per supernode the lower trapezoid of the symmetric gradient of its
 * frontal matrix.  The gradient with respect to the lower-stored entries of
 * X is out with the strictly-lower entries doubled.  Returns 1 on a zero
 * pivot of L11 (info as for chp_cholesky). */
int chp_cholesky_adjoint(chp_plan* plan, const double* lblkval, const double* gblkval,
                         double* out, void* stream, int64_t* info);

/* Projected inverse Y = P(X^{-1}) from the factor; replaces
 * projected_inverse() (projected_inverse.c:3-75, cbase.c:470-474). */
int chp_projected_inverse(chp_plan* plan, double* blkval, void* stream, int64_t* info);

/* Maximum-determinant positive definite completion: on exit blkval holds the
 * factor L with P((L L^T)^{-1}) = X; replaces completion() (completion.c:3-120,
 * update_factor.c:3-108) as called by cbase.c:521-621.  factored != 0 selects
 * Householder-factored update matrices.  Returns 1 on a failed factorization
 * (info as for chp_cholesky). */
int chp_completion(chp_plan* plan, double* blkval, int factored, void* stream, int64_t* info);

/* Hessian factors of the log-det barrier applied in place to nu device blkval
 * arrays ublkval[0..nu-1]; lblkval is the factor L, yblkval the projected
 * inverse Y.  adj: 0 G, 1 G^adj, -1 both (G^adj G, the reference's adj=None);
 * inv != 0 applies the inverses.  Replaces hessian() (hessian.c:3-374) as
 * called by cbase.c:675-857, including its second pass for adj=None
 * (cbase.c:833-842).  A failed update factorization is reported in info but,
 * as in the reference, not as an error. */
int chp_hessian(chp_plan* plan, const double* lblkval, const double* yblkval,
                double* const* ublkval, int nu, int adj, int inv, int factored,
                void* stream, int64_t* info);

/* B := alpha L^{-1} B (trans 'N') or alpha L^{-T} B (trans 'T'), B a device
 * column-major n x nrhs array with leading dimension ldb in the original
 * ordering; replaces trsm() (trsm.c:3-138, cbase.c:1042-1049).  Asynchronous
 * on `stream` (no host round trip once the schedule for nrhs exists); a zero
 * pivot gives non-finite entries in B, as the reference's dtrsm does, not an
 * error. */
int chp_trsm(chp_plan* plan, const double* blkval, double* B, int ldb, int nrhs,
             double alpha, int trans, void* stream);

/* B := alpha L B (trans 'N') or alpha L^T B (trans 'T'), B as for chp_trsm;
 * replaces trmm() (pybase/trmm.py:4-114, the reference's only
 * implementation; chompack.trmm is that function). */
int chp_trmm(chp_plan* plan, const double* blkval, double* B, int ldb, int nrhs,
             double alpha, int trans, void* stream);

/* *out = tr(X Y) for two chordal matrices on the plan's pattern; replaces
 * dot() (dot.c:3-22, cbase.c:106-174). */
int chp_dot(chp_plan* plan, const double* x, const double* y, double* out, void* stream);

/* *out = log det(L L^T) = 2 sum_j log L(j, j) from a factor; returns 1 when a
 * diagonal entry is not positive.  The reference computes it from
 * cspmatrix.diag() of the factor (symbolic.py:1113-1128); no kernel there. */
int chp_logdet(chp_plan* plan, const double* lblkval, double* out, void* stream);
/* out[j] = log L(j, j) for every (reordered) column j, out a device array of
 * length n; per-instance log-determinants of a block-diagonal batch are
 * segment sums of it (no reference counterpart; batching support). */
int chp_logdiag(chp_plan* plan, const double* lblkval, double* out, void* stream);

/* R = B - A X for a device CSR matrix (rowptr int64[n+1], colind int32,
 * vals float64; the full symmetric pattern) and column-major device arrays
 * X, B, R with nrhs columns: the FP64 residual of iterative refinement
 * (no reference counterpart).  work: NULL, or a device workspace of n * nrhs
 * doubles that receives a row-major copy of X, so that every nonzero reads
 * a contiguous row of right-hand sides (used when nrhs > 1). */
int chp_spmv_residual(int64_t n, const int64_t* rowptr, const int* colind, const double* vals,
                      const double* X, int ldx, const double* B, int ldb, double* R, int ldr,
                      int nrhs, double* work, void* stream);

/* The same residual without any allocation and without a work argument, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight refinement loops, hosts that forbid hipMalloc).  No counterpart in
 * the reference.  The row-major copy of X lives in a fixed device array of
 * CHP_RESID_NOALLOC_MAX doubles (64 MiB, held by each GPU the library runs
 * on until the process ends); n * nrhs above that returns -1 with a message.
 * Calls that overlap in time on one GPU share the array and race: serialize
 * them.  Results are bit-identical to chp_spmv_residual. */
#define CHP_RESID_NOALLOC_MAX 8388608
int chp_spmv_residual_noalloc(int64_t n, const int64_t* rowptr, const int* colind,
                              const double* vals, const double* X, int ldx, const double* B,
                              int ldb, double* R, int ldr, int nrhs, void* stream);

/* ---------------------------------------------------------------------------
 * Product-form Cholesky factors of I + V diag(a) V^T (pfcholesky.c), V, L, B
 * device column-major n x m arrays, a a host array of m doubles.  Each factor
 * L_i = diag(l_i) + strictly-lower(v_i b_i^T); the sequential row loops of
 * the reference run as scans, one workgroup per column.
 * ------------------------------------------------------------------------- */
/* Factorization in place (V's columns i+1.. are overwritten as in the
 * reference); replaces dpftrf (pfcholesky.c:52-87) as called by cbase.c
 * pfchol (cbase.c:862-880).  Returns 1 where the reference fails
 * ("factorization failed"). */
int chp_pf_factor(int64_t n, int m, const double* a, double* V, int ldv, double* L, int ldl,
                  double* B, int ldb, void* stream);
/* U := (L_1 ... L_m)^{-1} U (trans 'N') or (L_1 ... L_m)^{-T} U ('T') for nu
 * device columns of U; replaces dpfsv (pfcholesky.c:89-112) as called by
 * cbase.c pftrsm (cbase.c:885-911). */
int chp_pf_solve(int64_t n, int m, const double* V, int ldv, const double* L, int ldl,
                 const double* B, int ldb, double* U, int ldu, int nu, int trans, void* stream);
/* U := L_1 ... L_m U ('N') or (L_1 ... L_m)^T U ('T'); replaces dpfmv
 * (pfcholesky.c:114-136) as called by cbase.c pftrmm (cbase.c:916-942). */
int chp_pf_product(int64_t n, int m, const double* V, int ldv, const double* L, int ldl,
                   const double* B, int ldb, double* U, int ldu, int nu, int trans,
                   void* stream);

/* Maximum-determinant positive semidefinite completion; replaces
 * psdcompletion (pybase/psdcompletion.py:4-70; exported by
 * src/python/__init__.py:29,32).  X (device, n x n, ld ldx, zero on entry)
 * receives the dense completion of the cspmatrix (snptr, sncolptr,
 * snrowidx, blkptr: host symbolic arrays; blkval: device) in the reordered
 * numbering; tol: the pseudo-inverse threshold of the eigendecomposition
 * branch (the reference's kwargs tol, default 1e-15).  Cliques up to 2048
 * rows.  Synchronizes the stream. */
int chp_psdcompletion(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr,
                      const int64_t* snrowidx, const int64_t* blkptr, const double* blkval,
                      double* X, int64_t ldx, double tol, void* stream);
/* Euclidean distance matrix completion; replaces edmcompletion
 * (pybase/edmcompletion.py:4-74; src/python/__init__.py:29,32).  Same
 * arguments and conventions as chp_psdcompletion; every separator takes the
 * eigendecomposition pseudo-inverse of the Gram transform. */
int chp_edmcompletion(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr,
                      const int64_t* snrowidx, const int64_t* blkptr, const double* blkval,
                      double* X, int64_t ldx, double tol, void* stream);
/* Minimum-rank positive semidefinite completion; replaces mrcompletion
 * (pybase/mrcompletion.py:6-108; src/python/__init__.py:29,32).  Y (device,
 * n x maxr, ld ldy; maxr = the clique number) receives Y with
 * P(Y Y^T) = A in its first *rank columns (reordered numbering; the other
 * columns are zero); snpar: the supernodal parent array (root: snpar[k]
 * == k).  Cliques up to 2048 rows.  Synchronizes the stream. */
int chp_mrcompletion(int64_t n, int64_t nsn, const int64_t* snptr, const int64_t* sncolptr,
                     const int64_t* snrowidx, const int64_t* blkptr, const int64_t* snpar,
                     const double* blkval, double* Y, int64_t ldy, int maxr, int* rank,
                     void* stream);
/* out[a, b] = X[ip[a], permute_cols ? ip[b] : b] for a < rows, b < cols
 * (device arrays, column-major): the `reordered=False` return of both
 * completions (psdcompletion.py:67-70, mrcompletion.py:176-179). */
int chp_dense_permute(const double* X, int64_t ldx, const int64_t* ip, int64_t rows,
                      int64_t cols, int permute_cols, double* out, int64_t ldo, void* stream);

/* out := the factor's blkval with every (2,1) block B21 = L21 L11^{-1}
 * replaced by L21 = B21 L11, i.e. the blocks of the lower-triangular L of
 * P A P^T; replaces the per-supernode trmm of a factor's export in
 * cspmatrix.spmatrix() (symbolic.py:1052-1068).  lblkval and out are device
 * arrays of the plan's blkval length (16-byte aligned).  Synchronizes the
 * stream. */
int chp_factor_export(chp_plan* plan, const double* lblkval, double* out, void* stream);

/* dst[0:n] = src[0:n] for device float64 arrays (16-byte aligned); used to
 * restore a blkval array from a copy (no reference counterpart). */
int chp_copy(double* dst, const double* src, int64_t n, void* stream);

/* F[r[i], r[j]] += alpha U[i, j] for i >= j, with element (a, b) of F at
 * F + a*strides[0] + b*strides[1] and of U at U + a*strides[2] + b*strides[3];
 * r is a device int64 array of length N.  Replaces frontal_add_update
 * (cbase.c:11-34). */
int chp_frontal_add_update(double* F, int64_t N, const int64_t* relidx, const double* U,
                           const int64_t* strides, double alpha, void* stream);

/* U[i, j] = F[r[i], r[j]] for i >= j (same conventions); replaces
 * frontal_get_update (cbase.c:39-59). */
int chp_frontal_get_update(const double* F, int64_t N, const int64_t* relidx, double* U,
                           const int64_t* strides, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CHOMPACK_AMD_H */
