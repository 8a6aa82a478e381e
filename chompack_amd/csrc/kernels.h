// Kernel declarations shared between the kernel sources and the host
// scheduler (plan.hip).
#pragma once
#include "common.h"

#define NB 64          // inner panel width / diagonal block order of the large-front algorithms
#define NB2 256        // outer panel width (K of the big trailing updates)
#define TILE 64        // GEMM output tile (TILE x TILE)
// LDS of a tile-DAG workgroup: L and its inverse (64 x 65 each), the pivot
// exchange arrays (2 x 64 x 5), 1/diag (64), and the task index
#define DAG_LDS_DOUBLES (2 * NB * (NB + 1) + 2 * NB * 5 + NB)
#define DAG_LDS_BYTES (DAG_LDS_DOUBLES * 8 + 16)
#define SOLVE_RHS_CHUNK 64  // right-hand sides per workgroup of the slab gather/scatter (grid y)

__global__ void k_status_init(unsigned long long* status);

__global__ void k_chol_small(const int* fronts, int count, const FrontInfo* fi, double* blkval,
                             double* upd, const int* relidx, const int* chidx,
                             unsigned long long* status);
__global__ void k_extract(const int4* tiles, const FrontInfo* fi, double* blkval,
                          double* upd, const int* relidx, i64 bstride, i64 ustride,
                          const double* scratch, const i64* tptr);
__global__ void k_projinv_small(const int* fronts, int count, const FrontInfo* fi, double* blkval,
                                const double* upd, unsigned long long* status);
// chains of small fronts (nj <= CHAIN_NJ), one workgroup walking a
// ChainFront list: cholesky and llt child first, the projected inverse
// parent first (children's V extracted, the chain child's kept in LDS)
#define CHAIN_NJ 64
#define CHAIN_TPB 1024
__global__ void k_chol_chain(const ChainFront* cf, int count, const FrontInfo* fi, double* blkval,
                             double* upd, const int* relidx, const int* chidx,
                             unsigned long long* status);
__global__ void k_llt_chain(const ChainFront* cf, int count, const FrontInfo* fi, double* blkval,
                            double* upd, const int* relidx, const int* chidx);
__global__ void k_projinv_chain(const ChainFront* cf, int count, const FrontInfo* fi,
                                double* blkval, double* upd, const int* relidx, const int* chidx,
                                unsigned long long* status);
__global__ void k_llt_small(const int* fronts, int count, const FrontInfo* fi, double* blkval,
                            double* upd, const int* relidx, const int* chidx);

// large fronts
__global__ void k_extadd_tiles(const int4* tiles, const FrontInfo* fi, double* blkval, double* upd,
                               const int* relidx, const int* chidx, int zero22, i64 bstride,
                               i64 ustride);
__global__ void k_copy_top(const int2* tiles, const FrontInfo* fi, const double* blkval,
                           double* scratch, const i64* sptr);
__global__ void k_potrf_diag(const int* fronts, int c0, const FrontInfo* fi, double* blkval,
                             double* dinv, const i64* dinvptr, unsigned long long* status);
__global__ void k_panel_factor(const int4* tasks, const FrontInfo* fi, double* blkval,
                               double* dinv, const i64* dinvptr, unsigned long long* status);
__global__ void k_trtri_diag(const int2* blocks, const FrontInfo* fi, const double* blkval,
                             double* dinv, const i64* dinvptr, double* X, const i64* xptr,
                             unsigned long long* status);
// the tile-DAG factorization of one level's large fronts (persistent);
// list: int4 (ntasks, nfronts, queue counter index, 0), DagFront[nfronts],
// DagTask[ntasks]
// trace (optional, else null): 4 int64 per task, see chp_dag_trace
__global__ void k_chol_dag(const int4* list, double* blkval, double* upd, double* dinv, int* cnt,
                           unsigned long long* status, int spin_limit, i64* trace);
__global__ void k_fill_int(int* p, i64 n, int v);
__global__ void k_gemm_tiles(const GemmTile* tiles, const GemmTask* tasks, BufTab bt);
// launches of at most LAT_TILES tiles (split-K slices included): 512 threads, K split
// between two wave groups (the Launch's block size selects it)
#define LAT_TILES 256
// single-tile GEMM tasks with K at most this go to k_gemm_small (a workgroup each)
#define SMALL_TASK_K 32
__global__ void k_gemm_small(const int* ids, int count, const GemmTask* tasks, BufTab bt);
__global__ void k_gemm_tiles_lat(const GemmTile* tiles, const GemmTask* tasks, BufTab bt);
__global__ void k_splitk_reduce(const GemmTile* tiles, const GemmTask* tasks, BufTab bt);
// tasks with N <= NARROW_N (opA, opB in {0, 1}, no triangular output or B;
// a triangular A with opB 0): a 64 x N tile per workgroup, K split over its
// four waves
#define NARROW_N 16
__global__ void k_gemm_narrow(const GemmTile* tiles, const GemmTask* tasks, BufTab bt);
__global__ void k_gemm_tiles32(const GemmTile* tiles, const GemmTask* tasks, BufTab bt);
__global__ void k_gemm_tiles_bf3(const GemmTile* tiles, const GemmTask* tasks, BufTab bt);
// descriptor-addressed dense operations (completion, Hessian)
__global__ void k_diag_tasks(const DiagTask* tasks, BufTab bt, unsigned long long* status);
__global__ void k_matop(const MatOp* ops, BufTab bt);
__global__ void k_update_factor(const UFTask* tasks, BufTab bt, const int* relidx,
                                unsigned long long* status);
__global__ void k_completion_chain(const ChainItem* items, int nitems, const FrontInfo* fi,
                                   BufTab bt, const int* relidx, unsigned long long* status);
// large update-factor tasks on the whole device: the element copies (list:
// UFTask[ntask] then int4 (task, column, N1, 0) items), then launches that
// each run one block of reflections per task in LDS and apply the previous
// block's compact WY form to 16-column tiles of the columns left of it
#define UF_VCAP 12288  // len x nbk of a block (its V staged in LDS)
__global__ void k_uf_setup(const char* list, int ntask, BufTab bt, const int* relidx);
__global__ void k_uf_step(const UFBlock* items, BufTab bt, const int* relidx);
// N <= 64 and na <= 64; 64-thread workgroups
__global__ void k_update_factor_small(const UFTask* tasks, BufTab bt, const int* relidx,
                                      unsigned long long* status);

// triangular solve with the chordal factor
__global__ void k_solve_fwd(const int2* items, const FrontInfo* fi, const double* blkval,
                            const i64* snptr, const i64* perm, double* B, int ldb, int nrhs,
                            double alpha, double* work, const i64* wptr, const int* gptr,
                            const GEnt* gent);
__global__ void k_solve_bwd(const int2* items, const FrontInfo* fi, const double* blkval,
                            const i64* snptr, const int* relidx, const i64* perm,
                            double* B, int ldb, int nrhs, double alpha, double* work,
                            const i64* wptr);
__global__ void k_flag_clear(int* flag);
__global__ void k_l11_check(const CheckTile* tiles, const FrontInfo* fi, const double* blkval,
                            double* copy, int* changed);

__global__ void k_sgather_fwd(const int2* tiles, const FrontInfo* fi, const i64* snptr,
                              const i64* perm, const double* B, int ldb, int nrhs, double alpha,
                              double* work, const i64* wptr, const int* gptr, const GEnt* gent,
                              int bonly);
__global__ void k_trmm_fwd(const int2* items, const FrontInfo* fi, const double* blkval,
                           const i64* snptr, const i64* perm, double* B, int ldb, int nrhs,
                           double alpha, double* work, const i64* wptr, const int* gptr,
                           const GEnt* gent);
__global__ void k_trmm_bwd(const int2* items, const FrontInfo* fi, const double* blkval,
                           const i64* snptr, const int* relidx, const i64* perm, double* B,
                           int ldb, int nrhs, double alpha, double* work, const i64* wptr);
__global__ void k_trmm_finish(const STile* tiles, const FrontInfo* fi, const i64* snptr,
                              const i64* perm, double* B, int ldb, int nrhs, double* work,
                              const i64* wptr, const int* gptr, const GEnt* gent,
                              const double* tmp);
__global__ void k_sgather_bwd(const int2* tiles, const FrontInfo* fi, const i64* snptr,
                              const int* relidx, const i64* perm, const double* B, int ldb,
                              int nrhs, double alpha, double* work, const i64* wptr);
__global__ void k_sscatter(const STile* tiles, const FrontInfo* fi, const i64* snptr,
                           const i64* perm, double* B, int ldb, int nrhs, double* work,
                           const i64* wptr, const double* tmp, const double* tmp2,
                           int copyback);

// utilities
__global__ void k_dot_partial(const int* cols, int ncols, const FrontInfo* fi,
                              const int* colfront, const double* x, const double* y,
                              double* partial);
__global__ void k_copy(double* dst, const double* src, i64 n);
__global__ void k_to_rowmajor(i64 n, int nrhs, const double* X, int ldx, double* Xt);
__global__ void k_spmv_residual_rm(i64 n, const i64* rowptr, const int* colind,
                                   const double* vals, const double* Xt, int nrhs,
                                   const double* B, int ldb, double* R, int ldr);
__global__ void k_spmv_residual(i64 n, const i64* rowptr, const int* colind, const double* vals,
                                const double* X, int ldx, const double* B, int ldb, double* R,
                                int ldr, int nrhs);
__global__ void k_logdiag(const int* cols, int ncols, const FrontInfo* fi, const int* colfront,
                          const double* x, double* partial);
__global__ void k_frontal_add(double* F, i64 N, const i64* r, const double* U,
                              i64 fr, i64 fc, i64 ur, i64 uc, double alpha);
__global__ void k_frontal_get(const double* F, i64 N, const i64* r, double* U,
                              i64 fr, i64 fc, i64 ur, i64 uc);
