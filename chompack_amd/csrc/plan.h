// Device plan shared by the scheduler sources (plan.hip: cholesky, llt,
// projected inverse, solve; plan_dense.hip: completion and the Hessian maps).
#pragma once
#include "common.h"
#include "kernels.h"

#include <cstring>
#include <map>
#include <vector>

namespace chpi {

enum Kid {
    K_CHOL_SMALL, K_EXTADD, K_POTRF_DIAG, K_PANEL_FACTOR, K_GEMM, K_EXTRACT,
    K_PROJINV_SMALL, K_TRTRI_DIAG, K_LLT_SMALL, K_SOLVE_FWD, K_SOLVE_BWD,
    K_COPY_TOP, K_SPLITK_REDUCE, K_DIAG, K_MATOP, K_UPDATE_FACTOR,
    K_SGATHER_F, K_SGATHER_B, K_SSCATTER, K_TRMM_FWD, K_TRMM_BWD, K_TRMM_FIN, K_COMP_CHAIN,
    K_PROJINV_CHAIN, K_CHOL_DAG, K_FILL_INT, K_GEMM_SMALL, K_CHOL_CHAIN, K_LLT_CHAIN,
    K_UF_SETUP, K_UF_STEP, K_GEMM_NARROW,
    K_FORK, K_JOIN
};
const int NKIDS = 32;  // kernel kinds (K_FORK / K_JOIN are stream markers)

struct Launch {
    Kid kid;
    int grid;
    int block;
    size_t lds;
    i64 list;            // byte offset of the work list in the program's list buffer
    int iarg;
    double flops = 0.0;  // algorithmic flops of a GEMM launch
    int bsel = 0;        // buffer (BufTab index) used as blkval by front kernels
    int side = 0;        // 0: the sweep's stream; 1, 2: side stream (after a K_FORK of it)
    int fisel = 0;       // front table of the front kernels: 0 the compacted update
                         // offsets (d_fi), 1 one slot per front (d_fi_full)
    int gy = 1;          // extend-add / extract: instances along grid y, their
    i64 bstride = 0;     // blkval and update arenas bstride / ustride doubles apart
    i64 ustride = 0;
};

struct ListBuf {
    std::vector<char> bytes;
    template <class T>
    i64 add(const std::vector<T>& v) {
        i64 off = (i64)((bytes.size() + 255) / 256 * 256);
        bytes.resize(off + v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

// Work lists + GEMM tasks of a family of sweeps, with their device copies and
// split-K partial bookkeeping.
struct Program {
    ListBuf lists;
    std::vector<GemmTask> tasks;
    char* d_lists = nullptr;
    GemmTask* d_tasks = nullptr;
    size_t up_lists = 0, up_tasks = 0;  // sizes of the device copies
    i64 splitp_cur = 0, splitp_len = 0;
    std::vector<GemmTile> split_out;    // output tiles of split tasks, current launch
    i64 split_doubles() const { return splitp_len + 1; }
    // GEMMs built while split_side is set run on side stream 2: their split-K
    // partials take a region of their own, placed after the main stream's by
    // place_side_splits() once the program is complete
    bool split_side = false;
    i64 side_cur = 0, side_len = 0;
    std::vector<int> side_tasks;
    void place_side_splits() {
        for (int id : side_tasks) tasks[id].offP += splitp_len;
        splitp_len += side_len;
        side_tasks.clear();
        side_len = 0;
    }
};

// Solve schedule for one right-hand-side count (GEMM shapes and work offsets
// depend on it).
struct SolvePlan {
    int nrhs = 0;
    Program pg;
    std::vector<Launch> inv, fwd, bwd;  // X_k = L11^{-1}; the two sweeps
    std::vector<Launch> mfwd, mbwd;     // products with L ('N') and L^T ('T')
    bool mbuilt = false;
    i64 t2base = 0;                     // offset of T2 = B21^T x_A in d_tmp
    std::vector<CheckTile> chk;         // pieces of the L11 check
    CheckTile* d_chk = nullptr;
    int* d_changed = nullptr;           // set by k_l11_check when L11 differs from its copy
    double* d_l11 = nullptr;            // copy of the L11 the inverses were formed from
    bool inv_valid = false;             // d_inv holds the inverses of d_l11
    double* d_splitp = nullptr;
    // inverses of the large diagonal blocks, their workspace, solved blocks
    double* d_inv = nullptr;
    double* d_w = nullptr;
    double* d_tmp = nullptr;
    i64 inv_len = 0, w_len = 0, tmp_len = 0;
    // graphs of repeated solves, keyed by every kernel argument that varies
    struct Key {
        const void* blk = nullptr;
        const void* B = nullptr;
        int ldb = 0, trans = 0;
        double alpha = 0.0;
        bool operator==(const Key& o) const {
            return blk == o.blk && B == o.B && ldb == o.ldb && trans == o.trans && alpha == o.alpha;
        }
    };
    struct Graph { Key key; hipGraphExec_t exec = nullptr; };
    std::vector<Graph> graphs;  // most recent last
    Key last;                   // an eager call; the next identical call is captured
};

}  // namespace chpi

struct chp_plan {
    int device = 0;
    i64 n = 0, nsn = 0;
    int nlevels = 0;
    // tunable parameters (chp_plan_create_params): largest clique order of the
    // LDS-resident small-front kernels, the LDS a small-front workgroup may
    // use, the outer panel width of large fronts, the split-K cap
    int small_max = 64;
    size_t lds_cap = 64 * 1024;
    int nb2 = NB2;
    int splitk_max = 16;
    // device: compute units, LDS per compute unit
    int ncu = 256;
    size_t lds_per_cu = 64 * 1024;
    std::vector<FrontInfo> fi;
    // the same fronts with one update slot per front (updptr = prefix sums of
    // na^2): the layout of the Hessian scaling arena d_supd, which keeps every
    // slot of a sweep for the second scaling of hessian(adj=None)
    std::vector<FrontInfo> fi_full;
    FrontInfo* d_fi_full = nullptr;
    i64 upd_full_len = 0;
    std::vector<std::vector<int>> bylevel;
    std::vector<int> h_relidx, h_chidx;
    std::vector<i64> h_dinvptr, h_xptr, h_tptr, h_wunit, h_lltptr;
    i64* d_lltptr = nullptr;
    // device arrays
    FrontInfo* d_fi = nullptr;
    int* d_relidx = nullptr;
    int* d_chidx = nullptr;
    i64* d_perm = nullptr;
    i64* d_snptr = nullptr;
    i64* d_dinvptr = nullptr;
    i64* d_xptr = nullptr;
    i64* d_tptr = nullptr;
    i64* d_wunit = nullptr;
    int* d_gptr = nullptr;   // forward-solve gather map (GEnt)
    GEnt* d_gent = nullptr;
    int* d_colfront = nullptr;
    int* d_colidx = nullptr;
    double* d_upd = nullptr;
    double* d_dinv = nullptr;
    double* d_scratch = nullptr;
    double* d_partial = nullptr;
    double* d_work = nullptr;
    double* d_splitp = nullptr;
    i64 work_len = 0;
    unsigned long long* d_status = nullptr;
    i64 upd_len = 0, dinv_len = 0, scratch_len = 0, solve_len = 0;
    // chol: large fronts by the tile-DAG launch (FP64); chol_lp: the same
    // sweep with panel launches and GEMM launches for the Schur updates,
    // used by the low-precision factorizations
    std::vector<chpi::Launch> chol, chol_lp, projinv, llt;
    // tile counters and queue heads of the DAG launches (zeroed at the start
    // of every cholesky sweep); polls a dependency wait may take
    int* d_dagcnt = nullptr;
    i64 dagcnt_len = 0;
    int dag_wait_limit = 1 << 18;
    bool chol_dag = false;  // large fronts of the FP64 cholesky by the tile-DAG launch
    i64 dag_tasks = 0;           // tasks of all DAG launches of the sweep
    i64* d_dagtrace = nullptr;   // per-task trace (chp_dag_trace), when enabled
    bool dag_bad = false;  // a DAG queue order failed its host-side check
    bool building_lowprec = false;  // chol_lp under construction (push_gemm)
    chpi::Program main, dense;
    chpi::Program* pg = &main;  // program the builders append to
    int n_large_chol = 0, n_large_projinv = 0;
    i64 n_gemm_tiles = 0;
    // completion / Hessian sweeps, built on first use.  Index of scale:
    // 4 factored + 2 inv + adj; of y2k / m2t: inv.
    struct Seq { bool built = false; std::vector<chpi::Launch> seq; };
    struct LevelSeq {
        bool built = false;
        // per level, top-down: shared = the update matrices (or their factors),
        // kept in their own arena d_supd; shared_t = inverses into scratch
        std::vector<std::vector<chpi::Launch>> shared, shared_t, dir;
    };
    // the Hessian maps work on a chunk of nb directions at once; their
    // sequences are built per chunk size
    Seq comp[2];
    Seq chol_adj;  // reverse-mode sweep of the Cholesky factorization
    std::map<int, Seq> y2k[2], m2t[2];
    std::map<int, LevelSeq> scale[8];
    i64 blk_len = 0;            // blkptr[nsn]
    double* d_ustack = nullptr; // a chunk of directions, blk stride apart
    double* d_uupd = nullptr;   // their update arenas, upd stride apart
    i64 ustack_cap = 0, uupd_cap = 0;
    std::vector<chpi::SolvePlan*> solves;  // cached per nrhs
    double* d_dscratch = nullptr;
    double* d_ddinv = nullptr;
    double* d_dsplitp = nullptr;
    double* d_supd = nullptr;  // update arena of the Hessian's scaling sweeps (upd_len)
    i64 dscratch_len = 0, ddinv_len = 0;                       // required
    i64 dscratch_cap = 0, ddinv_cap = 0, dsplitp_cap = 0;      // allocated
    // HIP graphs of the sweeps, replayed on a plan-owned stream that is
    // joined to the caller's stream with events; cached per blkval pointer
    bool use_graphs = true;
    hipStream_t gstream = nullptr;
    hipEvent_t gev_in = nullptr, gev_out = nullptr;
    // side streams for independent work (K_FORK / K_JOIN with iarg = 1 or 2):
    // 1 carries a level's small fronts, 2 the stored-convention conversion and
    // the L11 inversions
    static const int NSIDE = 3;
    hipStream_t sstream[NSIDE] = {nullptr, nullptr, nullptr};
    hipEvent_t sev_fork[NSIDE] = {nullptr, nullptr, nullptr};
    hipEvent_t sev_join[NSIDE] = {nullptr, nullptr, nullptr};
    // A sweep's captured graph, and the choice between replaying it and
    // launching eagerly on the side streams (whichever measured faster: the
    // graph saves host launch work, eager launches give every side stream
    // its own hardware queue).  calls 1-2 run eagerly, 3-4 through the graph;
    // calls 2 and 4 are timed with events; from call 5 on `mode` is fixed.
    struct Cached {
        const void* key = nullptr;
        hipGraphExec_t exec = nullptr;
        int calls = 0, mode = 0;  // mode: 0 undecided, 1 eager, 2 graph
        float t_eager = 0.f, t_graph = 0.f;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int pending = 0;          // mode of a timed call whose end event is unread
    };
    Cached g_chol, g_projinv, g_llt, g_chol32, g_cholbf;
    // arithmetic of the Schur updates of the cholesky sweep in progress:
    // 0 FP64, 1 FP32, 2 BF16x3 split precision
    int schur_mode = 0;
    // profiling: per-kernel-kind time from HIP events around every launch
    bool prof = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<double> prof_rows;  // per launch: kind, grid, flops, ms
    double prof_ms[chpi::NKIDS] = {0};
    double prof_flops[chpi::NKIDS] = {0};
    i64 prof_cnt[chpi::NKIDS] = {0};
};

namespace chpi {

GemmTask mk(int bufA, i64 offA, int lda, int opA, int bufB, i64 offB, int ldb, int opB,
            int bufC, i64 offC, int ldc, int M, int N, int K, double alpha, double beta,
            int tri, int triA = 0, int triB = 0);
double task_flops(const GemmTask& t);
void add_gemm_tiles(chp_plan* P, std::vector<GemmTile>& tiles, GemmTask t, bool skip_upper);
// schur = 1 marks a Schur-complement update (FP32 arithmetic in the mixed sweep)
void push_gemm(chp_plan* P, std::vector<Launch>& seq, const std::vector<GemmTile>& tiles,
               int schur = 0);
// Device copy of a host vector (at least one element allocated).
template <class T>
int upload(T** dptr, const std::vector<T>& v) {
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    HIP_CHECK(hipMalloc((void**)dptr, bytes));
    if (!v.empty()) HIP_CHECK(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// Uploads a program's lists and tasks if they grew since the last upload.
int sync_program(Program& pg);
// Replays `seq` of program `pg` on stream s with buffer table bt.
int run_seq(chp_plan* P, const Program& pg, const std::vector<Launch>& seq, const BufTab& bt,
            hipStream_t s);
int read_status(chp_plan* P, hipStream_t s, i64* info);
// Creates the side streams and their fork / join events (once per plan).
int ensure_side_stream(chp_plan* P);
// Extend-add work list of the listed fronts: (front, c0, r0) for the target
// tiles of 16 columns x 256 rows that meet the lower triangle.
// Extend-add tiles of the fronts (x = front, y = c0, z = r0, w = index in the
// returned list of the first child's bounds), followed by the bounds: per
// tile and child (reverse child order) three int4, the first child index
// with relative index >= c0, c1, r0, r1, then (na, 0, relptr lo, relptr hi)
// and (updptr lo, updptr hi, 0, 0) of the child.  *ntiles receives the
// number of tiles.
std::vector<int4> extadd_tiles(const chp_plan* P, const std::vector<int>& fronts, int* ntiles);
// Extract work list of the listed fronts: (front, j0, i0) for the 16-column x
// 256-row blocks of V_k that meet the lower triangle.
std::vector<int4> extract_tiles(const chp_plan* P, const std::vector<int>& fronts);
// Children of the listed fronts (each front's children in chidx order).
std::vector<int> children_of(const chp_plan* P, const std::vector<int>& fr);
BufTab main_tab(chp_plan* P, double* blkval);
int alloc_doubles(double** d, i64 len);

// Dense operations on descriptors (plan_dense.hip): element (i, j) of a
// matrix at BufTab p[buf][off + i + j*ld], order n.
struct Mat {
    int buf;
    i64 off;
    int ld, n;
};
struct Bump {
    i64 cur = 0;
    i64 take(i64 len) {
        i64 o = cur;
        const i64 l = len < 1 ? 1 : len;
        cur += (l + 31) / 32 * 32;
        return o;
    }
};
template <class T>
void push_tasks(chp_plan* P, std::vector<Launch>& seq, Kid kid, const std::vector<T>& v,
                int bsel = 0);
void add_matop(std::vector<MatOp>& ops, MatOpKind kind, int sbuf, i64 src, int lds, int dbuf,
               i64 dst, int ldd, int m, int n, double alpha = 1.0);
// In-place lower Cholesky of each matrix (block inverses to buffer 3 via dv).
void emit_potrf(chp_plan* P, std::vector<Launch>& seq, const std::vector<Mat>& ms,
                const std::vector<int>& keys, Bump& dv);
// X = L^{-1}; W[q]: n x n workspace (ld n) in buffer 2.
void emit_trtri(chp_plan* P, std::vector<Launch>& seq, const std::vector<Mat>& Ls,
                const std::vector<Mat>& Xs, const std::vector<i64>& Ws,
                const std::vector<int>& keys);

}  // namespace chpi
