// Shared definitions for the chompack_amd HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

typedef int64_t i64;

// Per-supernode descriptor, resident in HBM (one entry per supernode k).
//   nn = |N_k| (supernode columns), na = |A_k| (separator rows), nj = nn + na.
//   blkptr: offset of k's column-major nj x nn block in blkval (ld = nj).
//   updptr: offset of k's na x na update / V_k matrix in the update arena
//           (column-major, ld = na, lower triangle meaningful).
//   relptr: offset of k's relative indices (int32) in relidx.
//   chbeg/chend: range of k's children in chidx (postorder).
//   parent: supernodal parent (== k for roots).
struct FrontInfo {
    int nn, na;
    int chbeg, chend;
    int parent, level;
    i64 blkptr, updptr, relptr;
};

// Failure status word: min over failing fronts of (k << 32) | info.  The
// word after it is an error word: nonzero when a dependency wait of the
// tile-DAG factorization ran out of its bound (an internal error).
#define CHP_STATUS_OK 0xffffffffffffffffULL

__device__ inline void chp_record_failure(unsigned long long* status, int k, int info) {
    unsigned long long v = ((unsigned long long)(unsigned)k << 32) | (unsigned)info;
    atomicMin(status, v);
}

// Error reporting for the C-ABI.
void chp_set_error(const std::string& msg);

#define HIP_CHECK(expr)                                                        \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            chp_set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return -2;                                                         \
        }                                                                      \
    } while (0)

// A GEMM-like task: C[i,j] = alpha * sum_k opA(i,k) opB(k,j) + beta * C[i,j]
// over an M x N block, used for the large-front Schur updates and the
// projected-inverse products.  Operands are given as (buffer, offset):
// buffer 0 = blkval (or the direction / matrix the sweep works on),
// 1 = update arena, 2 = scratch, 3 = diagonal-block inverses, 4 = split-K
// partials, 5 = factor L, 6 = projected inverse Y (see BufTab).
//   opA: 0 A(i,k) = A[i + k*lda]; 1 A(i,k) = A[k + i*lda];
//        2 symmetric, lower stored: A(i,k) = A[max + min*lda].
//   opB: 0 B(k,j) = B[k + j*ldb]; 1 B(k,j) = B[j + k*ldb];
//        2 symmetric, lower stored: B(k,j) = B[max + min*ldb].
//   tri: 0 full; 1 only C(i,j) with i >= j written; 2 as 1, and C(i,j) = 0
//        for i < j (tiles above the diagonal only write zeros);
//   triA: 0 none; 1 A is lower triangular (A(i,k) = 0 for k > i);
//         2 A is upper triangular in op form (A(i,k) = 0 for k < i);
//   triB (stored in `pad`): 1 B is lower triangular (B(k,j) = 0 for k < j);
//         2 B is upper triangular (B(k,j) = 0 for k > j).
//   ksplit > 1: deterministic split-K.  Tile slice ks covers
//        k in [ks*kchunk, min(K, (ks+1)*kchunk)) and writes alpha * partial to
//        buffer 4 at offP + ks*M*N (column-major, ld M); k_splitk_reduce then
//        forms C = beta*C + sum_ks partial in slice order.
struct GemmTask {
    i64 offA, offB, offC, offP;
    int lda, ldb, ldc;
    int M, N, K;
    int ksplit, kchunk;
    unsigned char bufA, bufB, bufC, opA, opB, tri, triA, pad;
    double alpha, beta;
};

// One 64 x 64 output tile (k-slice ks) of a GemmTask.
struct GemmTile {
    int task, tm, tn, ks;
};

// Base pointers of the buffers a launch addresses (kernel argument by value).
// gate: when not null, the dense kernels (k_diag_tasks, k_matop and the GEMM
// kernels) return at once unless *gate != 0, so that a launch sequence can be
// made conditional on a flag the device sets without a host round trip.
struct BufTab {
    double* p[8];
    const int* gate;
};

__device__ inline bool gated_off(const BufTab& bt) { return bt.gate != nullptr && *bt.gate == 0; }

// bt.p[i] for an index known only at run time, without the copy of the
// table into scratch memory that indexing the kernel argument would take.
__device__ inline double* buf_sel(const BufTab& bt, int i) {
    double* r = bt.p[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) r = (i == k) ? bt.p[k] : r;
    return r;
}

// One diagonal block (order nb <= 64) of a dense lower-triangular operation,
// addressed by descriptor: element (i, j) at p[buf][off + i + j*ld].
//   mode 0: Cholesky in place; the inverse of the factor goes row-major to
//           p[obuf][outoff + j + i*64] (64 x 64, zero padded).
//   mode 1: inverse of the lower-triangular block, written column-major to
//           p[obuf][outoff + i + j*ldo] (full nb x nb, zero above the diagonal).
//   mode 2: as mode 1, and the block's lower triangle is also copied to
//           p[buf2][off2 + i + j*ld2] (the output may overwrite the input).
// A failure records (key, col + LAPACK-style column) in the status word.
struct DiagTask {
    i64 off, outoff, off2;
    int ld, ldo, nb, key, col, ld2;
    unsigned char buf, obuf, mode, buf2;
    int pad;
};

// Element-wise operation on one 64 x 64 tile (origin i0, j0) of an m x n
// matrix: source element (i, j) at p[sbuf][src + i + j*lds], destination at
// p[dbuf][dst + i + j*ldd].  Kinds: MO_COPY dst = alpha src; MO_COPY_LOWER
// the same for i >= j only; MO_COPY_LOWER_Z also writes 0 for i < j;
// MO_RT dst(i, j) = src(n-1-j, n-1-i) for i >= j and 0 above (reversed
// transpose of a lower triangle, square n); MO_SYMM dst(i, j) = dst(j, i)
// for i < j; MO_ZERO_UPPER dst(i, j) = 0 for i < j; MO_ZERO dst = 0;
// MO_PHI dst(i, i) *= 1/2 and dst(i, j) = 0 for i < j (in place: the lower
// triangle with the diagonal halved); MO_SYMAVG dst(i, j) = alpha (src(i, j)
// + src(j, i)) for i >= j and 0 for i < j (square).
enum MatOpKind { MO_COPY, MO_COPY_LOWER, MO_COPY_LOWER_Z, MO_RT, MO_SYMM, MO_ZERO_UPPER, MO_ZERO,
                 MO_PHI, MO_SYMAVG };
struct MatOp {
    i64 src, dst;
    int lds, ldd, m, n, i0, j0;
    unsigned char sbuf, dbuf, kind, pad;
    int pad2;
    double alpha;
};

// update_factor for one child c of front k: F11/F21 = the nj x nn block of k
// at p[fbuf][foff] (ld nj), F22 = k's factor Lk (na x na, lower, ld na) at
// p[1][lkoff]; writes Li (N x N, lower, zero above) to p[1][outoff] and uses
// p[2][woff .. + na*N2] as workspace.  r = relidx + relpos (N entries).
struct UFTask {
    i64 foff, lkoff, outoff, woff, relpos;
    int nn, na, N, key;
    // defer_solve: stop once Li22 is in place (Li21 and the (1,1) block then
    // come from multi-workgroup triangular inverse and GEMM launches);
    // defer_rchol: leave A - Li21^T Li21 in Li11 (its reversed Cholesky runs
    // afterwards as a blocked multi-workgroup potrf)
    // ext_reduce: the copies and the reflections ran as launches of their own
    // (k_uf_setup, k_uf_step); the kernel starts at Li22
    unsigned char fbuf, defer_solve, defer_rchol, ext_reduce, pad[4];
};

// One item of a k_uf_step launch for a large update-factor task.  A block is
// a run of consecutive reflections js0 .. js0 + nbk - 1 whose bands cover
// rows lo .. lo + len - 1 of C (na x N2, ld na, the task's workspace in
// buffer 2); its T (UF_BMAX x UF_BMAX), tau and band starts go to buffer 2
// at toff.  role 0: the block's compact WY form applied to those of the 16
// columns c0 .. c0 + 15 of C that lie below the bound colmin and outside
// xlo .. xhi - 1; role 1: the block's reflections on its columns colmin ..
// colmin + nbk - 1, after the previous block (pjs0, pnbk, plo, plen, ptoff;
// pnbk = 0 if none) has been applied to them.
#define UF_BMAX 16
struct UFBlock {
    i64 woff, relpos, toff, ptoff;
    int nn, na, N, N2;
    int js0, nbk, lo, len;
    int colmin, c0, key, role;
    int pjs0, pnbk, plo, plen;
    int xlo, xhi;
};

// Tile-DAG factorization of a level's large fronts.  A front is cut into
// 64 x 64 tiles: row (and column) tiles 0 .. P-1 cover its nn supernode
// columns from 0, tiles P .. R-1 its na separator rows from nn.  Tile (i, j),
// i >= j, lives in blkval (ld nj) when j < P, otherwise in the front's
// update slot (ld na).  cbase: offset of the front's tile counters
// (R (R + 1) / 2, row-major lower triangle); dinvoff: its diagonal-block
// inverses (64 x 64 column-major per block); key: the supernode.
struct DagFront {
    i64 blkptr, updptr, dinvoff;
    int nn, na, P, R, cbase, key;
};
// One task: tile (i, j) and step k; kt = k | type << 16 with type 0 the
// factorization of diagonal tile (k, k) and the inverse of its factor, 1
// the solve of tile (i, k) with that factor, 2 the update of tile (i, j)
// with tiles (i, k) and (j, k).
struct DagTask {
    int f, i, j, kt;
};
enum { DAG_POTRF = 0, DAG_TRSM = 1, DAG_UPDATE = 2 };

// One item of a chain of small fronts: the completion step of `front` (kind
// 1), followed by the nch items of its children's update factors (kind 0;
// kind 2 for the child that is the chain's next front).
struct ChainItem {
    UFTask uf;
    int front, kind, nch, pad;
};

// One front of a chain launch of cholesky, llt or the projected inverse
// (k_chol_chain, k_llt_chain, k_projinv_chain), in the order the launch
// walks them.  `nbr` is the chain neighbour whose matrix the front takes
// from LDS (cholesky and llt: the previous front, a child of this one, with
// its update matrix; projected inverse: the previous front, this one's
// parent, having extracted this front's V), or -1; nbr_na and nbr_relptr
// describe the child's relative indices (cholesky and llt: of that child;
// projected inverse: of the next front, which is then kept in LDS).  `keep`
// says whether the matrix this front hands on stays in LDS for the next
// front (otherwise it goes to the update arena).
struct ChainFront {
    i64 blkptr, updptr, nbr_relptr;
    int front, nn, na, chbeg, chend;
    int nbr, nbr_na, keep;
};

// One piece of the solve's L11 comparison: columns c0 .. c0 + 15 and rows
// r0 .. r1 - 1 of front k's L11, whose copy starts at tofs (ld nn).
struct CheckTile {
    i64 tofs;
    int k, c0, r0, r1;
};

// Row tile of a front in the solve's scatter: rows r0 .. r0 + 63 of front k,
// whose solved block sits at offset tofs of the solve's block buffer.
struct STile {
    int k, r0;
    i64 tofs;
};

// Forward-solve gather map.  Slab row s = wunit[k] + i (row i of front k's
// slab) receives the update rows gent[gptr[s] .. gptr[s + 1]) of k's children,
// listed in reverse child order (the order in which the reference adds them,
// trsm.c:46-58).  An entry addresses row `row` of a child slab that starts at
// wunit `wu` and has leading dimension `ld`: element (row, j) of it is
// work[wu * nrhs + row + j * ld].
struct GEnt {
    i64 wu;
    int ld, row;
};
