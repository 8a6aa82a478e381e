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

// Failure status word: min over failing fronts of (k << 32) | info.
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
// buffer 0 = blkval, 1 = update arena, 2 = scratch, 3 = diagonal-block inverses.
//   opA: 0 A(i,k) = A[i + k*lda]; 1 A(i,k) = A[k + i*lda];
//        2 symmetric, lower stored: A(i,k) = A[max + min*lda].
//   opB: 0 B(k,j) = B[k + j*ldb]; 1 B(k,j) = B[j + k*ldb].
//   tri: 0 full; 1 only C(i,j) with i >= j written; 2 as 1, and C(i,j) = 0
//        for i < j (tiles above the diagonal only write zeros);
//   triA: 0 none; 1 A is lower triangular (A(i,k) = 0 for k > i);
//         2 A is upper triangular in op form (A(i,k) = 0 for k < i);
//   triB (stored in `pad`): 1 B is lower triangular (B(k,j) = 0 for k < j).
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
