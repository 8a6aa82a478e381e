// Export of a Cholesky factor to the lower-triangular L of P A P^T
// (symbolic.py:1052-1068): the stored (2,1) blocks B21 = L21 L11^{-1}
// (cholesky.c:71-72) are multiplied back by L11.  One workgroup per 64 x 64
// tile of a front's L21, K staged through LDS in chunks of 16; the product
// runs from the tile's first column on, since L11 is lower triangular.
#include "plan.h"
#include "../../include/chompack_amd.h"

#include <algorithm>

namespace {

constexpr int XT = 64;  // output tile
constexpr int XK = 16;  // K chunk

// tiles[i] = (front, r0, c0): rows [r0, r0 + 64) of B21, columns [c0, c0 + 64)
//   out[nn + r, c] = sum_{q >= c} B21[r, q] L11[q, c]
// out holds a copy of blkval on entry; only the L21 entries are written.
__global__ __launch_bounds__(256) void k_b21_l11(const int4* __restrict__ tiles,
                                                 const FrontInfo* __restrict__ fi,
                                                 const double* __restrict__ blkval,
                                                 double* __restrict__ out) {
    __shared__ double As[XK][XT + 1];  // B21 chunk, [q][row]
    __shared__ double Bs[XK][XT + 1];  // L11 chunk, [q][column]
    const int4 t = tiles[blockIdx.x];
    const FrontInfo f = fi[t.x];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int r0 = t.y, c0 = t.z;
    const double* blk = blkval + f.blkptr;
    const int tid = threadIdx.x, tr = tid & 15, tc = tid >> 4;  // 4 x 4 outputs per thread
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int k0 = c0; k0 < nn; k0 += XK) {
        for (int e = tid; e < XK * XT; e += 256) {
            const int ka = e / XT, i = e % XT;  // consecutive threads: consecutive rows
            const int row = r0 + i, qa = k0 + ka;
            As[ka][i] = (row < na && qa < nn) ? blk[nn + row + (i64)qa * nj] : 0.0;
            const int kb = e % XK, jc = e / XK;  // consecutive threads: consecutive q
            const int col = c0 + jc, qb = k0 + kb;
            Bs[kb][jc] = (col < nn && qb < nn && qb >= col) ? blk[qb + (i64)col * nj] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < XK; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = As[kk][tr + 16 * u];
                b[u] = Bs[kk][tc + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = r0 + tr + 16 * u, col = c0 + tc + 16 * v;
            if (row < na && col < nn) out[f.blkptr + nn + row + (i64)col * nj] = acc[u][v];
        }
}

}  // namespace

extern "C" {

int chp_factor_export(chp_plan* P, const double* lblkval, double* out, void* stream) {
    if (!P || !lblkval || !out) {
        chp_set_error("null argument");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    i64 len = 0;
    std::vector<int4> tiles;
    for (i64 k = 0; k < P->nsn; ++k) {
        const FrontInfo& f = P->fi[k];
        len = std::max(len, f.blkptr + (i64)f.nn * (f.nn + f.na));
        for (int r0 = 0; r0 < f.na; r0 += XT)
            for (int c0 = 0; c0 < f.nn; c0 += XT) tiles.push_back(make_int4((int)k, r0, c0, 0));
    }
    if (int rc = chp_copy(out, lblkval, len, stream)) return rc;
    if (tiles.empty()) return 0;
    int4* d_tiles = nullptr;
    HIP_CHECK(hipMalloc((void**)&d_tiles, tiles.size() * sizeof(int4)));
    int rc = 0;
    if (hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, s) !=
        hipSuccess)
        rc = -2;
    for (size_t b = 0; rc == 0 && b < tiles.size(); b += 65535u * 16) {
        const unsigned cnt = (unsigned)std::min<size_t>(tiles.size() - b, 65535u * 16);
        hipLaunchKernelGGL(k_b21_l11, dim3(cnt), dim3(256), 0, s, (const int4*)d_tiles + b,
                           (const FrontInfo*)P->d_fi, lblkval, out);
        if (hipGetLastError() != hipSuccess) rc = -2;
    }
    if (rc == 0 && hipStreamSynchronize(s) != hipSuccess) rc = -2;  // the tile list is freed
    (void)hipFree(d_tiles);
    if (rc) chp_set_error("chp_factor_export: HIP error");
    return rc;
}

}  // extern "C"
