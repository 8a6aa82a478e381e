// Product-form Cholesky factors of I + V diag(a) V^T (pfcholesky.c):
//   I + V diag(a) V^T = L_1 L_2 ... L_m L_m^T ... L_1^T,
//   L_i = diag(l_i) + strictly-lower(v_i b_i^T).
// The reference evaluates every factor and every solve or product with one
// as a sequential loop over the n rows carrying a scalar s.  Each such loop
// is the recurrence s_{k+1} = a_k s_k + c_k (a_k = 1 for the products), so
// one workgroup per column computes it as a scan: each thread composes the
// affine maps of a contiguous chunk of rows, the workgroup scans the chunk
// maps in LDS, and each thread then runs the reference's loop over its chunk
// from its incoming s.  Inside a chunk the arithmetic is the reference's;
// only the chunk-boundary values of s are summed in another order.
#include <hip/hip_runtime.h>

#include "common.h"
#include "../../include/chompack_amd.h"

#include <string>

namespace {

constexpr int PF_TPB = 1024;

enum PfMode { RSV_N = 0, RSV_T = 1, RMV_N = 2, RMV_T = 3 };

// Exclusive scan of the threads' maps (A, C) in scan position order q; returns
// the incoming s of position q (s = 0 before the first row).
__device__ double scan_maps(double A, double C, int q, double* sA, double* sC) {
    sA[q] = A;
    sC[q] = C;
    __syncthreads();
    for (int d = 1; d < PF_TPB; d <<= 1) {
        double a = 1.0, c = 0.0;
        const bool on = q >= d;
        if (on) { a = sA[q - d]; c = sC[q - d]; }
        __syncthreads();
        if (on) {  // map q after map q - d
            const double aq = sA[q], cq = sC[q];
            sA[q] = aq * a;
            sC[q] = aq * c + cq;
        }
        __syncthreads();
    }
    const double s = q > 0 ? sC[q - 1] : 0.0;  // inclusive up to q - 1, applied to 0
    __syncthreads();
    return s;
}

// One sweep with factor (v, l, b) over column blockIdx.x of X (ld ldx):
//   RSV_N: x := L^{-1} x   (ddrsv 'N', pfcholesky.c:12-17)
//   RSV_T: x := L^{-T} x   (ddrsv 'T', pfcholesky.c:18-23)
//   RMV_N: x := L x        (ddrmv 'N', pfcholesky.c:34-40)
//   RMV_T: x := L^T x      (ddrmv 'T', pfcholesky.c:41-47)
template <int MODE>
__global__ __launch_bounds__(PF_TPB) void k_pf_sweep(i64 n, const double* __restrict__ v,
                                                     const double* __restrict__ l,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ X, int ldx) {
    __shared__ double sA[PF_TPB], sC[PF_TPB];
    double* x = X + (i64)blockIdx.x * ldx;
    const bool rev = MODE == RSV_T || MODE == RMV_T;
    const int t = threadIdx.x;
    const i64 ch = (n + PF_TPB - 1) / PF_TPB;
    const i64 lo = min(n, (i64)t * ch), hi = min(n, lo + ch);
    const int q = rev ? PF_TPB - 1 - t : t;  // scan position of this chunk
    // the chunk's map from the old x
    double A = 1.0, C = 0.0;
    for (i64 e = 0; e < hi - lo; ++e) {
        const i64 k = rev ? hi - 1 - e : lo + e;
        const double xk = x[k];
        double a, c;
        if (MODE == RSV_N) { a = 1.0 - b[k] * v[k] / l[k]; c = b[k] * xk / l[k]; }
        else if (MODE == RSV_T) { a = 1.0 - v[k] * b[k] / l[k]; c = v[k] * xk / l[k]; }
        else if (MODE == RMV_N) { a = 1.0; c = b[k] * xk; }
        else { a = 1.0; c = v[k] * xk; }
        A = a * A;
        C = a * C + c;
    }
    double s = scan_maps(A, C, q, sA, sC);
    // the reference's loop over the chunk
    for (i64 e = 0; e < hi - lo; ++e) {
        const i64 k = rev ? hi - 1 - e : lo + e;
        const double xk = x[k];
        if (MODE == RSV_N) {
            const double y = (xk - v[k] * s) / l[k];
            x[k] = y;
            s += y * b[k];
        } else if (MODE == RSV_T) {
            const double y = (xk - b[k] * s) / l[k];
            x[k] = y;
            s += y * v[k];
        } else if (MODE == RMV_N) {
            const double tmp = xk * b[k];
            x[k] = l[k] * xk + v[k] * s;
            s += tmp;
        } else {
            const double tmp = xk * v[k];
            x[k] = l[k] * xk + b[k] * s;
            s += tmp;
        }
    }
}

// Factor i of dpftrf (pfcholesky.c:67-79): with alpha_0 = -a_i,
//   l_j = sqrt(1 - alpha_j v_j^2), b_j = -alpha_j v_j / l_j,
//   alpha_{j+1} = alpha_j + b_j^2,
// which is 1 / alpha_{j+1} = 1 / alpha_j - v_j^2: one prefix sum of v^2.
// *fail = 1 where 1 - alpha_j v_j^2 <= 0 (the reference's failure test).
__global__ __launch_bounds__(PF_TPB) void k_pf_factor(i64 n, const double* __restrict__ v,
                                                      double* __restrict__ l,
                                                      double* __restrict__ b, double alpha0,
                                                      int* __restrict__ fail) {
    __shared__ double sA[PF_TPB], sC[PF_TPB];
    const int t = threadIdx.x;
    const i64 ch = (n + PF_TPB - 1) / PF_TPB;
    const i64 lo = min(n, (i64)t * ch), hi = min(n, lo + ch);
    double C = 0.0;
    for (i64 k = lo; k < hi; ++k) C += v[k] * v[k];
    double S = scan_maps(1.0, C, t, sA, sC);  // sum of v^2 before the chunk
    const double r0 = 1.0 / alpha0;          // +-inf for alpha_0 = 0: alpha stays 0
    bool bad = false;
    for (i64 k = lo; k < hi; ++k) {
        const double alpha = 1.0 / (r0 - S);
        const double vk = v[k];
        const double tmp = 1.0 - alpha * vk * vk;
        if (tmp <= 0.0) bad = true;
        const double lk = sqrt(tmp);
        l[k] = lk;
        b[k] = -alpha * vk / lk;
        S += vk * vk;
    }
    if (bad) *fail = 1;
}

int pf_check(i64 n, int m, const void* V, const void* L, const void* B, int ldv, int ldl,
             int ldb) {
    if (n <= 0 || m < 0) { chp_set_error("n must be positive and m non-negative"); return -1; }
    if (!V || !L || !B) { chp_set_error("null argument"); return -1; }
    if (ldv < n || ldl < n || ldb < n) { chp_set_error("leading dimension smaller than n"); return -1; }
    return 0;
}

}  // namespace

extern "C" {

int chp_pf_factor(int64_t n, int m, const double* a, double* V, int ldv, double* L, int ldl,
                  double* B, int ldb, void* stream) {
    if (int rc = pf_check(n, m, V, L, B, ldv, ldl, ldb)) return rc;
    if (m > 0 && !a) { chp_set_error("null argument"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    int* dfail = nullptr;
    HIP_CHECK(hipMalloc((void**)&dfail, sizeof(int)));
    int rc = 0;
    if (hipMemsetAsync(dfail, 0, sizeof(int), s) != hipSuccess) rc = -2;
    for (int i = 0; i < m && rc == 0; ++i) {
        const double* v = V + (i64)i * ldv;
        double* l = L + (i64)i * ldl;
        double* b = B + (i64)i * ldb;
        hipLaunchKernelGGL(k_pf_factor, dim3(1), dim3(PF_TPB), 0, s, (i64)n, v, l, b, -a[i], dfail);
        if (i + 1 < m)
            hipLaunchKernelGGL(k_pf_sweep<RSV_N>, dim3(m - i - 1), dim3(PF_TPB), 0, s, (i64)n, v, l,
                               b, V + (i64)(i + 1) * ldv, ldv);
        if (hipGetLastError() != hipSuccess) rc = -2;
    }
    int fail = 0;
    if (rc == 0 && hipMemcpyAsync(&fail, dfail, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) rc = -2;
    if (rc == 0 && hipStreamSynchronize(s) != hipSuccess) rc = -2;
    (void)hipFree(dfail);
    if (rc) { chp_set_error("pf_factor: HIP error"); return rc; }
    return fail ? 1 : 0;
}

int chp_pf_solve(int64_t n, int m, const double* V, int ldv, const double* L, int ldl,
                 const double* B, int ldb, double* U, int ldu, int nu, int trans, void* stream) {
    if (int rc = pf_check(n, m, V, L, B, ldv, ldl, ldb)) return rc;
    if (trans != 'N' && trans != 'T') { chp_set_error("trans must be 'N' or 'T'"); return -1; }
    if (nu <= 0 || m == 0) return 0;
    if (!U || ldu < n) { chp_set_error("U is null or ldu is smaller than n"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    for (int q = 0; q < m; ++q) {
        const int i = trans == 'N' ? q : m - 1 - q;  // dpfsv order, pfcholesky.c:105-110
        const double* v = V + (i64)i * ldv;
        const double* l = L + (i64)i * ldl;
        const double* b = B + (i64)i * ldb;
        if (trans == 'N')
            hipLaunchKernelGGL(k_pf_sweep<RSV_N>, dim3(nu), dim3(PF_TPB), 0, s, (i64)n, v, l, b, U, ldu);
        else
            hipLaunchKernelGGL(k_pf_sweep<RSV_T>, dim3(nu), dim3(PF_TPB), 0, s, (i64)n, v, l, b, U, ldu);
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}

int chp_pf_product(int64_t n, int m, const double* V, int ldv, const double* L, int ldl,
                   const double* B, int ldb, double* U, int ldu, int nu, int trans, void* stream) {
    if (int rc = pf_check(n, m, V, L, B, ldv, ldl, ldb)) return rc;
    if (trans != 'N' && trans != 'T') { chp_set_error("trans must be 'N' or 'T'"); return -1; }
    if (nu <= 0 || m == 0) return 0;
    if (!U || ldu < n) { chp_set_error("U is null or ldu is smaller than n"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    for (int q = 0; q < m; ++q) {
        const int i = trans == 'N' ? m - 1 - q : q;  // dpfmv order, pfcholesky.c:129-134
        const double* v = V + (i64)i * ldv;
        const double* l = L + (i64)i * ldl;
        const double* b = B + (i64)i * ldb;
        if (trans == 'N')
            hipLaunchKernelGGL(k_pf_sweep<RMV_N>, dim3(nu), dim3(PF_TPB), 0, s, (i64)n, v, l, b, U, ldu);
        else
            hipLaunchKernelGGL(k_pf_sweep<RMV_T>, dim3(nu), dim3(PF_TPB), 0, s, (i64)n, v, l, b, U, ldu);
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}

}  // extern "C"
