// update_factor on the GPU: one 256-thread workgroup per child update
// (update_factor.c:3-108).  The child's update V = F[r, r] is produced in
// factored form V = Li^T Li from the parent's frontal matrix whose (2,2)
// block holds the parent's factor Lk (V_k = Lk^T Lk):
//   * A = F[r1, r1], B = F[r2, r1] are gathered into Li;
//   * C = Lk[:, r2 - nn] is reduced to lower-triangular form by Householder
//     reflections from its last column (the LAPACK dlarfg convention:
//     beta = -sign(alpha) ||[alpha; x]||), giving Li22.  Pivots can vanish
//     structurally; a pivot below PIVOT_NOISE ||x|| is taken as +0, the
//     exact-arithmetic value, so the signs of Li22's rows do not follow
//     rounding noise;
//   * Li21 solves Li22^T Li21 = B (back substitution, one column per thread);
//   * Li11 is the reversed Cholesky factor of A - Li21^T Li21
//     (A = Li11^T Li11, eliminated from the last row up).
// All work is in global memory (workspace C in the scratch buffer); the
// sequential steps are separated by workgroup barriers.
#include "common.h"
#include "kernels.h"

#define TPB 256
#define PIVOT_NOISE 1e-13

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(TPB) void k_update_factor(const UFTask* __restrict__ tasks,
                                                       BufTab bt, const int* __restrict__ relidx,
                                                       unsigned long long* status) {
    const UFTask t = tasks[blockIdx.x];
    const int nn = t.nn, na = t.na, N = t.N, nj = nn + na;
    const int* r = relidx + t.relpos;
    const double* F = bt.p[t.fbuf] + t.foff;
    const double* Lk = bt.p[1] + t.lkoff;
    double* Li = bt.p[1] + t.outoff;
    double* C = bt.p[2] + t.woff;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ int s_n1;
    __shared__ double red[4];
    __shared__ double s_tau, s_beta, s_scal, s_piv;
    if (tid == 0) {
        int c = 0;
        while (c < N && r[c] < nn) ++c;
        s_n1 = c;
    }
    __syncthreads();
    const int N1 = s_n1, N2 = N - N1;
    for (i64 e = tid; e < (i64)N * N; e += TPB) {
        const int i = (int)(e % N), j = (int)(e / N);
        Li[e] = (j < N1 && i >= j) ? F[r[i] + (i64)r[j] * nj] : 0.0;
    }
    for (i64 e = tid; e < (i64)na * N2; e += TPB) {
        const int i = (int)(e % na), j = (int)(e / na);
        const int cj = r[N1 + j] - nn;
        C[e] = (i >= cj) ? Lk[i + (i64)cj * na] : 0.0;
    }
    __syncthreads();
    if (N2 > 0) {
        for (int js = 0; js < N2; ++js) {
            const int col = N2 - 1 - js;
            const int jj = r[N - 1 - js] - nn;
            if (na - jj - (js + 1) == 0) continue;
            const int piv = na - 1 - js;
            double* cc = C + (i64)col * na;
            double s = 0.0;
            for (int i = jj + tid; i < piv; i += TPB) s += cc[i] * cc[i];
            s = wave_sum(s);
            if (lane == 0) red[wv] = s;
            __syncthreads();
            if (tid == 0) {
                const double xn = sqrt((red[0] + red[1]) + (red[2] + red[3]));
                double alpha = cc[piv];
                if (fabs(alpha) <= PIVOT_NOISE * xn) alpha = 0.0;
                if (xn == 0.0) {
                    s_tau = 0.0;
                } else {
                    const double beta = -copysign(hypot(alpha, xn), alpha);
                    s_tau = (beta - alpha) / beta;
                    s_beta = beta;
                    s_scal = 1.0 / (alpha - beta);
                }
            }
            __syncthreads();
            const double tau = s_tau;
            if (tau == 0.0) continue;
            const double scal = s_scal;
            for (int i = jj + tid; i < piv; i += TPB) cc[i] *= scal;
            if (tid == 0) cc[piv] = s_beta;
            __syncthreads();
            // H = I - tau v v^T, v = [cc[jj:piv]; 1], on rows jj..piv of columns < col
            for (int c = wv; c < col; c += 4) {
                double* cd = C + (i64)c * na;
                double d = 0.0;
                for (int i = jj + lane; i < piv; i += 64) d += cc[i] * cd[i];
                d = wave_sum(d) + cd[piv];
                const double f = tau * d;
                for (int i = jj + lane; i < piv; i += 64) cd[i] -= f * cc[i];
                if (lane == 0) cd[piv] -= f;
            }
            __syncthreads();
        }
        for (i64 e = tid; e < (i64)N2 * N2; e += TPB) {
            const int i = (int)(e % N2), j = (int)(e / N2);
            if (i >= j) Li[(N1 + i) + (i64)(N1 + j) * N] = C[(na - N2 + i) + (i64)j * na];
        }
        __syncthreads();
        const double* L22 = Li + N1 + (i64)N1 * N;
        for (int c = tid; c < N1; c += TPB) {
            double* x = Li + N1 + (i64)c * N;
            for (int i = N2 - 1; i >= 0; --i) {
                double v = x[i];
                for (int k = i + 1; k < N2; ++k) v -= L22[k + (i64)i * N] * x[k];
                x[i] = v / L22[i + (i64)i * N];
            }
        }
        __syncthreads();
        for (i64 e = tid; e < (i64)N1 * N1; e += TPB) {
            const int i = (int)(e % N1), j = (int)(e / N1);
            if (i < j) continue;
            double s0 = 0.0;
            const double* xi = Li + N1 + (i64)i * N;
            const double* xj = Li + N1 + (i64)j * N;
            for (int k = 0; k < N2; ++k) s0 += xi[k] * xj[k];
            Li[i + (i64)j * N] -= s0;
        }
        __syncthreads();
    }
    // reversed Cholesky of the (1,1) block: A = Lr^T Lr, last row first
    for (int p = N1 - 1; p >= 0; --p) {
        if (tid == 0) s_piv = Li[p + (i64)p * N];
        __syncthreads();
        const double d = s_piv;
        if (!(d > 0.0)) {
            if (tid == 0) chp_record_failure(status, t.key, N1 - p);
            return;
        }
        const double sd = sqrt(d), rs = 1.0 / sd;
        for (int j = tid; j < p; j += TPB) Li[p + (i64)j * N] *= rs;
        if (tid == 0) Li[p + (i64)p * N] = sd;
        __syncthreads();
        for (i64 e = tid; e < (i64)p * p; e += TPB) {
            const int i = (int)(e % p), j = (int)(e / p);
            if (i >= j) Li[i + (i64)j * N] -= Li[p + (i64)i * N] * Li[p + (i64)j * N];
        }
        __syncthreads();
    }
}
