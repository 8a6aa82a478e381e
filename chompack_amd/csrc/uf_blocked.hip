// update_factor (update_factor.c:3-108) for tasks with many reflections, on
// the whole device.  The single-workgroup kernel (k_update_factor) spends its
// time on two things when the child's separator is long: the reflections,
// each a chain of passes over a band of global memory, and applying each
// block of reflections to every column left of it, a product that one
// workgroup streams at a few percent of the device.  Here:
//   * k_uf_setup: the element copies of Li's first N1 columns from the front
//     and of C = Lk[:, r2 - nn] from the parent's factor, a workgroup per
//     column;
//   * k_uf_step, launched once per block of up to UF_BMAX consecutive
//     reflections and once more: one workgroup per task applies the previous
//     block's reflections to the block's own columns and then forms the
//     block's reflections with those columns (rows lo .. lo + len - 1) in
//     LDS, leaving the reflected columns, the block's tau and band starts,
//     and the upper triangular T of the compact WY form
//     H_{nb-1} ... H_0 = I - V T^T V^T (the dlarft forward recurrence) in
//     global memory; the other workgroups apply the previous block's form to
//     16-column tiles of the columns left of it, w = V^T c, y = T^T w,
//     c -= V y.  T alternates between two slots per task, so the step that
//     reads one block's T never writes it.
// The arithmetic of each reflection is k_update_factor's (dlarfg convention
// beta = -sign(alpha) ||[alpha; x]||, a pivot below PIVOT_NOISE ||x|| taken
// as +0); k_update_factor then continues from Li22 (UFTask::ext_reduce).
#include "common.h"
#include "kernels.h"

#define PIVOT_NOISE 1e-13
#define UB_TPB 256
#define UB_WAVES (UB_TPB / 64)
#define UB_SHORT 32  // reflections with bands up to this many rows are formed in registers
static_assert(UB_TPB == UF_BMAX * UF_BMAX, "one thread per entry of T");

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace {

__device__ inline double wsum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

// Li(:, j) for j < N (from the front for j < N1, zero beyond) and
// C(:, j) = Lk(:, r[N1 + j] - nn) below its diagonal; item (task, column, N1).
__global__ __launch_bounds__(UB_TPB) void k_uf_setup(const char* __restrict__ list, int ntask,
                                                     BufTab bt, const int* __restrict__ relidx) {
    const UFTask* tasks = reinterpret_cast<const UFTask*>(list);
    const int4 it = reinterpret_cast<const int4*>(list + (i64)ntask * sizeof(UFTask))[blockIdx.x];
    const UFTask t = tasks[it.x];
    const int j = it.y, N1 = it.z;
    const int nn = t.nn, na = t.na, N = t.N, nj = nn + na;
    const int* r = relidx + t.relpos;
    if (j < N) {
        const double* F = buf_sel(bt, t.fbuf) + t.foff;
        double* Li = bt.p[1] + t.outoff + (i64)j * N;
        const i64 fj = j < N1 ? (i64)r[j] * nj : 0;
        for (int i = threadIdx.x; i < N; i += UB_TPB)
            Li[i] = (j < N1 && i >= j) ? F[r[i] + fj] : 0.0;
    } else {
        const int jc = j - N;
        const int cj = r[N1 + jc] - nn;
        const double* Lk = bt.p[1] + t.lkoff + (i64)cj * na;
        double* C = bt.p[2] + t.woff + (i64)jc * na;
        for (int i = threadIdx.x; i < na; i += UB_TPB) C[i] = (i >= cj) ? Lk[i] : 0.0;
    }
}

// One block of reflections of one task (see the file comment); its columns
// have taken every earlier block's reflections.  Reflection
// b of the block is js = js0 + b, on column col = N2 - 1 - js of C, band
// rows jj .. piv - 1 with jj = r[N - 1 - js] - nn and pivot row
// piv = na - 1 - js.
__device__ void uf_reflect_block(const UFBlock& B, BufTab bt, const int* __restrict__ relidx,
                                 double* lds) {
    double* S = lds;  // S[b * len + i] = C(lo + i, col_b)
    double* red = S + UF_VCAP;
    double (*gram)[UF_BMAX] = reinterpret_cast<double (*)[UF_BMAX]>(red + UB_WAVES);
    double (*tmat)[UF_BMAX] = gram + UF_BMAX;
    double (*gpart)[UF_BMAX][UF_BMAX] = reinterpret_cast<double (*)[UF_BMAX][UF_BMAX]>(tmat + UF_BMAX);
    double* s_tau = &gpart[UB_WAVES][0][0];
    double* s_scal = s_tau + UF_BMAX;
    double* s_beta = s_scal + UF_BMAX;
    int* s_jj = reinterpret_cast<int*>(s_beta + UF_BMAX);
    const int nn = B.nn, na = B.na, N = B.N, N2 = B.N2;
    const int js0 = B.js0, nbk = B.nbk, lo = B.lo, len = B.len;
    const int* r = relidx + B.relpos;
    double* C = bt.p[2] + B.woff;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int b = 0; b < nbk; ++b) {  // eight loads in flight per thread
        const double* cb = C + lo + (i64)(N2 - 1 - (js0 + b)) * na;
        for (int i0 = 0; i0 < len; i0 += 8 * UB_TPB) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u * UB_TPB + tid;
                v[u] = cb[i < len ? i : 0];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u * UB_TPB + tid;
                if (i < len) S[b * len + i] = v[u];
            }
        }
    }
    if (tid < UF_BMAX) {
        s_tau[tid] = 0.0;
        s_scal[tid] = 0.0;  // (0: not a short-band reflection)
        s_jj[tid] = tid < nbk ? r[N - 1 - (js0 + tid)] - nn : 0;
    }
    __syncthreads();
    for (int b = 0; b < nbk; ++b) {
        const int js = js0 + b;
        const int jj = s_jj[b], piv = na - 1 - js;
        const int bl = piv - jj;
        if (bl <= 0) continue;
        double* cc = S + b * len;
        const int i0 = jj - lo, ip = piv - lo;
        if (bl <= UB_SHORT) {
            // short band: every thread holds it and forms the reflection; the
            // thread of each column left of col applies it; one barrier
            double x[UB_SHORT];
            double ss = 0.0;
#pragma unroll
            for (int u = 0; u < UB_SHORT; ++u) {
                x[u] = u < bl ? cc[i0 + min(u, bl - 1)] : 0.0;
                ss += x[u] * x[u];
            }
            const double xn = sqrt(ss);
            double alpha = cc[ip];
            if (fabs(alpha) <= PIVOT_NOISE * xn) alpha = 0.0;
            if (xn == 0.0) continue;
            const double beta = -copysign(hypot(alpha, xn), alpha);
            const double tau = (beta - alpha) / beta;
            const double scal = 1.0 / (alpha - beta);
#pragma unroll
            for (int u = 0; u < UB_SHORT; ++u) x[u] *= scal;
            const int b2 = b + 1 + tid;
            if (b2 < nbk) {
                double* cd = S + b2 * len;
                double rows[UB_SHORT];
#pragma unroll
                for (int u = 0; u < UB_SHORT; ++u) rows[u] = u < bl ? cd[i0 + min(u, bl - 1)] : 0.0;
                const double pv = cd[ip];
                double d = pv;
#pragma unroll
                for (int u = 0; u < UB_SHORT; ++u) d += x[u] * rows[u];
                const double f = tau * d;
#pragma unroll
                for (int u = 0; u < UB_SHORT; ++u)
                    if (u < bl) cd[i0 + u] = rows[u] - f * x[u];
                cd[ip] = pv - f;
            }
            // column b itself is scaled after the block's last reflection
            // (nothing reads it before): its tau, scal and beta are kept
            if (tid == 0) {
                s_tau[b] = tau;
                s_scal[b] = scal;
                s_beta[b] = beta;
            }
            __syncthreads();
            continue;
        }
        double s = 0.0, s1 = 0.0;
        {
            int i = i0 + tid;
            for (; i + UB_TPB < ip; i += 2 * UB_TPB) {
                s += cc[i] * cc[i];
                s1 += cc[i + UB_TPB] * cc[i + UB_TPB];
            }
            if (i < ip) s += cc[i] * cc[i];
        }
        s = wsum64(s + s1);
        if (lane == 0) red[wv] = s;
        __syncthreads();
        double ss = 0.0;
#pragma unroll
        for (int w = 0; w < UB_WAVES; ++w) ss += red[w];
        const double xn = sqrt(ss);
        double alpha = cc[ip];
        if (fabs(alpha) <= PIVOT_NOISE * xn) alpha = 0.0;
        __syncthreads();  // every thread has read the pivot and the partial sums
        if (xn == 0.0) continue;
        const double beta = -copysign(hypot(alpha, xn), alpha);
        const double tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        for (int i = i0 + tid; i < ip; i += UB_TPB) cc[i] *= scal;
        if (tid == 0) {
            cc[ip] = beta;
            s_tau[b] = tau;
        }
        __syncthreads();
        // H = I - tau v v^T, v = [cc[i0:ip]; 1], on rows i0..ip of the
        // block's columns left of col (reflections b + 1 ..): 16 threads per
        // column, each a strided share of the band, summed across the 16
        // lanes
        {
            const int q = tid >> 4, sl = tid & 15;
            const bool on = b + 1 + q < nbk;
            double* cd = S + (on ? b + 1 + q : b) * len;
            double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
            if (on) {  // four rows in flight per step
                int i = i0 + sl;
                for (; i + 48 < ip; i += 64) {
                    d0 += cc[i] * cd[i];
                    d1 += cc[i + 16] * cd[i + 16];
                    d2 += cc[i + 32] * cd[i + 32];
                    d3 += cc[i + 48] * cd[i + 48];
                }
                for (; i < ip; i += 16) d0 += cc[i] * cd[i];
            }
            d0 = (d0 + d1) + (d2 + d3);
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) d0 += __shfl_xor(d0, o, 64);
            const double f = on ? tau * (d0 + cd[ip]) : 0.0;
            if (on) {
                int i = i0 + sl;
                for (; i + 48 < ip; i += 64) {
                    const double c0 = cc[i], c1 = cc[i + 16], c2 = cc[i + 32], c3 = cc[i + 48];
                    const double e0 = cd[i], e1 = cd[i + 16], e2 = cd[i + 32], e3 = cd[i + 48];
                    cd[i] = e0 - f * c0;
                    cd[i + 16] = e1 - f * c1;
                    cd[i + 32] = e2 - f * c2;
                    cd[i + 48] = e3 - f * c3;
                }
                for (; i < ip; i += 16) cd[i] -= f * cc[i];
                if (sl == 0) cd[ip] -= f;
            }
        }
        __syncthreads();
    }
    // the short-band reflections' own columns: v = x scal, beta at the pivot
    for (int b = 0; b < nbk; ++b) {
        if (s_scal[b] == 0.0) continue;
        const int i0 = s_jj[b] - lo, ip = na - 1 - (js0 + b) - lo;
        double* cc = S + b * len;
        if (tid < ip - i0) cc[i0 + tid] *= s_scal[b];
        if (tid == 0) cc[ip] = s_beta[b];
    }
    __syncthreads();
    // the reflected columns back to C
    for (int b = 0; b < nbk; ++b) {
        double* cb = C + lo + (i64)(N2 - 1 - (js0 + b)) * na;
        for (int i = tid; i < len; i += UB_TPB) cb[i] = S[b * len + i];
    }
    if (B.colmin == 0) return;  // nothing left of the block
    // V_b: zero outside its band, 1 at its pivot (zero for tau_b = 0), in S
    for (int b = 0; b < nbk; ++b) {
        const int pb = na - 1 - (js0 + b), jb = s_jj[b];
        const bool on = s_tau[b] != 0.0;
        for (int i = tid; i < len; i += UB_TPB) {
            const int row = lo + i;
            const double v = S[b * len + i];
            S[b * len + i] = !on ? 0.0 : ((row == pb) ? 1.0 : ((row >= jb && row < pb) ? v : 0.0));
        }
    }
    __syncthreads();
    // gram = V^T V (16 x 16) on the matrix cores: v_mfma_f64_16x16x4f64 with
    // A = V^T and B = V, both operands lane l's V(row0 + l / 16, l % 16); the
    // waves take interleaved groups of four rows, their sums added through LDS
    {
        dbl4 acc = {0.0, 0.0, 0.0, 0.0};
        const int cb = lane & 15, rk = lane >> 4;
        for (int r0 = 4 * wv; r0 < len; r0 += 4 * UB_WAVES) {
            const int i = r0 + rk;
            const double v = (cb < nbk && i < len) ? S[cb * len + i] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v, v, acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) gpart[wv][(lane >> 4) + 4 * e][lane & 15] = acc[e];
    }
    __syncthreads();
    {
        const int a = tid >> 4, b = tid & 15;
        gram[a][b] = (gpart[0][a][b] + gpart[1][a][b]) + (gpart[2][a][b] + gpart[3][a][b]);
        tmat[a][b] = 0.0;
    }
    __syncthreads();
    // T (dlarft forward): T(b, b) = tau_b, T(a, b) = -tau_b sum_{a <= k < b}
    // T(a, k) gram(k, b); thread (a, k) forms one product, the 16 lanes of
    // row a sum them
    for (int b = 0; b < nbk; ++b) {
        const int a = tid >> 4, k = tid & 15;
        double pr = (a < b && k >= a && k < b) ? tmat[a][k] * gram[k][b] : 0.0;
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) pr += __shfl_xor(pr, o, 64);
        __syncthreads();  // row a of T read before it is written
        if (k == 0) {
            if (a < b) tmat[a][b] = -s_tau[b] * pr;
            else if (a == b) tmat[b][b] = s_tau[b];
        }
        __syncthreads();
    }
    double* Tg = bt.p[2] + B.toff;  // T, then tau, then the band starts
    Tg[tid] = tmat[tid / UF_BMAX][tid % UF_BMAX];
    if (tid < UF_BMAX) {
        Tg[UF_BMAX * UF_BMAX + tid] = s_tau[tid];
        Tg[UF_BMAX * UF_BMAX + UF_BMAX + tid] = (double)s_jj[tid];
    }
}

// The block's reflections applied to those of the columns c0 .. c0 + 15 of
// C below `bound` and outside xlo .. xhi - 1: w = V^T c, y = T^T w,
// c -= V y.  Thread t: column c0 + t / 16, rows lo + (t % 16) + 16 m.
__device__ void uf_apply_tile(const UFBlock& B, int c0, int bound, int xlo, int xhi, BufTab bt,
                              double* lds) {
    double* Vb = lds;  // Vb[b * len + i] = V_b(lo + i)
    double (*Wp)[UF_BMAX][17] = reinterpret_cast<double (*)[UF_BMAX][17]>(Vb + UF_VCAP);
    double (*Ts)[UF_BMAX] = reinterpret_cast<double (*)[UF_BMAX]>(&Wp[16][0][0]);
    double (*Wt)[17] = reinterpret_cast<double (*)[17]>(&Ts[UF_BMAX][0]);
    double (*Yt)[17] = Wt + UF_BMAX;
    const int na = B.na, N2 = B.N2, js0 = B.js0, nbk = B.nbk, lo = B.lo, len = B.len;
    double* C = bt.p[2] + B.woff;
    const double* Tg = bt.p[2] + B.toff;
    const int tid = threadIdx.x;
    const int cl = tid >> 4, rg = tid & 15;
    const int col = c0 + cl;
    for (int b = 0; b < nbk; ++b) {  // eight loads in flight per thread
        const int pb = na - 1 - (js0 + b), jb = (int)Tg[UF_BMAX * UF_BMAX + UF_BMAX + b];
        const bool on = Tg[UF_BMAX * UF_BMAX + b] != 0.0;
        const double* cb = C + (i64)(N2 - 1 - (js0 + b)) * na;
        for (int i0 = 0; i0 < len; i0 += 8 * UB_TPB) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int row = lo + i0 + u * UB_TPB + tid;
                v[u] = cb[(row >= jb && row < pb) ? row : 0];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u * UB_TPB + tid, row = lo + i;
                if (i < len)
                    Vb[b * len + i] = !on ? 0.0 : ((row == pb) ? 1.0 : ((row >= jb && row < pb) ? v[u] : 0.0));
            }
        }
    }
    Ts[tid / UF_BMAX][tid % UF_BMAX] = Tg[tid];
    __syncthreads();
    const bool live = col < bound && (col < xlo || col >= xhi);
    double* cc = C + (i64)col * na + lo;
    double w[UF_BMAX];
#pragma unroll
    for (int b = 0; b < UF_BMAX; ++b) w[b] = 0.0;
    if (live)
        for (int i = rg; i < len; i += 16) {
            const double x = cc[i];
#pragma unroll
            for (int b = 0; b < UF_BMAX; ++b)
                if (b < nbk) w[b] += Vb[b * len + i] * x;
        }
#pragma unroll
    for (int b = 0; b < UF_BMAX; ++b) Wp[rg][b][cl] = w[b];
    __syncthreads();
    {
        const int b = tid >> 4, c = tid & 15;  // W = V^T C: the 16 row groups summed
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < 16; ++g) s += Wp[g][b][c];
        Wt[b][c] = s;
    }
    __syncthreads();
    {
        const int b = tid >> 4, c = tid & 15;  // Y = T^T W (T upper triangular)
        double y = 0.0;
        for (int a = 0; a <= b; ++a) y += Ts[a][b] * Wt[a][c];
        Yt[b][c] = y;
    }
    __syncthreads();
    if (!live) return;
    double y[UF_BMAX];
#pragma unroll
    for (int b = 0; b < UF_BMAX; ++b) y[b] = Yt[b][cl];
    for (int i = rg; i < len; i += 16) {
        double x = cc[i];
#pragma unroll
        for (int b = 0; b < UF_BMAX; ++b)
            if (b < nbk) x -= Vb[b * len + i] * y[b];
        cc[i] = x;
    }
}

// LDS of a k_uf_step workgroup: the larger of the two roles' layouts
// (V or the block's columns, then the 16-row-group partial sums and the
// 16 x 16 T, W and Y of a tile; or the block's reduction scratch)
#define UF_STEP_LDS (UF_VCAP + 16 * UF_BMAX * 17 + UF_BMAX * UF_BMAX + 2 * UF_BMAX * 17)
static_assert(UF_STEP_LDS >= UF_VCAP + UB_WAVES + (2 + UB_WAVES) * UF_BMAX * UF_BMAX + 4 * UF_BMAX,
              "k_uf_step LDS");

// One step of the blocked reduction: role 0 items apply a block's compact WY
// form to a 16-column tile; role 1 items apply the previous block's form to
// the block's own columns, then run the block's reflections.
__global__ __launch_bounds__(UB_TPB) void k_uf_step(const UFBlock* __restrict__ items, BufTab bt,
                                                    const int* __restrict__ relidx) {
    __shared__ double lds[UF_STEP_LDS];
    const UFBlock B = items[blockIdx.x];
    if (B.role == 0) {
        uf_apply_tile(B, B.c0, B.colmin, B.xlo, B.xhi, bt, lds);
        return;
    }
    if (B.pnbk > 0) {
        UFBlock P = B;
        P.js0 = B.pjs0;
        P.nbk = B.pnbk;
        P.lo = B.plo;
        P.len = B.plen;
        P.toff = B.ptoff;
        uf_apply_tile(P, B.colmin, B.colmin + B.nbk, 0, 0, bt, lds);
        __syncthreads();  // the updated columns are read back from C below
    }
    uf_reflect_block(B, bt, relidx, lds);
}
