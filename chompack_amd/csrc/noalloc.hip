// Allocation-free variant of chp_spmv_residual for callers that cannot call
// hipMalloc (code under HIP graph capture, tight refinement loops, hosts that
// forbid allocation): the row-major copy of X that the multi-right-hand-side
// residual needs lives in a fixed device array instead of a caller-supplied
// or library-allocated workspace.
//
// The array is a file-scope device variable: every GPU the library runs on
// holds its own copy (RESID_SCRATCH_DOUBLES doubles, 64 MiB) from module load
// until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in symbol); its device address is
// looked up once per device with hipGetSymbolAddress and cached.
//
// All calls on one GPU share that GPU's array, so calls that overlap in time
// on one GPU (two streams, two host threads) race on it; serialize them.
// Nothing in the array needs to start at zero: each call overwrites the part
// it reads before reading it.
#include <hip/hip_runtime.h>

#include "common.h"
#include "../../include/chompack_amd.h"

#include <mutex>
#include <string>

namespace {

constexpr int64_t RESID_SCRATCH_DOUBLES = CHP_RESID_NOALLOC_MAX;  // 8,388,608 doubles = 64 MiB
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[RESID_SCRATCH_DOUBLES];

std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {nullptr};

// Device address of `scratch` on the current device (looked up on the first
// call on that device, with that device current).
double* scratch_on_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        chp_set_error("spmv_residual_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            chp_set_error(std::string("spmv_residual_noalloc: hipGetSymbolAddress: ") +
                          hipGetErrorString(e));
            return nullptr;
        }
        g_ptr[dev] = static_cast<double*>(p);
    }
    return g_ptr[dev];
}

}  // namespace

extern "C" {

int chp_spmv_residual_noalloc(int64_t n, const int64_t* rowptr, const int* colind,
                              const double* vals, const double* X, int ldx, const double* B,
                              int ldb, double* R, int ldr, int nrhs, void* stream) {
    if (n <= 0 || nrhs <= 0) return 0;
    if (n > RESID_SCRATCH_DOUBLES || (int64_t)nrhs * n > RESID_SCRATCH_DOUBLES) {
        chp_set_error("spmv_residual_noalloc: n * nrhs = " + std::to_string(n) + " * " +
                      std::to_string(nrhs) + " exceeds the fixed workspace of " +
                      std::to_string(RESID_SCRATCH_DOUBLES) + " doubles");
        return -1;
    }
    double* w = scratch_on_current_device();
    if (!w) return -2;
    return chp_spmv_residual(n, rowptr, colind, vals, X, ldx, B, ldb, R, ldr, nrhs,
                             nrhs > 1 ? w : nullptr, stream);
}

}  // extern "C"
