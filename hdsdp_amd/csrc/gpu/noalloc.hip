// noalloc.hip -- the Jacobi-preconditioned Schur CG for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// hdsdp_dev_kkt_solve keeps its CG vectors in device memory the library allocates with
// the solver state.  hdsdp_dev_kkt_solve_noalloc runs the same kernels in the same order
// on a caller-held Schur matrix, with its vectors in one file-scope device array,
// `scratch`: NOALLOC_VECS vectors of at most NOALLOC_MAX_M entries and the scalar slots
// the host reads back.  Each GPU the library runs on holds its own copy of the array
// (about 4 MiB) until the process ends; its device address is looked up once per device
// with hipGetSymbolAddress and only that address is handed to kernels and copies.
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU race.
#include "hdsdp_gpu.h"
#include "dev_common.h"
#include "vec.h"

#include <cstdio>
#include <mutex>

constexpr int NOALLOC_MAX_M = 65536;
// jac, r, rn, p, z, Mp, tmp, x
constexpr int NOALLOC_VECS = 8;
constexpr size_t SCRATCH_LEN = (size_t) NOALLOC_VECS * NOALLOC_MAX_M + hdk::SC_SLOTS;

static __device__ __align__(256) double scratch[SCRATCH_LEN];

namespace {

constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES];
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up on first use per device).
hdsdp_retcode scratch_on_device(double **p) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        fprintf(stderr, "[hdsdp_amd] kkt_solve_noalloc: no current HIP device\n");
        return HDSDP_RETCODE_FAILED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q) {
            fprintf(stderr, "[hdsdp_amd] kkt_solve_noalloc: hipGetSymbolAddress failed: %s\n", hipGetErrorString(e));
            return HDSDP_RETCODE_FAILED;
        }
        g_scratch[dev] = static_cast<double *>(q);
    }
    *p = g_scratch[dev];
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode read_sc(hipStream_t st, const double *sc_d, double *sc) {
    if (hipMemcpyAsync(sc, sc_d, sizeof(double) * hdk::SC_SLOTS, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return HDSDP_RETCODE_FAILED;
    return HDSDP_RETCODE_OK;
}

}  // namespace

extern "C" hdsdp_retcode hdsdp_dev_kkt_solve_noalloc(const double *M, int m, const double *rhs, double *sol,
                                                     double relTol, double absTol, int maxIter, int nRestartFreq,
                                                     void *stream, int *nIter) {
    if (nIter) *nIter = 0;
    if (m <= 0 || m > NOALLOC_MAX_M) {
        fprintf(stderr, "[hdsdp_amd] kkt_solve_noalloc: m = %d outside 1..%d (the scratch array's limit)\n", m,
                NOALLOC_MAX_M);
        return HDSDP_RETCODE_FAILED;
    }
    if (!M || !rhs || !sol || !(relTol > 0.0) || !(absTol > 0.0) || maxIter <= 0) return HDSDP_RETCODE_FAILED;
    double *base = nullptr;
    if (scratch_on_device(&base) != HDSDP_RETCODE_OK) return HDSDP_RETCODE_FAILED;
    hipStream_t st = static_cast<hipStream_t>(stream);
    double *jac = base, *r = jac + NOALLOC_MAX_M, *rn = r + NOALLOC_MAX_M, *p = rn + NOALLOC_MAX_M;
    double *z = p + NOALLOC_MAX_M, *Mp = z + NOALLOC_MAX_M, *tmp = Mp + NOALLOC_MAX_M, *x = tmp + NOALLOC_MAX_M;
    double *sc_d = x + NOALLOC_MAX_M;
    double sc[hdk::SC_SLOTS];
    auto symv = [&](const double *in, double *out) {
        hdk::coldot<double>(st, m, m, 0, 1.0, M, m, in, 0.0, out);
    };

    // conjGradSolve (linalg/hdsdp_linsolver.c:855-997) with the Jacobi preconditioner:
    // the control flow of hdsdp_dev_kkt_solve before its switch to the Cholesky factor
    const int restart = nRestartFreq > 20 ? nRestartFreq : 20;
    hdk::get_diag(st, m, M, m, jac);
    hdk::cg_start(st, m, rhs, r, x, sc_d);
    if (read_sc(st, sc_d, sc) != HDSDP_RETCODE_OK) return HDSDP_RETCODE_FAILED;
    double rnorm = sc[hdk::CG_RNORM];
    const double bnorm = sc[hdk::CG_BNORM];
    double tol = relTol * bnorm < absTol ? relTol * bnorm : absTol;
    if (tol < 0.1 * absTol) tol = 0.1 * absTol;
    if (rnorm < tol) return HDSDP_RETCODE_OK;  // tiny right-hand side: sol left as it was
    hdk::cg_jacobi(st, m, r, jac, z, p);
    symv(p, Mp);
    int iter = 0;
    bool converged = false;
    for (iter = 0; iter < maxIter; ++iter) {
        hdk::cg_step1(st, m, z, r, p, Mp, x, rn, sc_d);
        if (iter % restart == 5) {
            symv(x, tmp);
            hdk::cg_residual(st, m, rhs, tmp, r);
            hdk::cg_jacobi(st, m, r, jac, p, nullptr);
            symv(p, Mp);
            hdk::cg_jacobi(st, m, r, jac, z, nullptr);
            continue;
        }
        hdk::cg_jacobi(st, m, rn, jac, z, nullptr);
        hdk::cg_step2(st, m, rn, z, p, r, sc_d);
        symv(p, Mp);
        if (read_sc(st, sc_d, sc) != HDSDP_RETCODE_OK) return HDSDP_RETCODE_FAILED;
        rnorm = sc[hdk::CG_RNORM];
        if (rnorm != rnorm) return HDSDP_RETCODE_FAILED;
        if (iter > 20 && rnorm > 0.01 * bnorm) break;
        if (rnorm < tol) {
            converged = true;
            break;
        }
    }
    if (nIter) *nIter = iter;
    if (!converged) {
        fprintf(stderr,
                "[hdsdp_amd] kkt_solve_noalloc: Jacobi PCG did not reach the tolerance in %d iterations; the "
                "Cholesky preconditioner needs m^2 device memory (use hdsdp_dev_kkt_solve)\n",
                iter);
        return HDSDP_RETCODE_FAILED;
    }
    if (hipMemcpyAsync(sol, x, sizeof(double) * (size_t) m, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess)
        return HDSDP_RETCODE_FAILED;
    return HDSDP_RETCODE_OK;
}
