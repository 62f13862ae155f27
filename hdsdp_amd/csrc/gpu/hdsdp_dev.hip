// hdsdp_dev.hip -- the hdsdp_dev_* C ABI (include/hdsdp_gpu.h) on one MI355X.
//
// All n x n and m x m state of the SDP block lives in HBM: the three dual
// buffers, two Cholesky factors with their inverses (used as triangular
// solves by GEMV), S^-1, the Schur matrix M and its preconditioner.  Host
// control follows the reference routines cited per function; per call only
// m-vectors and scalars cross the host boundary.
#include "hdsdp_gpu.h"
#include "dev_common.h"
#include "schur.h"
#include "vec.h"
#include "bf16planes.h"
#include "eig_small.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

using hdk::DevCoeffs;

// preconditioner of the Schur CG
enum { PREC_JACOBI = 0, PREC_CHOL64 = 1, PREC_CHOL32 = 2 };

struct hdsdp_dev {
    int m, n, precision;
    hipStream_t st;
    hdk::LookAhead la;  // second stream + events for look-ahead factorisations
    std::vector<void *> allocs;
    DevCoeffs c;
    int cType;
    int conDenseKind;  // a constraint coefficient is DENSE or DSR1
    int psdpReady;     // hdsdp_dev_psdp_init has set the primal iterate
    std::vector<int> heavy;  // coefficient indices handled by the dense sandwich path
    int hasZero;
    // Gram formulations (all constraints of one kind): dense rank-one M = s s^T o (F^T W F)^2,
    // dense M = T^T T with T_p = L^-1 A_p L^-T
    int r1Gram, dnGram, nDn, r1Identity;
    double *H1;              // n x (nR1 + 1): W F
    double *Kr;              // nR1 x nR1 when the slot -> row map is not the identity
    double *Tall, *Tdg;      // packed strictly-lower T_p (K1 = n (n - 1) / 2 each), diagonals (n each)
    double *Dbuf, *Ubuf;     // gramBatch x n^2
    double *Gpart;           // gramParts x nDn^2
    int gramBatch, gramParts;
    long gramChunk;
    // mixed precision Gram builds (bf16 split products): FP32 copies and products
    float *LiTf;             // n x n, L^-T of the dual factor
    float *Ff, *Gf;          // n x nR1: rank-one factors and G = L^-1 F
    float *Tf, *Tdgf, *Dbuff, *Vf;  // mixed: packed T_p and diagonals, batch buffers
    // mixed dense Gram from bf16 planes (three per operand, split once by the producing
    // kernel): L^-T, a batch of A_p and of V_p = A_p L^-T (ld plane_ld(n)), the packed T_p
    // (ld K1p, a multiple of 32, entries K1..K1p zero) and their diagonals
    __bf16 *LiTp, *Dp, *Vp, *Tp, *Tdgp;
    __bf16 *Dall;  // planes of every dense constraint, split once at set-up (or null)
    long K1p, chunkP;
    int planes;
    // adaptive precision: the Gram blocks of M come from bf16 split products while
    // schurMixed; when the system they form cannot be factored or solved, they are
    // rebuilt in FP64 from the S^-1 and L^-1 of their build (Wm, Lim), the diagonal
    // terms added to M since its build (diag(M) - mDiag0) are carried over, and later
    // builds stay FP64
    int schurMixed, mGramPending;
    double *Wm, *Lim, *mDiag0, *mDiagCur;
    double *xCoef;  // m + 1 coefficients of the matrix-free Schur product
    double *splitW;  // k-split partials of the rank-nR1 buffer update
    size_t splitLen;

    double *S[3], *L[2], *Li[2], *LiT[2];
    long facVer[2], liVer[2];
    double *W, *W2, *X, *T1, *T2, *work;
    double *coef_d, *vecA_d, *vecB_d, *sc_d;
    int *info_d;
    // pinned staging
    double *hcoef, *hvecA, *hvecB, *hsc, *hvec_n, *hY, *hHist;
    int *hinfo;
    bool stagingBusy;

    // Lanczos (linalg/hdsdp_lanczos.c): the dual cone's solver (lzComputed, lwarm) and the
    // primal method's own (lzComputedP, lwarmP; interface/hdsdp_psdp.c:113-116)
    int lzMax, lzComputed, lzComputedP;
    double *V, *Hd, *lv, *lw, *lz1, *lz2, *lwarm, *lwarmP, *ly, *mv1, *mv2;
    // primal method (interface/hdsdp_psdp.c): the primal iterate is X; Xs its value at the
    // start (the fixed scaling), dXp the direction, XSX the buffer X S X (+ Xs dS Xs), Sold
    // the dual matrix of the iteration, LX / LXi / LXiT the Cholesky factor of X and inverses
    double *Xs, *dXp, *XSX, *Sold, *LX, *LXi, *LXiT;
    double *lzHist;  // (alpha_k, ||w_k||) of every Lanczos step, read at the check steps
    double *symvW;   // tile partials of the lower-triangle product with M (large m), else null
    std::vector<double> Hh;
    std::vector<double> lzStart, lzPert;  // Lanczos start vector and warm-start perturbation (fixed by n)

    // Schur system; mVer counts changes of M, cholVer is the mVer its FP64 factor
    // (Mchol, MLi, MLiT) was computed from (-1: none)
    long mVer, cholVer;
    double *M, *Mchol, *MLi, *MLiT;
    float *Mf, *MLif, *MLiTf, *cf1, *cf2, *workf;
    double *jac, *cr, *crn, *cp, *cz, *cMp, *cx, *crhs, *ctmp;
    double *precS;     // Jacobi scaling of the FP32 preconditioner
    double precShift;  // diagonal shift the FP32 factorisation needed
    int chol32Min;      // mixed: FP32 factor preconditioner for an FP64-built M once m >= this
    double *hsdCoef;    // coefficient vector e_0 (C alone) for the self-dual embedding's terms
    bool wFromDual;     // W holds S^-1 of the last Schur build (not the primal method's X)
    double mixedFloor;  // mixed: refinement against the exact operator stops at this fraction of ||b||
    double *cb, *cxa;   // mixed refinement: right-hand side and accumulated solution
    // the indefinite solver the Schur system switches to for good once its Cholesky
    // factorisation or the CG it preconditions fails: LU of M with partial pivoting
    bool useLU;
    double *LUm;
    int *luPiv, *luPerm, *luInfo;
    std::vector<int> hpiv;
    int precKind;
    int useJacobi, paramUseJacobi, maxIter, nRestart;
    double absTol, relTol;
    double stats[7];

    // device-time profile (events on st)
    bool prof;
    hipEvent_t evA[HDSDP_PROF_SLOTS], evB[HDSDP_PROF_SLOTS];
    double profMs[HDSDP_PROF_SLOTS], profCnt[HDSDP_PROF_SLOTS];
};

// bf16 terms per FP32 operand in the mixed-precision Gram products: three (six MFMA
// products, ~24 bits) keeps the Gram blocks FP32-accurate, so that M stays positive
// definite to a condition number near 1e7 before an FP64 rebuild is needed.
constexpr int BF16_TERMS = 3;

namespace {

template <typename T>
hdsdp_retcode dalloc(hdsdp_dev *d, T **p, size_t count) {
    void *q = nullptr;
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 8;
    if (hipMalloc(&q, bytes) != hipSuccess) {
        fprintf(stderr, "[hdsdp_amd] hipMalloc of %zu bytes failed\n", bytes);
        return HDSDP_RETCODE_MEMORY;
    }
    d->allocs.push_back(q);
    *p = (T *) q;
    return HDSDP_RETCODE_OK;
}

// device allocation cleared to zero on the library's stream
template <typename T>
hdsdp_retcode dalloc_zero(hdsdp_dev *d, T **p, size_t count) {
    hdsdp_retcode rc = dalloc(d, p, count);
    if (rc != HDSDP_RETCODE_OK) return rc;
    HIP_OK(hipMemsetAsync(*p, 0, sizeof(T) * std::max<size_t>(count, 1), d->st));
    return HDSDP_RETCODE_OK;
}

template <typename T>
hdsdp_retcode halloc(T **p, size_t count) {
    void *q = nullptr;
    if (hipHostMalloc(&q, (count ? count : 1) * sizeof(T), hipHostMallocDefault) != hipSuccess)
        return HDSDP_RETCODE_MEMORY;
    *p = (T *) q;
    return HDSDP_RETCODE_OK;
}

#define DCALL(expr)                                   \
    do {                                              \
        hdsdp_retcode r_ = (expr);                    \
        if (r_ != HDSDP_RETCODE_OK) return r_;        \
    } while (0)

hdsdp_retcode sync(hdsdp_dev *d) {
    HIP_OK(hipStreamSynchronize(d->st));
    d->stagingBusy = false;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode check_launch(void) {
    HIP_OK(hipGetLastError());
    return HDSDP_RETCODE_OK;
}

// upload `count` doubles from host memory through the pinned staging `stage`
hdsdp_retcode upload(hdsdp_dev *d, double *dst, const double *src, size_t count, double *stage) {
    if (d->stagingBusy) DCALL(sync(d));
    memcpy(stage, src, count * sizeof(double));
    HIP_OK(hipMemcpyAsync(dst, stage, count * sizeof(double), hipMemcpyHostToDevice, d->st));
    d->stagingBusy = true;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode download(hdsdp_dev *d, double *dst, const double *src, size_t count, double *stage) {
    HIP_OK(hipMemcpyAsync(stage, src, count * sizeof(double), hipMemcpyDeviceToHost, d->st));
    DCALL(sync(d));
    memcpy(dst, stage, count * sizeof(double));
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode read_scalars(hdsdp_dev *d) {
    HIP_OK(hipMemcpyAsync(d->hsc, d->sc_d, hdk::SC_SLOTS * sizeof(double), hipMemcpyDeviceToHost, d->st));
    return sync(d);
}

int fidx(int whichBuffer) { return whichBuffer == 0 ? 0 : 1; }

void prof_begin(hdsdp_dev *d, int k) {
    if (d->prof) (void) hipEventRecord(d->evA[k], d->st);
}

hdsdp_retcode prof_end(hdsdp_dev *d, int k) {
    if (!d->prof) return HDSDP_RETCODE_OK;
    HIP_OK(hipEventRecord(d->evB[k], d->st));
    HIP_OK(hipEventSynchronize(d->evB[k]));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, d->evA[k], d->evB[k]));
    d->profMs[k] += ms;
    d->profCnt[k] += 1.0;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode ensure_linv(hdsdp_dev *d, int f) {
    if (d->liVer[f] == d->facVer[f]) return HDSDP_RETCODE_OK;
    hdk::trtri<double>(d->st, d->n, d->L[f], d->n, d->Li[f], d->n, d->work);
    hdk::transpose<double>(d->st, d->n, d->Li[f], d->n, d->LiT[f], d->n);
    d->liVer[f] = d->facVer[f];
    return check_launch();
}

// The Lanczos operator x -> L^-1 (-D) L^-T x: the dual cone's with the factor of a dual
// buffer and D = dS (interface/hdsdp_conic_sdp.c:462-505), the primal method's with the
// factor of X and D = the primal direction (interface/hdsdp_psdp.c:58-66).
struct LzOp {
    const double *Li, *LiT, *D;  // L^-1, L^-T, D (n x n, both triangles)
    int *computed;               // calls so far (0: the fixed start vector)
    double *warm;                // warm start of the next call
};

void lz_mvec(hdsdp_dev *d, const LzOp &op, const double *in, double *out) {
    const int n = d->n;
    hdk::coldot<double>(d->st, n, n, 1, 1.0, op.Li, n, in, 0.0, d->mv1);
    hdk::coldot<double>(d->st, n, n, 0, -1.0, op.D, n, d->mv1, 0.0, d->mv2);
    hdk::coldot<double>(d->st, n, n, 2, 1.0, op.LiT, n, d->mv2, 0.0, out);
    d->stats[3] += 1;
}

// glibc rand start vector and warm-start perturbation (linalg/hdsdp_lanczos.c:33-53).
// Both restart the generator from srand(n) and reseed it per entry, so they are fixed
// by n: they are generated once per device (n reseeds of glibc's generator cost ~1.5 us
// each, ~3 ms per Lanczos call at n = 2000) and reused.  The perturbation entries are
// 1e-3 sqrt(sqrt(r)) times +-0.5 (an exact product), so adding the stored entry gives
// the same sum as adding the expression.
void lz_prepare(hdsdp_dev *d, int n, double *vVec) {
    if ((int) d->lzStart.size() != n) {
        d->lzStart.resize(n);
        srand(n);
        for (int i = 0; i < n; ++i) {
            srand(rand());
            d->lzStart[i] = sqrt(sqrt((rand() % 1627))) * (rand() % 2 - 0.5);
        }
    }
    memcpy(vVec, d->lzStart.data(), sizeof(double) * n);
}

void lz_perturb(hdsdp_dev *d, int n, double *vVec) {
    if ((int) d->lzPert.size() != n) {
        d->lzPert.resize(n);
        srand(n);
        for (int i = 0; i < n; ++i) {
            srand(rand());
            d->lzPert[i] = 1e-03 * sqrt(sqrt((rand() % 1627))) * (rand() % 2 - 0.5);
        }
    }
    for (int i = 0; i < n; ++i) vVec[i] += d->lzPert[i];
}

// HLanczosSolve (linalg/hdsdp_lanczos.c:161-292) with device mat-vecs.  The host
// needs a step's (alpha, ||w||) only for the tridiagonal at the check steps (every
// `check` steps and at a breakdown), so the steps run back to back and their scalars
// are read once per check window and processed step by step in the reference's order.
// A breakdown or a converged check inside a window has let the device run the rest
// of the window's steps; those only write Krylov columns beyond the ones the
// estimate combines, so the result is that of the step-by-step loop.
hdsdp_retcode lanczos(hdsdp_dev *d, const LzOp &op, double *maxStep) {
    const int n = d->n, maxDim = d->lzMax, nH = maxDim + 1;
    std::vector<double> &H = d->Hh;
    std::fill(H.begin(), H.end(), 0.0);
    if (*op.computed == 0) {
        lz_prepare(d, n, d->hvec_n);
    } else {
        HIP_OK(hipMemcpyAsync(d->hvec_n, op.warm, n * sizeof(double), hipMemcpyDeviceToHost, d->st));
        DCALL(sync(d));
        lz_perturb(d, n, d->hvec_n);
    }
    hdk::fill(d->st, d->V, (long) n * nH, 0.0);
    hdk::fill(d->st, d->Hd, (long) nH * nH, 0.0);
    HIP_OK(hipMemcpyAsync(d->lw, d->hvec_n, n * sizeof(double), hipMemcpyHostToDevice, d->st));
    hdk::lz_start(d->st, n, d->lw, d->lv, d->V);
    const int check = std::min(maxDim / 5, 3);
    double Td[32], Te[32], evals[2], Y[2 * 32];
    int kRead = 0;  // first step whose scalars the host has not processed
    bool done = false;
    for (int k = 0; k < maxDim && !done; ++k) {
        lz_mvec(d, op, d->lv, d->lw);
        hdk::lz_step(d->st, n, k, d->lw, d->V, d->Hd, nH, d->lv, d->sc_d, d->lzHist);
        if ((k + 1) % check != 0 && k < maxDim - 1) continue;
        HIP_OK(hipMemcpyAsync(d->hHist + 2 * kRead, d->lzHist + 2 * kRead, sizeof(double) * 2 * (k + 1 - kRead),
                              hipMemcpyDeviceToHost, d->st));
        DCALL(sync(d));
        for (int j = kRead; j <= k && !done; ++j) {
            const double alpha = d->hHist[2 * j], normPres = d->hHist[2 * j + 1];
            H[j + j * nH] = alpha;
            if (normPres > 0.0) H[(j + 1) + j * nH] = H[j + (j + 1) * nH] = normPres;
            if (!((j + 1) % check == 0 || j > maxDim - 1 || normPres == 0.0)) continue;
            const int kp1 = j + 1;
            for (int i = 0; i < kp1; ++i) {
                Td[i] = H[i + nH * i];
                if (i + 1 < kp1) Te[i] = H[(i + 1) + nH * i];
            }
            if (kp1 < 2 || hdsdp_eig_tridiag_top(kp1, Td, Te, 2, evals, Y) != 0) return HDSDP_RETCODE_FAILED;
            double resi = fabs(H[kp1 + j * nH] * Y[kp1 + j]);
            if (resi < 1e-04 || j >= maxDim - 1) {
                double e1 = evals[1], e2 = evals[0];
                memcpy(d->hY, Y, sizeof(double) * 2 * kp1);
                HIP_OK(hipMemcpyAsync(d->ly, d->hY, sizeof(double) * 2 * kp1, hipMemcpyHostToDevice, d->st));
                hdk::lz_comb(d->st, n, kp1, d->V, d->ly + kp1, d->lz1);
                lz_mvec(d, op, d->lz1, d->lz2);
                HIP_OK(hipMemcpyAsync(op.warm, d->lz2, n * sizeof(double), hipMemcpyDeviceToDevice, d->st));
                hdk::axpy_norm(d->st, n, -e1, d->lz1, d->lz2, d->sc_d, hdk::LZ_R1);
                hdk::lz_comb(d->st, n, kp1, d->V, d->ly, d->lz2);
                lz_mvec(d, op, d->lz2, d->lz1);
                hdk::axpy_norm(d->st, n, -e1, d->lz2, d->lz1, d->sc_d, hdk::LZ_R2);
                DCALL(read_scalars(d));
                double r1 = d->hsc[hdk::LZ_R1], r2 = d->hsc[hdk::LZ_R2];
                double diff = e1 - e2 - r2;
                double gamma = (diff > 0) ? diff : 1e-16;
                double r1sq = r1 * r1 / gamma;
                gamma = std::min(r1, r1sq);
                if (gamma < 1e-03 || gamma + e1 <= 0.5) {
                    *maxStep = (gamma + e1 <= 0.0) ? HDSDP_INFINITY : 1.0 / (gamma + e1);
                    done = true;
                } else {
                    if (normPres == 0.0) return HDSDP_RETCODE_FAILED;
                    *maxStep = 1.0 / (gamma + e1);
                }
            }
        }
        kRead = k + 1;
    }
    *op.computed += 1;
    return check_launch();
}

// FP64 Cholesky preconditioner: M = L L^T, kept as L^-1 and L^-T.
hdsdp_retcode build_chol64(hdsdp_dev *d) {
    const int m = d->m;
    if (!d->Mchol) {
        DCALL(dalloc(d, &d->Mchol, (size_t) m * m));
        DCALL(dalloc(d, &d->MLi, (size_t) m * m));
        DCALL(dalloc(d, &d->MLiT, std::max((size_t) m * m, hdk::POTRF_WORK)));
    }
    HIP_OK(hipMemcpyAsync(d->Mchol, d->M, sizeof(double) * (size_t) m * m, hipMemcpyDeviceToDevice, d->st));
    d->cholVer = -1;
    if (!hdk::potrf<double>(d->st, m, d->Mchol, m, d->MLiT, std::max((size_t) m * m, hdk::POTRF_WORK), d->info_d,
                            &d->la))
        return HDSDP_RETCODE_FAILED;
    HIP_OK(hipMemcpyAsync(d->hinfo, d->info_d, sizeof(int), hipMemcpyDeviceToHost, d->st));
    DCALL(sync(d));
    if (*d->hinfo != 0) return HDSDP_RETCODE_FAILED;
    hdk::trtri<double>(d->st, m, d->Mchol, m, d->MLi, m, d->MLiT);
    hdk::transpose<double>(d->st, m, d->MLi, m, d->MLiT, m);
    d->precKind = PREC_CHOL64;
    d->cholVer = d->mVer;
    return check_launch();
}

// FP32 Cholesky preconditioner of the equilibrated matrix S M S (S = diag(M)^-1/2):
// unit diagonal, so the FP32 factorisation sees the best-scaled form; when it still
// breaks down (M too ill-conditioned for FP32) a growing diagonal shift is added.
// The preconditioner only steers the FP64 CG, whose residuals decide convergence.
hdsdp_retcode build_chol32(hdsdp_dev *d, int *ok) {
    const int m = d->m;
    if (!d->Mf) {
        DCALL(dalloc(d, &d->Mf, (size_t) m * m));
        DCALL(dalloc(d, &d->MLif, (size_t) m * m));
        DCALL(dalloc(d, &d->MLiTf, (size_t) m * m));
        DCALL(dalloc(d, &d->cf1, (size_t) m));
        DCALL(dalloc(d, &d->cf2, (size_t) m));
        DCALL(dalloc(d, &d->workf, std::max((size_t) m * m, hdk::POTRF_WORK)));
        DCALL(dalloc(d, &d->precS, (size_t) m));
    }
    *ok = 0;
    hdk::jacobi_scale(d->st, m, d->M, m, d->precS);
    // M only grows more ill-conditioned along the path: start from the shift the
    // previous factorisation needed
    const double shifts[] = {0.0, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2};
    for (double shift : shifts) {
        if (shift < d->precShift) continue;
        hdk::convert_lower_scaled(d->st, m, d->M, m, d->precS, shift, d->Mf, m);
        if (!hdk::potrf<float>(d->st, m, d->Mf, m, d->workf, std::max((size_t) m * m, hdk::POTRF_WORK), d->info_d,
                               &d->la))
            return HDSDP_RETCODE_FAILED;
        HIP_OK(hipMemcpyAsync(d->hinfo, d->info_d, sizeof(int), hipMemcpyDeviceToHost, d->st));
        DCALL(sync(d));
        if (*d->hinfo == 0) {
            d->precShift = shift;
            *ok = 1;
            break;
        }
    }
    if (!*ok) return HDSDP_RETCODE_OK;
    hdk::trtri<float>(d->st, m, d->Mf, m, d->MLif, m, d->workf);
    hdk::transpose<float>(d->st, m, d->MLif, m, d->MLiTf, m);
    d->precKind = PREC_CHOL32;
    return check_launch();
}

hdsdp_retcode build_precond(hdsdp_dev *d) {
    const int m = d->m;
    if (d->useJacobi) {
        hdk::get_diag(d->st, m, d->M, m, d->jac);
        d->precKind = PREC_JACOBI;
        return check_launch();
    }
    d->stats[4] += 1;
    prof_begin(d, HDSDP_PROF_MFACTOR);
    // Mixed precision factors M in FP32 when its Gram blocks are bf16 products (an FP32-
    // accurate matrix) or when M is large enough for the FP32 factorisation to save more
    // than the extra CG steps of the coarser preconditioner cost (measured: m = 20000
    // saves 15 ms per iteration for 9 ms of CG; m = 8084 saves 1.7 ms for 7 ms of CG)
    if (d->precision == HDSDP_PRECISION_MIXED && (d->mGramPending || d->m >= d->chol32Min)) {
        int ok = 0;
        DCALL(build_chol32(d, &ok));
        if (!ok) DCALL(build_chol64(d));
    } else {
        DCALL(build_chol64(d));
    }
    return prof_end(d, HDSDP_PROF_MFACTOR);
}

// z = P^-1 r (and optionally p = z)
void apply_precond(hdsdp_dev *d, const double *r, double *z, double *p) {
    const int m = d->m;
    if (d->precKind == PREC_JACOBI) {
        hdk::cg_jacobi(d->st, m, r, d->jac, z, p);
        return;
    }
    if (d->precKind == PREC_CHOL32) {
        // z = S (L^-T L^-1) S r with the FP32 factor of S M S
        hdk::d2f_scaled(d->st, m, r, d->precS, d->cf1);
        hdk::coldot<float>(d->st, m, m, 2, 1.0f, d->MLiTf, m, d->cf1, 0.0f, d->cf2);
        hdk::coldot<float>(d->st, m, m, 1, 1.0f, d->MLif, m, d->cf2, 0.0f, d->cf1);
        hdk::f2d_scaled(d->st, m, d->cf1, d->precS, z);
    } else {
        hdk::coldot<double>(d->st, m, m, 2, 1.0, d->MLiT, m, r, 0.0, d->ctmp);
        hdk::coldot<double>(d->st, m, m, 1, 1.0, d->MLi, m, d->ctmp, 0.0, z);
    }
    if (p) hdk::cg_jacobi(d->st, m, z, nullptr, p, nullptr);
}

// y = M x.  Large M (m >= 16384 by default, HDSDP_SYMV_LOWER_MIN overrides) is read only
// in its lower triangle, half the bytes of the column-dot product over the full matrix;
// smaller M keep the column-dot product and its summation order.
void symv_M(hdsdp_dev *d, const double *x, double *y) {
    if (d->symvW) hdk::symv_lower(d->st, d->m, d->M, d->m, x, y, d->symvW);
    else hdk::coldot<double>(d->st, d->m, d->m, 0, 1.0, d->M, d->m, x, 0.0, y);
}

// y = M x in FP64 without M: (M x)_i = <A_i, W (sum_j x_j A_j) W> for the S^-1 = W of
// the build, plus the diagonal terms added to M since (diag(M) - its built diagonal).
// Used while the Gram blocks of M hold bf16-split products: it makes the residuals of the
// refinement those of the exact FP64 system.
void schur_apply_exact(hdsdp_dev *d, const double *x, double *y) {
    const int m = d->m, n = d->n;
    hdk::shift_coef(d->st, m, x, d->xCoef);
    hdk::assemble_buffer(d->st, d->c, d->xCoef, 0.0, d->T1, 0.0, d->work, d->splitW, d->splitLen);
    hdk::mirror_lower<double>(d->st, n, d->T1, n);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->Wm, n, 0, d->T1, n, 0, 0.0, d->T2, n, 0, 1);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->T2, n, 0, d->Wm, n, 0, 0.0, d->T1, n, 0, 1);
    hdk::traces(d->st, d->c, 1, m, d->T1, nullptr, 0.0, y, nullptr, false);
    hdk::r1_quadforms(d->st, d->c, d->T1, false, 1, m, 0.0, y, nullptr, d->H1);
    hdk::diag_delta_axpy(d->st, m, d->M, m, d->mDiag0, x, y);
}

int cg_solve(hdsdp_dev *d);

// conjGradSolve (linalg/hdsdp_linsolver.c:855-997) on crhs with the M on the device;
// result in cx.  Returns 0 on success, 1 if the solution is left unwritten (tiny rhs),
// -1 on failure.
int cg_solve_inner(hdsdp_dev *d) {
    const int m = d->m;
    const int restart = std::max(d->nRestart, 20);
    hdk::cg_start(d->st, m, d->crhs, d->cr, d->cx, d->sc_d);
    if (read_scalars(d) != HDSDP_RETCODE_OK) return -1;
    double rnorm = d->hsc[hdk::CG_RNORM], bnorm = d->hsc[hdk::CG_BNORM];
    double tol = std::min(d->absTol, bnorm * d->relTol);
    tol = std::max(tol, 0.1 * d->absTol);
    if (rnorm < tol) return 1;
    apply_precond(d, d->cr, d->cz, d->cp);
    symv_M(d, d->cp, d->cMp);
    int iter = 0, status = 0, refine = 0;
    double prevTrue = 0.0, rtrue = -1.0;
    for (iter = 0; iter < d->maxIter; ++iter) {
        hdk::cg_step1(d->st, m, d->cz, d->cr, d->cp, d->cMp, d->cx, d->crn, d->sc_d);
        d->stats[2] += 1;
        if (iter % restart == 5 && d->useJacobi) {
            symv_M(d, d->cx, d->ctmp);
            hdk::cg_residual(d->st, m, d->crhs, d->ctmp, d->cr);
            apply_precond(d, d->cr, d->cp, nullptr);
            symv_M(d, d->cp, d->cMp);
            apply_precond(d, d->cr, d->cz, nullptr);
            continue;
        }
        apply_precond(d, d->crn, d->cz, nullptr);
        hdk::cg_step2(d->st, m, d->crn, d->cz, d->cp, d->cr, d->sc_d);
        symv_M(d, d->cp, d->cMp);
        if (read_scalars(d) != HDSDP_RETCODE_OK) return -1;
        rnorm = d->hsc[hdk::CG_RNORM];
        if (rnorm != rnorm) return -1;
        if (iter > 20 && rnorm > 0.01 * bnorm) {
            status = 2;
            break;
        }
        if (rnorm < tol) {
            // With the FP32 preconditioner the recursively updated residual drifts
            // from b - M x: confirm convergence on the true FP64 residual and, if it
            // is not met, restart from x (iterative refinement).  Refinement stops at
            // tol or once the true residual stops halving (the rounding level of an
            // FP64 product with M).  The FP64 factor needs no such check (as in the FP64 path).
            if (d->precision != HDSDP_PRECISION_MIXED || d->useJacobi || d->precKind != PREC_CHOL32) break;
            if (refine >= 8) break;
            symv_M(d, d->cx, d->ctmp);
            hdk::cg_true_residual(d->st, m, d->crhs, d->ctmp, d->cr, d->sc_d);
            if (read_scalars(d) != HDSDP_RETCODE_OK) return -1;
            rtrue = d->hsc[hdk::CG_TRUE];
            if (rtrue != rtrue) return -1;
            if (rtrue < tol) break;
            if (refine > 0 && rtrue > 0.5 * prevTrue) break;
            prevTrue = rtrue;
            ++refine;
            apply_precond(d, d->cr, d->cz, d->cp);
            symv_M(d, d->cp, d->cMp);
        }
    }
    d->stats[5] = iter;
    static const int trace = getenv("HDSDP_TRACE_CG") != nullptr;
    if (trace)
        fprintf(stderr,
                "[hdsdp_amd] cg m=%d precond=%s iters=%d refine=%d rnorm=%.3e rtrue=%.3e bnorm=%.3e tol=%.3e "
                "status=%d\n",
                m, d->useJacobi ? "jacobi" : (d->precKind == PREC_CHOL32 ? "chol-fp32" : "chol-fp64"), iter, refine,
                rnorm, rtrue, bnorm, tol, status);
    if (iter >= d->maxIter || status == 2) {
        if (d->useJacobi) {
            d->useJacobi = 0;
            d->paramUseJacobi = 0;
            if (build_precond(d) != HDSDP_RETCODE_OK) return -1;
            return cg_solve(d);
        }
        if (d->precKind == PREC_CHOL32) {
            // the FP32 preconditioner was not good enough for this system: FP64 factor
            if (build_chol64(d) != HDSDP_RETCODE_OK) return -1;
            return cg_solve_inner(d);
        }
        return -1;
    }
    return 0;
}

// While the Gram blocks of M hold bf16-split products (FP32-accurate) and the
// factor preconditioner is in use, the system solved is the exact FP64 one: iterative
// refinement x += M~^-1 (b - M x), the residual from the matrix-free FP64 product and
// each correction from CG on the device M~ (products with M~, preconditioned by its
// factor).  Converged once ||b - M x|| < max(tol, mixedFloor ||b||), the level an FP64
// direct solve reaches (a relative 1e-15 or so of ||b|| once M is ill-conditioned);
// fails when a step does not halve the residual before that, and M is then rebuilt in
// FP64.
int cg_solve_refined(hdsdp_dev *d) {
    const int m = d->m;
    HIP_OK(hipMemcpyAsync(d->cb, d->crhs, sizeof(double) * m, hipMemcpyDeviceToDevice, d->st));
    hdk::axpy_norm(d->st, m, 0.0, d->cb, d->cb, d->sc_d, hdk::SC_AUX);
    if (read_scalars(d) != HDSDP_RETCODE_OK) return -1;
    const double bnorm = d->hsc[hdk::SC_AUX];
    const double tol = std::max(std::min(d->absTol, bnorm * d->relTol), 0.1 * d->absTol);
    const double accept = std::max(tol, d->mixedFloor * bnorm);
    if (bnorm < tol) return 1;
    hdk::fill(d->st, d->cxa, m, 0.0);
    static const int trace = getenv("HDSDP_TRACE_CG") != nullptr;
    double prev = HUGE_VAL, rtrue = bnorm;
    int rc = -1, step = 0;
    for (step = 0; step < 8; ++step) {
        const int inner = cg_solve_inner(d);
        if (inner < 0) break;
        if (inner == 0) hdk::axpy(d->st, m, 1.0, d->cx, d->cxa);
        schur_apply_exact(d, d->cxa, d->ctmp);
        hdk::cg_true_residual(d->st, m, d->cb, d->ctmp, d->crhs, d->sc_d);
        if (read_scalars(d) != HDSDP_RETCODE_OK) break;
        rtrue = d->hsc[hdk::CG_TRUE];
        if (trace)
            fprintf(stderr, "[hdsdp_amd] refine m=%d step=%d rtrue=%.3e (%.2e of bnorm) accept=%.3e\n", m, step, rtrue,
                    rtrue / bnorm, accept);
        if (rtrue != rtrue) break;
        if (rtrue < accept) {
            rc = 0;
            break;
        }
        if (inner == 1 || rtrue > 0.5 * prev) break;
        prev = rtrue;
    }
    HIP_OK(hipMemcpyAsync(d->cx, d->cxa, sizeof(double) * m, hipMemcpyDeviceToDevice, d->st));
    HIP_OK(hipMemcpyAsync(d->crhs, d->cb, sizeof(double) * m, hipMemcpyDeviceToDevice, d->st));
    return rc;
}

int cg_solve(hdsdp_dev *d) {
    d->useJacobi = d->paramUseJacobi;
    if (d->mGramPending && !d->useJacobi) return cg_solve_refined(d);
    return cg_solve_inner(d);
}

// M(p, q) = <T_p, T_q> over the dense constraints with T_p = L^-1 A_p L^-T (so that
// <A_p, W A_q W> = <T_p, T_q> for W = L^-T L^-1).  T_p is symmetric, so only its lower
// triangle is formed, and <T_p, T_q> = 2 <off_p, off_q> + <dg_p, dg_q> over the packed
// strictly-lower parts and the diagonals: a split-K Gram product over n (n - 1) / 2
// with alpha = 2 (exact) and one over n, summed in a fixed order.
void dense_gram(hdsdp_dev *d, const double *Li) {
    const int n = d->n, r = d->nDn, m = d->m;
    const long nn = (long) n * n, K1 = (long) n * (n - 1) / 2;
    for (int b0 = 0; b0 < r; b0 += d->gramBatch) {
        const int cnt = std::min(d->gramBatch, r - b0);
        hdk::unpack_dense(d->st, d->c, b0, cnt, d->Dbuf);
        // U = L^-1 A_p, then the lower triangle of T_p = U L^-T = L^-1 U^T: both products
        // have the lower-triangular L^-1 on the left, whose zero k ranges are skipped
        hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, Li, n, 0, d->Dbuf, n, nn, 0.0, d->Ubuf, n, nn, cnt,
                          2);
        hdk::gemm<double>(d->st, false, true, true, n, n, n, 1.0, Li, n, 0, d->Ubuf, n, nn, 0.0, d->Dbuf, n, nn, cnt,
                          2);
        hdk::pack_lower<double>(d->st, n, cnt, d->Dbuf, nn, d->Tall + (size_t) b0 * K1, d->Tdg + (size_t) b0 * n);
    }
    const long kc = d->gramChunk;
    const int full = (int) (K1 / kc);
    const long rem = K1 - (long) full * kc;
    const long rr = (long) r * r;
    const int nOff = full + (rem > 0 ? 1 : 0);
    if (full > 0)
        hdk::gemm<double>(d->st, true, false, true, r, r, (int) kc, 2.0, d->Tall, (int) K1, kc, d->Tall, (int) K1, kc,
                          0.0, d->Gpart, r, rr, full);
    if (rem > 0)
        hdk::gemm<double>(d->st, true, false, true, r, r, (int) rem, 2.0, d->Tall + (long) full * kc, (int) K1, 0,
                          d->Tall + (long) full * kc, (int) K1, 0, 0.0, d->Gpart + (long) full * rr, r, 0, 1);
    hdk::gemm<double>(d->st, true, false, true, r, r, n, 1.0, d->Tdg, n, 0, d->Tdg, n, 0, 0.0,
                      d->Gpart + (long) nOff * rr, r, 0, 1);
    hdk::gram_reduce(d->st, r, nOff + 1, d->Gpart, r, rr, d->c.denseCon, d->M, m);
}

// Rebuild the Gram blocks of M in FP64 from the factor of their build, re-apply the
// diagonal terms added since, and keep later builds in FP64.
hdsdp_retcode rebuild_gram_fp64(hdsdp_dev *d) {
    const int m = d->m, n = d->n;
    d->mVer += 1;
    d->schurMixed = 0;
    d->mGramPending = 0;
    hdk::get_diag(d->st, m, d->M, m, d->mDiagCur);
    if (d->r1Gram) {
        hdk::gemm<double>(d->st, false, false, false, n, d->c.nR1, n, 1.0, d->Wm, n, 0, d->c.rF, n, 0, 0.0, d->H1, n, 0,
                          1);
        if (d->r1Identity) hdk::r1_schur(d->st, d->c, d->H1, d->M, m, d->M, m);
        else hdk::r1_schur(d->st, d->c, d->H1, d->Kr, d->c.nR1, d->M, m);
    }
    if (d->dnGram) {
        if (!d->Tall) {
            const size_t nn = (size_t) n * n, K1 = (size_t) n * (n - 1) / 2;
            DCALL(dalloc(d, &d->Tall, K1 * d->nDn));
            DCALL(dalloc(d, &d->Tdg, (size_t) n * d->nDn));
            DCALL(dalloc(d, &d->Dbuf, nn * d->gramBatch));
            DCALL(dalloc(d, &d->Ubuf, nn * d->gramBatch));
        }
        dense_gram(d, d->Lim);
    }
    // the rewritten diagonal entries get back what was added to them after the build
    if (d->r1Gram) hdk::add_diag_delta(d->st, d->c.nR1, d->c.rCon, d->mDiagCur, d->mDiag0, d->M, m);
    if (d->dnGram) hdk::add_diag_delta(d->st, d->nDn, d->c.denseCon, d->mDiagCur, d->mDiag0, d->M, m);
    hdk::mirror_lower<double>(d->st, m, d->M, m);
    static const int trace = getenv("HDSDP_TRACE_CG") != nullptr;
    if (trace)
        fprintf(stderr, "[hdsdp_amd] Schur matrix of build %d rebuilt in FP64; later builds stay FP64\n",
                (int) d->stats[0]);
    return check_launch();
}

// Mixed precision: the same products on the bf16 matrix cores (FP32-equivalent split
// products), written as X^T Y with both operands contiguous along k:
//   V_p = A_p L^-T = (A_p)^T (L^-T),  T_p = L^-1 V_p = (L^-T)^T V_p,  M = T^T T.
void dense_gram_mixed_inloop(hdsdp_dev *d) {
    const int n = d->n, r = d->nDn, m = d->m;
    const long nn = (long) n * n, K1 = (long) n * (n - 1) / 2;
    hdk::d2f(d->st, nn, d->LiT[0], d->LiTf);
    for (int b0 = 0; b0 < r; b0 += d->gramBatch) {
        const int cnt = std::min(d->gramBatch, r - b0);
        hdk::unpack_dense_f(d->st, d->c, b0, cnt, d->Dbuff);
        // V = A_p L^-T: column j of L^-T is zero below row j (tri 1)
        hdk::gemm_bf16s_tn<float>(d->st, BF16_TERMS, false, n, n, n, 1.0, d->Dbuff, n, nn, d->LiTf, n, 0, 0.0, d->Vf, n,
                                  nn, cnt, 1);
        // lower triangle of T_p = (L^-T)^T V only (symmetric), into the free batch buffer,
        // then packed; column i of the left operand L^-T is zero below row i (tri 2)
        hdk::gemm_bf16s_tn<float>(d->st, BF16_TERMS, true, n, n, n, 1.0, d->LiTf, n, 0, d->Vf, n, nn, 0.0, d->Dbuff, n,
                                  nn, cnt, 2);
        hdk::pack_lower<float>(d->st, n, cnt, d->Dbuff, nn, d->Tf + (size_t) b0 * K1, d->Tdgf + (size_t) b0 * n);
    }
    const long kc = d->gramChunk;
    const int full = (int) (K1 / kc);
    const long rem = K1 - (long) full * kc;
    const long rr = (long) r * r;
    const int nOff = full + (rem > 0 ? 1 : 0);
    if (full > 0)
        hdk::gemm_bf16s_tn<double>(d->st, BF16_TERMS, true, r, r, (int) kc, 2.0, d->Tf, (int) K1, kc, d->Tf, (int) K1,
                                   kc, 0.0, d->Gpart, r, rr, full);
    if (rem > 0)
        hdk::gemm_bf16s_tn<double>(d->st, BF16_TERMS, true, r, r, (int) rem, 2.0, d->Tf + (long) full * kc, (int) K1, 0,
                                   d->Tf + (long) full * kc, (int) K1, 0, 0.0, d->Gpart + (long) full * rr, r, 0, 1);
    hdk::gemm_bf16s_tn<double>(d->st, BF16_TERMS, true, r, r, n, 1.0, d->Tdgf, n, 0, d->Tdgf, n, 0, 0.0,
                               d->Gpart + (long) nOff * rr, r, 0, 1);
    hdk::gram_reduce(d->st, r, nOff + 1, d->Gpart, r, rr, d->c.denseCon, d->M, m);
}

// Mixed precision: the products on the bf16 matrix cores (FP32-equivalent split
// products), written as X^T Y with both operands contiguous along k:
//   U_p^T = E_p^T L^-T,  T_p = U_p L^-T + (U_p L^-T)^T,  M = T^T T.
// Every operand is held as three bf16 planes written by the kernel that produces it
// (L^-T and A_p by split kernels, U_p^T by the first product's epilogue, T_p by the
// second's).
void dense_gram_mixed(hdsdp_dev *d) {
    const int n = d->n, r = d->nDn, m = d->m;
    const long nn = (long) n * n;
    if (!d->planes) {
        dense_gram_mixed_inloop(d);
        return;
    }
    const int ldn = (int) hdk::plane_ld(n);
    const long nnp = (long) ldn * n, bps = nnp * d->gramBatch, K1p = d->K1p, rr = (long) r * r;
    hdk::split_matrix(d->st, n, n, d->LiT[0], n, d->LiTp, ldn, nnp);
    for (int b0 = 0; b0 < r; b0 += d->gramBatch) {
        const int cnt = std::min(d->gramBatch, r - b0);
        // A_p planes: resident (split at set-up) or split for this batch
        const __bf16 *Ap = d->Dall ? d->Dall + (size_t) b0 * nnp : d->Dp;
        const long psA = d->Dall ? nnp * r : bps;
        if (!d->Dall) hdk::unpack_dense_split(d->st, d->c, b0, cnt, d->Dp, ldn, bps, nnp);
        // With A_p = E_p + E_p^T (E_p the lower triangle of A_p, diagonal halved) and the
        // lower-triangular U_p = L^-1 E_p, T_p = L^-1 A_p L^-T = U_p L^-T + (U_p L^-T)^T.
        // U_p^T = E_p^T L^-T (upper triangular, k from i to j: tri 3) is written as planes
        // (the tiles below the diagonal of the zero-filled buffer stay zero); then the
        // lower triangle of T_p, entry (i, j) = sum_k U^T(k, i) L^-T(k, j) + L^-T(k, i) U^T(k, j),
        // as one product over both sums (every operand column i zero below row i: tri 4),
        // written as packed planes of its strictly lower part and of its diagonal, the
        // operands of the Gram products below.  n^3 / 6 + n^3 / 3 multiply-adds per
        // constraint, against n^3 / 2 + n^3 / 3 for V_p = A_p L^-T, T_p = L^-1 V_p.
        hdk::gemm_bf16p(d->st, false, n, n, ldn, 1.0, Ap, ldn, psA, nnp, d->LiTp, ldn, nnp, 0, 0.0, d->Vp,
                        hdk::BF16P_OUT_PLANES, ldn, bps, nnp, cnt, 3);
        hdk::gemm_bf16p_packed(d->st, n, ldn, 1.0, d->Vp, ldn, bps, nnp, d->LiTp, ldn, nnp, 0,
                               d->Tp + (size_t) b0 * K1p, K1p, K1p * r, d->Tdgp + (size_t) b0 * ldn, ldn,
                               (long) ldn * r, cnt, 4, d->LiTp, nnp, 0, d->Vp, bps, nnp);
    }
    const long kc = d->chunkP;
    const int full = (int) (K1p / kc);
    const long rem = K1p - (long) full * kc;
    const int nOff = full + (rem > 0 ? 1 : 0);
    if (full > 0)
        hdk::gemm_bf16p(d->st, true, r, r, (int) kc, 2.0, d->Tp, (int) K1p, K1p * r, kc, d->Tp, (int) K1p, K1p * r, kc,
                        0.0, d->Gpart, hdk::BF16P_OUT_F64, r, 0, rr, full, 0);
    if (rem > 0)
        hdk::gemm_bf16p(d->st, true, r, r, (int) rem, 2.0, d->Tp + (long) full * kc, (int) K1p, K1p * r, 0,
                        d->Tp + (long) full * kc, (int) K1p, K1p * r, 0, 0.0, d->Gpart + (long) full * rr,
                        hdk::BF16P_OUT_F64, r, 0, 0, 1, 0);
    hdk::gemm_bf16p(d->st, true, r, r, ldn, 1.0, d->Tdgp, ldn, (long) ldn * r, 0, d->Tdgp, ldn, (long) ldn * r, 0, 0.0,
                    d->Gpart + (long) nOff * rr, hdk::BF16P_OUT_F64, r, 0, 0, 1, 0);
    hdk::gram_reduce(d->st, r, nOff + 1, d->Gpart, r, rr, d->c.denseCon, d->M, m);
}

// Mixed precision rank-one Gram: G = L^-1 F = (L^-T)^T F, K = G^T G, M = s s^T o K o K.
void r1_gram_mixed(hdsdp_dev *d) {
    const int n = d->n, r = d->c.nR1, m = d->m;
    hdk::d2f(d->st, (long) n * n, d->LiT[0], d->LiTf);
    hdk::gemm_bf16s_tn<float>(d->st, BF16_TERMS, false, n, r, n, 1.0, d->LiTf, n, 0, d->Ff, n, 0, 0.0, d->Gf, n, 0, 1);
    double *K = d->r1Identity ? d->M : d->Kr;
    const long ldk = d->r1Identity ? m : r;
    hdk::gemm_bf16s_tn<double>(d->st, BF16_TERMS, true, r, r, n, 1.0, d->Gf, n, 0, d->Gf, n, 0, 0.0, K, (int) ldk, 0, 1);
    hdk::r1_square(d->st, d->c, K, ldk, d->M, m);
}

}  // namespace

extern "C" {

const char *hdsdp_dev_backend_name(void) { return "hip-gfx950"; }

hdsdp_retcode hdsdp_dev_create(hdsdp_dev **pdev, int nRow, int nCol, int precision) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fprintf(stderr, "[hdsdp_amd] no HIP device: the MI355X library cannot run here\n");
        return HDSDP_RETCODE_FAILED;
    }
    hdsdp_dev *d = new hdsdp_dev();
    *pdev = d;
    d->m = nRow;
    d->n = nCol;
    d->precision = precision;
    HIP_OK(hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&d->la.s2, hipStreamNonBlocking));
    for (int k = 0; k < hdk::LA_EVENTS; ++k) HIP_OK(hipEventCreateWithFlags(&d->la.ev[k], hipEventDisableTiming));
    const size_t n = nCol, m = nRow, nn = n * n;
    for (int k = 0; k < 3; ++k) DCALL(dalloc(d, &d->S[k], nn));
    for (int k = 0; k < 2; ++k) {
        DCALL(dalloc(d, &d->L[k], nn));
        DCALL(dalloc(d, &d->Li[k], nn));
        DCALL(dalloc(d, &d->LiT[k], nn));
        d->facVer[k] = 0;
        d->liVer[k] = -1;
    }
    DCALL(dalloc(d, &d->W, nn));
    DCALL(dalloc(d, &d->W2, nn));
    DCALL(dalloc(d, &d->X, nn));
    DCALL(dalloc(d, &d->T1, std::max(nn, hdk::POTRF_WORK)));
    DCALL(dalloc(d, &d->T2, nn));
    DCALL(dalloc(d, &d->coef_d, m + 1));
    DCALL(dalloc(d, &d->vecA_d, m + 1));
    DCALL(dalloc(d, &d->vecB_d, m + 1));
    DCALL(dalloc(d, &d->sc_d, hdk::SC_SLOTS));
    DCALL(dalloc(d, &d->info_d, 1));
    DCALL(halloc(&d->hcoef, m + 1));
    DCALL(halloc(&d->hvecA, m + 1));
    DCALL(halloc(&d->hvecB, m + 1));
    DCALL(halloc(&d->hsc, hdk::SC_SLOTS));
    DCALL(halloc(&d->hvec_n, n));
    DCALL(halloc(&d->hY, 64));
    DCALL(halloc(&d->hinfo, 1));
    d->lzMax = 30;
    d->lzComputed = d->lzComputedP = 0;
    d->psdpReady = 0;
    d->lwarmP = d->Xs = d->dXp = d->XSX = d->Sold = d->LX = d->LXi = d->LXiT = nullptr;
    d->Hh.assign((size_t) (d->lzMax + 1) * (d->lzMax + 1), 0.0);
    DCALL(halloc(&d->hHist, 2 * (d->lzMax + 1)));
    DCALL(dalloc(d, &d->lzHist, 2 * (d->lzMax + 1)));
    DCALL(dalloc(d, &d->V, n * (d->lzMax + 1)));
    DCALL(dalloc(d, &d->Hd, (size_t) (d->lzMax + 1) * (d->lzMax + 1)));
    DCALL(dalloc(d, &d->lv, n));
    DCALL(dalloc(d, &d->lw, n));
    DCALL(dalloc(d, &d->lz1, n));
    DCALL(dalloc(d, &d->lz2, n));
    DCALL(dalloc(d, &d->lwarm, n));
    DCALL(dalloc(d, &d->ly, 64));
    DCALL(dalloc(d, &d->mv1, n));
    DCALL(dalloc(d, &d->mv2, n));
    DCALL(dalloc(d, &d->M, m * m));
    {
        long lowerMin = 16384;
        if (const char *v = getenv("HDSDP_SYMV_LOWER_MIN")) lowerMin = atol(v);
        d->symvW = nullptr;
        if (lowerMin > 0 && (long) m >= lowerMin) DCALL(dalloc(d, &d->symvW, hdk::symv_lower_work((int) m)));
    }
    DCALL(dalloc(d, &d->jac, m));
    DCALL(dalloc(d, &d->cr, m));
    DCALL(dalloc(d, &d->crn, m));
    DCALL(dalloc(d, &d->cp, m));
    DCALL(dalloc(d, &d->cz, m));
    DCALL(dalloc(d, &d->cMp, m));
    DCALL(dalloc(d, &d->cx, m));
    DCALL(dalloc(d, &d->crhs, m));
    DCALL(dalloc(d, &d->ctmp, m));
    DCALL(dalloc(d, &d->cb, m));
    DCALL(dalloc(d, &d->cxa, m));
    // CG defaults (linalg/hdsdp_linsolver.c:718-728)
    d->absTol = 1e-06;
    d->relTol = 1e-06;
    d->useJacobi = d->paramUseJacobi = 1;
    d->chol32Min = 16384;
    if (const char *c32 = getenv("HDSDP_MIXED_CHOL32_MIN")) d->chol32Min = atoi(c32);
    d->mixedFloor = 1e-13;
    if (const char *mf = getenv("HDSDP_MIXED_FLOOR")) d->mixedFloor = atof(mf);
    d->mVer = 0;
    d->cholVer = -1;
    d->maxIter = std::max(50, nRow / 20);
    d->nRestart = 20;
    hdk::fill(d->st, d->M, (long) m * m, 0.0);
    for (int k = 0; k < 3; ++k) hdk::fill(d->st, d->S[k], (long) nn, 0.0);
    d->prof = false;
    for (int k = 0; k < HDSDP_PROF_SLOTS; ++k) {
        d->evA[k] = d->evB[k] = nullptr;
        d->profMs[k] = d->profCnt[k] = 0.0;
    }
    return check_launch();
}

hdsdp_retcode hdsdp_dev_load(hdsdp_dev *d, const hdsdp_coeff_view *cv) {
    const int m = d->m, n = d->n;
    const long npk = (long) n * (n + 1) / 2;
    DevCoeffs &c = d->c;
    c.m = m;
    c.n = n;
    std::vector<int> ctype(m + 1), cBeg(m + 2, 0), dSlot(m + 1, -1), rSlot(m + 1, -1);
    std::vector<double> csign(m + 1, 0.0);
    std::vector<int> cRow, cCol, denseCon, rCon;
    std::vector<double> cVal;
    d->cType = cv[0].type;
    d->conDenseKind = 0;
    for (int k = 1; k <= m; ++k)
        if (cv[k].type == HDSDP_COEFF_DENSE || cv[k].type == HDSDP_COEFF_DSR1) d->conDenseKind = 1;
    d->hasZero = 0;
    for (int k = 0; k <= m; ++k) {
        ctype[k] = cv[k].type;
        csign[k] = cv[k].sign;
        if (k > 0 && cv[k].type == HDSDP_COEFF_ZERO) d->hasZero = 1;
        if (cv[k].type == HDSDP_COEFF_SPARSE) {
            for (int e = 0; e < cv[k].nnz; ++e) {
                cRow.push_back(cv[k].row[e]);
                cCol.push_back(cv[k].col[e]);
                cVal.push_back(cv[k].val[e]);
            }
        } else if (cv[k].type == HDSDP_COEFF_SPR1) {
            for (int e = 0; e < cv[k].nnz; ++e) {
                cRow.push_back(cv[k].row[e]);
                cCol.push_back(cv[k].row[e]);
                cVal.push_back(cv[k].val[e]);
            }
        }
        cBeg[k + 1] = (int) cRow.size();
    }
    // dense coefficients: constraints in order, C last (assembly order of the reference)
    for (int k = 1; k <= m; ++k)
        if (ctype[k] == HDSDP_COEFF_DENSE) denseCon.push_back(k);
    if (ctype[0] == HDSDP_COEFF_DENSE) denseCon.push_back(0);
    for (size_t q = 0; q < denseCon.size(); ++q) dSlot[denseCon[q]] = (int) q;
    for (int k = 1; k <= m; ++k)
        if (ctype[k] == HDSDP_COEFF_DSR1) rCon.push_back(k);
    for (size_t q = 0; q < rCon.size(); ++q) rSlot[rCon[q]] = (int) q;
    c.cDSR1 = (ctype[0] == HDSDP_COEFF_DSR1);
    if (c.cDSR1) rSlot[0] = (int) rCon.size();
    c.nDense = (int) denseCon.size();
    c.nR1 = (int) rCon.size();
    c.cR1Sign = cv[0].sign;

    // scatter map over SPARSE / SPR1 entries of the lower triangle, C last
    struct Ent { long pos; int con; int order; double val; };
    std::vector<Ent> ents;
    auto add_entries = [&](int k, int order) {
        if (ctype[k] == HDSDP_COEFF_SPARSE) {
            for (int e = cBeg[k]; e < cBeg[k + 1]; ++e)
                ents.push_back({cRow[e] + (long) cCol[e] * n, k, order, cVal[e]});
        } else if (ctype[k] == HDSDP_COEFF_SPR1) {
            for (int p = cBeg[k]; p < cBeg[k + 1]; ++p)
                for (int q = cBeg[k]; q <= p; ++q)
                    ents.push_back({cRow[p] + (long) cRow[q] * n, k, order, csign[k] * cVal[p] * cVal[q]});
        }
    };
    for (int k = 1; k <= m; ++k) add_entries(k, k);
    add_entries(0, m + 1);
    std::stable_sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
        return a.pos != b.pos ? a.pos < b.pos : a.order < b.order;
    });
    std::vector<long> posIdx;
    std::vector<int> posBeg, entCon;
    std::vector<double> entVal;
    for (size_t e = 0; e < ents.size(); ++e) {
        if (e == 0 || ents[e].pos != ents[e - 1].pos) {
            posIdx.push_back(ents[e].pos);
            posBeg.push_back((int) e);
        }
        entCon.push_back(ents[e].con);
        entVal.push_back(ents[e].val);
    }
    posBeg.push_back((int) ents.size());
    c.nPos = (int) posIdx.size();

    // small-constraint table for the pair kernel
    const int K = hdk::SMALL_K;
    std::vector<int> smKind(m + 1, hdk::SM_NONE), smNnz(m + 1, 0), smRow((m + 1) * K, 0), smCol((m + 1) * K, 0);
    std::vector<double> smVal((m + 1) * K, 0.0);
    d->heavy.clear();
    for (int k = 1; k <= m; ++k) {
        int t = ctype[k], nz = cBeg[k + 1] - cBeg[k];
        if (t == HDSDP_COEFF_ZERO) continue;
        if ((t == HDSDP_COEFF_SPARSE || t == HDSDP_COEFF_SPR1) && nz <= K) {
            smKind[k] = (t == HDSDP_COEFF_SPARSE) ? hdk::SM_SP : hdk::SM_R1;
            smNnz[k] = nz;
            for (int e = 0; e < nz; ++e) {
                smRow[k * K + e] = cRow[cBeg[k] + e];
                smCol[k * K + e] = cCol[cBeg[k] + e];
                smVal[k * K + e] = cVal[cBeg[k] + e];
            }
        } else {
            smKind[k] = hdk::SM_HEAVY;
            d->heavy.push_back(k);
        }
    }

    // single-entry small constraints (edges, diagonal entries): the LDS pair kernel
    {
        bool single = true;
        int cnt = 0;
        for (int k = 1; k <= m; ++k) {
            if (smKind[k] == hdk::SM_R1 || (smKind[k] == hdk::SM_SP && smNnz[k] != 1)) single = false;
            if (smKind[k] == hdk::SM_SP) ++cnt;
        }
        c.pairSingle = (single && cnt > 0 && n <= hdk::PAIR1_MAXN) ? 1 : 0;
    }

    // Gram paths: one kind of constraint throughout (heavy sparse rows allowed)
    int nSmall = 0, nR1k = 0, nDnk = 0;
    for (int k = 1; k <= m; ++k) {
        if (smKind[k] == hdk::SM_SP || smKind[k] == hdk::SM_R1) ++nSmall;
        if (ctype[k] == HDSDP_COEFF_DSR1) ++nR1k;
        if (ctype[k] == HDSDP_COEFF_DENSE) ++nDnk;
    }
    d->nDn = nDnk;
    d->r1Gram = (nR1k > 0 && nSmall == 0 && nDnk == 0);
    d->dnGram = (nDnk > 0 && nSmall == 0 && nR1k == 0);
    d->r1Identity = d->r1Gram && nR1k == m;
    if (d->r1Gram || d->dnGram) {
        std::vector<int> keep;
        for (int h : d->heavy) {
            if (d->r1Gram && ctype[h] == HDSDP_COEFF_DSR1) continue;
            if (d->dnGram && ctype[h] == HDSDP_COEFF_DENSE) continue;
            keep.push_back(h);
        }
        d->heavy.swap(keep);
    }

    auto up_i = [&](int **dst, const std::vector<int> &v) -> hdsdp_retcode {
        DCALL(dalloc(d, dst, v.size()));
        if (!v.empty()) HIP_OK(hipMemcpy(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
        return HDSDP_RETCODE_OK;
    };
    auto up_d = [&](double **dst, const std::vector<double> &v) -> hdsdp_retcode {
        DCALL(dalloc(d, dst, v.size()));
        if (!v.empty()) HIP_OK(hipMemcpy(*dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        return HDSDP_RETCODE_OK;
    };
    DCALL(up_i(&c.ctype, ctype));
    DCALL(up_d(&c.csign, csign));
    DCALL(up_i(&c.cBeg, cBeg));
    DCALL(up_i(&c.cRow, cRow));
    DCALL(up_i(&c.cCol, cCol));
    DCALL(up_d(&c.cVal, cVal));
    DCALL(up_i(&c.denseCon, denseCon));
    DCALL(up_i(&c.dSlot, dSlot));
    DCALL(up_i(&c.rCon, rCon));
    DCALL(up_i(&c.rSlot, rSlot));
    // coefficient columns of `len` doubles gathered on the host into 64 MB chunks, one
    // copy per chunk (one copy per constraint cost ~25 us each: 0.5 s for 20000 factors)
    auto gather_up = [&](double *dst, int count, size_t len, const int *con) -> hdsdp_retcode {
        if (count <= 0) return HDSDP_RETCODE_OK;
        const int per = (int) std::max<size_t>(1, ((size_t) 64 << 20) / (len * sizeof(double)));
        std::vector<double> buf((size_t) std::min(per, count) * len);
        for (int q0 = 0; q0 < count; q0 += per) {
            const int cnt = std::min(per, count - q0);
            for (int q = 0; q < cnt; ++q)
                memcpy(buf.data() + (size_t) q * len, cv[con ? con[q0 + q] : 0].val, len * sizeof(double));
            HIP_OK(hipMemcpy(dst + (size_t) q0 * len, buf.data(), (size_t) cnt * len * sizeof(double),
                             hipMemcpyHostToDevice));
        }
        return HDSDP_RETCODE_OK;
    };
    DCALL(dalloc(d, &c.dVal, (size_t) c.nDense * npk));
    DCALL(gather_up(c.dVal, c.nDense, npk, denseCon.data()));
    int nF = c.nR1 + (c.cDSR1 ? 1 : 0);
    DCALL(dalloc(d, &c.rF, (size_t) n * nF));
    DCALL(gather_up(c.rF, c.nR1, n, rCon.data()));
    if (c.cDSR1) DCALL(gather_up(c.rF + (size_t) c.nR1 * n, 1, n, nullptr));
    c.cR1F = c.rF + (size_t) c.nR1 * n;
    DCALL(dalloc(d, &c.r1w, c.nR1));
    c.packW = nullptr;
    c.trPart = nullptr;
    if (c.nDense > 0) {
        DCALL(dalloc(d, &c.packW, 2 * npk));
        DCALL(dalloc(d, &c.trPart, (size_t) 2 * hdk::TR_SPLIT * c.nDense));
    }
    {
        DCALL(dalloc(d, &c.posIdx, posIdx.size()));
        if (!posIdx.empty())
            HIP_OK(hipMemcpy(c.posIdx, posIdx.data(), posIdx.size() * sizeof(long), hipMemcpyHostToDevice));
    }
    DCALL(up_i(&c.posBeg, posBeg));
    DCALL(up_i(&c.entCon, entCon));
    DCALL(up_d(&c.entVal, entVal));
    DCALL(up_i(&c.smKind, smKind));
    DCALL(up_i(&c.smNnz, smNnz));
    DCALL(up_i(&c.smRow, smRow));
    DCALL(up_i(&c.smCol, smCol));
    DCALL(up_d(&c.smVal, smVal));
    size_t workSize = std::max((size_t) n * n, (size_t) n * std::max(c.nR1, 1));
    DCALL(dalloc(d, &d->work, workSize));
    d->splitW = nullptr;
    d->splitLen = 0;
    if (c.nR1 > 0) {
        const int parts = hdk::splitk_parts(n, n, c.nR1, true);
        if (parts > 1) {
            d->splitLen = (size_t) parts * n * n;
            DCALL(dalloc(d, &d->splitW, d->splitLen));
        }
    }
    d->H1 = d->Kr = d->Tall = d->Tdg = d->Dbuf = d->Ubuf = d->Gpart = nullptr;
    if (c.nR1 > 0 || c.cDSR1) DCALL(dalloc(d, &d->H1, (size_t) n * (c.nR1 + 1)));
    if (d->r1Gram && !d->r1Identity) DCALL(dalloc(d, &d->Kr, (size_t) c.nR1 * c.nR1));
    d->LiTf = d->Ff = d->Gf = d->Tf = d->Tdgf = d->Dbuff = d->Vf = nullptr;
    d->LiTp = d->Dp = d->Vp = d->Tp = d->Tdgp = d->Dall = nullptr;
    d->K1p = d->chunkP = 0;
    d->Wm = d->Lim = d->mDiag0 = d->mDiagCur = d->xCoef = nullptr;
    // bf16-split Gram blocks need the matrix-free FP64 product in the refinement.  For dense
    // constraints it costs two passes over the packed coefficients and two n^3 products,
    // far below the Gram build; for rank-one constraints it costs about 4 n^2 r + 4 n^3
    // flops against n r^2 for the FP64 Gram, so they use it only when r >= 32 n.
    // HDSDP_GRAM_BF16=1 (or 0) forces the choice where a Gram path exists.
    bool mixedGram = d->precision == HDSDP_PRECISION_MIXED &&
                     (d->dnGram || (d->r1Gram && (long) c.nR1 >= 32L * n));
    if (const char *force = getenv("HDSDP_GRAM_BF16"))
        mixedGram = d->precision == HDSDP_PRECISION_MIXED && (d->r1Gram || d->dnGram) && atoi(force) != 0;
    d->schurMixed = mixedGram ? 1 : 0;
    d->mGramPending = 0;
    // bf16 planes for the dense mixed Gram (split once by the producing kernels; 2.2x the
    // split-in-the-loop products on the n = 1000 Gram); HDSDP_BF16_PLANES=0 selects the
    // products that split their FP32 operands in the loop
    d->planes = 1;
    if (const char *pl = getenv("HDSDP_BF16_PLANES")) d->planes = atoi(pl) != 0;
    if (mixedGram) {
        if (d->r1Gram || !d->planes) DCALL(dalloc(d, &d->LiTf, (size_t) n * n));
        DCALL(dalloc(d, &d->Wm, (size_t) n * n));
        DCALL(dalloc(d, &d->Lim, (size_t) n * n));
        DCALL(dalloc(d, &d->mDiag0, (size_t) m));
        DCALL(dalloc(d, &d->mDiagCur, (size_t) m));
        DCALL(dalloc(d, &d->xCoef, (size_t) m + 1));
    }
    if (mixedGram && d->r1Gram) {
        DCALL(dalloc(d, &d->Ff, (size_t) n * c.nR1));
        DCALL(dalloc(d, &d->Gf, (size_t) n * c.nR1));
        hdk::d2f(d->st, (long) n * c.nR1, c.rF, d->Ff);
    }
    if (d->dnGram) {
        const size_t nn = (size_t) n * n;
        const int r = d->nDn;
        d->gramBatch = (int) std::max<size_t>(1, std::min<size_t>(r, (size_t) (512u << 20) / (nn * sizeof(double))));
        // split-K parts of the Gram product: enough 128-wide lower tiles for ~8 rounds over
        // the CUs, and a part count that divides n^2 where one exists, so that every part
        // goes into one batched launch (a remainder part runs alone, on too few tiles)
        long tiles = (long) ((r + 127) / 128) * ((r + 127) / 128 + 1) / 2;
        long want = std::min<long>(16, std::max<long>(1, (2048 + tiles - 1) / tiles));
        long K = (long) n * (n - 1) / 2, parts = want;
        for (long p = want; p <= 2 * want; ++p)
            if (K % p == 0 && K / p >= 256) {
                parts = p;
                break;
            }
        d->gramChunk = std::max<long>(1, (K % parts == 0) ? K / parts : std::max<long>(256, (K + parts - 1) / parts));
        d->gramParts = (int) ((K + d->gramChunk - 1) / d->gramChunk);
        const size_t K1 = std::max<size_t>(1, (size_t) n * (n - 1) / 2);
        if (mixedGram && d->planes) {
            // planes: operands zero outside their true extent (k padding of the products)
            const size_t ldn = (size_t) hdk::plane_ld(n), nnp = ldn * n;
            d->K1p = (long) (K1 + 31) / 32 * 32;
            const long tilesP = tiles, wantP = std::min<long>(16, std::max<long>(1, (2048 + tilesP - 1) / tilesP));
            d->chunkP = std::max<long>(32, (d->K1p / wantP + 31) / 32 * 32);
            // Gpart holds the parts of either build (the FP64 rebuild uses gramChunk)
            d->gramParts = std::max<int>(d->gramParts, (int) ((d->K1p + d->chunkP - 1) / d->chunkP));
            DCALL(dalloc_zero(d, &d->LiTp, 3 * nnp));
            // The constraints never change: their planes are split once and kept, 6 r n^2
            // bytes (12 GB for n = 1000, r = 2000), while that stays under
            // HDSDP_PLANES_RESIDENT_GB (default 96); otherwise each build splits a batch.
            double residentGb = 96.0;
            if (const char *rg = getenv("HDSDP_PLANES_RESIDENT_GB")) residentGb = atof(rg);
            const size_t allCount = 3 * nnp * (size_t) r;
            if ((double) allCount * sizeof(__bf16) <= residentGb * 1e9) {
                void *q = nullptr;
                if (hipMalloc(&q, allCount * sizeof(__bf16)) == hipSuccess) {
                    d->allocs.push_back(q);
                    d->Dall = (__bf16 *) q;
                    HIP_OK(hipMemsetAsync(d->Dall, 0, allCount * sizeof(__bf16), d->st));
                    hdk::unpack_dense_split(d->st, d->c, 0, r, d->Dall, (int) ldn, (long) nnp * r, (long) nnp);
                } else {
                    (void) hipGetLastError();  // the batch path below needs far less memory
                }
            }
            if (!d->Dall) DCALL(dalloc_zero(d, &d->Dp, 3 * nnp * d->gramBatch));
            DCALL(dalloc_zero(d, &d->Vp, 3 * nnp * d->gramBatch));
            DCALL(dalloc_zero(d, &d->Tp, 3 * (size_t) d->K1p * r));
            DCALL(dalloc_zero(d, &d->Tdgp, 3 * ldn * r));
        } else if (mixedGram) {
            DCALL(dalloc(d, &d->Tf, K1 * r));
            DCALL(dalloc(d, &d->Tdgf, (size_t) n * r));
            DCALL(dalloc(d, &d->Dbuff, nn * d->gramBatch));
            DCALL(dalloc(d, &d->Vf, nn * d->gramBatch));
        } else {
            DCALL(dalloc(d, &d->Tall, K1 * r));
            DCALL(dalloc(d, &d->Tdg, (size_t) n * r));
            DCALL(dalloc(d, &d->Dbuf, nn * d->gramBatch));
            DCALL(dalloc(d, &d->Ubuf, nn * d->gramBatch));
        }
        // split-K parts of the off-diagonal Gram and one for the diagonal Gram
        DCALL(dalloc(d, &d->Gpart, (size_t) r * r * (d->gramParts + 1)));
    }
    return check_launch();
}

hdsdp_retcode hdsdp_dev_scale_obj(hdsdp_dev *d, double s) {
    // rescale C where it is stored (sparse entries in the CSR and the scatter map,
    // dense packed block, or the rank-one sign)
    if (s == 1.0) return HDSDP_RETCODE_OK;
    DevCoeffs &c = d->c;
    if (d->cType == HDSDP_COEFF_SPR1 || d->cType == HDSDP_COEFF_DSR1) {
        double sg = 0.0;
        HIP_OK(hipMemcpy(&sg, c.csign, sizeof(double), hipMemcpyDeviceToHost));
        sg *= s;
        HIP_OK(hipMemcpy(c.csign, &sg, sizeof(double), hipMemcpyHostToDevice));
        c.cR1Sign *= s;
        if (d->cType == HDSDP_COEFF_DSR1) return HDSDP_RETCODE_OK;
    }
    if (d->cType == HDSDP_COEFF_DENSE) {
        long npk = (long) d->n * (d->n + 1) / 2;
        std::vector<int> dSlot(1);
        HIP_OK(hipMemcpy(dSlot.data(), c.dSlot, sizeof(int), hipMemcpyDeviceToHost));
        std::vector<double> buf(npk);
        double *p = c.dVal + (long) dSlot[0] * npk;
        HIP_OK(hipMemcpy(buf.data(), p, npk * sizeof(double), hipMemcpyDeviceToHost));
        for (auto &v : buf) v *= s;
        HIP_OK(hipMemcpy(p, buf.data(), npk * sizeof(double), hipMemcpyHostToDevice));
        return HDSDP_RETCODE_OK;
    }
    // SPARSE or SPR1 C: its entries are the last of each scatter group and the first CSR segment
    std::vector<int> cBeg(2);
    HIP_OK(hipMemcpy(cBeg.data(), c.cBeg, 2 * sizeof(int), hipMemcpyDeviceToHost));
    if (d->cType == HDSDP_COEFF_SPARSE && cBeg[1] > 0) {
        std::vector<double> v(cBeg[1]);
        HIP_OK(hipMemcpy(v.data(), c.cVal, v.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (auto &x : v) x *= s;
        HIP_OK(hipMemcpy(c.cVal, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (c.nPos > 0) {
        std::vector<int> posBeg(c.nPos + 1);
        HIP_OK(hipMemcpy(posBeg.data(), c.posBeg, posBeg.size() * sizeof(int), hipMemcpyDeviceToHost));
        int nEnt = posBeg[c.nPos];
        std::vector<int> con(nEnt);
        std::vector<double> val(nEnt);
        HIP_OK(hipMemcpy(con.data(), c.entCon, nEnt * sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(val.data(), c.entVal, nEnt * sizeof(double), hipMemcpyDeviceToHost));
        for (int e = 0; e < nEnt; ++e)
            if (con[e] == 0) val[e] *= s;
        HIP_OK(hipMemcpy(c.entVal, val.data(), nEnt * sizeof(double), hipMemcpyHostToDevice));
    }
    return HDSDP_RETCODE_OK;
}

void hdsdp_dev_destroy(hdsdp_dev **pdev) {
    if (!pdev || !*pdev) return;
    hdsdp_dev *d = *pdev;
    if (d->st) hipStreamSynchronize(d->st);
    for (void *p : d->allocs) hipFree(p);
    double *hp[] = {d->hcoef, d->hvecA, d->hvecB, d->hsc, d->hvec_n, d->hY, d->hHist};
    for (double *p : hp)
        if (p) hipHostFree(p);
    if (d->hinfo) hipHostFree(d->hinfo);
    for (int k = 0; k < HDSDP_PROF_SLOTS; ++k) {
        if (d->evA[k]) hipEventDestroy(d->evA[k]);
        if (d->evB[k]) hipEventDestroy(d->evB[k]);
    }
    for (int k = 0; k < hdk::LA_EVENTS; ++k)
        if (d->la.ev[k]) hipEventDestroy(d->la.ev[k]);
    if (d->la.s2) {
        hipStreamSynchronize(d->la.s2);
        hipStreamDestroy(d->la.s2);
    }
    if (d->st) hipStreamDestroy(d->st);
    delete d;
    *pdev = nullptr;
}

hdsdp_retcode hdsdp_dev_stats(hdsdp_dev *d, double *stats, int nStats) {
    for (int k = 0; k < nStats && k < 7; ++k) stats[k] = d->stats[k];
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_set_profiling(hdsdp_dev *d, int enable) {
    if (enable && !d->evA[0]) {
        for (int k = 0; k < HDSDP_PROF_SLOTS; ++k) {
            HIP_OK(hipEventCreate(&d->evA[k]));
            HIP_OK(hipEventCreate(&d->evB[k]));
        }
    }
    d->prof = enable != 0;
    return HDSDP_RETCODE_OK;
}

int hdsdp_dev_get_profile(hdsdp_dev *d, double *ms, double *counts, int nSlots) {
    int n = std::min(nSlots, (int) HDSDP_PROF_SLOTS);
    for (int k = 0; k < n; ++k) {
        if (ms) ms[k] = d->profMs[k];
        if (counts) counts[k] = d->profCnt[k];
    }
    return n;
}

// ---- dual buffers -----------------------------------------------------------------

hdsdp_retcode hdsdp_dev_update_buffer(hdsdp_dev *d, double cCoef, double aScal, const double *aCoef,
                                      double dEyeCoef, int which) {
    const int m = d->m;
    if (d->stagingBusy) DCALL(sync(d));
    d->hcoef[0] = cCoef;
    for (int i = 0; i < m; ++i) d->hcoef[i + 1] = aScal * aCoef[i];
    HIP_OK(hipMemcpyAsync(d->coef_d, d->hcoef, sizeof(double) * (m + 1), hipMemcpyHostToDevice, d->st));
    d->stagingBusy = true;
    prof_begin(d, HDSDP_PROF_BUFFER);
    hdk::assemble_buffer(d->st, d->c, d->coef_d, cCoef, d->S[which], dEyeCoef, d->work, d->splitW, d->splitLen);
    DCALL(check_launch());
    return prof_end(d, HDSDP_PROF_BUFFER);
}

hdsdp_retcode hdsdp_dev_factorize(hdsdp_dev *d, int which, int *isPsd) {
    const int n = d->n, f = fidx(which);
    prof_begin(d, HDSDP_PROF_SFACTOR);
    HIP_OK(hipMemcpyAsync(d->L[f], d->S[f], sizeof(double) * (size_t) n * n, hipMemcpyDeviceToDevice, d->st));
    if (!hdk::potrf<double>(d->st, n, d->L[f], n, d->T1, std::max((size_t) n * n, hdk::POTRF_WORK), d->info_d,
                            &d->la))
        return HDSDP_RETCODE_FAILED;
    HIP_OK(hipMemcpyAsync(d->hinfo, d->info_d, sizeof(int), hipMemcpyDeviceToHost, d->st));
    DCALL(prof_end(d, HDSDP_PROF_SFACTOR));
    DCALL(sync(d));
    d->facVer[f] += 1;
    d->stats[1] += 1;
    *isPsd = (*d->hinfo == 0);
    return check_launch();
}

hdsdp_retcode hdsdp_dev_axpy_check(hdsdp_dev *d, double dStep, int which, int *isPsd) {
    const int f = fidx(which);
    long nn = (long) d->n * d->n;
    hdk::axpy_buffer(d->st, nn, dStep, d->S[0], d->S[2], d->S[f]);
    return hdsdp_dev_factorize(d, f, isPsd);
}

hdsdp_retcode hdsdp_dev_logdet(hdsdp_dev *d, int which, double *logdet) {
    const int f = fidx(which);
    hdk::logdet_diag(d->st, d->n, d->L[f], d->n, d->sc_d + hdk::SC_LOGDET);
    DCALL(read_scalars(d));
    *logdet = d->hsc[hdk::SC_LOGDET];
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_ratio_test(hdsdp_dev *d, int which, double *maxStep) {
    const int f = fidx(which);
    if (d->n == 1) {
        double s = 0.0, ds = 0.0;
        HIP_OK(hipMemcpyAsync(d->hsc, d->S[f], sizeof(double), hipMemcpyDeviceToHost, d->st));
        HIP_OK(hipMemcpyAsync(d->hsc + 1, d->S[2], sizeof(double), hipMemcpyDeviceToHost, d->st));
        DCALL(sync(d));
        s = d->hsc[0];
        ds = d->hsc[1];
        *maxStep = (ds > 0.0) ? HDSDP_INFINITY : (-s / ds);
        return HDSDP_RETCODE_OK;
    }
    prof_begin(d, HDSDP_PROF_LANCZOS);
    DCALL(ensure_linv(d, f));
    hdk::mirror_lower<double>(d->st, d->n, d->S[2], d->n);
    const LzOp op{d->Li[f], d->LiT[f], d->S[2], &d->lzComputed, d->lwarm};
    DCALL(lanczos(d, op, maxStep));
    return prof_end(d, HDSDP_PROF_LANCZOS);
}

// ---- Schur complement ---------------------------------------------------------------

hdsdp_retcode hdsdp_dev_schur(hdsdp_dev *d, int typeKKT, double Rd, double *asinv, double *asinvrd,
                              double *traceSinv) {
    const int m = d->m, n = d->n;
    // HOMOGENEOUS (2) builds what INFEASIBLE builds; its extra terms come from
    // hdsdp_dev_hsd on the S^-1 formed here
    if (typeKKT == 3) {
        // PRIMAL: strategy M5 with W := X for every column (interface/hdsdp_conic_sdp.c:1741-1745,
        // 1813-1820); the reference's M5 takes no DENSE or DSR1 coefficient
        // (linalg/hdsdp_sdpdata.c:1961-1965, 1995-1999)
        if (d->conDenseKind || !d->psdpReady) return HDSDP_RETCODE_FAILED;
        prof_begin(d, HDSDP_PROF_SCHUR);
        HIP_OK(hipMemcpyAsync(d->W, d->X, sizeof(double) * (size_t) n * n, hipMemcpyDeviceToDevice, d->st));
    } else {
        if (typeKKT != 1) prof_begin(d, HDSDP_PROF_SCHUR);
        DCALL(ensure_linv(d, 0));
        // W = S^-1 = L^-T L^-1
        hdk::gemm<double>(d->st, true, false, false, n, n, n, 1.0, d->Li[0], n, 0, d->Li[0], n, 0, 0.0, d->W, n, 0,
                          1);
    }
    if (Rd != 0.0)
        hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->W, n, 0, d->W, n, 0, 0.0, d->W2, n, 0, 1);
    hdk::traces(d->st, d->c, 1, m, d->W, Rd != 0.0 ? d->W2 : nullptr, Rd, d->vecA_d, d->vecB_d, false);
    hdk::r1_quadforms(d->st, d->c, d->W, false, 1, m, Rd, d->vecA_d, Rd != 0.0 ? d->vecB_d : nullptr, d->H1);
    *traceSinv = 0.0;
    if (typeKKT != 1 && Rd != 0.0) hdk::trace_diag(d->st, n, d->W, d->sc_d + hdk::SC_AUX);
    if (typeKKT != 1) {
        if (d->hasZero) hdk::fill(d->st, d->M, (long) m * m, 0.0);
        prof_begin(d, HDSDP_PROF_PAIRS);
        hdk::schur_pairs(d->st, d->c, d->W, d->M, m);
        DCALL(prof_end(d, HDSDP_PROF_PAIRS));
        // once the system has switched to the direct LU solve, M is factored in FP64 as
        // built: no split products to rebuild
        const bool mixed = d->schurMixed != 0 && !d->useLU;
        d->mGramPending = mixed ? 1 : 0;
        d->mVer += 1;
        if (mixed) {
            // keep the factor this M is built from, for an FP64 rebuild
            HIP_OK(hipMemcpyAsync(d->Wm, d->W, sizeof(double) * (size_t) n * n, hipMemcpyDeviceToDevice, d->st));
            HIP_OK(hipMemcpyAsync(d->Lim, d->Li[0], sizeof(double) * (size_t) n * n, hipMemcpyDeviceToDevice, d->st));
        }
        if (d->r1Gram || d->dnGram) prof_begin(d, HDSDP_PROF_GRAM);
        if (mixed) prof_begin(d, HDSDP_PROF_GRAM16);
        if (d->r1Gram) {
            if (mixed) r1_gram_mixed(d);
            else if (d->r1Identity) hdk::r1_schur(d->st, d->c, d->H1, d->M, m, d->M, m);
            else hdk::r1_schur(d->st, d->c, d->H1, d->Kr, d->c.nR1, d->M, m);
        }
        if (d->dnGram) {
            if (mixed) dense_gram_mixed(d);
            else dense_gram(d, d->Li[0]);
        }
        if (mixed) DCALL(prof_end(d, HDSDP_PROF_GRAM16));
        if (d->r1Gram || d->dnGram) DCALL(prof_end(d, HDSDP_PROF_GRAM));
        for (int h : d->heavy) {
            // B_h = W A_h W, then M(h, k) = <A_k, B_h>
            hdk::densify(d->st, d->c, h, d->T1);
            hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->W, n, 0, d->T1, n, 0, 0.0, d->T2, n, 0, 1);
            hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->T2, n, 0, d->W, n, 0, 0.0, d->T1, n, 0, 1);
            hdk::traces(d->st, d->c, 1, m, d->T1, nullptr, 0.0, d->ctmp, nullptr, true);
            hdk::scatter_row(d->st, m, h - 1, d->ctmp, d->M, m);
        }
        if (mixed) hdk::get_diag(d->st, m, d->M, m, d->mDiag0);
        d->stats[0] += 1;
        DCALL(prof_end(d, HDSDP_PROF_SCHUR));
    }
    DCALL(check_launch());
    DCALL(download(d, asinv, d->vecA_d, m, d->hvecA));
    if (Rd != 0.0) {
        DCALL(download(d, asinvrd, d->vecB_d, m, d->hvecB));
    } else {
        memset(asinvrd, 0, sizeof(double) * m);
    }
    if (typeKKT != 1 && Rd != 0.0) {
        DCALL(read_scalars(d));
        *traceSinv = d->hsc[hdk::SC_AUX];
    }
    d->wFromDual = (typeKKT != 3);
    return HDSDP_RETCODE_OK;
}

// The self-dual embedding's terms (sdpDenseConeIGetHSDComponents,
// interface/hdsdp_conic_sdp.c:983-1029, and the dASinvCSinvVec entries that strategies
// M1-M5 add for KKT_TYPE_HOMOGENEOUS, 737-760 and alike) from the W = S^-1 of the last
// Schur build: with B = W C W,
//   dASinvCSinv[i] = <A_i, B>,  dCSinv = <C, W>,  dCSinvCSinv = <C, B>,
//   dCSinvRdSinv = Rd tr(B).
hdsdp_retcode hdsdp_dev_hsd(hdsdp_dev *d, double Rd, double *aSinvCSinv, double *cSinv, double *cSinvCSinv,
                            double *cSinvRdSinv) {
    const int m = d->m, n = d->n;
    if (!d->wFromDual) return HDSDP_RETCODE_FAILED;
    if (!d->hsdCoef) DCALL(dalloc(d, &d->hsdCoef, (size_t) m + 1));
    // C alone: the coefficient vector e_0
    if (d->stagingBusy) DCALL(sync(d));
    memset(d->hcoef, 0, sizeof(double) * (m + 1));
    d->hcoef[0] = 1.0;
    HIP_OK(hipMemcpyAsync(d->hsdCoef, d->hcoef, sizeof(double) * (m + 1), hipMemcpyHostToDevice, d->st));
    d->stagingBusy = true;
    hdk::assemble_buffer(d->st, d->c, d->hsdCoef, 1.0, d->T1, 0.0, d->work, d->splitW, d->splitLen);
    hdk::mirror_lower<double>(d->st, n, d->T1, n);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->W, n, 0, d->T1, n, 0, 0.0, d->T2, n, 0, 1);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->T2, n, 0, d->W, n, 0, 0.0, d->T1, n, 0, 1);
    // <A_k, B> for k = 0 (C) .. m, and <C, W>
    hdk::traces(d->st, d->c, 0, m + 1, d->T1, nullptr, 0.0, d->vecA_d, nullptr, false);
    hdk::r1_quadforms(d->st, d->c, d->T1, true, 0, m + 1, 0.0, d->vecA_d, nullptr, d->H1);
    hdk::traces(d->st, d->c, 0, 1, d->W, nullptr, 0.0, d->vecB_d, nullptr, false);
    hdk::r1_quadforms(d->st, d->c, d->W, true, 0, 1, 0.0, d->vecB_d, nullptr, d->H1);
    hdk::trace_diag(d->st, n, d->T1, d->sc_d + hdk::SC_AUX);
    DCALL(check_launch());
    DCALL(download(d, d->hvecB, d->vecA_d, m + 1, d->hvecA));
    memcpy(aSinvCSinv, d->hvecB + 1, sizeof(double) * m);
    *cSinvCSinv = d->hvecB[0];
    DCALL(download(d, d->hvecB, d->vecB_d, 1, d->hvecA));
    *cSinv = d->hvecB[0];
    DCALL(read_scalars(d));
    *cSinvRdSinv = Rd * d->hsc[hdk::SC_AUX];
    return HDSDP_RETCODE_OK;
}

// ---- Schur solve ----------------------------------------------------------------------

hdsdp_retcode hdsdp_dev_kkt_params(hdsdp_dev *d, double relTol, double absTol, int maxIter, int nRestart) {
    if (absTol > 0) d->absTol = absTol;
    if (relTol > 0) d->relTol = relTol;
    if (maxIter > 0) d->maxIter = maxIter;
    if (nRestart > 0) d->nRestart = nRestart;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_kkt_add_diag(hdsdp_dev *d, const double *diag) {
    d->mVer += 1;
    DCALL(upload(d, d->ctmp, diag, d->m, d->hvecA));
    hdk::add_diag(d->st, d->m, d->M, d->m, d->ctmp, 0.0);
    return check_launch();
}

hdsdp_retcode hdsdp_dev_kkt_regularize(hdsdp_dev *d, double reg) {
    d->mVer += 1;
    hdk::regularize(d->st, d->m, d->M, d->m, reg);
    return check_launch();
}

// LU factorisation of the FP64 M into LUm (HFpLinsysSwitchToIndefinite,
// linalg/hdsdp_linsolver.c:1236-1266: from the first call on, the system is solved
// directly).  The reference factors M by LDL^T with Bunch-Kaufman pivoting (dsytrf); LU
// with partial pivoting is as stable on a symmetric indefinite matrix, runs its trailing
// updates on the FP64 GEMM, and solves the same system.
hdsdp_retcode lu_factor(hdsdp_dev *d) {
    const int m = d->m;
    if (!d->LUm) {
        DCALL(dalloc(d, &d->LUm, (size_t) m * m));
        DCALL(dalloc(d, &d->luPiv, (size_t) m));
        DCALL(dalloc(d, &d->luPerm, (size_t) m));
        DCALL(dalloc(d, &d->luInfo, 1));
        d->hpiv.assign((size_t) m + 1, 0);
    }
    d->useLU = true;
    if (d->mGramPending) DCALL(rebuild_gram_fp64(d));
    hdk::mirror_lower<double>(d->st, m, d->M, m);
    HIP_OK(hipMemcpyAsync(d->LUm, d->M, sizeof(double) * (size_t) m * m, hipMemcpyDeviceToDevice, d->st));
    HIP_OK(hipMemsetAsync(d->luInfo, 0, sizeof(int), d->st));
    hdk::lu_factor(d->st, m, d->LUm, m, d->luPiv, d->luInfo);
    DCALL(check_launch());
    d->stats[4] += 1;
    HIP_OK(hipMemcpyAsync(d->hpiv.data(), d->luPiv, sizeof(int) * m, hipMemcpyDeviceToHost, d->st));
    HIP_OK(hipMemcpyAsync(d->hpiv.data() + m, d->luInfo, sizeof(int), hipMemcpyDeviceToHost, d->st));
    DCALL(sync(d));
    if (d->hpiv[m] != 0) return HDSDP_RETCODE_FAILED;
    // compose the interchanges: (P b)[i] = b[perm[i]]
    std::vector<int> perm(m);
    for (int i = 0; i < m; ++i) perm[i] = i;
    for (int c = 0; c < m; ++c) {
        const int p = d->hpiv[c];
        if (p < c || p >= m) return HDSDP_RETCODE_FAILED;
        std::swap(perm[c], perm[p]);
    }
    HIP_OK(hipMemcpyAsync(d->luPerm, perm.data(), sizeof(int) * m, hipMemcpyHostToDevice, d->st));
    return sync(d);
}

// HFpLinsysNumeric (linalg/hdsdp_linsolver.c:1367-1381): a Cholesky factorisation that
// fails switches the system to the direct indefinite solver
hdsdp_retcode hdsdp_dev_kkt_factorize(hdsdp_dev *d) {
    if (d->useLU) return lu_factor(d);
    hdk::mirror_lower<double>(d->st, d->m, d->M, d->m);
    hdsdp_retcode rc = build_precond(d);
    if (rc != HDSDP_RETCODE_OK && d->mGramPending) {
        DCALL(rebuild_gram_fp64(d));
        rc = build_precond(d);
    }
    if (rc == HDSDP_RETCODE_FAILED && !d->useJacobi) rc = lu_factor(d);
    return rc;
}

int hdsdp_dev_kkt_uses_ldl(hdsdp_dev *d) {
    return d->useLU ? 1 : 0;
}

// HFpLinsysSolve (linalg/hdsdp_linsolver.c:1410-1425): a CG solve that fails switches the
// system to the direct indefinite solver, which then solves this right-hand side
hdsdp_retcode hdsdp_dev_kkt_solve(hdsdp_dev *d, const double *rhs, double *sol) {
    DCALL(upload(d, d->crhs, rhs, d->m, d->hvecB));
    prof_begin(d, HDSDP_PROF_CG);
    int rc = 0;
    if (d->useLU) {
        hdk::lu_solve(d->st, d->m, d->LUm, d->m, d->luPerm, d->crhs, d->cx);
        DCALL(check_launch());
    } else {
        rc = cg_solve(d);
        if (rc < 0 && d->mGramPending) {
            DCALL(rebuild_gram_fp64(d));
            if (build_precond(d) == HDSDP_RETCODE_OK) rc = cg_solve(d);
        }
        if (rc < 0) {
            DCALL(upload(d, d->crhs, rhs, d->m, d->hvecB));
            if (lu_factor(d) == HDSDP_RETCODE_OK) {
                hdk::lu_solve(d->st, d->m, d->LUm, d->m, d->luPerm, d->crhs, d->cx);
                DCALL(check_launch());
                rc = 0;
            }
        }
    }
    DCALL(prof_end(d, HDSDP_PROF_CG));
    if (rc < 0) return HDSDP_RETCODE_FAILED;
    if (rc == 1) return HDSDP_RETCODE_OK;  // tiny right-hand side: solution left as it was
    return download(d, sol, d->cx, d->m, d->hvecB);
}

// z = (L L^T)^-1 r with the FP64 factor of M (MLi = L^-1, MLiT = L^-T); tmp: m-vector.
void apply_chol64(hdsdp_dev *d, const double *r, double *z, double *tmp) {
    const int m = d->m;
    hdk::coldot<double>(d->st, m, m, 2, 1.0, d->MLiT, m, r, 0.0, tmp);
    hdk::coldot<double>(d->st, m, m, 1, 1.0, d->MLi, m, tmp, 0.0, z);
}

// Direct FP64 solve of M x = rhs (FP64 Gram blocks, FP64 Cholesky, two refinement steps
// on the residual with M), for the centring and adjoint solves, which need the solution
// to working accuracy rather than to the interior-point tolerance.
//
// reuse: when the FP64 factor on the device belongs to an earlier M (the last centring
// Newton step factored the Schur matrix of the point before the centred one), refine
// against the current M with that factor as the preconditioner instead of factoring
// again: x = P^-1 b, then x += P^-1 (b - M x) until the correction stops shrinking (at
// most eight steps).  Kept only if the factor contracts the first correction tenfold;
// otherwise (a factor of a different matrix) M is factored afresh.
hdsdp_retcode solve_exact(hdsdp_dev *d, const double *rhs, double *sol, bool reuse) {
    const int m = d->m;
    if (d->mGramPending) DCALL(rebuild_gram_fp64(d));
    hdk::mirror_lower<double>(d->st, m, d->M, m);
    DCALL(upload(d, d->crhs, rhs, m, d->hvecB));
    if (reuse && d->cholVer >= 0 && d->cholVer != d->mVer) {
        apply_chol64(d, d->crhs, d->cx, d->ctmp);
        double z0 = -1.0, zprev = HUGE_VAL;
        bool ok = false;
        for (int it = 0; it < 8; ++it) {
            symv_M(d, d->cx, d->crn);
            hdk::cg_residual(d->st, m, d->crhs, d->crn, d->cr);
            apply_chol64(d, d->cr, d->cz, d->ctmp);
            hdk::axpy_norm(d->st, m, 0.0, d->cx, d->cz, d->sc_d, hdk::SC_AUX);   // ||z||
            hdk::axpy_norm(d->st, m, 1.0, d->cz, d->cx, d->sc_d, hdk::CG_TRUE);  // x += z, ||x||
            DCALL(read_scalars(d));
            const double zn = d->hsc[hdk::SC_AUX], xn = d->hsc[hdk::CG_TRUE];
            if (!(zn == zn) || !(xn == xn)) break;
            if (it == 0) z0 = zn;
            if (it == 1 && zn > 0.1 * z0) break;  // not a factor of a nearby matrix
            if (zn <= 1e-15 * xn || (it >= 1 && zn > 0.5 * zprev)) {
                ok = it >= 1;
                break;
            }
            zprev = zn;
        }
        if (ok) {
            d->stats[6] += 1;
            DCALL(check_launch());
            return download(d, sol, d->cx, m, d->hvecB);
        }
    }
    if (d->cholVer != d->mVer) {
        prof_begin(d, HDSDP_PROF_MFACTOR);
        DCALL(build_chol64(d));
        DCALL(prof_end(d, HDSDP_PROF_MFACTOR));
    }
    apply_chol64(d, d->crhs, d->cx, d->ctmp);
    for (int it = 0; it < 2; ++it) {
        symv_M(d, d->cx, d->crn);
        hdk::cg_residual(d->st, m, d->crhs, d->crn, d->cr);
        apply_chol64(d, d->cr, d->cz, d->ctmp);
        hdk::axpy(d->st, m, 1.0, d->cz, d->cx);
    }
    DCALL(check_launch());
    return download(d, sol, d->cx, m, d->hvecB);
}

hdsdp_retcode hdsdp_dev_kkt_solve_exact(hdsdp_dev *d, const double *rhs, double *sol) {
    return solve_exact(d, rhs, sol, false);
}

hdsdp_retcode hdsdp_dev_get_schur(hdsdp_dev *d, double *M) {
    const int m = d->m;
    hdk::mirror_lower<double>(d->st, m, d->M, m);
    DCALL(sync(d));
    HIP_OK(hipMemcpy(M, d->M, sizeof(double) * (size_t) m * m, hipMemcpyDeviceToHost));
    return HDSDP_RETCODE_OK;
}

// ---- primal recovery ----------------------------------------------------------------------

hdsdp_retcode hdsdp_dev_primal(hdsdp_dev *d, double mu, const double *y, const double *dy, int *ok) {
    (void) y;
    const int n = d->n;
    prof_begin(d, HDSDP_PROF_PRIMAL);
    DCALL(hdsdp_dev_factorize(d, 1, ok));
    if (!*ok) return HDSDP_RETCODE_OK;
    DCALL(hdsdp_dev_update_buffer(d, 0.0, 1.0, dy, 0.0, 2));
    hdk::mirror_lower<double>(d->st, n, d->S[2], n);
    DCALL(ensure_linv(d, 1));
    // X = mu L^-T (I + L^-1 dS L^-T) L^-1 for the checker factor L
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->Li[1], n, 0, d->S[2], n, 0, 0.0, d->T1, n, 0, 1);
    hdk::gemm<double>(d->st, false, true, false, n, n, n, 1.0, d->T1, n, 0, d->Li[1], n, 0, 0.0, d->T2, n, 0, 1);
    hdk::sym_avg(d->st, n, d->T2, 1.0, 1.0, d->T1);
    hdk::gemm<double>(d->st, true, false, false, n, n, n, 1.0, d->Li[1], n, 0, d->T1, n, 0, 0.0, d->T2, n, 0, 1);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->T2, n, 0, d->Li[1], n, 0, 0.0, d->T1, n, 0, 1);
    hdk::sym_avg(d->st, n, d->T1, mu, 0.0, d->X);
    DCALL(check_launch());
    return prof_end(d, HDSDP_PROF_PRIMAL);
}

hdsdp_retcode hdsdp_dev_primal_stats(hdsdp_dev *d, double *AX, double *traceCX, double *XdotS, double *traceX) {
    const int m = d->m;
    hdk::traces(d->st, d->c, 0, m + 1, d->X, nullptr, 0.0, d->vecA_d, nullptr, false);
    hdk::r1_quadforms(d->st, d->c, d->X, true, 0, m + 1, 0.0, d->vecA_d, nullptr, d->H1);
    hdk::xdots(d->st, d->n, d->X, d->S[0], d->sc_d + hdk::SC_AUX);
    DCALL(check_launch());
    DCALL(download(d, d->hvecB, d->vecA_d, m + 1, d->hvecA));
    *traceCX = d->hvecB[0];
    memcpy(AX, d->hvecB + 1, sizeof(double) * m);
    DCALL(read_scalars(d));
    *XdotS = d->hsc[hdk::SC_AUX];
    *traceX = d->hsc[hdk::SC_AUX + 1];
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_get_primal(hdsdp_dev *d, double *X) {
    DCALL(sync(d));
    HIP_OK(hipMemcpy(X, d->X, sizeof(double) * (size_t) d->n * d->n, hipMemcpyDeviceToHost));
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_get_dual(hdsdp_dev *d, double *S) {
    const int n = d->n;
    HIP_OK(hipMemcpyAsync(d->T1, d->S[0], sizeof(double) * (size_t) n * n, hipMemcpyDeviceToDevice, d->st));
    hdk::mirror_lower<double>(d->st, n, d->T1, n);
    DCALL(sync(d));
    HIP_OK(hipMemcpy(S, d->T1, sizeof(double) * (size_t) n * n, hipMemcpyDeviceToHost));
    return HDSDP_RETCODE_OK;
}

// ---- primal method (interface/hdsdp_psdp.c) ------------------------------------------------

namespace {

// Cholesky of X into LX with its inverse factors; *isPsd = 0 when X is not positive definite
// (HFpLinsysNumeric / HFpLinsysPsdCheck on the dense solver, linalg/hdsdp_linsolver.c:525-560).
hdsdp_retcode factor_x(hdsdp_dev *d, int *isPsd) {
    const int n = d->n;
    HIP_OK(hipMemcpyAsync(d->LX, d->X, sizeof(double) * (size_t) n * n, hipMemcpyDeviceToDevice, d->st));
    if (!hdk::potrf<double>(d->st, n, d->LX, n, d->T1, std::max((size_t) n * n, hdk::POTRF_WORK), d->info_d, &d->la))
        return HDSDP_RETCODE_FAILED;
    HIP_OK(hipMemcpyAsync(d->hinfo, d->info_d, sizeof(int), hipMemcpyDeviceToHost, d->st));
    DCALL(sync(d));
    *isPsd = (*d->hinfo == 0);
    if (!*isPsd) return HDSDP_RETCODE_OK;
    hdk::trtri<double>(d->st, n, d->LX, n, d->LXi, n, d->work);
    hdk::transpose<double>(d->st, n, d->LXi, n, d->LXiT, n);
    return check_launch();
}

}  // namespace

hdsdp_retcode hdsdp_dev_psdp_init(hdsdp_dev *d) {
    const int n = d->n;
    const size_t nn = (size_t) n * n;
    if (!d->Xs) {
        DCALL(dalloc(d, &d->Xs, nn));
        DCALL(dalloc(d, &d->dXp, nn));
        DCALL(dalloc(d, &d->XSX, nn));
        DCALL(dalloc(d, &d->Sold, nn));
        DCALL(dalloc(d, &d->LX, nn));
        DCALL(dalloc(d, &d->LXi, nn));
        DCALL(dalloc(d, &d->LXiT, nn));
        DCALL(dalloc(d, &d->lwarmP, n));
    }
    HIP_OK(hipMemcpyAsync(d->Xs, d->X, sizeof(double) * nn, hipMemcpyDeviceToDevice, d->st));
    d->lzComputedP = 0;
    d->psdpReady = 1;
    return check_launch();
}

hdsdp_retcode hdsdp_dev_psdp_rhs(hdsdp_dev *d, double mu, double *AX, double *rhs) {
    const int m = d->m, n = d->n;
    if (!d->psdpReady) return HDSDP_RETCODE_FAILED;
    // the dual matrix of this iteration, both triangles (HConeGetDual: interface/hdsdp_conic_sdp.c:2427-2438)
    HIP_OK(hipMemcpyAsync(d->Sold, d->S[0], sizeof(double) * (size_t) n * n, hipMemcpyDeviceToDevice, d->st));
    hdk::mirror_lower<double>(d->st, n, d->Sold, n);
    // A X (hdsdp_psdp.c:183-186)
    hdk::traces(d->st, d->c, 1, m, d->X, nullptr, 0.0, d->vecA_d, nullptr, false);
    hdk::r1_quadforms(d->st, d->c, d->X, false, 1, m, 0.0, d->vecA_d, nullptr, d->H1);
    // X S X (denseMatTripMultiplypy, hdsdp_psdp.c:21-52), then A(X S X) - mu A X (hdsdp_psdp.c:198-206)
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->Sold, n, 0, d->X, n, 0, 0.0, d->T1, n, 0, 1);
    hdk::gemm<double>(d->st, true, false, false, n, n, n, 1.0, d->X, n, 0, d->T1, n, 0, 0.0, d->XSX, n, 0, 1);
    hdk::traces(d->st, d->c, 1, m, d->XSX, nullptr, 0.0, d->vecB_d, nullptr, false);
    hdk::r1_quadforms(d->st, d->c, d->XSX, false, 1, m, 0.0, d->vecB_d, nullptr, d->H1);
    hdk::axpy(d->st, m, -mu, d->vecA_d, d->vecB_d);
    DCALL(check_launch());
    DCALL(download(d, AX, d->vecA_d, m, d->hvecA));
    return download(d, rhs, d->vecB_d, m, d->hvecB);
}

hdsdp_retcode hdsdp_dev_psdp_direction(hdsdp_dev *d, double mu, int first, double *maxStep) {
    const int n = d->n;
    const long nn = (long) n * n;
    if (!d->psdpReady) return HDSDP_RETCODE_FAILED;
    // X S X + Xs dS Xs with dS the dual step the ratio test left in DUALSTEP (hdsdp_psdp.c:213-218)
    hdk::mirror_lower<double>(d->st, n, d->S[2], n);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->S[2], n, 0, d->Xs, n, 0, 0.0, d->T1, n, 0, 1);
    hdk::gemm<double>(d->st, true, false, false, n, n, n, 1.0, d->Xs, n, 0, d->T1, n, 0, 1.0, d->XSX, n, 0, 1);
    // dX = X - (X S X + Xs dS Xs) / mu (hdsdp_psdp.c:220-224)
    HIP_OK(hipMemcpyAsync(d->dXp, d->X, sizeof(double) * nn, hipMemcpyDeviceToDevice, d->st));
    hdk::axpy(d->st, nn, -1.0 / mu, d->XSX, d->dXp);
    // the reference's products with dX read its lower triangle (fds_symv, linalg/dense_opts.c:46-51)
    hdk::mirror_lower<double>(d->st, n, d->dXp, n);
    if (first) {
        int isPsd = 0;
        DCALL(factor_x(d, &isPsd));
        if (!isPsd) return HDSDP_RETCODE_FAILED;
    }
    // the primal Lanczos ratio test on L_X^-1 (-dX) L_X^-T (hdsdp_psdp.c:227-231, 58-66)
    prof_begin(d, HDSDP_PROF_LANCZOS);
    const LzOp op{d->LXi, d->LXiT, d->dXp, &d->lzComputedP, d->lwarmP};
    DCALL(lanczos(d, op, maxStep));
    return prof_end(d, HDSDP_PROF_LANCZOS);
}

hdsdp_retcode hdsdp_dev_psdp_step(hdsdp_dev *d, double alpha, int *isPsd, double *traceCX, double *XdotS) {
    const int n = d->n;
    if (!d->psdpReady) return HDSDP_RETCODE_FAILED;
    // X += alpha dX, then the factor of the new X (hdsdp_psdp.c:237-267)
    hdk::axpy(d->st, (long) n * n, alpha, d->dXp, d->X);
    DCALL(factor_x(d, isPsd));
    // <C, X> and <X, S> with the dual matrix of this iteration (hdsdp_psdp.c:268-271)
    hdk::traces(d->st, d->c, 0, 1, d->X, nullptr, 0.0, d->vecA_d, nullptr, false);
    hdk::r1_quadforms(d->st, d->c, d->X, true, 0, 1, 0.0, d->vecA_d, nullptr, d->H1);
    hdk::xdots(d->st, n, d->X, d->Sold, d->sc_d + hdk::SC_AUX);
    DCALL(check_launch());
    DCALL(download(d, d->hvecB, d->vecA_d, 1, d->hvecA));
    *traceCX = d->hvecB[0];
    DCALL(read_scalars(d));
    *XdotS = d->hsc[hdk::SC_AUX];
    return HDSDP_RETCODE_OK;
}

// ---- adjoint -------------------------------------------------------------------------------

hdsdp_retcode hdsdp_dev_adjoint(hdsdp_dev *d, double mu, const double *dLdy, double *dLdb, double *dLdC) {
    const int m = d->m, n = d->n;
    std::vector<double> lam(m, 0.0);
    DCALL(solve_exact(d, dLdy, lam.data(), true));
    for (int i = 0; i < m; ++i) dLdb[i] = lam[i] / mu;
    // dLdC = W (sum_i lambda_i A_i) W with W = S^-1 of the dual factor
    DCALL(ensure_linv(d, 0));
    hdk::gemm<double>(d->st, true, false, false, n, n, n, 1.0, d->Li[0], n, 0, d->Li[0], n, 0, 0.0, d->W, n, 0, 1);
    DCALL(hdsdp_dev_update_buffer(d, 0.0, 1.0, lam.data(), 0.0, 2));
    hdk::mirror_lower<double>(d->st, n, d->S[2], n);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->W, n, 0, d->S[2], n, 0, 0.0, d->T1, n, 0, 1);
    hdk::gemm<double>(d->st, false, false, false, n, n, n, 1.0, d->T1, n, 0, d->W, n, 0, 0.0, d->T2, n, 0, 1);
    // the gradient with respect to a symmetric matrix: (G + G^T) / 2 (the two products
    // leave rounding-level asymmetry, large in absolute terms when S is near-singular)
    hdk::sym_avg(d->st, n, d->T2, 1.0, 0.0, d->T1);
    DCALL(sync(d));
    HIP_OK(hipMemcpy(dLdC, d->T1, sizeof(double) * (size_t) n * n, hipMemcpyDeviceToHost));
    return HDSDP_RETCODE_OK;
}

}  // extern "C"
