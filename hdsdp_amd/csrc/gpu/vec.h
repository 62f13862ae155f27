// vec.h -- vector kernels of the Schur CG and the Lanczos step length.
#pragma once

#include "dev_common.h"

namespace hdk {

// slots of the device scalar array read back by the host
enum {
    CG_RNORM = 0,
    CG_BNORM = 1,
    CG_RZ = 2,
    CG_ALPHA = 3,
    LZ_ALPHA = 4,
    LZ_NORM = 5,
    LZ_R1 = 6,
    LZ_R2 = 7,
    SC_LOGDET = 8,
    SC_AUX = 9,
    CG_TRUE = 11,
    SC_SLOTS = 16
};

void cg_start(hipStream_t s, int m, const double *rhs, double *r, double *x, double *sc);
void cg_jacobi(hipStream_t s, int m, const double *r, const double *jac, double *z, double *p);
void cg_step1(hipStream_t s, int m, const double *z, const double *r, const double *p, const double *Mp, double *x,
              double *rn, double *sc);
void cg_step2(hipStream_t s, int m, const double *rn, const double *z, double *p, double *r, double *sc);
void cg_residual(hipStream_t s, int m, const double *rhs, const double *Mx, double *r);
// r = rhs - Mx and sc[CG_TRUE] = ||r||
void cg_true_residual(hipStream_t s, int m, const double *rhs, const double *Mx, double *r, double *sc);

void lz_step(hipStream_t s, int n, int k, double *w, double *V, double *H, int ldh, double *v, double *sc,
             double *hist);
void lz_start(hipStream_t s, int n, const double *x, double *v, double *V);
void lz_comb(hipStream_t s, int n, int kp1, const double *V, const double *y, double *z);
void axpy_norm(hipStream_t s, int n, double a, const double *x, double *y, double *sc, int slot);

void axpy(hipStream_t s, long n, double a, const double *x, double *y);
void regularize(hipStream_t s, int m, double *M, long ldm, double r);
void get_diag(hipStream_t s, int m, const double *M, long ldm, double *d);
void f2d(hipStream_t s, long n, const float *x, double *y);
// y = x * sc (elementwise, with the precision conversion)
void f2d_scaled(hipStream_t s, long n, const float *x, const double *sc, double *y);
void d2f_scaled(hipStream_t s, long n, const double *x, const double *sc, float *y);
// sc[i] = 1 / sqrt(M_ii) (1 where M_ii <= 0)
void jacobi_scale(hipStream_t s, int m, const double *M, long ldm, double *sc);
void d2f(hipStream_t s, long n, const double *x, float *y);
// coef[0] = 0, coef[i + 1] = x[i] (coefficients of sum_i x_i A_i for the buffer assembly)
void shift_coef(hipStream_t s, int m, const double *x, double *coef);
// y[i] += (M_ii - ref[i]) x[i]
void diag_delta_axpy(hipStream_t s, int m, const double *M, long ldm, const double *ref, const double *x, double *y);

}  // namespace hdk
