// vec.hip -- m- and n-vector work of the Schur CG and the Lanczos step length.
//
// Vectors here have at most a few ten thousand entries, so each fused step
// runs in one workgroup of 1024 threads: the reductions are done in a fixed
// order (deterministic run to run) and no grid-wide synchronisation is needed.
// Scalars that the host controls on are written to a small device array that
// the host reads back once per step.
#include "vec.h"

namespace hdk {

__device__ inline double block_sum(double v, double *red) {
    const int tid = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double total = 0.0;
    if (tid == 0) {
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) total += red[w];
        red[16] = total;
    }
    __syncthreads();
    total = red[16];
    __syncthreads();
    return total;
}

// ---- conjugate gradient (linalg/hdsdp_linsolver.c:855-997) -----------------

// r = rhs, x = 0, sc[CG_RNORM] = sc[CG_BNORM] = ||rhs||
__global__ __launch_bounds__(1024) void cg_start_kernel(int m, const double *rhs, double *r, double *x, double *sc) {
    __shared__ double red[17];
    double acc = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        double v = rhs[i];
        r[i] = v;
        x[i] = 0.0;
        acc += v * v;
    }
    double nrm = sqrt(block_sum(acc, red));
    if (threadIdx.x == 0) {
        sc[CG_RNORM] = nrm;
        sc[CG_BNORM] = nrm;
    }
}

// z = r / jacobi (jacobi != NULL) else z = r; optional copy p = z
__global__ __launch_bounds__(1024) void cg_jacobi_kernel(int m, const double *r, const double *jac, double *z, double *p) {
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < m; i += blockDim.x * gridDim.x) {
        double v = jac ? r[i] / jac[i] : r[i];
        z[i] = v;
        if (p) p[i] = v;
    }
}

// rz = z.r; pMp = p.Mp; alpha = rz / pMp; x += alpha p; rn = r - alpha Mp
__global__ __launch_bounds__(1024) void cg_step1_kernel(int m, const double *z, const double *r, const double *p,
                                                        const double *Mp, double *x, double *rn, double *sc) {
    __shared__ double red[17];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        a += z[i] * r[i];
        b += p[i] * Mp[i];
    }
    double rz = block_sum(a, red);
    double pMp = block_sum(b, red);
    double alpha = rz / pMp;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        x[i] += alpha * p[i];
        rn[i] = r[i] - alpha * Mp[i];
    }
    if (threadIdx.x == 0) {
        sc[CG_RZ] = rz;
        sc[CG_ALPHA] = alpha;
    }
}

// beta = rn.z / rz; p = z + beta p; r = rn; sc[CG_RNORM] = ||r||
__global__ __launch_bounds__(1024) void cg_step2_kernel(int m, const double *rn, const double *z, double *p, double *r,
                                                        double *sc) {
    __shared__ double red[17];
    double a = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) a += rn[i] * z[i];
    double beta = block_sum(a, red) / sc[CG_RZ];
    double c = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        p[i] = z[i] + beta * p[i];
        double v = rn[i];
        r[i] = v;
        c += v * v;
    }
    double nrm = sqrt(block_sum(c, red));
    if (threadIdx.x == 0) sc[CG_RNORM] = nrm;
}

// r = rhs - Mx (Mx precomputed)
__global__ void cg_residual_kernel(int m, const double *rhs, const double *Mx, double *r) {
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < m; i += blockDim.x * gridDim.x) r[i] = rhs[i] - Mx[i];
}

__global__ __launch_bounds__(1024) void cg_true_residual_kernel(int m, const double *rhs, const double *Mx, double *r,
                                                                double *sc) {
    __shared__ double red[17];
    double acc = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const double v = rhs[i] - Mx[i];
        r[i] = v;
        acc += v * v;
    }
    const double nrm = sqrt(block_sum(acc, red));
    if (threadIdx.x == 0) sc[CG_TRUE] = nrm;
}

void cg_true_residual(hipStream_t s, int m, const double *rhs, const double *Mx, double *r, double *sc) {
    cg_true_residual_kernel<<<1, 1024, 0, s>>>(m, rhs, Mx, r, sc);
}

void cg_start(hipStream_t s, int m, const double *rhs, double *r, double *x, double *sc) {
    cg_start_kernel<<<1, 1024, 0, s>>>(m, rhs, r, x, sc);
}
void cg_jacobi(hipStream_t s, int m, const double *r, const double *jac, double *z, double *p) {
    cg_jacobi_kernel<<<(m + 1023) / 1024, 1024, 0, s>>>(m, r, jac, z, p);
}
void cg_step1(hipStream_t s, int m, const double *z, const double *r, const double *p, const double *Mp, double *x,
              double *rn, double *sc) {
    cg_step1_kernel<<<1, 1024, 0, s>>>(m, z, r, p, Mp, x, rn, sc);
}
void cg_step2(hipStream_t s, int m, const double *rn, const double *z, double *p, double *r, double *sc) {
    cg_step2_kernel<<<1, 1024, 0, s>>>(m, rn, z, p, r, sc);
}
void cg_residual(hipStream_t s, int m, const double *rhs, const double *Mx, double *r) {
    cg_residual_kernel<<<(m + 1023) / 1024, 1024, 0, s>>>(m, rhs, Mx, r);
}

// ---- Lanczos (linalg/hdsdp_lanczos.c:196-286) ---------------------------------

// One Lanczos orthogonalisation step on w = Op(v_k):
//   w -= H(k,k-1) V_{k-1};  alp = -w.V_k;  w += alp V_k;  H(k,k) = -alp
//   norm = ||w||;  v = w / norm (or 0);  if norm > 0: V_{k+1} = v, H(k+1,k) = H(k,k+1) = norm
// sc[LZ_ALPHA] = -alp, sc[LZ_NORM] = norm.
__global__ __launch_bounds__(1024) void lz_step_kernel(int n, int k, double *w, double *V, double *H, int ldh, double *v,
                                                       double *sc, double *hist) {
    __shared__ double red[17];
    const double *Vk = V + (long) k * n;
    if (k > 0) {
        const double *Vkm = V + (long) (k - 1) * n;
        double h = -H[k + (k - 1) * ldh];
        for (int i = threadIdx.x; i < n; i += blockDim.x) w[i] += h * Vkm[i];
        __syncthreads();
    }
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += w[i] * Vk[i];
    double alp = -block_sum(a, red);
    double c = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double t = w[i] + alp * Vk[i];
        w[i] = t;
        c += t * t;
    }
    double nrm = sqrt(block_sum(c, red));
    double inv = (nrm > 0.0) ? 1.0 / nrm : 0.0;
    double *Vn = V + (long) (k + 1) * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double t = (nrm > 0.0) ? w[i] * inv : 0.0;
        v[i] = t;
        if (nrm > 0.0) Vn[i] = t;
    }
    if (threadIdx.x == 0) {
        H[k + k * ldh] = -alp;
        if (nrm > 0.0) {
            H[(k + 1) + k * ldh] = nrm;
            H[k + (k + 1) * ldh] = nrm;
        }
        sc[LZ_ALPHA] = -alp;
        sc[LZ_NORM] = nrm;
        hist[2 * k] = -alp;
        hist[2 * k + 1] = nrm;
    }
}

// v = x / ||x|| (or 0); V_0 = v (start vector normalisation)
__global__ __launch_bounds__(1024) void lz_start_kernel(int n, const double *x, double *v, double *V) {
    __shared__ double red[17];
    double c = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) c += x[i] * x[i];
    double nrm = sqrt(block_sum(c, red));
    double inv = (nrm > 0.0) ? 1.0 / nrm : 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double t = (nrm > 0.0) ? x[i] * inv : 0.0;
        v[i] = t;
        V[i] = t;
    }
}

// z = V[:, 0:kp1] y
__global__ void lz_comb_kernel(int n, int kp1, const double *V, const double *y, double *z) {
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x) {
        double s = 0.0;
        for (int c = 0; c < kp1; ++c) s += V[i + (long) c * n] * y[c];
        z[i] = s;
    }
}

// y += a x; sc[slot] = ||y||
__global__ __launch_bounds__(1024) void axpy_norm_kernel(int n, double a, const double *x, double *y, double *sc, int slot) {
    __shared__ double red[17];
    double c = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double t = y[i] + a * x[i];
        y[i] = t;
        c += t * t;
    }
    double nrm = sqrt(block_sum(c, red));
    if (threadIdx.x == 0) sc[slot] = nrm;
}

void lz_step(hipStream_t s, int n, int k, double *w, double *V, double *H, int ldh, double *v, double *sc,
             double *hist) {
    lz_step_kernel<<<1, 1024, 0, s>>>(n, k, w, V, H, ldh, v, sc, hist);
}
void lz_start(hipStream_t s, int n, const double *x, double *v, double *V) {
    lz_start_kernel<<<1, 1024, 0, s>>>(n, x, v, V);
}
void lz_comb(hipStream_t s, int n, int kp1, const double *V, const double *y, double *z) {
    lz_comb_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, kp1, V, y, z);
}
void axpy_norm(hipStream_t s, int n, double a, const double *x, double *y, double *sc, int slot) {
    axpy_norm_kernel<<<1, 1024, 0, s>>>(n, a, x, y, sc, slot);
}

// ---- generic ------------------------------------------------------------------

__global__ void axpy_kernel(long n, double a, const double *x, double *y) {
    for (long i = threadIdx.x + blockIdx.x * (long) blockDim.x; i < n; i += (long) blockDim.x * gridDim.x)
        y[i] += a * x[i];
}

void axpy(hipStream_t s, long n, double a, const double *x, double *y) {
    long blocks = (n + 255) / 256;
    axpy_kernel<<<blocks > 4096 ? 4096 : blocks, 256, 0, s>>>(n, a, x, y);
}

// M_ii += reg where reg = min(r * min_i M_ii, 1e-5), dropped below 1e-14
// (interface/hdsdp_schur.c:348-373)
__global__ __launch_bounds__(1024) void regularize_kernel(int m, double *M, long ldm, double r) {
    __shared__ double red[17];
    double mn = 1e+30;
    for (int i = threadIdx.x; i < m; i += blockDim.x) mn = fmin(mn, M[i + i * ldm]);
    for (int off = 32; off > 0; off >>= 1) mn = fmin(mn, __shfl_down(mn, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 1e+30;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) t = fmin(t, red[w]);
        double reg = r * t;
        reg = fmin(reg, 1e-05);
        if (reg < 1e-14) reg = 0.0;
        red[16] = reg;
    }
    __syncthreads();
    double reg = red[16];
    for (int i = threadIdx.x; i < m; i += blockDim.x) M[i + i * ldm] += reg;
}

__global__ void shift_coef_kernel(int m, const double *x, double *coef) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > m) return;
    coef[i] = (i == 0) ? 0.0 : x[i - 1];
}

void shift_coef(hipStream_t s, int m, const double *x, double *coef) {
    shift_coef_kernel<<<(m + 256) / 256, 256, 0, s>>>(m, x, coef);
}

__global__ void diag_delta_axpy_kernel(int m, const double *M, long ldm, const double *ref, const double *x,
                                       double *y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    y[i] += (M[i + i * ldm] - ref[i]) * x[i];
}

void diag_delta_axpy(hipStream_t s, int m, const double *M, long ldm, const double *ref, const double *x, double *y) {
    diag_delta_axpy_kernel<<<(m + 255) / 256, 256, 0, s>>>(m, M, ldm, ref, x, y);
}

void regularize(hipStream_t s, int m, double *M, long ldm, double r) {
    regularize_kernel<<<1, 1024, 0, s>>>(m, M, ldm, r);
}

__global__ void get_diag_kernel(int m, const double *M, long ldm, double *d) {
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < m; i += blockDim.x * gridDim.x) d[i] = M[i + i * ldm];
}

void get_diag(hipStream_t s, int m, const double *M, long ldm, double *d) {
    get_diag_kernel<<<(m + 255) / 256, 256, 0, s>>>(m, M, ldm, d);
}

__global__ void f2d_kernel(long n, const float *x, double *y) {
    for (long i = threadIdx.x + blockIdx.x * (long) blockDim.x; i < n; i += (long) blockDim.x * gridDim.x) y[i] = x[i];
}
__global__ void d2f_kernel(long n, const double *x, float *y) {
    for (long i = threadIdx.x + blockIdx.x * (long) blockDim.x; i < n; i += (long) blockDim.x * gridDim.x)
        y[i] = (float) x[i];
}

__global__ void f2d_scaled_kernel(long n, const float *x, const double *sc, double *y) {
    for (long i = threadIdx.x + blockIdx.x * (long) blockDim.x; i < n; i += (long) blockDim.x * gridDim.x)
        y[i] = (double) x[i] * sc[i];
}
__global__ void d2f_scaled_kernel(long n, const double *x, const double *sc, float *y) {
    for (long i = threadIdx.x + blockIdx.x * (long) blockDim.x; i < n; i += (long) blockDim.x * gridDim.x)
        y[i] = (float) (x[i] * sc[i]);
}
__global__ void jacobi_scale_kernel(int m, const double *M, long ldm, double *sc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        const double d = M[i + (long) i * ldm];
        sc[i] = (d > 0.0) ? 1.0 / sqrt(d) : 1.0;
    }
}
void f2d_scaled(hipStream_t s, long n, const float *x, const double *sc, double *y) {
    f2d_scaled_kernel<<<(n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256, 256, 0, s>>>(n, x, sc, y);
}
void d2f_scaled(hipStream_t s, long n, const double *x, const double *sc, float *y) {
    d2f_scaled_kernel<<<(n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256, 256, 0, s>>>(n, x, sc, y);
}
void jacobi_scale(hipStream_t s, int m, const double *M, long ldm, double *sc) {
    jacobi_scale_kernel<<<(m + 255) / 256, 256, 0, s>>>(m, M, ldm, sc);
}

void f2d(hipStream_t s, long n, const float *x, double *y) {
    f2d_kernel<<<(n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256, 256, 0, s>>>(n, x, y);
}
void d2f(hipStream_t s, long n, const double *x, float *y) {
    d2f_kernel<<<(n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256, 256, 0, s>>>(n, x, y);
}

}  // namespace hdk
