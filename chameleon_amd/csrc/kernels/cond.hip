/*
 * cond.hip -- symmetric positive definite test matrices with a prescribed
 * spectrum, A = Q diag(lambda) Q^T, built on the device.
 *
 * The reference's conditioned generator (compute/pzlatms.c:74-300) takes Q
 * from the QR factorisation of a random matrix and lambda from
 * coreblas/compute/core_zlatm1.c (mode 3: geometric from 1 to 1/cond).
 * Here Q is a product of r random Householder reflectors
 * H = I - tau v v^T, applied from both sides, which keeps the spectrum
 * exactly lambda at O(r N^2) cost instead of O(N^3):
 *     H A H = A - v z^T - z v^T,  w = A v,  z = tau w - (tau^2/2)(v^T w) v.
 * The kernels work on the full (both triangles) FP64 tile mirror.
 */
#include "../internal.h"

namespace chx {
namespace k {

struct CondGeo {
    int NT, ldt, N;
};

/* A = diag(lambda_i), lambda_i = kappa^(-i/(N-1)); padding 0 (identity on
 * the padded diagonal of diagonal tiles) */
__global__ void cond_init_kernel(double *A, CondGeo g, double kappa)
{
    const size_t ts = (size_t)g.ldt * g.ldt;
    const size_t total = ts * g.NT * g.NT;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t t = e / ts, w = e % ts;
        const int m = (int)(t % g.NT), n = (int)(t / g.NT);
        const int i = (int)(w % g.ldt), j = (int)(w / g.ldt);
        const int gi = m * g.ldt + i, gj = n * g.ldt + j;
        double v = 0.0;
        if (gi == gj) {
            if (gi < g.N) v = (g.N > 1) ? pow(kappa, -(double)gi / (double)(g.N - 1)) : 1.0;
            else v = 1.0;
        }
        A[e] = v;
    }
}

/* w += A(m, n) v_n for one tile per workgroup (partial sums by atomics) */
__global__ __launch_bounds__(256) void cond_gemv_kernel(const double *A, CondGeo g, const double *v, double *w)
{
    const int m = blockIdx.x % g.NT, n = blockIdx.x / g.NT;
    const size_t ts = (size_t)g.ldt * g.ldt;
    const double *T = A + ((size_t)n * g.NT + m) * ts;
    const double *vn = v + (size_t)n * g.ldt;
    for (int i = threadIdx.x; i < g.ldt; i += blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < g.ldt; ++j) s += T[i + (size_t)j * g.ldt] * vn[j];
        atomicAdd(&w[(size_t)m * g.ldt + i], s);
    }
}

/* A -= v z^T + z v^T on the valid part */
__global__ void cond_rank2_kernel(double *A, CondGeo g, const double *v, const double *z)
{
    const size_t ts = (size_t)g.ldt * g.ldt;
    const size_t total = ts * g.NT * g.NT;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t t = e / ts, w = e % ts;
        const int m = (int)(t % g.NT), n = (int)(t / g.NT);
        const int gi = m * g.ldt + (int)(w % g.ldt), gj = n * g.ldt + (int)(w / g.ldt);
        A[e] -= v[gi] * z[gj] + z[gi] * v[gj];
    }
}

void cond_init(double *A, int NT, int ldt, int N, double kappa, hipStream_t s)
{
    CondGeo g = {NT, ldt, N};
    hipLaunchKernelGGL(cond_init_kernel, dim3(4096), dim3(256), 0, s, A, g, kappa);
}

void cond_gemv(const double *A, int NT, int ldt, int N, const double *v, double *w, hipStream_t s)
{
    CondGeo g = {NT, ldt, N};
    hipLaunchKernelGGL(cond_gemv_kernel, dim3(NT * NT), dim3(256), 0, s, A, g, v, w);
}

void cond_rank2(double *A, int NT, int ldt, int N, const double *v, const double *z, hipStream_t s)
{
    CondGeo g = {NT, ldt, N};
    hipLaunchKernelGGL(cond_rank2_kernel, dim3(8192), dim3(256), 0, s, A, g, v, z);
}

} /* namespace k */
} /* namespace chx */
