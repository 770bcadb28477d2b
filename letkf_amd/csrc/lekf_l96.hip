// lekf_l96.hip -- the Lorenz-96 model on the resident ensemble (gfx950):
// the forecast step of a served L96 cycle and its obs operator.
//
// tinteg_rk4 / lorenz96_core (lorenz96/model/lorenz96.f90:93-128,290-309):
//   q = dt * (x_{i-1} (x_{i+1} - x_{i-2}) - x_i + F), periodic in i;
//   x <- x + (q1 + 2 q2 + 2 q3 + q4) / 6 with the stages at x, x + q1/2,
//   x + q2/2, x + q3.
// Every operation is rounded as written (no contraction), so a member's
// trajectory is bit-identical to the reference model's.  One workgroup per
// member, the state in LDS, lanes over grid points.
// H (lorenz96/obs/h_ope.f90:14-39) selects grid points: H(x) of obs n is
// x at gidx[n].
#include "lekf_dev.hpp"
#include "lekf_l96.hpp"

// Every FP64 operation in this file is rounded as written: no FMA contraction
// (the reference's results are reproduced operation by operation).
#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double mul(double a, double b) { return rn_mul(a, b); }
__device__ __forceinline__ double add(double a, double b) { return rn_add(a, b); }
__device__ __forceinline__ double sub(double a, double b) { return rn_sub(a, b); }

// q = dt * f(x) for the points of this thread
__device__ __forceinline__ void l96_core(const double* x, double* q, int nx, double dt, double force,
                                         int tid, int nt)
{
    for (int i = tid; i < nx; i += nt) {
        const int im1 = i == 0 ? nx - 1 : i - 1;
        const int im2 = i < 2 ? i + nx - 2 : i - 2;
        const int ip1 = i == nx - 1 ? 0 : i + 1;
        const double t = add(sub(mul(x[im1], sub(x[ip1], x[im2])), x[i]), force);
        q[i] = mul(dt, t);
    }
}

__global__ __launch_bounds__(256) void l96_rk4(const double* __restrict__ xin,
                                               double* __restrict__ xout, int nx, int ne, int kt,
                                               double dt, double force)
{
    extern __shared__ double sm[];
    double* x = sm;
    double* xt = x + nx;
    double* q1 = xt + nx;
    double* q2 = q1 + nx;
    double* q3 = q2 + nx;
    double* q4 = q3 + nx;
    const int m = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < nx; i += nt) x[i] = xin[(size_t)i * ne + m];
    __syncthreads();
    for (int s = 0; s < kt; ++s) {
        l96_core(x, q1, nx, dt, force, tid, nt);
        __syncthreads();
        for (int i = tid; i < nx; i += nt) xt[i] = add(x[i], mul(0.5, q1[i]));
        __syncthreads();
        l96_core(xt, q2, nx, dt, force, tid, nt);
        __syncthreads();
        for (int i = tid; i < nx; i += nt) xt[i] = add(x[i], mul(0.5, q2[i]));
        __syncthreads();
        l96_core(xt, q3, nx, dt, force, tid, nt);
        __syncthreads();
        for (int i = tid; i < nx; i += nt) xt[i] = add(x[i], q3[i]);
        __syncthreads();
        l96_core(xt, q4, nx, dt, force, tid, nt);
        __syncthreads();
        for (int i = tid; i < nx; i += nt) {
            const double s4 = add(add(add(q1[i], mul(2.0, q2[i])), mul(2.0, q3[i])), q4[i]);
            x[i] = add(x[i], rn_div(s4, 6.0));
        }
        __syncthreads();
    }
    for (int i = tid; i < nx; i += nt) xout[(size_t)i * ne + m] = x[i];
}

// hxf(ny, ne) column-major from the background rows x[ij * ne + m]
__global__ void l96_hx(int ny, int ne, const int* __restrict__ gidx, const double* __restrict__ x,
                       double* __restrict__ hxf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ny * ne) return;
    const int n = (int)(t % ny), m = (int)(t / ny);
    hxf[t] = x[(size_t)gidx[n] * ne + m];
}

}  // namespace

hipError_t lekf_l96_rk4_launch(const double* xin, double* xout, int nx, int ne, int kt, double dt,
                               double force, hipStream_t st)
{
    if (nx < 4 || ne < 1 || kt < 0) return hipErrorInvalidValue;
    const size_t lds = 6 * (size_t)nx * sizeof(double);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(l96_rk4, dim3((unsigned)ne), dim3(256), lds, st, xin, xout, nx, ne, kt, dt,
                       force);
    return hipGetLastError();
}

hipError_t lekf_l96_hx_launch(int ny, int ne, const int* gidx, const double* x, double* hxf,
                              hipStream_t st)
{
    if (ny <= 0) return hipSuccess;
    const long n = (long)ny * ne;
    hipLaunchKernelGGL(l96_hx, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ny, ne, gidx, x,
                       hxf);
    return hipGetLastError();
}
