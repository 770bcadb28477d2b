// lekf_relax.hip -- relaxation of the analysis toward the background and the
// q spread limit (scale/letkf/letkf_tools.f90:215-268, weight_RTPP
// :1224-1236, weight_RTPS :1240-1268, relax_beta :1182-1220), gfx950.
//
// The reference builds the relaxed weight matrix in ensemble space:
//   T' = ((1 - a) Wa + a I + w-bar 1^T) b + (1 - b) I          (RTPP)
//   T' = (f Wa + w-bar 1^T) b + (1 - b) I,
//        f = a sqrt(var_b / (x'^T Pa x' (k-1))) - a + 1         (RTPS, per row)
// and Xa = x-mean + X' T'.  Every term is a function of the unrelaxed
// analysis row of the same point: X' Wa = xa - mean(xa), X' w-bar =
// mean(xa) - mean(xb), and x'^T Pa x' = |xa - mean(xa)|^2 / (k-1) because
// Wa^2 = (k-1) Pa.  So one pass over each point's rows after the solve gives
// the reference's result for every kernel family.  One wave per point; lane
// l holds members l, l+64, l+128, l+192.
#include "lekf_dev.hpp"
#include "lekf_relax.hpp"

namespace {

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(64) void relax_points(RelaxArgs A)
{
    const long b = blockIdx.x;
    const int lane = threadIdx.x;
    const int k = A.ne;
    const int lev = (int)(b / A.nij), ij = (int)(b - (long)lev * A.nij);
    const int n3 = A.nv3d, nrow = n3 + (lev == 0 ? A.nv2d : 0);
    // relax_beta: boundary buffer of a regional grid (1-based i, j, no halo)
    double beta = 1.0;
    if (A.width > 0.0) {
        const int i = ij % A.nx + 1, j = ij / A.nx + 1;
        const double di = fmin((double)i, (double)(A.nx + 1 - i)) * A.dx;
        const double dj = fmin((double)j, (double)(A.ny + 1 - j)) * A.dy;
        const double d = fmin(di, dj) / A.width;
        if (d < 1.0) beta = fmax(d, 0.0);
    }
    const double rk = 1.0 / (double)k;
    for (int r = 0; r < nrow; ++r) {
        const double* xb = (r < n3) ? A.x3b + ((size_t)b * n3 + r) * k
                                    : A.x2b + ((size_t)ij * A.nv2d + (r - n3)) * k;
        double* xa = (r < n3) ? A.x3a + ((size_t)b * n3 + r) * k
                              : A.x2a + ((size_t)ij * A.nv2d + (r - n3)) * k;
        double gb[4], ga[4];
        double sb = 0.0, sa = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int m = lane + 64 * t;
            gb[t] = m < k ? xb[m] : 0.0;
            ga[t] = m < k ? xa[m] : 0.0;
            sb += gb[t];
            sa += ga[t];
        }
        const double mb = wsum(sb) * rk, ma = wsum(sa) * rk;
        double vg = 0.0, va = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bool in = lane + 64 * t < k;
            gb[t] = in ? gb[t] - mb : 0.0;
            ga[t] = in ? ga[t] - ma : 0.0;
            vg = fma(gb[t], gb[t], vg);
            va = fma(ga[t], ga[t], va);
        }
        const double cw = ma - mb;  // X' w-bar
        double c1 = beta, c2 = 1.0 - beta;
        if (A.mode == LEKF_RELAX_RTPP) {
            c1 = beta * (1.0 - A.alpha);
            c2 = 1.0 - beta + beta * A.alpha;
        } else if (A.mode == LEKF_RELAX_RTPS) {
            vg = wsum(vg);
            va = wsum(va);
            // x'^T Pa x' = va / (k-1), so var_b / (x'^T Pa x' (k-1)) = vg / va
            const double f = (vg > 0.0 && va > 0.0) ? A.alpha * sqrt(vg / va) - A.alpha + 1.0 : 1.0;
            c1 = beta * f;
        }
        double y[4];
        double sq = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            y[t] = mb + (beta * cw + (c1 * ga[t] + c2 * gb[t]));
            if (lane + 64 * t < k) sq += y[t];
        }
        if (A.qmax > 0.0 && r == A.iq && r < n3) {
            const double qm = wsum(sq) * rk;
            double ss = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const double d = lane + 64 * t < k ? y[t] - qm : 0.0;
                ss = fma(d, d, ss);
            }
            const double qs = sqrt(wsum(ss) / (double)(k - 1)) / qm;
            if (qs > A.qmax)
#pragma unroll
                for (int t = 0; t < 4; ++t) y[t] = qm + (y[t] - qm) * A.qmax / qs;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (lane + 64 * t < k) xa[lane + 64 * t] = y[t];
    }
}

}  // namespace

hipError_t lekf_relax_launch(const RelaxArgs& A, hipStream_t st)
{
    if (A.npt <= 0) return hipSuccess;
    if (A.ne < 2 || A.ne > 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(relax_points, dim3((unsigned)A.npt), dim3(64), 0, st, A);
    return hipGetLastError();
}
