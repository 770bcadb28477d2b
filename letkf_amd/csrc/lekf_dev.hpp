// lekf_dev.hpp -- device-side types shared by the LETKF kernels (gfx950).
//
// One "job" is one analysis problem (one ensemble, one obs table) whose local
// analyses ("points") are solved in a launch.  Several jobs with ensemble sizes
// in the same padded bucket share a launch; every point is solved by one
// 64-lane wavefront with its reduction order fixed by the point alone, so a
// point's result does not depend on what else is in the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lekf_types.h"

typedef double lekf_d4 __attribute__((ext_vector_type(4)));
typedef float lekf_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 lekf_bf8 __attribute__((ext_vector_type(8)));
typedef float lekf_f2v __attribute__((ext_vector_type(2)));
typedef __bf16 lekf_b2v __attribute__((ext_vector_type(2)));

// speedy/common/common_obs_speedy.f90:17-24
#define LEKF_ID_U 2819
#define LEKF_ID_V 2820
#define LEKF_ID_T 3073
#define LEKF_ID_Q 3330
#define LEKF_ID_RH 3331
#define LEKF_ID_PS 14593
#define LEKF_ID_RAIN 9999

// common/common.f90:28,35 (pi is truncated in the reference; it moves the
// localisation cutoffs, so it is kept)
#define LEKF_PI_REF 3.1415926535
#define LEKF_RE_REF 6371.3e3

#define LEKF_WAVE 64

// Round-to-nearest FP64 operations that the compiler must not fuse into FMAs.
// Device code is compiled with fp-contract=fast, and the __d*_rn functions of
// the HIP headers are then plain operators that a neighbouring add may absorb
// into an FMA; code that must round every product and sum as the reference
// does (selection distances, obs operator, L96 model) uses these instead.
__device__ __forceinline__ double rn_mul(double a, double b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double rn_add(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double rn_sub(double a, double b)
{
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double rn_div(double a, double b)
{
#pragma clang fp contract(off)
    return a / b;
}

// Jacobi: rotate a pair while |a_pq| > tol * sqrt(a_pp a_qq); stop after a
// sweep with no rotation.
#define LEKF_JACOBI_TOL 1.0e-15
#define LEKF_JACOBI_MAX_SWEEPS 40

// job of flat point gp: the last job whose pt_begin <= gp (binary search;
// pt_begin increases with the job index)
// x of lane (lane ^ j), for the constant j of an unrolled loop: quad
// permutes for 1 and 2, row shifts (lane n reads n + 4 or n - 4) for 4, a
// rotation by half a row for 8; for 16 and 32 the permlane swaps, which
// return x of the row (half) above in the even rows (lower half) and x of
// the row (half) below in the odd rows (upper half)
__device__ __forceinline__ int lane_xor(int x, int j, int lane)
{
    switch (j) {
    case 16: {
        const auto p = __builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)x, false, false);
        return (int)((lane & 16) ? p[0] : p[1]);
    }
    case 32: {
        const auto p = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)x, false, false);
        return (int)((lane & 32) ? p[0] : p[1]);
    }
    case 1: return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    case 2: return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    case 4: {
        const int up = __builtin_amdgcn_mov_dpp(x, 0x104, 0xF, 0xF, false);  // row_shl:4
        const int dn = __builtin_amdgcn_mov_dpp(x, 0x114, 0xF, 0xF, false);  // row_shr:4
        return (lane & 4) ? dn : up;
    }
    case 8: return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);  // row_ror:8
    default: return __shfl_xor(x, j, 64);
    }
}

__device__ __forceinline__ int lekf_job_of(const LekfJob* jobs, int njobs, long gp)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].pt_begin <= gp)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
