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
