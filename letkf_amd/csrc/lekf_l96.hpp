// lekf_l96.hpp -- Lorenz-96 forecast and obs operator launchers.
#pragma once
#include <hip/hip_runtime.h>

// kt RK4 steps of every member: x rows [i * ne + m] in, out (may not alias)
hipError_t lekf_l96_rk4_launch(const double* xin, double* xout, int nx, int ne, int kt, double dt,
                               double force, hipStream_t st);
// hxf(ny, ne) column-major = x[gidx[n]][m]
hipError_t lekf_l96_hx_launch(int ny, int ne, const int* gidx, const double* x, double* hxf,
                              hipStream_t st);
