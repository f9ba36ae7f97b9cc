// lpm.h -- the reference's common header for drivers written against LPM-CPP
// (include/lpm.h), backed by the GPU adapter: a driver that includes lpm.h,
// assembly.h, solver_static.h, ... from this directory instead of LPM-CPP's
// include/ builds unchanged and runs on the GPU through liblpmg.so.
//
// Provides the reference's macros (PI, EPS, NDIM, IFORMAT), func_phi, MKL_INT,
// a no-op mkl_free_buffers and the adapter's classes in the global namespace.
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

#include <omp.h>

#include "../lpm_gpu.hpp"

#define PI 3.14159265358979323846
#define EPS 1e-6
#define NDIM 3
#define IFORMAT "%lli"

using namespace lpm_gpu;

// func_phi (lpm.h:49-53): the nonlocal weight 1 / (1 + (x/L)^8)^2
inline double func_phi(const double x, const double L) {
  const double t = 1 / (1 + std::pow(x / L, 8.0));
  return std::pow(t, 2.0);
}

// The reference calls this at exit to return MKL's buffers; the GPU path holds none.
inline void mkl_free_buffers() {}
