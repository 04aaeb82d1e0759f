// Device helpers shared by the Gauss-Seidel kernels (kernels.hip,
// tiles.hip): 5x5 block LU factorisation and solves, fixed row width,
// device-coherent X accessors.
#pragma once
#include <hip/hip_runtime.h>

namespace fpsi {

// ---------------------------------------------------------------------------
// in-register pivoted LU of a 5x5 block and its solve (Loci Mat
// decompose_lu_pivot / solve_lu_pivot; textbook Doolittle with partial
// pivoting, first maximum wins)
// ---------------------------------------------------------------------------
template <typename F>
__device__ __forceinline__ int lu5(F a[25]) {
  int perm[5] = {0, 1, 2, 3, 4};
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    int pr = k;
    F amax = (F)fabs((double)a[k + 5 * k]);
#pragma unroll
    for (int i = k + 1; i < 5; ++i) {
      const F v = (F)fabs((double)a[i + 5 * k]);
      if (v > amax) { amax = v; pr = i; }
    }
#pragma unroll
    for (int r = k + 1; r < 5; ++r) {
      const bool s = (r == pr);
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        const F t = a[k + 5 * j];
        a[k + 5 * j] = s ? a[r + 5 * j] : t;
        a[r + 5 * j] = s ? t : a[r + 5 * j];
      }
      const int tp = perm[k];
      perm[k] = s ? perm[r] : tp;
      perm[r] = s ? tp : perm[r];
    }
#pragma unroll
    for (int i = k + 1; i < 5; ++i) {
      const F l = a[i + 5 * k] / a[k + 5 * k];
      a[i + 5 * k] = l;
#pragma unroll
      for (int j = k + 1; j < 5; ++j) a[i + 5 * j] = a[i + 5 * j] - l * a[k + 5 * j];
    }
  }
  return perm[0] | (perm[1] << 3) | (perm[2] << 6) | (perm[3] << 9) | (perm[4] << 12);
}

template <typename F>
__device__ __forceinline__ void lusolve5(const F a[25], int pk, const double b[5], double x[5]) {
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const int pi = (pk >> (3 * i)) & 7;
    double tmp = pi == 0 ? b[0] : pi == 1 ? b[1] : pi == 2 ? b[2] : pi == 3 ? b[3] : b[4];
#pragma unroll
    for (int j = 0; j < i; ++j) tmp -= a[i + 5 * j] * x[j];
    x[i] = tmp;
  }
#pragma unroll
  for (int i = 4; i >= 0; --i) {
    double tmp = x[i];
#pragma unroll
    for (int j = i + 1; j < 5; ++j) tmp -= a[i + 5 * j] * x[j];
    x[i] = tmp / a[i + 5 * i];
  }
}

// ---------------------------------------------------------------------------
// cell assembly: cflpdt (src/solverSetup.loci:107-119), dtcfl/dtau
// (src/solverTimestep.loci:290-329), srcJ (src/solverLinearSystem.loci:106-120
// + boundary terms), fluid_D (:176-224), B = rhs in real_fj (:45-54), LU
// (src/linearSolverSGS.loci:43-50) and the Jacobi start X = D^-1 B (:57-69)
// ---------------------------------------------------------------------------
// LS: line-SGS diagonal.  fluid_D is kept unfactored in lu (identity
// pivots) and copied to FP64 and LU-factored without pivoting into lu64
// (src/linearSolverLSGS.loci:1633-1640, preprocessLine :1465-1470 for a line
// of one cell); the line-Jacobi start uses those factors (:1700-1733).
static constexpr int PIV_ID = 0 | (1 << 3) | (2 << 6) | (3 << 9) | (4 << 12);

__device__ __forceinline__ void lu5np(double a[25]) {
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int i = k + 1; i < 5; ++i) {
      const double l = a[i + 5 * k] / a[k + 5 * k];
      a[i + 5 * k] = l;
#pragma unroll
      for (int j = k + 1; j < 5; ++j) a[i + 5 * j] = a[i + 5 * j] - l * a[k + 5 * j];
    }
}

static constexpr int MAXR = 6;

template <bool COHERENT>
__device__ __forceinline__ double ldx(const double *p) {
  if (COHERENT) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return *p;
}
template <bool COHERENT>
__device__ __forceinline__ void stx(double *p, double v) {
  if (COHERENT) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

static constexpr int CPW = 12;  // cells per wave (60 of 64 lanes)

}  // namespace fpsi
