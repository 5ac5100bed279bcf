// Device helpers shared by the Krylov kernels: fixed-order block and grid
// reductions (wavefront shuffles, then the 4 waves of a 256-lane block in LDS),
// block-uniform flag reads and the preconditioner application.
#pragma once
#include <hip/hip_runtime.h>

#include "elastic_internal.h"

namespace elastic {

constexpr int kThreads = 256;

__device__ inline double block_sum(double v, double* sh) {
  // wavefront (64 lanes) shuffle reduction, then across the 4 waves in LDS
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;  // valid in thread 0
}

// Reads a flag once per block (uniform branch for every wave).
__device__ inline bool block_flag(const int32_t* f, double* sh) {
  if (threadIdx.x == 0) sh[5] = (double)*f;
  __syncthreads();
  const bool d = sh[5] != 0.0;
  __syncthreads();
  return d;
}

// Sum of kRedBlocks partials, same order in every block; result broadcast.
__device__ inline double grid_sum(const double* __restrict__ part, double* sh) {
  double v = 0.0;
  for (int q = threadIdx.x; q < kRedBlocks; q += kThreads) v += part[q];
  double r = block_sum(v, sh);
  if (threadIdx.x == 0) sh[4] = r;
  __syncthreads();
  r = sh[4];
  __syncthreads();
  return r;
}

// z = M^-1 r at row i: Jacobi (m = 1/ad) or, when bm != null, the 3x3 block
// inverse of consecutive equation triples (Precond.f:419-540); rows past the
// last full triple use 1/ad.
__device__ inline double apply_m(int i, int n, const double* __restrict__ r,
                                 const double* __restrict__ m, const double* __restrict__ bm) {
  if (bm == nullptr) return r[i] * m[i];
  const int b = i / 3, nb = n / 3;
  if (b >= nb) return r[i] * m[i];
  const int i0 = 3 * b, c = i - i0;
  const double* a = bm + 6 * b;  // c11 c22 c33 c21 c31 c32
  const double x1 = r[i0], x2 = r[i0 + 1], x3 = r[i0 + 2];
  if (c == 0) return a[0] * x1 + a[3] * x2 + a[4] * x3;
  if (c == 1) return a[3] * x1 + a[1] * x2 + a[5] * x3;
  return a[4] * x1 + a[5] * x2 + a[2] * x3;
}

// Progress guard of the mixed methods, run at every FP64 true-residual
// replacement with the new true residual norm tr.  When it has not dropped
// below 0.9 x the best one so far at three replacements in a row, the FP32
// operator is not accurate enough for this system (its perturbation times
// the condition number approaches 1), and the solve continues with the FP64
// operator from the current x.  Comparing with the best value, not the last
// one, also catches a true residual that oscillates without progress.  A
// replacement is also forced after kForceReplace iterations without one, so
// stagnation is always seen.  The FP64 iteration then restarts from the true
// residual (p = M^-1 r) and makes no more replacements: the search directions
// built with the FP32 operator, and replacements that keep them, slow CG down
// badly on such systems (a fan mesh of condition number 5.5e9 did not
// converge in 50,000 iterations without the restart, and in 8,126 with it).
constexpr int kForceReplace = 500;
// Systems of at least this many rows let the host skip the gated
// replacement passes in poll chunks where the trigger cannot come yet.
constexpr int kPredictRows = 200000;
__device__ inline void progress_guard(SolveScalars* sc, double tr) {
  if (tr > 0.9 * sc->last_true) {
    sc->stall += 1;
  } else {
    sc->stall = 0;
    sc->last_true = tr;  // the best true residual so far
  }
  if (sc->stall >= 3 && !sc->op64) {
    sc->op64 = 1;
    sc->op64_from = sc->iter;
  }
  sc->last_upd = sc->iter;
}
__device__ inline void progress_init(SolveScalars* sc, double r0) {
  sc->restart[0] = sc->restart[1] = 0;
  sc->last_true = r0;
  sc->stall = 0;
  sc->op64 = 0;
  sc->op64_from = -1;
  sc->last_upd = 0;
}

#define GRID_LOOP(i, n) \
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < (n); i += gridDim.x * kThreads)

}  // namespace elastic
