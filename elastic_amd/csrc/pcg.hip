// Device-resident Krylov iterations on gfx950.
//
// PCG follows Iters.f:252-524 step for step (x0 = 0, Jacobi or 3x3
// block-Jacobi M, stop when sqrt|r.M r| < tol sqrt|b.M b|).  One iteration is
// four kernels and no host synchronisation:
//   spmv tiles (q = A p)   ->  merge + p.q   ->  x,r update + r.z  ->  p update
// Dot products are reduced in two fixed-order stages (per-block partials over a
// fixed grid of kRedBlocks blocks, then every consumer block sums the partials
// in the same order), so each block sees the same alpha/beta and results are
// bit-reproducible run to run.  Per-thread products are formed in the vector
// type T; block and grid sums are kept in FP64 for both T = double and
// T = float.  Scalars live in device memory (SolveScalars); kernels return at
// once when the solve has converged, so the host only polls a flag every
// `check_every` iterations.
#include <algorithm>
#include <chrono>
#include <cmath>

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

// Reads the solve's done flag once per block (uniform branch for every wave).
__device__ inline bool block_done(const SolveScalars* sc, double* sh) {
  if (threadIdx.x == 0) sh[5] = (double)sc->done;
  __syncthreads();
  bool d = sh[5] != 0.0;
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
template <typename T>
__device__ inline T apply_m(int i, int n, const T* __restrict__ r, const T* __restrict__ m,
                            const T* __restrict__ bm) {
  if (bm == nullptr) return r[i] * m[i];
  const int b = i / 3, nb = n / 3;
  if (b >= nb) return r[i] * m[i];
  const int i0 = 3 * b, c = i - i0;
  const T* a = bm + 6 * b;  // c11 c22 c33 c21 c31 c32
  const T x1 = r[i0], x2 = r[i0 + 1], x3 = r[i0 + 2];
  if (c == 0) return a[0] * x1 + a[3] * x2 + a[4] * x3;
  if (c == 1) return a[3] * x1 + a[1] * x2 + a[5] * x3;
  return a[4] * x1 + a[5] * x2 + a[2] * x3;
}

// x = 0, r = b, p = z = M^-1 b, partial b.z
template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_init(int n, const T* __restrict__ b,
                                                        const T* __restrict__ m,
                                                        const T* __restrict__ bm, T* x, T* r,
                                                        T* p, double* part) {
  __shared__ double sh[8];
  double acc = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    T zi = apply_m(i, n, b, m, bm);
    x[i] = T(0);
    r[i] = b[i];
    p[i] = zi;
    acc += (double)b[i] * (double)zi;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_pcg_init2(const double* __restrict__ part, double tol, double* dslot,
                            SolveScalars* sc) {
  __shared__ double sh[8];
  double d = grid_sum(part, sh);
  if (threadIdx.x == 0) {
    *dslot = d;
    sc->conv = tol * sqrt(fabs(d));
    sc->norm_b = sqrt(fabs(d));
    sc->rz0 = d;
    sc->iter = 0;
    sc->status = 0;
    sc->done = (d == 0.0) ? 1 : 0;  // b = 0: x = 0 is the solution
  }
}

// q += merged external partials; partial p.q
template <typename T>
__global__ __launch_bounds__(kThreads) void k_merge_dot(int n, const int32_t* __restrict__ inc_ptr,
                                                         const int32_t* __restrict__ inc_idx,
                                                         const T* __restrict__ epart, T* q,
                                                         const T* __restrict__ p, double* part,
                                                         const SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_done(sc, sh)) return;
  double acc = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    T s = q[i];
    for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) s += epart[inc_idx[k]];
    q[i] = s;
    acc += (double)p[i] * (double)s;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha = d / p.q ; x += alpha p ; r -= alpha q ; partial r.(M^-1 r)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_update(int n, const double* __restrict__ part_pq,
                                                          const double* __restrict__ dslot, T* x,
                                                          const T* __restrict__ p, T* r,
                                                          const T* __restrict__ q,
                                                          double* part_rz, SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_done(sc, sh)) return;
  const double pq = grid_sum(part_pq, sh);
  const double alpha = *dslot / pq;
  const T a = (T)alpha;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    x[i] = x[i] + a * p[i];
    r[i] = r[i] - a * q[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->alpha = alpha;
    sc->pq = pq;
  }
  (void)part_rz;
}

// partial r.z (after every r entry of the update is written: separate kernel
// for the block preconditioner, fused for Jacobi)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_rz(int n, const T* __restrict__ r,
                                                  const T* __restrict__ m,
                                                  const T* __restrict__ bm, double* part,
                                                  const SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_done(sc, sh)) return;
  double acc = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    acc += (double)r[i] * (double)apply_m(i, n, r, m, bm);
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Jacobi-fused variant of update + r.z (z_i needs only r_i)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_update_jacobi(
    int n, const double* __restrict__ part_pq, const double* __restrict__ dslot, T* x,
    const T* __restrict__ p, T* r, const T* __restrict__ q, const T* __restrict__ m,
    double* part_rz, SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_done(sc, sh)) return;
  const double pq = grid_sum(part_pq, sh);
  const double alpha = *dslot / pq;
  const T a = (T)alpha;
  double acc = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    x[i] = x[i] + a * p[i];
    const T ri = r[i] - a * q[i];
    r[i] = ri;
    acc += (double)ri * (double)(ri * m[i]);
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part_rz[blockIdx.x] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->alpha = alpha;
    sc->pq = pq;
  }
}

// di = r.z ; beta = di / d ; p = z + beta p ; bookkeeping in block 0
template <typename T>
__global__ __launch_bounds__(kThreads) void k_pcg_pupdate(int n, const double* __restrict__ part_rz,
                                                           const double* __restrict__ dold,
                                                           double* dnew, T* p,
                                                           const T* __restrict__ r,
                                                           const T* __restrict__ m,
                                                           const T* __restrict__ bm,
                                                           SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_done(sc, sh)) return;
  const double di = grid_sum(part_rz, sh);
  const double beta = di / *dold;
  const T bt = (T)beta;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = apply_m(i, n, r, m, bm) + bt * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *dnew = di;
    sc->beta = beta;
    sc->di = di;
    sc->iter += 1;
    if (sqrt(fabs(di)) < sc->conv) sc->done = 1;
  }
}

// ------------------------------------------------------------- host side
// Runs PCG on device-resident data.  Returns iterations in sc (device).
template <typename T>
int pcg_device(Context* c, const PcgBuffers<T>& B, double tol, int maxit, double* dslots) {
  hipStream_t s = c->stream;
  const int n = c->neq;
  const dim3 G(kRedBlocks), TB(kThreads);
  double* part_a = c->red64;
  double* part_b = c->red64 + kRedBlocks;
  hipLaunchKernelGGL(k_pcg_init<T>, G, TB, 0, s, n, B.b, B.m, B.bm, B.x, B.r, B.p, part_a);
  hipLaunchKernelGGL(k_pcg_init2, dim3(1), TB, 0, s, part_a, tol, dslots, c->scal);
  int it = 0;
  while (it < maxit) {
    int chunk = std::min(c->check_every, maxit - it);
    for (int j = 0; j < chunk; ++j, ++it) {
      double* dcur = dslots + (it & 1);
      double* dnxt = dslots + ((it + 1) & 1);
      launch_spmv_tiles<T>(c->tiles, B.ad, B.al, B.p, B.q, B.epart, s, &c->scal->done);
      hipLaunchKernelGGL(k_merge_dot<T>, G, TB, 0, s, n, c->tiles.inc_ptr, c->tiles.inc_idx,
                         B.epart, B.q, B.p, part_a, c->scal);
      if (B.bm == nullptr) {
        hipLaunchKernelGGL(k_pcg_update_jacobi<T>, G, TB, 0, s, n, part_a, dcur, B.x, B.p, B.r,
                           B.q, B.m, part_b, c->scal);
      } else {
        hipLaunchKernelGGL(k_pcg_update<T>, G, TB, 0, s, n, part_a, dcur, B.x, B.p, B.r, B.q,
                           part_b, c->scal);
        hipLaunchKernelGGL(k_rz<T>, G, TB, 0, s, n, B.r, B.m, B.bm, part_b, c->scal);
      }
      hipLaunchKernelGGL(k_pcg_pupdate<T>, G, TB, 0, s, n, part_b, dcur, dnxt, B.p, B.r, B.m,
                         B.bm, c->scal);
    }
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->scal_host, c->scal, sizeof(SolveScalars),
                                        hipMemcpyDeviceToHost, s));
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
    if (c->scal_host->done) break;
  }
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  return ELASTIC_OK;
}

template int pcg_device<double>(Context*, const PcgBuffers<double>&, double, int, double*);
template int pcg_device<float>(Context*, const PcgBuffers<float>&, double, int, double*);

}  // namespace elastic
