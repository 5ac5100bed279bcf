// Vector kernels around the Krylov loops: precision casts, the Jacobi
// inverse, the post-solve statistics of Iters.f:435-457 and the FP64 steps of
// the mixed-precision iterative refinement.  All reductions use the fixed
// two-stage order of pcg.hip.
#include "elastic_internal.h"

namespace elastic {

constexpr int kT = 256;

__device__ inline double blk_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

template <typename S, typename D>
__global__ __launch_bounds__(kT) void k_cast(int64_t n, const S* __restrict__ a, D* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    b[i] = (D)a[i];
}

template <typename T>
__global__ __launch_bounds__(kT) void k_recip(int n, const T* __restrict__ ad, T* __restrict__ m) {
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) m[i] = T(1) / ad[i];
}

// y += merged partials (when mg.ptr != null); then 6 partial sums:
// x.y, x.x, (b-y).m(b-y), (b-y).(b-y), b.b, b.(m b)
__global__ __launch_bounds__(kT) void k_post(int n, MergeArgs mg,
                                              const double* __restrict__ x, double* y,
                                              const double* __restrict__ b,
                                              const double* __restrict__ m, double* part) {
  __shared__ double sh[8];
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    double s = y[i];
    s = merge_row(mg, i, s);
    y[i] = s;
    const double xi = x[i], bi = b[i], ri = bi - s;
    a[0] += xi * s;
    a[1] += xi * xi;
    a[2] += ri * (ri * m[i]);
    a[3] += ri * ri;
    a[4] += bi * bi;
    a[5] += bi * (bi * m[i]);
  }
  for (int j = 0; j < 6; ++j) {
    const double v = blk_sum(a[j], sh);
    if (threadIdx.x == 0) part[j * kRedBlocks + blockIdx.x] = v;
  }
}

// vectors between the CSRC node order and the compact layout order (one lane
// per layout node and component)
__global__ __launch_bounds__(kT) void k_permute(int nn, int nb, const int32_t* __restrict__ d2r,
                                                 const double* __restrict__ src,
                                                 double* __restrict__ dst, int to_layout) {
  const int64_t n = (int64_t)nn * nb;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const int d = (int)(i / nb), r = (int)(i - (int64_t)d * nb);
    const int64_t j = (int64_t)nb * d2r[d] + r;
    if (to_layout) dst[i] = src[j];
    else dst[j] = src[i];
  }
}

// out[j] = sum of part[j*kRedBlocks ...] for j < nsum (one block)
__global__ __launch_bounds__(kT) void k_sum_parts(const double* __restrict__ part, int nsum,
                                                   double* out) {
  __shared__ double sh[8];
  for (int j = 0; j < nsum; ++j) {
    double v = 0.0;
    for (int q = threadIdx.x; q < kRedBlocks; q += kT) v += part[j * kRedBlocks + q];
    v = blk_sum(v, sh);
    if (threadIdx.x == 0) out[j] = v;
  }
}


// ------------------------------------------------------------- launchers
static int grid_for(int64_t n) {
  int64_t g = (n + kT - 1) / kT;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

template <typename S, typename D>
void launch_cast(int64_t n, const S* a, D* b, hipStream_t s) {
  hipLaunchKernelGGL((k_cast<S, D>), dim3(grid_for(n)), dim3(kT), 0, s, n, a, b);
}
template void launch_cast<double, float>(int64_t, const double*, float*, hipStream_t);
template void launch_cast<float, double>(int64_t, const float*, double*, hipStream_t);
template void launch_cast<double, double>(int64_t, const double*, double*, hipStream_t);

template <typename T>
void launch_recip(int n, const T* ad, T* m, hipStream_t s) {
  hipLaunchKernelGGL(k_recip<T>, dim3(grid_for(n)), dim3(kT), 0, s, n, ad, m);
}
template void launch_recip<double>(int, const double*, double*, hipStream_t);
template void launch_recip<float>(int, const float*, float*, hipStream_t);

void launch_post(const MergeArgs& mg, int n, const double* x, double* y, const double* b,
                 const double* m, double* part, double* out6, hipStream_t s) {
  hipLaunchKernelGGL(k_post, dim3(kRedBlocks), dim3(kT), 0, s, n, mg, x, y, b, m, part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kT), 0, s, part, 6, out6);
}

void launch_permute(int nn, int nb, const int32_t* d2r, const double* src, double* dst,
                    bool to_layout, hipStream_t s) {
  hipLaunchKernelGGL(k_permute, dim3(grid_for((int64_t)nn * nb)), dim3(kT), 0, s, nn, nb, d2r, src,
                     dst, to_layout ? 1 : 0);
}

void launch_sum_parts(const double* part, int nsum, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kT), 0, s, part, nsum, out);
}


}  // namespace elastic
