// Symmetric CSRC product by gathers only (Matvec.f:528-621), for patterns the
// row tiles cannot hold (a CSRC row longer than a tile).
//
// y_i = ad_i x_i + sum_k al_k x_ja(k)  (row i of the stored lower triangle, in
// CSRC order) + sum over the rows r > i that store column i of al x_r (in
// ascending r).  The second sum walks the transposed lists of the pattern:
// colptr (neq+1) and, per entry, its CSRC position colk and its row colrow,
// sorted by position within a column.  This is exactly the order in which
// matvec_csrcsym1 accumulates y(i): the row sum t is formed first, y(i) = t,
// then the later rows scatter into y(i) in ascending order.  With every
// product and sum rounded on its own (contraction off) the FP64 result is
// bit-identical to the reference's.  One lane per row; no partials, no
// atomics, no merge.
#include "elastic_internal.h"

namespace elastic {

template <typename TV>
__global__ __launch_bounds__(256) void k_spmv_gather(int n, const int32_t* __restrict__ ia,
                                                      const int32_t* __restrict__ ja,
                                                      const TV* __restrict__ ad,
                                                      const TV* __restrict__ al,
                                                      const int32_t* __restrict__ colptr,
                                                      const int32_t* __restrict__ colk,
                                                      const int32_t* __restrict__ colrow,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y,
                                                      const int32_t* __restrict__ skip) {
#pragma clang fp contract(off)
  if (skip != nullptr && *skip != 0) return;  // converged solve: uniform for the grid
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double t = (double)ad[i] * x[i];
    const int k1 = ia[i] - 1, k2 = ia[i + 1] - 1;
    for (int k = k1; k < k2; ++k) t = t + (double)al[k] * x[ja[k] - 1];
    const int c1 = colptr[i], c2 = colptr[i + 1];
    for (int q = c1; q < c2; ++q) t = t + (double)al[colk[q]] * x[colrow[q]];
    y[i] = t;
  }
}

template <typename TV>
void launch_spmv_gather(int n, const int32_t* ia, const int32_t* ja, const TV* ad, const TV* al,
                        const int32_t* colptr, const int32_t* colk, const int32_t* colrow,
                        const double* x, double* y, hipStream_t s, const int32_t* skip) {
  const int g = std::max(1, std::min((n + 255) / 256, 16384));
  hipLaunchKernelGGL(k_spmv_gather<TV>, dim3(g), dim3(256), 0, s, n, ia, ja, ad, al, colptr, colk,
                     colrow, x, y, skip);
}

template void launch_spmv_gather<double>(int, const int32_t*, const int32_t*, const double*,
                                         const double*, const int32_t*, const int32_t*,
                                         const int32_t*, const double*, double*, hipStream_t,
                                         const int32_t*);
template void launch_spmv_gather<float>(int, const int32_t*, const int32_t*, const float*,
                                        const float*, const int32_t*, const int32_t*,
                                        const int32_t*, const double*, double*, hipStream_t,
                                        const int32_t*);

void build_gather(int neq, const int32_t* ia1, const int32_t* ja1, GatherLayout& L) {
  const int64_t nad = (int64_t)ia1[neq] - 1;
  L.colptr.assign((size_t)neq + 1, 0);
  for (int64_t k = 0; k < nad; ++k) L.colptr[ja1[k]]++;  // column ja1[k]-1 counts at +1
  for (int i = 0; i < neq; ++i) L.colptr[i + 1] += L.colptr[i];
  L.colk.resize(nad);
  L.colrow.resize(nad);
  std::vector<int32_t> fill(L.colptr.begin(), L.colptr.end() - 1);
  L.longest = 0;
  for (int r = 0; r < neq; ++r) {  // rows ascending: each column list ends up sorted
    L.longest = std::max(L.longest, ia1[r + 1] - ia1[r]);
    for (int64_t k = ia1[r] - 1; k < ia1[r + 1] - 1; ++k) {
      const int32_t p = fill[ja1[k] - 1]++;
      L.colk[p] = (int32_t)k;
      L.colrow[p] = r;
    }
  }
}

}  // namespace elastic
