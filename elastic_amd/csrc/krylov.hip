// 3x3 block-Jacobi set-up on the device and the non-PCG Krylov methods.
#include <cmath>

#include "elastic_internal.h"


namespace elastic {

// Block-Jacobi inverse of consecutive equation triples (Precond.f:163-358,
// nblock = 3, with the cofactor inverse of Precond.f:1676-1710).  slots holds
// the al positions of a21, a31, a32 of each triple (-1 when absent).
template <typename T>
__global__ void k_bdiag(int nb, const T* __restrict__ ad, const T* __restrict__ al,
                        const int64_t* __restrict__ slots, T* __restrict__ bm) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
    const int ii = 3 * i;
    const double a11 = ad[ii], a22 = ad[ii + 1], a33 = ad[ii + 2];
    const int64_t s21 = slots[3 * i], s31 = slots[3 * i + 1], s32 = slots[3 * i + 2];
    const double a21 = s21 >= 0 ? (double)al[s21] : 0.0;
    const double a31 = s31 >= 0 ? (double)al[s31] : 0.0;
    const double a32 = s32 >= 0 ? (double)al[s32] : 0.0;
    const double c11 = a22 * a33 - a32 * a32;
    const double c22 = a11 * a33 - a31 * a31;
    const double c33 = a11 * a22 - a21 * a21;
    const double c21 = a31 * a32 - a21 * a33;
    const double c31 = a21 * a32 - a31 * a22;
    const double c32 = a31 * a21 - a11 * a32;
    double det = a11 * a22 * a33 + a21 * a32 * a31 + a31 * a21 * a32 -
                 (a22 * a31 * a31 + a11 * a32 * a32 + a33 * a21 * a21);
    det = 1.0 / det;
    T* o = bm + 6 * i;
    o[0] = (T)(c11 * det);
    o[1] = (T)(c22 * det);
    o[2] = (T)(c33 * det);
    o[3] = (T)(c21 * det);
    o[4] = (T)(c31 * det);
    o[5] = (T)(c32 * det);
  }
}

int ensure_bdiag(elastic_ctx* x) {
  Context* c = &x->c;
  if (c->bdiag_set) return ELASTIC_OK;
  const int nb = c->neq / 3;
  if (nb == 0) return set_error(c, ELASTIC_BAD_ARG, "block preconditioner needs neq >= 3");
  std::vector<int64_t> slots(3 * (size_t)nb, -1);
  const int32_t* ia = x->ia.data();
  const int32_t* ja = x->ja.data();
  for (int i = 0; i < nb; ++i) {
    const int ii = 3 * i + 1;  // 1-based first equation of the triple
    for (int64_t k = ia[ii] - 1; k < ia[ii + 1] - 1; ++k)
      if (ja[k] == ii) slots[3 * i] = k;
    for (int64_t k = ia[ii + 1] - 1; k < ia[ii + 2] - 1; ++k) {
      if (ja[k] == ii) slots[3 * i + 1] = k;
      else if (ja[k] == ii + 1) slots[3 * i + 2] = k;
    }
  }
  int64_t* dslots = nullptr;
  ELASTIC_HIP_CHECK(c, hipMalloc((void**)&dslots, sizeof(int64_t) * slots.size()));
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(dslots, slots.data(), sizeof(int64_t) * slots.size(),
                                      hipMemcpyHostToDevice, c->stream));
  if (!c->A64.bm) ELASTIC_HIP_CHECK(c, hipMalloc((void**)&c->A64.bm, sizeof(double) * 6 * nb));
  int g = std::min((nb + 255) / 256, 4096);
  hipLaunchKernelGGL(k_bdiag<double>, dim3(g), dim3(256), 0, c->stream, nb, c->A64.ad, c->A64.al,
                     dslots, c->A64.bm);
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  hipFree(dslots);
  c->bdiag_set = true;
  return ELASTIC_OK;
}

int solve_krylov_other(Context* c, int solver, int precond, int precision, double tol, int maxit,
                       int ngram, elastic_stats* st) {
  (void)precond;
  (void)precision;
  (void)tol;
  (void)maxit;
  (void)ngram;
  (void)st;
  return set_error(c, ELASTIC_BAD_ARG, "solver " + std::to_string(solver) + " not available");
}

}  // namespace elastic
