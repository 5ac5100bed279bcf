// 3x3 block-Jacobi set-up on the device, BiCGSTAB and GMRES(k).
#include <cmath>

#include "device_util.h"
#include "elastic_internal.h"


namespace elastic {

// Block-Jacobi inverse of consecutive equation triples (Precond.f:163-358,
// nblock = 3, with the cofactor inverse of Precond.f:1676-1710).  slots holds
// the al positions of a21, a31, a32 of each triple (-1 when absent).
template <typename T>
__global__ void k_bdiag(int nb, const T* __restrict__ ad, const T* __restrict__ al,
                        const int64_t* __restrict__ slots, T* __restrict__ bm) {
  // every product and difference rounded on its own, as the reference's
  // x86-64 build evaluates them: the inverse is bit-identical to block_precond's
#pragma clang fp contract(off)
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

// 2x2 inverses of consecutive equation pairs (Precond.f:163-358, nblock = 2):
// (a22, a11, -a21) / det per pair; slot holds the al position of a21 (-1
// when absent).
__global__ void k_bdiag2(int nb, const double* __restrict__ ad, const double* __restrict__ al,
                         const int64_t* __restrict__ slots, double* __restrict__ bm) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
    const double a11 = ad[2 * i], a22 = ad[2 * i + 1];
    const double a21 = slots[i] >= 0 ? al[slots[i]] : 0.0;
    const double idet = 1.0 / (a11 * a22 - a21 * a21);
    bm[3 * i] = idet * a22;
    bm[3 * i + 1] = idet * a11;
    bm[3 * i + 2] = -idet * a21;
  }
}

// The block-Jacobi inverse of block size bsz (2 or 3) in c->A64.bm: 3 values
// per pair or 6 per triple, the reference's m(., i) of block_precond.
int ensure_bdiag(elastic_ctx* x, int bsz) {
  Context* c = &x->c;
  if (bsz != 2 && bsz != 3) return set_error(c, ELASTIC_BAD_ARG, "block size must be 2 or 3");
  if (c->bdiag_set && c->bdiag_nb == bsz) return ELASTIC_OK;
  const int nb = c->neq / bsz;
  if (nb == 0) return set_error(c, ELASTIC_BAD_ARG, "block preconditioner needs neq >= block size");
  const int per = bsz == 3 ? 3 : 1;  // a21 (a31, a32) positions per block
  std::vector<int64_t> slots((size_t)per * nb, -1);
  const int32_t* ia = x->ia.data();
  const int32_t* ja = x->ja.data();
  for (int i = 0; i < nb; ++i) {
    const int ii = bsz * i + 1;  // 1-based first equation of the block
    for (int64_t k = ia[ii] - 1; k < ia[ii + 1] - 1; ++k)
      if (ja[k] == ii) slots[(size_t)per * i] = k;
    if (bsz == 3)
      for (int64_t k = ia[ii + 1] - 1; k < ia[ii + 2] - 1; ++k) {
        if (ja[k] == ii) slots[3 * (size_t)i + 1] = k;
        else if (ja[k] == ii + 1) slots[3 * (size_t)i + 2] = k;
      }
  }
  int64_t* dslots = nullptr;
  ELASTIC_HIP_CHECK(c, hipMalloc((void**)&dslots, sizeof(int64_t) * slots.size()));
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(dslots, slots.data(), sizeof(int64_t) * slots.size(),
                                      hipMemcpyHostToDevice, c->stream));
  // room for either size: 6 per triple or 3 per pair
  if (!c->A64.bm)
    ELASTIC_HIP_CHECK(c, hipMalloc((void**)&c->A64.bm,
                                   sizeof(double) * std::max(6 * (c->neq / 3), 3 * (c->neq / 2))));
  int g = std::min((nb + 255) / 256, 4096);
  if (bsz == 3)
    hipLaunchKernelGGL(k_bdiag<double>, dim3(g), dim3(256), 0, c->stream, nb, c->A64.ad,
                       c->A64.al, dslots, c->A64.bm);
  else
    hipLaunchKernelGGL(k_bdiag2, dim3(g), dim3(256), 0, c->stream, nb, c->A64.ad, c->A64.al,
                       dslots, c->A64.bm);
  // compact layouts apply the 3x3 inverse in their node order (node blocks
  // of 6 values move with their nodes); A64.bm keeps the CSRC order
  if (c->bm_dev) hipFree(c->bm_dev);
  c->bm_dev = nullptr;
  if (bsz == 3 && c->blk.d2r && c->blk.nb == 3) {
    ELASTIC_HIP_CHECK(c, hipMalloc((void**)&c->bm_dev, sizeof(double) * 6 * (size_t)c->blk.nn));
    launch_permute(c->blk.nn, 6, c->blk.d2r, c->A64.bm, c->bm_dev, true, c->stream);
  }
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  hipFree(dslots);
  c->bdiag_set = true;
  c->bdiag_nb = bsz;
  return ELASTIC_OK;
}

}  // namespace elastic

// m(w, neq/bsz + mod(neq,bsz)) of block_precond (Precond.f:163-358), w = 6
// (bsz 3) or 3 (bsz 2): the device's block inverse, then one 1/ad entry per
// remaining equation.
extern "C" int elastic_get_block_precond(elastic_ctx* x, int32_t bsz, double* m) {
  using namespace elastic;
  if (!x || !m) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!c->values_set) return set_error(c, ELASTIC_NOT_READY, "values not set");
  hipSetDevice(c->device);
  int rc = ensure_bdiag(x, bsz);
  if (rc) return rc;
  const int w = bsz == 3 ? 6 : 3;
  const int nb = c->neq / bsz, rest = c->neq % bsz;
  std::vector<double> inv(rest > 0 ? rest : 1);
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(m, c->A64.bm, sizeof(double) * w * (size_t)nb,
                                      hipMemcpyDeviceToHost, c->stream));
  // 1/ad of the remaining equations from the CSRC-order diagonal (A64.m may
  // be in a compact layout's order); IEEE division, as the device's
  if (rest)
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(inv.data(), c->A64.ad + (size_t)bsz * nb,
                                        sizeof(double) * rest, hipMemcpyDeviceToHost, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < rest; ++i) {
    double* o = m + (size_t)w * ((size_t)nb + i);
    o[0] = 1.0 / inv[i];
    for (int k = 1; k < w; ++k) o[k] = 0.0;
  }
  return ELASTIC_OK;
}

extern "C" int elastic_get_bdiag(elastic_ctx* x, double* m) {
  return elastic_get_block_precond(x, 3, m);
}

namespace elastic {

// ================================================================ BiCGSTAB
// Iters.f:1079-1349: right-preconditioned BiCGSTAB from x0 = 0 with shadow
// residual r0 = b.  Stop when ||r|| < tol ||b|| at either half step
// (Iters.f:1226-1228, 1251-1253).  Vectors: x VX, r VR, p VP, v VQ, b VB,
// t VT, z VU, r0 VW.  The scalars r.r0 of consecutive iterations ping-pong
// between two device slots so that no kernel reads a slot another block of
// the same launch writes.

// x = 0; r = r0 = p = b; z = M b (Jacobi, else applied after); partial b.b
__global__ __launch_bounds__(kThreads) void k_bi_init(int n, const double* __restrict__ b,
                                                       const double* __restrict__ m, int jacobi,
                                                       double* x, double* r, double* r0,
                                                       double* p, double* z, double* part) {
  __shared__ double sh[8];
  double acc = 0.0;
  GRID_LOOP(i, n) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    r0[i] = bi;
    p[i] = bi;
    if (jacobi) z[i] = bi * m[i];
    acc += bi * bi;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_bi_init2(const double* __restrict__ part, double tol, double delta, int reliable,
                           double* rr, SolveScalars* sc) {
  __shared__ double sh[8];
  const double d = grid_sum(part, sh);
  if (threadIdx.x == 0) {
    rr[0] = d;  // r.r0 = b.b
    sc->conv = tol * sqrt(fabs(d));
    sc->norm_b = sqrt(fabs(d));
    sc->rmax = sqrt(fabs(d));
    sc->delta = delta;
    sc->iter = 0;
    sc->status = 0;
    sc->reliable = reliable;
    sc->upd = 0;
    sc->skip64 = 1;
    sc->nupd = 0;
    progress_init(sc, sqrt(fabs(d)));
    sc->done = (d == 0.0) ? 1 : 0;
  }
}

// z = M^-1 y for the block preconditioner (a pass of its own: a row of z
// reads the whole equation triple of y); gate: extra flag that must be set
__global__ __launch_bounds__(kThreads) void k_apply_m(int n, const double* __restrict__ y,
                                                       const double* __restrict__ m,
                                                       const double* __restrict__ bm, double* z,
                                                       const int32_t* gate,
                                                       const SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  if (gate && !block_flag(gate, sh)) return;
  GRID_LOOP(i, n) z[i] = apply_m(i, n, y, m, bm);
}

// y += merged partials; partial y.a (and y.a2 when a2 != null)
__global__ __launch_bounds__(kThreads) void k_merge_dot2(int n, MergeArgs mg,
                                                          double* y, const double* a, double* part_a,
                                                          const double* a2, double* part_b,
                                                          const SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  double s1 = 0.0, s2 = 0.0;
  GRID_LOOP(i, n) {
    double s = y[i];
    s = merge_row(mg, i, s);
    y[i] = s;
    s1 += s * a[i];
    if (a2) s2 += s * (a2 == y ? s : a2[i]);
  }
  const double t1 = block_sum(s1, sh);
  if (threadIdx.x == 0) part_a[blockIdx.x] = t1;
  if (a2) {
    const double t2 = block_sum(s2, sh);
    if (threadIdx.x == 0) part_b[blockIdx.x] = t2;
  }
}

// alpha = r.r0 / v.r0 ; x += alpha z ; r -= alpha v ; z = M r (Jacobi); partial r.r
__global__ __launch_bounds__(kThreads) void k_bi_half1(int n, const double* __restrict__ part_vr0,
                                                        const double* __restrict__ rr_cur,
                                                        double* x, double* z, double* r,
                                                        const double* __restrict__ v,
                                                        const double* __restrict__ m, int jacobi,
                                                        double* part, SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  const double vr0 = grid_sum(part_vr0, sh);
  // breakdown (every block sees the same sums): a zero or non-finite
  // denominator, or a non-finite r.r0 carried from the previous iteration
  if (vr0 == 0.0 || !isfinite(vr0) || !isfinite(*rr_cur)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->status = ELASTIC_BREAKDOWN;
      sc->done = 1;
    }
    return;
  }
  const double alpha = *rr_cur / vr0;
  double acc = 0.0;
  GRID_LOOP(i, n) {
    x[i] = x[i] + alpha * z[i];
    const double ri = r[i] - alpha * v[i];
    r[i] = ri;
    if (jacobi) z[i] = ri * m[i];
    acc += ri * ri;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha = alpha;
}

// omega = t.r / t.t ; x += omega z ; r -= omega t ; partials r.r and r.r0
__global__ __launch_bounds__(kThreads) void k_bi_half2(int n, const double* __restrict__ part_tr,
                                                        const double* __restrict__ part_tt,
                                                        double* x, const double* __restrict__ z,
                                                        double* r, const double* __restrict__ t,
                                                        const double* __restrict__ r0,
                                                        double* part_rr, double* part_rr0,
                                                        SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  const double tr = grid_sum(part_tr, sh);
  const double tt = grid_sum(part_tt, sh);
  if (tt == 0.0 || !isfinite(tt) || !isfinite(tr)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->status = ELASTIC_BREAKDOWN;
      sc->done = 1;
    }
    return;
  }
  const double w = tr / tt;
  double a1 = 0.0, a2 = 0.0;
  GRID_LOOP(i, n) {
    x[i] = x[i] + w * z[i];
    const double ri = r[i] - w * t[i];
    r[i] = ri;
    a1 += ri * ri;
    a2 += ri * r0[i];
  }
  const double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = s1;
  const double s2 = block_sum(a2, sh);
  if (threadIdx.x == 0) part_rr0[blockIdx.x] = s2;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->omega = w;
}

// convergence test on ||r|| (one block).  half = 1 counts the iteration;
// half = 2 also stores r.r0 for the next iteration.  With reliable updates
// the test only requests a true-residual replacement.
__global__ void k_bi_check(const double* __restrict__ part_rr, const double* __restrict__ part_rr0,
                           int half, double* rr_next, SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  const double d = grid_sum(part_rr, sh);
  double rr0 = 0.0;
  if (half == 2) rr0 = grid_sum(part_rr0, sh);
  if (threadIdx.x == 0) {
    if (half == 1) sc->iter += 1;
    if (half == 2) *rr_next = rr0;
    if (!isfinite(d) || (half == 2 && !isfinite(rr0))) {
      sc->status = ELASTIC_BREAKDOWN;
      sc->done = 1;
      return;
    }
    const double rn = sqrt(fabs(d));
    if (!sc->reliable) {
      if (rn < sc->conv) sc->done = 1;
    } else if (rn < sc->conv || rn < sc->delta * sc->rmax ||
               sc->iter - sc->last_upd >= kForceReplace) {
      sc->upd = 1;
      sc->skip64 = 0;
    }
  }
}

// true residual: r = b - (t + merged partials); z = M r (Jacobi) when asked;
// partials r.r and r.r0
__global__ __launch_bounds__(kThreads) void k_bi_rresid(int n, MergeArgs mg,
                                                         const double* __restrict__ t,
                                                         const double* __restrict__ b, double* r,
                                                         double* z, const double* __restrict__ m,
                                                         const double* __restrict__ r0,
                                                         double* part_rr, double* part_rr0,
                                                         const SolveScalars* sc) {
  __shared__ double sh[8];
  if (!block_flag(&sc->upd, sh) || block_flag(&sc->done, sh)) return;
  double a1 = 0.0, a2 = 0.0;
  GRID_LOOP(i, n) {
    double s = t[i];
    s = merge_row(mg, i, s);
    const double ri = b[i] - s;
    r[i] = ri;
    if (z) z[i] = ri * m[i];
    a1 += ri * ri;
    a2 += ri * r0[i];
  }
  const double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = s1;
  const double s2 = block_sum(a2, sh);
  if (threadIdx.x == 0) part_rr0[blockIdx.x] = s2;
}

__global__ void k_bi_rfinish(const double* __restrict__ part_rr, const double* __restrict__ part_rr0,
                             double* rr_next, SolveScalars* sc) {
  __shared__ double sh[8];
  if (!block_flag(&sc->upd, sh) || block_flag(&sc->done, sh)) return;
  const double d = grid_sum(part_rr, sh);
  const double rr0 = grid_sum(part_rr0, sh);
  if (threadIdx.x == 0) {
    sc->nupd += 1;
    sc->true_rmr = d;
    progress_guard(sc, sqrt(fabs(d)));
    if (sqrt(fabs(d)) < sc->conv) {
      sc->done = 1;
    } else {
      if (rr_next) *rr_next = rr0;
      sc->rmax = sqrt(fabs(d));
      sc->upd = 0;
      sc->skip64 = 1;
    }
  }
}

// beta = (r.r0 new / r.r0) (alpha / omega) ; p = r + beta (p - omega v) ; z = M p (Jacobi)
__global__ __launch_bounds__(kThreads) void k_bi_pupdate(int n, const double* __restrict__ rr_cur,
                                                          const double* __restrict__ rr_next,
                                                          double* p, const double* __restrict__ r,
                                                          const double* __restrict__ v, double* z,
                                                          const double* __restrict__ m, int jacobi,
                                                          SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  const double w = sc->omega;
  const double beta = (*rr_next / *rr_cur) * (sc->alpha / w);
  GRID_LOOP(i, n) {
    const double pi = r[i] + beta * (p[i] - w * v[i]);
    p[i] = pi;
    if (jacobi) z[i] = pi * m[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->beta = beta;
}

int bicgstab_run(Context* c, bool bdiag, bool fp32_operator, double tol, int maxit, double delta,
                 KrylovResult* out) {
  hipStream_t s = c->stream;
  const int n = c->neq;
  const dim3 G(kRedBlocks), TB(kThreads);
  double* P[6];
  for (int k = 0; k < 6; ++k) P[k] = c->red64 + (size_t)k * kRedBlocks;
  double* rr = c->red64 + 8 * kRedBlocks + 16;  // two ping-pong slots
  double *x = c->v64[VX], *r = c->v64[VR], *p = c->v64[VP], *v = c->v64[VQ], *b = c->v64[VB],
         *t = c->v64[VT], *z = c->v64[VU], *r0 = c->v64[VW];
  const double* m = c->A64.m;
  const double* bm = bdiag ? solver_bm(c) : nullptr;
  const int jac = bm == nullptr;
  const bool reliable = fp32_operator;
  SolveScalars* sc = c->scal;
  hipLaunchKernelGGL(k_bi_init, G, TB, 0, s, n, b, m, jac, x, r, r0, p, z, P[0]);
  hipLaunchKernelGGL(k_bi_init2, dim3(1), TB, 0, s, P[0], tol, delta, reliable ? 1 : 0, rr, sc);
  if (!jac) hipLaunchKernelGGL(k_apply_m, G, TB, 0, s, n, p, m, bm, z, nullptr, sc);
  // true-residual replacement (reliable mode), gated on sc->upd
  auto replace = [&](bool with_z, double* rr_next) {
    launch_spmv(c, false, x, t, c->epart64, &sc->skip64);
    hipLaunchKernelGGL(k_bi_rresid, G, TB, 0, s, n, merge_of(c, c->epart64), t, b, r,
                       (with_z && jac) ? z : nullptr, m, r0, P[4], P[5], sc);
    if (with_z && !jac) hipLaunchKernelGGL(k_apply_m, G, TB, 0, s, n, r, m, bm, z, &sc->upd, sc);
    hipLaunchKernelGGL(k_bi_rfinish, dim3(1), TB, 0, s, P[4], P[5], rr_next, sc);
  };
  int it = 0;
  const int every = std::max(1, c->check_every);
  bool done = false;
  bool op32 = fp32_operator;  // the progress guard may hand over to the FP64 operator
  while (it < maxit && !done) {
    const int chunk = std::min(every, maxit - it);
    for (int j = 0; j < chunk; ++j, ++it) {
      double* rr_cur = rr + (it & 1);
      double* rr_next = rr + ((it + 1) & 1);
      launch_spmv(c, op32, z, v, c->epart64, &sc->done);
      hipLaunchKernelGGL(k_merge_dot2, G, TB, 0, s, n, merge_of(c, c->epart64, epart_f32(c, op32)),
                         v, r0, P[0], (const double*)nullptr, P[1], sc);
      hipLaunchKernelGGL(k_bi_half1, G, TB, 0, s, n, P[0], rr_cur, x, z, r, v, m, jac, P[1], sc);
      if (!jac) hipLaunchKernelGGL(k_apply_m, G, TB, 0, s, n, r, m, bm, z, nullptr, sc);
      hipLaunchKernelGGL(k_bi_check, dim3(1), TB, 0, s, P[1], P[1], 1, rr_next, sc);
      if (reliable) replace(true, nullptr);
      launch_spmv(c, op32, z, t, c->epart64, &sc->done);
      hipLaunchKernelGGL(k_merge_dot2, G, TB, 0, s, n, merge_of(c, c->epart64, epart_f32(c, op32)),
                         t, r, P[2], (const double*)t, P[3], sc);
      hipLaunchKernelGGL(k_bi_half2, G, TB, 0, s, n, P[2], P[3], x, z, r, t, r0, P[4], P[5], sc);
      hipLaunchKernelGGL(k_bi_check, dim3(1), TB, 0, s, P[4], P[5], 2, rr_next, sc);
      if (reliable) replace(false, rr_next);
      hipLaunchKernelGGL(k_bi_pupdate, G, TB, 0, s, n, rr_cur, rr_next, p, r, v, z, m, jac, sc);
      if (!jac) hipLaunchKernelGGL(k_apply_m, G, TB, 0, s, n, p, m, bm, z, nullptr, sc);
    }
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->scal_host, sc, sizeof(SolveScalars),
                                        hipMemcpyDeviceToHost, s));
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
    done = c->scal_host->done != 0;
    if (c->scal_host->op64) op32 = false;
  }
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  const SolveScalars& h = *c->scal_host;
  out->iters = done ? h.iter : maxit;
  out->nupd = h.nupd;
  out->outer = h.nupd;
  out->done = done && h.status == 0;
  out->status = h.status ? h.status : (done ? ELASTIC_OK : ELASTIC_MAXIT);
  out->conv = h.conv;
  out->op64_from = h.op64_from;
  return ELASTIC_OK;
}

// ================================================================== GMRES
// Iters.f:528-808: restarted GMRES(k), left preconditioner M, x0 = 0,
// econv = tol ||M b||.  Arnoldi with modified Gram-Schmidt (one fused pass per
// basis vector: w -= h(j,i) g_j together with the partial dot for g_{j+1}),
// Givens rotations and the back substitution on the device (one block), the
// x update as one pass over the cycle's basis.  The host reads |e(i)| after
// every inner step.  Small arrays (column-major like the reference): h
// (k+1) x k, c(k), s(k), e(k+1), y(k).

struct GmSmall {
  double *h, *c, *s, *e, *y;
};
static GmSmall gm_small(double* base, int k) {
  return GmSmall{base, base + (size_t)(k + 1) * k, base + (size_t)(k + 1) * k + k,
                 base + (size_t)(k + 1) * k + 2 * k, base + (size_t)(k + 1) * k + 3 * k + 1};
}

// g = M u (u given) with partial g.g
__global__ __launch_bounds__(kThreads) void k_gm_applym_dot(int n, const double* __restrict__ u,
                                                             const double* __restrict__ m,
                                                             const double* __restrict__ bm,
                                                             double* g, double* part) {
  __shared__ double sh[8];
  double acc = 0.0;
  GRID_LOOP(i, n) {
    const double gi = apply_m(i, n, u, m, bm);
    g[i] = gi;
    acc += gi * gi;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// u = b - (t + merged partials)
__global__ __launch_bounds__(kThreads) void k_gm_resid(int n, MergeArgs mg,
                                                        const double* __restrict__ t,
                                                        const double* __restrict__ b, double* u) {
  GRID_LOOP(i, n) {
    double s = t[i];
    s = merge_row(mg, i, s);
    u[i] = b[i] - s;
  }
}

// u = w + merged partials (the raw product; M is applied by k_gm_mdot)
__global__ __launch_bounds__(kThreads) void k_gm_merge(int n, MergeArgs mg,
                                                        const double* __restrict__ w, double* u) {
  GRID_LOOP(i, n) {
    double s = w[i];
    s = merge_row(mg, i, s);
    u[i] = s;
  }
}

// w = M u ; partial w.g1
__global__ __launch_bounds__(kThreads) void k_gm_mdot(int n, const double* __restrict__ u,
                                                       const double* __restrict__ m,
                                                       const double* __restrict__ bm, double* w,
                                                       const double* __restrict__ g1,
                                                       double* part) {
  __shared__ double sh[8];
  double acc = 0.0;
  GRID_LOOP(i, n) {
    const double wi = apply_m(i, n, u, m, bm);
    w[i] = wi;
    acc += wi * g1[i];
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// one MGS step: beta = w.g_j (from the partials); w -= beta g_j; h(j,i) = beta;
// partial w.g_{j+1} (gnext) or w.w (gnext == null)
__global__ __launch_bounds__(kThreads) void k_gm_mgs(int n, const double* __restrict__ part_in,
                                                      double* w, const double* __restrict__ gj,
                                                      const double* __restrict__ gnext,
                                                      double* part_out, double* hji) {
  __shared__ double sh[8];
  const double beta = grid_sum(part_in, sh);
  double acc = 0.0;
  GRID_LOOP(i, n) {
    const double wi = w[i] - beta * gj[i];
    w[i] = wi;
    acc += wi * (gnext ? gnext[i] : wi);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) *hji = beta;
}

// e(0) = ||g1||: from the partials; sc->gm_h = e(0)
__global__ void k_gm_norm0(const double* __restrict__ part, GmSmall a, SolveScalars* sc) {
  __shared__ double sh[8];
  const double d = grid_sum(part, sh);
  if (threadIdx.x == 0) {
    a.e[0] = sqrt(d);
    sc->gm_h = a.e[0];
    sc->gm_e = a.e[0];
  }
}

// inner step i (1-based): h(i+1,i) = ||w||, previous rotations on column i,
// new rotation, e update (Iters.f:691-704)
__global__ void k_gm_givens(const double* __restrict__ part, int i, int k, GmSmall a,
                            SolveScalars* sc) {
  __shared__ double sh[8];
  const double d = grid_sum(part, sh);
  if (threadIdx.x == 0) {
    double* H = a.h + (size_t)(i - 1) * (k + 1);  // column i: H[j-1] = h(j,i)
    const double nrm = sqrt(d);
    H[i] = nrm;
    for (int j = 1; j <= i - 1; ++j) {
      const double a1 = a.c[j - 1] * H[j - 1] + a.s[j - 1] * H[j];
      const double a2 = -a.s[j - 1] * H[j - 1] + a.c[j - 1] * H[j];
      H[j - 1] = a1;
      H[j] = a2;
    }
    const double r = sqrt(H[i - 1] * H[i - 1] + H[i] * H[i]);
    a.c[i - 1] = H[i - 1] / r;
    a.s[i - 1] = H[i] / r;
    H[i - 1] = r;
    H[i] = 0.0;
    a.e[i] = -a.s[i - 1] * a.e[i - 1];
    a.e[i - 1] = a.c[i - 1] * a.e[i - 1];
    sc->gm_e = fabs(a.e[i]);
    sc->gm_h = nrm;
  }
}

// g = g / *den (den = 0: left unchanged)
__global__ __launch_bounds__(kThreads) void k_gm_scale(int n, double* g, const double* den) {
  const double d = *den;
  if (d == 0.0) return;
  GRID_LOOP(i, n) g[i] = g[i] / d;
}

// back substitution for y(1..ni) (Iters.f:710-717)
__global__ void k_gm_backsolve(int ni, int k, GmSmall a) {
  if (threadIdx.x != 0) return;
  auto H = [&](int r, int c) { return a.h[(size_t)(c - 1) * (k + 1) + (r - 1)]; };
  a.y[ni - 1] = a.e[ni - 1] / H(ni, ni);
  for (int i = ni - 1; i >= 1; --i) {
    double yi = 0.0;
    for (int j = i + 1; j <= ni; ++j) yi = yi - H(i, j) * a.y[j - 1];
    a.y[i - 1] = (yi + a.e[i - 1]) / H(i, i);
  }
}

// x += sum_j y(j) g_j (Iters.f:721-725), j in reference order
__global__ __launch_bounds__(kThreads) void k_gm_xupdate(int n, int ni, const double* __restrict__ g,
                                                          const double* __restrict__ y, double* x) {
  GRID_LOOP(i, n) {
    double xi = x[i];
    for (int j = 0; j < ni; ++j) xi = xi + y[j] * g[(size_t)j * n + i];
    x[i] = xi;
  }
}

int gmres_run(Context* c, bool bdiag, bool fp32_operator, double tol, int maxit, int k,
              KrylovResult* out) {
  hipStream_t s = c->stream;
  const int n = c->neq;
  if (k < 1) k = 50;
  if (k > 1000) return set_error(c, ELASTIC_BAD_ARG, "ngram must be <= 1000");
  if (c->gm_k != k) {
    if (c->gm_basis) hipFree(c->gm_basis);
    if (c->gm_small) hipFree(c->gm_small);
    c->gm_basis = c->gm_small = nullptr;
    c->gm_k = 0;
    if (hipMalloc((void**)&c->gm_basis, sizeof(double) * (size_t)(k + 1) * n) != hipSuccess ||
        hipMalloc((void**)&c->gm_small, sizeof(double) * ((size_t)(k + 1) * k + 4 * k + 2)) !=
            hipSuccess) {
      (void)hipGetLastError();
      return set_error(c, ELASTIC_OOM, "GMRES basis of " + std::to_string(k + 1) + " vectors");
    }
    c->gm_k = k;
  }
  const GmSmall a = gm_small(c->gm_small, k);
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(c->gm_small, 0,
                                      sizeof(double) * ((size_t)(k + 1) * k + 4 * k + 2), s));
  const dim3 G(kRedBlocks), TB(kThreads);
  double* P[2] = {c->red64, c->red64 + kRedBlocks};
  double *x = c->v64[VX], *b = c->v64[VB], *t = c->v64[VT], *u = c->v64[VU];
  const double* m = c->A64.m;
  const double* bm = bdiag ? solver_bm(c) : nullptr;
  auto g = [&](int j) { return c->gm_basis + (size_t)(j - 1) * n; };  // 1-based
  SolveScalars* sc = c->scal;
  SolveScalars* hs = c->scal_host;
  auto read_scalars = [&]() -> int {
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(hs, sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
    return ELASTIC_OK;
  };
  // x = 0; g1 = M b; econv = tol ||g1||
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(x, 0, sizeof(double) * n, s));
  hipLaunchKernelGGL(k_gm_applym_dot, G, TB, 0, s, n, b, m, bm, g(1), P[0]);
  hipLaunchKernelGGL(k_gm_norm0, dim3(1), TB, 0, s, P[0], a, sc);
  int rc = read_scalars();
  if (rc) return rc;
  const double econv = tol * hs->gm_h;
  out->conv = econv;
  int nit = 0, cycles = 0;
  bool conv = hs->gm_h == 0.0;
  bool op32 = fp32_operator;  // progress guard on the FP64 restart residuals
  double last_true = hs->gm_h;
  int stall = 0;
  out->op64_from = -1;
  for (int l = 1; l <= maxit && !conv; ++l) {
    // g1 = M (b - A x), FP64 operator (the true residual)
    launch_spmv(c, false, x, t, c->epart64);
    hipLaunchKernelGGL(k_gm_resid, G, TB, 0, s, n, merge_of(c, c->epart64), t, b, u);
    hipLaunchKernelGGL(k_gm_applym_dot, G, TB, 0, s, n, u, m, bm, g(1), P[0]);
    hipLaunchKernelGGL(k_gm_norm0, dim3(1), TB, 0, s, P[0], a, sc);
    if (fp32_operator && l > 1) {  // mixed: decide on the FP64 residual
      rc = read_scalars();
      if (rc) return rc;
      if (hs->gm_h <= econv) {
        conv = true;
        break;
      }
      if (hs->gm_h > 0.9 * last_true) {  // no progress on the best residual so far
        ++stall;
      } else {
        stall = 0;
        last_true = hs->gm_h;
      }
      if (op32 && stall >= 2) {
        op32 = false;
        out->op64_from = nit;
      }
    }
    hipLaunchKernelGGL(k_gm_scale, G, TB, 0, s, n, g(1), (const double*)a.e);
    ++cycles;
    int ni = 0;
    for (int i = 1; i <= k; ++i) {
      ++nit;
      ++ni;
      launch_spmv(c, op32, g(i), t, c->epart64);
      hipLaunchKernelGGL(k_gm_merge, G, TB, 0, s, n, merge_of(c, c->epart64, epart_f32(c, op32)),
                         t, u);
      double* w = g(i + 1);
      hipLaunchKernelGGL(k_gm_mdot, G, TB, 0, s, n, u, m, bm, w, g(1), P[0]);
      for (int j = 1; j <= i; ++j)
        hipLaunchKernelGGL(k_gm_mgs, G, TB, 0, s, n, P[(j - 1) & 1], w, g(j),
                           j < i ? (const double*)g(j + 1) : nullptr, P[j & 1],
                           a.h + (size_t)(i - 1) * (k + 1) + (j - 1));
      hipLaunchKernelGGL(k_gm_givens, dim3(1), TB, 0, s, P[i & 1], i, k, a, sc);
      hipLaunchKernelGGL(k_gm_scale, G, TB, 0, s, n, w, (const double*)&sc->gm_h);
      rc = read_scalars();
      if (rc) return rc;
      if (hs->gm_e <= econv) break;
    }
    hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), 0, s, ni, k, a);
    hipLaunchKernelGGL(k_gm_xupdate, G, TB, 0, s, n, ni, c->gm_basis, (const double*)a.y, x);
    if (!fp32_operator && hs->gm_e <= econv) conv = true;
  }
  if (fp32_operator && !conv && cycles > 0) {
    // the cycle limit came before the next cycle's FP64 test: test the last
    // update once on the FP64 residual
    launch_spmv(c, false, x, t, c->epart64);
    hipLaunchKernelGGL(k_gm_resid, G, TB, 0, s, n, merge_of(c, c->epart64), t, b, u);
    hipLaunchKernelGGL(k_gm_applym_dot, G, TB, 0, s, n, u, m, bm, g(1), P[0]);
    hipLaunchKernelGGL(k_gm_norm0, dim3(1), TB, 0, s, P[0], a, sc);
    rc = read_scalars();
    if (rc) return rc;
    conv = hs->gm_h <= econv;
  }
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  out->iters = nit;
  out->outer = cycles;
  out->done = conv;
  out->status = conv ? ELASTIC_OK : ELASTIC_MAXIT;
  return ELASTIC_OK;
}

}  // namespace elastic
