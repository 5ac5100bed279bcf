// Device-resident PCG on gfx950.
//
// PCG follows Iters.f:252-524 step for step (x0 = 0, Jacobi or 3x3
// block-Jacobi M, stop when sqrt|r.M r| < tol sqrt|b.M b|).  One iteration is
// four kernels and no host synchronisation on the node-blocked layout:
//   spmv tiles (q = A p, and each tile's share of p.q)  ->  sum of the shares
//     ->  merge of q's external partials + x, r update + r.z  ->  p update
// (the row-tile layout merges q and forms p.q in a kernel of its own).
// Dot products are reduced in two fixed-order stages (per-block partials over a
// fixed grid of kRedBlocks blocks, then every consumer block sums the partials
// in the same order), so every block sees the same alpha and beta and the
// results are reproducible run to run.  Vectors, dots and scalars are FP64.
// Scalars live in device memory (SolveScalars); kernels return at once when
// the solve has converged, so the host only polls a flag every `check_every`
// iterations.
//
// Mixed precision (TV = float): the operator of the iteration reads FP32
// matrix values (sums in FP64, or in FP32 when the context asks for it: then
// x is staged from an FP32 copy of p that the p update writes, and the own
// rows of q and the external partials are stored in FP32, all exactly the
// values the FP32 sums produce).  Whenever the recursively updated residual
// has dropped by `delta` since the last replacement, it is replaced by the
// FP64 true residual r = b - A64 x (an FP64 product with the FP64 values)
// and the search direction is kept, so the Krylov iteration is not
// restarted; the recurrences are the flexible ones (alpha = p.r / p.q,
// beta = -z.q / p.q), and convergence is decided on the true residual only.
#include <algorithm>
#include <cmath>

#include "device_util.h"
#include "elastic_internal.h"

namespace elastic {

// x = 0, r = b, p = z = M^-1 b, partial b.z
__global__ __launch_bounds__(kThreads) void k_pcg_init(int n, const double* __restrict__ b,
                                                        const double* __restrict__ m,
                                                        const double* __restrict__ bm, double* x,
                                                        double* r, double* p, double* part,
                                                        double* part_pr, float* p32,
                                                        float* r32, float* dx32) {
  __shared__ double sh[8];
  double acc = 0.0, apr = 0.0;
  GRID_LOOP(i, n) {
    const double zi = apply_m(i, n, b, m, bm);
    x[i] = 0.0;
    if (dx32) dx32[i] = 0.0f;
    r[i] = b[i];
    p[i] = zi;
    if (p32) p32[i] = (float)zi;
    acc += b[i] * zi;
    if (r32) {
      r32[i] = (float)b[i];
      apr += (double)(float)b[i] * (double)(float)zi;
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (part_pr) {  // p.r = z.b (of the FP32 vectors when the iteration keeps those)
    const double t = r32 ? block_sum(apr, sh) : s;
    if (threadIdx.x == 0) part_pr[blockIdx.x] = t;
  }
}

__global__ void k_pcg_init2(const double* __restrict__ part, double tol, double delta,
                            int reliable, double* dslot, SolveScalars* sc) {
  __shared__ double sh[8];
  const double d = grid_sum(part, sh);
  if (threadIdx.x == 0) {
    *dslot = d;
    sc->conv = tol * sqrt(fabs(d));
    sc->norm_b = sqrt(fabs(d));
    sc->rz0 = d;
    sc->rmax = sqrt(fabs(d));
    sc->delta = delta;
    sc->true_rmr = d;
    sc->iter = 0;
    sc->status = 0;
    sc->reliable = reliable;
    sc->upd = 0;
    sc->skip64 = 1;
    sc->nupd = 0;
    progress_init(sc, sqrt(fabs(d)));
    sc->pq_tiles = 0.0;
    sc->done = (d == 0.0) ? 1 : 0;  // b = 0: x = 0 is the solution
  }
}

// q += merged external partials; partial p.q, and p.r for the flexible
// recurrences (when part_pr != null)
__global__ __launch_bounds__(kThreads) void k_merge_dot(int n, MergeArgs mg,
                                                         double* q, const double* __restrict__ p,
                                                         double* part, const double* __restrict__ r,
                                                         double* part_pr, const SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  double acc = 0.0, apr = 0.0;
  GRID_LOOP(i, n) {
    double s = q[i];
    s = merge_row(mg, i, s);
    q[i] = s;
    acc += p[i] * s;
    if (part_pr) apr += p[i] * r[i];
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (part_pr) {
    const double t = block_sum(apr, sh);
    if (threadIdx.x == 0) part_pr[blockIdx.x] = t;
  }
}

// alpha = d / p.q (or p.r / p.q) ; x += alpha p ; r -= alpha q ; with Jacobi
// also partial r.(m r) and, for the flexible recurrences, (m r).q
__global__ __launch_bounds__(kThreads) void k_pcg_update(int n, const double* __restrict__ part_pq,
                                                          const double* __restrict__ part_pr,
                                                          const double* __restrict__ dslot,
                                                          double* x, const double* __restrict__ p,
                                                          double* r, const double* __restrict__ q,
                                                          const double* __restrict__ m,
                                                          int jacobi, double* part_rz,
                                                          double* part_zq, SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  const double pq = grid_sum(part_pq, sh);
  const double num = part_pr ? grid_sum(part_pr, sh) : *dslot;
  const double alpha = num / pq;
  double acc = 0.0, azq = 0.0;
  GRID_LOOP(i, n) {
    const double qi = q[i];
    x[i] = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * qi;
    r[i] = ri;
    const double zi = ri * m[i];
    acc += ri * zi;
    azq += zi * qi;
  }
  if (jacobi) {
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = s;
    if (part_zq) {
      const double t = block_sum(azq, sh);
      if (threadIdx.x == 0) part_zq[blockIdx.x] = t;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->alpha = alpha;
    sc->pq = pq;
  }
}

// k_pcg_update on the node-blocked layout (n = NB * nn), one lane per node:
// the node's merge list is walked once for its NB rows (contiguous partials),
// and the vector loads are issued before that gather.  p.q comes from the
// product's tiles (sc->pq_tiles).  EP / QT: types of the stored partials and
// of the product's own rows (FP32 after a product with FP32 sums that was
// asked for FP32 rows; q is then not written back).  VT: type of p, r and the
// Jacobi scale m (FP32 when the mixed iteration keeps its vectors in FP32: p
// is then the vector the product staged, so x += alpha p moves along exactly
// that vector).  XT: type of the stored x; float when the iteration adds its
// steps to FP32 increments dx (x = x64 + dx), which every true-residual
// replacement folds into the FP64 x (k_rel_resid) and clears.  bm (3-dof
// layouts, FP32 vectors): the 3x3 node-block inverse of block_precond
// (c11 c22 c33 c21 c31 c32 per node, Precond.f:419-540) gives z = M^-1 r of
// the node here, so r.z and z.q need no pass of their own.
template <int NB, typename EP, typename QT, typename VT, typename XT = double>
__global__ __launch_bounds__(kThreads) void k_pcg_update_nodes(int nn, MergeArgs mg,
                                                                const double* __restrict__ part_pr,
                                                                const double* __restrict__ dslot,
                                                                XT* x,
                                                                const VT* __restrict__ p,
                                                                VT* r, double* q, int keep_q,
                                                                const VT* __restrict__ m,
                                                                int jacobi, double* part_rz,
                                                                double* part_zq, SolveScalars* sc,
                                                                const double* __restrict__ bm) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  // the loads of each lane's first node are issued before alpha's
  // block-wide sum, and every node's loads before the previous node's
  // stores, so that their latency overlaps
  const int stride = gridDim.x * kThreads;
  int J = blockIdx.x * kThreads + threadIdx.x;
  double qv[NB], xv[NB], pv[NB], rv[NB], mv[NB];
  int k0 = 0, k1 = 0;
  auto load_node = [&](int Jn) {
    const int64_t i0 = (int64_t)NB * Jn;
    k0 = mg.ptr[Jn];
    k1 = mg.ptr[Jn + 1];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      qv[c] = (double)reinterpret_cast<const QT*>(q)[i0 + c];
      xv[c] = (double)x[i0 + c];
      pv[c] = (double)p[i0 + c];
      rv[c] = (double)r[i0 + c];
      mv[c] = jacobi ? (double)m[i0 + c] : 0.0;  // (the block preconditioner's r.z is k_rz's)
    }
  };
  if (J < nn) load_node(J);
  const double pq = sc->pq_tiles;
  const double num = part_pr ? grid_sum(part_pr, sh) : *dslot;
  const double alpha = num / pq;
  double acc = 0.0, azq = 0.0;
  for (; J < nn; J += stride) {
    const int64_t i0 = (int64_t)NB * J;
    for (int k = k0; k < k1; ++k) {
      const EP* e = reinterpret_cast<const EP*>(mg.epart) + (int64_t)NB * mg.idx[k];
#pragma unroll
      for (int c = 0; c < NB; ++c) qv[c] += (double)e[c];
    }
    double xo[NB], ro[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      xo[c] = xv[c] + alpha * pv[c];
      const VT rs = (VT)(rv[c] - alpha * qv[c]);
      ro[c] = (double)rs;
    }
    if (NB == 3 && bm) {  // z = B^-1 r of the node (the block inverse's products)
      const double* a = bm + 6 * (int64_t)J;
      const double z[3] = {a[0] * ro[0] + a[3] * ro[1] + a[4] * ro[2],
                           a[3] * ro[0] + a[1] * ro[1] + a[5] * ro[2],
                           a[4] * ro[0] + a[5] * ro[1] + a[2] * ro[2]};
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        acc += ro[c] * z[c];
        azq += z[c] * qv[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const double zi = ro[c] * mv[c];
        acc += ro[c] * zi;
        azq += zi * qv[c];
      }
    }
    double qo[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) qo[c] = qv[c];
    if (J + stride < nn) load_node(J + stride);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      if (sizeof(QT) == 8 && keep_q) q[i0 + c] = qo[c];
      x[i0 + c] = (XT)xo[c];
      r[i0 + c] = (VT)ro[c];
    }
  }
  if (jacobi || bm) {
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = s;
    if (part_zq) {
      const double t = block_sum(azq, sh);
      if (threadIdx.x == 0) part_zq[blockIdx.x] = t;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->alpha = alpha;
    sc->pq = pq;
  }
}

// partial r.(M^-1 r) for the block preconditioner (after r is complete), and
// (M^-1 r).q for the flexible recurrences
__global__ __launch_bounds__(kThreads) void k_rz(int n, const double* __restrict__ r,
                                                  const double* __restrict__ m,
                                                  const double* __restrict__ bm, double* part,
                                                  const double* __restrict__ q, double* part_zq,
                                                  const int32_t* gate, const SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  if (gate && !block_flag(gate, sh)) return;
  double acc = 0.0, azq = 0.0;
  GRID_LOOP(i, n) {
    const double zi = apply_m(i, n, r, m, bm);
    acc += r[i] * zi;
    if (part_zq) azq += zi * q[i];
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (part_zq) {
    const double t = block_sum(azq, sh);
    if (threadIdx.x == 0) part_zq[blockIdx.x] = t;
  }
}

// di = r.z ; beta = di / d (or -z.q / p.q) ; p = z + beta p ; partial p.r
// when part_pr != null ; bookkeeping in block 0
__global__ __launch_bounds__(kThreads) void k_pcg_pupdate(int n, const double* __restrict__ part_rz,
                                                           const double* __restrict__ part_zq,
                                                           const double* __restrict__ dold,
                                                           double* dnew, double* p,
                                                           const double* __restrict__ r,
                                                           const double* __restrict__ m,
                                                           const double* __restrict__ bm,
                                                           double* part_pr, float* p32,
                                                           const float* __restrict__ r32,
                                                           const float* __restrict__ m32,
                                                           SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  // FP32 vectors: the first kPre grid-stride groups of r, m and p are loaded
  // before beta's block-wide sums, so that their latency overlaps
  constexpr int kPre = 4;
  const int n4 = (r32 && !bm) ? (n >> 2) : 0;  // (the node-block path loads per node)
  const int stride = gridDim.x * kThreads;
  int k = blockIdx.x * kThreads + threadIdx.x;
  float4 rf[kPre], mf[kPre], pf[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u)
    if (k + u * stride < n4) {
      rf[u] = reinterpret_cast<const float4*>(r32)[k + u * stride];
      mf[u] = reinterpret_cast<const float4*>(m32)[k + u * stride];
      pf[u] = reinterpret_cast<const float4*>(p32)[k + u * stride];
    }
  const double di = grid_sum(part_rz, sh);
  const double beta = part_zq ? -grid_sum(part_zq, sh) / sc->pq : di / *dold;
  double apr = 0.0;
  if (r32 && bm) {
    // FP32 vectors with the 3x3 node-block inverse (3-dof layouts, n = 3 nn):
    // one lane per node, p = fl32(B^-1 r + beta p) in FP64 arithmetic, and
    // p.r of the stored values
    const int nn = n / 3;
    for (int J = blockIdx.x * kThreads + threadIdx.x; J < nn; J += gridDim.x * kThreads) {
      const int64_t i0 = 3 * (int64_t)J;
      const double r0 = (double)r32[i0], r1 = (double)r32[i0 + 1], r2 = (double)r32[i0 + 2];
      const double* a = bm + 6 * (int64_t)J;
      const double z[3] = {a[0] * r0 + a[3] * r1 + a[4] * r2, a[3] * r0 + a[1] * r1 + a[5] * r2,
                           a[4] * r0 + a[5] * r1 + a[2] * r2};
      const double rv[3] = {r0, r1, r2};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float pi = (float)(z[c] + beta * (double)p32[i0 + c]);
        p32[i0 + c] = pi;
        apr += (double)pi * rv[c];
      }
    }
  } else if (r32) {
    // the iteration's vectors are FP32 (Jacobi): p = fl32(m r + beta p), in
    // FP64 arithmetic, and p.r of the stored values; four rows per lane
    while (k < n4) {
#pragma unroll
      for (int u = 0; u < kPre; ++u) {
        const int kk = k + u * stride;
        if (kk < n4) {
          const float4 o =
              make_float4((float)((double)rf[u].x * (double)mf[u].x + beta * (double)pf[u].x),
                          (float)((double)rf[u].y * (double)mf[u].y + beta * (double)pf[u].y),
                          (float)((double)rf[u].z * (double)mf[u].z + beta * (double)pf[u].z),
                          (float)((double)rf[u].w * (double)mf[u].w + beta * (double)pf[u].w));
          reinterpret_cast<float4*>(p32)[kk] = o;
          apr += (double)o.x * (double)rf[u].x + (double)o.y * (double)rf[u].y +
                 (double)o.z * (double)rf[u].z + (double)o.w * (double)rf[u].w;
        }
      }
      k += kPre * stride;
#pragma unroll
      for (int u = 0; u < kPre; ++u)
        if (k + u * stride < n4) {
          rf[u] = reinterpret_cast<const float4*>(r32)[k + u * stride];
          mf[u] = reinterpret_cast<const float4*>(m32)[k + u * stride];
          pf[u] = reinterpret_cast<const float4*>(p32)[k + u * stride];
        }
    }
    for (int i = 4 * n4 + blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
      const float pi = (float)((double)r32[i] * (double)m32[i] + beta * (double)p32[i]);
      p32[i] = pi;
      apr += (double)pi * (double)r32[i];
    }
  } else if (bm == nullptr) {
    // Jacobi: four rows per lane as two 16-byte loads of r, m and p each, so
    // that enough bytes are in flight at 4 waves per SIMD; vectors are
    // 256-byte aligned (hipMalloc), rows 4k..4k+3 are 32-byte aligned
    const int n4 = n >> 2;
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < n4; k += gridDim.x * kThreads) {
      const double2* r2 = reinterpret_cast<const double2*>(r) + 2 * k;
      const double2* m2 = reinterpret_cast<const double2*>(m) + 2 * k;
      double2* p2 = reinterpret_cast<double2*>(p) + 2 * k;
      const double2 ra = r2[0], rb = r2[1], ma = m2[0], mb = m2[1], pa = p2[0], pb = p2[1];
      const double2 qa = {ra.x * ma.x + beta * pa.x, ra.y * ma.y + beta * pa.y};
      const double2 qb = {rb.x * mb.x + beta * pb.x, rb.y * mb.y + beta * pb.y};
      p2[0] = qa;
      p2[1] = qb;
      if (p32)  // 16-byte aligned: 4 floats from a 256-byte aligned base
        reinterpret_cast<float4*>(p32)[k] = make_float4((float)qa.x, (float)qa.y, (float)qb.x,
                                                        (float)qb.y);
      apr += qa.x * ra.x + qa.y * ra.y + qb.x * rb.x + qb.y * rb.y;
    }
    for (int i = 4 * n4 + blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
      const double pi = r[i] * m[i] + beta * p[i];
      p[i] = pi;
      if (p32) p32[i] = (float)pi;
      apr += pi * r[i];
    }
  } else {
    GRID_LOOP(i, n) {
      const double pi = apply_m(i, n, r, m, bm) + beta * p[i];
      p[i] = pi;
      if (p32) p32[i] = (float)pi;
      apr += pi * r[i];
    }
  }
  if (part_pr) {  // p.r of the next step (flexible recurrences, tile dots)
    const double t = block_sum(apr, sh);
    if (threadIdx.x == 0) part_pr[blockIdx.x] = t;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *dnew = di;
    sc->beta = beta;
    sc->di = di;
    sc->iter += 1;
    const double rn = sqrt(fabs(di));
    if (!sc->reliable) {
      if (rn < sc->conv) sc->done = 1;
    } else if (rn < sc->delta * sc->rmax || rn < sc->conv ||
               sc->iter - sc->last_upd >= kForceReplace) {
      sc->upd = 1;  // true-residual replacement at the end of this chunk
      sc->skip64 = 0;
    }
  }
}

// ---------------------------------------------------- residual replacement
__global__ void k_rel_force(SolveScalars* sc) {
  if (threadIdx.x == 0 && !sc->done) {
    sc->upd = 1;
    sc->skip64 = 0;
  }
}

// r = b - (t + merged partials); with Jacobi also partial r.(m r); with
// part_pr also the partial p.r of the replaced residual.  With r32 the
// iteration's vectors are FP32: the residual is stored there, rounded, and p
// is read from p32; the norm is formed from the FP64 residual.  With dx32 the
// product t was formed on x + dx32 (the FP32 increments of x since the last
// replacement): they are added to x here and cleared.
__global__ __launch_bounds__(kThreads) void k_rel_resid(int n, MergeArgs mg,
                                                         const double* __restrict__ t,
                                                         const double* __restrict__ b,
                                                         const double* __restrict__ m, double* r,
                                                         int jacobi, double* part,
                                                         const double* __restrict__ p,
                                                         const float* __restrict__ p32,
                                                         float* r32, double* part_pr,
                                                         double* x, float* dx32,
                                                         const SolveScalars* sc,
                                                         const double* __restrict__ bm) {
  __shared__ double sh[8];
  if (!block_flag(&sc->upd, sh) || block_flag(&sc->done, sh)) return;
  double acc = 0.0, apr = 0.0;
  if (r32 && bm) {  // FP32 vectors, 3x3 node-block inverse: r.(B^-1 r) per node (n = 3 nn)
    for (int J = blockIdx.x * kThreads + threadIdx.x; J < n / 3; J += gridDim.x * kThreads) {
      double rv[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int i = 3 * J + c;
        if (dx32) {
          x[i] = x[i] + (double)dx32[i];
          dx32[i] = 0.0f;
        }
        double s = t[i];
        s = merge_row(mg, i, s);
        rv[c] = b[i] - s;
        const float rs = (float)rv[c];
        r32[i] = rs;
        if (part_pr) apr += (double)p32[i] * (double)rs;
      }
      const double* a = bm + 6 * (int64_t)J;
      acc += rv[0] * (a[0] * rv[0] + a[3] * rv[1] + a[4] * rv[2]) +
             rv[1] * (a[3] * rv[0] + a[1] * rv[1] + a[5] * rv[2]) +
             rv[2] * (a[4] * rv[0] + a[5] * rv[1] + a[2] * rv[2]);
    }
  }
  for (int i = (r32 && bm) ? n : blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    if (dx32) {
      x[i] = x[i] + (double)dx32[i];
      dx32[i] = 0.0f;
    }
    double s = t[i];
    s = merge_row(mg, i, s);
    const double ri = b[i] - s;
    acc += ri * (ri * m[i]);
    if (r32) {
      const float rs = (float)ri;
      r32[i] = rs;
      if (part_pr) apr += (double)p32[i] * (double)rs;
    } else {
      r[i] = ri;
      if (part_pr) apr += p[i] * ri;
    }
  }
  if (jacobi || (r32 && bm)) {
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
  if (part_pr) {
    const double s = block_sum(apr, sh);
    if (threadIdx.x == 0) part_pr[blockIdx.x] = s;
  }
}

// true-residual test; otherwise the iteration resumes from the replaced
// residual with its search direction kept
__global__ void k_rel_finish(const double* __restrict__ part, double* dslot, SolveScalars* sc) {
  __shared__ double sh[8];
  if (!block_flag(&sc->upd, sh) || block_flag(&sc->done, sh)) return;
  const double rz = grid_sum(part, sh);
  if (threadIdx.x == 0) {
    sc->true_rmr = rz;
    sc->nupd += 1;
    progress_guard(sc, sqrt(fabs(rz)));
    if (sqrt(fabs(rz)) < sc->conv) {
      sc->done = 1;
      sc->skip64 = 1;  // gated products later in the chunk return at once
    } else {
      *dslot = rz;
      sc->rmax = sqrt(fabs(rz));
      sc->upd = 0;
      sc->skip64 = 1;
    }
  }
}

// restart after the hand-over to the FP64 operator: p = M^-1 r from the true
// residual just computed (partial p.r for the next update when asked), and
// no more replacements (the stop test is made on the recursive residual, as
// FP64 PCG makes it)
__global__ __launch_bounds__(kThreads) void k_pcg_restart(int n, const double* __restrict__ r,
                                                           const double* __restrict__ m,
                                                           const double* __restrict__ bm,
                                                           double* p, double* part_pr,
                                                           SolveScalars* sc) {
  __shared__ double sh[8];
  if (block_flag(&sc->done, sh)) return;
  double apr = 0.0;
  GRID_LOOP(i, n) {
    const double zi = apply_m(i, n, r, m, bm);
    p[i] = zi;
    apr += zi * r[i];
  }
  if (part_pr) {
    const double t = block_sum(apr, sh);
    if (threadIdx.x == 0) part_pr[blockIdx.x] = t;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->reliable = 0;
    sc->upd = 0;
  }
}

// ------------------------------------------------------------- host side
int pcg_run(Context* c, const PcgVecs& V, bool fp32_operator, double tol, int maxit,
            bool reliable, double delta, double* dslots) {
  hipStream_t s = c->stream;
  const int n = c->neq;
  const dim3 G(kRedBlocks), TB(kThreads);
  double* part_a = c->red64;
  double* part_b = c->red64 + kRedBlocks;
  // flexible recurrences with the FP32 operator and residual replacement
  double* part_pr = reliable ? c->red64 + 2 * kRedBlocks : nullptr;
  double* part_zq = reliable ? c->red64 + 3 * kRedBlocks : nullptr;
  const int jac = V.bm == nullptr;
  SolveScalars* sc = c->scal;
  // node-blocked layout: the product's tiles form p.q themselves
  const bool tdots = c->tdot != nullptr && c->blk.ntiles > 0;
  const DotOut dot{c->tdot, &sc->pq_tiles};
  // p.r after the p update (and after a replacement), read by the next update
  double* pr_next = tdots ? part_pr : nullptr;
  // FP32 copy of p for the FP32-sum operator's staging (node-blocked layouts)
  float* p32 = (tdots && fp32_operator && c->accum32) ? c->p32 : nullptr;
  // With Jacobi and the FP32-sum operator the iteration keeps p, r and the
  // Jacobi scale in FP32 (x, the true residuals and all sums stay FP64).  If
  // the progress guard hands over to the FP64 operator, p and r are copied to
  // their FP64 vectors and the iteration continues there.
  // With the 3x3 node-block inverse (3-dof layouts) the FP32 iteration applies
  // it per node inside the update, the p update and the replacement pass.
  const bool bm3 = !jac && c->layout_nb == 3 && tdots;
  bool v32 = p32 != nullptr && (jac || bm3) && reliable && epart_f32(c, true);
  float* r32 = v32 ? p32 + vec32_stride(n) : nullptr;
  float* m32 = (v32 && jac) ? p32 + 2 * vec32_stride(n) : nullptr;
  // On large systems the FP32 iteration also keeps its steps of x as FP32
  // increments dx32 (x = V.x + dx32): the update then moves 4 bytes of x per
  // row instead of 8 each way.  The increments since the last replacement
  // are a small fraction of x (the residual fell by at most delta since),
  // and every replacement forms its product on V.x + dx32 and folds dx32
  // into V.x, so x, the true residuals and the stop test stay FP64.
  float* dx32 = (v32 && n >= kPredictRows) ? p32 + 3 * vec32_stride(n) : nullptr;
  if (v32 && jac) launch_cast<double, float>(n, V.m, m32, s);
  hipLaunchKernelGGL(k_pcg_init, G, TB, 0, s, n, V.b, V.m, V.bm, V.x, V.r, V.p, part_a, pr_next,
                     p32, r32, dx32);
  hipLaunchKernelGGL(k_pcg_init2, dim3(1), TB, 0, s, part_a, tol, delta, reliable ? 1 : 0, dslots,
                     sc);
  auto replace = [&](int iters_so_far) {
    double* dnext = dslots + (iters_so_far & 1);
    launch_spmv(c, false, V.x, V.t, V.epart, &sc->skip64, nullptr, false, nullptr, dx32);
    hipLaunchKernelGGL(k_rel_resid, G, TB, 0, s, n, merge_of(c, V.epart),
                       V.t, V.b, V.m, V.r, jac, part_b, (const double*)V.p,
                       (const float*)(v32 ? p32 : nullptr), v32 ? r32 : nullptr, pr_next, V.x,
                       dx32, sc, (const double*)(v32 && !jac ? V.bm : nullptr));
    if (!jac && !v32)
      hipLaunchKernelGGL(k_rz, G, TB, 0, s, n, V.r, V.m, V.bm, part_b, (const double*)nullptr,
                         (double*)nullptr, &sc->upd, sc);
    hipLaunchKernelGGL(k_rel_finish, dim3(1), TB, 0, s, part_b, dnext, sc);
  };
  int it = 0;
  const int every = std::max(2, c->check_every & ~1);
  bool done = false;
  bool op32 = fp32_operator;  // the progress guard may hand over to the FP64 operator
  bool replacing = reliable;  // replacements stop at the hand-over's restart
  // sampled event timing of the products (measurement option)
  const int stride = c->time_stride;
  if (stride > 0 && c->tev.size() < (size_t)2 * every) {
    for (hipEvent_t e : c->tev) hipEventDestroy(e);
    c->tev.assign(2 * every, nullptr);
    // timing only: no system-scope fence (no cache writeback/invalidation)
    // when an event completes, which would slow the kernel after it
    for (auto& e : c->tev)
      ELASTIC_HIP_CHECK(c, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  }
  // Gated replacement passes cost their launches even when they skip.  On
  // large systems a replacement that the iteration asked for (the recursive
  // residual fell by delta, or kForceReplace iterations passed) runs right
  // after the poll that sees the flag, and the gated pass is enqueued every
  // second iteration only in chunks where the stopping threshold can be
  // reached within two chunks, so that convergence is confirmed on the true
  // residual without delay.  The decay rate of the recursive residual is
  // taken over the last kRateWindow polls: CG's residual is not monotone, so
  // a rate over one chunk often shows growth and would keep the gate open.
  // Small systems enqueue the gated pass every second iteration.
  const bool predict = reliable && n >= kPredictRows;
  bool gate = true;
  constexpr int kRateWindow = 4;
  double rn_hist[kRateWindow + 1];
  int it_hist[kRateWindow + 1];
  int nhist = 0;
  while (it < maxit) {
    const int chunk = std::min(every, maxit - it);
    const int it0 = it;
    for (int j = 0; j < chunk; ++j, ++it) {
      double* dcur = dslots + (it & 1);
      double* dnxt = dslots + ((it + 1) & 1);
      const bool timed = stride > 0 && it % stride == 0;
      if (timed) ELASTIC_HIP_CHECK(c, hipEventRecord(c->tev[2 * j], s));
      if (tdots) {
        // q is read again only by k_rz's (M^-1 r).q (FP64 vectors with the
        // block preconditioner)
        const int keep_q = (!jac && reliable && !v32) ? 1 : 0;
        const bool e32 = epart_f32(c, op32);
        // q in FP32 when the product sums in FP32 and nothing reads q afterwards
        const bool q32 = e32 && !keep_q;
        launch_spmv(c, op32, V.p, V.q, V.epart, &sc->done, &dot, q32, e32 ? p32 : nullptr);
        if (timed) ELASTIC_HIP_CHECK(c, hipEventRecord(c->tev[2 * j + 1], s));
        // q is read again only by the block preconditioner's (M^-1 r).q
        const MergeArgs mg = merge_of(c, V.epart, e32);
        const double* bmv = (v32 && !jac) ? V.bm : nullptr;  // node-block inverse, FP32 vectors
#define ELASTIC_UPD(NB, EP, QT, VT, P, R, M, BM)                                               \
  hipLaunchKernelGGL((k_pcg_update_nodes<NB, EP, QT, VT>), G, TB, 0, s, n / NB, mg,            \
                     (const double*)part_pr, dcur, V.x, (const VT*)(P), (VT*)(R), V.q, keep_q,  \
                     (const VT*)(M), jac, part_b, part_zq, sc, (const double*)(BM))
#define ELASTIC_UPD_DX(NB)                                                                     \
  hipLaunchKernelGGL((k_pcg_update_nodes<NB, float, float, float, float>), G, TB, 0, s, n / NB, \
                     mg, (const double*)part_pr, dcur, dx32, (const float*)p32, r32, V.q,      \
                     keep_q, (const float*)m32, jac, part_b, part_zq, sc, bmv)
        if (c->layout_nb == 3) {
          if (q32 && v32 && dx32) ELASTIC_UPD_DX(3);
          else if (q32 && v32) ELASTIC_UPD(3, float, float, float, p32, r32, m32, bmv);
          else if (q32) ELASTIC_UPD(3, float, float, double, V.p, V.r, V.m, nullptr);
          else if (e32) ELASTIC_UPD(3, float, double, double, V.p, V.r, V.m, nullptr);
          else ELASTIC_UPD(3, double, double, double, V.p, V.r, V.m, nullptr);
        } else {
          if (q32 && v32 && dx32) ELASTIC_UPD_DX(2);
          else if (q32 && v32) ELASTIC_UPD(2, float, float, float, p32, r32, m32, nullptr);
          else if (q32) ELASTIC_UPD(2, float, float, double, V.p, V.r, V.m, nullptr);
          else if (e32) ELASTIC_UPD(2, float, double, double, V.p, V.r, V.m, nullptr);
          else ELASTIC_UPD(2, double, double, double, V.p, V.r, V.m, nullptr);
        }
#undef ELASTIC_UPD
#undef ELASTIC_UPD_DX
      } else {
        launch_spmv(c, op32, V.p, V.q, V.epart, &sc->done);
        if (timed) ELASTIC_HIP_CHECK(c, hipEventRecord(c->tev[2 * j + 1], s));
        hipLaunchKernelGGL(k_merge_dot, G, TB, 0, s, n, merge_of(c, V.epart), V.q, V.p, part_a,
                           (const double*)V.r, part_pr, sc);
        hipLaunchKernelGGL(k_pcg_update, G, TB, 0, s, n, (const double*)part_a,
                           (const double*)part_pr, dcur, V.x, V.p, V.r, (const double*)V.q, V.m,
                           jac, part_b, part_zq, sc);
      }
      if (!jac && !(v32 && tdots))
        hipLaunchKernelGGL(k_rz, G, TB, 0, s, n, V.r, V.m, V.bm, part_b, (const double*)V.q,
                           part_zq, (const int32_t*)nullptr, sc);
      hipLaunchKernelGGL(k_pcg_pupdate, G, TB, 0, s, n, part_b, (const double*)part_zq, dcur, dnxt,
                         V.p, V.r, V.m, V.bm, pr_next, p32, (const float*)r32,
                         (const float*)m32, sc);
      // the replacement kernels gate themselves on the flag set by k_pcg_pupdate
      if (replacing && gate && ((it + 1) & 1) == 0) replace(it + 1);
    }
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->scal_host, sc, sizeof(SolveScalars),
                                        hipMemcpyDeviceToHost, s));
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
    if (stride > 0) {  // products of iterations that ran (later ones return at once)
      for (int j = 0; j < chunk; ++j)
        if ((it0 + j) % stride == 0 && it0 + j < c->scal_host->iter) {
          float ms = 0.f;
          ELASTIC_HIP_CHECK(c, hipEventElapsedTime(&ms, c->tev[2 * j], c->tev[2 * j + 1]));
          c->tprod_ms += ms;
          c->tprod_n += 1;
        }
    }
    if (c->scal_host->done) {
      done = true;
      break;
    }
    if (c->scal_host->op64 && op32) {
      op32 = false;
      if (v32) {  // the FP64 operator and kernels read the FP64 vectors from here on
        launch_cast<float, double>(n, p32, V.p, s);
        launch_cast<float, double>(n, r32, V.r, s);
        v32 = false;
        r32 = m32 = nullptr;
      }
      // restart: r = b - A64 x, p = M^-1 r, FP64 PCG from here on (the
      // replacement folds the FP32 increments of x, which end there)
      hipLaunchKernelGGL(k_rel_force, dim3(1), dim3(64), 0, s, sc);
      replace(it);
      dx32 = nullptr;
      hipLaunchKernelGGL(k_pcg_restart, G, TB, 0, s, n, (const double*)V.r, V.m, V.bm, V.p,
                         pr_next, sc);
      replacing = false;
    }
    if (predict && replacing) {
      const SolveScalars& h = *c->scal_host;
      if (h.upd) replace(it);  // asked for during the chunk: runs before the next one
      const double rn = std::sqrt(std::fabs(h.di));
      if (nhist == kRateWindow + 1) {
        for (int k = 0; k < kRateWindow; ++k) {
          rn_hist[k] = rn_hist[k + 1];
          it_hist[k] = it_hist[k + 1];
        }
        --nhist;
      }
      rn_hist[nhist] = rn;
      it_hist[nhist] = h.iter;
      ++nhist;
      gate = true;
      if (rn > h.conv && nhist >= 2 && h.iter > it_hist[0] && rn_hist[0] > 0.0) {
        const double rate = std::pow(rn / rn_hist[0], 1.0 / (double)(h.iter - it_hist[0]));
        if (rate < 1.0)
          gate = std::log(h.conv / rn) / std::log(rate) <= 2.0 * every;
        else  // no decay over the window: gate only close to the threshold
          gate = rn < 10.0 * h.conv;
      }
    }
  }
  if (!done && replacing) {  // maxit: leave the true residual of the final x
    hipLaunchKernelGGL(k_rel_force, dim3(1), dim3(64), 0, s, sc);
    replace(it);
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->scal_host, sc, sizeof(SolveScalars),
                                        hipMemcpyDeviceToHost, s));
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
  }
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  return ELASTIC_OK;
}


}  // namespace elastic
