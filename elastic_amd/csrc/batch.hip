// Multi-instance PCG for serving: many independent systems (each held by its
// own context: structure, values, b and vectors resident in HBM) iterate
// together.  Every phase of a PCG iteration is ONE launch over all active
// instances, one workgroup per item of a device work list:
//   SpMV          work item = (instance, tile)          block_tile of spmv_block.h
//                 (q = A p and each tile's share of p.q)
//   shares sum    one workgroup per instance: p.q in tile order
//   vector phases work item = (instance, chunk of whole nodes, <= 2048 rows)
// The grid of a list is a power-of-two size class of its length (workgroups
// past the length return at once), so a chunk of iterations is captured once
// as a HIP graph and replayed while instances are admitted and retired: only
// the work lists and the instance table (device memory) change between
// replays, and a new capture is needed only when a size class changes.
//
// The arithmetic of an instance never depends on the other instances: its
// tiles, its chunks and the order of its reductions (per-chunk partials summed
// in chunk order) are functions of its own size, so a result is bit for bit
// the same alone and in any mix.  The method per instance is the PCG of
// pcg.hip (Iters.f:252-524, x0 = 0, M-norm stopping rule) with Jacobi or,
// when its context asks for it ("batch_precond" 6), the node-block Jacobi
// of block_precond / op_block_precond (2x2 in 2-D, 3x3 in 3-D), in FP64
// or with the FP32 operator (FP32 values; staging and sums in FP64, or in
// FP32 when the context asks for it) and FP64 true-residual replacement.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "device_util.h"
#include "spmv_block.h"

namespace elastic {

int ensure_bdiag(elastic_ctx* x, int bsz);  // krylov.hip

constexpr int kChunk = 2048;         // rows per vector work item at most (256 lanes x 8)
constexpr int kMaxSlots = 255;       // slot index lives in the top 8 bits of a work item
// node blocks per lane: 2-D tiles of <= 512 blocks, 3-D tiles of <= 1024
template <int NB>
constexpr int batch_per() { return NB == 3 ? 4 : 2; }

struct BInst {
  BlkArgs a;
  const double *v64, *d64;
  const float *v32, *d32;
  MergeArgs mg;   // external partials of the instance's product -> node rows
  double* epart;
  double* tdot;   // ntiles shares of p.q written by the product's tiles
  double *x, *r, *p, *q, *b, *t;
  // Instances whose products sum in FP32 keep p, r and the Jacobi scale as
  // FP32 vectors while they use the FP32 operator (x, the true residuals and
  // all sums stay FP64); null otherwise.  Every true-residual replacement also
  // refreshes the FP64 r and p, so a hand-over to the FP64 operator (which
  // only happens there) continues from current FP64 vectors.
  float *p32, *r32, *m32, *dx32;
  const double* m;
  // node-block Jacobi (null: Jacobi with m): per node the nb(nb+1)/2 values
  // of block_precond's inverse (2x2: a11 a22 a21, 3x3: c11 c22 c33 c21 c31 c32)
  const double* bm;
  double* part;   // 5 * nch partials: (unused), r.Mr, true-residual r.Mr, p.r, (Mr).q
  double* dslot;  // 2 ping-pong slots of r.Mr
  SolveScalars* sc;
  int n, nb, nch, cr, fp32, acc32;  // cr: rows per chunk (nb * (kChunk / nb))
  int live, ntiles;  // live: the slot holds an admitted instance (checked first)
};

// ---------------------------------------------------------------- kernels
// z = B_J^-1 r of node J for the node-block Jacobi (op_block_precond's
// products, Precond.f:419-540)
template <int NB>
__device__ inline void node_apply(const double* __restrict__ bm, int J, const double* r,
                                  double* z) {
  if constexpr (NB == 2) {
    const double* a = bm + 3 * (int64_t)J;
    z[0] = a[0] * r[0] + a[2] * r[1];
    z[1] = a[2] * r[0] + a[1] * r[1];
  } else {
    const double* a = bm + 6 * (int64_t)J;
    z[0] = a[0] * r[0] + a[3] * r[1] + a[4] * r[2];
    z[1] = a[3] * r[0] + a[1] * r[1] + a[5] * r[2];
    z[2] = a[4] * r[0] + a[5] * r[1] + a[2] * r[2];
  }
}

__device__ inline int wl_slot(uint32_t e) { return (int)(e >> 24); }
__device__ inline int wl_index(uint32_t e) { return (int)(e & 0xFFFFFFu); }

// Sum of an instance's nch partials in chunk order (same value in every block).
__device__ inline double inst_sum(const double* __restrict__ part, int nch, double* sh) {
  double v = 0.0;
  for (int q = threadIdx.x; q < nch; q += kThreads) v += part[q];
  double r = block_sum(v, sh);
  if (threadIdx.x == 0) sh[4] = r;
  __syncthreads();
  r = sh[4];
  __syncthreads();
  return r;
}

// The product of one operator variant (value type TV, sum type TA, node block
// NB) over its work list, one work item per workgroup (a loop over items would
// keep per-item values live across iterations: 120 instead of 80 VGPRs for
// the FP64 3-D variant), so each variant compiles with the registers of the
// single-system kernel; the grid is a size class of the list (workgroups
// beyond the list's count return at once): q = A p and the tiles' shares of
// p.q.  The FP64 lists hold the FP64 instances, the FP32 lists the mixed
// instances, which run here while they are on the FP32 operator.  The FP64
// work of mixed instances (after the progress guard's hand-over, and the
// true-residual products) goes through device-built lists (k_bt_dyn_*).
//
// Every lane reads the instance's flags itself (uniform addresses that no
// kernel writes while this one runs, so all lanes see the same values); the
// tile's loads are issued first and the flags are checked while they are in
// flight.
template <typename TV, typename TA, int NB>
__global__ __launch_bounds__(256) void k_bt_spmv(const BInst* __restrict__ insts,
                                                 const uint32_t* __restrict__ work,
                                                 const int32_t* __restrict__ count) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool f64 = sizeof(TV) == 8;
  constexpr int PER = batch_per<NB>();
  const int w = blockIdx.x;
  if (w >= *count) return;
  const uint32_t e = work[w];
  const BInst& I = insts[wl_slot(e)];
  const int t = wl_index(e);
  if (!I.live || t >= I.ntiles) return;  // not a work item of an admitted instance
  const SolveScalars* sc = I.sc;
  const TV* vals = f64 ? (const TV*)I.v64 : (const TV*)I.v32;
  const TV* diag = f64 ? (const TV*)I.d64 : (const TV*)I.d32;
  TileRegs<TV, NB, PER> R;
  load_tile<TV, NB, PER>(I.a, t, vals, diag, R);
  const bool run = f64 ? !sc->done : (!sc->done && !sc->op64);
  if (!run) return;
  // the FP32-sum variant stages x from the FP32 copy of p and, for instances
  // with FP32 vectors, stores its own rows of q in FP32; one tile per
  // workgroup (FRESH)
  compute_tile<TV, TA, NB, PER, true, NoHook, 256, false, false, true>(
      I.a, t, R, I.p, I.q, I.epart, smem, I.tdot, (sizeof(TA) == 4 && I.p32) ? 1 : 0,
      sizeof(TA) == 4 ? I.p32 : nullptr);
}

// Device-built lists of mixed instances with FP64 work, so that a captured
// graph serves them without a grid per instance: list 2 * mode + (nb == 3),
// mode 0: instances the progress guard handed over to the FP64 operator
// (q = A64 p), mode 1: instances that asked for a true residual (t = A64 x,
// then r = b - t).  Built by one workgroup in slot-list order: the slots and
// the running totals of their tiles and chunks.
struct DynList {
  int32_t n;
  int32_t slot[kMaxSlots];
  int32_t tile_end[kMaxSlots];   // inclusive running total of ntiles
  int32_t chunk_end[kMaxSlots];  // of nch
};
constexpr int kDynLists = 4;
constexpr int kDynGrid = 1024;   // workgroups of the list kernels (loop over items)

__device__ inline bool dyn_wants(const BInst& I, int mode) {
  if (!I.live || !I.fp32) return false;
  const SolveScalars* sc = I.sc;
  return mode == 0 ? (sc->op64 && !sc->done) : (sc->upd && !sc->done);
}

__global__ __launch_bounds__(256) void k_bt_dyn_list(const BInst* __restrict__ insts,
                                                     const int32_t* __restrict__ slots,
                                                     const int32_t* __restrict__ count,
                                                     DynList* lists, int mode) {
  __shared__ int32_t f[2][256], tl[2][256], ch[2][256];
  const int ns = *count;  // <= kMaxSlots < 256
  const int w = threadIdx.x;
  int flag[2] = {0, 0}, nt = 0, nc = 0, k = -1;
  if (w < ns) {
    k = slots[w];
    const BInst& I = insts[k];
    if (dyn_wants(I, mode)) {
      flag[I.nb == 3 ? 1 : 0] = 1;
      nt = I.ntiles;
      nc = I.nch;
    }
  }
  // inclusive scans over the slot list (Hillis-Steele in LDS, fixed order)
  for (int l = 0; l < 2; ++l) {
    f[l][w] = flag[l];
    tl[l][w] = flag[l] ? nt : 0;
    ch[l][w] = flag[l] ? nc : 0;
  }
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    int a[2][3];
    for (int l = 0; l < 2; ++l) {
      a[l][0] = w >= d ? f[l][w - d] : 0;
      a[l][1] = w >= d ? tl[l][w - d] : 0;
      a[l][2] = w >= d ? ch[l][w - d] : 0;
    }
    __syncthreads();
    for (int l = 0; l < 2; ++l) {
      f[l][w] += a[l][0];
      tl[l][w] += a[l][1];
      ch[l][w] += a[l][2];
    }
    __syncthreads();
  }
  for (int l = 0; l < 2; ++l) {
    DynList& L = lists[2 * mode + l];
    if (flag[l]) {
      const int pos = f[l][w] - 1;
      L.slot[pos] = k;
      L.tile_end[pos] = tl[l][w];
      L.chunk_end[pos] = ch[l][w];
    }
    if (w == 255) L.n = f[l][255];
  }
}

// Item w of a list (tiles or chunks): the entry whose running total first
// exceeds w (entries in LDS; every lane walks the same path).
__device__ inline int dyn_find(const int32_t* end, int n, int w, int k) {
  while (k < n && end[k] <= w) ++k;
  return k;
}

// FP64 products of the listed instances, each workgroup looping over the
// list's tiles (mode 0: q = A64 p and the tiles' shares of p.q, mode 1:
// t = A64 x).  The flags are checked again per item: a list built before an
// instance finished (or before its slot was reused) skips it.
template <int NB>
__global__ __launch_bounds__(256) void k_bt_dyn_spmv(const BInst* __restrict__ insts,
                                                     const DynList* __restrict__ lists, int mode) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_slot[kMaxSlots], s_end[kMaxSlots];
  constexpr int PER = batch_per<NB>();
  const DynList& L = lists[2 * mode + (NB == 3 ? 1 : 0)];
  const int n = L.n;
  if (n == 0) return;
  const int total = L.tile_end[n - 1];
  if ((int)blockIdx.x >= total) return;
  for (int k = threadIdx.x; k < n; k += 256) {
    s_slot[k] = L.slot[k];
    s_end[k] = L.tile_end[k];
  }
  __syncthreads();
  int k = 0;
  for (int w = blockIdx.x; w < total; w += gridDim.x) {
    k = dyn_find(s_end, n, w, k);
    const BInst& I = insts[s_slot[k]];
    const int t = w - (k ? s_end[k - 1] : 0);
    if (t >= I.ntiles || !dyn_wants(I, mode)) continue;
    TileRegs<double, NB, PER> R;
    load_tile<double, NB, PER>(I.a, t, I.v64, I.d64, R);
    if (mode == 0)
      compute_tile<double, double, NB, PER, true>(I.a, t, R, I.p, I.q, I.epart, smem, I.tdot);
    else if (I.dx32)  // the product of x + dx32 (the FP32 increments since the last fold)
      compute_tile<double, double, NB, PER, false, NoHook, 256, true>(
          I.a, t, R, I.x, I.t, I.epart, smem, nullptr, 0, nullptr, NoHook(), I.dx32);
    else
      compute_tile<double, double, NB, PER>(I.a, t, R, I.x, I.t, I.epart, smem);
  }
}

// operator variants of the static lists: v = 3 * (nb == 3) + {0: FP64,
// 1: FP32 values FP64 sums, 2: FP32 values and sums}
constexpr int kVariants = 6;
static const void* variant_kernel(int v) {
  switch (v) {
    case 0: return (const void*)k_bt_spmv<double, double, 2>;
    case 1: return (const void*)k_bt_spmv<float, double, 2>;
    case 2: return (const void*)k_bt_spmv<float, float, 2>;
    case 3: return (const void*)k_bt_spmv<double, double, 3>;
    case 4: return (const void*)k_bt_spmv<float, double, 3>;
    default: return (const void*)k_bt_spmv<float, float, 3>;
  }
}
static void launch_variant(int v, dim3 G, dim3 TB, size_t lds, hipStream_t s, const BInst* insts,
                           const uint32_t* work, const int32_t* count) {
  switch (v) {
    case 0: hipLaunchKernelGGL((k_bt_spmv<double, double, 2>), G, TB, lds, s, insts, work, count); break;
    case 1: hipLaunchKernelGGL((k_bt_spmv<float, double, 2>), G, TB, lds, s, insts, work, count); break;
    case 2: hipLaunchKernelGGL((k_bt_spmv<float, float, 2>), G, TB, lds, s, insts, work, count); break;
    case 3: hipLaunchKernelGGL((k_bt_spmv<double, double, 3>), G, TB, lds, s, insts, work, count); break;
    case 4: hipLaunchKernelGGL((k_bt_spmv<float, double, 3>), G, TB, lds, s, insts, work, count); break;
    default: hipLaunchKernelGGL((k_bt_spmv<float, float, 3>), G, TB, lds, s, insts, work, count); break;
  }
}

// reads an instance flag into a block-uniform value: true when the work item
// (chunk c) is to be skipped; items of slots that hold no admitted instance
// and chunks beyond the instance's size are always skipped
__device__ inline bool uniform_done(const BInst& I, int c, double* sh, bool want_upd = false) {
  if (threadIdx.x == 0) {
    if (!I.live || c >= I.nch) {
      sh[5] = 1.0;
    } else {
      const SolveScalars* sc = I.sc;
      sh[5] = want_upd ? (double)(!sc->upd || sc->done) : (double)sc->done;
    }
  }
  __syncthreads();
  const bool d = sh[5] != 0.0;
  __syncthreads();
  return d;
}

// per instance (one workgroup each): p.q = the tiles' shares in tile order
__global__ __launch_bounds__(kThreads) void k_bt_sum_shares(const BInst* __restrict__ insts,
                                                            const int32_t* __restrict__ slots,
                                                            const int32_t* __restrict__ count) {
  __shared__ double sh[8];
  const int ns = *count;
  for (int w = blockIdx.x; w < ns; w += gridDim.x) {
    const BInst& I = insts[slots[w]];
    if (uniform_done(I, 0, sh)) continue;
    // lane l adds shares l, l + 256, ... in index order, eight loads in flight
    double v = 0.0;
    for (int k0 = threadIdx.x; k0 < I.ntiles; k0 += 8 * kThreads) {
      double t[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + j * kThreads;
        t[j] = k < I.ntiles ? I.tdot[k] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) v += t[j];
    }
    const double pq = block_sum(v, sh);
    if (threadIdx.x == 0) I.sc->pq_tiles = pq;
  }
}

// alpha = d / p.q (mixed: p.r / p.q) ; q = own rows + merged external
// partials (node by node, not stored) ; x += alpha p ; r -= alpha q ;
// partials r.(m r) and (mixed) (m r).q.  EP: type of the stored partials
// (FP32 after a product with FP32 sums).  V32: the instance keeps its
// iteration vectors in FP32 (p, r, the Jacobi scale, its product's own rows
// of q, and the steps of x as FP32 increments dx32 on the FP64 x, which
// every true-residual replacement folds into x).
template <int NB, typename EP, bool V32, typename AlphaFn>
__device__ inline double bt_update_chunk(const BInst& I, int c, double* sh, AlphaFn alpha_fn) {
  typedef typename std::conditional<V32, float, double>::type VT;
  const VT* P = V32 ? (const VT*)I.p32 : (const VT*)I.p;
  VT* R = V32 ? (VT*)I.r32 : (VT*)I.r;
  const VT* M = V32 ? (const VT*)I.m32 : (const VT*)I.m;
  const VT* Q = (const VT*)I.q;  // own rows as the product stored them
  VT* X = V32 ? (VT*)I.dx32 : (VT*)I.x;
  const int J0 = c * (I.cr / NB), J1 = min(I.n / NB, J0 + I.cr / NB);
  const double alpha = alpha_fn();
  double acc = 0.0, azq = 0.0;
  for (int J = J0 + threadIdx.x; J < J1; J += kThreads) {
    const int64_t i0 = (int64_t)NB * J;
    const int k0 = I.mg.ptr[J], k1 = I.mg.ptr[J + 1];
    double qv[NB], xv[NB], pv[NB], rv[NB], mv[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      qv[r] = (double)Q[i0 + r];
      xv[r] = (double)X[i0 + r];
      pv[r] = (double)P[i0 + r];
      rv[r] = (double)R[i0 + r];
      mv[r] = I.bm ? 0.0 : (double)M[i0 + r];  // node-block Jacobi applies bm below
    }
    for (int k = k0; k < k1; ++k) {
      const EP* e = reinterpret_cast<const EP*>(I.epart) + (int64_t)NB * I.mg.idx[k];
#pragma unroll
      for (int r = 0; r < NB; ++r) qv[r] += (double)e[r];
    }
    double rn[NB], zv[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      X[i0 + r] = (VT)(xv[r] + alpha * pv[r]);
      const VT rs = (VT)(rv[r] - alpha * qv[r]);
      R[i0 + r] = rs;
      rn[r] = (double)rs;
      zv[r] = rn[r] * mv[r];
    }
    if (I.bm) node_apply<NB>(I.bm, J, rn, zv);  // node-block Jacobi (FP64 vectors)
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      acc += rn[r] * zv[r];
      azq += zv[r] * qv[r];
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) I.part[I.nch + c] = s;
  if (I.fp32) {
    const double t = block_sum(azq, sh);
    if (threadIdx.x == 0) I.part[4 * I.nch + c] = t;
  }
  return alpha;
}

__global__ __launch_bounds__(kThreads) void k_bt_update(const BInst* __restrict__ insts,
                                                        const uint32_t* __restrict__ work,
                                                        const int32_t* __restrict__ count,
                                                        int cur) {
  __shared__ double sh[8];
  const int w = blockIdx.x;  // one work item per workgroup
  if (w >= *count) return;
  {
    const uint32_t e = work[w];
    const BInst& I = insts[wl_slot(e)];
    const int c = wl_index(e);
    if (uniform_done(I, c, sh)) return;
    auto alpha_fn = [&]() {
      const double num = I.fp32 ? inst_sum(I.part + 3 * I.nch, I.nch, sh) : I.dslot[cur];
      return num / I.sc->pq_tiles;
    };
    // the FP32-sum variant served this instance's product (see k_bt_spmv);
    // FP32 vectors only with it (p32 is set for such instances only)
    const bool e32 = I.fp32 && I.acc32 && !I.sc->op64;
    const bool v32 = e32 && I.p32 != nullptr;
    double alpha;
    if (I.nb == 3) {
      if (v32) alpha = bt_update_chunk<3, float, true>(I, c, sh, alpha_fn);
      else if (e32) alpha = bt_update_chunk<3, float, false>(I, c, sh, alpha_fn);
      else alpha = bt_update_chunk<3, double, false>(I, c, sh, alpha_fn);
    } else {
      if (v32) alpha = bt_update_chunk<2, float, true>(I, c, sh, alpha_fn);
      else if (e32) alpha = bt_update_chunk<2, float, false>(I, c, sh, alpha_fn);
      else alpha = bt_update_chunk<2, double, false>(I, c, sh, alpha_fn);
    }
    if (c == 0 && threadIdx.x == 0) {
      I.sc->alpha = alpha;
      I.sc->pq = I.sc->pq_tiles;
    }
  }
}

// beta = r.Mr / d (mixed: -(Mr).q / p.q) ; p = M r + beta p ; (mixed) partial
// p.r of the new direction ; chunk 0 keeps the books
__global__ __launch_bounds__(kThreads) void k_bt_pupdate(const BInst* __restrict__ insts,
                                                         const uint32_t* __restrict__ work,
                                                         const int32_t* __restrict__ count,
                                                         int cur) {
  __shared__ double sh[8];
  const int w = blockIdx.x;  // one work item per workgroup
  if (w >= *count) return;
  {
    const uint32_t e = work[w];
    const BInst& I = insts[wl_slot(e)];
    const int c = wl_index(e);
    if (uniform_done(I, c, sh)) return;
    const int i0 = c * I.cr, i1 = min(I.n, i0 + I.cr);
    const bool v32 = I.p32 && !I.sc->op64;
    // FP32 vectors: the lane's rows (at most kChunk / kThreads) are loaded
    // before beta's sums, so that their latency overlaps
    constexpr int kRows = kChunk / kThreads;
    float rf[kRows], mf[kRows], pf[kRows];
    if (v32) {
#pragma unroll
      for (int u = 0; u < kRows; ++u) {
        const int i = i0 + threadIdx.x + u * kThreads;
        if (i < i1) {
          rf[u] = I.r32[i];
          mf[u] = I.m32[i];
          pf[u] = I.p32[i];
        }
      }
    }
    const double di = inst_sum(I.part + I.nch, I.nch, sh);
    // restart (beta = 0) right after the hand-over to the FP64 operator
    const bool restart = I.sc->restart[cur] != 0;
    const double beta =
        restart ? 0.0
        : I.fp32 ? -inst_sum(I.part + 4 * I.nch, I.nch, sh) / I.sc->pq : di / I.dslot[cur];
    double apr = 0.0;
    if (v32) {  // FP32 vectors (see BInst)
#pragma unroll
      for (int u = 0; u < kRows; ++u) {
        const int i = i0 + threadIdx.x + u * kThreads;
        if (i < i1) {
          const double ri = (double)rf[u];
          const float pi = (float)(ri * (double)mf[u] + beta * (double)pf[u]);
          I.p32[i] = pi;
          apr += (double)pi * ri;
        }
      }
    } else if (I.bm) {  // node-block Jacobi: one lane per node of the chunk
      const int nb = I.nb;
      for (int J = i0 / nb + threadIdx.x; J < i1 / nb; J += kThreads) {
        double rv[3], zv[3];
        for (int r = 0; r < nb; ++r) rv[r] = I.r[nb * J + r];
        if (nb == 3) node_apply<3>(I.bm, J, rv, zv);
        else node_apply<2>(I.bm, J, rv, zv);
        for (int r = 0; r < nb; ++r) {
          const double pi = zv[r] + beta * I.p[nb * J + r];
          I.p[nb * J + r] = pi;
          apr += pi * rv[r];
        }
      }
    } else {
      for (int i = i0 + threadIdx.x; i < i1; i += kThreads) {
        const double ri = I.r[i];
        const double pi = ri * I.m[i] + beta * I.p[i];
        I.p[i] = pi;
        apr += pi * ri;
      }
    }
    if (I.fp32) {  // flexible recurrences of the mixed method
      const double t = block_sum(apr, sh);
      if (threadIdx.x == 0) I.part[3 * I.nch + c] = t;
    }
    if (c == 0 && threadIdx.x == 0) {
      SolveScalars* sc = I.sc;
      sc->restart[cur ^ 1] = 0;  // the previous step's flag (no chunk reads it now)
      if (restart) {  // plain FP64 PCG from here: no more replacements
        sc->reliable = 0;
        sc->upd = 0;
      }
      I.dslot[cur ^ 1] = di;
      sc->beta = beta;
      sc->di = di;
      sc->iter += 1;
      const double rn = sqrt(fabs(di));
      if (!sc->reliable) {
        if (rn < sc->conv) sc->done = 1;
      } else if (rn < sc->delta * sc->rmax || rn < sc->conv ||
                 sc->iter - sc->last_upd >= kForceReplace) {
        sc->upd = 1;
        sc->skip64 = 0;
      }
      if (!sc->done && sc->iter >= sc->maxit) {
        sc->done = 1;
        sc->status = ELASTIC_MAXIT;
      }
    }
  }
}

// true residual r = b - (t + merged partials) of chunk c of an instance that
// asked for it; partials r.(m r) and p.r.  Called by every lane of the
// workgroup (block sums inside).
__device__ inline void rel_resid_chunk(const BInst& I, int c, double* sh) {
  const int i0 = c * I.cr, i1 = min(I.n, i0 + I.cr);
  const bool v32 = I.p32 && !I.sc->op64;
  double acc = 0.0, apr = 0.0;
  if (I.bm) {  // node-block Jacobi (FP64 vectors): one lane per node
    const int nb = I.nb;
    for (int J = i0 / nb + threadIdx.x; J < i1 / nb; J += kThreads) {
      double rv[3], zv[3];
      for (int r = 0; r < nb; ++r) {
        const int i = nb * J + r;
        rv[r] = I.b[i] - merge_row(I.mg, i, I.t[i]);
        I.r[i] = rv[r];
        apr += I.p[i] * rv[r];
      }
      if (nb == 3) node_apply<3>(I.bm, J, rv, zv);
      else node_apply<2>(I.bm, J, rv, zv);
      for (int r = 0; r < nb; ++r) acc += rv[r] * zv[r];
    }
  }
  for (int i = i0 + threadIdx.x; i < (I.bm ? i0 : i1); i += kThreads) {
    double s = I.t[i];
    s = merge_row(I.mg, i, s);
    const double ri = I.b[i] - s;
    I.r[i] = ri;
    acc += ri * (ri * I.m[i]);
    if (v32) {  // FP32 vectors: store the rounded residual; refresh the FP64 p
      if (I.dx32) {  // the product was formed on x + dx32: fold the increments
        I.x[i] = I.x[i] + (double)I.dx32[i];
        I.dx32[i] = 0.0f;
      }
      const float rs = (float)ri;
      I.r32[i] = rs;
      const double pi = (double)I.p32[i];
      I.p[i] = pi;
      apr += pi * (double)rs;
    } else {
      apr += I.p[i] * ri;
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) I.part[2 * I.nch + c] = s;
  const double t = block_sum(apr, sh);  // only mixed instances ask for true residuals
  if (threadIdx.x == 0) I.part[3 * I.nch + c] = t;
}

// the true-residual chunks of the instances on list 2 + nb3 (mode 1), each
// workgroup looping over the list's chunks
__global__ __launch_bounds__(kThreads) void k_bt_dyn_resid(const BInst* __restrict__ insts,
                                                           const DynList* __restrict__ lists,
                                                           int nb3) {
  __shared__ double sh[8];
  __shared__ int32_t s_slot[kMaxSlots], s_end[kMaxSlots];
  const DynList& L = lists[2 + nb3];
  const int n = L.n;
  if (n == 0) return;
  const int total = L.chunk_end[n - 1];
  if ((int)blockIdx.x >= total) return;
  for (int k = threadIdx.x; k < n; k += kThreads) {
    s_slot[k] = L.slot[k];
    s_end[k] = L.chunk_end[k];
  }
  __syncthreads();
  int k = 0;
  for (int w = blockIdx.x; w < total; w += gridDim.x) {
    k = dyn_find(s_end, n, w, k);
    const BInst& I = insts[s_slot[k]];
    const int c = w - (k ? s_end[k - 1] : 0);
    if (c >= I.nch || !dyn_wants(I, 1)) continue;
    rel_resid_chunk(I, c, sh);
  }
}

// per instance: accept the true residual (done) or resume from it
__global__ __launch_bounds__(kThreads) void k_bt_rel_finish(const BInst* __restrict__ insts,
                                                            const int32_t* __restrict__ slots,
                                                            const int32_t* __restrict__ count,
                                                            int nxt) {
  __shared__ double sh[8];
  const int ns = *count;
  for (int w = blockIdx.x; w < ns; w += gridDim.x) {
    const BInst& I = insts[slots[w]];
    if (uniform_done(I, 0, sh, true)) continue;
    const double rz = inst_sum(I.part + 2 * I.nch, I.nch, sh);
    if (threadIdx.x == 0) {
      SolveScalars* sc = I.sc;
      sc->true_rmr = rz;
      sc->nupd += 1;
      const int32_t was64 = sc->op64;
      progress_guard(sc, sqrt(fabs(rz)));
      if (!was64 && sc->op64) sc->restart[nxt] = 1;  // read by the next p update
      if (sqrt(fabs(rz)) < sc->conv) {
        sc->done = 1;
        sc->status = 0;
      } else {
        I.dslot[nxt] = rz;
        sc->rmax = sqrt(fabs(rz));
        sc->upd = 0;
        sc->skip64 = 1;
      }
    }
  }
}

// admission of one instance (eager, grid = its chunks): x = 0, r = b, p = M b,
// partial b.Mb (= p.r, the first numerator of the mixed method)
__global__ __launch_bounds__(kThreads) void k_bt_init(const BInst* __restrict__ insts, int slot) {
  __shared__ double sh[8];
  const BInst& I = insts[slot];
  const int c = blockIdx.x;
  const int i0 = c * I.cr, i1 = min(I.n, i0 + I.cr);
  double acc = 0.0, apr = 0.0;
  if (I.bm) {  // node-block Jacobi (FP64 vectors): p = B^-1 b, one lane per node
    const int nb = I.nb;
    for (int J = i0 / nb + threadIdx.x; J < i1 / nb; J += kThreads) {
      double bv[3], zv[3];
      for (int r = 0; r < nb; ++r) bv[r] = I.b[nb * J + r];
      if (nb == 3) node_apply<3>(I.bm, J, bv, zv);
      else node_apply<2>(I.bm, J, bv, zv);
      for (int r = 0; r < nb; ++r) {
        const int i = nb * J + r;
        I.x[i] = 0.0;
        I.r[i] = bv[r];
        I.p[i] = zv[r];
        acc += bv[r] * zv[r];
      }
    }
  }
  for (int i = i0 + threadIdx.x; i < (I.bm ? i0 : i1); i += kThreads) {
    const double bi = I.b[i], zi = bi * I.m[i];
    I.x[i] = 0.0;
    I.r[i] = bi;
    I.p[i] = zi;
    acc += bi * zi;
    if (I.p32) {  // FP32 vectors (see BInst)
      I.p32[i] = (float)zi;
      I.dx32[i] = 0.0f;
      I.r32[i] = (float)bi;
      I.m32[i] = (float)I.m[i];
      apr += (double)(float)bi * (double)(float)zi;
    }
  }
  const double s = block_sum(acc, sh);
  const double t = I.p32 ? block_sum(apr, sh) : s;
  if (threadIdx.x == 0) {
    I.part[I.nch + c] = s;
    I.part[3 * I.nch + c] = t;
  }
}

__global__ void k_bt_init2(const BInst* __restrict__ insts, int slot, double tol, int maxit,
                           double delta) {
  __shared__ double sh[8];
  const BInst& I = insts[slot];
  const double d = inst_sum(I.part + I.nch, I.nch, sh);
  if (threadIdx.x == 0) {
    SolveScalars* sc = I.sc;
    I.dslot[0] = d;  // steps start at even iteration parity
    I.dslot[1] = d;
    sc->conv = tol * sqrt(fabs(d));
    sc->norm_b = sqrt(fabs(d));
    sc->rz0 = d;
    sc->rmax = sqrt(fabs(d));
    sc->delta = delta;
    sc->true_rmr = d;
    sc->iter = 0;
    sc->maxit = maxit;
    sc->status = 0;
    sc->reliable = I.fp32;
    sc->upd = 0;
    sc->skip64 = 1;
    sc->nupd = 0;
    sc->pq_tiles = 0.0;
    progress_init(sc, sqrt(fabs(d)));
    sc->done = (d == 0.0) ? 1 : 0;
  }
}

// x += dx32, dx32 = 0: the FP32 increments of an instance that stopped
// without a final replacement (maxit, or a batch closed under it)
__global__ __launch_bounds__(kThreads) void k_bt_fold(int n, double* x, float* dx32) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    x[i] = x[i] + (double)dx32[i];
    dx32[i] = 0.0f;
  }
}

// done flags / iterations of the active slots for the host
struct BStatus {
  int32_t done, status, iter, nupd, op64_from;
};
__global__ void k_bt_status(const BInst* __restrict__ insts, const int32_t* __restrict__ slots,
                            const int32_t* __restrict__ count, BStatus* out) {
  const int ns = *count;
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < ns; w += gridDim.x * blockDim.x) {
    const BInst& I = insts[slots[w]];
    if (!I.live) {
      out[w] = BStatus{0, 0, 0, 0, -1};
      continue;
    }
    const SolveScalars* sc = I.sc;
    out[w] = BStatus{sc->done, sc->status, sc->iter, sc->nupd, sc->op64_from};
  }
}

}  // namespace elastic

using namespace elastic;

// ------------------------------------------------------------------ host
struct elastic_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int max_slots = 0, iters = 16;
  bool use_graph = true;
  double delta = 0.1;
  struct Slot {
    bool used = false;
    elastic_ctx* ctx = nullptr;
    int64_t tag = 0;
    int ntiles = 0, nch = 0;
    size_t lds = 0;
    std::chrono::steady_clock::time_point t0;
  };
  std::vector<Slot> slots;
  std::vector<BInst> h_insts;
  BInst* d_insts = nullptr;
  // one device buffer per product variant: a captured graph holds these
  // pointers, so a list may only change in place (a reallocation re-captures)
  uint32_t* d_sw[kVariants] = {nullptr};
  size_t sw_cap[kVariants] = {0};
  uint32_t* d_vwork = nullptr;
  // counts: the kVariants product lists, then vector items, then slots
  int32_t *d_slist = nullptr, *d_counts = nullptr;
  size_t vwork_cap = 0;
  std::vector<uint32_t> h_sw[kVariants], h_vwork;
  std::vector<int32_t> h_slist, h_counts;
  // non-empty variant lists, + bit 8: mixed instances present, bits 9 / 10:
  // 2-D / 3-D mixed instances present (their device-built lists are served)
  unsigned mask = 0, graph_mask = ~0u;
  DynList* d_dyn = nullptr;  // kDynLists device-built lists (see DynList)
  // grid of each product variant: a power-of-two size class of its list
  // (one workgroup per item); a captured graph holds the classes it was
  // captured with
  int sgrid[kVariants] = {0}, graph_sgrid[kVariants] = {0};
  int ncu = 0;
  int vgrid = 0, graph_vgrid = 0;  // size class of the vector list (one chunk per workgroup)
  BStatus *d_status = nullptr, *h_status = nullptr;  // h_status pinned
  bool dirty = true, stepped = false;
  size_t lds = 0, graph_lds = 0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
};

static int berr(elastic_batch* b, int code, const std::string& m) {
  b->err = m;
  return code;
}
#define BATCH_HIP(b, expr)                                                                   \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return berr((b), ELASTIC_HIP_ERROR, std::string(#expr) + ": " +   \
                                                                   hipGetErrorString(e_));   \
  } while (0)

// LDS of one tile's product with FP64 sums (FP32 sums take half)
static size_t tile_lds(int nb, int ncols_max, int nblk_max) {
  return tile_smem(sizeof(double), nb, ncols_max, nblk_max);
}
// the dynamic LDS of variant v's launch: the batch's largest tile with the
// variant's sum type
static size_t variant_lds(const elastic_batch* b, int v) { return v % 3 == 2 ? b->lds / 2 : b->lds; }
// static LDS of the list kernels (slots and running totals)
constexpr size_t kDynStaticLds = 2 * kMaxSlots * sizeof(int32_t) + 64;

static void launch_dyn_spmv(elastic_batch* b, int nb3, int mode) {
  if (nb3)
    hipLaunchKernelGGL((k_bt_dyn_spmv<3>), dim3(kDynGrid), dim3(256), b->lds, b->stream,
                       b->d_insts, b->d_dyn, mode);
  else
    hipLaunchKernelGGL((k_bt_dyn_spmv<2>), dim3(kDynGrid), dim3(256), b->lds, b->stream,
                       b->d_insts, b->d_dyn, mode);
}

// one PCG iteration of every active instance; `it` is the iteration index in
// the step (its parity selects the r.Mr slot)
static void enqueue_iteration(elastic_batch* b, int it) {
  hipStream_t s = b->stream;
  const dim3 TB(kThreads);
  const int cur = it & 1;
  const int32_t* vcount = b->d_counts + kVariants;
  const int32_t* scount = vcount + 1;
  // mixed instances: FP64 products of those the progress guard handed over.
  // A hand-over happens only in the true-residual pass that ends odd
  // iterations, and a step starts at an even one, so the list is rebuilt on
  // the device at every even iteration: a hand-over inside a captured step is
  // served from the next iteration on.
  const bool mixed = (b->mask & 0x100u) != 0;
  if (mixed) {
    if (cur == 0)
      hipLaunchKernelGGL(k_bt_dyn_list, dim3(1), dim3(256), 0, s, b->d_insts, b->d_slist, scount,
                         b->d_dyn, 0);
    for (int nb3 = 0; nb3 < 2; ++nb3)
      if (b->mask & (nb3 ? 0x400u : 0x200u))
        launch_dyn_spmv(b, nb3, 0);
  }
  for (int v = 0; v < kVariants; ++v)
    if (b->mask & (1u << v))
      launch_variant(v, dim3(b->sgrid[v]), TB, variant_lds(b, v), s, b->d_insts, b->d_sw[v],
                     b->d_counts + v);
  hipLaunchKernelGGL(k_bt_sum_shares, dim3(kMaxSlots), TB, 0, s, b->d_insts, b->d_slist, scount);
  hipLaunchKernelGGL(k_bt_update, dim3(b->vgrid), TB, 0, s, b->d_insts, b->d_vwork, vcount, cur);
  hipLaunchKernelGGL(k_bt_pupdate, dim3(b->vgrid), TB, 0, s, b->d_insts, b->d_vwork, vcount, cur);
  if (cur == 1 && mixed) {  // every second iteration: true residuals where asked
    hipLaunchKernelGGL(k_bt_dyn_list, dim3(1), dim3(256), 0, s, b->d_insts, b->d_slist, scount,
                       b->d_dyn, 1);
    for (int nb3 = 0; nb3 < 2; ++nb3)
      if (b->mask & (nb3 ? 0x400u : 0x200u)) {
        launch_dyn_spmv(b, nb3, 1);
        hipLaunchKernelGGL(k_bt_dyn_resid, dim3(kDynGrid), TB, 0, s, b->d_insts, b->d_dyn, nb3);
      }
    hipLaunchKernelGGL(k_bt_rel_finish, dim3(kMaxSlots), TB, 0, s, b->d_insts, b->d_slist,
                       scount, 0);
  }
}

static int set_lds_limit(elastic_batch* b) {
  if (b->lds + kDynStaticLds > 64 * 1024) {
    for (int v = 0; v < kVariants; ++v)
      BATCH_HIP(b, hipFuncSetAttribute(variant_kernel(v),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    BATCH_HIP(b, hipFuncSetAttribute((const void*)k_bt_dyn_spmv<2>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024));
    BATCH_HIP(b, hipFuncSetAttribute((const void*)k_bt_dyn_spmv<3>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024));
  }
  return ELASTIC_OK;
}

static int upload_lists(elastic_batch* b) {
  for (auto& l : b->h_sw) l.clear();
  b->h_vwork.clear();
  b->h_slist.clear();
  size_t lds = 0;
  bool mixed = false;
  unsigned mixed_nb = 0;
  for (int k = 0; k < b->max_slots; ++k) {
    const auto& S = b->slots[k];
    if (!S.used) continue;
    const BInst& I = b->h_insts[k];
    b->h_slist.push_back(k);
    const int v0 = I.nb == 3 ? 3 : 0;
    // FP64 instances on the FP64 list; mixed ones on an FP32 list (their FP64
    // work goes through the device-built lists)
    const int v = I.fp32 ? v0 + (I.acc32 ? 2 : 1) : v0;
    for (int t = 0; t < S.ntiles; ++t) b->h_sw[v].push_back(((uint32_t)k << 24) | (uint32_t)t);
    if (I.fp32) {
      mixed = true;
      mixed_nb |= I.nb == 3 ? 0x400u : 0x200u;
    }
    for (int c = 0; c < S.nch; ++c) b->h_vwork.push_back(((uint32_t)k << 24) | (uint32_t)c);
    lds = std::max(lds, S.lds);
  }
  unsigned mask = mixed ? 0x100u | mixed_nb : 0u;
  for (int v = 0; v < kVariants; ++v)
    if (!b->h_sw[v].empty()) mask |= 1u << v;
  // grid size classes: the count rounded up to 16 steps per octave (at most
  // 1/16 of the next power of two idle), kept while the count stays above
  // 3/4 of it (fewer re-captures as instances leave)
  auto size_class = [](int n, int cur) {
    if (n <= 0) return 0;
    int g = 64;
    while (g < n) g <<= 1;
    const int step = std::max(8, g / 16);
    const int c = (n + step - 1) / step * step;
    return (cur >= n && 4 * n > 3 * cur) ? cur : c;
  };
  for (int v = 0; v < kVariants; ++v) b->sgrid[v] = size_class((int)b->h_sw[v].size(), b->sgrid[v]);
  b->vgrid = size_class((int)b->h_vwork.size(), b->vgrid);
  auto grow = [&](uint32_t*& p, size_t& cap, size_t need) -> int {
    if (need <= cap) return ELASTIC_OK;
    size_t nc = std::max(need, 2 * cap + 4096);
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc((void**)&p, nc * sizeof(uint32_t)) != hipSuccess)
      return berr(b, ELASTIC_OOM, "batch work list");
    cap = nc;
    b->graph_lds = 0;  // pointers changed: capture again
    return ELASTIC_OK;
  };
  int rc = 0;
  for (int v = 0; v < kVariants && !rc; ++v) rc = grow(b->d_sw[v], b->sw_cap[v], b->h_sw[v].size());
  if (rc) return rc;
  rc = grow(b->d_vwork, b->vwork_cap, b->h_vwork.size());
  if (rc) return rc;
  b->h_counts.assign(kVariants + 2, 0);
  for (int v = 0; v < kVariants; ++v) b->h_counts[v] = (int32_t)b->h_sw[v].size();
  b->h_counts[kVariants] = (int32_t)b->h_vwork.size();
  b->h_counts[kVariants + 1] = (int32_t)b->h_slist.size();
  hipStream_t s = b->stream;
  for (int v = 0; v < kVariants; ++v)
    if (!b->h_sw[v].empty())
      BATCH_HIP(b, hipMemcpyAsync(b->d_sw[v], b->h_sw[v].data(), b->h_sw[v].size() * 4,
                                  hipMemcpyHostToDevice, s));
  if (!b->h_vwork.empty())
    BATCH_HIP(b, hipMemcpyAsync(b->d_vwork, b->h_vwork.data(), b->h_vwork.size() * 4,
                                hipMemcpyHostToDevice, s));
  if (!b->h_slist.empty())
    BATCH_HIP(b, hipMemcpyAsync(b->d_slist, b->h_slist.data(), b->h_slist.size() * 4,
                                hipMemcpyHostToDevice, s));
  BATCH_HIP(b, hipMemcpyAsync(b->d_counts, b->h_counts.data(), b->h_counts.size() * 4,
                              hipMemcpyHostToDevice, s));
  // the host vectors must outlive the copies: wait here (admission time only)
  BATCH_HIP(b, hipStreamSynchronize(s));
  if (lds > b->lds) b->lds = lds;
  b->mask = mask;
  b->dirty = false;
  return ELASTIC_OK;
}

static int capture(elastic_batch* b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  int rc = set_lds_limit(b);
  if (rc) return rc;
  BATCH_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < b->iters; ++it) enqueue_iteration(b, it);
  BATCH_HIP(b, hipStreamEndCapture(b->stream, &b->graph));
  BATCH_HIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  b->graph_lds = b->lds;
  b->graph_mask = b->mask;
  for (int v = 0; v < kVariants; ++v) b->graph_sgrid[v] = b->sgrid[v];
  b->graph_vgrid = b->vgrid;
  return ELASTIC_OK;
}

// folds an instance's FP32 increments of x (no-op without them)
static void fold_increments(const BInst& I, hipStream_t s) {
  if (!I.dx32) return;
  const int g = std::max(1, std::min((I.n + kThreads - 1) / kThreads, 2048));
  hipLaunchKernelGGL(k_bt_fold, dim3(g), dim3(kThreads), 0, s, I.n, I.x, I.dx32);
}

extern "C" {

int elastic_batch_create(elastic_batch** out, int32_t device, int32_t max_slots,
                         int32_t iters_per_step, int32_t use_graph) {
  if (!out || max_slots < 1 || max_slots > kMaxSlots || iters_per_step < 2) return ELASTIC_BAD_ARG;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return ELASTIC_HIP_ERROR;
  auto* b = new elastic_batch();
  b->device = device;
  b->max_slots = max_slots;
  b->iters = iters_per_step & ~1;  // even: steps start at even parity
  b->use_graph = use_graph != 0;
  if (const char* e = getenv("ELASTIC_RELIABLE_DELTA")) b->delta = atof(e);
  // ELASTIC_BATCH_GRAPH=0: steps are eager launches (for profilers)
  if (const char* e = getenv("ELASTIC_BATCH_GRAPH")) b->use_graph = b->use_graph && atoi(e) != 0;
  b->slots.resize(max_slots);
  b->h_insts.resize(max_slots);
  if (hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      b->ncu < 1)
    b->ncu = 256;
  std::memset(b->h_insts.data(), 0, sizeof(BInst) * max_slots);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&b->d_insts, sizeof(BInst) * max_slots) != hipSuccess ||
      hipMalloc((void**)&b->d_slist, sizeof(int32_t) * max_slots) != hipSuccess ||
      hipMalloc((void**)&b->d_counts, sizeof(int32_t) * (kVariants + 2)) != hipSuccess ||
      hipMalloc((void**)&b->d_status, sizeof(BStatus) * max_slots) != hipSuccess ||
      hipMalloc((void**)&b->d_dyn, sizeof(DynList) * kDynLists) != hipSuccess ||
      hipMemset(b->d_dyn, 0, sizeof(DynList) * kDynLists) != hipSuccess ||
      hipHostMalloc((void**)&b->h_status, sizeof(BStatus) * max_slots) != hipSuccess) {
    elastic_batch_destroy(b);
    return ELASTIC_HIP_ERROR;
  }
  if (hipMemcpy(b->d_insts, b->h_insts.data(), sizeof(BInst) * max_slots,
                hipMemcpyHostToDevice) != hipSuccess) {
    elastic_batch_destroy(b);
    return ELASTIC_HIP_ERROR;
  }
  *out = b;
  return ELASTIC_OK;
}

int elastic_batch_destroy(elastic_batch* b) {
  if (!b) return ELASTIC_OK;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  // contexts of instances never retired go back to the CSRC order (with the
  // x their iteration reached)
  for (size_t k = 0; k < b->slots.size(); ++k) {
    auto& S = b->slots[k];
    if (S.used && S.ctx) {
      fold_increments(b->h_insts[k], S.ctx->c.stream);
      (void)leave_layout_order(&S.ctx->c);
      hipStreamSynchronize(S.ctx->c.stream);
    }
  }
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (void* p : {(void*)b->d_insts, (void*)b->d_vwork, (void*)b->d_slist, (void*)b->d_counts,
                  (void*)b->d_status, (void*)b->d_dyn})
    if (p) hipFree(p);
  for (uint32_t* p : b->d_sw)
    if (p) hipFree(p);
  if (b->h_status) hipHostFree(b->h_status);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return ELASTIC_OK;
}

const char* elastic_batch_last_error(const elastic_batch* b) {
  return b ? b->err.c_str() : "null batch";
}

int elastic_batch_submit(elastic_batch* b, elastic_ctx* x, int32_t precision, double tol,
                         int32_t maxit, int64_t tag) {
  if (!b || !x) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!c->values_set) return berr(b, ELASTIC_NOT_READY, "context has no values");
  if (c->layout_nb < 2) return berr(b, ELASTIC_BAD_ARG, "batched solves need a node-blocked layout");
  if (c->blk.nblk_max > 256 * (c->layout_nb == 3 ? batch_per<3>() : batch_per<2>()))
    return berr(b, ELASTIC_BAD_ARG, "tiles longer than the batch product holds (512 blocks in 2-D, 1024 in 3-D)");
  if (tile_lds(c->layout_nb, c->blk.ncols_max, c->blk.nblk_max) + kDynStaticLds > 156 * 1024)
    return berr(b, ELASTIC_BAD_ARG, "tiles with more local columns than the batch product's LDS holds");
  if (c->zero_diag_row > 0) return berr(b, ELASTIC_ZERO_DIAG, "zero diagonal coefficient");
  if (precision != ELASTIC_FP64 && precision != ELASTIC_MIXED) return ELASTIC_BAD_ARG;
  if (maxit <= 0) return ELASTIC_BAD_ARG;
  if (tol <= 0.0) tol = 2.220446049250313e-16;
  if (!c->tdot) return berr(b, ELASTIC_NOT_READY, "context has no tile shares buffer");
  const int cr = c->layout_nb * (kChunk / c->layout_nb);  // whole nodes per chunk
  const int nch = (c->neq + cr - 1) / cr;
  if (5 * nch > 8 * kRedBlocks) return berr(b, ELASTIC_BAD_ARG, "system too large for a batch slot");
  if (c->blk.ntiles >= (1 << 24) || nch >= (1 << 24)) return ELASTIC_BAD_ARG;
  int k = -1;
  for (int i = 0; i < b->max_slots; ++i)
    if (!b->slots[i].used) {
      k = i;
      break;
    }
  if (k < 0) return berr(b, ELASTIC_BUSY, "no free slot");
  hipSetDevice(b->device);
  // node-block Jacobi when the context asks for it (FP64 vectors then)
  const double* bm = nullptr;
  if (c->batch_precond == ELASTIC_PRECOND_BDIAG3) {
    int rc = ensure_bdiag(x, c->layout_nb);
    if (rc) return berr(b, rc, c->err);
    bm = solver_bm(c);
  }
  // a compact layout iterates in its node order until the instance retires
  {
    int rc = enter_layout_order(c);
    if (rc) return berr(b, rc, c->err);
  }
  // the context's own work (uploads, assembly) must be complete
  BATCH_HIP(b, hipStreamSynchronize(c->stream));
  BInst I;
  std::memset(&I, 0, sizeof(I));
  I.a = blk_args(c->blk);
  I.v64 = c->blk.v64;
  I.d64 = c->blk.d64;
  I.v32 = c->blk.v32;
  I.d32 = c->blk.d32;
  I.mg = merge_of(c, c->epart64);
  I.epart = c->epart64;
  I.tdot = c->tdot;
  I.x = c->v64[VX];
  I.r = c->v64[VR];
  I.p = c->v64[VP];
  I.q = c->v64[VQ];
  I.b = c->v64[VB];
  I.t = c->v64[VT];
  // FP32 vectors only for instances whose products sum in FP32
  const bool v32 = precision == ELASTIC_MIXED && c->accum32 && !bm;
  I.p32 = v32 ? c->p32 : nullptr;
  I.r32 = v32 ? c->p32 + vec32_stride(c->neq) : nullptr;
  I.m32 = v32 ? c->p32 + 2 * vec32_stride(c->neq) : nullptr;
  I.dx32 = v32 ? c->p32 + 3 * vec32_stride(c->neq) : nullptr;
  I.m = c->A64.m;
  I.bm = bm;
  I.part = c->red64;
  I.dslot = c->red64 + 8 * kRedBlocks + 8;
  I.sc = c->scal;
  I.n = c->neq;
  I.nb = c->layout_nb;
  I.nch = nch;
  I.cr = cr;
  I.fp32 = precision == ELASTIC_MIXED;
  I.acc32 = c->accum32 ? 1 : 0;
  I.live = 1;
  I.ntiles = c->blk.ntiles;
  b->h_insts[k] = I;
  BATCH_HIP(b, hipMemcpyAsync(b->d_insts + k, &b->h_insts[k], sizeof(BInst), hipMemcpyHostToDevice,
                              b->stream));
  hipLaunchKernelGGL(k_bt_init, dim3(nch), dim3(kThreads), 0, b->stream, b->d_insts, k);
  hipLaunchKernelGGL(k_bt_init2, dim3(1), dim3(kThreads), 0, b->stream, b->d_insts, k, tol, maxit,
                     b->delta);
  auto& S = b->slots[k];
  S.used = true;
  S.ctx = x;
  S.tag = tag;
  S.ntiles = c->blk.ntiles;
  S.nch = nch;
  S.lds = tile_lds(c->layout_nb, c->blk.ncols_max, c->blk.nblk_max);
  S.t0 = std::chrono::steady_clock::now();
  b->dirty = true;
  return ELASTIC_OK;
}

int elastic_batch_active(elastic_batch* b, int32_t* n) {
  if (!b || !n) return ELASTIC_BAD_ARG;
  int a = 0;
  for (auto& S : b->slots) a += S.used;
  *n = a;
  return ELASTIC_OK;
}

int elastic_batch_step(elastic_batch* b) {
  if (!b) return ELASTIC_BAD_ARG;
  hipSetDevice(b->device);
  if (b->dirty) {
    int rc = upload_lists(b);
    if (rc) return rc;
  }
  if (b->h_slist.empty()) return ELASTIC_OK;
  if (b->use_graph) {
    bool grids = b->vgrid != b->graph_vgrid;
    for (int v = 0; v < kVariants; ++v) grids |= (b->mask & (1u << v)) && b->sgrid[v] != b->graph_sgrid[v];
    if (!b->gexec || b->graph_lds < b->lds || b->graph_mask != b->mask || grids) {
      int rc = capture(b);
      if (rc) return rc;
    }
    BATCH_HIP(b, hipGraphLaunch(b->gexec, b->stream));
  } else {
    int rc = set_lds_limit(b);
    if (rc) return rc;
    for (int it = 0; it < b->iters; ++it) enqueue_iteration(b, it);
  }
  hipLaunchKernelGGL(k_bt_status, dim3(1), dim3(256), 0, b->stream, b->d_insts, b->d_slist,
                     b->d_counts + kVariants + 1, b->d_status);
  BATCH_HIP(b, hipMemcpyAsync(b->h_status, b->d_status, sizeof(BStatus) * b->h_slist.size(),
                              hipMemcpyDeviceToHost, b->stream));
  b->stepped = true;
  return ELASTIC_OK;
}

int elastic_batch_poll(elastic_batch* b, int64_t* tags, elastic_stats* stats, int32_t cap,
                       int32_t* ndone) {
  if (!b || !ndone || (cap > 0 && !tags)) return ELASTIC_BAD_ARG;
  *ndone = 0;
  hipSetDevice(b->device);
  BATCH_HIP(b, hipStreamSynchronize(b->stream));
  BATCH_HIP(b, hipGetLastError());
  if (!b->stepped) return ELASTIC_OK;
  int nd = 0;
  for (size_t w = 0; w < b->h_slist.size() && nd < cap; ++w) {
    const BStatus& st = b->h_status[w];
    if (!st.done) continue;
    const int k = b->h_slist[w];
    auto& S = b->slots[k];
    if (!S.used) continue;
    tags[nd] = S.tag;
    // an instance stopped by maxit still holds FP32 increments of x (a
    // converged one folded them at its last replacement)
    if (st.status != 0) {
      fold_increments(b->h_insts[k], S.ctx->c.stream);
      BATCH_HIP(b, hipStreamSynchronize(S.ctx->c.stream));
    }
    if (stats) {
      elastic_stats& o = stats[nd];
      std::memset(&o, 0, sizeof(o));
      o.iters = st.iter;
      o.outer = st.nupd;
      o.status = st.status;
      o.spmv_count = st.iter + st.nupd;
      o.op64_from = st.op64_from;
      o.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                              S.t0).count();
      // post-solve figures on the instance's own context (FP64 operator)
      Context* c = &S.ctx->c;
      hipMemcpy(c->scal_host, c->scal, sizeof(SolveScalars), hipMemcpyDeviceToHost);
      o.conv = c->scal_host->conv;
      int rc = post_solve_stats(c, &o);
      if (rc) return berr(b, rc, c->err);
    }
    {  // x back into the CSRC order (compact layouts)
      Context* c = &S.ctx->c;
      int rc = leave_layout_order(c);
      if (rc) return berr(b, rc, c->err);
      BATCH_HIP(b, hipStreamSynchronize(c->stream));
    }
    // the slot holds no instance any more: kernels skip it even if a list
    // still named it (the lists are rebuilt before the next step)
    // (synchronous: the batch stream is idle here, and a later submit into
    // this slot rewrites the host copy)
    b->h_insts[k].live = 0;
    BATCH_HIP(b, hipMemcpy(&b->d_insts[k].live, &b->h_insts[k].live, sizeof(int),
                           hipMemcpyHostToDevice));
    S.used = false;
    S.ctx = nullptr;
    ++nd;
  }
  if (nd) b->dirty = true;
  b->stepped = false;  // the statuses of this step are consumed
  *ndone = nd;
  return ELASTIC_OK;
}

}  // extern "C"
