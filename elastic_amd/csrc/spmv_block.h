// Node-blocked symmetric product y = (D + L + L^T) x on gfx950 (same product
// as Matvec.f:528-621, on the block layout of layout_block.cpp): the tile
// routine shared by the single-system kernel (spmv_block.hip) and the
// multi-instance serving kernel (batch.hip).
//
// One workgroup per tile of node rows.  Every lane owns up to PER node blocks:
// it loads their nb*nb values (one coalesced plane at a time) and the block
// word, then forms A_IJ x_J for its row node and A_IJ^T x_I for its column
// node from the LDS-staged x.  Row pieces are summed across the wavefront by a
// segmented scan on the row key (DPP lane moves), transposed products land in
// LDS at their column-sorted position and each local column node is summed by
// one lane in a fixed order: deterministic, no atomics.  Values are TV
// (double, or float for the mixed-precision operator); staged x and sums are
// T.
#pragma once
#include <hip/hip_runtime.h>

#include "elastic_internal.h"

namespace elastic {


struct BlkArgs {
  const int32_t *row0, *nrows, *nblk, *ncols, *ext0, *seg0, *rp0;
  const int64_t* blk0;
  const uint32_t* word;
  const uint16_t *rowptr, *segptr;
  const int32_t* ext;
  int nn, nblk_max, ncols_max;
};

inline BlkArgs blk_args(const DevBlocks& b) {
  return BlkArgs{b.row0, b.nrows,  b.nblk,   b.ncols, b.ext0, b.seg0,    b.rp0,        b.blk0,
                 b.word, b.rowptr, b.segptr, b.ext,   b.nn,   b.nblk_max, b.ncols_max};
}

// Value planes of a tile: FP64 values are stored plane by plane (a lane's load
// is 8 bytes); FP32 values are stored with planes 2k and 2k+1 interleaved so a
// lane loads both with one 8-byte load (the last odd plane stays single).
template <typename TV, int NB>
struct Planes;
// A uniform pointer to global memory held in scalar registers: loads through
// it take the lane's 32-bit offset in one vector register, whichever plane
// they read.
template <typename P>
using global_ptr = const __attribute__((address_space(1))) P*;
template <typename P>
__device__ __forceinline__ global_ptr<P> scalar_ptr(const P* p) {
  const uint64_t u = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (global_ptr<P>)(((uint64_t)hi << 32) | lo);
}
template <int NB>
struct Planes<double, NB> {
  __device__ static inline void load(const double* v, int nbk, int q, double* out) {
#pragma unroll
    for (int p = 0; p < NB * NB; ++p) out[p] = __builtin_nontemporal_load(v + (int64_t)p * nbk + q);
  }
  // the same loads, v uniform and q >= 0: plane bases in scalar registers
  __device__ static inline void load_s(const double* v, int nbk, unsigned q, double* out) {
#pragma unroll
    for (int p = 0; p < NB * NB; ++p)
      out[p] = __builtin_nontemporal_load(scalar_ptr(v + (int64_t)p * nbk) + q);
  }
  __device__ __host__ static inline int64_t index(int nbk, int q, int p) { return (int64_t)p * nbk + q; }
};
template <int NB>
struct Planes<float, NB> {
  __device__ static inline void load(const float* v, int nbk, int q, float* out) {
    constexpr int NP = NB * NB / 2;
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2* v2 = reinterpret_cast<const f2*>(v);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const f2 f = __builtin_nontemporal_load(v2 + (int64_t)k * nbk + q);
      out[2 * k] = f.x;
      out[2 * k + 1] = f.y;
    }
    if (NB * NB % 2) out[NB * NB - 1] = __builtin_nontemporal_load(v + (int64_t)2 * NP * nbk + q);
  }
  __device__ static inline void load_s(const float* v, int nbk, unsigned q, float* out) {
    constexpr int NP = NB * NB / 2;
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2* v2 = reinterpret_cast<const f2*>(v);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const f2 f = __builtin_nontemporal_load(scalar_ptr(v2 + (int64_t)k * nbk) + q);
      out[2 * k] = f.x;
      out[2 * k + 1] = f.y;
    }
    if (NB * NB % 2)
      out[NB * NB - 1] = __builtin_nontemporal_load(scalar_ptr(v + (int64_t)2 * NP * nbk) + q);
  }
  __device__ __host__ static inline int64_t index(int nbk, int q, int p) {
    constexpr int NP = NB * NB / 2;
    if (p < 2 * NP) return (int64_t)2 * ((int64_t)(p / 2) * nbk + q) + (p & 1);
    return (int64_t)2 * NP * nbk + q;
  }
};

// Segmented inclusive scan of y over the wavefront, segments = runs of equal
// keys, with the keys ascending along the lanes (a tile's blocks are sorted by
// row node).  DPP moves on the VALU instead of LDS permutes: shifts by 1, 2,
// 4 and 8 lanes inside each row of 16 lanes, then the last lane of rows 0 and
// 2 added to rows 1 and 3, then lane 31 added to rows 2 and 3.  With ascending
// keys a lane whose key equals the source lane's key has every lane between
// them in its run, so each step adds a run's partial sum exactly once.
template <int CTRL, int RMASK>
__device__ __forceinline__ int dpp_move(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, RMASK, 0xF, false);
}
template <int CTRL, int RMASK>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(dpp_move<CTRL, RMASK>(0, __float_as_int(v)));
}
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_move<CTRL, RMASK>(0, (int)(b & 0xFFFFFFFFll));
  const int hi = dpp_move<CTRL, RMASK>(0, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// one step: lanes without a source (outside the row, or rows the mask leaves
// out) read key -1, which matches no key
template <int CTRL, int RMASK, typename T, int NB>
__device__ __forceinline__ void seg_step(int key, T (&y)[NB]) {
  const int ok = dpp_move<CTRL, RMASK>(-1, key);
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    const T o = dpp_move<CTRL, RMASK>(y[r]);
    if (ok == key) y[r] += o;
  }
}
template <typename T, int NB>
__device__ __forceinline__ void seg_scan(int key, T (&y)[NB]) {
  seg_step<0x111, 0xF>(key, y);  // row_shr:1
  seg_step<0x112, 0xF>(key, y);  // row_shr:2
  seg_step<0x114, 0xF>(key, y);  // row_shr:4
  seg_step<0x118, 0xF>(key, y);  // row_shr:8
  seg_step<0x142, 0xA>(key, y);  // row_bcast:15 into rows 1 and 3
  seg_step<0x143, 0xC>(key, y);  // row_bcast:31 into rows 2 and 3
}

// Staged x entries a lane gathers into registers (entries tid + NT k, k <
// kXR); a tile's entries beyond kXR * NT are gathered by a loop of their own.
constexpr int kXR = 3;

// Registers of one tile: the words and values of the lane's PER blocks (valid
// for blocks below the tile's block count) and, for lanes below the tile's
// row count, the diagonal block of that row node; gx (loaded with GX only):
// the node of each x entry the lane stages, where that entry is an external
// column.
template <typename TV, int NB, int PER>
struct TileRegs {
  uint32_t wr[PER];
  TV vr[PER][NB * NB];
  TV dg[NB * (NB + 1) / 2];
  int32_t gx[kXR];
};

// Issues the loads of tile t (values nontemporal: read once per product).
// NT: lanes of the workgroup (256 or 512); lane tid owns blocks tid + NT i.
// GX: the nodes of the x entries the lane stages are loaded first: they
// return while the value loads are still in flight, so compute_tile can issue
// the x gathers behind the values without waiting for them.  Every load is
// then issued by every lane, from a clamped index where the lane has nothing
// to load (compute_tile does not use such a result), so the number of loads
// in flight is the same on every path and each use waits only for the loads
// it needs.  (The single-system kernel uses GX; its workgroups serve one
// tile each.)
template <typename TV, int NB, int PER, int NT = 256, bool GX = false>
__device__ __forceinline__ void load_tile(const BlkArgs& a, int t, const TV* __restrict__ vals,
                                          const TV* __restrict__ diag, TileRegs<TV, NB, PER>& R) {
  constexpr int NB2 = NB * NB;
  const int tid = threadIdx.x;
  const int nbk = a.nblk[t], nr = a.nrows[t];
  const int64_t base = a.blk0[t];
  const TV* v = vals + base * NB2;
  if constexpr (GX) {
    const int nx = a.ncols[t] * NB;
    const int32_t* ext = a.ext + a.ext0[t];
#pragma unroll
    for (int k = 0; k < kXR; ++k) {
      // lanes past the end take the last entry again (in bounds, not stored)
      const int node = min(tid + NT * k, nx - 1) / NB;
      const int e = node - nr;  // external column, or negative for an own row
      // own rows (node I0 + node, formed where it is used) load a word that
      // is not used, so that nothing waits for these loads here
      R.gx[k] = *(e >= 0 ? ext + e : a.row0 + t);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      // blocks q >= nbk: not used; a tile without blocks reads the padding
      // block that follows the last tile's
      const unsigned qc = min((unsigned)(tid + NT * i), (unsigned)(max(nbk, 1) - 1));
      R.wr[i] = __builtin_nontemporal_load(a.word + base + qc);
      Planes<TV, NB>::load_s(v, max(nbk, 1), qc, R.vr[i]);
    }
    const unsigned g = a.row0[t] + min(tid, nr - 1);  // used for tid < nr only
#pragma unroll
    for (int k = 0; k < NB * (NB + 1) / 2; ++k) R.dg[k] = scalar_ptr(diag + (int64_t)k * a.nn)[g];
  } else {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int q = tid + NT * i;
      if (q < nbk) {
        R.wr[i] = __builtin_nontemporal_load(a.word + base + q);
        Planes<TV, NB>::load(v, nbk, q, R.vr[i]);
      } else {
        R.wr[i] = 0xFFFFFFFFu;
#pragma unroll
        for (int p = 0; p < NB2; ++p) R.vr[i][p] = TV(0);
      }
    }
    if (tid < nr) {
      const int g = a.row0[t] + tid;
#pragma unroll
      for (int k = 0; k < NB * (NB + 1) / 2; ++k) R.dg[k] = diag[(int64_t)k * a.nn + g];
    }
  }
}

// Shared memory of one tile's product with sums of type T (bytes): staged x,
// row pieces and transposed products (NB values of T per local column, per
// two pieces of each row, per block), then the tile's row and segment
// pointers (int32, one T-sized slot each, so the FP32 size is half the FP64
// size).
inline size_t tile_smem(size_t t_size, int nb, int ncols_max, int nblk_max) {
  return t_size * ((size_t)nb * ((size_t)ncols_max + 2 * kBlkRowsMax + (size_t)nblk_max) +
                   (kBlkRowsMax + 1) + ((size_t)ncols_max + 1));
}

// The product of tile t from its registers (a workgroup of NT lanes); smem
// holds tile_smem(sizeof(T), NB, ncols_max, nblk_max) bytes.  T is
// the type of the staged x, the products and the sums (double, or float for
// the FP32 operator with FP32 accumulation); x, y and the external partials
// are FP64 in memory either way.  Starts with a barrier, so tiles can follow
// each other in one workgroup; FRESH (a workgroup that serves this tile only)
// leaves out the barriers that guard an earlier tile's LDS.
struct NoHook {
  __device__ void operator()() const {}
};

// Block sum of the lanes' shares of x.(A x) into part[t] (fixed order): each
// wavefront sums by DPP moves (the steps of seg_scan without keys: lanes
// without a source add zero), so lane 63 holds the wavefront's sum.  FRESH:
// the workgroup serves this tile only, so red holds no earlier tile's sums.
template <int NT = 256, bool FRESH = false>
__device__ __forceinline__ void tile_dot_store(double* part, int t, double v) {
  __shared__ double red[NT / 64];
  v += dpp_move<0x111, 0xF>(v);
  v += dpp_move<0x112, 0xF>(v);
  v += dpp_move<0x114, 0xF>(v);
  v += dpp_move<0x118, 0xF>(v);
  v += dpp_move<0x142, 0xA>(v);
  v += dpp_move<0x143, 0xC>(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if constexpr (!FRESH) __syncthreads();  // red of a previous tile is free
  if (lane == 63) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = (red[0] + red[1]) + (red[2] + red[3]);
    if constexpr (NT == 512) s += (red[4] + red[5]) + (red[6] + red[7]);
    part[t] = s;
  }
}

// DOT: also the tile's share of x.(A x) into tdot[t]: x.y over its own rows
// plus x_J.partial_J over its external columns, with x as staged in LDS (exact
// for T = double; rounded to FP32 like the operator's own input otherwise),
// products and sums in FP64.  The shares
// are added in tile order by a kernel of their own (k_sum_shares), so x.(A x)
// needs no further pass over y and no floating-point atomics.
template <typename TV, typename T, int NB, int PER, bool DOT = false, typename Hook = NoHook,
          int NT = 256, bool XD = false, bool GX = false, bool FRESH = GX>
__device__ __forceinline__ void compute_tile(const BlkArgs& a, int t,
                                             const TileRegs<TV, NB, PER>& R,
                                             const double* __restrict__ x, double* __restrict__ y,
                                             double* __restrict__ epart, unsigned char* smem,
                                             double* __restrict__ tdot = nullptr,
                                             int y32 = 0, const float* __restrict__ x32 = nullptr,
                                             Hook hook = Hook(),
                                             const float* __restrict__ dx = nullptr) {
  constexpr int NB2 = NB * NB;
  T* xs = reinterpret_cast<T*>(smem);       // ncols_max * NB
  T* pp = xs + a.ncols_max * NB;            // 2 * kBlkRowsMax * NB (row pieces)
  T* tr = pp + 2 * kBlkRowsMax * NB;        // nblk_max * NB (transposed products)
  // the tile's row pointers (nr + 1) and column segment pointers (ncols + 1)
  int32_t* rps = reinterpret_cast<int32_t*>(tr + a.nblk_max * NB);
  int32_t* sps = rps + kBlkRowsMax + 1;
  const int tid = threadIdx.x, lane = tid & 63;
  const int I0 = a.row0[t], nr = a.nrows[t], nbk = a.nblk[t], ncols = a.ncols[t];
  const int ext0 = a.ext0[t];
  const int32_t* ext = a.ext + ext0;
  const uint16_t* rp = a.rowptr + a.rp0[t];
  const uint16_t* sp = a.segptr + a.seg0[t];
  const uint32_t* wr = R.wr;
  const auto& vr = R.vr;
  // x of the tile's local columns: the first entries are loaded into
  // registers, then `hook` runs (a caller may issue more loads there, which
  // stay in flight behind these), then they go to LDS; x32: an FP32 copy of x
  // (FP32 staging takes the same values from it).  XD (FP64 staging only): x
  // is held as x + dx, an FP64 base plus FP32 increments, and the product is
  // formed on their FP64 sum (a separate instance, so the other products keep
  // their registers)
  const int nx = ncols * NB;
  int xdone;  // entries staged from registers
  if constexpr (GX) {
    // kXR * NT entries, from the nodes load_tile fetched; all of a lane's
    // loads are issued before any is used
    auto stage = [&](auto src) {
      using S = typename std::remove_const<typename std::remove_pointer<decltype(src)>::type>::type;
      S xv[kXR];
      float dxv[kXR];
#pragma unroll
      for (int k = 0; k < kXR; ++k) {
        const int c = min(tid + NT * k, nx - 1);
        const int node = c / NB;
        const int g = node < nr ? I0 + node : R.gx[k];
        const int64_t i = (int64_t)NB * g + (c - node * NB);
        xv[k] = src[i];
        if constexpr (XD) dxv[k] = dx[i];
      }
      hook();
      if constexpr (!FRESH) __syncthreads();  // the LDS of a previous tile is free
#pragma unroll
      for (int k = 0; k < kXR; ++k) {
        const int c = tid + NT * k;
        if (c < nx) {
          double v = (double)xv[k];
          if constexpr (XD) v += (double)dxv[k];
          xs[c] = (T)v;
        }
      }
    };
    if (x32) stage(x32);
    else stage(x);
    xdone = kXR * NT;
  } else {
    constexpr int XR = 8;
    double xv[XR];
#pragma unroll
    for (int k = 0; k < XR; ++k) {
      const int c = tid + NT * k;
      if (c < nx) {
        const int node = c / NB, comp = c - node * NB;
        const int g = node < nr ? I0 + node : ext[node - nr];
        xv[k] = x32 ? (double)x32[(int64_t)NB * g + comp] : x[(int64_t)NB * g + comp];
        if constexpr (XD) xv[k] += (double)dx[(int64_t)NB * g + comp];
      }
    }
    hook();
    if constexpr (!FRESH) __syncthreads();  // the LDS of a previous tile is free
#pragma unroll
    for (int k = 0; k < XR; ++k) {
      const int c = tid + NT * k;
      if (c < nx) xs[c] = (T)xv[k];
    }
    xdone = XR * NT;
  }
  for (int c = tid + xdone; c < nx; c += NT) {
    const int node = c / NB, comp = c - node * NB;
    const int g = node < nr ? I0 + node : ext[node - nr];
    double v = x32 ? (double)x32[(int64_t)NB * g + comp] : x[(int64_t)NB * g + comp];
    if constexpr (XD) v += (double)dx[(int64_t)NB * g + comp];
    xs[c] = (T)v;
  }
  // the tile's row and segment pointers, kept in LDS for the scans and the
  // column sums (one wait here instead of one at each use)
  {
    // both loads issued before either is used (clamped indices: in bounds)
    const int rv = rp[min(tid, nr)], sv = sp[min(tid, ncols)];
    if (tid <= nr) rps[tid] = rv;  // nr <= kBlkRowsMax < NT
    if (tid <= ncols) sps[tid] = sv;
    for (int c = tid + NT; c <= ncols; c += NT) sps[c] = sp[c];
  }
  __syncthreads();
  // the products of all the lane's blocks first (their values are then free),
  // then the row scans, which are independent of each other
  T yl[PER][NB];
  int key[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const bool valid = tid + NT * i < nbk;
    const uint32_t w = wr[i];
    key[i] = valid ? (int)(w >> 25) : 255;
    const int lc = (w >> 10) & 0x7FFF, tp = w & 0x3FF;
#pragma unroll
    for (int r = 0; r < NB; ++r) yl[i][r] = T(0);
    if (valid) {
      T xc[NB], xr[NB];
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        xc[r] = xs[lc * NB + r];
        xr[r] = xs[key[i] * NB + r];
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        T s = T(0);
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) s += (T)vr[i][r * NB + cc] * xc[cc];
        yl[i][r] = s;
      }
#pragma unroll
      for (int cc = 0; cc < NB; ++cc) {
        T s = T(0);
#pragma unroll
        for (int r = 0; r < NB; ++r) s += (T)vr[i][r * NB + cc] * xr[r];
        tr[tp * NB + cc] = s;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) seg_scan(key[i], yl[i]);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int q = tid + NT * i;
    const bool valid = q < nbk;
    // a node row has at most 64 blocks (checked when the layout is built), so
    // it ends in its own chunk or in the next one: two pieces at most
    const bool row_end = valid && (q + 1 == rps[key[i] + 1]);
    if (valid && (lane == 63 || row_end)) {
      const int slot = 2 * key[i] + (row_end ? 1 : 0);
#pragma unroll
      for (int r = 0; r < NB; ++r) pp[slot * NB + r] = yl[i][r];
    }
  }
  __syncthreads();
  double dacc = 0.0;  // this lane's share of x.(Ax) (DOT)
  for (int c = tid; c < ncols; c += NT) {
    T s[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) s[r] = T(0);
    if (c < nr) {
      // node diagonal block (nb diagonal entries, then a21, a31, a32)
      // c < nr <= 127 < NT, so c == tid: the lane loaded this node's diagonal
      T d[NB * (NB + 1) / 2];
#pragma unroll
      for (int k = 0; k < NB * (NB + 1) / 2; ++k) d[k] = (T)R.dg[k];
      const T* xi = xs + c * NB;
      if constexpr (NB == 3) {
        s[0] = d[0] * xi[0] + d[3] * xi[1] + d[4] * xi[2];
        s[1] = d[3] * xi[0] + d[1] * xi[1] + d[5] * xi[2];
        s[2] = d[4] * xi[0] + d[5] * xi[1] + d[2] * xi[2];
      } else {
        s[0] = d[0] * xi[0] + d[2] * xi[1];
        s[1] = d[2] * xi[0] + d[1] * xi[1];
      }
      const int b0 = rps[c], b1 = rps[c + 1];
      if (b1 > b0) {
        if ((b0 | 63) < b1 - 1)  // the row crosses a chunk boundary
#pragma unroll
          for (int r = 0; r < NB; ++r) s[r] += pp[(2 * c) * NB + r];
#pragma unroll
        for (int r = 0; r < NB; ++r) s[r] += pp[(2 * c + 1) * NB + r];
      }
    }
    {
      // the column's transposed products in sorted order; four entries' LDS
      // reads are issued together, then added in order
      int q = sps[c];
      const int qe = sps[c + 1];
      for (; q + 4 <= qe; q += 4) {
        T v[4][NB];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < NB; ++r) v[j][r] = tr[(q + j) * NB + r];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < NB; ++r) s[r] += v[j][r];
      }
      for (; q < qe; ++q)
#pragma unroll
        for (int r = 0; r < NB; ++r) s[r] += tr[q * NB + r];
    }
    if (c < nr) {
      if (sizeof(T) == 4 && y32) {  // FP32 sums, y asked in FP32 (exact)
        float* yf = reinterpret_cast<float*>(y);
#pragma unroll
        for (int r = 0; r < NB; ++r) yf[(int64_t)NB * (I0 + c) + r] = (float)s[r];
      } else {
#pragma unroll
        for (int r = 0; r < NB; ++r) y[(int64_t)NB * (I0 + c) + r] = (double)s[r];
      }
    } else if constexpr (sizeof(T) == 4) {  // FP32 sums: FP32 partials (exact)
      float* ef = reinterpret_cast<float*>(epart);
#pragma unroll
      for (int r = 0; r < NB; ++r) ef[(int64_t)NB * (ext0 + c - nr) + r] = s[r];
    } else {
#pragma unroll
      for (int r = 0; r < NB; ++r) epart[(int64_t)NB * (ext0 + c - nr) + r] = (double)s[r];
    }
    if constexpr (DOT) {
#pragma unroll
      for (int r = 0; r < NB; ++r) dacc += (double)s[r] * (double)xs[c * NB + r];
    }
  }
  if constexpr (DOT) tile_dot_store<NT, FRESH>(tdot, t, dacc);
}

// One tile: its loads, then its product.
template <typename TV, typename T, int NB, int PER, bool DOT = false, int NT = 256,
          bool XD = false, bool GX = false>
__device__ __forceinline__ void block_tile(const BlkArgs& a, int t, const TV* __restrict__ vals,
                                           const TV* __restrict__ diag,
                                           const double* __restrict__ x, double* __restrict__ y,
                                           double* __restrict__ epart, unsigned char* smem,
                                           double* __restrict__ tdot = nullptr, int y32 = 0,
                                           const float* __restrict__ x32 = nullptr,
                                           const float* __restrict__ dx = nullptr) {
  TileRegs<TV, NB, PER> R;
  load_tile<TV, NB, PER, NT, GX>(a, t, vals, diag, R);
  compute_tile<TV, T, NB, PER, DOT, NoHook, NT, XD, GX>(a, t, R, x, y, epart, smem, tdot, y32,
                                                        x32, NoHook(), dx);
}

}  // namespace elastic
