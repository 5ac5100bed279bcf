// Symmetric CSRC product y = (D + L + L^T) x on gfx950 (Matvec.f:528-621).
//
// One workgroup per row tile (TileLayout).  The tile's x entries (own rows and
// external columns) are staged in LDS; every nonzero is read once from HBM
// (value + one packed index word, coalesced) and produces its lower product
// (into LDS in row order) and its transposed product (into LDS at its
// column-sorted position).  Each local column is then summed by one thread in
// a fixed order, so the result is deterministic and no atomics are used.
// External-column sums go to `epart` and are added to their rows by the merge
// kernel, which is fused into the solver's dot-product kernels.
#include "elastic_internal.h"

namespace elastic {

struct TileArgs {
  const int32_t *row0, *nrows, *nnz, *ncols, *ext0, *rp0, *seg0;
  const int64_t* nz0;
  const uint32_t* word;
  const uint16_t *rowptr, *segptr;
  const int32_t* ext;
  int nnz_max, ncols_max;
};

static TileArgs tile_args(const DevTiles& t) {
  return TileArgs{t.row0, t.nrows, t.nnz,  t.ncols,  t.ext0,   t.rp0, t.seg0,
                  t.nz0,  t.word,  t.rowptr, t.segptr, t.ext, t.nnz_max, t.ncols_max};
}

// PER = nonzeros per thread (tile nnz <= 256 * PER).  All of a tile's value and
// index loads are issued before any is used, so a workgroup keeps its whole
// tile in flight; the lower products of a row are summed with a segmented
// inclusive scan across the wavefront (rows are contiguous in the stream), and
// the lane ending a row piece stores it at the piece's last position.  A row's
// pieces end at every q = 64m - 1 inside the row and at its last nonzero, so
// the row sum reads at most ceil(len/64) + 1 values, in a fixed order.
// TV: stored value type (double, or float for the mixed-precision operator);
// T: vector and accumulation type.
template <typename TV, typename T, int PER>
__global__ __launch_bounds__(256) void k_spmv_tiles(TileArgs a, const TV* __restrict__ ad,
                                                     const TV* __restrict__ al,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     T* __restrict__ epart,
                                                     const int32_t* __restrict__ skip) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (skip != nullptr && *skip != 0) return;  // converged solve: uniform for the grid
  T* xs = reinterpret_cast<T*>(smem);
  T* pp = xs + a.ncols_max;
  T* tr = pp + a.nnz_max;
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  const int row0 = a.row0[t], nrows = a.nrows[t], nnz = a.nnz[t], ncols = a.ncols[t];
  const int ext0 = a.ext0[t];
  const int64_t nz0 = a.nz0[t];
  const int32_t* ext = a.ext + ext0;
  const uint32_t* w = a.word + nz0;
  const TV* v = al + nz0;
  uint32_t wr[PER];
  TV vr[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int q = tid + 256 * i;
    if (q < nnz) {
      wr[i] = __builtin_nontemporal_load(w + q);
      vr[i] = __builtin_nontemporal_load(v + q);
    } else {
      wr[i] = 0xFFFFFFFFu;
      vr[i] = TV(0);
    }
  }
  for (int c = tid; c < ncols; c += 256) xs[c] = (c < nrows) ? x[row0 + c] : x[ext[c - nrows]];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int q = tid + 256 * i;
    const bool valid = q < nnz;
    const uint32_t wq = wr[i];
    const int key = valid ? (int)(wq >> 24) : 256;
    const int lc = (wq >> 12) & 0xFFF, tp = wq & 0xFFF;
    T s = T(0);
    if (valid) {
      const T vv = (T)vr[i];
      s = vv * xs[lc];
      tr[tp] = vv * xs[key];
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const T o = __shfl_up(s, d, 64);
      const int ok = __shfl_up(key, d, 64);
      if (lane >= d && ok == key) s += o;
    }
    const int nkey = __shfl_down(key, 1, 64);
    if (valid && (lane == 63 || nkey != key)) pp[q] = s;
  }
  __syncthreads();
  const uint16_t* rp = a.rowptr + a.rp0[t];
  const uint16_t* sp = a.segptr + a.seg0[t];
  for (int c = tid; c < ncols; c += 256) {
    T s = T(0);
    if (c < nrows) {
      s = (T)ad[row0 + c] * xs[c];
      const int b0 = rp[c], b1 = rp[c + 1];
      if (b1 > b0) {
        for (int q = b0 | 63; q < b1 - 1; q += 64) s += pp[q];
        s += pp[b1 - 1];
      }
    }
    for (int q = sp[c]; q < sp[c + 1]; ++q) s += tr[q];
    if (c < nrows) y[row0 + c] = s;
    else epart[ext0 + c - nrows] = s;
  }
}

__global__ __launch_bounds__(256) void k_merge(int n, MergeArgs mg, double* __restrict__ y) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) y[i] = merge_row(mg, i, y[i]);
}

template <typename TV, typename T>
void launch_spmv_tiles(const DevTiles& t, const TV* ad, const TV* al, const T* x, T* y, T* epart,
                       hipStream_t s, const int32_t* skip) {
  const size_t lds = sizeof(T) * ((size_t)t.ncols_max + 2 * (size_t)t.nnz_max);
  const int per = (t.nnz_max + 255) / 256;
#define ELASTIC_SPMV_LAUNCH(P)                                                                 \
  do {                                                                                         \
    static bool raised = false;                                                                \
    if (lds > 65536 && !raised) {                                                              \
      hipFuncSetAttribute((const void*)k_spmv_tiles<TV, T, P>,                                     \
                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);             \
      raised = true;                                                                           \
    }                                                                                          \
    hipLaunchKernelGGL((k_spmv_tiles<TV, T, P>), dim3(t.ntiles), dim3(256), lds, s, tile_args(t),  \
                       ad, al, x, y, epart, skip);                                             \
  } while (0)
  if (per <= 2) ELASTIC_SPMV_LAUNCH(2);
  else if (per <= 4) ELASTIC_SPMV_LAUNCH(4);
  else if (per <= 8) ELASTIC_SPMV_LAUNCH(8);
  else if (per <= 12) ELASTIC_SPMV_LAUNCH(12);
  else ELASTIC_SPMV_LAUNCH(15);
#undef ELASTIC_SPMV_LAUNCH
}

void launch_merge(Context* c, const double* epart, double* y, bool ep32) {
  const int neq = c->neq;
  int blocks = std::max(1, std::min((neq + 255) / 256, 4096));
  hipLaunchKernelGGL(k_merge, dim3(blocks), dim3(256), 0, c->stream, neq, merge_of(c, epart, ep32), y);
}

template void launch_spmv_tiles<double, double>(const DevTiles&, const double*, const double*,
                                                const double*, double*, double*, hipStream_t,
                                                const int32_t*);
template void launch_spmv_tiles<float, double>(const DevTiles&, const float*, const float*,
                                               const double*, double*, double*, hipStream_t,
                                               const int32_t*);



}  // namespace elastic
