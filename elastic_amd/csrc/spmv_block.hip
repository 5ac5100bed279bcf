// Node-blocked symmetric product: the single-system kernel, the value pack
// after assembly and the launch dispatch (the tile routine is in
// spmv_block.h).
#include "spmv_block.h"

namespace elastic {

// DOT: also the tiles' shares of x.(A x) into tdot; NT lanes per tile; XD:
// the product of x + dx (FP32 increments dx).  FP32 values and sums: the
// register budget is held to eight waves per SIMD, the occupancy the tile's
// LDS allows.
template <typename TV, typename T, int NB, int PER, bool DOT, int NT, bool XD = false>
__global__ __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(sizeof(TV) == 4 && sizeof(T) == 4 ? 8 : 1)))
void k_spmv_blocks(BlkArgs a, const TV* __restrict__ vals, const TV* __restrict__ diag,
                   const double* __restrict__ x, double* __restrict__ y,
                   double* __restrict__ epart, const int32_t* __restrict__ skip,
                   double* __restrict__ tdot, int y32, const float* __restrict__ x32,
                   const float* __restrict__ dx) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (skip != nullptr && *skip != 0) return;
  block_tile<TV, T, NB, PER, DOT, NT, XD, true>(a, blockIdx.x, vals, diag, x, y, epart, smem,
                                                tdot, y32, x32, dx);
}

// *out = the n shares added in a fixed order (one workgroup of 1024 lanes):
// lane l sums shares l, l + 1024, ... (sixteen loads in flight at a time, so
// up to 16,384 shares take one round trip; added in index order); then a
// fixed tree over the lanes.
__global__ __launch_bounds__(1024) void k_sum_shares(const double* __restrict__ part, int n,
                                                      double* out,
                                                      const int32_t* __restrict__ skip) {
  __shared__ double red[16];
  if (skip != nullptr && *skip != 0) return;
  double v = 0.0;
  for (int k0 = threadIdx.x; k0 < n; k0 += 16 * 1024) {
    double w[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int k = k0 + 1024 * j;
      w[j] = k < n ? part[k] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) v += w[j];
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < 16; ++w) s += red[w];
    *out = s;
  }
}

// Gathers the CSRC al/ad into the block planes (one workgroup per tile) and
// the node diagonal blocks (grid over nodes).  Block q takes the values of
// entry k of CSRC node row R's lower blocks (src_row / src_k), transposed
// when bit 31 of src_k is set (compact node order); row r of that CSRC block
// starts at ia[NB R + r] - 1 + NB k.
template <typename TV, int NB>
__global__ __launch_bounds__(256) void k_pack_blocks(BlkArgs a, const int32_t* __restrict__ ia,
                                                      const int32_t* __restrict__ src_row,
                                                      const int32_t* __restrict__ src_k,
                                                      const double* __restrict__ al,
                                                      TV* __restrict__ vals) {
  const int t = blockIdx.x;
  const int nbk = a.nblk[t];
  const int64_t base = a.blk0[t];
  TV* v = vals + base * (NB * NB);
  for (int q = threadIdx.x; q < nbk; q += 256) {
    const int R = src_row[base + q];
    const int kk = src_k[base + q];
    const int k = kk & 0x7FFFFFFF;
    const bool tr = kk < 0;
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int64_t row_start = (int64_t)ia[NB * R + r] - 1 + (int64_t)NB * k;
#pragma unroll
      for (int cc = 0; cc < NB; ++cc) {
        // CSRC entry (r, cc) is layout entry (r, cc), or (cc, r) when transposed
        const int p = tr ? cc * NB + r : r * NB + cc;
        v[Planes<TV, NB>::index(nbk, q, p)] = (TV)al[row_start + cc];
      }
    }
  }
}

// d2r: layout node -> CSRC node (null: the same order)
template <typename TV, int NB>
__global__ __launch_bounds__(256) void k_pack_diag(int nn, const int32_t* __restrict__ ia,
                                                    const int32_t* __restrict__ d2r,
                                                    const double* __restrict__ ad,
                                                    const double* __restrict__ al,
                                                    TV* __restrict__ diag) {
  for (int d = blockIdx.x * 256 + threadIdx.x; d < nn; d += gridDim.x * 256) {
    const int I = d2r ? d2r[d] : d;
    const int nbr = (ia[NB * I + 1] - ia[NB * I]) / NB;  // lower blocks of the node row
#pragma unroll
    for (int r = 0; r < NB; ++r) diag[(int64_t)r * nn + d] = (TV)ad[NB * I + r];
    const int64_t p1 = (int64_t)ia[NB * I + 1] - 1 + (int64_t)NB * nbr;  // row 1 own entries
    diag[(int64_t)NB * nn + d] = (TV)al[p1];                               // a21
    if constexpr (NB == 3) {
      const int64_t p2 = (int64_t)ia[NB * I + 2] - 1 + (int64_t)NB * nbr;
      diag[(int64_t)4 * nn + d] = (TV)al[p2];      // a31
      diag[(int64_t)5 * nn + d] = (TV)al[p2 + 1];  // a32
    }
  }
}

// Jacobi scale in the layout order: m[NB d + r] = 1 / ad[NB d2r[d] + r]
__global__ __launch_bounds__(256) void k_recip_perm(int nn, int nb, const int32_t* __restrict__ d2r,
                                                     const double* __restrict__ ad,
                                                     double* __restrict__ m) {
  const int64_t n = (int64_t)nn * nb;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int d = (int)(i / nb), r = (int)(i - (int64_t)d * nb);
    m[i] = 1.0 / ad[(int64_t)nb * d2r[d] + r];
  }
}

template <typename TV, typename TA, int NB>
static void spmv_blocks_launch(const DevBlocks& b, const TV* vals, const TV* diag,
                               const double* x, double* y, double* epart, hipStream_t s,
                               const int32_t* skip, const DotOut* dot, int y32,
                               const float* x32, int lanes, const float* dx = nullptr) {
  const size_t lds = tile_smem(sizeof(TA), NB, b.ncols_max, b.nblk_max);
  const int nt = lanes == 512 && b.nblk_max > 256 ? 512 : 256;
  const int per = (b.nblk_max + nt - 1) / nt;
#define ELASTIC_BLK_LAUNCH(K, P, D, T, G, ...)                                                  \
  do {                                                                                          \
    static bool raised = false;                                                                 \
    if (lds > 65536 && !raised) {                                                               \
      (void)hipFuncSetAttribute((const void*)K<TV, TA, NB, P, D, T>,                            \
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);        \
      raised = true;                                                                            \
    }                                                                                           \
    hipLaunchKernelGGL((K<TV, TA, NB, P, D, T>), dim3(G), dim3(T), lds, s, __VA_ARGS__);        \
  } while (0)
  double* td = dot ? dot->part : nullptr;
#define ELASTIC_BLK_PER(D)                                                                               \
  do {                                                                                                   \
    if (nt == 512) ELASTIC_BLK_LAUNCH(k_spmv_blocks, 2, D, 512, b.ntiles, blk_args(b), vals, diag, x, y, epart, skip, td, y32, x32, dx); \
    else if (per <= 1) ELASTIC_BLK_LAUNCH(k_spmv_blocks, 1, D, 256, b.ntiles, blk_args(b), vals, diag, x, y, epart, skip, td, y32, x32, dx); \
    else if (per <= 2) ELASTIC_BLK_LAUNCH(k_spmv_blocks, 2, D, 256, b.ntiles, blk_args(b), vals, diag, x, y, epart, skip, td, y32, x32, dx); \
    else ELASTIC_BLK_LAUNCH(k_spmv_blocks, 4, D, 256, b.ntiles, blk_args(b), vals, diag, x, y, epart, skip, td, y32, x32, dx); \
  } while (0)
  if constexpr (sizeof(TV) == 8 && sizeof(TA) == 8) {
    if (dx != nullptr && dot == nullptr) {  // the product of x + dx (residual replacement)
      const int per256 = (b.nblk_max + 255) / 256;  // 256 lanes whatever the lanes option
      const auto kx = per256 <= 1 ? k_spmv_blocks<TV, TA, NB, 1, false, 256, true>
                    : per256 <= 2 ? k_spmv_blocks<TV, TA, NB, 2, false, 256, true>
                                  : k_spmv_blocks<TV, TA, NB, 4, false, 256, true>;
      if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)kx, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
      hipLaunchKernelGGL(kx, dim3(b.ntiles), dim3(256), lds, s, blk_args(b), vals, diag, x, y,
                         epart, skip, (double*)nullptr, 0, (const float*)nullptr, dx);
      return;
    }
  }
  if (dot) {
    ELASTIC_BLK_PER(true);
    if (dot->result)  // (null: the shares only)
      hipLaunchKernelGGL(k_sum_shares, dim3(1), dim3(1024), 0, s, (const double*)dot->part,
                         b.ntiles, dot->result, skip);
  } else {
    ELASTIC_BLK_PER(false);
  }
#undef ELASTIC_BLK_PER
#undef ELASTIC_BLK_LAUNCH
}

void launch_spmv(Context* c, bool fp32, const double* x, double* y, double* epart,
                 const int32_t* skip, const DotOut* dot, bool y32, const float* x32,
                 const float* dx) {
  hipStream_t s = c->stream;
  const DevBlocks& b = c->blk;
  const int l32 = c->lanes[1], l64 = c->lanes[0];
  if (c->layout_nb == 3) {
    if (fp32 && c->accum32) spmv_blocks_launch<float, float, 3>(b, b.v32, b.d32, x, y, epart, s, skip, dot, y32 ? 1 : 0, x32, l32);
    else if (fp32) spmv_blocks_launch<float, double, 3>(b, b.v32, b.d32, x, y, epart, s, skip, dot, 0, nullptr, l32);
    else spmv_blocks_launch<double, double, 3>(b, b.v64, b.d64, x, y, epart, s, skip, dot, 0, nullptr, l64, dx);
  } else if (c->layout_nb == 2) {
    if (fp32 && c->accum32) spmv_blocks_launch<float, float, 2>(b, b.v32, b.d32, x, y, epart, s, skip, dot, y32 ? 1 : 0, x32, l32);
    else if (fp32) spmv_blocks_launch<float, double, 2>(b, b.v32, b.d32, x, y, epart, s, skip, dot, 0, nullptr, l32);
    else spmv_blocks_launch<double, double, 2>(b, b.v64, b.d64, x, y, epart, s, skip, dot, 0, nullptr, l64, dx);
  } else if (c->gather) {
    const DevGather& g = c->gth;
    if (fp32)
      launch_spmv_gather<float>(c->neq, g.ia, g.ja, c->A32.ad, c->A32.al, g.colptr, g.colk,
                                g.colrow, x, y, s, skip);
    else
      launch_spmv_gather<double>(c->neq, g.ia, g.ja, c->A64.ad, c->A64.al, g.colptr, g.colk,
                                 g.colrow, x, y, s, skip);
  } else {
    if (fp32)
      launch_spmv_tiles<float, double>(c->tiles, c->A32.ad, c->A32.al, x, y, epart, s, skip);
    else
      launch_spmv_tiles<double, double>(c->tiles, c->A64.ad, c->A64.al, x, y, epart, s, skip);
  }
}

void launch_block_pack(Context* c) {
  if (c->layout_nb == 1) return;
  hipStream_t s = c->stream;
  const DevBlocks& b = c->blk;
  const int g = std::max(1, std::min((b.nn + 255) / 256, 8192));
#define ELASTIC_PACK(NB)                                                                       \
  do {                                                                                         \
    hipLaunchKernelGGL((k_pack_blocks<double, NB>), dim3(b.ntiles), dim3(256), 0, s,           \
                       blk_args(b), b.ia, b.src_row, b.src_k, c->A64.al, b.v64);               \
    hipLaunchKernelGGL((k_pack_blocks<float, NB>), dim3(b.ntiles), dim3(256), 0, s,            \
                       blk_args(b), b.ia, b.src_row, b.src_k, c->A64.al, b.v32);               \
    hipLaunchKernelGGL((k_pack_diag<double, NB>), dim3(g), dim3(256), 0, s, b.nn, b.ia, b.d2r, \
                       c->A64.ad, c->A64.al, b.d64);                                           \
    hipLaunchKernelGGL((k_pack_diag<float, NB>), dim3(g), dim3(256), 0, s, b.nn, b.ia, b.d2r,  \
                       c->A64.ad, c->A64.al, b.d32);                                           \
  } while (0)
  if (b.nb == 3) ELASTIC_PACK(3);
  else ELASTIC_PACK(2);
#undef ELASTIC_PACK
}

void launch_recip_layout(Context* c) {
  const DevBlocks& b = c->blk;
  const int64_t n = (int64_t)b.nn * b.nb;
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_recip_perm, dim3(g), dim3(256), 0, c->stream, b.nn, b.nb, b.d2r,
                     c->A64.ad, c->A64.m);
}

}  // namespace elastic
