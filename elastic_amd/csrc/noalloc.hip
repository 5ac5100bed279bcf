// The symmetric CSRC product on caller-held device arrays, with no device
// allocation: for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops).  It computes what elastic_spmv computes
// (Matvec.f:528-621) without a context: the transposed lists of the pattern
// that the gather-only product walks (spmv_gather.hip) are built on every call
// in `scratch`, a file-scope device array of fixed size, instead of a layout
// built once at set-up.
//
// Per call, on the caller's stream:
//   1. column counts of the pattern (integer atomics into cnt, cleared first
//      by a fill kernel; no hipMemsetAsync, so the call can be captured);
//   2. colptr = exclusive scan of the counts (three launches: block scans,
//      the scan of the block totals, the offsets);
//   3. every stored entry placed in its column's list (integer atomics on a
//      cursor copied from colptr), then each list sorted by CSRC position, so
//      the order, and the FP64 result, never depend on the atomics;
//   4. the gather product: y_i = ad_i x_i + row i in CSRC order + column i in
//      ascending rows, bit-identical to matvec_csrcsym1.
#include <mutex>

#include "elastic_internal.h"

namespace elastic {
namespace {

// 128 MiB of int32 words per GPU: cnt (neq), colptr (neq+1), colk (nad),
// colrow (nad), the block totals of the scan.
constexpr int64_t kScratchWords = int64_t(32) << 20;
constexpr int kScanBlock = 1024;  // entries per block of the scan (256 lanes x 4)
constexpr int kMaxDevices = 64;

static __device__ __align__(256) int32_t scratch[kScratchWords];

// device address of `scratch` on each device (hipGetSymbolAddress once per
// device, with that device current)
int32_t* g_scratch[kMaxDevices] = {nullptr};
std::once_flag g_once[kMaxDevices];
hipError_t g_status[kMaxDevices];

int64_t scratch_words(int64_t neq, int64_t nad) {
  const int64_t nblk = (neq + 1 + kScanBlock - 1) / kScanBlock;
  return 2 * neq + 1 + 2 * nad + nblk + 1;
}

__global__ __launch_bounds__(256) void k_fill(int64_t n, int32_t* __restrict__ p, int32_t v) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_count_cols(int64_t nad, int n,
                                                     const int32_t* __restrict__ ja,
                                                     int32_t* __restrict__ cnt) {
  for (int64_t k = blockIdx.x * 256ll + threadIdx.x; k < nad; k += (int64_t)gridDim.x * 256) {
    const int c = ja[k] - 1;
    if (c >= 0 && c < n) atomicAdd(&cnt[c], 1);  // a column outside the system is dropped
  }
}

// exclusive scan of in[0..n) into out[0..n), one kScanBlock chunk per
// workgroup; the chunk's total goes to tot[block]
__global__ __launch_bounds__(256) void k_scan_blocks(int n, const int32_t* __restrict__ in,
                                                      int32_t* __restrict__ out,
                                                      int32_t* __restrict__ tot) {
  __shared__ int32_t sh[256];
  const int base = blockIdx.x * kScanBlock + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = base + j < n ? in[base + j] : 0;
    s += v[j];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {  // inclusive scan of the lane totals
    const int o = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
    __syncthreads();
    sh[threadIdx.x] += o;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - s;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 255) tot[blockIdx.x] = sh[255];
}

// exclusive scan of the nb block totals in place (one workgroup, sequential
// chunks of 256); total of all at tot[nb]
__global__ __launch_bounds__(256) void k_scan_totals(int nb, int32_t* __restrict__ tot) {
  __shared__ int32_t sh[256];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 256) {
    const int i = b0 + threadIdx.x;
    const int v = i < nb ? tot[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int o = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
      __syncthreads();
      sh[threadIdx.x] += o;
      __syncthreads();
    }
    if (i < nb) tot[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 255) carry += sh[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) tot[nb] = carry;
}

// colptr[i] += offset of its block; colptr[n] = the total; cursor = colptr
__global__ __launch_bounds__(256) void k_scan_add(int n, int32_t* __restrict__ colptr,
                                                   const int32_t* __restrict__ tot, int nb,
                                                   int32_t* __restrict__ cursor) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const int v = colptr[i] + tot[i / kScanBlock];
    colptr[i] = v;
    cursor[i] = v;
  }
  if (i == n) colptr[n] = tot[nb];
}

// every stored entry into its column's list (any order inside a list)
__global__ __launch_bounds__(256) void k_place(int n, const int32_t* __restrict__ ia,
                                                const int32_t* __restrict__ ja,
                                                int32_t* __restrict__ cursor,
                                                int32_t* __restrict__ colk,
                                                int32_t* __restrict__ colrow) {
  for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256)
    for (int k = ia[r] - 1; k < ia[r + 1] - 1; ++k) {
      const int c = ja[k] - 1;
      if (c < 0 || c >= n) continue;  // as in the count
      const int p = atomicAdd(&cursor[c], 1);
      colk[p] = k;
      colrow[p] = r;
    }
}

// each column's list sorted by CSRC position (insertion sort: the lists of
// finite element patterns are short); positions ascend with rows
__global__ __launch_bounds__(256) void k_sort_cols(int n, const int32_t* __restrict__ colptr,
                                                    int32_t* __restrict__ colk,
                                                    int32_t* __restrict__ colrow) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
    const int a = colptr[j], b = colptr[j + 1];
    for (int q = a + 1; q < b; ++q) {
      const int k = colk[q], r = colrow[q];
      int p = q - 1;
      while (p >= a && colk[p] > k) {
        colk[p + 1] = colk[p];
        colrow[p + 1] = colrow[p];
        --p;
      }
      colk[p + 1] = k;
      colrow[p + 1] = r;
    }
  }
}

int grid_of(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

}  // namespace
}  // namespace elastic

extern "C" int elastic_spmv_csrc_noalloc(int32_t neq, int64_t nad, const int32_t* ia,
                                         const int32_t* ja, const double* ad, const double* al,
                                         const double* x, double* y, void* stream) {
  using namespace elastic;
  if (neq <= 0 || nad < 0 || !ia || (nad > 0 && (!ja || !al)) || !ad || !x || !y)
    return ELASTIC_BAD_ARG;
  if (nad > INT32_MAX - 1 || scratch_words(neq, nad) > kScratchWords) {
    fprintf(stderr,
            "elastic_spmv_csrc_noalloc: neq %d, nad %lld need %lld scratch words; the fixed "
            "scratch holds %lld (2 neq + 2 nad + neq/1024 + 2 <= %lld)\n",
            neq, (long long)nad, (long long)scratch_words(neq, nad), (long long)kScratchWords,
            (long long)kScratchWords);
    return ELASTIC_BAD_ARG;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return ELASTIC_HIP_ERROR;
  std::call_once(g_once[dev], [dev]() {
    void* p = nullptr;
    g_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_status[dev] == hipSuccess) g_scratch[dev] = static_cast<int32_t*>(p);
  });
  if (g_status[dev] != hipSuccess || !g_scratch[dev]) {
    fprintf(stderr, "elastic_spmv_csrc_noalloc: hipGetSymbolAddress(scratch): %s\n",
            hipGetErrorString(g_status[dev]));
    return ELASTIC_HIP_ERROR;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int32_t* cnt = g_scratch[dev];  // later the placement cursor
  int32_t* colptr = cnt + neq;
  int32_t* colk = colptr + neq + 1;
  int32_t* colrow = colk + nad;
  int32_t* tot = colrow + nad;
  const int nb = (neq + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_fill, dim3(grid_of(neq)), dim3(256), 0, s, (int64_t)neq, cnt, 0);
  if (nad > 0)
    hipLaunchKernelGGL(k_count_cols, dim3(grid_of(nad)), dim3(256), 0, s, nad, (int)neq, ja, cnt);
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, s, neq, (const int32_t*)cnt, colptr,
                     tot);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(256), 0, s, nb, tot);
  hipLaunchKernelGGL(k_scan_add, dim3((neq + 1 + 255) / 256), dim3(256), 0, s, neq, colptr,
                     (const int32_t*)tot, nb, cnt);
  if (nad > 0) {
    hipLaunchKernelGGL(k_place, dim3(grid_of(neq)), dim3(256), 0, s, neq, ia, ja, cnt, colk,
                       colrow);
    hipLaunchKernelGGL(k_sort_cols, dim3(grid_of(neq)), dim3(256), 0, s, neq,
                       (const int32_t*)colptr, colk, colrow);
  }
  launch_spmv_gather<double>(neq, ia, ja, ad, al, colptr, colk, colrow, x, y, s, nullptr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "elastic_spmv_csrc_noalloc: %s\n", hipGetErrorString(e));
    return ELASTIC_HIP_ERROR;
  }
  return ELASTIC_OK;
}
