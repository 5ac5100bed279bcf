// Internal declarations of the MI355X (gfx950) implementation of the Elastic
// solve path.  Public entry points are in include/elastic_hip.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/elastic_hip.h"

namespace elastic {

// ---------------------------------------------------------------- host set-up
void node_graph(int nnode, int numel, int nen, const int32_t* ix, std::vector<int32_t>& xadj,
                std::vector<int32_t>& adj);
void rcm_order(int nnode, int numel, int nen, const int32_t* ix, int32_t* perm);
int number_equations(int32_t* id, const int32_t* perm, int nnode, int ndf);
void csrc_pattern(const int32_t* id, const int32_t* ix, int nnode, int numel, int nen, int ndf,
                  int neq, std::vector<int32_t>& ia, std::vector<int32_t>& ja);

// ------------------------------------------------------------- SpMV row tiles
// The symmetric CSRC product y = (D + L + L^T) x is computed per tile of
// consecutive rows.  A tile owns rows [row0, row0+nrows) and the slice
// [nz0, nz0+nnz) of the CSRC value stream (the reference's al, unchanged).
// Its "local columns" are its own rows (0..nrows-1) followed by the distinct
// columns below row0 that its rows reference (the "external" columns, sorted).
// Every nonzero carries one 32-bit word:
//   bits 24..31  local row            (nrows <= 256)
//   bits 12..23  local column          (ncols <= 4096)
//   bits  0..11  position of the nonzero in the tile's column-sorted order
// so the transposed products land in LDS already grouped by column and are
// summed in a fixed order: no atomics, bit-reproducible.  The sums for
// external columns are written to a partials array and added to their rows by
// the merge kernel through a per-row list (deterministic order).
constexpr int kTileRowsMax = 256;
constexpr int kTileColsMax = 4096;

struct TileLayout {
  int neq = 0;
  int64_t nad = 0;
  int ntiles = 0;
  int nnz_max = 0;    // largest tile nnz (LDS sizing)
  int ncols_max = 0;  // largest tile local-column count
  int64_t next_total = 0;
  // per tile
  std::vector<int32_t> row0, nrows, nnz, ncols, ext0;
  std::vector<int64_t> nz0;
  std::vector<int32_t> rp0, seg0;  // offsets into rowptr / segptr
  // streams
  std::vector<uint32_t> word;      // nad packed index words
  std::vector<uint16_t> rowptr;    // per tile nrows+1 local offsets
  std::vector<uint16_t> segptr;    // per tile ncols+1 local offsets (column-sorted order)
  std::vector<int32_t> ext;        // external columns (0-based global rows)
  // merge lists: row i receives partials inc_idx[inc_ptr[i] .. inc_ptr[i+1])
  std::vector<int32_t> inc_ptr, inc_idx;
};

// Builds the tiles from the 1-based CSRC pattern; tile_nnz bounds the nonzeros
// per tile (<= 3840 so that nrows + ncols fit the index fields).
void build_tiles(int neq, const int32_t* ia1, const int32_t* ja1, int tile_nnz, TileLayout& L);

// -------------------------------------------------------------- device state
template <typename T>
struct DevMatrix {
  T* ad = nullptr;  // diagonal
  T* al = nullptr;  // strictly lower values, CSRC order
  T* m = nullptr;   // Jacobi 1/ad
  T* bm = nullptr;  // 3x3 block-Jacobi inverse (6 per block), optional
};

struct DevTiles {
  int ntiles = 0, nnz_max = 0, ncols_max = 0;
  int32_t *row0 = nullptr, *nrows = nullptr, *nnz = nullptr, *ncols = nullptr, *ext0 = nullptr;
  int64_t* nz0 = nullptr;
  int32_t *rp0 = nullptr, *seg0 = nullptr;
  uint32_t* word = nullptr;
  uint16_t *rowptr = nullptr, *segptr = nullptr;
  int32_t* ext = nullptr;
  int32_t *inc_ptr = nullptr, *inc_idx = nullptr;
  int64_t next_total = 0;
};

// Scalars of one Krylov solve, resident in device memory.
struct SolveScalars {
  double d, di, alpha, beta, conv, pq, rz0, norm_b;
  int32_t iter, done, maxit, status;
};

constexpr int kRedBlocks = 1024;  // fixed reduction grid (deterministic sums)

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int neq = 0;
  int64_t nad = 0;
  int tile_nnz = 2048;
  TileLayout layout;
  DevTiles tiles;
  DevMatrix<double> A64;
  DevMatrix<float> A32;
  bool values_set = false, values32 = false, bdiag_set = false;
  // workspace
  double *epart64 = nullptr, *red64 = nullptr;
  float* epart32 = nullptr;
  double *v64[8] = {nullptr};
  float* v32[6] = {nullptr};
  SolveScalars* scal = nullptr;
  SolveScalars* scal_host = nullptr;  // pinned
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  // assembly state (optional)
  void* mesh = nullptr;  // MeshDev (assembly.hip)
  // autotuning/measurement
  double last_spmv_ms = 0.0;
  int check_every = 16;
  int zero_diag_row = -1;  // first row with ad == 0 (1-based), -1 if none
};

int set_error(Context* c, int code, const std::string& msg);
struct MeshDev;
void free_mesh(MeshDev* m);
#define ELASTIC_HIP_CHECK(ctx, expr)                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_error((ctx), ELASTIC_HIP_ERROR,                                       \
                       std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)

// kernels (spmv.hip / vector.hip)
template <typename T>
void launch_spmv_tiles(const DevTiles& t, const T* ad, const T* al, const T* x, T* y, T* epart,
                       hipStream_t s, const int32_t* skip = nullptr);
template <typename T>
void launch_merge(const DevTiles& t, int neq, const T* epart, T* y, hipStream_t s);
template <typename S, typename D>
void launch_cast(int64_t n, const S* a, D* b, hipStream_t s);
template <typename T>
void launch_recip(int n, const T* ad, T* m, hipStream_t s);
void launch_post(const DevTiles& t, int n, const double* epart, const double* x, double* y,
                 const double* b, const double* m, double* part, double* out4, hipStream_t s);
void launch_sum_parts(const double* part, int nsum, double* out, hipStream_t s);
void launch_ir_correct(int n, const float* e, double* x, hipStream_t s);
void launch_ir_residual(const DevTiles& t, int n, const double* epart, const double* y,
                        const double* b, const double* m, double* r, float* r32, double* part,
                        hipStream_t s);

// pcg.hip: buffers of one PCG solve in precision T (all device pointers)
template <typename T>
struct PcgBuffers {
  T *x, *r, *p, *q, *b, *m, *bm, *ad, *al, *epart;
};
template <typename T>
int pcg_device(Context* c, const PcgBuffers<T>& B, double tol, int maxit, double* dslots);

}  // namespace elastic

// the opaque handle of the C-ABI
struct elastic_ctx {
  elastic::Context c;
  std::vector<int32_t> ia, ja;  // host copy of the pattern (block preconditioner, drop-in)
};
