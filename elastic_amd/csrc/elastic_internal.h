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

// ---------------------------------------------------- node-blocked row tiles
// (layout_block.cpp; see the comment there)
constexpr int kBlkRowsMax = 127;
constexpr int kBlkTileMax = 1024;

struct BlockLayout {
  int nb = 1, nn = 0, ntiles = 0, nblk_max = 0, ncols_max = 0;
  int64_t nblocks = 0, next_nodes = 0;
  bool compact = false;  // nodes in the compact tile order (d2r), else the CSRC order
  std::vector<int32_t> row0, nrows, nblk, ncols, ext0, seg0, rp0;
  std::vector<int64_t> blk0;
  std::vector<uint32_t> word;
  std::vector<uint16_t> rowptr, segptr;
  std::vector<int32_t> ext, inc_ptr, inc_idx;
  // layout node d is CSRC node d2r[d] (compact order only)
  std::vector<int32_t> d2r;
  // source of block q in the CSRC values: the lower blocks of CSRC node row
  // src_row[q], entry src_k[q] & 0x7FFFFFFF of that row; bit 31 of src_k: the
  // layout block is the transpose of that CSRC block
  std::vector<int32_t> src_row, src_k;
};

// 3 (or 2) when the CSRC pattern is a pattern of node blocks, else 1
int detect_node_blocks(int neq, const int32_t* ia1, const int32_t* ja1);
// compact: order the nodes in compact tiles grown breadth-first over the node
// graph (a tile's lower neighbours outside it are then few) instead of the
// CSRC order; falls back to the CSRC order when a node would exceed 64 lower
// blocks in the compact order
void build_block_tiles(int neq, int nb, const int32_t* ia1, const int32_t* ja1, int tile_blocks,
                       BlockLayout& L, bool compact = false);

struct DevBlocks {
  int nb = 1, nn = 0, ntiles = 0, nblk_max = 0, ncols_max = 0;
  int64_t nblocks = 0, next_nodes = 0;
  int32_t *row0 = nullptr, *nrows = nullptr, *nblk = nullptr, *ncols = nullptr, *ext0 = nullptr,
          *seg0 = nullptr, *rp0 = nullptr;
  int64_t* blk0 = nullptr;
  uint32_t* word = nullptr;
  uint16_t *rowptr = nullptr, *segptr = nullptr;
  int32_t *ext = nullptr, *inc_ptr = nullptr, *inc_idx = nullptr;
  int32_t* ia = nullptr;  // device copy of ia (value pack)
  // compact order: layout node d is CSRC node d2r[d] (null: the same order);
  // per block the CSRC source of its values (BlockLayout::src_row / src_k)
  int32_t *d2r = nullptr, *src_row = nullptr, *src_k = nullptr;
  // values: plane p of block q of tile t at blk0[t]*nb*nb + p*nblk[t] + q;
  // node diagonal blocks: entry c of node I at c*nn + I (nb diagonal entries,
  // then the in-node lower entries a21, a31, a32)
  double *v64 = nullptr, *d64 = nullptr;
  float *v32 = nullptr, *d32 = nullptr;
};

// ------------------------------------------------------ gather-only product
// (spmv_gather.hip) for patterns the row tiles cannot hold: the transposed
// lists of the CSRC pattern, built on the host.  colptr (neq+1, 0-based
// offsets), then per entry its CSRC position colk and row colrow (0-based),
// rows ascending within a column.
struct GatherLayout {
  std::vector<int32_t> colptr, colk, colrow;
  int longest = 0;  // longest CSRC row
};
void build_gather(int neq, const int32_t* ia1, const int32_t* ja1, GatherLayout& L);

struct DevGather {
  int32_t *ia = nullptr, *ja = nullptr;  // the CSRC pattern (1-based values)
  int32_t *colptr = nullptr, *colk = nullptr, *colrow = nullptr;
  int longest = 0;
};
template <typename TV>
void launch_spmv_gather(int n, const int32_t* ia, const int32_t* ja, const TV* ad, const TV* al,
                        const int32_t* colptr, const int32_t* colk, const int32_t* colrow,
                        const double* x, double* y, hipStream_t s, const int32_t* skip);

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
  // reliable-update mixed precision (FP32 iteration, FP64 residual replacement)
  double rmax, delta, true_rmr;
  int32_t reliable, upd, skip64, nupd;
  // BiCGSTAB omega; GMRES: |e(i)| of the last inner step and its norm h(i+1,i)
  double omega, gm_e, gm_h;
  // progress guard of the mixed methods (device_util.h: progress_guard)
  double last_true;
  int32_t stall, op64, op64_from, last_upd;
  // p.(A p) formed by the product's tiles (node-blocked layouts)
  double pq_tiles;
  // batch: restart flags by iteration parity (set when the progress guard
  // hands over to the FP64 operator; the next p update takes beta = 0)
  int32_t restart[2];
};

constexpr int kRedBlocks = 1024;  // fixed reduction grid (deterministic sums)

// Default node blocks per tile: large 3-dof systems take the largest tiles
// (fewer external partials to write and merge; the FP32 operator's fastest
// size on hex8 100^3); smaller systems and 2-dof nodes keep 512 (more tiles,
// so more workgroups in flight; the batch's 2-D product holds 2 blocks per
// lane).
constexpr int kTileBlocks3 = 1024, kTileBlocks2 = 512;
constexpr int kLargeTileNodes = 32768;

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int neq = 0;
  int64_t nad = 0;
  int tile_nnz = 2048;
  int tile_blocks = 0;         // node blocks per tile; 0: by node size (kTileBlocks3/2)
  // lanes per tile of the node-blocked product (256 or 512), FP64 [0] and
  // FP32-value [1] operators
  int lanes[2] = {256, 256};
  bool tune_tiles = true;      // elastic_autotune also tries tile sizes (else lanes only)
  int layout_nb = 1;           // 1: scalar tiles; 2 or 3: node-blocked tiles
  bool accum32 = false;        // FP32 operator: FP32 staging and sums (else FP64)
  // -1: accum32 follows the system (on for 3-dof layouts of at least
  // kLargeTileNodes nodes, where FP32 sums pay); else set by the caller
  int accum32_set = -1;
  bool allow_blocks = true;
  std::string layout_note;  // why the layout differs from the default (fallbacks), or empty
  TileLayout layout;
  DevTiles tiles;
  BlockLayout blayout;
  DevBlocks blk;
  bool gather = false;  // the gather-only product (layout_nb stays 1)
  DevGather gth;
  const int32_t* inc_ptr = nullptr;  // merge lists of the active layout
  const int32_t* inc_idx = nullptr;
  DevMatrix<double> A64;
  DevMatrix<float> A32;
  bool values_set = false, values32 = false, bdiag_set = false;
  int bdiag_nb = 0;         // block size of the inverse in A64.bm (2 or 3)
  int batch_precond = 2;    // preconditioner of batched solves (2 Jacobi, 6 node blocks)
  // workspace
  double *epart64 = nullptr, *red64 = nullptr;
  double* tdot = nullptr;  // per-tile shares of p.(A p) (node-blocked layouts)
  // FP32 vectors of the mixed PCG with the FP32-sum operator (node-blocked):
  // p at p32, r at p32 + vec32_stride(neq), the Jacobi scale after that,
  // then the FP32 increments of x since the last true-residual replacement
  float* p32 = nullptr;
  double *v64[8] = {nullptr};
  SolveScalars* scal = nullptr;
  SolveScalars* scal_host = nullptr;  // pinned
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  // assembly state (optional)
  void* mesh = nullptr;  // MeshDev (assembly.hip)
  void* j2 = nullptr;    // J2 plasticity state (assembly.hip)
  // GMRES(k) Krylov basis ((k+1) vectors of neq) and its small arrays
  double* gm_basis = nullptr;
  double* gm_small = nullptr;
  int gm_k = 0;
  // autotuning/measurement
  double last_spmv_ms = 0.0;
  int check_every = 16;
  // sampled event timing of PCG's iteration products: every time_stride-th
  // product of an iteration is bracketed by events (0: off); totals accumulate
  // over solves until read (elastic_product_timing)
  int time_stride = 0;
  std::vector<hipEvent_t> tev;  // 2 per sampled product of a poll chunk
  double tprod_ms = 0.0;
  int64_t tprod_n = 0;
  int zero_diag_row = -1;  // first row with ad == 0 (1-based), -1 if none
  // Node order of the node-blocked layout: -1 follows the system (compact
  // tiles for 3-dof layouts of at least kLargeTileNodes nodes), 0 the CSRC
  // order, 1 compact tiles (3-dof layouts).  With
  // compact tiles every vector inside a solve is in the layout's node order:
  // b and x cross between orders when a solve starts and ends (enter/leave
  // of the layout order), so v64[VB] and v64[VX] are in the CSRC order
  // outside a solve, and the Jacobi scale A64.m and bm_dev are in the layout
  // order always.
  int node_order = -1;
  double *xd = nullptr, *bd = nullptr;  // x and b in the layout order (compact only)
  double* bm_dev = nullptr;             // 3x3 block inverse in the layout order
  bool in_layout_order = false;         // v64[VX] / v64[VB] hold xd / bd (inside a solve)
};

// How the external partials of a product reach their rows.
//   nb == 1 (scalar tiles): row i adds epart[idx[k]], k in ptr[i]..ptr[i+1]-1
//   nb > 1 (node blocks): the tiles write nb values per (tile, external
//     node) in tile order; row i = nb*J + r adds epart[nb*idx[k] + r],
//     k in ptr[J]..ptr[J+1]-1 (one index per partial node)
// Either way the partials of a row are added in tile order.
struct MergeArgs {
  const int32_t* ptr;
  const int32_t* idx;
  const double* epart;
  int nb;
  int ep32;  // the partials are FP32 (written by a product with FP32 sums)
};
__device__ inline double merge_row(const MergeArgs& m, int i, double s) {
  if (!m.ptr) return s;
  if (m.nb <= 1) {
    for (int k = m.ptr[i]; k < m.ptr[i + 1]; ++k) s += m.epart[m.idx[k]];
    return s;
  }
  const int J = i / m.nb, r = i - J * m.nb;
  if (m.ep32) {
    const float* e = reinterpret_cast<const float*>(m.epart);
    for (int k = m.ptr[J]; k < m.ptr[J + 1]; ++k) s += (double)e[(int64_t)m.nb * m.idx[k] + r];
    return s;
  }
  for (int k = m.ptr[J]; k < m.ptr[J + 1]; ++k) s += m.epart[(int64_t)m.nb * m.idx[k] + r];
  return s;
}

// device vector slots of a context (c->v64[...])
enum VecSlot { VX = 0, VR, VP, VQ, VB, VT, VU, VW };

int set_error(Context* c, int code, const std::string& msg);
// a product layout has been built (structure set)
inline bool layout_ready(const Context* c) {
  return c->tiles.ntiles > 0 || c->blk.ntiles > 0 || c->gth.colptr != nullptr;
}
// Partials of a product with FP32 sums (node-blocked layouts) are stored as
// FP32 (exactly the values summed); every other product stores FP64.
// distance between the FP32 vectors held at Context::p32 (256-byte aligned)
inline size_t vec32_stride(int n) { return ((size_t)n + 63) & ~(size_t)63; }
inline bool epart_f32(const Context* c, bool fp32_operator) {
  return fp32_operator && c->accum32 && c->layout_nb > 1;
}
// merge lists of the context; ep32: the partials were written by a product
// with FP32 sums (epart_f32)
inline MergeArgs merge_of(const Context* c, const double* epart, bool ep32 = false) {
  return MergeArgs{c->inc_ptr, c->inc_idx, epart, c->layout_nb, ep32 ? 1 : 0};
}
// arrays derived from new A64 values (asynchronous on c->stream): the
// Jacobi inverse, the FP32 operator (scalar tiles) or the block planes
void derive_operator(Context* c);
// x.Kx, ||x||, ||b - Ax||, ||b - Ax||_M and rel_res of the context's x and b
int post_solve_stats(Context* c, elastic_stats* st);

// krylov.hip: BiCGSTAB (Iters.f:1079-1349) and GMRES(k) (Iters.f:528-808) on
// device-resident vectors (b in v64[VB], x left in v64[VX]).  fp32_operator
// selects the FP32-value operator inside the iteration; the FP64 operator
// then recomputes the true residual (BiCGSTAB: reliable replacement as in
// PCG; GMRES: every restart) and convergence is decided on it.
struct KrylovResult {
  int iters = 0, outer = 0, nupd = 0, status = 0, op64_from = -1;
  bool done = false;
  double conv = 0.0;
};
int bicgstab_run(Context* c, bool bdiag, bool fp32_operator, double tol, int maxit, double delta,
                 KrylovResult* out);
int gmres_run(Context* c, bool bdiag, bool fp32_operator, double tol, int maxit, int k,
              KrylovResult* out);
struct MeshDev;
void free_mesh(MeshDev* m);
void free_j2(void* j);
#define ELASTIC_HIP_CHECK(ctx, expr)                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_error((ctx), ELASTIC_HIP_ERROR,                                       \
                       std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)

// kernels (spmv.hip / vector.hip)
template <typename TV, typename T>
void launch_spmv_tiles(const DevTiles& t, const TV* ad, const TV* al, const T* x, T* y, T* epart,
                       hipStream_t s, const int32_t* skip = nullptr);
void launch_merge(Context* c, const double* epart, double* y, bool ep32 = false);
template <typename S, typename D>
void launch_cast(int64_t n, const S* a, D* b, hipStream_t s);
// y = A x with the active layout; fp32 selects the FP32-value operator.  The
// external partials are left in epart for the merge lists c->inc_ptr/inc_idx.
// Optional output of a node-blocked product: x.(A x), formed from per-tile
// shares (part, one per tile) that a one-workgroup kernel adds in tile order
// into *result (result == null: the shares only).
struct DotOut {
  double* part;
  double* result;
};
void launch_spmv(Context* c, bool fp32, const double* x, double* y, double* epart,
                 const int32_t* skip = nullptr, const DotOut* dot = nullptr, bool y32 = false,
                 const float* x32 = nullptr, const float* dx = nullptr);
// (y32: a product with FP32 sums on the node-blocked layout stores its own
// rows of y as FP32 too, exactly; the caller then reads y as float.  x32: an
// FP32 copy of x that such a product stages from instead of x.  dx: FP32
// increments that the FP64 product of the node-blocked layout adds to x when
// it stages it, i.e. the product of x + dx)
// gathers the CSRC values into the node-blocked planes (after every assembly)
void launch_block_pack(Context* c);
// the Jacobi scale 1/ad in the compact layout order (into A64.m)
void launch_recip_layout(Context* c);
// vectors between the CSRC order and the compact layout order:
//   to_layout: dst[nb*d + r] = src[nb*d2r[d] + r];  else the inverse
void launch_permute(int nn, int nb, const int32_t* d2r, const double* src, double* dst,
                    bool to_layout, hipStream_t s);
// v64[VB] -> layout order and v64[VX] / v64[VB] swapped with xd / bd (enter),
// and back with x -> CSRC order (leave); no-ops for layouts in the CSRC order
int enter_layout_order(Context* c);
int leave_layout_order(Context* c, bool keep_x = true);
// the block inverse that solves apply: bm_dev (compact order) or A64.bm
inline const double* solver_bm(const Context* c) { return c->bm_dev ? c->bm_dev : c->A64.bm; }
template <typename T>
void launch_recip(int n, const T* ad, T* m, hipStream_t s);
void launch_post(const MergeArgs& mg, int n, const double* x, double* y, const double* b,
                 const double* m, double* part, double* out6, hipStream_t s);
void launch_sum_parts(const double* part, int nsum, double* out, hipStream_t s);

// pcg.hip: one PCG solve on device-resident FP64 vectors.  The operator's
// values are TV (double, or float for the mixed-precision method); with
// `reliable` the recursively updated residual is replaced by the FP64 true
// residual b - A64 x (ad64/al64) whenever it has dropped by `delta`, and the
// stopping test is made on that true residual.
struct PcgVecs {
  double *x, *r, *p, *q, *b, *m, *bm, *t, *epart;
};
int pcg_run(Context* c, const PcgVecs& V, bool fp32_operator, double tol, int maxit,
            bool reliable, double delta, double* dslots);
}  // namespace elastic

// the opaque handle of the C-ABI
struct elastic_ctx {
  elastic::Context c;
  std::vector<int32_t> ia, ja;  // host copy of the pattern (block preconditioner, drop-in)
};
