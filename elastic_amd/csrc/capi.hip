// C-ABI of the library (include/elastic_hip.h): contexts, host set-up,
// system upload, products, solves and the drop-in for the reference's solv.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "elastic_internal.h"


namespace elastic {

int set_error(Context* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

template <typename T>
static int dalloc(Context* c, T** p, size_t n) {
  if (*p) {
    hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(c, ELASTIC_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return ELASTIC_OK;
}

template <typename T>
static int upload(Context* c, T** p, const T* host, size_t n) {
  int rc = dalloc(c, p, n);
  if (rc) return rc;
  if (n) ELASTIC_HIP_CHECK(c, hipMemcpyAsync(*p, host, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return ELASTIC_OK;
}

static void free_tiles(DevTiles& t) {
  void* ptrs[] = {t.row0, t.nrows, t.nnz, t.ncols, t.ext0, t.nz0, t.rp0, t.seg0, t.word,
                  t.rowptr, t.segptr, t.ext, t.inc_ptr, t.inc_idx};
  for (void* p : ptrs)
    if (p) hipFree(p);
  t = DevTiles();
}

static void free_blocks(DevBlocks& b) {
  void* ptrs[] = {b.row0,   b.nrows,  b.nblk, b.ncols,   b.ext0,    b.seg0, b.rp0,
                  b.blk0,   b.word,   b.rowptr, b.segptr, b.ext, b.inc_ptr, b.inc_idx,
                  b.ia,     b.v64,  b.d64,     b.v32, b.d32, b.d2r, b.src_row, b.src_k};
  for (void* p : ptrs)
    if (p) hipFree(p);
  b = DevBlocks();
}

static void free_gather(DevGather& g) {
  void* ptrs[] = {g.ia, g.ja, g.colptr, g.colk, g.colrow};
  for (void* p : ptrs)
    if (p) hipFree(p);
  g = DevGather();
}

// device layout of the product (tiles or node blocks, FP32 copies, partials)
static void free_layout(Context* c) {
  free_tiles(c->tiles);
  free_blocks(c->blk);
  free_gather(c->gth);
  c->gather = false;
  c->inc_ptr = c->inc_idx = nullptr;
  // (a layout is never freed inside a solve: x and b are in the CSRC order)
  void* ptrs[] = {c->A32.ad, c->A32.al, c->epart64, c->tdot, c->p32, c->xd, c->bd, c->bm_dev};
  for (void* p : ptrs)
    if (p) hipFree(p);
  c->A32 = DevMatrix<float>();
  c->epart64 = nullptr;
  c->tdot = nullptr;
  c->p32 = nullptr;
  c->xd = c->bd = c->bm_dev = nullptr;
  c->bdiag_set = false;  // its layout-order copy went with the layout
}

static void free_system(Context* c) {
  free_layout(c);
  void* ptrs[] = {c->A64.ad, c->A64.al, c->A64.m, c->A64.bm};
  for (void* p : ptrs)
    if (p) hipFree(p);
  c->A64 = DevMatrix<double>();
  for (auto& v : c->v64)
    if (v) hipFree(v), v = nullptr;
  if (c->gm_basis) hipFree(c->gm_basis);
  if (c->gm_small) hipFree(c->gm_small);
  c->gm_basis = c->gm_small = nullptr;
  c->gm_k = 0;
  c->values_set = c->values32 = c->bdiag_set = false;
}

}  // namespace elastic

using namespace elastic;


// Builds the device layout of the product from x->ia / x->ja (node-blocked
// when the pattern allows it, else scalar row tiles) with the context's tile
// sizes; values are packed separately (derive_operator).  A pattern that the
// node-blocked tiles cannot hold (a node row of more than 64 lower blocks, or
// longer than the block tile) falls back to scalar tiles, and the scalar
// tiles grow to the longest CSRC row; a row of more than
// kTileColsMax - kTileRowsMax lower entries takes the gather-only product
// (spmv_gather.hip), which holds any pattern.
// node blocks per tile the layout is built with (tile_blocks 0: by node size)
static int effective_tile_blocks(const Context* c, int nb) {
  if (c->tile_blocks > 0) return c->tile_blocks;
  return nb == 3 && c->neq / 3 >= kLargeTileNodes ? kTileBlocks3 : kTileBlocks2;
}

static bool getenv_is(const char* name, const char* value) {
  const char* v = getenv(name);
  return v && std::string(v) == value;
}

static int build_layout(elastic_ctx* x) {
  Context* c = &x->c;
  const int neq = c->neq;
  const int64_t nad = c->nad;
  const int32_t* ia = x->ia.data();
  const int32_t* ja = x->ja.data();
  int nb = 1;
  const char* lay = getenv("ELASTIC_LAYOUT");
  if (c->allow_blocks && !(lay && std::string(lay) == "scalar")) nb = detect_node_blocks(neq, ia, ja);
  c->layout_note.clear();
  if (nb > 1) {
    // compact node order for large 3-dof layouts unless the caller chose (a
    // node's 3 equations are the triple that the 3x3 block preconditioner
    // inverts in either order).  Small systems keep the CSRC order, whose
    // sums follow the rows as the reference's loops do: BiCGSTAB's iteration
    // count moves with the order of its sums (hex8 5^3: 74 iterations in the
    // CSRC order, 80 in the compact one).
    const bool compact =
        nb == 3 && (c->node_order == 1 || (c->node_order < 0 && neq / 3 >= kLargeTileNodes));
    try {
      build_block_tiles(neq, nb, ia, ja, effective_tile_blocks(c, nb), c->blayout, compact);
    } catch (const std::exception& e) {
      c->layout_note = std::string("scalar tiles: ") + e.what();
      c->blayout = BlockLayout();
      nb = 1;
    }
  }
  if (nb == 1) {
    int longest = 0;
    for (int i = 0; i < neq; ++i) longest = std::max(longest, ia[i + 1] - ia[i]);
    if (longest > kTileColsMax - kTileRowsMax || getenv_is("ELASTIC_LAYOUT", "gather")) {
      c->layout_note += (c->layout_note.empty() ? "" : "; ") + std::string("gather product (row of ") +
                        std::to_string(longest) + " entries)";
      GatherLayout G;
      try {
        build_gather(neq, ia, ja, G);
      } catch (const std::bad_alloc&) {
        return set_error(c, ELASTIC_OOM, "host memory for the transposed pattern");
      }
      DevGather& g = c->gth;
      int rc = upload(c, &g.ia, ia, (size_t)neq + 1);
      rc = rc ? rc : upload(c, &g.ja, ja, (size_t)nad);
      rc = rc ? rc : upload(c, &g.colptr, G.colptr.data(), G.colptr.size());
      rc = rc ? rc : upload(c, &g.colk, G.colk.data(), G.colk.size());
      rc = rc ? rc : upload(c, &g.colrow, G.colrow.data(), G.colrow.size());
      rc = rc ? rc : dalloc(c, &c->A32.ad, neq);
      rc = rc ? rc : dalloc(c, &c->A32.al, (size_t)nad);
      rc = rc ? rc : dalloc(c, &c->epart64, 1);
      if (rc) return rc;
      g.longest = G.longest;
      c->gather = true;
      c->layout_nb = 1;
      c->inc_ptr = c->inc_idx = nullptr;
      if (c->accum32_set < 0) c->accum32 = false;
      ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
      return ELASTIC_OK;
    }
    const int tn = std::max(c->tile_nnz, longest);
    if (tn != c->tile_nnz)
      c->layout_note += (c->layout_note.empty() ? "" : "; ") +
                        std::string("row tiles of ") + std::to_string(tn) + " nonzeros";
    try {
      build_tiles(neq, ia, ja, tn, c->layout);
    } catch (const std::exception& e) {
      return set_error(c, ELASTIC_BAD_ARG, e.what());
    }
  }
  c->layout_nb = nb;
  // FP32 sums by default where they pay: large 3-dof systems (hex8 100^3:
  // mixed solve 254 ms with them, 419 ms without; tet4 80^3: 605 / 880 ms)
  if (c->accum32_set < 0) c->accum32 = nb == 3 && neq / 3 >= kLargeTileNodes;
  int rc = 0;
  int64_t nepart = 0;
  if (nb == 1) {
    TileLayout& L = c->layout;
    DevTiles& t = c->tiles;
    rc = rc ? rc : upload(c, &t.row0, L.row0.data(), L.row0.size());
    rc = rc ? rc : upload(c, &t.nrows, L.nrows.data(), L.nrows.size());
    rc = rc ? rc : upload(c, &t.nnz, L.nnz.data(), L.nnz.size());
    rc = rc ? rc : upload(c, &t.ncols, L.ncols.data(), L.ncols.size());
    rc = rc ? rc : upload(c, &t.ext0, L.ext0.data(), L.ext0.size());
    rc = rc ? rc : upload(c, &t.nz0, L.nz0.data(), L.nz0.size());
    rc = rc ? rc : upload(c, &t.rp0, L.rp0.data(), L.rp0.size());
    rc = rc ? rc : upload(c, &t.seg0, L.seg0.data(), L.seg0.size());
    rc = rc ? rc : upload(c, &t.word, L.word.data(), L.word.size());
    rc = rc ? rc : upload(c, &t.rowptr, L.rowptr.data(), L.rowptr.size());
    rc = rc ? rc : upload(c, &t.segptr, L.segptr.data(), L.segptr.size());
    rc = rc ? rc : upload(c, &t.ext, L.ext.data(), L.ext.size());
    rc = rc ? rc : upload(c, &t.inc_ptr, L.inc_ptr.data(), L.inc_ptr.size());
    rc = rc ? rc : upload(c, &t.inc_idx, L.inc_idx.data(), L.inc_idx.size());
    t.ntiles = L.ntiles;
    t.nnz_max = L.nnz_max;
    t.ncols_max = L.ncols_max;
    t.next_total = L.next_total;
    c->inc_ptr = t.inc_ptr;
    c->inc_idx = t.inc_idx;
    nepart = L.next_total;
  } else {
    BlockLayout& L = c->blayout;
    DevBlocks& b = c->blk;
    rc = rc ? rc : upload(c, &b.row0, L.row0.data(), L.row0.size());
    rc = rc ? rc : upload(c, &b.nrows, L.nrows.data(), L.nrows.size());
    rc = rc ? rc : upload(c, &b.nblk, L.nblk.data(), L.nblk.size());
    rc = rc ? rc : upload(c, &b.ncols, L.ncols.data(), L.ncols.size());
    rc = rc ? rc : upload(c, &b.ext0, L.ext0.data(), L.ext0.size());
    rc = rc ? rc : upload(c, &b.seg0, L.seg0.data(), L.seg0.size());
    rc = rc ? rc : upload(c, &b.rp0, L.rp0.data(), L.rp0.size());
    rc = rc ? rc : upload(c, &b.blk0, L.blk0.data(), L.blk0.size());
    rc = rc ? rc : upload(c, &b.word, L.word.data(), L.word.size());
    rc = rc ? rc : upload(c, &b.rowptr, L.rowptr.data(), L.rowptr.size());
    rc = rc ? rc : upload(c, &b.segptr, L.segptr.data(), L.segptr.size());
    rc = rc ? rc : upload(c, &b.ext, L.ext.data(), L.ext.size());
    rc = rc ? rc : upload(c, &b.inc_ptr, L.inc_ptr.data(), L.inc_ptr.size());
    rc = rc ? rc : upload(c, &b.inc_idx, L.inc_idx.data(), L.inc_idx.size());
    rc = rc ? rc : upload(c, &b.ia, ia, (size_t)neq + 1);
    rc = rc ? rc : upload(c, &b.src_row, L.src_row.data(), L.src_row.size());
    rc = rc ? rc : upload(c, &b.src_k, L.src_k.data(), L.src_k.size());
    if (L.compact) {
      rc = rc ? rc : upload(c, &b.d2r, L.d2r.data(), L.d2r.size());
      rc = rc ? rc : dalloc(c, &c->xd, (size_t)neq);
      rc = rc ? rc : dalloc(c, &c->bd, (size_t)neq);
    }
    // one padding block: a tile without blocks still loads in bounds
    const size_t nvals = ((size_t)L.nblocks + 1) * nb * nb, ndiag = (size_t)L.nn * (nb * (nb + 1) / 2);
    rc = rc ? rc : dalloc(c, &b.v64, nvals);
    rc = rc ? rc : dalloc(c, &b.v32, nvals);
    rc = rc ? rc : dalloc(c, &b.d64, ndiag);
    rc = rc ? rc : dalloc(c, &b.d32, ndiag);
    b.nb = nb;
    b.nn = L.nn;
    b.ntiles = L.ntiles;
    b.nblk_max = L.nblk_max;
    b.ncols_max = L.ncols_max;
    b.nblocks = L.nblocks;
    b.next_nodes = L.next_nodes;
    c->inc_ptr = b.inc_ptr;
    c->inc_idx = b.inc_idx;  // node-granular lists
    nepart = (int64_t)nb * L.next_nodes;
    rc = rc ? rc : dalloc(c, &c->tdot, (size_t)L.ntiles);
    rc = rc ? rc : dalloc(c, &c->p32, 4 * vec32_stride(neq));
  }
  if (nb == 1) {  // FP32 operator of the scalar tiles (the blocked layout has its planes)
    rc = rc ? rc : dalloc(c, &c->A32.ad, neq);
    rc = rc ? rc : dalloc(c, &c->A32.al, (size_t)nad);
  }
  rc = rc ? rc : dalloc(c, &c->epart64, (size_t)nepart);
  if (rc) return rc;
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  // the host streams are no longer needed
  c->layout.word.clear();
  c->layout.word.shrink_to_fit();
  c->blayout.word.clear();
  c->blayout.word.shrink_to_fit();
  for (auto* v : {&c->blayout.src_row, &c->blayout.src_k, &c->blayout.d2r}) {
    v->clear();
    v->shrink_to_fit();
  }
  return ELASTIC_OK;
}

namespace elastic {

// Solves on a compact layout run on vectors in the layout's node order: b
// crosses into it when a solve starts and x back when it ends (one gather
// each, against the hundreds of products of a solve).
int enter_layout_order(Context* c) {
  if (!c->blk.d2r || c->in_layout_order) return ELASTIC_OK;
  launch_permute(c->blk.nn, c->blk.nb, c->blk.d2r, c->v64[VB], c->bd, true, c->stream);
  std::swap(c->v64[VB], c->bd);
  std::swap(c->v64[VX], c->xd);
  c->in_layout_order = true;
  return ELASTIC_OK;
}

int leave_layout_order(Context* c, bool keep_x) {
  if (!c->in_layout_order) return ELASTIC_OK;
  std::swap(c->v64[VB], c->bd);
  std::swap(c->v64[VX], c->xd);
  c->in_layout_order = false;
  if (keep_x) launch_permute(c->blk.nn, c->blk.nb, c->blk.d2r, c->xd, c->v64[VX], false, c->stream);
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  return ELASTIC_OK;
}

}  // namespace elastic

// Rebuilds the layout with new tile sizes.  If that fails, the previous
// sizes and layout are restored (the values are re-derived), so the context
// stays usable and the call reports the failure; if even the restore fails,
// the values are dropped so that solves report NOT_READY.
static int relayout(elastic_ctx* x, int tile_blocks, int tile_nnz) {
  Context* c = &x->c;
  const int ob = c->tile_blocks, on = c->tile_nnz;
  hipSetDevice(c->device);
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  c->tile_blocks = tile_blocks;
  c->tile_nnz = tile_nnz;
  free_layout(c);
  int rc = build_layout(x);
  if (rc) {
    const std::string why = c->err;
    c->tile_blocks = ob;
    c->tile_nnz = on;
    free_layout(c);
    const int rc2 = build_layout(x);
    if (rc2) {
      free_layout(c);
      c->values_set = c->values32 = c->bdiag_set = false;
      return set_error(c, rc2, why + "; rebuilding the previous layout failed too: " + c->err);
    }
    if (c->values_set) derive_operator(c);
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return set_error(c, rc, why + " (previous layout kept)");
  }
  if (c->values_set) derive_operator(c);
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

extern "C" {

int elastic_create(elastic_ctx** out, int32_t device) {
  if (!out) return ELASTIC_BAD_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ELASTIC_HIP_ERROR;
  if (device < 0 || device >= ndev) return ELASTIC_BAD_ARG;
  if (hipSetDevice(device) != hipSuccess) return ELASTIC_HIP_ERROR;
  auto* x = new elastic_ctx();
  Context* c = &x->c;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess ||
      hipMalloc((void**)&c->scal, sizeof(SolveScalars)) != hipSuccess ||
      hipHostMalloc((void**)&c->scal_host, sizeof(SolveScalars), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&c->red64, sizeof(double) * (8 * kRedBlocks + 64)) != hipSuccess) {
    delete x;
    return ELASTIC_HIP_ERROR;
  }
  if (const char* ac = getenv("ELASTIC_ACCUM")) {
    c->accum32 = atoi(ac) == 32;
    c->accum32_set = 1;
  }
  if (const char* tb = getenv("ELASTIC_TILE_BLOCKS"))
    c->tile_blocks = std::max(32, std::min(kBlkTileMax, atoi(tb)));
  if (const char* l = getenv("ELASTIC_LANES64")) c->lanes[0] = atoi(l) == 512 ? 512 : 256;
  if (const char* l = getenv("ELASTIC_LANES32")) c->lanes[1] = atoi(l) == 512 ? 512 : 256;
  const char* ce = getenv("ELASTIC_CHECK_EVERY");
  if (ce) c->check_every = std::max(1, atoi(ce));
  *out = x;
  return ELASTIC_OK;
}

int elastic_destroy(elastic_ctx* x) {
  if (!x) return ELASTIC_OK;
  Context* c = &x->c;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  free_system(c);
  if (c->mesh) free_mesh((MeshDev*)c->mesh);
  c->mesh = nullptr;
  if (c->j2) free_j2(c->j2);
  c->j2 = nullptr;
  if (c->scal) hipFree(c->scal);
  if (c->scal_host) hipHostFree(c->scal_host);
  if (c->red64) hipFree(c->red64);
  for (hipEvent_t e : {c->ev0, c->ev1, c->ev2, c->ev3})
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : c->tev) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete x;
  return ELASTIC_OK;
}

const char* elastic_last_error(const elastic_ctx* x) { return x ? x->c.err.c_str() : "null context"; }

// ------------------------------------------------------------ host set-up
int elastic_reord(int32_t nnode, int32_t numel, int32_t nen, const int32_t* ix, int32_t* perm) {
  if (nnode <= 0 || numel < 0 || nen <= 0 || !ix || !perm) return ELASTIC_BAD_ARG;
  try {
    rcm_order(nnode, numel, nen, ix, perm);
  } catch (const std::bad_alloc&) {
    return ELASTIC_OOM;
  }
  return ELASTIC_OK;
}

int elastic_numeq(int32_t nnode, int32_t ndf, const int32_t* perm, int32_t* id, int32_t* neq) {
  if (nnode <= 0 || ndf <= 0 || !perm || !id || !neq) return ELASTIC_BAD_ARG;
  *neq = number_equations(id, perm, nnode, ndf);
  return ELASTIC_OK;
}

int elastic_csrstruct(int32_t nnode, int32_t numel, int32_t nen, int32_t ndf, int32_t neq,
                      const int32_t* ix, const int32_t* id, int32_t* ia, int32_t* ja,
                      int64_t* nad) {
  if (nnode <= 0 || neq < 0 || !ix || !id || !ia || !nad) return ELASTIC_BAD_ARG;
  try {
    std::vector<int32_t> vi, vj;
    csrc_pattern(id, ix, nnode, numel, nen, ndf, neq, vi, vj);
    std::copy(vi.begin(), vi.end(), ia);
    *nad = (int64_t)vj.size();
    if (ja) std::copy(vj.begin(), vj.end(), ja);
  } catch (const std::bad_alloc&) {
    return ELASTIC_OOM;
  }
  return ELASTIC_OK;
}

// ----------------------------------------------------------- device system
int elastic_set_tile_blocks(elastic_ctx* x, int32_t tile_blocks) {
  if (!x || tile_blocks < 32 || tile_blocks > kBlkTileMax) return ELASTIC_BAD_ARG;
  x->c.tile_blocks = tile_blocks;
  return ELASTIC_OK;
}

int elastic_set_tile_nnz(elastic_ctx* x, int32_t tile_nnz) {
  if (!x || tile_nnz < 64 || tile_nnz > kTileColsMax - kTileRowsMax) return ELASTIC_BAD_ARG;
  x->c.tile_nnz = tile_nnz;
  return ELASTIC_OK;
}

int elastic_set_structure(elastic_ctx* x, int32_t neq, int64_t nad, const int32_t* ia,
                          const int32_t* ja) {
  if (!x || neq <= 0 || nad < 0 || !ia || (nad > 0 && !ja)) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  hipSetDevice(c->device);
  if (ia[0] != 1 || (int64_t)ia[neq] - 1 != nad)
    return set_error(c, ELASTIC_BAD_ARG, "ia must start at 1 and end at nad+1");
  for (int i = 0; i < neq; ++i) {
    if (ia[i + 1] < ia[i]) return set_error(c, ELASTIC_BAD_ARG, "ia not monotone");
    for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; ++k)
      if (ja[k] < 1 || ja[k] > i || (k > ia[i] - 1 && ja[k] <= ja[k - 1]))
        return set_error(c, ELASTIC_BAD_ARG, "ja must be strictly lower and ascending per row");
  }
  hipStreamSynchronize(c->stream);
  free_system(c);
  x->ia.assign(ia, ia + neq + 1);
  x->ja.assign(ja, ja + nad);
  c->neq = neq;
  c->nad = nad;
  int rc = build_layout(x);
  if (rc) {
    free_system(c);
    x->ia.clear();
    x->ja.clear();
    c->neq = 0;
    c->nad = 0;
    return rc;
  }
  rc = rc ? rc : dalloc(c, &c->A64.ad, neq);
  rc = rc ? rc : dalloc(c, &c->A64.al, (size_t)nad);
  rc = rc ? rc : dalloc(c, &c->A64.m, neq);
  for (auto& v : c->v64) rc = rc ? rc : dalloc(c, &v, neq);
  if (rc) {
    free_system(c);
    x->ia.clear();
    x->ja.clear();
    c->neq = 0;
    c->nad = 0;
    return rc;
  }
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

namespace elastic {
int solve_resident(elastic_ctx* x, int solver, int precond, int precision, double tol, int maxit,
                   int ngram, elastic_stats* st);
}

// Average device time of the products of a short resident PCG run of
// `iters` iterations (b = ones; the caller's b is kept and put back; x is
// overwritten), sampled with HIP events on every iteration.
static int time_in_solve(elastic_ctx* x, int precision, int iters, double* ms) {
  Context* c = &x->c;
  const size_t nb = sizeof(double) * (size_t)c->neq;
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->v64[VW], c->v64[VB], nb, hipMemcpyDeviceToDevice, c->stream));
  const std::vector<double> ones((size_t)c->neq, 1.0);
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->v64[VB], ones.data(), nb, hipMemcpyHostToDevice, c->stream));
  const int stride = c->time_stride;
  const double tsum = c->tprod_ms;
  const int64_t tn = c->tprod_n;
  c->time_stride = 1;
  c->tprod_ms = 0.0;
  c->tprod_n = 0;
  elastic_stats st;
  int rc = solve_resident(x, ELASTIC_SOLVER_PCG, ELASTIC_PRECOND_DIAG, precision, 1e-300, iters,
                          50, &st);
  *ms = c->tprod_n ? c->tprod_ms / (double)c->tprod_n : 1e300;
  c->time_stride = stride;
  c->tprod_ms = tsum;
  c->tprod_n = tn;
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->v64[VB], c->v64[VW], nb, hipMemcpyDeviceToDevice, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return rc;
}

// Autotuner: for each candidate layout (node-blocked: tile size x lanes per
// tile; scalar rows: tile size) runs a short resident PCG of the given
// precision and keeps the candidate whose products are fastest on average.
// The tile size is shared by both operators (one layout); the lanes are
// kept per precision.  With the option "tune_tiles" = 0 only the lanes are
// tuned.
int elastic_autotune(elastic_ctx* x, int32_t precision, int32_t reps, int32_t* best,
                     double* best_ms) {
  if (!x || reps <= 0) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!c->values_set) return set_error(c, ELASTIC_NOT_READY, "values not set");
  if (c->gather) return set_error(c, ELASTIC_BAD_ARG, "the gather product has no tiles to tune");
  hipSetDevice(c->device);
  const bool blocked = c->layout_nb > 1;
  const int pi = precision == ELASTIC_MIXED ? 1 : 0;
  auto current = [&]() { return blocked ? effective_tile_blocks(c, c->layout_nb) : c->tile_nnz; };
  // the current layout is timed first and a candidate replaces the best
  // only when faster by more than 2%: timing noise does not move the layout
  std::vector<int> tiles = {current()};
  if (c->tune_tiles)
    for (int v : blocked ? std::vector<int>{256, 512, 768, 1024}
                         : std::vector<int>{1024, 2048, 3072, 3840})
      if (v != tiles[0]) tiles.push_back(v);
  const int iters = std::max(16, (int)reps);
  const bool verbose = getenv("ELASTIC_VERBOSE") != nullptr;
  int bt = -1, bl = 256;
  double tmin = 1e300;
  for (int v : tiles) {
    if (v != current()) {
      int rc = blocked ? relayout(x, v, c->tile_nnz) : relayout(x, c->tile_blocks, v);
      if (rc) return rc;
    }
    for (int lanes : {256, 512}) {
      if (lanes == 512 && (!blocked || c->blk.nblk_max <= 256)) continue;
      c->lanes[pi] = lanes;
      double ms = 0.0;
      int rc = time_in_solve(x, precision, iters, &ms);
      if (rc && rc != ELASTIC_MAXIT) return rc;
      if (verbose) fprintf(stderr, "elastic_autotune: tile %d lanes %d product %.4f ms\n", v, lanes, ms);
      if (ms < 0.98 * tmin) tmin = ms, bt = v, bl = lanes;
    }
  }
  c->lanes[pi] = bl;
  if (bt > 0 && bt != current()) {
    int rc = blocked ? relayout(x, bt, c->tile_nnz) : relayout(x, c->tile_blocks, bt);
    if (rc) return rc;
  }
  if (best) *best = current();
  if (best_ms) *best_ms = tmin;
  return ELASTIC_OK;
}

// Current value of a tuning knob (names of elastic_set_option, plus
// "lanes64"/"lanes32" and "layout_nb").
int elastic_get_option(elastic_ctx* x, const char* name, int64_t* value) {
  if (!x || !name || !value) return ELASTIC_BAD_ARG;
  const Context* c = &x->c;
  const std::string n(name);
  if (n == "accum32") *value = c->accum32 ? 1 : 0;
  else if (n == "batch_precond") *value = c->batch_precond;
  else if (n == "check_every") *value = c->check_every;
  else if (n == "time_products") *value = c->time_stride;
  else if (n == "tile_blocks") *value = effective_tile_blocks(c, c->layout_nb > 1 ? c->layout_nb : 3);
  else if (n == "tile_nnz") *value = c->tile_nnz;
  else if (n == "lanes64") *value = c->lanes[0];
  else if (n == "lanes32") *value = c->lanes[1];
  else if (n == "tune_tiles") *value = c->tune_tiles ? 1 : 0;
  else if (n == "layout_nb") *value = c->gather ? 0 : c->layout_nb;
  else if (n == "node_order") *value = c->blk.d2r ? 1 : 0;
  else return set_error((Context*)c, ELASTIC_BAD_ARG, "unknown option " + n);
  return ELASTIC_OK;
}

// Tuning knobs of a context (no counterpart in the reference).
int elastic_set_option(elastic_ctx* x, const char* name, int64_t value) {
  if (!x || !name) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  const std::string n(name);
  if (n == "accum32") {
    c->accum32 = value != 0;
    c->accum32_set = 1;
  } else if (n == "batch_precond") {
    // preconditioner of the context's batched solves: 2 Jacobi, 6 block
    // Jacobi with the layout's node blocks (2x2 in 2-D, 3x3 in 3-D)
    if (value != ELASTIC_PRECOND_DIAG && value != ELASTIC_PRECOND_BDIAG3) return ELASTIC_BAD_ARG;
    c->batch_precond = (int)value;
  } else if (n == "lanes64" || n == "lanes32") {
    if (value != 256 && value != 512) return ELASTIC_BAD_ARG;
    c->lanes[n == "lanes32" ? 1 : 0] = (int)value;
  } else if (n == "tune_tiles") {
    c->tune_tiles = value != 0;
  } else if (n == "node_order") {
    // 0: the node-blocked tiles follow the CSRC node order; 1: compact tiles
    // (3-dof layouts; 2-dof layouts keep the CSRC order); until set, compact
    // for 3-dof layouts of at least kLargeTileNodes nodes
    if (value != 0 && value != 1) return ELASTIC_BAD_ARG;
    c->node_order = (int)value;
    if (c->neq > 0 && layout_ready(c)) return relayout(x, c->tile_blocks, c->tile_nnz);
  } else if (n == "check_every") {
    if (value < 1 || value > 4096) return ELASTIC_BAD_ARG;
    c->check_every = (int)value;
  } else if (n == "time_products") {
    if (value < 0 || value > 4096) return ELASTIC_BAD_ARG;
    c->time_stride = (int)value;
  } else if (n == "tile_blocks" || n == "tile_nnz") {
    const bool blocks = n == "tile_blocks";
    if (blocks && (value < 32 || value > kBlkTileMax)) return ELASTIC_BAD_ARG;
    if (!blocks && (value < 64 || value > kTileColsMax - kTileRowsMax)) return ELASTIC_BAD_ARG;
    const int tb = blocks ? (int)value : c->tile_blocks;
    const int tn = blocks ? c->tile_nnz : (int)value;
    if (c->neq > 0 && layout_ready(c)) return relayout(x, tb, tn);  // now
    c->tile_blocks = tb;
    c->tile_nnz = tn;
  } else {
    return set_error(c, ELASTIC_BAD_ARG, "unknown option " + n);
  }
  return ELASTIC_OK;
}

int elastic_set_values(elastic_ctx* x, const double* ad, const double* al) {
  if (!x || !ad || (x->c.nad > 0 && !al)) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!layout_ready(c)) return set_error(c, ELASTIC_NOT_READY, "structure not set");
  hipSetDevice(c->device);
  c->zero_diag_row = -1;
  for (int i = 0; i < c->neq; ++i)
    if (ad[i] == 0.0) {
      c->zero_diag_row = i + 1;
      break;
    }
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->A64.ad, ad, sizeof(double) * c->neq, hipMemcpyHostToDevice, c->stream));
  if (c->nad)
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->A64.al, al, sizeof(double) * c->nad, hipMemcpyHostToDevice, c->stream));
  derive_operator(c);
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  c->values_set = true;
  c->bdiag_set = false;
  return ELASTIC_OK;
}

int elastic_get_values(elastic_ctx* x, double* ad, double* al) {
  if (!x || !ad) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!c->values_set) return set_error(c, ELASTIC_NOT_READY, "values not set");
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(ad, c->A64.ad, sizeof(double) * c->neq, hipMemcpyDeviceToHost, c->stream));
  if (al && c->nad)
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(al, c->A64.al, sizeof(double) * c->nad, hipMemcpyDeviceToHost, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

static int spmv_resident(Context* c, int precision, const double* x64, double* y64) {
  launch_spmv(c, precision == ELASTIC_MIXED, x64, y64, c->epart64);
  launch_merge(c, c->epart64, y64, epart_f32(c, precision == ELASTIC_MIXED));
  return ELASTIC_OK;
}

int elastic_spmv(elastic_ctx* x, int32_t precision, const double* xv, double* yv) {
  if (!x || !xv || !yv) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!c->values_set) return set_error(c, ELASTIC_NOT_READY, "values not set");
  hipSetDevice(c->device);
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->v64[VT], xv, sizeof(double) * c->neq, hipMemcpyHostToDevice, c->stream));
  double* yd = c->v64[VU];
  if (c->blk.d2r) {  // compact layout: the product runs in its node order
    launch_permute(c->blk.nn, c->blk.nb, c->blk.d2r, c->v64[VT], c->v64[VW], true, c->stream);
    spmv_resident(c, precision, c->v64[VW], c->v64[VU]);
    launch_permute(c->blk.nn, c->blk.nb, c->blk.d2r, c->v64[VU], c->v64[VT], false, c->stream);
    yd = c->v64[VT];
  } else {
    spmv_resident(c, precision, c->v64[VT], c->v64[VU]);
  }
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(yv, yd, sizeof(double) * c->neq, hipMemcpyDeviceToHost, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  return ELASTIC_OK;
}

int elastic_time_spmv(elastic_ctx* x, int32_t precision, int32_t reps, double* total_ms,
                      double* kernel_ms) {
  if (!x || reps <= 0) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!c->values_set) return set_error(c, ELASTIC_NOT_READY, "values not set");
  hipSetDevice(c->device);
  const bool f32 = precision == ELASTIC_MIXED;
  // the product as PCG launches it: on node-blocked layouts with the tiles'
  // shares of x.(Ax) (kernel time), then their sum and the merge (total)
  double* dres = c->red64 + 8 * kRedBlocks + 32;  // clear of the other scalar slots
  const DotOut shares{c->tdot, nullptr}, full{c->tdot, dres};
  auto tiles = [&]() {
    launch_spmv(c, f32, c->v64[VT], c->v64[VU], c->epart64, nullptr, c->tdot ? &shares : nullptr);
  };
  auto merge = [&]() {
    launch_spmv(c, f32, c->v64[VT], c->v64[VU], c->epart64, nullptr, c->tdot ? &full : nullptr);
    launch_merge(c, c->epart64, c->v64[VU], epart_f32(c, f32));
  };
  for (int i = 0; i < 3; ++i) merge();  // warm-up
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < reps; ++i) tiles();
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
  for (int i = 0; i < reps; ++i) merge();
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev2, c->stream));
  ELASTIC_HIP_CHECK(c, hipEventSynchronize(c->ev2));
  float a = 0, b = 0;
  hipEventElapsedTime(&a, c->ev0, c->ev1);
  hipEventElapsedTime(&b, c->ev1, c->ev2);
  if (kernel_ms) *kernel_ms = a / reps;
  if (total_ms) *total_ms = b / reps;
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  return ELASTIC_OK;
}

int elastic_product_timing(elastic_ctx* x, double* avg_ms, int64_t* count, int32_t reset) {
  if (!x || !avg_ms || !count) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  *count = c->tprod_n;
  *avg_ms = c->tprod_n ? c->tprod_ms / (double)c->tprod_n : 0.0;
  if (reset) {
    c->tprod_ms = 0.0;
    c->tprod_n = 0;
  }
  return ELASTIC_OK;
}

int elastic_layout_info(elastic_ctx* x, int64_t* info) {
  if (!x || !info) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (c->gather) {  // no tiles: the longest row, and 0 as the block size
    info[0] = 0;
    info[1] = c->gth.longest;
    info[2] = 0;
    info[3] = 0;
  } else if (c->layout_nb > 1) {
    info[0] = c->blk.ntiles;
    info[1] = c->blk.nblk_max;
    info[2] = c->blk.next_nodes;
    info[3] = c->layout_nb;
  } else {
    info[0] = c->tiles.ntiles;
    info[1] = c->tiles.nnz_max;
    info[2] = c->tiles.next_total;
    info[3] = 1;
  }
  return ELASTIC_OK;
}

int elastic_upload_rhs(elastic_ctx* x, const double* b) {
  if (!x || !b) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  if (!layout_ready(c)) return set_error(c, ELASTIC_NOT_READY, "structure not set");
  hipSetDevice(c->device);
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(c->v64[VB], b, sizeof(double) * c->neq, hipMemcpyHostToDevice, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

int elastic_download_x(elastic_ctx* x, double* xv) {
  if (!x || !xv) return ELASTIC_BAD_ARG;
  Context* c = &x->c;
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(xv, c->v64[VX], sizeof(double) * c->neq, hipMemcpyDeviceToHost, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ solves
namespace elastic {

int ensure_bdiag(elastic_ctx* x, int bsz);

// statistics of Iters.f:435-457 on the resident x (FP64 operator)
// Post-solve figures of Iters.f:435-457 on the FP64 operator: x.Kx, ||x||,
// ||b - Ax|| and ||b - Ax||_M; *bnorm2 / *bnormm receive ||b|| and ||b||_M.
static int post_stats(Context* c, elastic_stats* st, double* bnorm2, double* bnormm) {
  double* out = c->red64 + 8 * kRedBlocks;
  launch_spmv(c, false, c->v64[VX], c->v64[VQ], c->epart64);
  launch_post(merge_of(c, c->epart64), c->neq, c->v64[VX], c->v64[VQ], c->v64[VB], c->A64.m,
              c->red64, out, c->stream);
  double h[6];
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  st->xkx = h[0];
  st->norm_x = std::sqrt(h[1]);
  st->norm_m_r = std::sqrt(std::fabs(h[2]));
  st->norm_r = std::sqrt(h[3]);
  *bnorm2 = std::sqrt(h[4]);
  *bnormm = std::sqrt(std::fabs(h[5]));
  return ELASTIC_OK;
}

void derive_operator(Context* c) {
  if (c->blk.d2r) launch_recip_layout(c);  // m in the layout's node order
  else launch_recip<double>(c->neq, c->A64.ad, c->A64.m, c->stream);
  if (c->layout_nb == 1) {
    launch_cast<double, float>(c->neq, c->A64.ad, c->A32.ad, c->stream);
    launch_cast<double, float>(c->nad, c->A64.al, c->A32.al, c->stream);
  } else {
    launch_block_pack(c);
  }
}

// post-solve figures and the M-norm relative residual (used by the batch)
int post_solve_stats(Context* c, elastic_stats* st) {
  double bn2, bnm;
  int rc = post_stats(c, st, &bn2, &bnm);
  if (rc) return rc;
  st->rel_res = bnm > 0 ? st->norm_m_r / bnm : 0.0;
  return ELASTIC_OK;
}

static int solve_pcg(Context* c, int precond, int precision, double tol, int maxit,
                     elastic_stats* st) {
  hipStream_t s = c->stream;
  double* dslots = c->red64 + 8 * kRedBlocks + 8;
  const bool bd = precond == ELASTIC_PRECOND_BDIAG3;
  PcgVecs V{c->v64[VX], c->v64[VR], c->v64[VP],      c->v64[VQ], c->v64[VB],
            c->A64.m,   bd ? (double*)solver_bm(c) : nullptr, c->v64[VT], c->epart64};
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev0, s));
  int rc;
  if (precision == ELASTIC_MIXED) {
    double delta = 0.1;
    if (const char* e = getenv("ELASTIC_RELIABLE_DELTA")) delta = atof(e);
    rc = pcg_run(c, V, true, tol, maxit, true, delta, dslots);
  } else {
    rc = pcg_run(c, V, false, tol, maxit, false, 0.0, dslots);
  }
  if (rc) return rc;
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev1, s));
  ELASTIC_HIP_CHECK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  const SolveScalars& h = *c->scal_host;
  st->t_device_ms = ms;
  st->iters = h.iter;
  st->outer = precision == ELASTIC_MIXED ? h.nupd : 1;
  st->op64_from = precision == ELASTIC_MIXED ? h.op64_from : -1;
  st->spmv_count = h.iter + (precision == ELASTIC_MIXED ? h.nupd : 0);
  st->conv = h.conv;
  if (!h.done) st->status = ELASTIC_MAXIT;
  double bn2, bnm;
  rc = post_stats(c, st, &bn2, &bnm);
  if (rc) return rc;
  st->rel_res = bnm > 0 ? st->norm_m_r / bnm : 0.0;
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  return ELASTIC_OK;
}

// BiCGSTAB / GMRES(k): the iterations, then the same post-solve figures.
// rel_res is ||b - Ax|| / ||b|| for BiCGSTAB (its stopping norm) and the
// M-norm ratio for GMRES.
static int solve_other(Context* c, int solver, int precond, int precision, double tol, int maxit,
                       int ngram, elastic_stats* st) {
  hipStream_t s = c->stream;
  const bool bd = precond == ELASTIC_PRECOND_BDIAG3;
  const bool mixed = precision == ELASTIC_MIXED;
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev0, s));
  KrylovResult kr;
  int rc;
  // The automatic FP32 sums are PCG's: BiCGSTAB took 3,195 iterations with
  // them against 472 without on hex8 40^3, so these methods sum in FP64
  // unless the caller set the option.
  const bool acc_saved = c->accum32;
  if (c->accum32_set < 0) c->accum32 = false;
  struct Restore {
    Context* c;
    bool v;
    ~Restore() { c->accum32 = v; }
  } restore{c, acc_saved};
  if (solver == ELASTIC_SOLVER_BICGSTAB) {
    double delta = 0.1;
    if (const char* e = getenv("ELASTIC_RELIABLE_DELTA")) delta = atof(e);
    rc = bicgstab_run(c, bd, mixed, tol, maxit, delta, &kr);
  } else if (solver == ELASTIC_SOLVER_GMRES) {
    rc = gmres_run(c, bd, mixed, tol, maxit, ngram, &kr);
  } else {
    return set_error(c, ELASTIC_BAD_ARG, "solver " + std::to_string(solver) +
                                             " not available (1 pcg, 2 gmres, 4 bicgstab)");
  }
  if (rc) return rc;
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev1, s));
  ELASTIC_HIP_CHECK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  st->t_device_ms = ms;
  st->iters = kr.iters;
  st->outer = kr.outer;
  st->op64_from = mixed ? kr.op64_from : -1;
  st->status = kr.status;
  st->conv = kr.conv;
  st->spmv_count = solver == ELASTIC_SOLVER_BICGSTAB ? 2 * kr.iters + kr.nupd : kr.iters + kr.outer;
  double bn2, bnm;
  rc = post_stats(c, st, &bn2, &bnm);
  if (rc) return rc;
  if (solver == ELASTIC_SOLVER_BICGSTAB)
    st->rel_res = bn2 > 0 ? st->norm_r / bn2 : 0.0;
  else
    st->rel_res = bnm > 0 ? st->norm_m_r / bnm : 0.0;
  if (st->status == ELASTIC_BREAKDOWN)
    return set_error(c, ELASTIC_BREAKDOWN, "zero or non-finite denominator in the Krylov recurrence");
  return ELASTIC_OK;
}

int solve_resident(elastic_ctx* x, int solver, int precond, int precision, double tol, int maxit,
                   int ngram, elastic_stats* st) {
  Context* c = &x->c;
  if (!c->values_set) return set_error(c, ELASTIC_NOT_READY, "values not set");
  if (c->zero_diag_row > 0)
    return set_error(c, ELASTIC_ZERO_DIAG,
                     "zero diagonal coefficient at equation " + std::to_string(c->zero_diag_row));
  if (tol <= 0.0) tol = 2.220446049250313e-16;  // tol 0 -> machine epsilon (Rdata.f:2056)
  if (maxit <= 0) return set_error(c, ELASTIC_BAD_ARG, "maxit must be positive");
  if (precond == ELASTIC_PRECOND_BDIAG3) {
    int rc = ensure_bdiag(x, 3);
    if (rc) return rc;
  } else if (precond != ELASTIC_PRECOND_DIAG) {
    return set_error(c, ELASTIC_BAD_ARG, "preconditioner must be 2 (diag) or 6 (3x3 blocks)");
  }
  memset(st, 0, sizeof(*st));
  // compact layouts: b into the layout order, x back when the solve ends
  // (also when it fails)
  int rc = enter_layout_order(c);
  if (rc) return rc;
  struct Leave {
    Context* c;
    ~Leave() { (void)leave_layout_order(c); }
  } leave{c};
  if (solver == ELASTIC_SOLVER_PCG) return solve_pcg(c, precond, precision, tol, maxit, st);
  return solve_other(c, solver, precond, precision, tol, maxit, ngram, st);
}

}  // namespace elastic

extern "C" {

int elastic_solve_resident(elastic_ctx* x, int32_t solver, int32_t precond, int32_t precision,
                           double tol, int32_t maxit, int32_t ngram, elastic_stats* st) {
  if (!x) return ELASTIC_BAD_ARG;
  elastic_stats local;
  if (!st) st = &local;
  hipSetDevice(x->c.device);
  int rc = solve_resident(x, solver, precond, precision, tol, maxit, ngram, st);
  if (rc) return rc;
  return st->status;
}

int elastic_solve(elastic_ctx* x, int32_t solver, int32_t precond, int32_t precision,
                  const double* b, double* xv, double tol, int32_t maxit, int32_t ngram,
                  elastic_stats* st) {
  if (!x || !b || !xv) return ELASTIC_BAD_ARG;
  auto t0 = std::chrono::steady_clock::now();
  elastic_stats local;
  if (!st) st = &local;
  Context* c = &x->c;
  if (!c->values_set) return set_error(c, ELASTIC_NOT_READY, "values not set");
  // zero diagonal check (Iters.f:337-342) on the device copy
  int rc = elastic_upload_rhs(x, b);
  if (rc) return rc;
  rc = solve_resident(x, solver, precond, precision, tol, maxit, ngram, st);
  if (rc) return rc;
  rc = elastic_download_x(x, xv);
  if (rc) return rc;
  st->t_total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return st->status;
}

}  // extern "C"

// ------------------------------------------------- drop-in for Solv.f:13
namespace {
struct DropIn {
  elastic_ctx* ctx = nullptr;
};
DropIn g_dropin;

// true when the caller's pattern is the one the context holds: the sizes and
// every entry of ip and ja (a Fortran pool may hand the same addresses to a
// different mesh of the same size)
bool same_pattern(const elastic_ctx* x, int32_t neq, int32_t nad, const int32_t* ip,
                  const int32_t* ja) {
  if (x->c.neq != neq || x->c.nad != nad || x->ia.size() != (size_t)neq + 1 ||
      x->ja.size() != (size_t)nad)
    return false;
  return std::equal(ip, ip + neq + 1, x->ia.begin()) && std::equal(ja, ja + nad, x->ja.begin());
}
}  // namespace

extern "C" int32_t elastic_solv_(int32_t* neq, int32_t* nad, int32_t* ip, int32_t* ja, double* ad,
                                 double* au, double* al, double* m, double* b, double* x,
                                 double* tol, int32_t* maxit, int32_t* ngram, int32_t* unsym,
                                 int32_t* solver, int32_t* neqf1i, int32_t* neqf2i, int32_t* neq3i,
                                 int32_t* neq4i, int32_t* neq_doti, int64_t* i_fmapi,
                                 int64_t* i_xfi, int64_t* i_rcvsi, int64_t* i_dspli) {
  (void)au;  // symmetric storage: au is al (Datastruct.f:89-90)
  (void)neqf1i, (void)neqf2i, (void)neq3i, (void)neq4i, (void)neq_doti;
  (void)i_fmapi, (void)i_xfi, (void)i_rcvsi, (void)i_dspli;
  if (!neq || !nad || !ip || !ja || !ad || !al || !m || !b || !x || !tol || !maxit || !solver)
    return ELASTIC_BAD_ARG;
  if (unsym && *unsym) {
    fprintf(stderr, "Solver %d nao disponivel para matriz nao-simetrica !\n", *solver);
    return ELASTIC_BAD_ARG;
  }
  DropIn& g = g_dropin;
  int rc;
  if (!g.ctx) {
    rc = elastic_create(&g.ctx, 0);
    if (rc) return rc;
  }
  if (!same_pattern(g.ctx, *neq, *nad, ip, ja)) {
    rc = elastic_set_structure(g.ctx, *neq, *nad, ip, ja);
    if (rc) {
      fprintf(stderr, "elastic_solv: %s\n", elastic_last_error(g.ctx));
      return rc;
    }
  }
  rc = elastic_set_values(g.ctx, ad, al);
  if (rc) return rc;
  for (int i = 0; i < *neq; ++i) m[i] = 1.0 / ad[i];  // pre_diag (Solv.f:65-69)
  int precision = ELASTIC_FP64;
  if (const char* e = getenv("ELASTIC_PRECISION")) precision = atoi(e) == 32 ? ELASTIC_MIXED : ELASTIC_FP64;
  elastic_stats st;
  rc = elastic_solve(g.ctx, *solver, ELASTIC_PRECOND_DIAG, precision, b, x, *tol, *maxit,
                     ngram ? *ngram : 50, &st);
  if (getenv("ELASTIC_VERBOSE")) {
    printf(" (PCG) solver [gfx950]:\n");
    printf("     Solver tol           = %20.6E\n", *tol);
    printf("     tol * ||b||m         = %20.6E\n", st.conv);
    printf("     Number of equations  = %20d\n", *neq);
    printf("     nad                  = %20d\n", *nad);
    printf("     Number of iterations = %20d\n", st.iters);
    printf("     x * Kx               = %20.10E\n", st.xkx);
    printf("     || x ||              = %20.10E\n", st.norm_x);
    printf("     || b - Ax ||         = %20.10E\n", st.norm_r);
    printf("     || b - Ax ||m        = %20.10E\n", st.norm_m_r);
    printf("     time (s)             = %20.4f\n", st.t_total_ms * 1e-3);
  }
  if (rc == ELASTIC_MAXIT)
    printf(" *** WARNING: No convergence reached after %9d iterations !\n", *maxit);
  else if (rc)
    fprintf(stderr, "elastic_solv: %s\n", elastic_last_error(g.ctx));
  return rc;
}
