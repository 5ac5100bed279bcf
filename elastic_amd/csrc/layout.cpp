// Row-tile layout of the symmetric CSRC product (see TileLayout in
// elastic_internal.h).  Built once per structure on the host, in parallel
// over tiles.
#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "elastic_internal.h"

namespace elastic {

void build_tiles(int neq, const int32_t* ia1, const int32_t* ja1, int tile_nnz, TileLayout& L) {
  if (tile_nnz < 64) tile_nnz = 64;
  if (tile_nnz > kTileColsMax - kTileRowsMax) tile_nnz = kTileColsMax - kTileRowsMax;
  L = TileLayout();
  L.neq = neq;
  L.nad = (int64_t)ia1[neq] - 1;
  // 1. greedy partition of consecutive rows by nonzeros and rows
  int r = 0;
  while (r < neq) {
    int r0 = r;
    int64_t nz = 0;
    while (r < neq && r - r0 < kTileRowsMax) {
      int len = ia1[r + 1] - ia1[r];
      if (len > tile_nnz)
        throw std::runtime_error("CSRC row longer than the tile size; raise tile_nnz");
      if (r > r0 && nz + len > tile_nnz) break;
      nz += len;
      ++r;
    }
    L.row0.push_back(r0);
    L.nrows.push_back(r - r0);
    L.nz0.push_back((int64_t)ia1[r0] - 1);
    L.nnz.push_back((int32_t)nz);
  }
  const int nt = L.ntiles = (int)L.row0.size();
  L.ncols.assign(nt, 0);
  L.ext0.assign(nt, 0);
  L.rp0.assign(nt, 0);
  L.seg0.assign(nt, 0);
  L.word.assign((size_t)L.nad, 0u);
  // 2. per tile: external columns, local indices, column-sorted positions
  std::vector<std::vector<int32_t>> ext(nt);
  std::vector<std::vector<uint16_t>> seg(nt);
  unsigned nthr = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::vector<std::string> errs(nthr);
  for (unsigned w = 0; w < nthr; ++w)
    pool.emplace_back([&, w]() {
      std::vector<int32_t> lcol, order;
      std::vector<int32_t> cnt;
      for (int t = (int)w; t < nt; t += (int)nthr) {
        const int r0 = L.row0[t], nr = L.nrows[t], nnz = L.nnz[t];
        const int64_t k0 = L.nz0[t];
        std::vector<int32_t>& ex = ext[t];
        ex.clear();
        for (int64_t k = k0; k < k0 + nnz; ++k) {
          int c = ja1[k] - 1;
          if (c < r0) ex.push_back(c);
        }
        std::sort(ex.begin(), ex.end());
        ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
        const int ncols = nr + (int)ex.size();
        if (ncols > kTileColsMax) {
          errs[w] = "tile local columns exceed 4096";
          return;
        }
        L.ncols[t] = ncols;
        lcol.resize(nnz);
        for (int q = 0; q < nnz; ++q) {
          int c = ja1[k0 + q] - 1;
          lcol[q] = (c >= r0) ? c - r0
                              : nr + (int)(std::lower_bound(ex.begin(), ex.end(), c) - ex.begin());
        }
        // column-sorted order (counting sort keeps row order inside a column)
        cnt.assign(ncols + 1, 0);
        for (int q = 0; q < nnz; ++q) cnt[lcol[q] + 1]++;
        for (int c = 0; c < ncols; ++c) cnt[c + 1] += cnt[c];
        seg[t].assign(cnt.begin(), cnt.end());
        order.assign(nnz, 0);
        std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
        for (int q = 0; q < nnz; ++q) order[q] = pos[lcol[q]]++;
        int lr = 0;
        for (int q = 0; q < nnz; ++q) {
          while (ia1[r0 + lr + 1] - 1 <= k0 + q) ++lr;  // row of nonzero q
          L.word[k0 + q] = ((uint32_t)lr << 24) | ((uint32_t)lcol[q] << 12) | (uint32_t)order[q];
        }
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (!e.empty()) throw std::runtime_error(e);
  // 3. concatenate per-tile streams
  int64_t next = 0, nseg = 0;
  for (int t = 0; t < nt; ++t) {
    L.ext0[t] = (int32_t)next;
    next += (int64_t)ext[t].size();
    L.seg0[t] = (int32_t)nseg;
    nseg += L.ncols[t] + 1;
    L.rp0[t] = L.row0[t] + t;  // nrows+1 entries per tile
    L.nnz_max = std::max(L.nnz_max, L.nnz[t]);
    L.ncols_max = std::max(L.ncols_max, L.ncols[t]);
  }
  if (next > INT32_MAX || nseg > INT32_MAX) throw std::runtime_error("layout too large");
  L.next_total = next;
  L.ext.resize(next);
  L.segptr.resize(nseg);
  L.rowptr.resize((size_t)neq + nt);
  for (int t = 0; t < nt; ++t) {
    std::copy(ext[t].begin(), ext[t].end(), L.ext.begin() + L.ext0[t]);
    std::copy(seg[t].begin(), seg[t].end(), L.segptr.begin() + L.seg0[t]);
    const int r0 = L.row0[t];
    for (int q = 0; q <= L.nrows[t]; ++q)
      L.rowptr[L.rp0[t] + q] = (uint16_t)(ia1[r0 + q] - ia1[r0]);
  }
  // 4. merge lists: partial (tile t, ext e) is added to row ext[ext0[t]+e];
  //    rows list their partials in tile order
  L.inc_ptr.assign((size_t)neq + 1, 0);
  for (int64_t q = 0; q < next; ++q) L.inc_ptr[L.ext[q] + 1]++;
  for (int i = 0; i < neq; ++i) L.inc_ptr[i + 1] += L.inc_ptr[i];
  L.inc_idx.resize(next);
  std::vector<int32_t> fill(L.inc_ptr.begin(), L.inc_ptr.end() - 1);
  for (int64_t q = 0; q < next; ++q) L.inc_idx[fill[L.ext[q]]++] = (int32_t)q;
}

}  // namespace elastic
