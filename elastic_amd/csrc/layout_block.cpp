// Node-blocked row tiles of the symmetric CSRC product.
//
// When every node carries nb consecutive free equations (nb = 3 in 3-D, 2 in
// 2-D: meshes whose restrained nodes are restrained in all directions), the
// reference's CSRC pattern is a pattern of nb x nb node blocks: row nb*I+a
// holds, for every lower neighbour node J < I, the columns nb*J..nb*J+nb-1,
// followed by the a lower columns of node I itself (Csr.f:200-305 after the
// per-row sort).  The blocked tiles keep one 32-bit word per node block
// (instead of one per nonzero) and the nb*nb values of a block as nb*nb
// planes, so a block is read once, coalesced, and applied as A_IJ x_J to row
// node I and as A_IJ^T x_I to column node J.  The diagonal node block (ad and
// the in-node lower entries) is kept per node.  The values are gathered from
// the CSRC al/ad on the device after every assembly (pack kernel), so al stays
// the canonical storage.
//
// Block word: bits 25..31 local row node (<= 127 per tile), bits 10..24 local
// column node (<= 32767), bits 0..9 position of the block in the tile's
// column-sorted order (<= 1024 blocks per tile).
#include <algorithm>
#include <stdexcept>
#include <thread>

#include "elastic_internal.h"

namespace elastic {

int detect_node_blocks(int neq, const int32_t* ia1, const int32_t* ja1) {
  for (int nb : {3, 2}) {
    if (neq % nb) continue;
    bool ok = true;
    const int nn = neq / nb;
    for (int I = 0; I < nn && ok; ++I) {
      const int r0 = nb * I;
      const int len0 = ia1[r0 + 1] - ia1[r0];
      if (len0 % nb) {
        ok = false;
        break;
      }
      const int32_t* c0 = ja1 + ia1[r0] - 1;
      for (int k = 0; k < len0 && ok; k += nb) {
        const int c = c0[k] - 1;
        if (c % nb) ok = false;
        for (int b = 1; b < nb && ok; ++b)
          if (c0[k + b] - 1 != c + b) ok = false;
        if (c >= r0) ok = false;
      }
      for (int a = 1; a < nb && ok; ++a) {
        const int r = r0 + a;
        const int len = ia1[r + 1] - ia1[r];
        if (len != len0 + a) {
          ok = false;
          break;
        }
        const int32_t* ca = ja1 + ia1[r] - 1;
        for (int k = 0; k < len0 && ok; ++k)
          if (ca[k] != c0[k]) ok = false;
        for (int b = 0; b < a && ok; ++b)
          if (ca[len0 + b] - 1 != r0 + b) ok = false;
      }
    }
    if (ok) return nb;
  }
  return 1;
}

// Lower lists of the layout's node rows (CSR over layout nodes): the lower
// neighbours of each node in the layout order, ascending, with the CSRC
// source of each block's values.
struct LowerLists {
  std::vector<int64_t> ptr;
  std::vector<int32_t> col, src_row, src_k;
};

// The CSRC order: row I lists its CSRC lower blocks as stored.
static void csrc_lists(int nn, int nb, const int32_t* ia1, const int32_t* ja1, LowerLists& LL) {
  LL.ptr.assign((size_t)nn + 1, 0);
  for (int I = 0; I < nn; ++I) LL.ptr[I + 1] = LL.ptr[I] + (ia1[nb * I + 1] - ia1[nb * I]) / nb;
  const int64_t nbk = LL.ptr[nn];
  LL.col.resize(nbk);
  LL.src_row.resize(nbk);
  LL.src_k.resize(nbk);
  for (int I = 0; I < nn; ++I) {
    const int32_t* c0 = ja1 + ia1[nb * I] - 1;
    for (int64_t q = LL.ptr[I]; q < LL.ptr[I + 1]; ++q) {
      const int k = (int)(q - LL.ptr[I]);
      LL.col[q] = (c0[(int64_t)k * nb] - 1) / nb;
      LL.src_row[q] = I;
      LL.src_k[q] = k;
    }
  }
}

// The compact order.  Tiles are grown one after the other, breadth-first over
// the node graph from the lowest unassigned node (in the CSRC order, itself a
// bandwidth-reducing order): a node joins the current tile when its blocks
// (its neighbours already placed, in earlier tiles or in this one) still fit
// the tile.  A node's lower neighbours in the layout order are exactly the
// neighbours placed before it, so a tile's rows reference few nodes outside
// it: a breadth-first ball is compact, where a run of consecutive nodes of a
// bandwidth-reducing order is a thin strip of one level set (on hex8 60^3
// the external column nodes per row node drop from 3.15 to 1.28).  Returns
// false when a node would have more than 64 lower blocks.
static bool compact_lists(int nn, int nb, const int32_t* ia1, const int32_t* ja1,
                          int tile_blocks, LowerLists& LL, std::vector<int32_t>& d2r,
                          std::vector<int32_t>& tile_start) {
  // full node adjacency from the lower pattern (rows ascending)
  std::vector<int64_t> xadj((size_t)nn + 1, 0);
  for (int I = 0; I < nn; ++I) {
    const int32_t* c0 = ja1 + ia1[nb * I] - 1;
    const int nbr = (ia1[nb * I + 1] - ia1[nb * I]) / nb;
    xadj[I + 1] += nbr;
    for (int k = 0; k < nbr; ++k) xadj[(c0[(int64_t)k * nb] - 1) / nb + 1]++;
  }
  for (int I = 0; I < nn; ++I) xadj[I + 1] += xadj[I];
  std::vector<int32_t> adj(xadj[nn]);
  {
    // each list ascending: the lower neighbours, then the upper ones
    std::vector<int64_t> fill(xadj.begin(), xadj.end() - 1);
    for (int I = 0; I < nn; ++I) {  // lower neighbours of I: ascending, all < I
      const int32_t* c0 = ja1 + ia1[nb * I] - 1;
      const int nbr = (ia1[nb * I + 1] - ia1[nb * I]) / nb;
      for (int k = 0; k < nbr; ++k) adj[fill[I]++] = (c0[(int64_t)k * nb] - 1) / nb;
    }
    for (int I = 0; I < nn; ++I) {  // upper neighbours of J: rows I > J, ascending
      const int32_t* c0 = ja1 + ia1[nb * I] - 1;
      const int nbr = (ia1[nb * I + 1] - ia1[nb * I]) / nb;
      for (int k = 0; k < nbr; ++k) {
        const int J = (c0[(int64_t)k * nb] - 1) / nb;
        adj[fill[J]++] = I;
      }
    }
  }
  std::vector<int32_t> where((size_t)nn, -1), queued((size_t)nn, -1);
  std::vector<int32_t> queue;
  d2r.clear();
  d2r.reserve(nn);
  tile_start.clear();
  int nxt = 0, d = 0, tile = 0;
  while (true) {
    while (nxt < nn && where[nxt] >= 0) ++nxt;
    if (nxt >= nn) break;
    tile_start.push_back(d);
    int blocks = 0, rows = 0;
    queue.clear();
    size_t head = 0;
    queue.push_back(nxt);
    queued[nxt] = tile;
    while (head < queue.size()) {
      const int v = queue[head++];
      int nbk = 0;
      for (int64_t k = xadj[v]; k < xadj[v + 1]; ++k) nbk += where[adj[k]] >= 0;
      if (nbk > 64) return false;
      if (rows >= kBlkRowsMax || (rows > 0 && blocks + nbk > tile_blocks)) continue;
      if (nbk > tile_blocks) return false;
      where[v] = d++;
      d2r.push_back(v);
      ++rows;
      blocks += nbk;
      for (int64_t k = xadj[v]; k < xadj[v + 1]; ++k) {
        const int w = adj[k];
        if (where[w] < 0 && queued[w] != tile) {
          queued[w] = tile;
          queue.push_back(w);
        }
      }
    }
    ++tile;
  }
  // lower lists in the layout order, with each block's CSRC source: for a
  // neighbour w of v placed before it, the CSRC block of the pair lies in the
  // row of the larger CSRC node; it is the transpose when that row is w's
  LL.ptr.assign((size_t)nn + 1, 0);
  for (int e = 0; e < nn; ++e) {
    const int v = d2r[e];
    int cnt = 0;
    for (int64_t k = xadj[v]; k < xadj[v + 1]; ++k) cnt += where[adj[k]] < e;
    LL.ptr[e + 1] = LL.ptr[e] + cnt;
  }
  LL.col.resize(LL.ptr[nn]);
  LL.src_row.resize(LL.ptr[nn]);
  LL.src_k.resize(LL.ptr[nn]);
  std::vector<std::pair<int32_t, int32_t>> nbrs;
  for (int e = 0; e < nn; ++e) {
    const int v = d2r[e];
    nbrs.clear();
    for (int64_t k = xadj[v]; k < xadj[v + 1]; ++k)
      if (where[adj[k]] < e) nbrs.emplace_back(where[adj[k]], adj[k]);
    std::sort(nbrs.begin(), nbrs.end());
    int64_t q = LL.ptr[e];
    for (const auto& pr : nbrs) {
      const int w = pr.second;
      const int R = std::max(v, w), C = std::min(v, w);
      // position of C among R's CSRC lower blocks (ascending)
      const int32_t* c0 = ja1 + ia1[nb * R] - 1;
      const int nbr = (ia1[nb * R + 1] - ia1[nb * R]) / nb;
      int lo = 0, hi = nbr;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if ((c0[(int64_t)mid * nb] - 1) / nb < C) lo = mid + 1;
        else hi = mid;
      }
      LL.col[q] = pr.first;
      LL.src_row[q] = R;
      LL.src_k[q] = lo | (w > v ? (int32_t)0x80000000u : 0);
      ++q;
    }
  }
  return true;
}

void build_block_tiles(int neq, int nb, const int32_t* ia1, const int32_t* ja1, int tile_blocks,
                       BlockLayout& L, bool compact) {
  if (tile_blocks < 32) tile_blocks = 32;
  if (tile_blocks > kBlkTileMax) tile_blocks = kBlkTileMax;
  L = BlockLayout();
  L.nb = nb;
  L.nn = neq / nb;
  const int nn = L.nn;
  LowerLists LL;
  std::vector<int32_t> tile_start;
  L.compact = compact && compact_lists(nn, nb, ia1, ja1, tile_blocks, LL, L.d2r, tile_start);
  if (!L.compact) {
    L.d2r.clear();
    csrc_lists(nn, nb, ia1, ja1, LL);
  }
  // node-row block counts
  std::vector<int32_t> nbr(nn);
  for (int I = 0; I < nn; ++I) {
    nbr[I] = (int32_t)(LL.ptr[I + 1] - LL.ptr[I]);
    if (nbr[I] > 64) throw std::runtime_error("node row with more than 64 lower blocks");
  }
  int64_t b0 = 0;
  auto add_tile = [&](int I0, int I1) {
    int cnt = 0;
    for (int I = I0; I < I1; ++I) cnt += nbr[I];
    L.row0.push_back(I0);
    L.nrows.push_back(I1 - I0);
    L.blk0.push_back(b0);
    L.nblk.push_back(cnt);
    b0 += cnt;
  };
  if (L.compact) {  // the tiles grown with the order
    for (size_t t = 0; t < tile_start.size(); ++t)
      add_tile(tile_start[t], t + 1 < tile_start.size() ? tile_start[t + 1] : nn);
  } else {  // greedy tiles of consecutive node rows
    int I = 0;
    while (I < nn) {
      const int I0 = I;
      int cnt = 0;
      while (I < nn && I - I0 < kBlkRowsMax) {
        if (nbr[I] > tile_blocks) throw std::runtime_error("node row longer than the block tile");
        if (I > I0 && cnt + nbr[I] > tile_blocks) break;
        cnt += nbr[I];
        ++I;
      }
      add_tile(I0, I);
    }
  }
  L.nblocks = b0;
  const int nt = L.ntiles = (int)L.row0.size();
  L.ncols.assign(nt, 0);
  L.ext0.assign(nt, 0);
  L.seg0.assign(nt, 0);
  L.rp0.assign(nt, 0);
  L.word.assign((size_t)b0 + 1, 0u);  // one padding word: tiles load in bounds
  L.src_row.assign((size_t)b0, 0);
  L.src_k.assign((size_t)b0, 0);
  std::vector<std::vector<int32_t>> ext(nt);
  std::vector<std::vector<uint16_t>> seg(nt);
  unsigned nthr = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::vector<std::string> errs(nthr);
  for (unsigned w = 0; w < nthr; ++w)
    pool.emplace_back([&, w]() {
      std::vector<int32_t> cols, lcol, order, cnt, pos;
      for (int t = (int)w; t < nt; t += (int)nthr) {
        const int I0 = L.row0[t], nr = L.nrows[t], nbk = L.nblk[t];
        cols.clear();
        for (int r = I0; r < I0 + nr; ++r)
          for (int64_t q = LL.ptr[r]; q < LL.ptr[r + 1]; ++q) cols.push_back(LL.col[q]);
        std::vector<int32_t>& ex = ext[t];
        ex.clear();
        for (int J : cols)
          if (J < I0) ex.push_back(J);
        std::sort(ex.begin(), ex.end());
        ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
        const int ncols = nr + (int)ex.size();
        if (ncols > 32767) {
          errs[w] = "block tile local columns exceed 32767";
          return;
        }
        L.ncols[t] = ncols;
        lcol.resize(nbk);
        for (int q = 0; q < nbk; ++q) {
          const int J = cols[q];
          lcol[q] = J >= I0 ? J - I0
                            : nr + (int)(std::lower_bound(ex.begin(), ex.end(), J) - ex.begin());
        }
        cnt.assign(ncols + 1, 0);
        for (int q = 0; q < nbk; ++q) cnt[lcol[q] + 1]++;
        for (int c = 0; c < ncols; ++c) cnt[c + 1] += cnt[c];
        seg[t].assign(cnt.begin(), cnt.end());
        pos.assign(cnt.begin(), cnt.end() - 1);
        order.assign(nbk, 0);
        for (int q = 0; q < nbk; ++q) order[q] = pos[lcol[q]]++;
        int q = 0;
        for (int r = 0; r < nr; ++r)
          for (int64_t s = LL.ptr[I0 + r]; s < LL.ptr[I0 + r + 1]; ++s, ++q) {
            L.word[L.blk0[t] + q] = ((uint32_t)r << 25) | ((uint32_t)lcol[q] << 10) | (uint32_t)order[q];
            L.src_row[L.blk0[t] + q] = LL.src_row[s];
            L.src_k[L.blk0[t] + q] = LL.src_k[s];
          }
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (!e.empty()) throw std::runtime_error(e);
  int64_t next = 0, nseg = 0;
  for (int t = 0; t < nt; ++t) {
    L.ext0[t] = (int32_t)next;
    next += (int64_t)ext[t].size();
    L.seg0[t] = (int32_t)nseg;
    nseg += L.ncols[t] + 1;
    L.rp0[t] = L.row0[t] + t;
    L.nblk_max = std::max(L.nblk_max, L.nblk[t]);
    L.ncols_max = std::max(L.ncols_max, L.ncols[t]);
  }
  if (next > INT32_MAX / 3 || nseg > INT32_MAX) throw std::runtime_error("block layout too large");
  L.next_nodes = next;
  L.ext.resize(next);
  L.segptr.resize(nseg);
  L.rowptr.resize((size_t)nn + nt);
  for (int t = 0; t < nt; ++t) {
    std::copy(ext[t].begin(), ext[t].end(), L.ext.begin() + L.ext0[t]);
    std::copy(seg[t].begin(), seg[t].end(), L.segptr.begin() + L.seg0[t]);
    int acc = 0;
    for (int r = 0; r <= L.nrows[t]; ++r) {
      L.rowptr[L.rp0[t] + r] = (uint16_t)acc;
      if (r < L.nrows[t]) acc += nbr[L.row0[t] + r];
    }
  }
  // merge lists at node granularity: node J receives the partials q (tile
  // order) whose external node is J, listed at inc_ptr[J] .. inc_ptr[J+1]-1
  L.inc_ptr.assign((size_t)nn + 1, 0);
  for (int64_t q = 0; q < next; ++q) L.inc_ptr[(size_t)L.ext[q] + 1]++;
  for (int J = 0; J < nn; ++J) L.inc_ptr[J + 1] += L.inc_ptr[J];
  L.inc_idx.resize(next);
  std::vector<int32_t> fill(L.inc_ptr.begin(), L.inc_ptr.end() - 1);
  for (int64_t q = 0; q < next; ++q) L.inc_idx[fill[L.ext[q]]++] = (int32_t)q;
}

}  // namespace elastic
