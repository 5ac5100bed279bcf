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

void build_block_tiles(int neq, int nb, const int32_t* ia1, const int32_t* ja1, int tile_blocks,
                       BlockLayout& L) {
  if (tile_blocks < 32) tile_blocks = 32;
  if (tile_blocks > kBlkTileMax) tile_blocks = kBlkTileMax;
  L = BlockLayout();
  L.nb = nb;
  L.nn = neq / nb;
  const int nn = L.nn;
  // node-row block counts
  std::vector<int32_t> nbr(nn);
  for (int I = 0; I < nn; ++I) {
    nbr[I] = (ia1[nb * I + 1] - ia1[nb * I]) / nb;
    if (nbr[I] > 64) throw std::runtime_error("node row with more than 64 lower blocks");
  }
  // greedy tiles of consecutive node rows
  int I = 0;
  int64_t b0 = 0;
  while (I < nn) {
    const int I0 = I;
    int cnt = 0;
    while (I < nn && I - I0 < kBlkRowsMax) {
      if (nbr[I] > tile_blocks) throw std::runtime_error("node row longer than the block tile");
      if (I > I0 && cnt + nbr[I] > tile_blocks) break;
      cnt += nbr[I];
      ++I;
    }
    L.row0.push_back(I0);
    L.nrows.push_back(I - I0);
    L.blk0.push_back(b0);
    L.nblk.push_back(cnt);
    b0 += cnt;
  }
  L.nblocks = b0;
  const int nt = L.ntiles = (int)L.row0.size();
  L.ncols.assign(nt, 0);
  L.ext0.assign(nt, 0);
  L.seg0.assign(nt, 0);
  L.rp0.assign(nt, 0);
  L.word.assign((size_t)b0, 0u);
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
        for (int r = I0; r < I0 + nr; ++r) {
          const int32_t* c0 = ja1 + ia1[nb * r] - 1;
          for (int k = 0; k < nbr[r]; ++k) cols.push_back((c0[k * nb] - 1) / nb);
        }
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
          for (int k = 0; k < nbr[I0 + r]; ++k, ++q) {
            L.word[L.blk0[t] + q] = ((uint32_t)r << 25) | ((uint32_t)lcol[q] << 10) | (uint32_t)order[q];
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
