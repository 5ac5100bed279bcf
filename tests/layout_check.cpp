// CPU check of the node-blocked tile layout (elastic_amd/csrc/layout_block.cpp):
// builds the layout of a CSRC pattern, applies it as the GPU product does
// (block words, external columns, diagonal blocks, values gathered from the
// CSRC arrays through each block's source, transposed where flagged) in FP64
// on the host, and compares y with the plain symmetric CSRC product y =
// (D + L + L^T) x (Matvec.f:528-621).  Also checks the layout's invariants.
// Input (binary): neq, nad (int32), compact, tile_blocks (int32), ia(neq+1),
// ja(nad) (int32, 1-based), ad(neq), al(nad), x(neq) (double).
// Output (stdout): "nb ntiles next_nodes compact max_rel_err" or an error.
#include <cmath>
#include <cstdio>
#include <vector>

#include "elastic_internal.h"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[4];
  if (fread(hdr, 4, 4, f) != 4) return 2;
  const int neq = hdr[0], nad = hdr[1], compact = hdr[2], tile_blocks = hdr[3];
  std::vector<int32_t> ia(neq + 1), ja(nad);
  std::vector<double> ad(neq), al(nad), x(neq);
  if (fread(ia.data(), 4, neq + 1, f) != (size_t)neq + 1 || fread(ja.data(), 4, nad, f) != (size_t)nad ||
      fread(ad.data(), 8, neq, f) != (size_t)neq || fread(al.data(), 8, nad, f) != (size_t)nad ||
      fread(x.data(), 8, neq, f) != (size_t)neq)
    return 2;
  fclose(f);
  // reference: the CSRC product
  std::vector<double> yr(neq);
  for (int i = 0; i < neq; ++i) yr[i] = ad[i] * x[i];
  for (int i = 0; i < neq; ++i)
    for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
      const int j = ja[k] - 1;
      yr[i] += al[k] * x[j];
      yr[j] += al[k] * x[i];
    }
  const int nb = elastic::detect_node_blocks(neq, ia.data(), ja.data());
  if (nb == 1) {
    printf("not node-blocked\n");
    return 3;
  }
  elastic::BlockLayout L;
  elastic::build_block_tiles(neq, nb, ia.data(), ja.data(), tile_blocks, L, compact != 0);
  const int nn = L.nn;
  auto d2r = [&](int d) { return L.compact ? L.d2r[d] : d; };
  // d2r is a permutation
  std::vector<int> seen(nn, 0);
  for (int d = 0; d < nn; ++d)
    if (d2r(d) < 0 || d2r(d) >= nn || seen[d2r(d)]++) {
      printf("d2r is not a permutation\n");
      return 4;
    }
  // every CSRC block used once: count per (row node, k)
  std::vector<int64_t> rowoff(nn + 1, 0);
  for (int I = 0; I < nn; ++I) rowoff[I + 1] = rowoff[I] + (ia[nb * I + 1] - ia[nb * I]) / nb;
  std::vector<int> used(rowoff[nn], 0);
  std::vector<double> y(neq, 0.0);  // layout order
  for (int t = 0; t < L.ntiles; ++t) {
    const int I0 = L.row0[t], nr = L.nrows[t], nbk = L.nblk[t];
    if (nr > elastic::kBlkRowsMax || nbk > std::max(tile_blocks, 64)) {
      printf("tile %d too large\n", t);
      return 4;
    }
    auto gnode = [&](int lc) { return lc < nr ? I0 + lc : L.ext[L.ext0[t] + lc - nr]; };
    for (int q = 0; q < nbk; ++q) {
      const int64_t g = L.blk0[t] + q;
      const uint32_t w = L.word[g];
      const int lr = (int)(w >> 25), lc = (int)((w >> 10) & 0x7FFF);
      const int I = I0 + lr, J = gnode(lc);
      if (J >= I) {
        printf("block not lower in the layout order\n");
        return 4;
      }
      const int R = L.src_row[g], kk = L.src_k[g] & 0x7FFFFFFF;
      const bool tr = L.src_k[g] < 0;
      if (R < 0 || R >= nn || kk >= rowoff[R + 1] - rowoff[R]) {
        printf("bad block source\n");
        return 4;
      }
      used[rowoff[R] + kk]++;
      // the source pair must be the layout pair
      const int C = (ja[ia[nb * R] - 1 + nb * kk] - 1) / nb;
      const int a = tr ? C : R, b = tr ? R : C;  // CSRC nodes of layout row / column
      if (a != d2r(I) || b != d2r(J)) {
        printf("block source is not the layout pair\n");
        return 4;
      }
      double A[9];
      for (int r = 0; r < nb; ++r)
        for (int cc = 0; cc < nb; ++cc) {
          const double v = al[ia[nb * R + r] - 1 + nb * kk + cc];
          if (tr) A[cc * nb + r] = v;
          else A[r * nb + cc] = v;
        }
      // y_I += A x_J ; y_J += A^T x_I (layout order vectors)
      for (int r = 0; r < nb; ++r)
        for (int cc = 0; cc < nb; ++cc) {
          y[nb * I + r] += A[r * nb + cc] * x[nb * d2r(J) + cc];
          y[nb * J + cc] += A[r * nb + cc] * x[nb * d2r(I) + r];
        }
    }
  }
  for (size_t k = 0; k < used.size(); ++k)
    if (used[k] != 1) {
      printf("CSRC block used %d times\n", used[k]);
      return 4;
    }
  // diagonal node blocks, then back to the CSRC order
  double err = 0.0, nrm = 0.0;
  for (int d = 0; d < nn; ++d) {
    const int I = d2r(d);
    const int nbr = (ia[nb * I + 1] - ia[nb * I]) / nb;
    for (int r = 0; r < nb; ++r) {
      double s = y[nb * d + r] + ad[nb * I + r] * x[nb * I + r];
      for (int cc = 0; cc < nb; ++cc) {
        if (cc == r) continue;
        const int hi = std::max(r, cc), lo = std::min(r, cc);
        const double v = al[ia[nb * I + hi] - 1 + nb * nbr + lo];
        s += v * x[nb * I + cc];
      }
      err = std::max(err, std::fabs(s - yr[nb * I + r]));
      nrm = std::max(nrm, std::fabs(yr[nb * I + r]));
    }
  }
  printf("%d %d %lld %d %.3e\n", nb, L.ntiles, (long long)L.next_nodes, L.compact ? 1 : 0,
         nrm > 0 ? err / nrm : err);
  return 0;
}
