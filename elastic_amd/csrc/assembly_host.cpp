// Host set-up of the GPU assembly: element colouring and CSRC slot maps.
//
// Colouring follows Colormsh.f:160-247 (greedy, node driven: for colour c,
// every node gives colour c to the last uncoloured element around it that has
// no colour-c neighbour, and blocks that element's neighbours for colour c),
// and setcolor (Colormsh.f:255-308) lists the elements of each colour in
// ascending order.  Elements of one colour share no node, so their
// contributions to al/ad/b never meet and are added without atomics; every
// CSRC entry receives its element contributions in colour order, which is the
// order of the reference's OpenMP assembly (Pform.f:108-184).
//
// Slot map: for every element and every unordered pair of its local dofs
// {i >= j} (nst(nst+1)/2 pairs) one int32 target:
//   -1                 no equation (restrained dof)
//   <= -2              diagonal: ad index = -t - 2
//   >= 0               al index (0 .. nad-1; nad < 2^31, as the reference's
//                      int32 ja allows)
// (al is filled with s(row, col) for the local dof whose equation is the row,
// as Assbly.f:116-143 does: s(i,j) when eq(i) > eq(j), else s(j,i); the
// kernels read the equation numbers, so the slot carries no orientation).
#include <algorithm>
#include <stdexcept>
#include <thread>

#include "elastic_internal.h"

namespace elastic {

int color_elements(const int32_t* ix, int nnode, int numel, int nen, int nenv,
                   std::vector<int32_t>& icolor) {
  const int ld = nen + 1;
  std::vector<int32_t> nincid(nnode + 1, 0);
  for (int e = 0; e < numel; ++e)
    for (int j = 0; j < nenv; ++j) {
      int node = ix[(size_t)e * ld + j];
      if (node > 0) nincid[node]++;
    }
  std::vector<int64_t> start(nnode + 2, 0);
  for (int n = 1; n <= nnode; ++n) start[n + 1] = start[n] + nincid[n];
  std::vector<int32_t> incid(start[nnode + 1]);
  std::vector<int32_t> fill(nnode + 1, 0);
  for (int e = 0; e < numel; ++e)
    for (int j = 0; j < nenv; ++j) {
      int node = ix[(size_t)e * ld + j];
      if (node > 0) incid[start[node] + fill[node]++] = e + 1;
    }
  icolor.assign(numel + 1, 0);  // 1-based element numbers
  // The reference stops at 200 colours (maxcolors, Colormsh.f:176); the same
  // greedy rule continues here, so meshes with more than 200 elements around
  // one node still assemble (every colour takes at least one element, so at
  // most numel colours are needed).  The first 200 colours are the reference's.
  for (int c = 1; c <= numel; ++c) {
    for (int node = 1; node <= nnode; ++node) {
      bool found = false;
      int first = 0;
      for (int64_t q = start[node]; q < start[node + 1]; ++q) {
        int ik = incid[q];
        if (icolor[ik] == c) found = true;
        if (icolor[ik] == 0) first = ik;
      }
      if (found || first == 0) continue;
      icolor[first] = c;
      for (int j = 0; j < nenv; ++j) {
        int ixj = ix[(size_t)(first - 1) * ld + j];
        if (ixj <= 0) continue;
        for (int64_t q = start[ixj]; q < start[ixj + 1]; ++q)
          if (icolor[incid[q]] == 0) icolor[incid[q]] = -1;
      }
    }
    bool quit = true;
    for (int e = 1; e <= numel; ++e)
      if (icolor[e] == -1) {
        icolor[e] = 0;
        quit = false;
      }
    if (quit) return c;
  }
  return numel > 0 ? -1 : 0;  // not reached for numel > 0
}

// slots: numel * npair targets, pairs enumerated i = 0..nst-1, j = 0..i.
void element_slots(const int32_t* ix, const int32_t* id, int numel, int nen, int ndf, int neq,
                   const int32_t* ia1, const int32_t* ja1, std::vector<int32_t>& slots) {
  const int nst = nen * ndf, npair = nst * (nst + 1) / 2, ldx = nen + 1;
  slots.assign((size_t)numel * npair, -1);
  unsigned nthr = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned w = 0; w < nthr; ++w)
    pool.emplace_back([&, w]() {
      std::vector<int32_t> ldv(nst);
      for (int e = (int)w; e < numel; e += (int)nthr) {
        for (int a = 0; a < nen; ++a) {
          int no = ix[(size_t)e * ldx + a];
          for (int d = 0; d < ndf; ++d) ldv[a * ndf + d] = id[(size_t)(no - 1) * ndf + d];
        }
        int32_t* out = slots.data() + (size_t)e * npair;
        int p = 0;
        for (int i = 0; i < nst; ++i)
          for (int j = 0; j <= i; ++j, ++p) {
            int ki = ldv[i], kj = ldv[j];
            if (ki <= 0 || kj <= 0 || ki > neq || kj > neq) continue;
            if (ki == kj) {
              out[p] = -(ki - 1) - 2;
              continue;
            }
            int row = std::max(ki, kj), col = std::min(ki, kj);
            const int32_t* b = ja1 + ia1[row - 1] - 1;
            const int32_t* en = ja1 + ia1[row] - 1;
            const int32_t* f = std::lower_bound(b, en, col);
            if (f == en || *f != col) continue;  // not in the pattern (cannot happen)
            const int64_t slot = (int64_t)(f - ja1);
            if (slot > INT32_MAX) throw std::runtime_error("al too large for slot maps");
            out[p] = (int32_t)slot;
          }
      }
    });
  for (auto& t : pool) t.join();
}

}  // namespace elastic
