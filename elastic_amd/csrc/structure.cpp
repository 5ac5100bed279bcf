// Host-side set-up of the equation system, as the reference driver does it
// once per mesh (Main.f:417-494): node graph, reverse Cuthill-McKee order,
// equation numbering and the CSRC (strictly lower) pattern.
//
// The outputs are bit-identical to the reference's (RCM tie-breaking included):
//   graph    Graph.f:20-257   sorted, de-duplicated node adjacency
//   reord    Reord.f:1-434    SPARSPAK genrcm (fnroot/rcm/rootls/degree)
//   numeq    Numeq.f:139-191  equations node by node in RCM order
//   csr      Csr.f:13-305     csria/csrja + per-row sort, lower only
// Index values are 1-based at the C-ABI, like the Fortran arrays they replace.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "elastic_internal.h"

namespace elastic {

// Node adjacency (0-based storage, 1-based node values): xadj[n], n = 0..nnode.
void node_graph(int nnode, int numel, int nen, const int32_t* ix, std::vector<int32_t>& xadj,
                std::vector<int32_t>& adj) {
  const int ld = nen + 1;
  std::vector<int64_t> cnt(nnode + 1, 0);
  for (int e = 0; e < numel; ++e)
    for (int j = 0; j < nen; ++j) {
      int no = ix[(size_t)e * ld + j];
      if (no <= 0 || no > nnode) continue;
      cnt[no] += nen - 1;
    }
  std::vector<int64_t> start(nnode + 2, 0);
  for (int n = 1; n <= nnode; ++n) start[n + 1] = start[n] + cnt[n];
  std::vector<int32_t> raw(start[nnode + 1]);
  std::vector<int64_t> fill(nnode + 1, 0);
  for (int e = 0; e < numel; ++e)
    for (int j = 0; j < nen; ++j) {
      int a = ix[(size_t)e * ld + j];
      if (a <= 0 || a > nnode) continue;
      for (int k = 0; k < nen; ++k) {
        int b = ix[(size_t)e * ld + k];
        if (b <= 0 || b > nnode || b == a) continue;
        raw[start[a] + fill[a]++] = b;
      }
    }
  xadj.assign(nnode + 1, 0);
  adj.clear();
  adj.reserve(raw.size() / 2 + 1);
  for (int n = 1; n <= nnode; ++n) {
    auto b = raw.begin() + start[n], e = b + fill[n];
    std::sort(b, e);
    auto last = std::unique(b, e);
    adj.insert(adj.end(), b, last);
    xadj[n] = (int32_t)adj.size();
  }
}

namespace {

// SPARSPAK routines on 1-based arrays (xadj is 1-based: xadj[1..n+1]).
struct Rcm {
  std::vector<int32_t>& xadj;  // sized n+2, 1-based, values 1-based into adjncy
  const std::vector<int32_t>& adjncy;  // 1-based via adjncy[k-1]
  std::vector<int32_t> mask, xls, deg;

  int A(int k) const { return adjncy[k - 1]; }

  // Reord.f:278-354
  void rootls(int root, int* nlvl, int32_t* ls) {
    mask[root] = 0;
    ls[1] = root;
    *nlvl = 0;
    int lvlend = 0, ccsize = 1;
    for (;;) {
      int lbegin = lvlend + 1;
      lvlend = ccsize;
      xls[++*nlvl] = lbegin;
      for (int i = lbegin; i <= lvlend; ++i) {
        int node = ls[i];
        for (int j = xadj[node]; j < xadj[node + 1]; ++j) {
          int nbr = A(j);
          if (!mask[nbr]) continue;
          ls[++ccsize] = nbr;
          mask[nbr] = 0;
        }
      }
      if (ccsize == lvlend) break;
    }
    xls[*nlvl + 1] = lvlend + 1;
    for (int i = 1; i <= ccsize; ++i) mask[ls[i]] = 1;
  }

  // Reord.f:355-434 (xadj sign used as the visited mark)
  int degree(int root, int32_t* ls) {
    ls[1] = root;
    xadj[root] = -xadj[root];
    int lvlend = 0, ccsize = 1;
    for (;;) {
      int lbegin = lvlend + 1;
      lvlend = ccsize;
      for (int i = lbegin; i <= lvlend; ++i) {
        int node = ls[i];
        int jstrt = -xadj[node], jstop = std::abs(xadj[node + 1]) - 1;
        int ideg = 0;
        for (int j = jstrt; j <= jstop; ++j) {
          int nbr = A(j);
          if (!mask[nbr]) continue;
          ++ideg;
          if (xadj[nbr] < 0) continue;
          xadj[nbr] = -xadj[nbr];
          ls[++ccsize] = nbr;
        }
        deg[node] = ideg;
      }
      if (ccsize == lvlend) break;
    }
    for (int i = 1; i <= ccsize; ++i) xadj[ls[i]] = -xadj[ls[i]];
    return ccsize;
  }

  // Reord.f:101-169
  void fnroot(int* root, int* nlvl, int32_t* ls) {
    rootls(*root, nlvl, ls);
    int ccsize = xls[*nlvl + 1] - 1;
    if (*nlvl == 1 || *nlvl == ccsize) return;
    for (;;) {
      int jstrt = xls[*nlvl];
      int mindeg = ccsize;
      *root = ls[jstrt];
      if (ccsize != jstrt) {
        for (int j = jstrt; j <= ccsize; ++j) {
          int node = ls[j], ndeg = 0;
          for (int k = xadj[node]; k < xadj[node + 1]; ++k)
            if (mask[A(k)] > 0) ++ndeg;
          if (ndeg < mindeg) {
            *root = node;
            mindeg = ndeg;
          }
        }
      }
      int nunlvl;
      rootls(*root, &nunlvl, ls);
      if (nunlvl <= *nlvl) return;
      *nlvl = nunlvl;
      if (*nlvl >= ccsize) return;
    }
  }

  // Reord.f:170-277
  int rcm(int root, int32_t* perm) {
    int ccsize = degree(root, perm);
    mask[root] = 0;
    if (ccsize <= 1) return ccsize;
    int lvlend = 0, lnbr = 1;
    do {
      int lbegin = lvlend + 1;
      lvlend = lnbr;
      for (int i = lbegin; i <= lvlend; ++i) {
        int node = perm[i];
        int fnbr = lnbr + 1;
        for (int j = xadj[node]; j < xadj[node + 1]; ++j) {
          int nbr = A(j);
          if (!mask[nbr]) continue;
          mask[nbr] = 0;
          perm[++lnbr] = nbr;
        }
        if (fnbr >= lnbr) continue;
        // stable insertion sort of the new neighbours by degree
        for (int k = fnbr + 1; k <= lnbr; ++k) {
          int nbr = perm[k], l = k - 1;
          while (l >= fnbr && deg[perm[l]] > deg[nbr]) {
            perm[l + 1] = perm[l];
            --l;
          }
          perm[l + 1] = nbr;
        }
      }
    } while (lnbr > lvlend);
    std::reverse(perm + 1, perm + ccsize + 1);
    return ccsize;
  }
};

}  // namespace

// perm[k-1] = old node at new position k (Reord.f:1-44 with afl = true)
void rcm_order(int nnode, int numel, int nen, const int32_t* ix, int32_t* perm) {
  std::vector<int32_t> xadj0, adj;
  node_graph(nnode, numel, nen, ix, xadj0, adj);
  std::vector<int32_t> xadj(nnode + 2);
  for (int n = 1; n <= nnode + 1; ++n) xadj[n] = xadj0[n - 1] + 1;
  Rcm r{xadj, adj, std::vector<int32_t>(nnode + 2, 1), std::vector<int32_t>(nnode + 2, 0),
        std::vector<int32_t>(nnode + 2, 0)};
  std::vector<int32_t> p(nnode + 2, 0);  // 1-based
  int num = 1;
  for (int i = 1; i <= nnode; ++i) {
    if (!r.mask[i]) continue;
    int root = i, nlvl;
    r.fnroot(&root, &nlvl, p.data() + num - 1);
    num += r.rcm(root, p.data() + num - 1);
    if (num > nnode) break;
  }
  for (int k = 0; k < nnode; ++k) perm[k] = p[k + 1];
}

int number_equations(int32_t* id, const int32_t* perm, int nnode, int ndf) {
  int neq = 0;
  for (int i = 0; i < nnode; ++i) {
    int32_t* row = id + (size_t)(perm[i] - 1) * ndf;
    for (int j = 0; j < ndf; ++j) row[j] = (row[j] == 0) ? ++neq : 0;
  }
  return neq;
}

// CSRC strictly-lower pattern: ia (neq+1, 1-based values), ja (1-based).
void csrc_pattern(const int32_t* id, const int32_t* ix, int nnode, int numel, int nen, int ndf,
                  int neq, std::vector<int32_t>& ia, std::vector<int32_t>& ja) {
  std::vector<int32_t> xadj, adj;
  node_graph(nnode, numel, nen, ix, xadj, adj);
  ia.assign(neq + 1, 0);
  // Row lengths (Csr.f:92-199): every lower-numbered dof of the node itself
  // and of its graph neighbours.
  for (int n = 1; n <= nnode; ++n) {
    const int32_t* dn = id + (size_t)(n - 1) * ndf;
    for (int ii = 0; ii < ndf; ++ii) {
      int ei = dn[ii];
      if (ei <= 0 || ei > neq) continue;
      int c = 0;
      for (int kk = 0; kk < ndf; ++kk)
        if (dn[kk] > 0 && dn[kk] < ei) ++c;
      for (int k = xadj[n - 1]; k < xadj[n]; ++k) {
        const int32_t* dj = id + (size_t)(adj[k] - 1) * ndf;
        for (int jj = 0; jj < ndf; ++jj)
          if (dj[jj] > 0 && dj[jj] < ei) ++c;
      }
      ia[ei] = c;
    }
  }
  ia[0] = 1;
  for (int i = 1; i <= neq; ++i) ia[i] += ia[i - 1];
  ja.assign((size_t)ia[neq] - 1, 0);
  for (int n = 1; n <= nnode; ++n) {
    const int32_t* dn = id + (size_t)(n - 1) * ndf;
    for (int ii = 0; ii < ndf; ++ii) {
      int ei = dn[ii];
      if (ei <= 0 || ei > neq) continue;
      size_t w = (size_t)ia[ei - 1] - 1, b = w;
      for (int kk = 0; kk < ndf; ++kk)
        if (dn[kk] > 0 && dn[kk] < ei) ja[w++] = dn[kk];
      for (int k = xadj[n - 1]; k < xadj[n]; ++k) {
        const int32_t* dj = id + (size_t)(adj[k] - 1) * ndf;
        for (int jj = 0; jj < ndf; ++jj)
          if (dj[jj] > 0 && dj[jj] < ei) ja[w++] = dj[jj];
      }
      std::sort(ja.begin() + b, ja.begin() + w);  // sortgraph (Csr.f:81)
    }
  }
}

}  // namespace elastic
