/*
 * oracle.c -- CPU restatement of the Elastic solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP product never links or calls this file;
 * tests/, the smoke check and the CPU-baseline leg of bench.py use it as the
 * checker.  It is pinned against the reference itself (fixtures written by the
 * Fortran reference through oracle/ref_harness.f, see tests/golden/) and against
 * closed forms (tests/test_oracle_*.py).
 *
 * Arithmetic follows the reference statement by statement (same operand order
 * and parenthesisation, no contraction: built with -ffp-contract=off), so the
 * serial paths reproduce the reference's FP64 results bit for bit.  Index
 * values are 1-based like the Fortran arrays; the helper macros below give
 * 1-based access to 0-based C storage.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define A1(a, i) ((a)[(i)-1])                      /* a(i) */
#define A2(a, ld, i, j) ((a)[((j)-1) * (ld) + (i)-1]) /* a(i,j), leading dim ld */

/* Gauss-Legendre 2-point abscissa and weight as tabulated in Elmlib.f:2985-3016
 * (pg(1:2,2) = -/+0.577350269189626, wg(1:2,2) = 1). */
static const double PG2[2] = {-.577350269189626, .577350269189626};
static const double WG2[2] = {1.0, 1.0};
/* Elmlib02.f:1277-1279 */
static const double DIV6 = 0.166666666666667;
static const double DIV144 = 0.364444444444444e-02;
static const double JZERO = 1.0e-14; /* Elmlib.f:1992, 2661 */

/* ------------------------------------------------------------------ graph */
/* Graph.f:20-257.  aponta/graph0 collect the other nodes of every element of a
 * node; graph1 sorts (bubblesort) and drops repeats.  The result is each node's
 * neighbour set in ascending order, which is what is built here directly. */
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int or_graph(int nnode, int numel, int nen, const int32_t* ix, int32_t* ip, int32_t** ips) {
  /* aponta (Graph.f:72-108): upper bound of neighbours per node */
  int64_t* cnt = (int64_t*)calloc((size_t)nnode + 2, sizeof(int64_t));
  for (int e = 1; e <= numel; e++)
    for (int j = 1; j <= nen; j++) {
      int no = A2(ix, nen + 1, j, e);
      if (no == 0 || no > nnode) continue;
      cnt[no + 1] += nen - 1;
    }
  cnt[1] = 1;
  for (int i = 2; i <= nnode + 1; i++) cnt[i] += cnt[i - 1];
  int64_t ipos = cnt[nnode + 1] - 1;
  int32_t* nov = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ipos > 0 ? ipos : 1));
  int32_t* fill = (int32_t*)calloc((size_t)nnode + 1, sizeof(int32_t));
  /* graph0 (Graph.f:109-149) */
  for (int e = 1; e <= numel; e++)
    for (int j = 1; j <= nen; j++) {
      int noj = A2(ix, nen + 1, j, e);
      if (noj == 0 || noj > nnode) continue;
      for (int k = 1; k <= nen; k++) {
        int nok = A2(ix, nen + 1, k, e);
        if (nok == 0 || nok > nnode || nok == noj) continue;
        nov[cnt[noj] - 1 + fill[noj]] = nok;
        fill[noj]++;
      }
    }
  /* graph1 (Graph.f:150-208): sort, unique, compact */
  int64_t w = 0;
  A1(ip, 1) = 1;
  for (int i = 1; i <= nnode; i++) {
    int32_t* seg = nov + cnt[i] - 1;
    int n = fill[i];
    qsort(seg, (size_t)n, sizeof(int32_t), cmp_i32);
    for (int q = 0; q < n; q++) {
      if (q + 1 < n && seg[q + 1] == seg[q]) continue;
      nov[w++] = seg[q];
    }
    A1(ip, i + 1) = (int32_t)(w + 1);
  }
  free(cnt);
  free(fill);
  *ips = nov;
  return (int)w;
}

/* ------------------------------------------------------------------- RCM */
/* SPARSPAK family as restated from Reord.f:278-354 (rootls),
 * 355-434 (degree), 101-169 (fnroot), 170-277 (rcm), 45-100 (genrcm).
 * All arrays 1-based through the A1 macro. */
static void rootls(int root, int32_t* xadj, const int32_t* adjncy, int32_t* mask, int* nlvl,
                   int32_t* xls, int32_t* ls) {
  A1(mask, root) = 0;
  A1(ls, 1) = root;
  *nlvl = 0;
  int lvlend = 0, ccsize = 1;
  for (;;) {
    int lbegin = lvlend + 1;
    lvlend = ccsize;
    (*nlvl)++;
    A1(xls, *nlvl) = lbegin;
    for (int i = lbegin; i <= lvlend; i++) {
      int node = A1(ls, i);
      int jstrt = A1(xadj, node), jstop = A1(xadj, node + 1) - 1;
      for (int j = jstrt; j <= jstop; j++) {
        int nbr = A1(adjncy, j);
        if (A1(mask, nbr) == 0) continue;
        ccsize++;
        A1(ls, ccsize) = nbr;
        A1(mask, nbr) = 0;
      }
    }
    if (ccsize - lvlend <= 0) break;
  }
  A1(xls, *nlvl + 1) = lvlend + 1;
  for (int i = 1; i <= ccsize; i++) A1(mask, A1(ls, i)) = 1;
}

static void degree(int root, int32_t* xadj, const int32_t* adjncy, const int32_t* mask,
                   int32_t* deg, int* ccsize_out, int32_t* ls) {
  A1(ls, 1) = root;
  A1(xadj, root) = -A1(xadj, root);
  int lvlend = 0, ccsize = 1;
  for (;;) {
    int lbegin = lvlend + 1;
    lvlend = ccsize;
    for (int i = lbegin; i <= lvlend; i++) {
      int node = A1(ls, i);
      int jstrt = -A1(xadj, node);
      int jstop = abs(A1(xadj, node + 1)) - 1;
      int ideg = 0;
      for (int j = jstrt; j <= jstop; j++) {
        int nbr = A1(adjncy, j);
        if (A1(mask, nbr) == 0) continue;
        ideg++;
        if (A1(xadj, nbr) < 0) continue;
        A1(xadj, nbr) = -A1(xadj, nbr);
        ccsize++;
        A1(ls, ccsize) = nbr;
      }
      A1(deg, node) = ideg;
    }
    if (ccsize - lvlend <= 0) break;
  }
  for (int i = 1; i <= ccsize; i++) {
    int node = A1(ls, i);
    A1(xadj, node) = -A1(xadj, node);
  }
  *ccsize_out = ccsize;
}

static void fnroot(int* root, int32_t* xadj, const int32_t* adjncy, int32_t* mask, int* nlvl,
                   int32_t* xls, int32_t* ls) {
  rootls(*root, xadj, adjncy, mask, nlvl, xls, ls);
  int ccsize = A1(xls, *nlvl + 1) - 1;
  if (*nlvl == 1 || *nlvl == ccsize) return;
  for (;;) {
    int jstrt = A1(xls, *nlvl);
    int mindeg = ccsize;
    *root = A1(ls, jstrt);
    if (ccsize != jstrt) {
      for (int j = jstrt; j <= ccsize; j++) {
        int node = A1(ls, j);
        int ndeg = 0;
        for (int k = A1(xadj, node); k <= A1(xadj, node + 1) - 1; k++)
          if (A1(mask, A1(adjncy, k)) > 0) ndeg++;
        if (ndeg >= mindeg) continue;
        *root = node;
        mindeg = ndeg;
      }
    }
    int nunlvl;
    rootls(*root, xadj, adjncy, mask, &nunlvl, xls, ls);
    if (nunlvl <= *nlvl) return;
    *nlvl = nunlvl;
    if (*nlvl >= ccsize) return;
  }
}

static void rcm(int root, int32_t* xadj, const int32_t* adjncy, int32_t* mask, int32_t* perm,
                int* ccsize_out, int32_t* deg) {
  int ccsize;
  degree(root, xadj, adjncy, mask, deg, &ccsize, perm);
  A1(mask, root) = 0;
  *ccsize_out = ccsize;
  if (ccsize <= 1) return;
  int lvlend = 0, lnbr = 1;
  do {
    int lbegin = lvlend + 1;
    lvlend = lnbr;
    for (int i = lbegin; i <= lvlend; i++) {
      int node = A1(perm, i);
      int jstrt = A1(xadj, node), jstop = A1(xadj, node + 1) - 1;
      int fnbr = lnbr + 1;
      for (int j = jstrt; j <= jstop; j++) {
        int nbr = A1(adjncy, j);
        if (A1(mask, nbr) == 0) continue;
        lnbr++;
        A1(mask, nbr) = 0;
        A1(perm, lnbr) = nbr;
      }
      if (fnbr >= lnbr) continue;
      /* stable linear insertion by increasing degree (Reord.f:249-262) */
      int k = fnbr;
      do {
        int l = k;
        k++;
        int nbr = A1(perm, k);
        while (l >= fnbr) {
          int lperm = A1(perm, l);
          if (A1(deg, lperm) <= A1(deg, nbr)) break;
          A1(perm, l + 1) = lperm;
          l--;
        }
        A1(perm, l + 1) = nbr;
      } while (k < lnbr);
    }
  } while (lnbr > lvlend);
  /* reverse (Reord.f:267-276) */
  int k = ccsize / 2, l = ccsize;
  for (int i = 1; i <= k; i++) {
    int lperm = A1(perm, l);
    A1(perm, l) = A1(perm, i);
    A1(perm, i) = lperm;
    l--;
  }
}

static void genrcm(int neq, int32_t* xadj, const int32_t* adjncy, int32_t* perm, int32_t* mask,
                   int32_t* xls) {
  for (int i = 1; i <= neq; i++) A1(mask, i) = 1;
  int num = 1;
  for (int i = 1; i <= neq; i++) {
    if (A1(mask, i) == 0) continue;
    int root = i, nlvl, ccsize;
    fnroot(&root, xadj, adjncy, mask, &nlvl, xls, perm + (num - 1));
    rcm(root, xadj, adjncy, mask, perm + (num - 1), &ccsize, xls);
    num += ccsize;
    if (num > neq) return;
  }
}

void or_reord(const int32_t* ix, int32_t* perm, int nnode, int numel, int nen) {
  int32_t* ip = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nnode + 1));
  int32_t* ips = NULL;
  or_graph(nnode, numel, nen, ix, ip, &ips);
  int32_t* mask = (int32_t*)calloc((size_t)nnode + 1, sizeof(int32_t));
  int32_t* xls = (int32_t*)calloc((size_t)nnode + 2, sizeof(int32_t));
  memset(perm, 0, sizeof(int32_t) * (size_t)nnode);
  genrcm(nnode, ip, ips, perm, mask, xls);
  free(mask);
  free(xls);
  free(ips);
  free(ip);
}

/* ------------------------------------------------------------------ numeq */
int or_numeq(int32_t* id, const int32_t* perm, int nnode, int ndf) {
  int neq = 0;
  for (int i = 1; i <= nnode; i++) {
    int n = A1(perm, i);
    for (int j = 1; j <= ndf; j++) {
      if (A2(id, ndf, j, n) == 0) A2(id, ndf, j, n) = ++neq;
      else A2(id, ndf, j, n) = 0;
    }
  }
  return neq;
}

/* -------------------------------------------------------------- CSRC build */
int or_csrstruct(const int32_t* id, const int32_t* ix, const int32_t* perm, int nnode,
                 int numel, int nen, int ndf, int neq, int32_t* ia, int32_t** ja_out) {
  int32_t* ip = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nnode + 1));
  int32_t* ips = NULL;
  or_graph(nnode, numel, nen, ix, ip, &ips);
  /* csria (Csr.f:92-199): count strictly-lower columns per row */
  for (int i = 1; i <= neq + 1; i++) A1(ia, i) = 0;
  for (int no = 1; no <= nnode; no++) {
    int i = A1(perm, no);
    for (int ii = 1; ii <= ndf; ii++) {
      int neqi = A2(id, ndf, ii, i);
      if (!(neqi > 0 && neqi <= neq)) continue;
      int c = 0;
      for (int kk = 1; kk <= ndf; kk++) {
        int neqj = A2(id, ndf, kk, i);
        if (neqj > 0 && neqj < neqi) c++;
      }
      for (int k = A1(ip, i); k <= A1(ip, i + 1) - 1; k++) {
        int j = A1(ips, k);
        for (int jj = 1; jj <= ndf; jj++) {
          int neqj = A2(id, ndf, jj, j);
          if (neqj > 0 && neqj < neqi) c++;
        }
      }
      A1(ia, neqi) = c;
    }
  }
  for (int i = neq; i >= 1; i--) A1(ia, i + 1) = A1(ia, i);
  A1(ia, 1) = 1;
  for (int i = 1; i <= neq; i++) A1(ia, i + 1) += A1(ia, i);
  int nad = A1(ia, neq + 1) - 1;
  /* csrja (Csr.f:200-305) then sortgraph (Csr.f:81) */
  int32_t* ja = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nad > 0 ? nad : 1));
  for (int no = 1; no <= nnode; no++) {
    int i = A1(perm, no);
    for (int ii = 1; ii <= ndf; ii++) {
      int neqi = A2(id, ndf, ii, i);
      if (!(neqi > 0 && neqi <= neq)) continue;
      int n = A1(ia, neqi) - 1;
      for (int kk = 1; kk <= ndf; kk++) {
        int neqj = A2(id, ndf, kk, i);
        if (neqj > 0 && neqj < neqi) ja[n++] = neqj;
      }
      for (int k = A1(ip, i); k <= A1(ip, i + 1) - 1; k++) {
        int j = A1(ips, k);
        for (int jj = 1; jj <= ndf; jj++) {
          int neqj = A2(id, ndf, jj, j);
          if (neqj > 0 && neqj < neqi) ja[n++] = neqj;
        }
      }
      int b0 = A1(ia, neqi) - 1;
      qsort(ja + b0, (size_t)(n - b0), sizeof(int32_t), cmp_i32);
    }
  }
  free(ips);
  free(ip);
  *ja_out = ja;
  return nad;
}

/* --------------------------------------------------------- element helpers */
/* Elmlib.f:815-886 */
static void sfquad4(double* h, double* hx, double* hy, double r, double s, int afl) {
  double opr = 1.0 + r, ops = 1.0 + s, omr = 1.0 - r, oms = 1.0 - s;
  if (afl) {
    h[0] = opr * ops * 0.25;
    h[1] = omr * ops * 0.25;
    h[2] = omr * oms * 0.25;
    h[3] = opr * oms * 0.25;
  }
  hx[0] = ops * 0.25;
  hx[1] = -ops * 0.25;
  hx[2] = -oms * 0.25;
  hx[3] = oms * 0.25;
  hy[0] = opr * 0.25;
  hy[1] = omr * 0.25;
  hy[2] = -omr * 0.25;
  hy[3] = -opr * 0.25;
}

/* Elmlib.f:1151-1259 */
static void sfhexa8(double* h, double* hx, double* hy, double* hz, double r, double s, double t,
                    int afl) {
  double pR = 1.0 + r, pS = 1.0 + s, pT = 1.0 + t, mR = 1.0 - r, mS = 1.0 - s, mT = 1.0 - t;
  if (afl) {
    h[0] = pR * pS * pT * 0.125;
    h[1] = mR * pS * pT * 0.125;
    h[2] = mR * mS * pT * 0.125;
    h[3] = pR * mS * pT * 0.125;
    h[4] = pR * pS * mT * 0.125;
    h[5] = mR * pS * mT * 0.125;
    h[6] = mR * mS * mT * 0.125;
    h[7] = pR * mS * mT * 0.125;
  }
  hx[0] = pS * pT * 0.125;
  hx[1] = -pS * pT * 0.125;
  hx[2] = -mS * pT * 0.125;
  hx[3] = mS * pT * 0.125;
  hx[4] = pS * mT * 0.125;
  hx[5] = -pS * mT * 0.125;
  hx[6] = -mS * mT * 0.125;
  hx[7] = mS * mT * 0.125;
  hy[0] = pR * pT * 0.125;
  hy[1] = mR * pT * 0.125;
  hy[2] = -mR * pT * 0.125;
  hy[3] = -pR * pT * 0.125;
  hy[4] = pR * mT * 0.125;
  hy[5] = mR * mT * 0.125;
  hy[6] = -mR * mT * 0.125;
  hy[7] = -pR * mT * 0.125;
  hz[0] = pR * pS * 0.125;
  hz[1] = mR * pS * 0.125;
  hz[2] = mR * mS * 0.125;
  hz[3] = pR * mS * 0.125;
  hz[4] = -pR * pS * 0.125;
  hz[5] = -mR * pS * 0.125;
  hz[6] = -mR * mS * 0.125;
  hz[7] = -pR * mS * 0.125;
}

/* Elmlib.f:1769-1852 (afl = true); x has leading dimension ndm. */
static int jacob2d(double* hx, double* hy, const double* x, int nen, int ndm, double* det) {
  double xj[2][2];
  for (int j = 0; j < 2; j++) {
    xj[0][j] = 0.0;
    xj[1][j] = 0.0;
    for (int k = 0; k < nen; k++) {
      xj[0][j] = xj[0][j] + hx[k] * x[k * ndm + j];
      xj[1][j] = xj[1][j] + hy[k] * x[k * ndm + j];
    }
  }
  *det = xj[0][0] * xj[1][1] - xj[1][0] * xj[0][1];
  if (*det <= JZERO) return -1;
  double d = *det;
  double i11 = xj[1][1] / d, i12 = -xj[0][1] / d, i21 = -xj[1][0] / d, i22 = xj[0][0] / d;
  for (int k = 0; k < nen; k++) {
    double a = hx[k], b = hy[k];
    hx[k] = i11 * a + i12 * b;
    hy[k] = i21 * a + i22 * b;
  }
  return 0;
}

/* Elmlib.f:1942-2037 (afl = true) */
static int jacob3d(double* hx, double* hy, double* hz, const double* x, int nen, double* det) {
  double xj[3][3];
  for (int j = 0; j < 3; j++) {
    xj[0][j] = 0.0;
    xj[1][j] = 0.0;
    xj[2][j] = 0.0;
    for (int k = 0; k < nen; k++) {
      xj[0][j] = xj[0][j] + hx[k] * x[k * 3 + j];
      xj[1][j] = xj[1][j] + hy[k] * x[k * 3 + j];
      xj[2][j] = xj[2][j] + hz[k] * x[k * 3 + j];
    }
  }
#define XJ(a, b) xj[(a)-1][(b)-1]
  *det = XJ(1, 1) * XJ(2, 2) * XJ(3, 3) + XJ(1, 2) * XJ(2, 3) * XJ(3, 1) +
         XJ(1, 3) * XJ(2, 1) * XJ(3, 2) - XJ(3, 1) * XJ(2, 2) * XJ(1, 3) -
         XJ(2, 1) * XJ(1, 2) * XJ(3, 3) - XJ(1, 1) * XJ(3, 2) * XJ(2, 3);
  if (*det <= JZERO) return -1;
  double d = *det, ji[3][3];
#define JI(a, b) ji[(a)-1][(b)-1]
  JI(1, 1) = (XJ(2, 2) * XJ(3, 3) - XJ(2, 3) * XJ(3, 2)) / d;
  JI(2, 1) = -(XJ(2, 1) * XJ(3, 3) - XJ(2, 3) * XJ(3, 1)) / d;
  JI(3, 1) = (XJ(2, 1) * XJ(3, 2) - XJ(2, 2) * XJ(3, 1)) / d;
  JI(1, 2) = -(XJ(1, 2) * XJ(3, 3) - XJ(1, 3) * XJ(3, 2)) / d;
  JI(2, 2) = (XJ(1, 1) * XJ(3, 3) - XJ(1, 3) * XJ(3, 1)) / d;
  JI(3, 2) = -(XJ(1, 1) * XJ(3, 2) - XJ(1, 2) * XJ(3, 1)) / d;
  JI(1, 3) = (XJ(1, 2) * XJ(2, 3) - XJ(1, 3) * XJ(2, 2)) / d;
  JI(2, 3) = -(XJ(1, 1) * XJ(2, 3) - XJ(1, 3) * XJ(2, 1)) / d;
  JI(3, 3) = (XJ(1, 1) * XJ(2, 2) - XJ(1, 2) * XJ(2, 1)) / d;
  for (int k = 0; k < nen; k++) {
    double a = hx[k], b = hy[k], c = hz[k];
    hx[k] = JI(1, 1) * a + JI(1, 2) * b + JI(1, 3) * c;
    hy[k] = JI(2, 1) * a + JI(2, 2) * b + JI(2, 3) * c;
    hz[k] = JI(3, 1) * a + JI(3, 2) * b + JI(3, 3) * c;
  }
#undef JI
#undef XJ
  return 0;
}

/* Elmlib.f:2574-2635, x leading dimension ndm, det only (afl = false) */
static int jtria3_det(const double* x, int ndm, double* det) {
  double j11 = x[0] - x[2 * ndm + 0], j12 = x[1] - x[2 * ndm + 1];
  double j21 = x[ndm + 0] - x[2 * ndm + 0], j22 = x[ndm + 1] - x[2 * ndm + 1];
  *det = j11 * j22 - j12 * j21;
  return (*det <= JZERO) ? -1 : 0;
}

/* Elmlib.f:2636-2713 */
static int jtetra4(const double* x, double* hx, double* hy, double* hz, double* det, int afl) {
#define X(i, n) x[((n)-1) * 3 + (i)-1]
  double xj[3][3];
#define XJ(a, b) xj[(a)-1][(b)-1]
  XJ(1, 1) = X(1, 1) - X(1, 4);
  XJ(1, 2) = X(2, 1) - X(2, 4);
  XJ(1, 3) = X(3, 1) - X(3, 4);
  XJ(2, 1) = X(1, 2) - X(1, 4);
  XJ(2, 2) = X(2, 2) - X(2, 4);
  XJ(2, 3) = X(3, 2) - X(3, 4);
  XJ(3, 1) = X(1, 3) - X(1, 4);
  XJ(3, 2) = X(2, 3) - X(2, 4);
  XJ(3, 3) = X(3, 3) - X(3, 4);
  *det = XJ(1, 1) * XJ(2, 2) * XJ(3, 3) + XJ(1, 2) * XJ(2, 3) * XJ(3, 1) +
         XJ(1, 3) * XJ(2, 1) * XJ(3, 2) - XJ(3, 1) * XJ(2, 2) * XJ(1, 3) -
         XJ(1, 2) * XJ(2, 1) * XJ(3, 3) - XJ(1, 1) * XJ(3, 2) * XJ(2, 3);
  if (*det <= JZERO) return -1;
  if (afl) {
    double di = 1.0 / *det;
    hx[0] = (XJ(2, 2) * XJ(3, 3) - XJ(2, 3) * XJ(3, 2)) * di;
    hx[1] = (XJ(1, 3) * XJ(3, 2) - XJ(1, 2) * XJ(3, 3)) * di;
    hx[2] = (XJ(1, 2) * XJ(2, 3) - XJ(1, 3) * XJ(2, 2)) * di;
    hy[0] = (XJ(2, 3) * XJ(3, 1) - XJ(2, 1) * XJ(3, 3)) * di;
    hy[1] = (XJ(1, 1) * XJ(3, 3) - XJ(1, 3) * XJ(3, 1)) * di;
    hy[2] = (XJ(1, 3) * XJ(2, 1) - XJ(1, 1) * XJ(2, 3)) * di;
    hz[0] = (XJ(2, 1) * XJ(3, 2) - XJ(2, 2) * XJ(3, 1)) * di;
    hz[1] = (XJ(1, 2) * XJ(3, 1) - XJ(1, 1) * XJ(3, 2)) * di;
    hz[2] = (XJ(1, 1) * XJ(2, 2) - XJ(1, 2) * XJ(2, 1)) * di;
    hx[3] = -hx[0] - hx[1] - hx[2];
    hy[3] = -hy[0] - hy[1] - hy[2];
    hz[3] = -hz[0] - hz[1] - hz[2];
  }
#undef XJ
#undef X
  return 0;
}

/* Elmlib.f:2038-2087 and 2159-2185 (rotate, transp = false) */
static void rotmatrix(const double* x, double r[3][3]) {
  double v1x = x[3] - x[0], v1y = x[4] - x[1], v1z = x[5] - x[2];
  double v2x = x[6] - x[0], v2y = x[7] - x[1], v2z = x[8] - x[2];
  double v3x = v1y * v2z - v1z * v2y;
  double v3y = v1z * v2x - v1x * v2z;
  double v3z = v1x * v2y - v1y * v2x;
  v2x = v1z * v3y - v1y * v3z;
  v2y = v1x * v3z - v1z * v3x;
  v2z = v1y * v3x - v1x * v3y;
  double n1 = sqrt(v1x * v1x + v1y * v1y + v1z * v1z);
  double n2 = sqrt(v2x * v2x + v2y * v2y + v2z * v2z);
  double n3 = sqrt(v3x * v3x + v3y * v3y + v3z * v3z);
  r[0][0] = v1x / n1;
  r[0][1] = v1y / n1;
  r[0][2] = v1z / n1;
  r[1][0] = v2x / n2;
  r[1][1] = v2y / n2;
  r[1][2] = v2z / n2;
  r[2][0] = v3x / n3;
  r[2][1] = v3y / n3;
  r[2][2] = v3z / n3;
}
static void rotate(double* x, double r[3][3]) {
  double x1 = r[0][0] * x[0] + r[0][1] * x[1] + r[0][2] * x[2];
  double x2 = r[1][0] * x[0] + r[1][1] * x[1] + r[1][2] * x[2];
  double x3 = r[2][0] * x[0] + r[2][1] * x[1] + r[2][2] * x[2];
  x[0] = x1;
  x[1] = x2;
  x[2] = x3;
}

/* Pload.f:119-261, the time-independent types used on this path (1, 2 at
 * t = 0, 40). */
static void tload(int nc, double* c, double* vc, const or_loads* ld) {
  *c = 0.0;
  if (nc <= 0) return;
  int ty = ld->type[nc - 1];
  if (ty == 1 || ty == 2) *c = ld->f[nc - 1][0];
  else if (ty == 40)
    for (int i = 0; i < ld->nparc[nc - 1]; i++) vc[i] = ld->f[nc - 1][i];
}

/* Elmlib.f:336-369 */
static void lku(const double* s, const double* u, double* p, int nst) {
  for (int i = 0; i < nst; i++) p[i] = 0.0;
  for (int j = 0; j < nst; j++)
    for (int i = 0; i < nst; i++) p[i] = p[i] + s[j * nst + i] * u[j];
}

#define S(i, j) s[((j)-1) * nst + (i)-1]

/* Elmlib02.f:50-627: tria3, plane strain (iel 2) or plane stress (iel 3) */
static int elmt_tria3(int ps_stress, int isw, const double* e, const int32_t* iq,
                      const double* x, const double* u, double* p, double* s, int ndm, int nst,
                      const or_loads* ld) {
  static const int side[3][2] = {{1, 2}, {2, 3}, {3, 1}};
  if (isw == 2) {
    double xj11 = x[0] - x[2 * ndm], xj12 = x[1] - x[2 * ndm + 1];
    double xj21 = x[ndm] - x[2 * ndm], xj22 = x[ndm + 1] - x[2 * ndm + 1];
    double det = xj11 * xj22 - xj12 * xj21;
    if (det <= 0.0) return -1;
    double ji11 = xj22 / det, ji12 = -xj12 / det, ji21 = -xj21 / det, ji22 = xj11 / det;
    double hx[3] = {ji11, ji12, -ji11 - ji12}, hy[3] = {ji21, ji22, -ji21 - ji22};
    double ym = e[0], ps = e[1], a, b, c;
    if (!ps_stress) {
      double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
      a = (ym * a1) / (a3 * a2);
      b = ps / a1;
      c = 0.5 * (a2 / a1);
    } else {
      double a1 = 1.0 - ps, a2 = 1.0 - ps * ps;
      a = ym / a2;
      b = ps;
      c = 0.5 * a1;
    }
    double d11 = a, d12 = a * b, d22 = a, d33 = a * c;
    for (int q = 0; q < nst * nst; q++) s[q] = 0.0;
    double wt1 = 0.5 * det;
    for (int j = 1; j <= 3; j++) {
      int k1 = 2 * j - 1, k2 = k1 + 1;
      for (int i = 1; i <= 3; i++) {
        int l1 = 2 * i - 1, l2 = l1 + 1;
        double hxi = hx[i - 1], hyi = hy[i - 1], hxj = hx[j - 1], hyj = hy[j - 1];
        S(l1, k1) = (hxi * d11 * hxj + hyi * d33 * hyj) * wt1;
        S(l1, k2) = (hxi * d12 * hyj + hyi * d33 * hxj) * wt1;
        S(l2, k1) = (hyi * d12 * hxj + hxi * d33 * hyj) * wt1;
        S(l2, k2) = (hyi * d22 * hyj + hxi * d33 * hxj) * wt1;
      }
    }
    lku(s, u, p, nst);
    return 0;
  }
  if (isw == 4) {
    int tp = 0;
    for (int i = 0; i < 3; i++) tp += iq[i];
    if (tp <= 0) return 0;
    for (int j = 0; j < 3; j++) {
      if (iq[j] <= 0) continue;
      const double* a = x + (side[j][0] - 1) * ndm;
      const double* b = x + (side[j][1] - 1) * ndm;
      double det = (b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]);
      if (det <= 0.0) return -1;
      det = sqrt(det);
      double c, ff[3] = {0, 0, 0};
      tload(iq[j], &c, ff, ld);
      double wt1 = 0.5 * det;
      for (int i = 0; i < 2; i++) {
        int no = side[j][i];
        int l1 = 2 * no - 1, l2 = l1 + 1;
        p[l1 - 1] = p[l1 - 1] - ff[0] * wt1;
        p[l2 - 1] = p[l2 - 1] - ff[1] * wt1;
      }
    }
  }
  return 0;
}

/* Elmlib02.f:629-1230: quad4, plane strain (iel 4) or plane stress (iel 5) */
static int elmt_quad4(int ps_stress, int isw, const double* e, const int32_t* iq,
                      const double* x, const double* u, double* p, double* s, int ndm, int nst,
                      const or_loads* ld) {
  static const int side[4][2] = {{1, 2}, {2, 3}, {3, 4}, {4, 1}};
  if (isw == 2) {
    double ym = e[0], ps = e[1], a, b, c, thic = 1.0;
    if (!ps_stress) {
      double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
      a = (ym * a1) / (a3 * a2);
      b = ps / a1;
      c = 0.5 * (a2 / a1);
    } else {
      thic = e[2];
      if (thic == 0) thic = 1.0;
      double a1 = 1.0 - ps, a2 = 1.0 - ps * ps;
      a = ym / a2;
      b = ps;
      c = 0.5 * a1;
    }
    double d11 = a, d12 = a * b, d22 = a, d33 = a * c;
    for (int q = 0; q < nst * nst; q++) s[q] = 0.0;
    for (int ly = 0; ly < 2; ly++) {
      double si = PG2[ly];
      for (int lx = 0; lx < 2; lx++) {
        double ri = PG2[lx], h[4], hx[4], hy[4], det;
        sfquad4(h, hx, hy, ri, si, 0);
        if (jacob2d(hx, hy, x, 4, ndm, &det)) return -1;
        double w = WG2[lx] * WG2[ly] * det;
        if (ps_stress) w = w * thic;
        double wt1 = w;
        for (int i = 1; i <= 4; i++) {
          int l1 = 2 * i - 1, l2 = l1 + 1;
          for (int j = 1; j <= 4; j++) {
            int k1 = 2 * j - 1, k2 = k1 + 1;
            double hxi = hx[i - 1], hyi = hy[i - 1], hxj = hx[j - 1], hyj = hy[j - 1];
            S(l1, k1) = S(l1, k1) + (hxi * d11 * hxj + hyi * d33 * hyj) * wt1;
            S(l1, k2) = S(l1, k2) + (hxi * d12 * hyj + hyi * d33 * hxj) * wt1;
            S(l2, k1) = S(l2, k1) + (hyi * d12 * hxj + hxi * d33 * hyj) * wt1;
            S(l2, k2) = S(l2, k2) + (hyi * d22 * hyj + hxi * d33 * hxj) * wt1;
          }
        }
      }
    }
    lku(s, u, p, nst);
    return 0;
  }
  if (isw == 4) {
    int tp = 0;
    for (int i = 0; i < 4; i++) tp += iq[i];
    if (tp <= 0) return 0;
    for (int j = 0; j < 4; j++) {
      if (iq[j] <= 0) continue;
      const double* a = x + (side[j][0] - 1) * ndm;
      const double* b = x + (side[j][1] - 1) * ndm;
      double det = (b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]);
      if (det <= 0.0) return -1;
      det = sqrt(det);
      double c, ff[3] = {0, 0, 0};
      tload(iq[j], &c, ff, ld);
      double wt1 = 0.5 * det;
      for (int i = 0; i < 2; i++) {
        int no = side[j][i];
        int l1 = 2 * no - 1, l2 = l1 + 1;
        p[l1 - 1] = p[l1 - 1] - ff[0] * wt1;
        p[l2 - 1] = p[l2 - 1] - ff[1] * wt1;
      }
    }
  }
  return 0;
}

/* Elmlib02.f:1232-1573: tet4 (iel 6) */
static int elmt_tet4(int isw, const double* e, const int32_t* iq, const double* x,
                     const double* u, double* p, double* s, int nst, const or_loads* ld,
                     const double* gravity) {
  static const int face[4][3] = {{2, 3, 4}, {1, 4, 3}, {1, 2, 4}, {1, 3, 2}};
  double hx[4], hy[4], hz[4], det;
  if (isw == 2) {
    double ym = e[0], ps = e[1];
    double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
    double a = (ym * a1) / (a3 * a2), b = ps / a1, c = 0.5 * (a2 / a1);
    for (int q = 0; q < nst * nst; q++) s[q] = 0.0;
    if (jtetra4(x, hx, hy, hz, &det, 1)) return -1;
    double wt1 = a * det * DIV6;
    for (int i = 1; i <= 4; i++) {
      int l1 = 3 * i - 2, l2 = l1 + 1, l3 = l2 + 1;
      double xi = hx[i - 1], yi = hy[i - 1], zi = hz[i - 1];
      for (int j = 1; j <= 4; j++) {
        int k1 = 3 * j - 2, k2 = k1 + 1, k3 = k2 + 1;
        double xj = hx[j - 1], yj = hy[j - 1], zj = hz[j - 1];
        S(l1, k1) = (xi * xj + c * yi * yj + c * zi * zj) * wt1;
        S(l1, k2) = (b * xi * yj + c * yi * xj) * wt1;
        S(l1, k3) = (b * xi * zj + c * zi * xj) * wt1;
        S(l2, k1) = (b * yi * xj + c * xi * yj) * wt1;
        S(l2, k2) = (yi * yj + c * xi * xj + c * zi * zj) * wt1;
        S(l2, k3) = (b * yi * zj + c * zi * yj) * wt1;
        S(l3, k1) = (b * zi * xj + c * xi * zj) * wt1;
        S(l3, k2) = (b * zi * yj + c * yi * zj) * wt1;
        S(l3, k3) = (zi * zj + c * yi * yj + c * xi * xj) * wt1;
      }
    }
    lku(s, u, p, nst);
    return 0;
  }
  if (isw == 4) {
    double density = e[2];
    if (density != 0.0) {
      density = e[2] * 1.0e-06;
      double gl[3] = {gravity[0] * density, gravity[1] * density, gravity[2] * density};
      if (jtetra4(x, hx, hy, hz, &det, 0)) return -1;
      double wt1 = det * DIV144;
      for (int i = 1; i <= 4; i++) {
        int l1 = 3 * i - 2;
        p[l1 - 1] = -wt1 * gl[0];
        p[l1] = -wt1 * gl[1];
        p[l1 + 1] = -wt1 * gl[2];
      }
    }
    int tp = 0;
    for (int i = 0; i < 4; i++) tp += iq[i];
    if (tp <= 0) return 0;
    for (int j = 0; j < 4; j++) {
      if (iq[j] <= 0) continue;
      double xf[9], r[3][3];
      for (int i = 0; i < 3; i++)
        for (int d = 0; d < 3; d++) xf[i * 3 + d] = x[(face[j][i] - 1) * 3 + d];
      rotmatrix(xf, r);
      for (int i = 0; i < 3; i++) rotate(xf + 3 * i, r);
      double c, ff[3] = {0, 0, 0};
      tload(iq[j], &c, ff, ld);
      if (jtria3_det(xf, 3, &det)) return -1;
      double wt1 = det * DIV6;
      for (int i = 0; i < 3; i++) {
        int no = face[j][i], l1 = 3 * no - 2;
        p[l1 - 1] = p[l1 - 1] - ff[0] * wt1;
        p[l1] = p[l1] - ff[1] * wt1;
        p[l1 + 1] = p[l1 + 1] - ff[2] * wt1;
      }
    }
  }
  return 0;
}

/* Elmlib02.f:1577-2023: hex8 (iel 7) */
static int elmt_hex8(int isw, const double* e, const int32_t* iq, const double* x,
                     const double* u, double* p, double* s, int nst, const or_loads* ld,
                     const double* gravity) {
  static const int face[6][4] = {{1, 2, 3, 4}, {5, 6, 7, 8}, {1, 5, 6, 2},
                                 {4, 3, 7, 8}, {1, 4, 8, 5}, {2, 6, 7, 3}};
  double h[8], hx[8], hy[8], hz[8], det;
  if (isw == 2) {
    double ym = e[0], ps = e[1];
    double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
    double a = (ym * a1) / (a3 * a2), b = ps / a1, c = 0.5 * (a2 / a1);
    for (int q = 0; q < nst * nst; q++) s[q] = 0.0;
    for (int lz = 0; lz < 2; lz++) {
      double ti = PG2[lz];
      for (int ly = 0; ly < 2; ly++) {
        double si = PG2[ly];
        for (int lx = 0; lx < 2; lx++) {
          double ri = PG2[lx];
          sfhexa8(h, hx, hy, hz, ri, si, ti, 0);
          if (jacob3d(hx, hy, hz, x, 8, &det)) return -1;
          double w = WG2[lx] * WG2[ly] * WG2[lz] * det;
          double wt1 = w * a;
          for (int i = 1; i <= 8; i++) {
            int l1 = 3 * i - 2, l2 = l1 + 1, l3 = l2 + 1;
            double xi = hx[i - 1], yi = hy[i - 1], zi = hz[i - 1];
            for (int j = 1; j <= 8; j++) {
              int k1 = 3 * j - 2, k2 = k1 + 1, k3 = k2 + 1;
              double xj = hx[j - 1], yj = hy[j - 1], zj = hz[j - 1];
              S(l1, k1) = S(l1, k1) + (xi * xj + c * (yi * yj + zi * zj)) * wt1;
              S(l1, k2) = S(l1, k2) + (b * xi * yj + c * yi * xj) * wt1;
              S(l1, k3) = S(l1, k3) + (b * xi * zj + c * zi * xj) * wt1;
              S(l2, k1) = S(l2, k1) + (b * yi * xj + c * xi * yj) * wt1;
              S(l2, k2) = S(l2, k2) + (yi * yj + c * (xi * xj + zi * zj)) * wt1;
              S(l2, k3) = S(l2, k3) + (b * yi * zj + c * zi * yj) * wt1;
              S(l3, k1) = S(l3, k1) + (b * zi * xj + c * xi * zj) * wt1;
              S(l3, k2) = S(l3, k2) + (b * zi * yj + c * yi * zj) * wt1;
              S(l3, k3) = S(l3, k3) + (zi * zj + c * (yi * yj + xi * xj)) * wt1;
            }
          }
        }
      }
    }
    lku(s, u, p, nst);
    return 0;
  }
  if (isw == 4) {
    double density = e[2];
    if (density != 0.0) {
      density = e[2] * 1.0e-06;
      double gl[3] = {gravity[0] * density, gravity[1] * density, gravity[2] * density};
      for (int lz = 0; lz < 2; lz++)
        for (int ly = 0; ly < 2; ly++)
          for (int lx = 0; lx < 2; lx++) {
            sfhexa8(h, hx, hy, hz, PG2[lx], PG2[ly], PG2[lz], 1);
            if (jacob3d(hx, hy, hz, x, 8, &det)) return -1;
            double w = WG2[lx] * WG2[ly] * WG2[lz] * det;
            for (int i = 1; i <= 8; i++) {
              int l1 = 3 * i - 2;
              double wt1 = h[i - 1] * w;
              p[l1 - 1] = p[l1 - 1] - wt1 * gl[0];
              p[l1] = p[l1] - wt1 * gl[1];
              p[l1 + 1] = p[l1 + 1] - wt1 * gl[2];
            }
          }
    }
    int tp = 0;
    for (int i = 0; i < 6; i++) tp += iq[i];
    if (tp <= 0) return 0;
    for (int j = 0; j < 6; j++) {
      if (iq[j] <= 0) continue;
      double xf[12], r[3][3];
      for (int i = 0; i < 4; i++)
        for (int d = 0; d < 3; d++) xf[i * 3 + d] = x[(face[j][i] - 1) * 3 + d];
      rotmatrix(xf, r);
      for (int i = 0; i < 4; i++) rotate(xf + 3 * i, r);
      double c, ff[3] = {0, 0, 0};
      tload(iq[j], &c, ff, ld);
      for (int ly = 0; ly < 2; ly++)
        for (int lx = 0; lx < 2; lx++) {
          double hq[4], hqx[4], hqy[4];
          sfquad4(hq, hqx, hqy, PG2[lx], PG2[ly], 1);
          if (jacob2d(hqx, hqy, xf, 4, 3, &det)) return -1;
          double w = WG2[lx] * WG2[ly] * det;
          for (int i = 0; i < 4; i++) {
            int no = face[j][i];
            double wt1 = hq[i] * w;
            int l1 = 3 * no - 2;
            p[l1 - 1] = p[l1 - 1] - ff[0] * wt1;
            p[l1] = p[l1] - ff[1] * wt1;
            p[l1 + 1] = p[l1 + 1] - ff[2] * wt1;
          }
        }
    }
  }
  return 0;
}
#undef S

int or_element(int iel, int isw, const double* e, const int32_t* iq, const double* xl,
               const double* ul, double* p, double* s, int ndm, int nst, const or_loads* ld,
               const double* gravity) {
  switch (iel) {
    case 2: return elmt_tria3(0, isw, e, iq, xl, ul, p, s, ndm, nst, ld);
    case 3: return elmt_tria3(1, isw, e, iq, xl, ul, p, s, ndm, nst, ld);
    case 4: return elmt_quad4(0, isw, e, iq, xl, ul, p, s, ndm, nst, ld);
    case 5: return elmt_quad4(1, isw, e, iq, xl, ul, p, s, ndm, nst, ld);
    case 6: return elmt_tet4(isw, e, iq, xl, ul, p, s, nst, ld, gravity);
    case 7: return elmt_hex8(isw, e, iq, xl, ul, p, s, nst, ld, gravity);
  }
  return -2;
}

/* ------------------------------------------------------------- assembly */
int or_pform(const int32_t* ix, const int32_t* iq, const int32_t* ie, const double* e,
             const double* x, const int32_t* id, const int32_t* ia, const int32_t* ja,
             double* al, double* ad, double* b, const double* u, int numel, int nen,
             int nenv, int ndf, int ndm, int neq, int lhs, int rhs, int isw,
             const or_loads* ld, const double* gravity) {
  int nst = nen * ndf;
  if (lhs) {
    int nad = A1(ia, neq + 1) - 1;
    memset(al, 0, sizeof(double) * (size_t)nad);
    memset(ad, 0, sizeof(double) * (size_t)neq);
  }
  int32_t ldv[64];
  double ul[64], pl[64], sl[64 * 64], xl[3 * 8];
  for (int nel = 1; nel <= numel; nel++) {
    int kk = 0;
    for (int i = 1; i <= nen; i++) {
      int no = A2(ix, nen + 1, i, nel);
      for (int j = 1; j <= ndf; j++) {
        ldv[kk] = A2(id, ndf, j, no);
        pl[kk] = 0.0;
        ul[kk] = A2(u, ndf, j, no);
        kk++;
      }
    }
    for (int i = 1; i <= nenv; i++) {
      int no = A2(ix, nen + 1, i, nel);
      for (int j = 1; j <= ndm; j++) xl[(i - 1) * ndm + j - 1] = A2(x, ndm, j, no);
    }
    int ma = A2(ix, nen + 1, nen + 1, nel);
    int iel = A1(ie, ma);
    if (or_element(iel, isw, e + (size_t)(ma - 1) * OR_PROP, iq + (size_t)(nel - 1) * 7, xl,
                   ul, pl, sl, ndm, nst, ld, gravity))
      return nel;
    /* Assbly.f:66-146 (csr) */
    for (int i = 1; i <= nst; i++) {
      int k = ldv[i - 1];
      if (k <= 0 || k > neq) continue;
      if (rhs) A1(b, k) = A1(b, k) - pl[i - 1];
      if (lhs) {
        A1(ad, k) = A1(ad, k) + sl[(i - 1) * nst + i - 1];
        for (int j = 1; j <= nst; j++) {
          int l = ldv[j - 1];
          if (l <= 0) continue;
          for (int q = A1(ia, k); q <= A1(ia, k + 1) - 1; q++)
            if (l == A1(ja, q)) A1(al, q) = A1(al, q) + sl[(j - 1) * nst + i - 1];
        }
      }
    }
  }
  return 0;
}

/* ------------------------------------------ nodal stresses (isw = 3) */
/* deform2d / deform3d (Elmlib.f:2368-2472): strains from nodal derivatives */
static void deform2d(const double* hx, const double* hy, const double* u, double* eps, int nen) {
  double ux = 0.0, uy = 0.0, vx = 0.0, vy = 0.0;
  for (int i = 1; i <= nen; i++) {
    int j1 = 2 * i - 1, j2 = j1 + 1;
    ux = ux + hx[i - 1] * u[j1 - 1];
    uy = uy + hy[i - 1] * u[j1 - 1];
    vx = vx + hx[i - 1] * u[j2 - 1];
    vy = vy + hy[i - 1] * u[j2 - 1];
  }
  eps[0] = ux;
  eps[1] = vy;
  eps[2] = uy + vx;
}

static void deform3d(const double* hx, const double* hy, const double* hz, const double* u,
                     double* eps, int nen) {
  double ux = 0.0, uy = 0.0, uz = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
  for (int i = 1; i <= nen; i++) {
    int j1 = 3 * i - 2, j2 = j1 + 1, j3 = j2 + 1;
    ux = ux + hx[i - 1] * u[j1 - 1];
    uy = uy + hy[i - 1] * u[j1 - 1];
    uz = uz + hz[i - 1] * u[j1 - 1];
    vx = vx + hx[i - 1] * u[j2 - 1];
    vy = vy + hy[i - 1] * u[j2 - 1];
    vz = vz + hz[i - 1] * u[j2 - 1];
    wx = wx + hx[i - 1] * u[j3 - 1];
    wy = wy + hy[i - 1] * u[j3 - 1];
    wz = wz + hz[i - 1] * u[j3 - 1];
  }
  eps[0] = ux;
  eps[1] = vy;
  eps[2] = wz;
  eps[3] = uy + vx;
  eps[4] = vz + wy;
  eps[5] = uz + wx;
}

/* stress2d_m (Elmlib.f:2502-2539): sxx syy szz sxy; szz = ps (sxx + syy) in
 * plane strain, 0 in plane stress */
static void stress2d_m(double d11, double d12, double d22, double d33, double ps, int ept,
                       const double* eps, double* t) {
  t[0] = d11 * eps[0] + d12 * eps[1];
  t[1] = d12 * eps[0] + d22 * eps[1];
  t[2] = ept ? ps * (t[0] + t[1]) : 0.0;
  t[3] = d33 * eps[2];
}

/* stress3d (Elmlib.f:2540-2573): sxx syy szz sxy syz sxz */
static void stress3d(double a, double b, double c, const double* eps, double* t) {
  t[0] = (eps[0] + b * eps[1] + b * eps[2]) * a;
  t[1] = (b * eps[0] + eps[1] + b * eps[2]) * a;
  t[2] = (b * eps[0] + b * eps[1] + eps[2]) * a;
  t[3] = eps[3] * c * a;
  t[4] = eps[4] * c * a;
  t[5] = eps[5] * c * a;
}

/* the constitutive constants a, b, c of the elements (Elmlib02.f:1693-1702
 * and the plane-stress form, Elmlib02.f:1033-1048) */
static void abc(int plane_stress, double ym, double ps, double* a, double* b, double* c) {
  if (!plane_stress) {
    double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
    *a = (ym * a1) / (a3 * a2);
    *b = ps / a1;
    *c = 0.5 * (a2 / a1);
  } else {
    double a1 = 1.0 - ps, a2 = 1.0 - ps * ps;
    *a = ym / a2;
    *b = ps;
    *c = 0.5 * a1;
  }
}

/* isw = 3 of the element library: the element's nodal stresses p(ntn, nen).
 * tria3 (Elmlib02.f:143-190 then 300: one constant stress copied to its 3
 * nodes), quad4 (Elmlib02.f:716+ and 1018+ label 300: at each node's natural
 * coordinates), tet4 (Elmlib02.f:1312+ label 300: constant, 4 copies), hex8
 * (Elmlib02.f:1791-1815: at each node's natural coordinates). */
int or_element_stress(int iel, const double* e, const double* x, const double* u, double* p,
                      int ndm) {
  static const double rq[4] = {1.0, -1.0, -1.0, 1.0}, sq[4] = {1.0, 1.0, -1.0, -1.0};
  static const double rh[8] = {1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0};
  static const double sh[8] = {1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0};
  static const double th[8] = {1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0};
  double a, b, c, eps[6], det;
  double ym = e[0], ps = e[1];
  if (iel == 2 || iel == 3) {
    double xj11 = x[0] - x[2 * ndm], xj12 = x[1] - x[2 * ndm + 1];
    double xj21 = x[ndm] - x[2 * ndm], xj22 = x[ndm + 1] - x[2 * ndm + 1];
    det = xj11 * xj22 - xj12 * xj21;
    if (det <= 0.0) return -1;
    double ji11 = xj22 / det, ji12 = -xj12 / det, ji21 = -xj21 / det, ji22 = xj11 / det;
    double hx[3] = {ji11, ji12, -ji11 - ji12}, hy[3] = {ji21, ji22, -ji21 - ji22};
    abc(iel == 3, ym, ps, &a, &b, &c);
    deform2d(hx, hy, u, eps, 3);
    stress2d_m(a, a * b, a, a * c, ps, iel == 2, eps, p);
    for (int k = 0; k < 4; k++) p[4 + k] = p[8 + k] = p[k];
    return 0;
  }
  if (iel == 4 || iel == 5) {
    abc(iel == 5, ym, ps, &a, &b, &c);
    for (int i = 0; i < 4; i++) {
      double h[4], hx[4], hy[4];
      sfquad4(h, hx, hy, rq[i], sq[i], 0);
      if (jacob2d(hx, hy, x, 4, ndm, &det)) return -1;
      deform2d(hx, hy, u, eps, 4);
      stress2d_m(a, a * b, a, a * c, ps, iel == 4, eps, p + 4 * i);
    }
    return 0;
  }
  if (iel == 6) {
    double hx[4], hy[4], hz[4];
    abc(0, ym, ps, &a, &b, &c);
    if (jtetra4(x, hx, hy, hz, &det, 1)) return -1;
    deform3d(hx, hy, hz, u, eps, 4);
    stress3d(a, b, c, eps, p);
    for (int k = 0; k < 6; k++) p[6 + k] = p[12 + k] = p[18 + k] = p[k];
    return 0;
  }
  if (iel == 7) {
    abc(0, ym, ps, &a, &b, &c);
    for (int i = 0; i < 8; i++) {
      double h[8], hx[8], hy[8], hz[8];
      sfhexa8(h, hx, hy, hz, rh[i], sh[i], th[i], 0);
      if (jacob3d(hx, hy, hz, x, 8, &det)) return -1;
      deform3d(hx, hy, hz, u, eps, 8);
      stress3d(a, b, c, eps, p + 6 * i);
    }
    return 0;
  }
  return -2;
}

/* tform_mec (Pform.f:258-421) with isw = 3: element nodal stresses added
 * to their nodes in element order, divided by the node's element count,
 * plus the initial stress tx0 (may be NULL).  u(ndf, nnode), tx(ntn,
 * nnode), ntn = 4 (2-D) or 6 (3-D).  Returns 0, or the 1-based element
 * whose Jacobian is not positive at a node. */
int or_tform(const int32_t* ix, const double* x, const double* e, const int32_t* ie,
             const double* u, const double* tx0, double* tx, int nnode, int numel, int nen,
             int ndm, int ndf, int ntn) {
  int32_t* ic = (int32_t*)calloc((size_t)nnode + 1, sizeof(int32_t));
  for (size_t q = 0; q < (size_t)nnode * ntn; q++) tx[q] = 0.0;
  double ul[64], pl[64], xl[3 * 8];
  int rc = 0;
  for (int nel = 1; nel <= numel && !rc; nel++) {
    for (int q = 0; q < nen * ntn; q++) pl[q] = 0.0;
    int kk = 0;
    for (int i = 1; i <= nen; i++) {
      int no = A2(ix, nen + 1, i, nel);
      for (int j = 1; j <= ndf; j++) ul[kk++] = A2(u, ndf, j, no);
    }
    for (int i = 1; i <= nen; i++) {
      int no = A2(ix, nen + 1, i, nel);
      for (int j = 1; j <= ndm; j++) xl[(i - 1) * ndm + j - 1] = A2(x, ndm, j, no);
    }
    int ma = A2(ix, nen + 1, nen + 1, nel);
    int iel = A1(ie, ma);
    if (or_element_stress(iel, e + (size_t)(ma - 1) * OR_PROP, xl, ul, pl, ndm)) {
      rc = nel;
      break;
    }
    for (int i = 1; i <= nen; i++) {
      int no = A2(ix, nen + 1, i, nel);
      if (no <= 0) continue;
      ic[no] += 1;
      for (int j = 1; j <= ntn; j++) A2(tx, ntn, j, no) = A2(tx, ntn, j, no) + pl[(i - 1) * ntn + j - 1];
    }
  }
  for (int i = 1; i <= nnode && !rc; i++)
    for (int j = 1; j <= ntn; j++)
      A2(tx, ntn, j, i) = A2(tx, ntn, j, i) / (double)ic[i] + (tx0 ? A2(tx0, ntn, j, i) : 0.0);
  free(ic);
  return rc;
}

void or_pload(const int32_t* id, double* f, double* u, double* b, const int32_t* nload,
              int nnode, int ndf, const or_loads* ld) {
  for (int i = 1; i <= nnode; i++)
    for (int j = 1; j <= ndf; j++) {
      int nc = A2(nload, ndf, j, i);
      if (nc > 0) {
        double c, vc[3];
        tload(nc, &c, vc, ld);
        A2(f, ndf, j, i) = c;
      }
    }
  for (int i = 1; i <= nnode; i++)
    for (int j = 1; j <= ndf; j++) {
      int k = A2(id, ndf, j, i);
      if (k > 0) A1(b, k) = A2(f, ndf, j, i);
      else A2(u, ndf, j, i) = A2(f, ndf, j, i);
    }
}

void or_update(int nnode, int ndf, const int32_t* id, double* u, const double* x) {
  for (int i = 1; i <= nnode; i++)
    for (int j = 1; j <= ndf; j++) {
      int k = A2(id, ndf, j, i);
      if (k > 0) A2(u, ndf, j, i) = A1(x, k);
    }
}

/* ------------------------------------------------------- preconditioners */
void or_pre_diag(double* m, const double* ad, int neq) {
  for (int i = 0; i < neq; i++) m[i] = 1.0 / ad[i];
}

void or_block_precond3(const double* ad, const double* al, const int32_t* ia,
                       const int32_t* ja, double* m, int neq) {
  int ndiv = neq / 3, rest = neq % 3;
  for (int i = 0; i < 6 * (ndiv + rest); i++) m[i] = 0.0;
  for (int i = 1; i <= ndiv; i++) {
    int ii = 3 * i - 2;
    double a11 = A1(ad, ii), a22 = A1(ad, ii + 1), a33 = A1(ad, ii + 2);
    double a21 = 0.0, a31 = 0.0, a32 = 0.0;
    for (int j = A1(ia, ii + 1); j <= A1(ia, ii + 2) - 1; j++)
      if (A1(ja, j) == ii) a21 = A1(al, j);
    for (int j = A1(ia, ii + 2); j <= A1(ia, ii + 3) - 1; j++) {
      int jj = A1(ja, j);
      if (jj == ii) a31 = A1(al, j);
      else if (jj == ii + 1) a32 = A1(al, j);
    }
    /* Precond.f:1676-1710 */
    double c11 = a22 * a33 - a32 * a32;
    double c22 = a11 * a33 - a31 * a31;
    double c33 = a11 * a22 - a21 * a21;
    double c21 = a31 * a32 - a21 * a33;
    double c31 = a21 * a32 - a31 * a22;
    double c32 = a31 * a21 - a11 * a32;
    double det = a11 * a22 * a33 + a21 * a32 * a31 + a31 * a21 * a32 -
                 (a22 * a31 * a31 + a11 * a32 * a32 + a33 * a21 * a21);
    det = 1.0 / det;
    double* mi = m + 6 * (i - 1);
    mi[0] = c11 * det;
    mi[1] = c22 * det;
    mi[2] = c33 * det;
    mi[3] = c21 * det;
    mi[4] = c31 * det;
    mi[5] = c32 * det;
  }
  for (int i = 1; i <= rest; i++) m[6 * (ndiv + i - 1)] = 1.0 / A1(ad, 3 * ndiv + i);
}

void or_op_block_precond3(const double* m, const double* x, double* y, int neq) {
  int ndiv = neq / 3, rest = neq % 3;
  for (int i = 1; i <= ndiv; i++) {
    int ii = 3 * i - 3;
    const double* a = m + 6 * (i - 1);
    double x1 = x[ii], x2 = x[ii + 1], x3 = x[ii + 2];
    y[ii] = a[0] * x1 + a[3] * x2 + a[4] * x3;
    y[ii + 1] = a[3] * x1 + a[1] * x2 + a[5] * x3;
    y[ii + 2] = a[4] * x1 + a[5] * x2 + a[2] * x3;
  }
  for (int i = 1; i <= rest; i++) {
    int ii = 3 * ndiv + i - 1;
    y[ii] = m[6 * (ndiv + i - 1)] * x[ii];
  }
}

/* Precond.f:163-358 nblock=2: 2x2 inverses of consecutive equation pairs,
 * m(3, neq/2 + rest) = (a22, a11, -a21)/det, then 1/ad for the last equation. */
void or_block_precond2(const double* ad, const double* al, const int32_t* ia,
                       const int32_t* ja, double* m, int neq) {
  int ndiv = neq / 2, rest = neq % 2;
  for (int i = 0; i < 3 * (ndiv + rest); i++) m[i] = 0.0;
  for (int i = 1; i <= ndiv; i++) {
    int ii = 2 * i - 1;
    double a11 = A1(ad, ii), a22 = A1(ad, ii + 1), a21 = 0.0;
    for (int j = A1(ia, ii + 1); j <= A1(ia, ii + 2) - 1; j++)
      if (A1(ja, j) == ii) a21 = A1(al, j);
    double idet = 1.0 / (a11 * a22 - a21 * a21);
    double* mi = m + 3 * (i - 1);
    mi[0] = idet * a22;
    mi[1] = idet * a11;
    mi[2] = -idet * a21;
  }
  for (int i = 1; i <= rest; i++) m[3 * (ndiv + i - 1)] = 1.0 / A1(ad, 2 * ndiv + i);
}

/* Precond.f:419-540 nblock=2 */
void or_op_block_precond2(const double* m, const double* x, double* y, int neq) {
  int ndiv = neq / 2, rest = neq % 2;
  for (int i = 1; i <= ndiv; i++) {
    int ii = 2 * i - 2;
    const double* a = m + 3 * (i - 1);
    double x1 = x[ii], x2 = x[ii + 1];
    y[ii] = a[0] * x1 + a[2] * x2;
    y[ii + 1] = a[2] * x1 + a[1] * x2;
  }
  for (int i = 1; i <= rest; i++) {
    int ii = 2 * ndiv + i - 1;
    y[ii] = m[3 * (ndiv + i - 1)] * x[ii];
  }
}

/* ------------------------------------------------------------------ SpMV */
void or_matvec_csrcsym1(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                        const double* al, const double* x, double* y) {
  for (int i = 1; i <= neq; i++) {
    double xi = A1(x, i);
    double t = A1(ad, i) * xi;
    int k1 = A1(ia, i), k2 = A1(ia, i + 1);
    if (k2 != k1) {
      if ((k2 - k1) % 2 != 0) {
        int jak1 = A1(ja, k1);
        t = t + A1(al, k1) * A1(x, jak1);
        A1(y, jak1) = A1(y, jak1) + A1(al, k1) * xi;
        k1++;
      }
      for (int k = k1; k <= k2 - 1; k += 2) {
        int jak = A1(ja, k), jak1 = A1(ja, k + 1);
        double s = A1(al, k), s1 = A1(al, k + 1);
        t = t + s * A1(x, jak) + s1 * A1(x, jak1);
        A1(y, jak) = A1(y, jak) + s * xi;
        A1(y, jak1) = A1(y, jak1) + s1 * xi;
      }
    }
    A1(y, i) = t;
  }
}

double or_dot(const double* a, const double* b, int n) {
  double t = 0.0;
  for (int i = 0; i < n; i++) t = t + a[i] * b[i];
  return t;
}

/* --------------------------------------------------------------- PCG */
void or_pcg(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
            const double* al, const double* m, const double* b, double* x, double tol,
            int maxit, or_stats* st) {
  memset(st, 0, sizeof(*st));
  for (int i = 0; i < neq; i++)
    if (ad[i] == 0.0) {
      st->status = 2;
      return;
    }
  double* z = (double*)malloc(sizeof(double) * (size_t)neq);
  double* r = (double*)malloc(sizeof(double) * (size_t)neq);
  double* p = (double*)malloc(sizeof(double) * (size_t)neq);
  for (int i = 0; i < neq; i++) x[i] = 0.0;
  for (int i = 0; i < neq; i++) z[i] = b[i] * m[i];
  double d = or_dot(b, z, neq);
  double conv = tol * sqrt(fabs(d));
  or_matvec_csrcsym1(neq, ia, ja, ad, al, x, z);
  for (int i = 0; i < neq; i++) {
    r[i] = b[i] - z[i];
    z[i] = r[i] * m[i];
    p[i] = z[i];
  }
  d = or_dot(r, z, neq);
  int j, conv_ok = 0;
  for (j = 1; j <= maxit; j++) {
    or_matvec_csrcsym1(neq, ia, ja, ad, al, p, z);
    double alpha = d / or_dot(z, p, neq);
    for (int i = 0; i < neq; i++) {
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] - alpha * z[i];
      z[i] = r[i] * m[i];
    }
    double di = or_dot(r, z, neq);
    double beta = di / d;
    for (int i = 0; i < neq; i++) p[i] = z[i] + beta * p[i];
    d = di;
    if (sqrt(fabs(d)) < conv) {
      conv_ok = 1;
      break;
    }
  }
  if (!conv_ok) st->status = 1;
  st->iters = conv_ok ? j : maxit;
  or_matvec_csrcsym1(neq, ia, ja, ad, al, x, z);
  st->xkx = or_dot(x, z, neq);
  st->norm_x = sqrt(or_dot(x, x, neq));
  for (int i = 0; i < neq; i++) {
    r[i] = b[i] - z[i];
    z[i] = r[i] * m[i];
  }
  st->norm_m_r = sqrt(fabs(or_dot(r, z, neq)));
  st->norm_r = sqrt(or_dot(r, r, neq));
  st->conv = conv;
  free(z);
  free(r);
  free(p);
}

/* Iters.f:252-524 with the Jacobi step z(i) = r(i)*m(i) replaced by the 3x3
 * block-Jacobi application op_block_precond (Precond.f:419-540) of the m(6,.)
 * that block_precond (Precond.f:163-358) builds: the preconditioned CG that
 * the reference's precond code 6 names (its call site is dead).  Stopping
 * rule and post-solve figures as in or_pcg, with M the block inverse. */
static void op_block(int nblock, const double* bm, const double* x, double* y, int neq) {
  if (nblock == 2) or_op_block_precond2(bm, x, y, neq);
  else or_op_block_precond3(bm, x, y, neq);
}

void or_pcg_bdiag3(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                   const double* al, const double* bm, const double* b, double* x, double tol,
                   int maxit, or_stats* st) {
  or_pcg_block(3, neq, ia, ja, ad, al, bm, b, x, tol, maxit, st);
}

void or_pcg_block(int nblock, int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                  const double* al, const double* bm, const double* b, double* x, double tol,
                  int maxit, or_stats* st) {
  memset(st, 0, sizeof(*st));
  for (int i = 0; i < neq; i++)
    if (ad[i] == 0.0) {
      st->status = 2;
      return;
    }
  double* z = (double*)malloc(sizeof(double) * (size_t)neq);
  double* r = (double*)malloc(sizeof(double) * (size_t)neq);
  double* p = (double*)malloc(sizeof(double) * (size_t)neq);
  for (int i = 0; i < neq; i++) x[i] = 0.0;
  op_block(nblock, bm, b, z, neq);
  double d = or_dot(b, z, neq);
  double conv = tol * sqrt(fabs(d));
  or_matvec_csrcsym1(neq, ia, ja, ad, al, x, z);
  for (int i = 0; i < neq; i++) r[i] = b[i] - z[i];
  op_block(nblock, bm, r, z, neq);
  for (int i = 0; i < neq; i++) p[i] = z[i];
  d = or_dot(r, z, neq);
  int j, conv_ok = 0;
  for (j = 1; j <= maxit; j++) {
    or_matvec_csrcsym1(neq, ia, ja, ad, al, p, z);
    double alpha = d / or_dot(z, p, neq);
    for (int i = 0; i < neq; i++) {
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] - alpha * z[i];
    }
    op_block(nblock, bm, r, z, neq);
    double di = or_dot(r, z, neq);
    double beta = di / d;
    for (int i = 0; i < neq; i++) p[i] = z[i] + beta * p[i];
    d = di;
    if (sqrt(fabs(d)) < conv) {
      conv_ok = 1;
      break;
    }
  }
  if (!conv_ok) st->status = 1;
  st->iters = conv_ok ? j : maxit;
  or_matvec_csrcsym1(neq, ia, ja, ad, al, x, z);
  st->xkx = or_dot(x, z, neq);
  st->norm_x = sqrt(or_dot(x, x, neq));
  for (int i = 0; i < neq; i++) r[i] = b[i] - z[i];
  op_block(nblock, bm, r, z, neq);
  st->norm_m_r = sqrt(fabs(or_dot(r, z, neq)));
  st->norm_r = sqrt(or_dot(r, r, neq));
  st->conv = conv;
  free(z);
  free(r);
  free(p);
}

/* Matrix_partition.f:98-133 (partition_csrc_bynonzeros, one process) and
 * 196-211 (compute_effective_work).  Bounds are 1-based rows. */
static void partition_rows(const int32_t* ia, const int32_t* ja, int neq, int nth, int* tbeg,
                           int* tend, int* thgt) {
  int nad = A1(ia, neq + 1) - 1;
  int mean = (2 * nad + neq) / nth + 1;
  int line = 2;
  tbeg[0] = 1;
  for (int i = 0; i < nth - 1; i++) {
    int size = 0;
    for (;;) {
      int tam = 2 * (A1(ia, line) - A1(ia, line - 1)) + 1;
      if (size + tam <= mean) {
        size += tam;
        tend[i] = line - 1;
        tbeg[i + 1] = line;
        line++;
      } else
        break;
    }
  }
  tend[nth - 1] = neq;
  for (int t = 0; t < nth; t++) {
    int h = tbeg[t];
    for (int i = tbeg[t]; i <= tend[t]; i++) {
      int c = A1(ja, A1(ia, i));
      if (c < h) h = c;
    }
    thgt[t] = h;
  }
}

void or_pcg_omp(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                const double* al, const double* m, const double* b, double* x, double tol,
                int maxit, int nth, or_stats* st) {
  memset(st, 0, sizeof(*st));
  for (int i = 0; i < neq; i++)
    if (ad[i] == 0.0) {
      st->status = 2;
      return;
    }
  if (nth < 1) nth = 1;
  int* tbeg = (int*)malloc(sizeof(int) * (size_t)nth);
  int* tend = (int*)malloc(sizeof(int) * (size_t)nth);
  int* thgt = (int*)malloc(sizeof(int) * (size_t)nth);
  partition_rows(ia, ja, neq, nth, tbeg, tend, thgt);
  double* ty = (double*)malloc(sizeof(double) * (size_t)nth * (size_t)neq);
  double* z = (double*)malloc(sizeof(double) * (size_t)neq);
  double* r = (double*)malloc(sizeof(double) * (size_t)neq);
  double* p = (double*)malloc(sizeof(double) * (size_t)neq);
  double dsh = 0.0, red[2] = {0, 0};
  int iters = maxit, conv_ok = 0;
  double conv = 0.0;
#pragma omp parallel num_threads(nth)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    /* Matvec_omp.f:150-212 as a closure over the team */
#define OMP_MATVEC(XV, YV)                                                          \
  do {                                                                              \
    for (int q = 0; q < nth; q++) {                                                 \
      double* tq = ty + (size_t)q * neq;                                            \
      _Pragma("omp for schedule(static)") for (int jj = thgt[q]; jj < tbeg[q]; jj++) \
        tq[jj - 1] = 0.0;                                                           \
    }                                                                               \
    double* tm = ty + (size_t)tid * neq;                                            \
    for (int i = tbeg[tid]; i <= tend[tid]; i++) {                                  \
      double xi = (XV)[i - 1], t = ad[i - 1] * xi;                                  \
      (YV)[i - 1] = 0.0;                                                            \
      for (int k = A1(ia, i); k <= A1(ia, i + 1) - 1; k++) {                        \
        int c = A1(ja, k);                                                          \
        double s = A1(al, k);                                                       \
        t = t + s * (XV)[c - 1];                                                    \
        tm[c - 1] = tm[c - 1] + s * xi;                                             \
      }                                                                             \
      tm[i - 1] = t;                                                                \
    }                                                                               \
    _Pragma("omp barrier") for (int q = 0; q < nth; q++) {                          \
      double* tq = ty + (size_t)q * neq;                                            \
      _Pragma("omp for schedule(static)") for (int jj = thgt[q]; jj <= tend[q]; jj++) \
        (YV)[jj - 1] = (YV)[jj - 1] + tq[jj - 1];                                   \
    }                                                                               \
  } while (0)
#define OMP_DOT(A, B, OUT)                                                     \
  do {                                                                         \
    _Pragma("omp single") dsh = 0.0;                                           \
    _Pragma("omp for schedule(static) reduction(+ : dsh)") for (int i = 0; i < neq; i++) \
      dsh += (A)[i] * (B)[i];                                                  \
    OUT = dsh;                                                                 \
    _Pragma("omp barrier")                                                     \
  } while (0)
    double d, di, alpha, beta, tmp;
#pragma omp for schedule(static)
    for (int i = 0; i < neq; i++) x[i] = 0.0;
#pragma omp for schedule(static)
    for (int i = 0; i < neq; i++) z[i] = b[i] * m[i];
    OMP_DOT(b, z, d);
    double cv = tol * sqrt(fabs(d));
    OMP_MATVEC(x, z);
#pragma omp for schedule(static)
    for (int i = 0; i < neq; i++) {
      r[i] = b[i] - z[i];
      z[i] = r[i] * m[i];
      p[i] = z[i];
    }
    OMP_DOT(r, z, d);
    int j;
    for (j = 1; j <= maxit; j++) {
      OMP_MATVEC(p, z);
      OMP_DOT(z, p, tmp);
      alpha = d / tmp;
#pragma omp for schedule(static)
      for (int i = 0; i < neq; i++) {
        x[i] = x[i] + alpha * p[i];
        r[i] = r[i] - alpha * z[i];
        z[i] = r[i] * m[i];
      }
      OMP_DOT(r, z, di);
      beta = di / d;
#pragma omp for schedule(static)
      for (int i = 0; i < neq; i++) p[i] = z[i] + beta * p[i];
      d = di;
      if (sqrt(fabs(d)) < cv) break;
    }
    OMP_MATVEC(x, z);
    double xkx, nx, nmr, nr;
    OMP_DOT(x, z, xkx);
    OMP_DOT(x, x, nx);
#pragma omp for schedule(static)
    for (int i = 0; i < neq; i++) {
      r[i] = b[i] - z[i];
      z[i] = r[i] * m[i];
    }
    OMP_DOT(r, z, nmr);
    OMP_DOT(r, r, nr);
#pragma omp single
    {
      conv_ok = j <= maxit;
      iters = conv_ok ? j : maxit;
      conv = cv;
      red[0] = xkx;
      st->norm_x = sqrt(nx);
      st->norm_m_r = sqrt(fabs(nmr));
      st->norm_r = sqrt(nr);
    }
#undef OMP_MATVEC
#undef OMP_DOT
  }
  st->iters = iters;
  st->status = conv_ok ? 0 : 1;
  st->xkx = red[0];
  st->conv = conv;
  free(ty);
  free(z);
  free(r);
  free(p);
  free(tbeg);
  free(tend);
  free(thgt);
}

/* -------------------------------------------------------------- GMRES */
void or_gmres(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
              const double* al, const double* m, const double* b, double* x, int k,
              double tol, int maxit, or_stats* st) {
  memset(st, 0, sizeof(*st));
  double* g = (double*)malloc(sizeof(double) * (size_t)neq * (size_t)(k + 1));
  double* h = (double*)calloc((size_t)(k + 1) * (size_t)k, sizeof(double));
  double* y = (double*)malloc(sizeof(double) * (size_t)k);
  double* c = (double*)malloc(sizeof(double) * (size_t)k);
  double* s = (double*)malloc(sizeof(double) * (size_t)k);
  double* e = (double*)malloc(sizeof(double) * (size_t)(k + 1));
#define G(i, j) g[(size_t)((j)-1) * neq + (i)-1]
#define H(i, j) h[(size_t)((j)-1) * (k + 1) + (i)-1]
  for (int i = 1; i <= neq; i++) {
    A1(x, i) = 0.0;
    G(i, 1) = A1(b, i) * A1(m, i);
  }
  double norm = sqrt(or_dot(&G(1, 1), &G(1, 1), neq));
  double econv = tol * norm;
  int nit = 0, ni = 0, l;
  for (l = 1; l <= maxit; l++) {
    or_matvec_csrcsym1(neq, ia, ja, ad, al, x, &G(1, 1));
    for (int i = 1; i <= neq; i++) G(i, 1) = (A1(b, i) - G(i, 1)) * A1(m, i);
    e[0] = sqrt(or_dot(&G(1, 1), &G(1, 1), neq));
    for (int i = 1; i <= neq; i++) G(i, 1) = G(i, 1) / e[0];
    ni = 0;
    for (int i = 1; i <= k; i++) {
      nit++;
      ni++;
      or_matvec_csrcsym1(neq, ia, ja, ad, al, &G(1, i), &G(1, i + 1));
      for (int j = 1; j <= neq; j++) G(j, i + 1) = G(j, i + 1) * A1(m, j);
      for (int j = 1; j <= i; j++) {
        double beta = or_dot(&G(1, i + 1), &G(1, j), neq);
        for (int ic = 1; ic <= neq; ic++) G(ic, i + 1) = G(ic, i + 1) - beta * G(ic, j);
        H(j, i) = beta;
      }
      double nrm = sqrt(or_dot(&G(1, i + 1), &G(1, i + 1), neq));
      H(i + 1, i) = nrm;
      for (int ic = 1; ic <= neq; ic++) G(ic, i + 1) = G(ic, i + 1) / nrm;
      for (int j = 1; j <= i - 1; j++) {
        double a1 = c[j - 1] * H(j, i) + s[j - 1] * H(j + 1, i);
        double a2 = -s[j - 1] * H(j, i) + c[j - 1] * H(j + 1, i);
        H(j, i) = a1;
        H(j + 1, i) = a2;
      }
      double r = sqrt(H(i, i) * H(i, i) + H(i + 1, i) * H(i + 1, i));
      c[i - 1] = H(i, i) / r;
      s[i - 1] = H(i + 1, i) / r;
      H(i, i) = r;
      H(i + 1, i) = 0.0;
      e[i] = -s[i - 1] * e[i - 1];
      e[i - 1] = c[i - 1] * e[i - 1];
      if (fabs(e[i]) <= econv) break;
    }
    y[ni - 1] = e[ni - 1] / H(ni, ni);
    for (int i = ni - 1; i >= 1; i--) {
      y[i - 1] = 0.0;
      for (int j = i + 1; j <= ni; j++) y[i - 1] = y[i - 1] - H(i, j) * y[j - 1];
      y[i - 1] = (y[i - 1] + e[i - 1]) / H(i, i);
    }
    for (int i = 1; i <= neq; i++)
      for (int j = 1; j <= ni; j++) A1(x, i) = A1(x, i) + y[j - 1] * G(i, j);
    if (fabs(e[ni]) <= econv) break;
  }
  st->status = (fabs(e[ni]) <= econv) ? 0 : 1;
  st->iters = nit;
  or_matvec_csrcsym1(neq, ia, ja, ad, al, x, &G(1, 1));
  st->xkx = or_dot(x, &G(1, 1), neq);
  st->norm_x = sqrt(or_dot(x, x, neq));
  for (int i = 1; i <= neq; i++) G(i, 2) = A1(b, i) - G(i, 1);
  st->norm_r = sqrt(or_dot(&G(1, 2), &G(1, 2), neq));
  st->norm_m_r = fabs(e[ni]);
  st->conv = econv;
#undef G
#undef H
  free(g);
  free(h);
  free(y);
  free(c);
  free(s);
  free(e);
}

/* ----------------------------------------------------------- BiCGSTAB */
void or_pbicgstab(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                  const double* al, const double* m, const double* b, double* x,
                  double tol, int maxit, or_stats* st) {
  memset(st, 0, sizeof(*st));
  for (int i = 0; i < neq; i++)
    if (ad[i] == 0.0) {
      st->status = 2;
      return;
    }
  size_t n = (size_t)neq;
  double* t = (double*)malloc(sizeof(double) * n);
  double* v = (double*)malloc(sizeof(double) * n);
  double* r = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* r0 = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < neq; i++) x[i] = 0.0;
  double d = or_dot(b, b, neq);
  double conv = tol * sqrt(fabs(d));
  or_matvec_csrcsym1(neq, ia, ja, ad, al, x, z);
  for (int i = 0; i < neq; i++) {
    r0[i] = b[i] - z[i];
    p[i] = r0[i];
    r[i] = r0[i];
    z[i] = p[i] * m[i];
  }
  int j, ok = 0;
  for (j = 1; j <= maxit; j++) {
    or_matvec_csrcsym1(neq, ia, ja, ad, al, z, v);
    double rr0 = or_dot(r, r0, neq);
    double alpha = rr0 / or_dot(v, r0, neq);
    for (int i = 0; i < neq; i++) {
      x[i] = x[i] + alpha * z[i];
      r[i] = r[i] - alpha * v[i];
      z[i] = r[i] * m[i];
    }
    d = or_dot(r, r, neq);
    if (sqrt(fabs(d)) < conv) {
      ok = 1;
      break;
    }
    or_matvec_csrcsym1(neq, ia, ja, ad, al, z, t);
    double w = or_dot(t, r, neq) / or_dot(t, t, neq);
    for (int i = 0; i < neq; i++) {
      x[i] = x[i] + w * z[i];
      r[i] = r[i] - w * t[i];
    }
    d = or_dot(r, r, neq);
    if (sqrt(fabs(d)) < conv) {
      ok = 1;
      break;
    }
    double beta = (or_dot(r, r0, neq) / rr0) * (alpha / w);
    for (int i = 0; i < neq; i++) {
      p[i] = r[i] + beta * (p[i] - w * v[i]);
      z[i] = p[i] * m[i];
    }
  }
  st->status = ok ? 0 : 1;
  st->iters = ok ? j : maxit;
  or_matvec_csrcsym1(neq, ia, ja, ad, al, x, z);
  st->xkx = or_dot(x, z, neq);
  st->norm_x = sqrt(or_dot(x, x, neq));
  for (int i = 0; i < neq; i++) r[i] = b[i] - z[i];
  st->norm_r = sqrt(or_dot(r, r, neq));
  st->conv = conv;
  free(t);
  free(v);
  free(r);
  free(p);
  free(z);
  free(r0);
}

/* ------------------------------------------------------------- colouring */
int or_colormsh(const int32_t* ix, int nnode, int numel, int nen, int nenv, int32_t* icolor) {
  /* nodegrade + elmincid (Colormsh.f:65-150) */
  int32_t* nincid = (int32_t*)calloc((size_t)nnode + 1, sizeof(int32_t));
  int maxgrade = 0;
  for (int i = 1; i <= numel; i++)
    for (int j = 1; j <= nenv; j++) {
      int node = A2(ix, nen + 1, j, i);
      if (node > 0 && ++nincid[node] > maxgrade) maxgrade = nincid[node];
    }
  int32_t* incid = (int32_t*)malloc(sizeof(int32_t) * (size_t)maxgrade * (size_t)(nnode + 1));
  memset(nincid, 0, sizeof(int32_t) * ((size_t)nnode + 1));
  for (int i = 1; i <= numel; i++)
    for (int j = 1; j <= nenv; j++) {
      int node = A2(ix, nen + 1, j, i);
      if (node > 0) {
        int ipos = ++nincid[node];
        incid[(size_t)node * maxgrade + ipos - 1] = i;
      }
    }
  /* colormsh (Colormsh.f:151-247) */
  for (int i = 1; i <= numel; i++) A1(icolor, i) = 0;
  int numcolors = -1;
  for (int c = 1; c <= 200; c++) {
    for (int node = 1; node <= nnode; node++) {
      int l = nincid[node], found = 0;
      for (int q = 0; q < l; q++)
        if (A1(icolor, incid[(size_t)node * maxgrade + q]) == c) found = 1;
      if (found) continue;
      int first = 0;
      for (int q = 0; q < l; q++) {
        int ik = incid[(size_t)node * maxgrade + q];
        if (A1(icolor, ik) == 0) first = ik;
      }
      if (first == 0) continue;
      A1(icolor, first) = c;
      for (int j = 1; j <= nenv; j++) {
        int ixj = A2(ix, nen + 1, j, first);
        if (ixj <= 0) continue;
        for (int q = 0; q < nincid[ixj]; q++) {
          int ik = incid[(size_t)ixj * maxgrade + q];
          if (A1(icolor, ik) == 0) A1(icolor, ik) = -1;
        }
      }
    }
    int quit = 1;
    for (int n = 1; n <= numel; n++)
      if (A1(icolor, n) == -1) {
        A1(icolor, n) = 0;
        quit = 0;
      }
    if (quit) {
      numcolors = c;
      break;
    }
  }
  free(incid);
  free(nincid);
  return numcolors;
}
