// Element stiffness and load assembly on gfx950 (Pform.f:1-254, Assbly.f:66-146,
// Elmlib02.f elmt02..07, Elmlib.f shape functions and Jacobians).
//
// Stiffness (isw = 2), one launch per colour.  hex8 with u = 0 (the
// assembly of K: b - K u = b) takes k_stiffness_hex8: 16 elements per
// workgroup, Gauss-point geometry formed once and shared in LDS, one lane per
// (element, node pair, row) forming only the entries the CSRC stores.  The
// other cases take k_stiffness: one 64-lane wavefront per element, every
// lane evaluating the Jacobian of each Gauss point itself, the element matrix
// s(nst,nst) in LDS and p = s u in the reference's order (Elmlib.f:336-369).
// Either way the operations and their order are the reference's (FP
// contraction off), so every entry is bit-identical to its element matrix,
// and the entries are added to ad/al/b through the element's slot map with
// plain read-modify-writes (elements of one colour share no node).
// Loads (isw = 4): one lane per element of a colour, face/edge tractions and
// self weight as Elmlib02.f computes them.
#include <cmath>

#include "elastic_internal.h"

#pragma clang fp contract(off)

namespace elastic {

// element kinds
enum { K_TRIA3 = 0, K_QUAD4 = 1, K_TET4 = 2, K_HEX8 = 3 };

template <int KIND>
struct ElemTraits;
template <>
struct ElemTraits<K_TRIA3> { static constexpr int NEN = 3, NDF = 2, NDM = 2; };
template <>
struct ElemTraits<K_QUAD4> { static constexpr int NEN = 4, NDF = 2, NDM = 2; };
template <>
struct ElemTraits<K_TET4> { static constexpr int NEN = 4, NDF = 3, NDM = 3; };
template <>
struct ElemTraits<K_HEX8> { static constexpr int NEN = 8, NDF = 3, NDM = 3; };

__constant__ double kPG2[2] = {-.577350269189626, .577350269189626};
constexpr double kDIV6 = 0.166666666666667;
constexpr double kDIV144 = 0.364444444444444e-02;
constexpr double kJZERO = 1.0e-14;

struct AsmArgs {
  const int32_t* ix;     // (nen+1) per element
  const double* x;       // ndm per node
  const int32_t* id;     // equation numbers, ndf per node
  const int32_t* ie;     // element type per material
  const double* e;       // 15 per material
  const int32_t* eload;  // 7 per element
  const int32_t* order;  // elements (0-based) grouped by colour
  const int32_t* slots;  // npair per element
  const int32_t* ltype;  // per load
  const int32_t* lnparc;
  const double* lf;      // 3 per load
  const double* u;       // ndf per node (may be null)
  double *ad, *al, *b;
  int32_t* err;          // first element (1-based) with a non-positive Jacobian
  int nen, ndf, ndm, neq, nloads;
  double g0, g1, g2;     // gravity (Main.f:205-207)
};

// ---------------------------------------------------------- shape functions
__device__ inline void sfquad4_d(double r, double s, double* h, double* hx, double* hy) {
  double opr = 1.0 + r, ops = 1.0 + s, omr = 1.0 - r, oms = 1.0 - s;
  h[0] = opr * ops * 0.25;
  h[1] = omr * ops * 0.25;
  h[2] = omr * oms * 0.25;
  h[3] = opr * oms * 0.25;
  hx[0] = ops * 0.25;
  hx[1] = -ops * 0.25;
  hx[2] = -oms * 0.25;
  hx[3] = oms * 0.25;
  hy[0] = opr * 0.25;
  hy[1] = omr * 0.25;
  hy[2] = -omr * 0.25;
  hy[3] = -opr * 0.25;
}

__device__ inline void sfhexa8_d(double r, double s, double t, double* h, double* hx, double* hy,
                                 double* hz) {
  double pR = 1.0 + r, pS = 1.0 + s, pT = 1.0 + t, mR = 1.0 - r, mS = 1.0 - s, mT = 1.0 - t;
  h[0] = pR * pS * pT * 0.125;
  h[1] = mR * pS * pT * 0.125;
  h[2] = mR * mS * pT * 0.125;
  h[3] = pR * mS * pT * 0.125;
  h[4] = pR * pS * mT * 0.125;
  h[5] = mR * pS * mT * 0.125;
  h[6] = mR * mS * mT * 0.125;
  h[7] = pR * mS * mT * 0.125;
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

// jacob2d_m (Elmlib.f:1769-1852); x has leading dimension ldx
__device__ inline bool jacob2d_d(double* hx, double* hy, const double* x, int nen, int ldx,
                                 double* det) {
  double j00 = 0.0, j01 = 0.0, j10 = 0.0, j11 = 0.0;
  for (int k = 0; k < nen; ++k) {
    j00 = j00 + hx[k] * x[k * ldx + 0];
    j10 = j10 + hy[k] * x[k * ldx + 0];
  }
  for (int k = 0; k < nen; ++k) {
    j01 = j01 + hx[k] * x[k * ldx + 1];
    j11 = j11 + hy[k] * x[k * ldx + 1];
  }
  *det = j00 * j11 - j10 * j01;
  if (*det <= kJZERO) return false;
  const double d = *det;
  const double i11 = j11 / d, i12 = -j01 / d, i21 = -j10 / d, i22 = j00 / d;
  for (int k = 0; k < nen; ++k) {
    double a = hx[k], b = hy[k];
    hx[k] = i11 * a + i12 * b;
    hy[k] = i21 * a + i22 * b;
  }
  return true;
}

// jacob3d_m (Elmlib.f:1942-2037)
__device__ inline bool jacob3d_d(double* hx, double* hy, double* hz, const double* x, int nen,
                                 double* det) {
  double xj[3][3];
  for (int j = 0; j < 3; ++j) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = 0; k < nen; ++k) {
      a = a + hx[k] * x[k * 3 + j];
      b = b + hy[k] * x[k * 3 + j];
      c = c + hz[k] * x[k * 3 + j];
    }
    xj[0][j] = a;
    xj[1][j] = b;
    xj[2][j] = c;
  }
#define XJ(p, q) xj[(p)-1][(q)-1]
  *det = XJ(1, 1) * XJ(2, 2) * XJ(3, 3) + XJ(1, 2) * XJ(2, 3) * XJ(3, 1) +
         XJ(1, 3) * XJ(2, 1) * XJ(3, 2) - XJ(3, 1) * XJ(2, 2) * XJ(1, 3) -
         XJ(2, 1) * XJ(1, 2) * XJ(3, 3) - XJ(1, 1) * XJ(3, 2) * XJ(2, 3);
  if (*det <= kJZERO) return false;
  const double d = *det;
  const double i11 = (XJ(2, 2) * XJ(3, 3) - XJ(2, 3) * XJ(3, 2)) / d;
  const double i21 = -(XJ(2, 1) * XJ(3, 3) - XJ(2, 3) * XJ(3, 1)) / d;
  const double i31 = (XJ(2, 1) * XJ(3, 2) - XJ(2, 2) * XJ(3, 1)) / d;
  const double i12 = -(XJ(1, 2) * XJ(3, 3) - XJ(1, 3) * XJ(3, 2)) / d;
  const double i22 = (XJ(1, 1) * XJ(3, 3) - XJ(1, 3) * XJ(3, 1)) / d;
  const double i32 = -(XJ(1, 1) * XJ(3, 2) - XJ(1, 2) * XJ(3, 1)) / d;
  const double i13 = (XJ(1, 2) * XJ(2, 3) - XJ(1, 3) * XJ(2, 2)) / d;
  const double i23 = -(XJ(1, 1) * XJ(2, 3) - XJ(1, 3) * XJ(2, 1)) / d;
  const double i33 = (XJ(1, 1) * XJ(2, 2) - XJ(1, 2) * XJ(2, 1)) / d;
#undef XJ
  for (int k = 0; k < nen; ++k) {
    double a = hx[k], b = hy[k], c = hz[k];
    hx[k] = i11 * a + i12 * b + i13 * c;
    hy[k] = i21 * a + i22 * b + i23 * c;
    hz[k] = i31 * a + i32 * b + i33 * c;
  }
  return true;
}

// jtetra4 (Elmlib.f:2636-2713)
__device__ inline bool jtetra4_d(const double* x, double* hx, double* hy, double* hz, double* det,
                                 bool afl) {
#define X(i, n) x[((n)-1) * 3 + (i)-1]
  const double j11 = X(1, 1) - X(1, 4), j12 = X(2, 1) - X(2, 4), j13 = X(3, 1) - X(3, 4);
  const double j21 = X(1, 2) - X(1, 4), j22 = X(2, 2) - X(2, 4), j23 = X(3, 2) - X(3, 4);
  const double j31 = X(1, 3) - X(1, 4), j32 = X(2, 3) - X(2, 4), j33 = X(3, 3) - X(3, 4);
#undef X
  *det = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j31 * j22 * j13 -
         j12 * j21 * j33 - j11 * j32 * j23;
  if (*det <= kJZERO) return false;
  if (afl) {
    const double di = 1.0 / *det;
    hx[0] = (j22 * j33 - j23 * j32) * di;
    hx[1] = (j13 * j32 - j12 * j33) * di;
    hx[2] = (j12 * j23 - j13 * j22) * di;
    hy[0] = (j23 * j31 - j21 * j33) * di;
    hy[1] = (j11 * j33 - j13 * j31) * di;
    hy[2] = (j13 * j21 - j11 * j23) * di;
    hz[0] = (j21 * j32 - j22 * j31) * di;
    hz[1] = (j12 * j31 - j11 * j32) * di;
    hz[2] = (j11 * j22 - j12 * j21) * di;
    hx[3] = -hx[0] - hx[1] - hx[2];
    hy[3] = -hy[0] - hy[1] - hy[2];
    hz[3] = -hz[0] - hz[1] - hz[2];
  }
  return true;
}

// ------------------------------------------------------- stiffness entries
// s(ndf*a+r, ndf*bn+c) contribution of one Gauss point (3-D, a,b,c of the
// constitutive matrix; Elmlib02.f:1715-1777 hex8, 1353-1383 tet4)
__device__ inline double term3(bool hex, int r, int c, double xi, double yi, double zi, double xj,
                               double yj, double zj, double b, double cc, double wt1) {
  if (r == 0 && c == 0)
    return hex ? (xi * xj + cc * (yi * yj + zi * zj)) * wt1
               : (xi * xj + cc * yi * yj + cc * zi * zj) * wt1;
  if (r == 1 && c == 1)
    return hex ? (yi * yj + cc * (xi * xj + zi * zj)) * wt1
               : (yi * yj + cc * xi * xj + cc * zi * zj) * wt1;
  if (r == 2 && c == 2)
    return hex ? (zi * zj + cc * (yi * yj + xi * xj)) * wt1
               : (zi * zj + cc * yi * yj + cc * xi * xj) * wt1;
  if (r == 0 && c == 1) return (b * xi * yj + cc * yi * xj) * wt1;
  if (r == 0 && c == 2) return (b * xi * zj + cc * zi * xj) * wt1;
  if (r == 1 && c == 0) return (b * yi * xj + cc * xi * yj) * wt1;
  if (r == 1 && c == 2) return (b * yi * zj + cc * zi * yj) * wt1;
  if (r == 2 && c == 0) return (b * zi * xj + cc * xi * zj) * wt1;
  return (b * zi * yj + cc * yi * zj) * wt1;  // (2,1)
}

// 2-D (Elmlib02.f:207-224 tria3, 780-800 quad4)
__device__ inline double term2(int r, int c, double xi, double yi, double xj, double yj, double d11,
                               double d12, double d22, double d33, double wt1) {
  if (r == 0 && c == 0) return (xi * d11 * xj + yi * d33 * yj) * wt1;
  if (r == 0 && c == 1) return (xi * d12 * yj + yi * d33 * xj) * wt1;
  if (r == 1 && c == 0) return (yi * d12 * xj + xi * d33 * yj) * wt1;
  return (yi * d22 * yj + xi * d33 * xj) * wt1;
}

template <int KIND>
__global__ __launch_bounds__(64) void k_stiffness(AsmArgs a, int c0, int lhs, int rhs) {
  using TR = ElemTraits<KIND>;
  constexpr int NEN = TR::NEN, NDF = TR::NDF, NDM = TR::NDM, NST = NEN * NDF;
  constexpr int NPAIR = NST * (NST + 1) / 2, NENT = NST * NST;
  constexpr int PER = (NENT + 63) / 64;
  __shared__ double xl[NEN * NDM], ul[NST], s[NENT], pl[NST];
  __shared__ int32_t ldv[NST];
  __shared__ int bad;
  const int el = a.order[c0 + blockIdx.x];
  const int tid = threadIdx.x;
  const int32_t* ixe = a.ix + (size_t)el * (NEN + 1);
  if (tid == 0) bad = 0;
  if (tid < NEN * NDM) {
    int n = tid / NDM, d = tid % NDM;
    xl[tid] = a.x[(size_t)(ixe[n] - 1) * NDM + d];
  }
  if (tid < NST) {
    int n = tid / NDF, d = tid % NDF;
    ldv[tid] = a.id[(size_t)(ixe[n] - 1) * NDF + d];
    ul[tid] = a.u ? a.u[(size_t)(ixe[n] - 1) * NDF + d] : 0.0;
  }
  const int ma = ixe[NEN];
  const int iel = a.ie[ma - 1];
  const double* em = a.e + (size_t)(ma - 1) * 15;
  const double ym = em[0], ps = em[1];
  __syncthreads();
  // constitutive constants (Elmlib02.f:1693-1702 and the 2-D variants)
  double ca = 0, cb = 0, cc = 0, d11 = 0, d12 = 0, d22 = 0, d33 = 0, thic = 1.0;
  const bool stress = (iel == 3 || iel == 5);
  if (!stress) {
    const double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
    ca = (ym * a1) / (a3 * a2);
    cb = ps / a1;
    cc = 0.5 * (a2 / a1);
  } else {
    thic = em[2];
    if (thic == 0) thic = 1.0;
    const double a1 = 1.0 - ps, a2 = 1.0 - ps * ps;
    ca = ym / a2;
    cb = ps;
    cc = 0.5 * a1;
  }
  d11 = ca;
  d12 = ca * cb;
  d22 = ca;
  d33 = ca * cc;
  double acc[PER];
  for (int q = 0; q < PER; ++q) acc[q] = 0.0;
  double hx[NEN], hy[NEN], hz[NEN], h[NEN];
  if (KIND == K_HEX8) {
    for (int lz = 0; lz < 2; ++lz)
      for (int ly = 0; ly < 2; ++ly)
        for (int lx = 0; lx < 2; ++lx) {
          double det;
          sfhexa8_d(kPG2[lx], kPG2[ly], kPG2[lz], h, hx, hy, hz);
          if (!jacob3d_d(hx, hy, hz, xl, NEN, &det)) bad = 1;
          const double w = 1.0 * 1.0 * 1.0 * det;
          const double wt1 = w * ca;
#pragma unroll
          for (int q = 0; q < PER; ++q) {
            const int en = tid + 64 * q;
            if (en < NENT) {
              const int i = en % NST, j = en / NST;
              const int na = i / NDF, r = i % NDF, nb = j / NDF, c = j % NDF;
              acc[q] = acc[q] + term3(true, r, c, hx[na], hy[na], hz[na], hx[nb], hy[nb], hz[nb],
                                      cb, cc, wt1);
            }
          }
        }
  } else if (KIND == K_TET4) {
    double det;
    if (!jtetra4_d(xl, hx, hy, hz, &det, true)) bad = 1;
    const double wt1 = ca * det * kDIV6;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int en = tid + 64 * q;
      if (en < NENT) {
        const int i = en % NST, j = en / NST;
        const int na = i / NDF, r = i % NDF, nb = j / NDF, c = j % NDF;
        acc[q] = term3(false, r, c, hx[na], hy[na], hz[na], hx[nb], hy[nb], hz[nb], cb, cc, wt1);
      }
    }
  } else if (KIND == K_QUAD4) {
    for (int ly = 0; ly < 2; ++ly)
      for (int lx = 0; lx < 2; ++lx) {
        double det;
        sfquad4_d(kPG2[lx], kPG2[ly], h, hx, hy);
        if (!jacob2d_d(hx, hy, xl, NEN, NDM, &det)) bad = 1;
        double w = 1.0 * 1.0 * det;
        if (stress) w = w * thic;
        const double wt1 = w;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int en = tid + 64 * q;
          if (en < NENT) {
            const int i = en % NST, j = en / NST;
            const int na = i / NDF, r = i % NDF, nb = j / NDF, c = j % NDF;
            acc[q] = acc[q] + term2(r, c, hx[na], hy[na], hx[nb], hy[nb], d11, d12, d22, d33, wt1);
          }
        }
      }
  } else {  // K_TRIA3 (Elmlib02.f:143-229: determinant test against 0)
    const double xj11 = xl[0] - xl[4], xj12 = xl[1] - xl[5];
    const double xj21 = xl[2] - xl[4], xj22 = xl[3] - xl[5];
    const double det = xj11 * xj22 - xj12 * xj21;
    if (det <= 0.0) bad = 1;
    const double ji11 = xj22 / det, ji12 = -xj12 / det, ji21 = -xj21 / det, ji22 = xj11 / det;
    hx[0] = ji11;
    hx[1] = ji12;
    hx[2] = -ji11 - ji12;
    hy[0] = ji21;
    hy[1] = ji22;
    hy[2] = -ji21 - ji22;
    const double wt1 = 0.5 * det;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int en = tid + 64 * q;
      if (en < NENT) {
        const int i = en % NST, j = en / NST;
        const int na = i / NDF, r = i % NDF, nb = j / NDF, c = j % NDF;
        acc[q] = term2(r, c, hx[na], hy[na], hx[nb], hy[nb], d11, d12, d22, d33, wt1);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int en = tid + 64 * q;
    if (en < NENT) s[en] = acc[q];  // s(i,j) at i + NST*j (column-major)
  }
  __syncthreads();
  if (bad) {
    if (tid == 0) a.err[0] = el + 1;  // any offending element is reported
    return;
  }
  // p = s u (Elmlib.f:336-369), b -= p
  if (rhs && tid < NST) {
    double p = 0.0;
    for (int j = 0; j < NST; ++j) p = p + s[j * NST + tid] * ul[j];
    pl[tid] = p;
    const int k = ldv[tid];
    if (k > 0 && k <= a.neq) a.b[k - 1] = a.b[k - 1] - p;
  }
  if (!lhs) return;
  const int32_t* sl = a.slots + (size_t)el * NPAIR;
  for (int pp = tid; pp < NPAIR; pp += 64) {
    const int32_t t = sl[pp];
    if (t == -1) continue;
    // pair pp <-> (i >= j): i = floor((sqrt(8pp+1)-1)/2)
    int i = (int)((sqrtf(8.0f * pp + 1.0f) - 1.0f) * 0.5f);
    while ((i + 1) * (i + 2) / 2 <= pp) ++i;
    while (i * (i + 1) / 2 > pp) --i;
    const int j = pp - i * (i + 1) / 2;
    if (t <= -2) {
      const int k = -t - 2;
      a.ad[k] = a.ad[k] + s[i * NST + i];
    } else {
      // the row is the dof with the larger equation number (Assbly.f:116-143)
      const double v = ldv[i] > ldv[j] ? s[j * NST + i] : s[i * NST + j];  // s(i,j) : s(j,i)
      a.al[t] = a.al[t] + v;
    }
  }
}

// Stiffness of hex8 elements (elmt07_mec isw 2, Elmlib02.f:1691-1785), kHexE
// elements per 256-lane workgroup.  The shape-function derivatives and the
// weight det * a of each (element, Gauss point) are formed once, by one lane
// (sfhexa8_m + jacob3d_m, Elmlib.f:1151-1259, 1942-2037), and kept in LDS.
// Each lane then forms the three entries (r, 0..2) of one node pair (na >= nb)
// of one element that the CSRC stores: for every pair of dofs only the entry
// of the row with the larger equation number, s(i,j) or s(j,i), the way
// csr (Assbly.f:116-143) picks it, summed over the Gauss points in the
// reference's order with its operations (FP contraction off), so every value
// is bit-identical to the reference's element matrix.  The sums go to ad/al
// through the element's slot map (elements of one colour share no node).
// b -= s u needs the whole matrix and is left to k_stiffness.
constexpr int kHexE = 16;
constexpr int kHexSub = 108;  // (na >= nb) node pairs x 3 rows

__global__ __launch_bounds__(256) void k_stiffness_hex8(AsmArgs a, int c0, int cnt) {
  constexpr int NEN = 8, NST = 24, NPAIR = NST * (NST + 1) / 2;
  __shared__ double dX[kHexE][8][3][NEN];  // derivative d of node n at Gauss point g
  __shared__ double wt[kHexE][8];          // det * a at Gauss point g
  __shared__ double xl[kHexE][NEN * 3];
  __shared__ double cbs[kHexE], ccs[kHexE], cas[kHexE];
  __shared__ int32_t ldv[kHexE][NST];
  __shared__ int32_t elno[kHexE];
  __shared__ int bad[kHexE];
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * kHexE;
  const int ne = min(kHexE, cnt - e0);
  // element data: coordinates and equation numbers
  for (int q = tid; q < ne * NST; q += 256) {
    const int e = q / NST, k = q - e * NST;
    const int el = a.order[c0 + e0 + e];
    const int32_t* ixe = a.ix + (size_t)el * (NEN + 1);
    const size_t g = (size_t)(ixe[k / 3] - 1) * 3 + k % 3;
    xl[e][k] = a.x[g];
    ldv[e][k] = a.id[g];
    if (k == 0) {
      elno[e] = el;
      bad[e] = 0;
      const double* em = a.e + (size_t)(ixe[NEN] - 1) * 15;
      const double ym = em[0], ps = em[1];
      const double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
      cas[e] = (ym * a1) / (a3 * a2);
      cbs[e] = ps / a1;
      ccs[e] = 0.5 * (a2 / a1);
    }
  }
  __syncthreads();
  // one lane per (element, Gauss point): derivatives and weight
  if (tid < ne * 8) {
    const int e = tid >> 3, gp = tid & 7;
    const int lx = gp & 1, ly = (gp >> 1) & 1, lz = gp >> 2;
    double h[NEN], hx[NEN], hy[NEN], hz[NEN], det;
    sfhexa8_d(kPG2[lx], kPG2[ly], kPG2[lz], h, hx, hy, hz);
    if (!jacob3d_d(hx, hy, hz, xl[e], NEN, &det)) bad[e] = 1;
    const double w = 1.0 * 1.0 * 1.0 * det;
    wt[e][gp] = w * cas[e];
#pragma unroll
    for (int n = 0; n < NEN; ++n) {
      dX[e][gp][0][n] = hx[n];
      dX[e][gp][1][n] = hy[n];
      dX[e][gp][2][n] = hz[n];
    }
  }
  __syncthreads();
  for (int q = tid; q < ne * kHexSub; q += 256) {
    const int e = q / kHexSub, sub = q - e * kHexSub;
    if (bad[e]) continue;
    const int pr = sub / 3, r = sub - 3 * pr;  // node pair pr, row r of na
    int na = (int)((sqrtf(8.0f * pr + 1.0f) - 1.0f) * 0.5f);  // pr = na(na+1)/2 + nb
    while ((na + 1) * (na + 2) / 2 <= pr) ++na;
    while (na * (na + 1) / 2 > pr) --na;
    const int nb = pr - na * (na + 1) / 2;
    const int i = 3 * na + r;
    const int el = elno[e];
    const int32_t* sl = a.slots + (size_t)el * NPAIR + i * (i + 1) / 2 + 3 * nb;
    const int cmax = na == nb ? r : 2;  // lower triangle: j = 3 nb + c <= i
    int32_t t[3];
    bool up[3];  // the entry is s(j,i): the row is dof j
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      t[c] = c <= cmax ? sl[c] : -1;
      up[c] = ldv[e][i] < ldv[e][3 * nb + c];
    }
    if (t[0] == -1 && t[1] == -1 && t[2] == -1) continue;
    const double cb = cbs[e], cc = ccs[e];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int gp = 0; gp < 8; ++gp) {
      const double xi = dX[e][gp][0][na], yi = dX[e][gp][1][na], zi = dX[e][gp][2][na];
      const double xj = dX[e][gp][0][nb], yj = dX[e][gp][1][nb], zj = dX[e][gp][2][nb];
      const double w1 = wt[e][gp];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (t[c] == -1) continue;
        acc[c] = acc[c] + (up[c] ? term3(true, c, r, xj, yj, zj, xi, yi, zi, cb, cc, w1)
                                 : term3(true, r, c, xi, yi, zi, xj, yj, zj, cb, cc, w1));
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (t[c] == -1) continue;
      if (t[c] <= -2) {
        const int k = -t[c] - 2;
        a.ad[k] = a.ad[k] + acc[c];
      } else {
        a.al[t[c]] = a.al[t[c]] + acc[c];
      }
    }
  }
  if (tid < ne && bad[tid]) a.err[0] = elno[tid] + 1;  // any offending element is reported
}

// ---------------------------------------------------------- loads (isw = 4)
__device__ inline void tload_vec(const AsmArgs& a, int nc, double* ff) {
  ff[0] = ff[1] = ff[2] = 0.0;
  if (nc <= 0 || nc > a.nloads) return;
  if (a.ltype[nc - 1] != 40) return;
  for (int q = 0; q < a.lnparc[nc - 1] && q < 3; ++q) ff[q] = a.lf[3 * (nc - 1) + q];
}

__device__ inline void rotmat_d(const double* x, double r[3][3]) {
  double v1x = x[3] - x[0], v1y = x[4] - x[1], v1z = x[5] - x[2];
  double v2x = x[6] - x[0], v2y = x[7] - x[1], v2z = x[8] - x[2];
  double v3x = v1y * v2z - v1z * v2y, v3y = v1z * v2x - v1x * v2z, v3z = v1x * v2y - v1y * v2x;
  v2x = v1z * v3y - v1y * v3z;
  v2y = v1x * v3z - v1z * v3x;
  v2z = v1y * v3x - v1x * v3y;
  double n1 = sqrt(v1x * v1x + v1y * v1y + v1z * v1z);
  double n2 = sqrt(v2x * v2x + v2y * v2y + v2z * v2z);
  double n3 = sqrt(v3x * v3x + v3y * v3y + v3z * v3z);
  r[0][0] = v1x / n1, r[0][1] = v1y / n1, r[0][2] = v1z / n1;
  r[1][0] = v2x / n2, r[1][1] = v2y / n2, r[1][2] = v2z / n2;
  r[2][0] = v3x / n3, r[2][1] = v3y / n3, r[2][2] = v3z / n3;
}
__device__ inline void rot_d(double* x, double r[3][3]) {
  double x1 = r[0][0] * x[0] + r[0][1] * x[1] + r[0][2] * x[2];
  double x2 = r[1][0] * x[0] + r[1][1] * x[1] + r[1][2] * x[2];
  double x3 = r[2][0] * x[0] + r[2][1] * x[1] + r[2][2] * x[2];
  x[0] = x1, x[1] = x2, x[2] = x3;
}

template <int KIND>
__global__ __launch_bounds__(64) void k_loads(AsmArgs a, int c0, int cnt) {
  using TR = ElemTraits<KIND>;
  constexpr int NEN = TR::NEN, NDF = TR::NDF, NDM = TR::NDM, NST = NEN * NDF;
  const int q0 = blockIdx.x * 64 + threadIdx.x;
  if (q0 >= cnt) return;
  const int el = a.order[c0 + q0];
  const int32_t* ixe = a.ix + (size_t)el * (NEN + 1);
  const int32_t* iq = a.eload + (size_t)el * 7;
  const int ma = ixe[NEN];
  const double* em = a.e + (size_t)(ma - 1) * 15;
  double x[NEN * NDM], p[NST];
  for (int n = 0; n < NEN; ++n)
    for (int d = 0; d < NDM; ++d) x[n * NDM + d] = a.x[(size_t)(ixe[n] - 1) * NDM + d];
  for (int i = 0; i < NST; ++i) p[i] = 0.0;
  bool any = false, bad = false;
  if (KIND == K_HEX8 || KIND == K_TET4) {
    double density = em[2];
    if (density != 0.0) {
      any = true;
      density = em[2] * 1.0e-06;
      const double gl[3] = {a.g0 * density, a.g1 * density, a.g2 * density};
      double h[8], hx[8], hy[8], hz[8], det;
      if (KIND == K_HEX8) {
        for (int lz = 0; lz < 2; ++lz)
          for (int ly = 0; ly < 2; ++ly)
            for (int lx = 0; lx < 2; ++lx) {
              sfhexa8_d(kPG2[lx], kPG2[ly], kPG2[lz], h, hx, hy, hz);
              if (!jacob3d_d(hx, hy, hz, x, 8, &det)) bad = true;
              const double w = 1.0 * 1.0 * 1.0 * det;
              for (int i = 0; i < 8; ++i) {
                const double wt1 = h[i] * w;
                p[3 * i] = p[3 * i] - wt1 * gl[0];
                p[3 * i + 1] = p[3 * i + 1] - wt1 * gl[1];
                p[3 * i + 2] = p[3 * i + 2] - wt1 * gl[2];
              }
            }
      } else {
        if (!jtetra4_d(x, hx, hy, hz, &det, false)) bad = true;
        const double wt1 = det * kDIV144;
        for (int i = 0; i < 4; ++i) {
          p[3 * i] = -wt1 * gl[0];
          p[3 * i + 1] = -wt1 * gl[1];
          p[3 * i + 2] = -wt1 * gl[2];
        }
      }
    }
  }
  if (KIND == K_HEX8) {
    const int face[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 4, 5, 1},
                            {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 5, 6, 2}};
    for (int j = 0; j < 6; ++j) {
      if (iq[j] <= 0) continue;
      any = true;
      double xf[12], r[3][3], ff[3];
      for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) xf[3 * i + d] = x[3 * face[j][i] + d];
      rotmat_d(xf, r);
      for (int i = 0; i < 4; ++i) rot_d(xf + 3 * i, r);
      tload_vec(a, iq[j], ff);
      for (int ly = 0; ly < 2; ++ly)
        for (int lx = 0; lx < 2; ++lx) {
          double h[4], hx[4], hy[4], det;
          sfquad4_d(kPG2[lx], kPG2[ly], h, hx, hy);
          if (!jacob2d_d(hx, hy, xf, 4, 3, &det)) bad = true;
          const double w = 1.0 * 1.0 * det;
          for (int i = 0; i < 4; ++i) {
            const int no = face[j][i];
            const double wt1 = h[i] * w;
            p[3 * no] = p[3 * no] - ff[0] * wt1;
            p[3 * no + 1] = p[3 * no + 1] - ff[1] * wt1;
            p[3 * no + 2] = p[3 * no + 2] - ff[2] * wt1;
          }
        }
    }
  } else if (KIND == K_TET4) {
    const int face[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    for (int j = 0; j < 4; ++j) {
      if (iq[j] <= 0) continue;
      any = true;
      double xf[9], r[3][3], ff[3];
      for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) xf[3 * i + d] = x[3 * face[j][i] + d];
      rotmat_d(xf, r);
      for (int i = 0; i < 3; ++i) rot_d(xf + 3 * i, r);
      tload_vec(a, iq[j], ff);
      const double j11 = xf[0] - xf[6], j12 = xf[1] - xf[7];
      const double j21 = xf[3] - xf[6], j22 = xf[4] - xf[7];
      const double det = j11 * j22 - j12 * j21;
      if (det <= kJZERO) bad = true;
      const double wt1 = det * kDIV6;
      for (int i = 0; i < 3; ++i) {
        const int no = face[j][i];
        p[3 * no] = p[3 * no] - ff[0] * wt1;
        p[3 * no + 1] = p[3 * no + 1] - ff[1] * wt1;
        p[3 * no + 2] = p[3 * no + 2] - ff[2] * wt1;
      }
    }
  } else {
    const int nsides = NEN;  // tria3: 3 sides, quad4: 4 sides
    for (int j = 0; j < nsides; ++j) {
      if (iq[j] <= 0) continue;
      any = true;
      const int n0 = j, n1 = (j + 1) % NEN;
      const double dx = x[2 * n1] - x[2 * n0], dy = x[2 * n1 + 1] - x[2 * n0 + 1];
      double det = dx * dx + dy * dy;
      if (det <= 0.0) bad = true;
      det = sqrt(det);
      double ff[3];
      tload_vec(a, iq[j], ff);
      const double wt1 = 0.5 * det;
      for (int i = 0; i < 2; ++i) {
        const int no = i == 0 ? n0 : n1;
        p[2 * no] = p[2 * no] - ff[0] * wt1;
        p[2 * no + 1] = p[2 * no + 1] - ff[1] * wt1;
      }
    }
  }
  if (bad) {
    a.err[0] = el + 1;
    return;
  }
  if (!any) return;
  for (int i = 0; i < NST; ++i) {
    const int k = a.id[(size_t)(ixe[i / NDF] - 1) * NDF + (i % NDF)];
    if (k > 0 && k <= a.neq) a.b[k - 1] = a.b[k - 1] - p[i];
  }
}

// ------------------------------------------------------------- host driver
struct MeshDev {
  int nnode = 0, numel = 0, nen = 0, ndf = 0, ndm = 0, kind = -1, nloads = 0, ncolors = 0;
  std::vector<int> color_start;  // ncolors + 1 offsets into order
  int32_t *ix = nullptr, *id = nullptr, *ie = nullptr, *eload = nullptr, *order = nullptr,
          *slots = nullptr, *ltype = nullptr, *lnparc = nullptr, *err = nullptr;
  double *x = nullptr, *e = nullptr, *lf = nullptr, *u = nullptr, *b = nullptr;
  // nodal stresses: (element, local node) pairs of every node in ascending
  // element order, and the buffers of the stress pass (allocated on first use)
  int32_t *inc_ptr = nullptr, *inc = nullptr;
  double *pl = nullptr, *tx = nullptr, *tx0 = nullptr;
};

int color_elements(const int32_t* ix, int nnode, int numel, int nen, int nenv,
                   std::vector<int32_t>& icolor);
void element_slots(const int32_t* ix, const int32_t* id, int numel, int nen, int ndf, int neq,
                   const int32_t* ia1, const int32_t* ja1, std::vector<int32_t>& slots);

void free_mesh(MeshDev* m) {
  if (!m) return;
  void* ptrs[] = {m->ix, m->id, m->ie, m->eload, m->order, m->slots, m->ltype,
                  m->lnparc, m->err, m->x, m->e, m->lf, m->u, m->b,
                  m->inc_ptr, m->inc, m->pl, m->tx, m->tx0};
  for (void* p : ptrs)
    if (p) hipFree(p);
  delete m;
}

static bool getenv_flag(const char* name) {
  const char* v = getenv(name);
  return v && v[0] == '1';
}

template <typename T>
static hipError_t up(T** d, const T* h, size_t n, hipStream_t s) {
  hipError_t e = hipMalloc((void**)d, sizeof(T) * std::max<size_t>(n, 1));
  if (e != hipSuccess) return e;
  if (n && h) return hipMemcpyAsync(*d, h, sizeof(T) * n, hipMemcpyHostToDevice, s);
  return hipSuccess;
}

static AsmArgs make_args(Context* c, MeshDev* m) {
  AsmArgs a;
  a.ix = m->ix, a.x = m->x, a.id = m->id, a.ie = m->ie, a.e = m->e, a.eload = m->eload;
  a.order = m->order, a.slots = m->slots, a.ltype = m->ltype, a.lnparc = m->lnparc, a.lf = m->lf;
  a.u = nullptr, a.ad = c->A64.ad, a.al = c->A64.al, a.b = m->b, a.err = m->err;
  a.nen = m->nen, a.ndf = m->ndf, a.ndm = m->ndm, a.neq = c->neq, a.nloads = m->nloads;
  a.g0 = 0.0, a.g1 = 0.0, a.g2 = -9.81;
  return a;
}

template <int KIND>
static void launch_stiffness(const AsmArgs& a, const MeshDev* m, int lhs, int rhs, hipStream_t s) {
  // hex8 stiffness with u = 0 (b - K u = b): the pair kernel
  const bool pairs = KIND == K_HEX8 && lhs && (!rhs || a.u == nullptr) &&
                     !getenv_flag("ELASTIC_ASM_WAVE");
  for (int c = 0; c < m->ncolors; ++c) {
    int c0 = m->color_start[c], cnt = m->color_start[c + 1] - c0;
    if (cnt <= 0) continue;
    if (pairs)
      hipLaunchKernelGGL(k_stiffness_hex8, dim3((cnt + kHexE - 1) / kHexE), dim3(256), 0, s, a,
                         c0, cnt);
    else
      hipLaunchKernelGGL(k_stiffness<KIND>, dim3(cnt), dim3(64), 0, s, a, c0, lhs, rhs);
  }
}

template <int KIND>
static void launch_loads(const AsmArgs& a, const MeshDev* m, hipStream_t s) {
  for (int c = 0; c < m->ncolors; ++c) {
    int c0 = m->color_start[c], cnt = m->color_start[c + 1] - c0;
    if (cnt > 0) hipLaunchKernelGGL(k_loads<KIND>, dim3((cnt + 63) / 64), dim3(64), 0, s, a, c0, cnt);
  }
}

// ------------------------------------------------- nodal stresses (isw = 3)
// tform_mec (Pform.f:258-421) with the elements' isw = 3 branches (label 300
// of elmt02..07, Elmlib02.f): the element's stresses at its nodes, then per
// node the sum over its elements in element order, divided by their count,
// plus the initial stress.  Same operations in the same order as the
// reference (FP contraction off), so tx is bit-identical to tform_mec's.
struct StressArgs {
  const int32_t* ix;
  const double* x;
  const int32_t* ie;
  const double* e;
  const double* u;   // ndf per node
  double* pl;        // ntn per (element, local node)
  int32_t* err;      // an element (1-based) with a non-positive Jacobian at a node
  int numel;
};

// deform2d / deform3d (Elmlib.f:2368-2472)
template <int NEN>
__device__ inline void deform2d_d(const double* hx, const double* hy, const double* u,
                                  double* eps) {
  double ux = 0.0, uy = 0.0, vx = 0.0, vy = 0.0;
  for (int i = 0; i < NEN; ++i) {
    ux = ux + hx[i] * u[2 * i];
    uy = uy + hy[i] * u[2 * i];
    vx = vx + hx[i] * u[2 * i + 1];
    vy = vy + hy[i] * u[2 * i + 1];
  }
  eps[0] = ux;
  eps[1] = vy;
  eps[2] = uy + vx;
}

template <int NEN>
__device__ inline void deform3d_d(const double* hx, const double* hy, const double* hz,
                                  const double* u, double* eps) {
  double ux = 0.0, uy = 0.0, uz = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
  for (int i = 0; i < NEN; ++i) {
    ux = ux + hx[i] * u[3 * i];
    uy = uy + hy[i] * u[3 * i];
    uz = uz + hz[i] * u[3 * i];
    vx = vx + hx[i] * u[3 * i + 1];
    vy = vy + hy[i] * u[3 * i + 1];
    vz = vz + hz[i] * u[3 * i + 1];
    wx = wx + hx[i] * u[3 * i + 2];
    wy = wy + hy[i] * u[3 * i + 2];
    wz = wz + hz[i] * u[3 * i + 2];
  }
  eps[0] = ux;
  eps[1] = vy;
  eps[2] = wz;
  eps[3] = uy + vx;
  eps[4] = vz + wy;
  eps[5] = uz + wx;
}

// stress2d_m (Elmlib.f:2502-2539) and stress3d (Elmlib.f:2540-2573)
__device__ inline void stress2d_m_d(double d11, double d12, double d22, double d33, double ps,
                                    bool ept, const double* eps, double* t) {
  t[0] = d11 * eps[0] + d12 * eps[1];
  t[1] = d12 * eps[0] + d22 * eps[1];
  t[2] = ept ? ps * (t[0] + t[1]) : 0.0;
  t[3] = d33 * eps[2];
}

__device__ inline void stress3d_d(double a, double b, double c, const double* eps, double* t) {
  t[0] = (eps[0] + b * eps[1] + b * eps[2]) * a;
  t[1] = (b * eps[0] + eps[1] + b * eps[2]) * a;
  t[2] = (b * eps[0] + b * eps[1] + eps[2]) * a;
  t[3] = eps[3] * c * a;
  t[4] = eps[4] * c * a;
  t[5] = eps[5] * c * a;
}

// One lane per (element, local node): the stresses at that node.
template <int KIND>
__global__ __launch_bounds__(256) void k_stress_el(StressArgs a) {
  using TR = ElemTraits<KIND>;
  constexpr int NEN = TR::NEN, NDF = TR::NDF, NDM = TR::NDM, NTN = NDM == 3 ? 6 : 4;
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= (int64_t)a.numel * NEN) return;
  const int el = (int)(q / NEN), node = (int)(q - (int64_t)el * NEN);
  const int32_t* ixe = a.ix + (size_t)el * (NEN + 1);
  double xl[NEN * NDM], ul[NEN * NDF];
  for (int n = 0; n < NEN; ++n) {
    const size_t g = (size_t)(ixe[n] - 1);
    for (int d = 0; d < NDM; ++d) xl[n * NDM + d] = a.x[g * NDM + d];
    for (int d = 0; d < NDF; ++d) ul[n * NDF + d] = a.u[g * NDF + d];
  }
  const int ma = ixe[NEN];
  const int iel = a.ie[ma - 1];
  const double* em = a.e + (size_t)(ma - 1) * 15;
  const double ym = em[0], ps = em[1];
  const bool pstress = (iel == 3 || iel == 5);
  double ca, cb, cc;
  if (!pstress) {
    const double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
    ca = (ym * a1) / (a3 * a2);
    cb = ps / a1;
    cc = 0.5 * (a2 / a1);
  } else {
    const double a1 = 1.0 - ps, a2 = 1.0 - ps * ps;
    ca = ym / a2;
    cb = ps;
    cc = 0.5 * a1;
  }
  double t[NTN], eps[6], det;
  bool ok = true;
  if constexpr (KIND == K_TRIA3) {
    const double xj11 = xl[0] - xl[4], xj12 = xl[1] - xl[5];
    const double xj21 = xl[2] - xl[4], xj22 = xl[3] - xl[5];
    det = xj11 * xj22 - xj12 * xj21;
    ok = det > 0.0;
    const double ji11 = xj22 / det, ji12 = -xj12 / det, ji21 = -xj21 / det, ji22 = xj11 / det;
    const double hx[3] = {ji11, ji12, -ji11 - ji12}, hy[3] = {ji21, ji22, -ji21 - ji22};
    deform2d_d<3>(hx, hy, ul, eps);
    stress2d_m_d(ca, ca * cb, ca, ca * cc, ps, !pstress, eps, t);
  } else if constexpr (KIND == K_QUAD4) {
    const double rn = (node == 0 || node == 3) ? 1.0 : -1.0, sn = node < 2 ? 1.0 : -1.0;
    double h[4], hx[4], hy[4];
    sfquad4_d(rn, sn, h, hx, hy);
    ok = jacob2d_d(hx, hy, xl, 4, 2, &det);
    deform2d_d<4>(hx, hy, ul, eps);
    stress2d_m_d(ca, ca * cb, ca, ca * cc, ps, !pstress, eps, t);
  } else if constexpr (KIND == K_TET4) {
    double hx[4], hy[4], hz[4];
    ok = jtetra4_d(xl, hx, hy, hz, &det, true);
    deform3d_d<4>(hx, hy, hz, ul, eps);
    stress3d_d(ca, cb, cc, eps, t);
  } else {
    const int c4 = node & 3;
    const double rn = (c4 == 0 || c4 == 3) ? 1.0 : -1.0, sn = c4 < 2 ? 1.0 : -1.0;
    const double tn = node < 4 ? 1.0 : -1.0;
    double h[8], hx[8], hy[8], hz[8];
    sfhexa8_d(rn, sn, tn, h, hx, hy, hz);
    ok = jacob3d_d(hx, hy, hz, xl, 8, &det);
    deform3d_d<8>(hx, hy, hz, ul, eps);
    stress3d_d(ca, cb, cc, eps, t);
  }
  if (!ok) atomicCAS(a.err, 0, el + 1);
  for (int j = 0; j < NTN; ++j) a.pl[q * NTN + j] = t[j];
}

// One lane per node: the node's element rows added in element order (the
// incidence list), / count, + initial stress (tx0 may be null).
__global__ __launch_bounds__(256) void k_stress_nodes(int nnode, int ntn,
                                                       const int32_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ inc,
                                                       const double* __restrict__ pl,
                                                       const double* __restrict__ tx0,
                                                       double* __restrict__ tx) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= nnode) return;
  double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int k0 = ptr[n], k1 = ptr[n + 1];
  for (int k = k0; k < k1; ++k) {
    const double* p = pl + (size_t)inc[k] * ntn;
    for (int j = 0; j < ntn; ++j) t[j] = t[j] + p[j];
  }
  const double ic = (double)(k1 - k0);
  for (int j = 0; j < ntn; ++j)
    tx[(size_t)n * ntn + j] = t[j] / ic + (tx0 ? tx0[(size_t)n * ntn + j] : 0.0);
}

// update_mec (Predict.f:183-209) on the device: u(j,i) = x(id(j,i)) where the
// dof is free
__global__ __launch_bounds__(256) void k_update_u(int n, const int32_t* __restrict__ id,
                                                   const double* __restrict__ x, double* u) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < n && id[q] > 0) u[q] = x[id[q] - 1];
}

}  // namespace elastic

using namespace elastic;

extern "C" int elastic_set_mesh(elastic_ctx* xc, int32_t nnode, int32_t numel, int32_t nen,
                                int32_t ndf, int32_t ndm, int32_t numat, const int32_t* ix,
                                const double* x, const int32_t* id, const int32_t* ie,
                                const double* e, const int32_t* eload, int32_t nloads,
                                const int32_t* ltype, const int32_t* lnparc,
                                const double* lvals) {
  if (!xc || nnode <= 0 || numel <= 0 || !ix || !x || !id || !ie || !e || numat <= 0)
    return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  if (!layout_ready(c)) return set_error(c, ELASTIC_NOT_READY, "structure not set");
  hipSetDevice(c->device);
  // element kind from the materials used (one kind per mesh)
  int kind = -1;
  for (int el = 0; el < numel; ++el) {
    int ma = ix[(size_t)el * (nen + 1) + nen];
    if (ma < 1 || ma > numat) return set_error(c, ELASTIC_BAD_ARG, "material number out of range");
    int iel = ie[ma - 1], k;
    switch (iel) {
      case 2: case 3: k = K_TRIA3; break;
      case 4: case 5: k = K_QUAD4; break;
      case 6: k = K_TET4; break;
      case 7: k = K_HEX8; break;
      default: return set_error(c, ELASTIC_BAD_ARG, "element type " + std::to_string(iel) + " not on the GPU path");
    }
    if (kind == -1) kind = k;
    else if (kind != k) return set_error(c, ELASTIC_BAD_ARG, "one element kind per mesh");
  }
  const int want_nen[] = {3, 4, 4, 8}, want_ndf[] = {2, 2, 3, 3}, want_ndm[] = {2, 2, 3, 3};
  if (nen != want_nen[kind] || ndf != want_ndf[kind] || ndm != want_ndm[kind])
    return set_error(c, ELASTIC_BAD_ARG, "nen/ndf/ndm do not match the element type");
  // every index the device kernels follow stays in range
  for (int el = 0; el < numel; ++el)
    for (int i = 0; i < nen; ++i) {
      const int no = ix[(size_t)el * (nen + 1) + i];
      if (no < 1 || no > nnode)
        return set_error(c, ELASTIC_BAD_ARG, "node number out of range in element " + std::to_string(el + 1));
    }
  for (size_t q = 0; q < (size_t)nnode * ndf; ++q)
    if (id[q] > c->neq) return set_error(c, ELASTIC_BAD_ARG, "equation number above neq");
  std::vector<int32_t> icolor;
  int ncol = color_elements(ix, nnode, numel, nen, nen, icolor);
  if (ncol < 0) return set_error(c, ELASTIC_BAD_ARG, "element colouring failed");
  std::vector<int32_t> order;
  order.reserve(numel);
  std::vector<int> cstart(ncol + 1, 0);
  for (int cc = 1; cc <= ncol; ++cc) {
    cstart[cc - 1] = (int)order.size();
    for (int el = 1; el <= numel; ++el)
      if (icolor[el] == cc) order.push_back(el - 1);
  }
  cstart[ncol] = (int)order.size();
  std::vector<int32_t> slots;
  try {
    element_slots(ix, id, numel, nen, ndf, c->neq, xc->ia.data(), xc->ja.data(), slots);
  } catch (const std::exception& ex) {
    return set_error(c, ELASTIC_BAD_ARG, ex.what());
  }
  if (c->mesh) free_mesh((MeshDev*)c->mesh);
  MeshDev* m = new MeshDev();
  c->mesh = m;
  m->nnode = nnode, m->numel = numel, m->nen = nen, m->ndf = ndf, m->ndm = ndm, m->kind = kind;
  m->nloads = nloads, m->ncolors = ncol, m->color_start = cstart;
  std::vector<int32_t> lt(std::max(nloads, 1), 0), ln(std::max(nloads, 1), 0);
  std::vector<double> lv(3 * std::max(nloads, 1), 0.0);
  for (int i = 0; i < nloads; ++i) {
    lt[i] = ltype[i];
    ln[i] = lnparc[i];
    for (int q = 0; q < 3; ++q) lv[3 * i + q] = lvals[3 * i + q];
  }
  std::vector<int32_t> el7((size_t)numel * 7, 0);
  if (eload) std::copy(eload, eload + (size_t)numel * 7, el7.begin());
  hipStream_t s = c->stream;
  hipError_t r = hipSuccess;
  auto chk = [&](hipError_t q) { if (r == hipSuccess) r = q; };
  chk(up(&m->ix, ix, (size_t)numel * (nen + 1), s));
  chk(up(&m->x, x, (size_t)nnode * ndm, s));
  chk(up(&m->id, id, (size_t)nnode * ndf, s));
  chk(up(&m->ie, ie, (size_t)numat, s));
  chk(up(&m->e, e, (size_t)numat * 15, s));
  chk(up(&m->eload, el7.data(), el7.size(), s));
  chk(up(&m->order, order.data(), order.size(), s));
  chk(up(&m->slots, slots.data(), slots.size(), s));
  chk(up(&m->ltype, lt.data(), lt.size(), s));
  chk(up(&m->lnparc, ln.data(), ln.size(), s));
  chk(up(&m->lf, lv.data(), lv.size(), s));
  chk(up<double>(&m->u, nullptr, (size_t)nnode * ndf, s));
  chk(up<double>(&m->b, nullptr, (size_t)c->neq, s));
  chk(up<int32_t>(&m->err, nullptr, 1, s));
  {  // node -> (element, local node) in ascending element order (tform_mec's order)
    std::vector<int32_t> ptr((size_t)nnode + 1, 0), inc;
    for (int el = 0; el < numel; ++el)
      for (int i = 0; i < nen; ++i) {
        const int no = ix[(size_t)el * (nen + 1) + i];
        if (no > 0 && no <= nnode) ptr[no] += 1;
      }
    for (int n = 0; n < nnode; ++n) ptr[n + 1] += ptr[n];
    inc.resize(std::max<size_t>(ptr[nnode], 1));
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int el = 0; el < numel; ++el)
      for (int i = 0; i < nen; ++i) {
        const int no = ix[(size_t)el * (nen + 1) + i];
        if (no > 0 && no <= nnode) inc[fill[no - 1]++] = el * nen + i;
      }
    chk(up(&m->inc_ptr, ptr.data(), ptr.size(), s));
    chk(up(&m->inc, inc.data(), inc.size(), s));
    chk(hipStreamSynchronize(s));  // the host vectors go out of scope
  }
  chk(hipStreamSynchronize(s));
  if (r != hipSuccess) {
    free_mesh(m);
    c->mesh = nullptr;
    return set_error(c, r == hipErrorOutOfMemory ? ELASTIC_OOM : ELASTIC_HIP_ERROR,
                     std::string("set_mesh: ") + hipGetErrorString(r));
  }
  return ELASTIC_OK;
}

// Pform.f:1-254: isw 2 (lhs: K into ad/al; rhs: b -= K_e u_e) or isw 4 (loads
// into b).  u is u(ndf,nnode) (may be NULL: zero); b is b(neq), in/out on the
// host (may be NULL when rhs = 0).
extern "C" int elastic_assemble(elastic_ctx* xc, int32_t isw, int32_t lhs, int32_t rhs,
                                const double* u, double* b) {
  if (!xc) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  if (!m) return set_error(c, ELASTIC_NOT_READY, "mesh not set");
  if (rhs && !b) return set_error(c, ELASTIC_BAD_ARG, "rhs needs b");
  if (isw != 2 && isw != 4) return set_error(c, ELASTIC_BAD_ARG, "isw must be 2 or 4");
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  AsmArgs a = make_args(c, m);
  if (u) {
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(m->u, u, sizeof(double) * m->nnode * m->ndf,
                                        hipMemcpyHostToDevice, s));
    a.u = m->u;
  }
  if (rhs)
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(m->b, b, sizeof(double) * c->neq, hipMemcpyHostToDevice, s));
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(m->err, 0, sizeof(int32_t), s));
  if (isw == 2) {
    if (lhs) {
      ELASTIC_HIP_CHECK(c, hipMemsetAsync(c->A64.ad, 0, sizeof(double) * c->neq, s));
      ELASTIC_HIP_CHECK(c, hipMemsetAsync(c->A64.al, 0, sizeof(double) * std::max<int64_t>(c->nad, 1), s));
    }
    switch (m->kind) {
      case K_TRIA3: launch_stiffness<K_TRIA3>(a, m, lhs, rhs, s); break;
      case K_QUAD4: launch_stiffness<K_QUAD4>(a, m, lhs, rhs, s); break;
      case K_TET4: launch_stiffness<K_TET4>(a, m, lhs, rhs, s); break;
      default: launch_stiffness<K_HEX8>(a, m, lhs, rhs, s); break;
    }
  } else if (rhs) {
    switch (m->kind) {
      case K_TRIA3: launch_loads<K_TRIA3>(a, m, s); break;
      case K_QUAD4: launch_loads<K_QUAD4>(a, m, s); break;
      case K_TET4: launch_loads<K_TET4>(a, m, s); break;
      default: launch_loads<K_HEX8>(a, m, s); break;
    }
  }
  int32_t err = 0;
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(&err, m->err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (rhs)
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(b, m->b, sizeof(double) * c->neq, hipMemcpyDeviceToHost, s));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  if (err) return set_error(c, ELASTIC_NEG_JACOBIAN, "non-positive Jacobian in element " + std::to_string(err));
  if (isw == 2 && lhs) {
    // derived arrays of the new values (Jacobi inverse, FP32 copies)
    std::vector<double> ad(c->neq);
    ELASTIC_HIP_CHECK(c, hipMemcpyAsync(ad.data(), c->A64.ad, sizeof(double) * c->neq, hipMemcpyDeviceToHost, s));
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
    c->zero_diag_row = -1;
    for (int i = 0; i < c->neq; ++i)
      if (ad[i] == 0.0) {
        c->zero_diag_row = i + 1;
        break;
      }
    derive_operator(c);
    ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
    c->values_set = true;
    c->bdiag_set = false;
  }
  return ELASTIC_OK;
}

// tform_mec (Pform.f:258-421, isw = 3): nodal stresses tx(ntn, nnode) of the
// mesh from u(ndf, nnode) (NULL: zero); from_x: the free dofs of u are first
// replaced by the context's resident solution x (update_mec).  tx0 (ntn,
// nnode) may be NULL.
extern "C" int elastic_stress(elastic_ctx* xc, const double* u, int32_t from_x, const double* tx0,
                              double* tx) {
  if (!xc || !tx) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  if (!m) return set_error(c, ELASTIC_NOT_READY, "mesh not set");
  if (from_x && !c->v64[VX]) return set_error(c, ELASTIC_NOT_READY, "no resident solution");
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const int ntn = m->ndm == 3 ? 6 : 4;
  const size_t nu = (size_t)m->nnode * m->ndf, nt = (size_t)m->nnode * ntn;
  const size_t npl = (size_t)m->numel * m->nen * ntn;
  hipError_t r = hipSuccess;
  if (!m->pl) r = up<double>(&m->pl, nullptr, npl, s);
  if (r == hipSuccess && !m->tx) r = up<double>(&m->tx, nullptr, nt, s);
  if (r == hipSuccess && tx0 && !m->tx0) r = up<double>(&m->tx0, nullptr, nt, s);
  if (r != hipSuccess)
    return set_error(c, r == hipErrorOutOfMemory ? ELASTIC_OOM : ELASTIC_HIP_ERROR, "stress buffers");
  if (u) ELASTIC_HIP_CHECK(c, hipMemcpyAsync(m->u, u, sizeof(double) * nu, hipMemcpyHostToDevice, s));
  else ELASTIC_HIP_CHECK(c, hipMemsetAsync(m->u, 0, sizeof(double) * nu, s));
  if (tx0) ELASTIC_HIP_CHECK(c, hipMemcpyAsync(m->tx0, tx0, sizeof(double) * nt, hipMemcpyHostToDevice, s));
  if (from_x)
    hipLaunchKernelGGL(k_update_u, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, s, (int)nu,
                       (const int32_t*)m->id, (const double*)c->v64[VX], m->u);
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(m->err, 0, sizeof(int32_t), s));
  StressArgs a{m->ix, m->x, m->ie, m->e, m->u, m->pl, m->err, m->numel};
  const dim3 ge((unsigned)(((int64_t)m->numel * m->nen + 255) / 256));
  switch (m->kind) {
    case K_TRIA3: hipLaunchKernelGGL(k_stress_el<K_TRIA3>, ge, dim3(256), 0, s, a); break;
    case K_QUAD4: hipLaunchKernelGGL(k_stress_el<K_QUAD4>, ge, dim3(256), 0, s, a); break;
    case K_TET4: hipLaunchKernelGGL(k_stress_el<K_TET4>, ge, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(k_stress_el<K_HEX8>, ge, dim3(256), 0, s, a); break;
  }
  hipLaunchKernelGGL(k_stress_nodes, dim3((m->nnode + 255) / 256), dim3(256), 0, s, m->nnode, ntn,
                     (const int32_t*)m->inc_ptr, (const int32_t*)m->inc, (const double*)m->pl,
                     (const double*)(tx0 ? m->tx0 : nullptr), m->tx);
  int32_t err = 0;
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(&err, m->err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(tx, m->tx, sizeof(double) * nt, hipMemcpyDeviceToHost, s));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  if (err)
    return set_error(c, ELASTIC_NEG_JACOBIAN,
                     "non-positive Jacobian at a node of element " + std::to_string(err));
  return ELASTIC_OK;
}

// Average device time (ms) of one stiffness assembly (isw 2, lhs) over reps.
extern "C" int elastic_time_assembly(elastic_ctx* xc, int32_t reps, double* ms) {
  if (!xc || reps <= 0 || !ms) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  if (!m) return set_error(c, ELASTIC_NOT_READY, "mesh not set");
  hipStream_t s = c->stream;
  AsmArgs a = make_args(c, m);
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev0, s));
  for (int i = 0; i < reps; ++i) {
    ELASTIC_HIP_CHECK(c, hipMemsetAsync(c->A64.ad, 0, sizeof(double) * c->neq, s));
    ELASTIC_HIP_CHECK(c, hipMemsetAsync(c->A64.al, 0, sizeof(double) * std::max<int64_t>(c->nad, 1), s));
    switch (m->kind) {
      case K_TRIA3: launch_stiffness<K_TRIA3>(a, m, 1, 0, s); break;
      case K_QUAD4: launch_stiffness<K_QUAD4>(a, m, 1, 0, s); break;
      case K_TET4: launch_stiffness<K_TET4>(a, m, 1, 0, s); break;
      default: launch_stiffness<K_HEX8>(a, m, 1, 0, s); break;
    }
  }
  ELASTIC_HIP_CHECK(c, hipEventRecord(c->ev1, s));
  ELASTIC_HIP_CHECK(c, hipEventSynchronize(c->ev1));
  float t = 0;
  hipEventElapsedTime(&t, c->ev0, c->ev1);
  *ms = t / reps;
  return ELASTIC_OK;
}

// ======================================================= J2 plasticity (hex8)
// Small-strain J2 plasticity with linear isotropic hardening (current yield
// stress sy + H * alpha, alpha the equivalent plastic strain, the hardening
// law of Plasticity01.f:322, 389), integrated by the closed-form radial
// return with its consistent tangent (Simo & Hughes 1998, Box 3.2).  The
// reference reaches no plasticity from its driver, so there is no reference
// result to match; Voigt order xx, yy, zz, xy, yz, zx, engineering shears.
// One 64-lane wavefront per element of a colour: lanes 0..7 integrate the 8
// Gauss points (same points, shape functions and Jacobian as elmt07_mec),
// then all lanes form K_t = sum B^T C B det and f = sum B^T sigma det, and the
// element is added to ad/al and b (b -= f) through the slot map.
namespace elastic {

struct J2Dev {
  const double *ep_n, *al_n;  // committed state: 6 + 1 per Gauss point
  double *ep_t, *al_t;        // state of the last assembly
  const double* u;            // nodal displacements ndf*nnode
  double sy, H;
};

__device__ inline void j2_point(const double* eps, const double* epn, double aln, double E,
                                double nu, double sy, double H, double* sig, double* C,
                                double* ep, double* al) {
  const double mu = E / (2.0 * (1.0 + nu));
  const double kap = E / (3.0 * (1.0 - 2.0 * nu));
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double ee[6];
  for (int k = 0; k < 6; ++k) ee[k] = eps[k] - epn[k];
  const double tr = ee[0] + ee[1] + ee[2];
  double st[6];
  for (int k = 0; k < 3; ++k) st[k] = lam * tr + 2.0 * mu * ee[k];
  for (int k = 3; k < 6; ++k) st[k] = mu * ee[k];
  const double p = (st[0] + st[1] + st[2]) / 3.0;
  double s[6] = {st[0] - p, st[1] - p, st[2] - p, st[3], st[4], st[5]};
  const double sn = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] +
                         2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
  const double sq23 = sqrt(2.0 / 3.0);
  const double f = sn - sq23 * (sy + H * aln);
  const double w3[6] = {1.0, 1.0, 1.0, 0.5, 0.5, 0.5};
  if (f <= 0.0) {
    for (int k = 0; k < 6; ++k) sig[k] = st[k], ep[k] = epn[k];
    *al = aln;
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c)
        C[r * 6 + c] = (r < 3 && c < 3 ? lam : 0.0) + (r == c ? 2.0 * mu * w3[r] : 0.0);
    return;
  }
  const double dg = f / (2.0 * mu + 2.0 * H / 3.0);
  double n[6];
  for (int k = 0; k < 6; ++k) n[k] = s[k] / sn;
  for (int k = 0; k < 6; ++k) sig[k] = st[k] - 2.0 * mu * dg * n[k];
  for (int k = 0; k < 3; ++k) ep[k] = epn[k] + dg * n[k];
  for (int k = 3; k < 6; ++k) ep[k] = epn[k] + dg * 2.0 * n[k];
  *al = aln + sq23 * dg;
  const double th = 1.0 - 2.0 * mu * dg / sn;
  const double thb = 1.0 / (1.0 + H / (3.0 * mu)) - (1.0 - th);
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      const double pd = (r == c ? w3[r] : 0.0) - (r < 3 && c < 3 ? 1.0 / 3.0 : 0.0);
      C[r * 6 + c] = (r < 3 && c < 3 ? kap : 0.0) + 2.0 * mu * th * pd -
                     2.0 * mu * thb * n[r] * n[c];
    }
}

// B column of dof r of a node: rows and the derivative each row takes
__device__ inline void bcol(int r, int* row, int* der) {
  if (r == 0) row[0] = 0, der[0] = 0, row[1] = 3, der[1] = 1, row[2] = 5, der[2] = 2;
  else if (r == 1) row[0] = 1, der[0] = 1, row[1] = 3, der[1] = 0, row[2] = 4, der[2] = 2;
  else row[0] = 2, der[0] = 2, row[1] = 4, der[1] = 1, row[2] = 5, der[2] = 0;
}

__global__ __launch_bounds__(64) void k_j2_hex8(AsmArgs a, J2Dev j, int c0, int lhs) {
  constexpr int NEN = 8, NST = 24, NPAIR = NST * (NST + 1) / 2, NENT = NST * NST;
  __shared__ double xl[NEN * 3], ul[NST], dX[8][3][NEN], Cg[8][36], Sg[8][6], s[NENT];
  __shared__ int32_t ldv[NST];
  __shared__ int bad;
  const int el = a.order[c0 + blockIdx.x];
  const int tid = threadIdx.x;
  const int32_t* ixe = a.ix + (size_t)el * (NEN + 1);
  if (tid == 0) bad = 0;
  if (tid < NEN * 3) xl[tid] = a.x[(size_t)(ixe[tid / 3] - 1) * 3 + tid % 3];
  if (tid < NST) {
    const size_t g = (size_t)(ixe[tid / 3] - 1) * 3 + tid % 3;
    ldv[tid] = a.id[g];
    ul[tid] = j.u[g];
  }
  const double* em = a.e + (size_t)(ixe[NEN] - 1) * 15;
  const double E = em[0], nu = em[1];
  __syncthreads();
  if (tid < 8) {
    const int gp = tid, lx = gp & 1, ly = (gp >> 1) & 1, lz = gp >> 2;
    double h[NEN], hx[NEN], hy[NEN], hz[NEN], det;
    sfhexa8_d(kPG2[lx], kPG2[ly], kPG2[lz], h, hx, hy, hz);
    if (!jacob3d_d(hx, hy, hz, xl, NEN, &det)) bad = 1;
    double eps[6] = {0, 0, 0, 0, 0, 0};
    for (int n = 0; n < NEN; ++n) {
      dX[gp][0][n] = hx[n];
      dX[gp][1][n] = hy[n];
      dX[gp][2][n] = hz[n];
      const double ux = ul[3 * n], uy = ul[3 * n + 1], uz = ul[3 * n + 2];
      eps[0] += hx[n] * ux;
      eps[1] += hy[n] * uy;
      eps[2] += hz[n] * uz;
      eps[3] += hy[n] * ux + hx[n] * uy;
      eps[4] += hz[n] * uy + hy[n] * uz;
      eps[5] += hz[n] * ux + hx[n] * uz;
    }
    const size_t q = (size_t)el * 8 + gp;
    double epn[6], sig[6], C[36], ep[6], al;
    for (int k = 0; k < 6; ++k) epn[k] = j.ep_n[q * 6 + k];
    j2_point(eps, epn, j.al_n[q], E, nu, j.sy, j.H, sig, C, ep, &al);
    for (int k = 0; k < 6; ++k) Sg[gp][k] = sig[k] * det;
    for (int k = 0; k < 36; ++k) Cg[gp][k] = C[k] * det;
    for (int k = 0; k < 6; ++k) j.ep_t[q * 6 + k] = ep[k];
    j.al_t[q] = al;
  }
  __syncthreads();
  if (bad) {
    if (tid == 0) a.err[0] = el + 1;
    return;
  }
  // internal force, b -= f
  if (tid < NST) {
    const int n = tid / 3, r = tid % 3;
    int row[3], der[3];
    bcol(r, row, der);
    double f = 0.0;
    for (int gp = 0; gp < 8; ++gp)
      for (int k = 0; k < 3; ++k) f += dX[gp][der[k]][n] * Sg[gp][row[k]];
    const int k = ldv[tid];
    if (k > 0 && k <= a.neq) a.b[k - 1] = a.b[k - 1] - f;
  }
  if (!lhs) return;
  for (int en = tid; en < NENT; en += 64) {
    const int i = en % NST, jj = en / NST;
    const int na = i / 3, r = i % 3, nb = jj / 3, c = jj % 3;
    int ri[3], di[3], rj[3], dj[3];
    bcol(r, ri, di);
    bcol(c, rj, dj);
    double acc = 0.0;
    for (int gp = 0; gp < 8; ++gp)
      for (int p = 0; p < 3; ++p) {
        const double bi = dX[gp][di[p]][na];
        double t = 0.0;
        for (int q2 = 0; q2 < 3; ++q2) t += Cg[gp][ri[p] * 6 + rj[q2]] * dX[gp][dj[q2]][nb];
        acc += bi * t;
      }
    s[en] = acc;  // s(i,j) at i + NST*j
  }
  __syncthreads();
  const int32_t* sl = a.slots + (size_t)el * NPAIR;
  for (int pp = tid; pp < NPAIR; pp += 64) {
    const int32_t t = sl[pp];
    if (t == -1) continue;
    int i = (int)((sqrtf(8.0f * pp + 1.0f) - 1.0f) * 0.5f);
    while ((i + 1) * (i + 2) / 2 <= pp) ++i;
    while (i * (i + 1) / 2 > pp) --i;
    const int jx = pp - i * (i + 1) / 2;
    if (t <= -2) {
      const int k = -t - 2;
      a.ad[k] = a.ad[k] + s[i * NST + i];
    } else {
      a.al[t] = a.al[t] + (ldv[i] > ldv[jx] ? s[jx * NST + i] : s[i * NST + jx]);
    }
  }
}

// u(dof) = lambda * ubar(dof) on restrained dofs (id == 0)
__global__ void k_j2_prescribe(int n, const int32_t* __restrict__ id, const double* __restrict__ ubar,
                               double lambda, double* u) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (id[i] == 0) u[i] = lambda * ubar[i];
}

// u(dof) += x(id(dof)) on free dofs
// u += s du on the free dofs (s = 1: the Newton correction; other s: a
// line-search move along the same du)
__global__ void k_j2_correct(int n, const int32_t* __restrict__ id, const double* __restrict__ x,
                             double s, double* u) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (id[i] > 0) u[i] = u[i] + (s == 1.0 ? x[id[i] - 1] : s * x[id[i] - 1]);
}

// per-block partials: sum of b^2 (mode 0); max alpha and count of alpha > 0 (mode 1)
__global__ __launch_bounds__(256) void k_j2_reduce(int64_t n, const double* __restrict__ v, int mode,
                                                   double* part) {
  __shared__ double sh[256], sc[256];
  double acc = 0.0, cnt = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double x = v[i];
    if (mode == 0) acc += x * x;
    else {
      acc = fmax(acc, x);
      cnt += x > 0.0 ? 1.0 : 0.0;
    }
  }
  sh[threadIdx.x] = acc;
  sc[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sh[threadIdx.x] = mode == 0 ? sh[threadIdx.x] + sh[threadIdx.x + o]
                                  : fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
      sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0];
    part[2 * blockIdx.x + 1] = sc[0];
  }
}

struct J2State {
  double *ep_n = nullptr, *al_n = nullptr, *ep_t = nullptr, *al_t = nullptr, *u = nullptr,
         *ubar = nullptr, *part = nullptr;
  double sy = 0, H = 0;
  int64_t ngp = 0;
};

void free_j2(void* p) {
  J2State* j = (J2State*)p;
  if (!j) return;
  for (void* q : {(void*)j->ep_n, (void*)j->al_n, (void*)j->ep_t, (void*)j->al_t, (void*)j->u,
                  (void*)j->ubar, (void*)j->part})
    if (q) hipFree(q);
  delete j;
}

constexpr int kJ2RedBlocks = 512;

static int j2_reduce(Context* c, J2State* j, const double* v, int64_t n, int mode, double* r0,
                     double* r1) {
  hipLaunchKernelGGL(k_j2_reduce, dim3(kJ2RedBlocks), dim3(256), 0, c->stream, n, v, mode, j->part);
  std::vector<double> h(2 * kJ2RedBlocks);
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(h.data(), j->part, sizeof(double) * h.size(),
                                      hipMemcpyDeviceToHost, c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  double a = 0.0, b = 0.0;
  for (int k = 0; k < kJ2RedBlocks; ++k) {
    a = mode == 0 ? a + h[2 * k] : std::max(a, h[2 * k]);
    b += h[2 * k + 1];
  }
  *r0 = a;
  if (r1) *r1 = b;
  return ELASTIC_OK;
}

}  // namespace elastic

extern "C" int elastic_j2_setup(elastic_ctx* xc, double sy, double H, const double* ubar) {
  if (!xc || sy <= 0.0 || H < 0.0) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  if (!m) return set_error(c, ELASTIC_NOT_READY, "mesh not set");
  if (m->kind != K_HEX8) return set_error(c, ELASTIC_BAD_ARG, "J2 plasticity is built for hex8");
  hipSetDevice(c->device);
  if (c->j2) free_j2(c->j2);
  J2State* j = new J2State();
  c->j2 = j;
  j->sy = sy;
  j->H = H;
  j->ngp = (int64_t)m->numel * 8;
  const size_t nu = (size_t)m->nnode * 3;
  if (hipMalloc((void**)&j->ep_n, sizeof(double) * 6 * j->ngp) != hipSuccess ||
      hipMalloc((void**)&j->ep_t, sizeof(double) * 6 * j->ngp) != hipSuccess ||
      hipMalloc((void**)&j->al_n, sizeof(double) * j->ngp) != hipSuccess ||
      hipMalloc((void**)&j->al_t, sizeof(double) * j->ngp) != hipSuccess ||
      hipMalloc((void**)&j->u, sizeof(double) * nu) != hipSuccess ||
      hipMalloc((void**)&j->ubar, sizeof(double) * nu) != hipSuccess ||
      hipMalloc((void**)&j->part, sizeof(double) * 2 * kJ2RedBlocks) != hipSuccess) {
    (void)hipGetLastError();
    free_j2(j);
    c->j2 = nullptr;
    return set_error(c, ELASTIC_OOM, "J2 state");
  }
  hipStream_t s = c->stream;
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(j->ep_n, 0, sizeof(double) * 6 * j->ngp, s));
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(j->al_n, 0, sizeof(double) * j->ngp, s));
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(j->u, 0, sizeof(double) * nu, s));
  if (ubar) ELASTIC_HIP_CHECK(c, hipMemcpyAsync(j->ubar, ubar, sizeof(double) * nu, hipMemcpyHostToDevice, s));
  else ELASTIC_HIP_CHECK(c, hipMemsetAsync(j->ubar, 0, sizeof(double) * nu, s));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
  return ELASTIC_OK;
}

extern "C" int elastic_j2_increment(elastic_ctx* xc, double lambda) {
  if (!xc) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  J2State* j = (J2State*)c->j2;
  if (!m || !j) return set_error(c, ELASTIC_NOT_READY, "J2 not set up");
  hipSetDevice(c->device);
  const int n = m->nnode * 3;
  hipLaunchKernelGGL(k_j2_prescribe, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, m->id,
                     (const double*)j->ubar, lambda, j->u);
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

extern "C" int elastic_j2_assemble(elastic_ctx* xc, int32_t lhs, double* rnorm) {
  if (!xc) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  J2State* j = (J2State*)c->j2;
  if (!m || !j) return set_error(c, ELASTIC_NOT_READY, "J2 not set up");
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  AsmArgs a = make_args(c, m);
  a.b = c->v64[VB];  // the residual stays on the device, as the solve's b
  J2Dev d{j->ep_n, j->al_n, j->ep_t, j->al_t, j->u, j->sy, j->H};
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(m->err, 0, sizeof(int32_t), s));
  ELASTIC_HIP_CHECK(c, hipMemsetAsync(a.b, 0, sizeof(double) * c->neq, s));
  if (lhs) {
    ELASTIC_HIP_CHECK(c, hipMemsetAsync(c->A64.ad, 0, sizeof(double) * c->neq, s));
    ELASTIC_HIP_CHECK(c, hipMemsetAsync(c->A64.al, 0, sizeof(double) * std::max<int64_t>(c->nad, 1), s));
  }
  for (int k = 0; k < m->ncolors; ++k) {
    const int c0 = m->color_start[k], cnt = m->color_start[k + 1] - c0;
    if (cnt > 0) hipLaunchKernelGGL(k_j2_hex8, dim3(cnt), dim3(64), 0, s, a, d, c0, lhs);
  }
  int32_t err = 0;
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(&err, m->err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
  ELASTIC_HIP_CHECK(c, hipGetLastError());
  if (err) return set_error(c, ELASTIC_NEG_JACOBIAN, "non-positive Jacobian in element " + std::to_string(err));
  if (lhs) {
    derive_operator(c);
    c->values_set = true;
    c->bdiag_set = false;
    c->zero_diag_row = -1;
  }
  double r2 = 0.0;
  int rc = j2_reduce(c, j, a.b, c->neq, 0, &r2, nullptr);
  if (rc) return rc;
  if (rnorm) *rnorm = std::sqrt(r2);
  return ELASTIC_OK;
}

extern "C" int elastic_j2_step(elastic_ctx* xc, double s) {
  if (!xc || !std::isfinite(s)) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  J2State* j = (J2State*)c->j2;
  if (!m || !j) return set_error(c, ELASTIC_NOT_READY, "J2 not set up");
  hipSetDevice(c->device);
  const int n = m->nnode * 3;
  hipLaunchKernelGGL(k_j2_correct, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, m->id,
                     (const double*)c->v64[VX], s, j->u);
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

extern "C" int elastic_j2_correct(elastic_ctx* xc) { return elastic_j2_step(xc, 1.0); }

extern "C" int elastic_j2_commit(elastic_ctx* xc, double* alpha_max, int64_t* nplastic) {
  if (!xc) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  J2State* j = (J2State*)c->j2;
  if (!j) return set_error(c, ELASTIC_NOT_READY, "J2 not set up");
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(j->ep_n, j->ep_t, sizeof(double) * 6 * j->ngp, hipMemcpyDeviceToDevice, s));
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(j->al_n, j->al_t, sizeof(double) * j->ngp, hipMemcpyDeviceToDevice, s));
  double amax = 0.0, cnt = 0.0;
  int rc = j2_reduce(c, j, j->al_n, j->ngp, 1, &amax, &cnt);
  if (rc) return rc;
  if (alpha_max) *alpha_max = amax;
  if (nplastic) *nplastic = (int64_t)cnt;
  return ELASTIC_OK;
}

extern "C" int elastic_j2_get_u(elastic_ctx* xc, double* u) {
  if (!xc || !u) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  MeshDev* m = (MeshDev*)c->mesh;
  J2State* j = (J2State*)c->j2;
  if (!m || !j) return set_error(c, ELASTIC_NOT_READY, "J2 not set up");
  hipSetDevice(c->device);
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(u, j->u, sizeof(double) * m->nnode * 3, hipMemcpyDeviceToHost,
                                      c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}

extern "C" int elastic_j2_get_state(elastic_ctx* xc, int32_t trial, double* ep, double* alpha) {
  if (!xc) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  J2State* j = (J2State*)c->j2;
  if (!j) return set_error(c, ELASTIC_NOT_READY, "J2 not set up");
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  if (ep) ELASTIC_HIP_CHECK(c, hipMemcpyAsync(ep, trial ? j->ep_t : j->ep_n, sizeof(double) * 6 * j->ngp, hipMemcpyDeviceToHost, s));
  if (alpha) ELASTIC_HIP_CHECK(c, hipMemcpyAsync(alpha, trial ? j->al_t : j->al_n, sizeof(double) * j->ngp, hipMemcpyDeviceToHost, s));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(s));
  return ELASTIC_OK;
}

extern "C" int elastic_j2_get_residual(elastic_ctx* xc, double* r) {
  if (!xc || !r) return ELASTIC_BAD_ARG;
  Context* c = &xc->c;
  if (!c->j2) return set_error(c, ELASTIC_NOT_READY, "J2 not set up");
  hipSetDevice(c->device);
  ELASTIC_HIP_CHECK(c, hipMemcpyAsync(r, c->v64[VB], sizeof(double) * c->neq, hipMemcpyDeviceToHost,
                                      c->stream));
  ELASTIC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return ELASTIC_OK;
}
