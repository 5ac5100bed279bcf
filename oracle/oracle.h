/*
 * oracle.h -- CPU restatement of the Elastic solve path (test infrastructure).
 *
 * This library exists only to check the HIP path: tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it; the product never does.
 * Every routine states the reference file:line it restates.  Arrays follow
 * the reference's Fortran conventions: 1-based index values, column-major
 * element arrays (ix(nen+1,numel), x(ndm,nnode), id(ndf,nnode)), CSRC
 * strictly-lower storage with ia(1)=1.
 */
#ifndef ELASTIC_ORACLE_H
#define ELASTIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PROP 15 /* termprop.fi: prop = 15 material constants */

/* Load table (load.fi): type (1 const, 2 linear, 40 vector traction), number
 * of parcels and up to 3 values each. */
typedef struct {
  int nload;
  int type[200];
  int nparc[200];
  double f[200][3];
} or_loads;

typedef struct {
  int iters;
  double xkx, norm_x, norm_r, norm_m_r, conv;
  int status; /* 0 ok, 1 maxit, 2 zero diagonal */
} or_stats;

/* Graph.f:20-257: node adjacency, sorted ascending, duplicates and self removed.
 * ip (nnode+1) is filled; returns the adjacency via *ips (malloc'ed, caller frees). */
int or_graph(int nnode, int numel, int nen, const int32_t* ix, int32_t* ip, int32_t** ips);

/* Reord.f:1-44 (+ genrcm/fnroot/rcm/rootls/degree 45-434): perm(k) = old node at
 * position k. */
void or_reord(const int32_t* ix, int32_t* perm, int nnode, int numel, int nen);

/* Numeq.f:139-191: id holds 0 (free) / nonzero (restrained) on input, equation
 * numbers (0 if restrained) on output.  Returns neq. */
int or_numeq(int32_t* id, const int32_t* perm, int nnode, int ndf);

/* Csr.f:13-305 via Datastruct.f:71-93 (lower=T, diag=F, upper=F).  Returns nad;
 * ia (neq+1) filled, *ja malloc'ed. */
int or_csrstruct(const int32_t* id, const int32_t* ix, const int32_t* perm, int nnode,
                 int numel, int nen, int ndf, int neq, int32_t* ia, int32_t** ja);

/* Element routines, isw=2 (stiffness s(nst,nst) column-major + p = s u) and
 * isw=4 (surface loads into p).  Return 0, or -1 on a non-positive Jacobian.
 * iel: 2 tria3 plane strain, 3 tria3 plane stress, 4 quad4 plane strain,
 * 5 quad4 plane stress, 6 tet4, 7 hex8 (Elmlib.f:1-166 dispatch). */
int or_element(int iel, int isw, const double* e, const int32_t* iq, const double* xl,
               const double* ul, double* p, double* s, int ndm, int nst,
               const or_loads* ld, const double* gravity);

/* Pform.f:1-254 serial branch + Assbly.f:66-146 (csr scatter).  lhs zeroes and
 * assembles ad/al; rhs updates b -= p.  Returns 0 or the 1-based element
 * number whose Jacobian was non-positive. */
int or_pform(const int32_t* ix, const int32_t* iq, const int32_t* ie, const double* e,
             const double* x, const int32_t* id, const int32_t* ia, const int32_t* ja,
             double* al, double* ad, double* b, const double* u, int numel, int nen,
             int nenv, int ndf, int ndm, int neq, int lhs, int rhs, int isw,
             const or_loads* ld, const double* gravity);

/* Pload.f:2-57 (time-constant loads, t = 0). */
/* isw = 3 of the element library (Elmlib02.f label 300 of elmt02..07):
 * nodal stresses p(ntn, nen) of one element */
int or_element_stress(int iel, const double* e, const double* x, const double* u, double* p,
                      int ndm);
/* tform_mec (Pform.f:258-421), isw = 3: averaged nodal stresses tx(ntn, nnode) */
int or_tform(const int32_t* ix, const double* x, const double* e, const int32_t* ie,
             const double* u, const double* tx0, double* tx, int nnode, int numel, int nen,
             int ndm, int ndf, int ntn);
void or_pload(const int32_t* id, double* f, double* u, double* b, const int32_t* nload,
              int nnode, int ndf, const or_loads* ld);

/* Predict.f:183-209 */
void or_update(int nnode, int ndf, const int32_t* id, double* u, const double* x);

/* Precond.f:22-44 (pd = 0 branch, as Solv.f:65-69 calls it). */
void or_pre_diag(double* m, const double* ad, int neq);

/* Precond.f:163-358 nblock=3 + Precond.f:1676-1710: m(6, neq/3 + rest). */
void or_block_precond3(const double* ad, const double* al, const int32_t* ia,
                       const int32_t* ja, double* m, int neq);
/* Precond.f:419-540 nblock=3 */
void or_op_block_precond3(const double* m, const double* x, double* y, int neq);

/* Matvec.f:528-621: y = (D + L + L^T) x in ascending row order. */
void or_matvec_csrcsym1(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                        const double* al, const double* x, double* y);

/* Matvec.f:1237-1266 (one process). */
double or_dot(const double* a, const double* b, int n);

/* Iters.f:252-524 with matvec_csrcsym1 and dot_par (Solv.f:121-132). */
void or_pcg(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
            const double* al, const double* m, const double* b, double* x, double tol,
            int maxit, or_stats* st);

/* Precond.f:163-358 nblock=2: m(3, neq/2 + rest) */
void or_block_precond2(const double* ad, const double* al, const int32_t* ia,
                       const int32_t* ja, double* m, int neq);
/* Precond.f:419-540 nblock=2 */
void or_op_block_precond2(const double* m, const double* x, double* y, int neq);

/* Iters.f:252-524 with z = M^-1 r applied by op_block_precond (Precond.f:419-540)
 * on bm = m(nblock(nblock+1)/2, .) of or_block_precond2/3 (block Jacobi,
 * precond code 6 with block size 2 or 3). */
void or_pcg_block(int nblock, int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                  const double* al, const double* bm, const double* b, double* x, double tol,
                  int maxit, or_stats* st);
/* or_pcg_block with nblock = 3 */
void or_pcg_bdiag3(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                   const double* al, const double* bm, const double* b, double* x, double tol,
                   int maxit, or_stats* st);

/* Iters_omp.f:17-331 + Matvec_omp.f:150-212 + Matrix_partition.f:98-211 with
 * nth threads (OpenMP).  The CPU baseline. */
void or_pcg_omp(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                const double* al, const double* m, const double* b, double* x, double tol,
                int maxit, int nth, or_stats* st);

/* Iters.f:528-808 (restarted GMRES(k), left Jacobi, x0 = 0). */
void or_gmres(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
              const double* al, const double* m, const double* b, double* x, int k,
              double tol, int maxit, or_stats* st);

/* Iters.f:1079-1349 (right-preconditioned BiCGSTAB, x0 = 0). */
void or_pbicgstab(int neq, const int32_t* ia, const int32_t* ja, const double* ad,
                  const double* al, const double* m, const double* b, double* x,
                  double tol, int maxit, or_stats* st);

/* Colormsh.f:9-308: greedy element colouring; icolor(numel) 1-based colours;
 * returns the number of colours (or -1 above 200). */
int or_colormsh(const int32_t* ix, int nnode, int numel, int nen, int nenv, int32_t* icolor);

#ifdef __cplusplus
}
#endif
#endif
