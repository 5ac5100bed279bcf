/*
 * elastic_hip.h -- C-ABI of the MI355X (gfx950) implementation of the Elastic
 * solve path: CSRC assembly and PCG / BiCGSTAB / GMRES on one GPU.
 *
 * Conventions (those of the reference's Fortran, so that a Fortran caller can
 * bind every entry point with ISO_C_BINDING and pass its own arrays):
 *   - plain host pointers and sizes; index arrays int32 with 1-based values;
 *   - CSRC symmetric storage: ia(neq+1) with ia(1)=1, ja(nad) strictly lower
 *     columns sorted per row, al(nad) values, ad(neq) diagonal
 *     (Datastruct.f:71-93, Csr.f:13-305);
 *   - the caller owns host arrays; the library owns the device copies inside a
 *     context; one host thread per context;
 *   - every entry point returns a status code instead of stopping the program.
 * Unless stated otherwise an entry point is synchronous (it returns when its
 * results are in the caller's arrays).
 */
#ifndef ELASTIC_HIP_H
#define ELASTIC_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ELASTIC_OK 0
#define ELASTIC_ZERO_DIAG 1      /* Iters.f:337-342 */
#define ELASTIC_MAXIT 2          /* Iters.f:429-431 */
#define ELASTIC_NEG_JACOBIAN 3   /* Elmlib.f:2004, 2685; Elmlib02.f:152 */
#define ELASTIC_OOM 4            /* Malloc.f:348 */
#define ELASTIC_HIP_ERROR 5
#define ELASTIC_BAD_ARG 6
#define ELASTIC_NOT_READY 7      /* values or structure not set */
#define ELASTIC_BREAKDOWN 8      /* zero denominator inside a Krylov method */
#define ELASTIC_BUSY 9           /* every slot of a batch is taken (try again later) */

/* solver codes (Main.f:193, Solv.f:467-498) */
#define ELASTIC_SOLVER_PCG 1
#define ELASTIC_SOLVER_GMRES 2
#define ELASTIC_SOLVER_BICGSTAB 4
/* preconditioner codes (Main.f:190, precond.fi) */
#define ELASTIC_PRECOND_NONE 1
#define ELASTIC_PRECOND_DIAG 2
#define ELASTIC_PRECOND_BDIAG3 6
/* arithmetic */
#define ELASTIC_FP64 64          /* FP64 values, vectors and sums */
#define ELASTIC_MIXED 32         /* FP32 operator, FP64 vectors and true residual */

typedef struct elastic_ctx elastic_ctx;

/* Results of one solve, the quantities pcg prints (Iters.f:435-507). */
typedef struct {
  int32_t iters;        /* Krylov iterations */
  int32_t outer;        /* FP64 residual replacements (MIXED) or restart cycles (GMRES) */
  int32_t status;
  int32_t spmv_count;   /* products with A (any precision) */
  double xkx;           /* x . K x */
  double norm_x;        /* ||x|| */
  double norm_r;        /* ||b - A x|| (FP64, explicit) */
  double norm_m_r;      /* ||b - A x||_M with M = diag(1/ad) */
  double conv;          /* tol * ||b||_M, the stopping threshold (Iters.f:355-360) */
  double rel_res;       /* norm_m_r / ||b||_M */
  double t_device_ms;   /* device time of the iterations (HIP events, inputs resident) */
  double t_total_ms;    /* wall time of the call incl. b upload and x download */
  int32_t op64_from;    /* MIXED: iteration from which the FP64 operator took over
                           because the FP32 one stopped making progress; -1: never */
  int32_t reserved;
} elastic_stats;

/* ------------------------------------------------------------ context */
/* Creates a context on HIP device `device`. */
int elastic_create(elastic_ctx** ctx, int32_t device);
int elastic_destroy(elastic_ctx* ctx);
/* Human-readable text of the last error on this context (never NULL). */
const char* elastic_last_error(const elastic_ctx* ctx);

/* --------------------------------------------------- host-side set-up */
/* These replace the reference's host steps and return exactly its arrays. */
/* Graph.f:20-257 + Reord.f:1-434 (reord with afl=.true.): perm(k) = old node
 * at position k.  ix is ix(nen+1,numel) (element nodes, then material). */
int elastic_reord(int32_t nnode, int32_t numel, int32_t nen, const int32_t* ix, int32_t* perm);
/* Numeq.f:139-191: id(ndf,nnode) 0 = free / 1 = restrained on input, equation
 * numbers on output; *neq receives the equation count. */
int elastic_numeq(int32_t nnode, int32_t ndf, const int32_t* perm, int32_t* id, int32_t* neq);
/* Csr.f:13-305 via Datastruct.f:71-93: first call with ja == NULL returns *nad
 * and fills ia(neq+1); second call fills ja(nad). */
int elastic_csrstruct(int32_t nnode, int32_t numel, int32_t nen, int32_t ndf, int32_t neq,
                      const int32_t* ix, const int32_t* id, int32_t* ia, int32_t* ja,
                      int64_t* nad);

/* ------------------------------------------------- system on the GPU */
/* Uploads the CSRC pattern once and builds the device layout (row tiles,
 * packed local indices, merge lists).  Replaces the structure half of
 * datastruct (Datastruct.f:71-93) for the device. */
int elastic_set_structure(elastic_ctx* ctx, int32_t neq, int64_t nad, const int32_t* ia,
                          const int32_t* ja);
/* Uploads matrix values (as the reference assembled them on the host). */
int elastic_set_values(elastic_ctx* ctx, const double* ad, const double* al);
/* Copies the device values back (after elastic_assemble, so that the Fortran
 * caller's CSRC stays authoritative). */
int elastic_get_values(elastic_ctx* ctx, double* ad, double* al);
/* y = (D + L + L^T) x, Matvec.f:528-621.  precision ELASTIC_FP64 or
 * ELASTIC_MIXED (the FP32 operator of the mixed method: values rounded to
 * FP32, products summed in FP64, or in FP32 with "accum32"; y and the
 * partial sums between tiles in FP64). */
int elastic_spmv(elastic_ctx* ctx, int32_t precision, const double* x, double* y);

/* ------------------------------------------------------- GPU assembly */
/* Uploads the mesh for GPU assembly (after elastic_set_structure): element
 * nodes ix(nen+1,numel), coordinates x(ndm,nnode), equation numbers
 * id(ndf,nnode) as numeq left them, element type per material ie(numat) (2,3
 * tria3, 4,5 quad4, 6 tet4, 7 hex8; one kind per mesh), material constants
 * e(15,numat), element load numbers eload(7,numel) (may be NULL) and the load
 * table (type, parcels, 3 values each; type 40 = traction vector, Pload.f:246).
 * Builds the element colouring (Colormsh.f:9-308) and the CSRC slot map of
 * every element. */
int elastic_set_mesh(elastic_ctx* ctx, int32_t nnode, int32_t numel, int32_t nen, int32_t ndf,
                     int32_t ndm, int32_t numat, const int32_t* ix, const double* x,
                     const int32_t* id, const int32_t* ie, const double* e,
                     const int32_t* eload, int32_t nloads, const int32_t* load_type,
                     const int32_t* load_nparc, const double* load_vals);
/* pform_mec (Pform.f:1-254) on the GPU: isw 2 with lhs assembles K into the
 * device ad/al (zeroed first) and with rhs sets b = b - K_e u_e; isw 4 adds
 * the element loads to b.  u(ndf,nnode) may be NULL (zero); b(neq) is in/out
 * on the host.  Use elastic_get_values to copy K back. */
int elastic_assemble(elastic_ctx* ctx, int32_t isw, int32_t lhs, int32_t rhs, const double* u,
                     double* b);
/* tform_mec (Pform.f:258-421) with the elements' isw = 3 (label 300 of
 * elmt02..07, Elmlib02.f) on the GPU: the nodal stresses tx(ntn, nnode),
 * ntn = 4 in 2-D (sxx syy szz sxy) or 6 in 3-D (sxx syy szz sxy syz sxz),
 * each node's element values averaged in element order plus the initial
 * stress tx0(ntn, nnode) (NULL: zero).  u(ndf, nnode) on the host (NULL:
 * zero); from_x != 0 first replaces its free dofs by the context's resident
 * solution x, as update_mec (Predict.f:183-209) does before tform_mec
 * (Main.f:946).  Bit-identical to the reference's tform_mec.  A
 * non-positive Jacobian at an element node returns ELASTIC_NEG_JACOBIAN. */
int elastic_stress(elastic_ctx* ctx, const double* u, int32_t from_x, const double* tx0,
                   double* tx);
/* Average device time (ms) of one stiffness assembly (isw 2, lhs) over reps. */
int elastic_time_assembly(elastic_ctx* ctx, int32_t reps, double* ms);

/* ------------------------------------------- J2 plasticity (hex8, GPU) */
/* No counterpart on the reference's called path (Plasticity01.f is not
 * reached from its driver).  Small-strain J2 plasticity with linear isotropic
 * hardening (yield stress sy + H * equivalent plastic strain, the law of
 * Plasticity01.f:322), closed-form radial return and consistent tangent,
 * for Newton steps that re-assemble K_t on the GPU and solve with
 * elastic_solve_resident or a batch.  State lives on the device: committed
 * and trial plastic strain (6) and equivalent plastic strain per Gauss point,
 * nodal displacements u(3,nnode).  After elastic_set_mesh (hex8 mesh, E and
 * nu from its material): */
/* allocates the state (zero) and stores ubar(3,nnode), the prescribed
 * displacements of restrained dofs at load factor 1. */
int elastic_j2_setup(elastic_ctx* ctx, double sy, double H, const double* ubar);
/* u = lambda * ubar on restrained dofs (free dofs keep their values). */
int elastic_j2_increment(elastic_ctx* ctx, double lambda);
/* Return map at every Gauss point from u and the committed state; K_t into
 * the device ad/al when lhs (and the operator derived from it), the residual
 * b = -f_int into the device b of elastic_solve_resident; *rnorm = ||b||. */
int elastic_j2_assemble(elastic_ctx* ctx, int32_t lhs, double* rnorm);
/* u += x (the device x of the last resident solve) on free dofs. */
int elastic_j2_correct(elastic_ctx* ctx);
/* u += s x on free dofs: a line-search move along the last correction
 * (s = 1 is elastic_j2_correct). */
int elastic_j2_step(elastic_ctx* ctx, double s);
/* committed state <- state of the last assembly; returns the largest
 * equivalent plastic strain and the number of yielded Gauss points. */
int elastic_j2_commit(elastic_ctx* ctx, double* alpha_max, int64_t* nplastic);
int elastic_j2_get_u(elastic_ctx* ctx, double* u);
/* the residual b = -f_int of the last elastic_j2_assemble (neq values) */
int elastic_j2_get_residual(elastic_ctx* ctx, double* r);
/* plastic strains (6 per Gauss point, element-major) and equivalent plastic
 * strains of the committed (trial = 0) or last assembled (trial = 1) state;
 * either pointer may be NULL. */
int elastic_j2_get_state(elastic_ctx* ctx, int32_t trial, double* ep, double* alpha);

/* 3x3 block-Jacobi set-up of preconditioner code 6, block_precond with
 * nblock = 3 (Precond.f:163-358, cofactor inverse Precond.f:1676-1710): fills
 * m(6, neq/3 + mod(neq,3)) in the reference's layout (c11 c22 c33 c21 c31
 * c32 of each equation triple over the determinant, then 1/ad for each of
 * the last mod(neq,3) equations), bit-identical to the reference's, from
 * the device values that the solves apply. */
int elastic_get_bdiag(elastic_ctx* ctx, double* m);
/* block_precond (Precond.f:163-358) with block size bsz = 2 or 3: fills
 * m(w, neq/bsz + mod(neq,bsz)), w = bsz(bsz+1)/2 (2x2: a22/det, a11/det,
 * -a21/det of each equation pair; 3x3 as elastic_get_bdiag), then 1/ad for
 * each remaining equation; bit-identical to the reference's. */
int elastic_get_block_precond(elastic_ctx* ctx, int32_t bsz, double* m);

/* ----------------------------------------------------------- solving */
/* Solve A x = b.  Replaces solv -> pcg (Solv.f:13-139, Iters.f:252-524) for
 * solver 1, gmres (Iters.f:528-808) for 2 and pbicgstab (Iters.f:1079-1349)
 * for 4.  x0 = 0 (fnew = .true.); the stopping rule is the reference's:
 * PCG ||r||_M < tol ||b||_M, GMRES ||M r|| < tol ||M b||, BiCGSTAB
 * ||r|| < tol ||b||.  MIXED applies the FP32-value operator inside the
 * iteration and replaces the recursive residual by the FP64 true residual
 * b - A x whenever it has dropped by a factor 10 since the last replacement;
 * the stopping rule is tested on that FP64 residual.  st may be NULL. */
int elastic_solve(elastic_ctx* ctx, int32_t solver, int32_t precond, int32_t precision,
                  const double* b, double* x, double tol, int32_t maxit, int32_t ngram,
                  elastic_stats* st);

/* Drop-in for the reference's solver call, Solv.f:13-15:
 *   call solv(neq,nad,ip,ja,ad,au,al,m,b,x,tol,maxit,ngram,unsym,solver,
 *             neqf1i,neqf2i,neq3i,neq4i,neq_doti,i_fmapi,i_xfi,i_rcvsi,i_dspli)
 * Fortran passes every argument by reference (integer = int32, real*8 =
 * double, logical = int32, pool offsets = int64).  Behaviour: m is
 * overwritten with 1/ad (Solv.f:65-69); x receives the solution; the MPI
 * arguments are accepted and ignored (one process).  Uses one internal context
 * per device-0 process, rebuilt when neq, nad or any entry of ip/ja changes.
 * Returns the status (a Fortran caller may ignore it or declare a function). */
int32_t elastic_solv_(int32_t* neq, int32_t* nad, int32_t* ip, int32_t* ja, double* ad,
                      double* au, double* al, double* m, double* b, double* x, double* tol,
                      int32_t* maxit, int32_t* ngram, int32_t* unsym, int32_t* solver,
                      int32_t* neqf1i, int32_t* neqf2i, int32_t* neq3i, int32_t* neq4i,
                      int32_t* neq_doti, int64_t* i_fmapi, int64_t* i_xfi, int64_t* i_rcvsi,
                      int64_t* i_dspli);

/* No counterpart in the reference: y = (D + L + L^T) x (the product of
 * elastic_spmv, Matvec.f:528-621) for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Every pointer is a device pointer to arrays the caller
 * holds: ia(neq+1), ja(nad) (1-based CSRC pattern, which must be valid: the
 * call does not check it), ad(neq), al(nad), x(neq), y(neq).  The work is
 * enqueued on `stream` (a hipStream_t; NULL for the null stream) and the call
 * returns at once.  No context, no hipMalloc/hipFree: the transposed pattern
 * is built per call in a fixed 128 MiB device array of the library (one per
 * GPU, held until the process ends), which bounds the size:
 * 2 neq + 2 nad + neq/1024 + 2 <= 33,554,432 (ELASTIC_BAD_ARG beyond, with a
 * message on stderr).  The array keeps the last call's contents; calls that
 * overlap on one GPU (other streams or threads) race on it.  The FP64 result
 * is the reference's matvec_csrcsym1 bit for bit. */
int elastic_spmv_csrc_noalloc(int32_t neq, int64_t nad, const int32_t* ia, const int32_t* ja,
                              const double* ad, const double* al, const double* x, double* y,
                              void* stream);

/* ------------------------------------------------ resident benchmarking */
/* Uploads b to the device copy used by elastic_solve_resident. */
int elastic_upload_rhs(elastic_ctx* ctx, const double* b);
/* Solves with b already resident (elastic_upload_rhs) and leaves x on the
 * device; st->t_device_ms is the device time of the iterations alone. */
int elastic_solve_resident(elastic_ctx* ctx, int32_t solver, int32_t precond, int32_t precision,
                           double tol, int32_t maxit, int32_t ngram, elastic_stats* st);
int elastic_download_x(elastic_ctx* ctx, double* x);
/* Average device time (ms) of `reps` symmetric SpMV launches on resident data
 * (precision as in elastic_spmv), measured with HIP events on the library's
 * stream; also returns the tile-kernel time alone in *kernel_ms. */
int elastic_time_spmv(elastic_ctx* ctx, int32_t precision, int32_t reps, double* total_ms,
                      double* kernel_ms);
/* Layout facts: ntiles, max tile nnz, external partials, merge entries. */
int elastic_layout_info(elastic_ctx* ctx, int64_t* info4);
/* Tile size of the scalar row tiles (nonzeros per tile, 64..3840); takes
 * effect at the next elastic_set_structure. */
int elastic_set_tile_nnz(elastic_ctx* ctx, int32_t tile_nnz);
/* Tile size of the node-blocked layout (node blocks per tile, 32..1024; by
 * default 1024 for 3-dof systems of at least 32,768 nodes, else 512);
 * takes effect at the next elastic_set_structure. */
int elastic_set_tile_blocks(elastic_ctx* ctx, int32_t tile_blocks);
/* Tuning knobs (no counterpart in the reference): "tile_blocks" /
 * "tile_nnz" (tile size; rebuilds the layout at once when a structure is
 * set), "accum32" (1: the FP32 operator stages x and sums in FP32, and the
 * mixed PCG with Jacobi keeps p, r and the Jacobi scale in FP32, faster but
 * less accurate, which ill-conditioned systems may not tolerate; 0: in
 * FP64; until set, 1 for PCG on 3-dof systems of at least 32,768 nodes,
 * else 0),
 * "lanes64" / "lanes32" (256 or 512 lanes per tile of the node-blocked FP64 /
 * FP32-value product, 4 or 2 blocks per lane; default 256), "tune_tiles"
 * (0: elastic_autotune keeps the tile size),
 * "check_every" (iterations between convergence polls),
 * "batch_precond" (2: the context's batched solves use Jacobi, the default;
 * 6: node-block Jacobi, the inverses of block_precond with the layout's node
 * size, 2x2 in 2-D and 3x3 in 3-D, applied as op_block_precond does),
 * "time_products" (S > 0: PCG brackets the product of every S-th iteration
 * with HIP events on its stream; 0, the default: off),
 * "node_order" (1: the node-blocked tiles of a 3-dof layout are compact
 * breadth-first clusters of the node graph, in whose node order the solves
 * iterate -- b and x are exchanged with the caller in the CSRC order as
 * always; 0: tiles of consecutive CSRC node rows; until set, 1 for 3-dof
 * systems of at least 32,768 nodes, else 0; rebuilds the layout at once when
 * a structure is set). */
int elastic_set_option(elastic_ctx* ctx, const char* name, int64_t value);
/* Average device time of the PCG iteration products sampled under the
 * "time_products" option, and how many were sampled, since the last reset
 * (reset != 0 clears the totals after reading).  Measurement only; no
 * counterpart in the reference (its matvectime timer, Iters.f:383-386, is
 * the closest). */
int elastic_product_timing(elastic_ctx* ctx, double* avg_ms, int64_t* count, int32_t reset);
/* Autotuner (no counterpart in the reference): with values resident, runs
 * a short resident PCG of the given precision (max(16, reps) iterations, b
 * = ones; the caller's resident b is put back, the resident x is
 * overwritten) for each candidate layout -- node-blocked: 256, 512, 768 and
 * 1024 blocks per tile, each with 256 and 512 lanes per tile (4 or 2 blocks
 * per lane); scalar rows: 1024 to 3840 nonzeros per tile -- and keeps the
 * candidate whose products are fastest on average (HIP events around every
 * product of the run).  The tile size is shared by both operators; the lanes
 * are kept for the given precision.  With the option "tune_tiles" = 0 only
 * the lanes are tuned.  Returns the tile size and the product time.  Not
 * while the context is in a batch. */
int elastic_autotune(elastic_ctx* ctx, int32_t precision, int32_t reps, int32_t* best,
                     double* best_ms);
/* Current value of a knob of elastic_set_option, or "lanes64" / "lanes32"
 * (lanes per tile of the FP64 / FP32-value product) or "layout_nb" (node
 * block size of the layout, 1 scalar rows, 0 the gather product);
 * "node_order" reads the order of the layout in use (1 compact). */
int elastic_get_option(elastic_ctx* ctx, const char* name, int64_t* value);

/* ------------------------------------------------------ batched serving */
/* No counterpart in the reference (its solver serves one system per call).
 * A batch iterates many independent systems together: every PCG phase is one
 * launch over all active instances, and a step of iters_per_step iterations
 * is captured once as a HIP graph (use_graph != 0) and replayed while
 * instances join and leave.  Each instance is a context holding its
 * structure, values (node-blocked layout) and b (elastic_upload_rhs); its
 * method is the PCG of elastic_solve (Iters.f:252-524, Jacobi), FP64 or
 * MIXED, and its result is bit for bit the same alone and in any mix.  x is
 * left in the context (elastic_download_x).  One host thread drives a batch;
 * other threads may prepare contexts meanwhile. */
typedef struct elastic_batch elastic_batch;
int elastic_batch_create(elastic_batch** b, int32_t device, int32_t max_slots,
                         int32_t iters_per_step, int32_t use_graph);
int elastic_batch_destroy(elastic_batch* b);
const char* elastic_batch_last_error(const elastic_batch* b);
/* Admits ctx as a new instance (x0 = 0, stop when ||r||_M < tol ||b||_M or
 * at maxit); ELASTIC_BUSY when every slot is taken.  The context must
 * not be used by the caller until the instance is retired by a poll. */
int elastic_batch_submit(elastic_batch* b, elastic_ctx* ctx, int32_t precision, double tol,
                         int32_t maxit, int64_t tag);
/* Enqueues one step for every active instance (asynchronous). */
int elastic_batch_step(elastic_batch* b);
/* Waits for the last step and retires finished instances: their tags (and,
 * when stats != NULL, their statistics, with t_total_ms = time since submit)
 * go to the first *ndone entries; their slots are free again. */
int elastic_batch_poll(elastic_batch* b, int64_t* tags, elastic_stats* stats, int32_t cap,
                       int32_t* ndone);
int elastic_batch_active(elastic_batch* b, int32_t* nactive);

#ifdef __cplusplus
}
#endif
#endif
