! ISO_C_BINDING interfaces of libelastic_hip.so (declared in
! include/elastic_hip.h) for Fortran callers of the Elastic solve path.
!
! Arguments follow the C prototypes: arrays are passed as the caller's
! arrays (1-based index values, as the reference's CSRC holds them), scalars
! with VALUE where the C side takes them by value.  Every function returns
! the library's status code (0 = ok; see the ELASTIC_* parameters).
module elastic_hip
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: ELASTIC_OK = 0, ELASTIC_ZERO_DIAG = 1, &
       ELASTIC_MAXIT = 2, ELASTIC_NEG_JACOBIAN = 3, ELASTIC_OOM = 4, &
       ELASTIC_HIP_ERROR = 5, ELASTIC_BAD_ARG = 6, ELASTIC_NOT_READY = 7, &
       ELASTIC_BREAKDOWN = 8, ELASTIC_BUSY = 9
  integer(c_int32_t), parameter :: ELASTIC_SOLVER_PCG = 1, ELASTIC_SOLVER_GMRES = 2, &
       ELASTIC_SOLVER_BICGSTAB = 4
  integer(c_int32_t), parameter :: ELASTIC_PRECOND_NONE = 1, ELASTIC_PRECOND_DIAG = 2, &
       ELASTIC_PRECOND_BDIAG3 = 6
  integer(c_int32_t), parameter :: ELASTIC_FP64 = 64, ELASTIC_MIXED = 32

  type, bind(C) :: elastic_stats
    integer(c_int32_t) :: iters, outer, status, spmv_count
    real(c_double) :: xkx, norm_x, norm_r, norm_m_r, conv, rel_res, t_device_ms, t_total_ms
    integer(c_int32_t) :: op64_from, reserved
  end type elastic_stats

  interface
    integer(c_int) function elastic_create(ctx, device) bind(C, name="elastic_create")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(out) :: ctx
      integer(c_int32_t), value :: device
    end function
    integer(c_int) function elastic_destroy(ctx) bind(C, name="elastic_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function elastic_last_error(ctx) bind(C, name="elastic_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function elastic_set_structure(ctx, neq, nad, ia, ja) &
         bind(C, name="elastic_set_structure")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: neq
      integer(c_int64_t), value :: nad
      integer(c_int32_t), intent(in) :: ia(*), ja(*)
    end function
    integer(c_int) function elastic_set_values(ctx, ad, al) bind(C, name="elastic_set_values")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ad(*), al(*)
    end function
    integer(c_int) function elastic_get_values(ctx, ad, al) bind(C, name="elastic_get_values")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ad(*), al(*)
    end function
    integer(c_int) function elastic_spmv(ctx, precision, x, y) bind(C, name="elastic_spmv")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: precision
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function
    integer(c_int) function elastic_solve(ctx, solver, precond, precision, b, x, tol, maxit, &
         ngram, st) bind(C, name="elastic_solve")
      import :: c_int, c_ptr, c_int32_t, c_double, elastic_stats
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: solver, precond, precision, maxit, ngram
      real(c_double), intent(in) :: b(*)
      real(c_double), intent(out) :: x(*)
      real(c_double), value :: tol
      type(elastic_stats), intent(out) :: st
    end function
    integer(c_int) function elastic_reord(nnode, numel, nen, ix, perm) bind(C, name="elastic_reord")
      import :: c_int, c_int32_t
      integer(c_int32_t), value :: nnode, numel, nen
      integer(c_int32_t), intent(in) :: ix(*)
      integer(c_int32_t), intent(out) :: perm(*)
    end function
    integer(c_int) function elastic_numeq(nnode, ndf, perm, id, neq) bind(C, name="elastic_numeq")
      import :: c_int, c_int32_t
      integer(c_int32_t), value :: nnode, ndf
      integer(c_int32_t), intent(in) :: perm(*)
      integer(c_int32_t), intent(inout) :: id(*)
      integer(c_int32_t), intent(out) :: neq
    end function
    ! first call with ja = c_null_ptr returns nad and fills ia; the second fills ja
    integer(c_int) function elastic_csrstruct(nnode, numel, nen, ndf, neq, ix, id, ia, ja, nad) &
         bind(C, name="elastic_csrstruct")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      integer(c_int32_t), value :: nnode, numel, nen, ndf, neq
      integer(c_int32_t), intent(in) :: ix(*), id(*)
      integer(c_int32_t), intent(out) :: ia(*)
      type(c_ptr), value :: ja
      integer(c_int64_t), intent(out) :: nad
    end function
    integer(c_int) function elastic_set_mesh(ctx, nnode, numel, nen, ndf, ndm, numat, ix, x, id, &
         ie, e, eload, nloads, load_type, load_nparc, load_vals) bind(C, name="elastic_set_mesh")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nnode, numel, nen, ndf, ndm, numat, nloads
      integer(c_int32_t), intent(in) :: ix(*), id(*), ie(*), eload(*), load_type(*), load_nparc(*)
      real(c_double), intent(in) :: x(*), e(*), load_vals(*)
    end function
    integer(c_int) function elastic_assemble(ctx, isw, lhs, rhs, u, b) bind(C, name="elastic_assemble")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: isw, lhs, rhs
      real(c_double), intent(in) :: u(*)
      real(c_double), intent(inout) :: b(*)
    end function
    ! tform_mec (Pform.f:258-421) on the GPU: nodal stresses tx(ntn, nnode);
    ! tx0 is c_loc of the initial stresses or c_null_ptr
    integer(c_int) function elastic_stress(ctx, u, from_x, tx0, tx) bind(C, name="elastic_stress")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
      integer(c_int32_t), value :: from_x
      type(c_ptr), value :: tx0
      real(c_double), intent(out) :: tx(*)
    end function
    integer(c_int) function elastic_upload_rhs(ctx, b) bind(C, name="elastic_upload_rhs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: b(*)
    end function
    integer(c_int) function elastic_solve_resident(ctx, solver, precond, precision, tol, maxit, &
         ngram, st) bind(C, name="elastic_solve_resident")
      import :: c_int, c_ptr, c_int32_t, c_double, elastic_stats
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: solver, precond, precision, maxit, ngram
      real(c_double), value :: tol
      type(elastic_stats), intent(out) :: st
    end function
    integer(c_int) function elastic_download_x(ctx, x) bind(C, name="elastic_download_x")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: x(*)
    end function
    ! the drop-in with the reference's solv argument list (all by reference)
    integer(c_int32_t) function elastic_solv(neq, nad, ip, ja, ad, au, al, m, b, x, tol, &
         maxit, ngram, unsym, solver, neqf1i, neqf2i, neq3i, neq4i, neq_doti, i_fmapi, &
         i_xfi, i_rcvsi, i_dspli) bind(C, name="elastic_solv_")
      import :: c_int32_t, c_int64_t, c_double
      integer(c_int32_t) :: neq, nad, ip(*), ja(*), maxit, ngram, unsym, solver
      integer(c_int32_t) :: neqf1i, neqf2i, neq3i, neq4i, neq_doti
      integer(c_int64_t) :: i_fmapi, i_xfi, i_rcvsi, i_dspli
      real(c_double) :: ad(*), au(*), al(*), m(*), b(*), x(*), tol
    end function
    integer(c_int) function elastic_time_assembly(ctx, reps, ms) bind(C, name="elastic_time_assembly")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: reps
      real(c_double), intent(out) :: ms
    end function

    ! J2 plasticity on hex8 meshes (state on the device; after elastic_set_mesh)
    integer(c_int) function elastic_j2_setup(ctx, sy, h, ubar) bind(C, name="elastic_j2_setup")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: sy, h
      real(c_double), intent(in) :: ubar(*)
    end function
    integer(c_int) function elastic_j2_increment(ctx, lambda) bind(C, name="elastic_j2_increment")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: lambda
    end function
    integer(c_int) function elastic_j2_assemble(ctx, lhs, rnorm) bind(C, name="elastic_j2_assemble")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: lhs
      real(c_double), intent(out) :: rnorm
    end function
    integer(c_int) function elastic_j2_correct(ctx) bind(C, name="elastic_j2_correct")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function elastic_j2_step(ctx, s) bind(C, name="elastic_j2_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: s
    end function
    integer(c_int) function elastic_j2_commit(ctx, alpha_max, nplastic) bind(C, name="elastic_j2_commit")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: alpha_max
      integer(c_int64_t), intent(out) :: nplastic
    end function
    integer(c_int) function elastic_j2_get_u(ctx, u) bind(C, name="elastic_j2_get_u")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: u(*)
    end function
    integer(c_int) function elastic_j2_get_residual(ctx, r) bind(C, name="elastic_j2_get_residual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: r(*)
    end function
    ! ep and alpha are c_loc of the caller's arrays, or c_null_ptr
    integer(c_int) function elastic_j2_get_state(ctx, trial, ep, alpha) bind(C, name="elastic_j2_get_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: trial
      type(c_ptr), value :: ep, alpha
    end function

    ! block_precond (Precond.f:163-358): m(6, neq/3 + mod(neq,3)) or, for
    ! bsz = 2, m(3, neq/2 + mod(neq,2))
    integer(c_int) function elastic_get_bdiag(ctx, m) bind(C, name="elastic_get_bdiag")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: m(*)
    end function
    integer(c_int) function elastic_get_block_precond(ctx, bsz, m) &
         bind(C, name="elastic_get_block_precond")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: bsz
      real(c_double), intent(out) :: m(*)
    end function

    ! device pointers and a hipStream_t (c_null_ptr: the null stream)
    integer(c_int) function elastic_spmv_csrc_noalloc(neq, nad, ia, ja, ad, al, x, y, stream) &
         bind(C, name="elastic_spmv_csrc_noalloc")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      integer(c_int32_t), value :: neq
      integer(c_int64_t), value :: nad
      type(c_ptr), value :: ia, ja, ad, al, x, y, stream
    end function

    ! measurement and tuning (names are C strings: 'accum32'//c_null_char)
    integer(c_int) function elastic_time_spmv(ctx, precision, reps, total_ms, kernel_ms) &
         bind(C, name="elastic_time_spmv")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: precision, reps
      real(c_double), intent(out) :: total_ms, kernel_ms
    end function
    integer(c_int) function elastic_layout_info(ctx, info4) bind(C, name="elastic_layout_info")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: info4(4)
    end function
    integer(c_int) function elastic_set_tile_nnz(ctx, tile_nnz) bind(C, name="elastic_set_tile_nnz")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tile_nnz
    end function
    integer(c_int) function elastic_set_tile_blocks(ctx, tile_blocks) &
         bind(C, name="elastic_set_tile_blocks")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tile_blocks
    end function
    integer(c_int) function elastic_set_option(ctx, name, val) bind(C, name="elastic_set_option")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t), value :: val
    end function
    integer(c_int) function elastic_get_option(ctx, name, val) bind(C, name="elastic_get_option")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t), intent(out) :: val
    end function
    integer(c_int) function elastic_product_timing(ctx, avg_ms, count, reset) &
         bind(C, name="elastic_product_timing")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: avg_ms
      integer(c_int64_t), intent(out) :: count
      integer(c_int32_t), value :: reset
    end function
    integer(c_int) function elastic_autotune(ctx, precision, reps, best, best_ms) &
         bind(C, name="elastic_autotune")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: precision, reps
      integer(c_int32_t), intent(out) :: best
      real(c_double), intent(out) :: best_ms
    end function

    ! batched serving: many contexts iterated together (one host thread)
    integer(c_int) function elastic_batch_create(b, device, max_slots, iters_per_step, use_graph) &
         bind(C, name="elastic_batch_create")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(out) :: b
      integer(c_int32_t), value :: device, max_slots, iters_per_step, use_graph
    end function
    integer(c_int) function elastic_batch_destroy(b) bind(C, name="elastic_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    type(c_ptr) function elastic_batch_last_error(b) bind(C, name="elastic_batch_last_error")
      import :: c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function elastic_batch_submit(b, ctx, precision, tol, maxit, tag) &
         bind(C, name="elastic_batch_submit")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: b, ctx
      integer(c_int32_t), value :: precision, maxit
      real(c_double), value :: tol
      integer(c_int64_t), value :: tag
    end function
    integer(c_int) function elastic_batch_step(b) bind(C, name="elastic_batch_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function elastic_batch_poll(b, tags, stats, cap, ndone) &
         bind(C, name="elastic_batch_poll")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, elastic_stats
      type(c_ptr), value :: b
      integer(c_int64_t), intent(out) :: tags(*)
      type(elastic_stats), intent(out) :: stats(*)
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: ndone
    end function
    integer(c_int) function elastic_batch_active(b, nactive) bind(C, name="elastic_batch_active")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), intent(out) :: nactive
    end function
  end interface
end module elastic_hip
