! Replacement of the reference's solver call (Solv.f:13-15) by the GPU
! library: same name, same 24 arguments, same effects on m and x.  Link this
! object instead of the reference's solv (Solv.f:13-440) and add
! -lelastic_hip; Main.f:820-826 then needs no change.  A nonzero status stops
! the program through stop_mef, as the reference does on a zero diagonal
! (Iters.f:337-342) or when maxit is reached (Iters.f:429-431).
subroutine solv(neq, nad, ip, ja, ad, au, al, m, b, x, tol, maxit, ngram, unsym, solver, &
                neqf1i, neqf2i, neq3i, neq4i, neq_doti, i_fmapi, i_xfi, i_rcvsi, i_dspli)
  use elastic_hip
  implicit none
  integer :: neq, nad, ip(*), ja(*), maxit, ngram, solver
  integer :: neqf1i, neqf2i, neq3i, neq4i, neq_doti
  integer(8) :: i_fmapi, i_xfi, i_rcvsi, i_dspli
  real(8) :: ad(*), au(*), al(*), m(*), b(*), x(*), tol
  logical :: unsym
  integer(c_int32_t) :: iunsym, rc
  iunsym = 0
  if (unsym) iunsym = 1
  rc = elastic_solv(neq, nad, ip, ja, ad, au, al, m, b, x, tol, maxit, ngram, iunsym, solver, &
                    neqf1i, neqf2i, neq3i, neq4i, neq_doti, i_fmapi, i_xfi, i_rcvsi, i_dspli)
  if (rc /= ELASTIC_OK) then
    print '(a,i3)', ' solv (gfx950): stopped with status ', rc
    call stop_mef()
  endif
end subroutine solv
