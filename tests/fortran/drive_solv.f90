! Test driver for the Fortran drop-in: reads a CSRC system written by the test
! (stream file: neq, nad, ia(neq+1), ja(nad), ad(neq), al(nad), b(neq),
! tol, maxit), calls solv exactly as Main.f:820-826 does (sequential MPI
! arguments), and writes m(neq) and x(neq) to a second stream file.
program drive_solv
  implicit none
  integer :: neq, nad, maxit, u, ngram, solver
  integer :: neqf1i, neqf2i, neq3i, neq4i, neq_doti
  integer(8) :: i_fmapi, i_xfi, i_rcvsi, i_dspli
  integer, allocatable :: ia(:), ja(:)
  real(8), allocatable :: ad(:), al(:), m(:), b(:), x(:)
  real(8) :: tol
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) neq, nad
  allocate(ia(neq + 1), ja(max(nad, 1)), ad(neq), al(max(nad, 1)), m(neq), b(neq), x(neq))
  read(u) ia, ja(1:nad), ad, al(1:nad), b, tol, maxit
  close(u)
  x = 0.d0
  ngram = 50
  solver = 1
  neqf1i = 0; neqf2i = 0; neq3i = 0; neq4i = 0; neq_doti = neq
  i_fmapi = 1; i_xfi = 1; i_rcvsi = 1; i_dspli = 1
  call solv(neq, nad, ia, ja, ad, al, al, m, b, x, tol, maxit, ngram, .false., solver, &
            neqf1i, neqf2i, neq3i, neq4i, neq_doti, i_fmapi, i_xfi, i_rcvsi, i_dspli)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) m, x
  close(u)
end program drive_solv

! the reference's stop_mef (Paralelo.f:1123) ends the run; here: exit status 3
subroutine stop_mef()
  stop 3
end subroutine stop_mef
