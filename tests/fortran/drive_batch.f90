! Test driver for the batched-serving and tuning entry points from Fortran:
! reads a CSRC system (the stream file of drive_solv), puts it into two
! contexts, sets and reads back a tuning knob, admits both contexts into one
! batch (FP64 and MIXED), steps and polls until both are retired, and writes
! x(FP64), x(MIXED), the knob read back and the two iteration counts.
program drive_batch
  use iso_c_binding
  use elastic_hip
  implicit none
  integer :: neq, nad, maxit, u, k, rc, ndone, nact, got
  integer, allocatable :: ia(:), ja(:)
  real(8), allocatable :: ad(:), al(:), b(:), x64(:), x32(:)
  real(8) :: tol
  type(c_ptr) :: ctx(2), bat
  integer(c_int64_t) :: tags(4), knob
  type(elastic_stats) :: st(4)
  real(8) :: iters(2)
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) neq, nad
  allocate(ia(neq + 1), ja(max(nad, 1)), ad(neq), al(max(nad, 1)), b(neq), x64(neq), x32(neq))
  read(u) ia, ja(1:nad), ad, al(1:nad), b, tol, maxit
  close(u)
  do k = 1, 2
    call check(elastic_create(ctx(k), 0), 'create')
    call check(elastic_set_structure(ctx(k), neq, int(nad, c_int64_t), ia, ja), 'structure')
    call check(elastic_set_values(ctx(k), ad, al), 'values')
    call check(elastic_upload_rhs(ctx(k), b), 'rhs')
  end do
  call check(elastic_set_option(ctx(2), 'check_every'//c_null_char, 8_c_int64_t), 'set_option')
  call check(elastic_get_option(ctx(2), 'check_every'//c_null_char, knob), 'get_option')
  call check(elastic_batch_create(bat, 0, 4, 16, 1), 'batch_create')
  call check(elastic_batch_submit(bat, ctx(1), ELASTIC_FP64, tol, maxit, 1_c_int64_t), 'submit')
  call check(elastic_batch_submit(bat, ctx(2), ELASTIC_MIXED, tol, maxit, 2_c_int64_t), 'submit')
  got = 0
  iters = 0.d0
  do while (got < 2)
    call check(elastic_batch_step(bat), 'step')
    call check(elastic_batch_poll(bat, tags, st, 4, ndone), 'poll')
    do k = 1, ndone
      if (st(k)%status /= ELASTIC_OK) stop 4
      iters(tags(k)) = st(k)%iters
    end do
    got = got + ndone
  end do
  call check(elastic_batch_active(bat, nact), 'active')
  if (nact /= 0) stop 5
  call check(elastic_download_x(ctx(1), x64), 'download')
  call check(elastic_download_x(ctx(2), x32), 'download')
  call check(elastic_batch_destroy(bat), 'batch_destroy')
  do k = 1, 2
    call check(elastic_destroy(ctx(k)), 'destroy')
  end do
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) x64, x32, real(knob, 8), iters
  close(u)
contains
  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= ELASTIC_OK) then
      print '(a,a,a,i3)', ' drive_batch: ', what, ' returned ', status
      stop 3
    endif
  end subroutine check
end program drive_batch
