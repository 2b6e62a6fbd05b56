! Test infrastructure: an independent Fortran check of the integrals file layout. It restates, in
! its own code, the reference's writer and reader of and.integrals.bin for one subdomain
! (WriteIntegralsToDiskStokes / ReadIntegralsFileStokes, src/lib_Stokes.f90:2742-2872;
! WriteIntegralsToDisk / ReadIntegralsFile, src/lib_int.f90:136-221; WrSMat / RdSMat and
! wrvec / rdvec with 8192-value records, src/lib_util.f90:201-287), so that the Fortran runtime's
! own unformatted records decide whether the library writes and reads the layout right.
!   integrals_check write stokes|laplace FILE NSP NN NQ   writes the test pattern
!   integrals_check read  stokes|laplace FILE NSP NN NQ   reads a file, prints OK or MISMATCH
! Pattern: matrix k (1-based in the file's order), entry (i, j) = 1e6 k + 1e3 i + j; errors (0.25, 0.5).
program integrals_check
  implicit none
  integer, parameter :: rk = selected_real_kind(8)
  character(len=16) :: mode, kind
  character(len=512) :: path
  character(len=32) :: arg
  integer :: nsp, nn, nq, nmat, k, ncol, lun, id, a, b, c, bad
  real(rk) :: e1, e2
  real(rk), allocatable :: mat(:, :)
  call get_command_argument(1, mode)
  call get_command_argument(2, kind)
  call get_command_argument(3, path)
  call get_command_argument(4, arg); read (arg, *) nsp
  call get_command_argument(5, arg); read (arg, *) nn
  call get_command_argument(6, arg); read (arg, *) nq
  nmat = 2
  if (trim(kind) == 'stokes') nmat = 18
  lun = 12
  bad = 0
  if (trim(mode) == 'write') then
    open (lun, file=trim(path), form='unformatted', status='replace')
    if (trim(kind) == 'stokes') write (lun) 1407
    write (lun) nsp, nn, nq
    write (lun) 0.25_rk, 0.5_rk
    do k = 1, nmat
      ncol = ncols(k)
      allocate (mat(nsp, ncol))
      call pattern(mat, k, nsp, ncol)
      call wr_mat(mat, nsp, ncol, lun)
      deallocate (mat)
    end do
    close (lun)
    print '(a)', 'WRITTEN'
  else
    open (lun, file=trim(path), form='unformatted', status='old')
    if (trim(kind) == 'stokes') then
      read (lun) id
      if (id /= 1407) bad = bad + 1
    end if
    read (lun) a, b, c
    if (a /= nsp .or. b /= nn .or. c /= nq) bad = bad + 1
    read (lun) e1, e2
    if (e1 /= 0.25_rk .or. e2 /= 0.5_rk) bad = bad + 1
    do k = 1, nmat
      ncol = ncols(k)
      allocate (mat(nsp, ncol))
      call rd_mat(mat, nsp, ncol, lun)
      call check(mat, k, nsp, ncol, bad)
      deallocate (mat)
    end do
    close (lun)
    if (bad == 0) then
      print '(a)', 'OK'
    else
      print '(a, i0)', 'MISMATCH ', bad
    end if
  end if
contains
  integer function ncols(k)
    integer, intent(in) :: k
    ! Stokes: T (6), G (6), prT (3), prG (3); Laplace: H, G
    ncols = nq
    if (nmat == 2 .and. k == 1) ncols = nn
    if (nmat == 18 .and. (k <= 6 .or. (k >= 13 .and. k <= 15))) ncols = nn
  end function
  subroutine pattern(m, k, nr, nc)
    integer, intent(in) :: k, nr, nc
    real(rk), intent(out) :: m(nr, nc)
    integer :: i, j
    do j = 1, nc
      do i = 1, nr
        m(i, j) = 1.0e6_rk * k + 1.0e3_rk * i + real(j, rk)
      end do
    end do
  end subroutine
  subroutine check(m, k, nr, nc, bad)
    integer, intent(in) :: k, nr, nc
    real(rk), intent(in) :: m(nr, nc)
    integer, intent(inout) :: bad
    integer :: i, j
    do j = 1, nc
      do i = 1, nr
        if (m(i, j) /= 1.0e6_rk * k + 1.0e3_rk * i + real(j, rk)) bad = bad + 1
      end do
    end do
  end subroutine
  subroutine wr_mat(m, nr, nc, io)  ! column by column, 8192 values per record, then the rest
    integer, intent(in) :: nr, nc, io
    real(rk), intent(in) :: m(nr, nc)
    integer :: i, j, blk, nb, k0
    do j = 1, nc
      nb = nr / 8192
      do blk = 1, nb
        k0 = 8192 * (blk - 1)
        write (io) (m(i, j), i = k0 + 1, k0 + 8192)
      end do
      k0 = 8192 * nb
      write (io) (m(i, j), i = k0 + 1, nr)
    end do
  end subroutine
  subroutine rd_mat(m, nr, nc, io)
    integer, intent(in) :: nr, nc, io
    real(rk), intent(out) :: m(nr, nc)
    integer :: i, j, blk, nb, k0
    do j = 1, nc
      nb = nr / 8192
      do blk = 1, nb
        k0 = 8192 * (blk - 1)
        read (io) (m(i, j), i = k0 + 1, k0 + 8192)
      end do
      k0 = 8192 * nb
      read (io) (m(i, j), i = k0 + 1, nr)
    end do
  end subroutine
end program
