! Minimal Fortran host for the device library, calling it the way the reference's drivers would:
!   resistance_host <mesh.msh> [fp64|mixed]
! 1. reads a Gmsh 2.2 ASCII surface mesh (3-node triangles, first tag = wall id),
! 2. assembles the Laplace H/G matrices on the device in the reference's column-major layout
!    (sdFormIntegralMatrices, src/lib_int.f90:225-323), builds the Dirichlet system for u = x in
!    1-based CSR, and solves it with lsqr_precon/lsqr_solve called through the reference's
!    implicit F77 interface (src/lib_CRS.f90:77-85); prints max |q + n_x| (exact flux q = -n_x),
! 3. computes the 6x6 resistance tensor of wall 1 (StokesInlet3sidesForceCRS /
!    TorqueOmegaCRS, src/lib_particle.f90:176-476) in the exterior form and prints it;
! 4. the velocity at the mesh's centre for a uniform boundary flow (domain values).
program resistance_host
  use iso_c_binding
  use andromeda_hip
  implicit none
  character(len=512) :: path, mode
  real(c_double), allocatable, target :: xyz(:, :), H(:, :), G(:, :), a(:), q(:), b(:), x(:)
  integer(c_int32_t), allocatable, target :: con(:, :), bcid(:), ja(:)
  integer(c_int64_t), allocatable :: ia(:)
  real(c_double), target :: normmul(1)
  real(c_double) :: err(2), FT(36), eps, nrm(3), e1(3), e2(3), worst
  real(c_double), allocatable :: ub(:, :), qb(:, :)
  real(c_double) :: pt(3), vel(3), pd(1), ed(1)
  type(andr_mesh) :: mesh
  type(andr_resistance_opts) :: opts
  type(andr_resistance_report) :: rep
  type(c_ptr) :: ctx
  integer :: nn, ne, nsp, i, j, k, rc
  integer(c_int32_t) :: m32, n32, maxit, nit, ierr
  integer(c_int64_t) :: nnz

  if (command_argument_count() < 1) stop "usage: resistance_host mesh.msh [fp64|mixed]"
  call get_command_argument(1, path)
  mode = "fp64"
  if (command_argument_count() >= 2) call get_command_argument(2, mode)
  call read_gmsh(trim(path), xyz, con, bcid)
  nn = size(xyz, 2)
  ne = size(con, 2)
  nsp = nn + ne
  normmul = 1.0d0
  mesh = andr_mesh(nn, ne, maxval(bcid), c_loc(xyz), c_loc(con), c_loc(bcid), c_loc(normmul), 1, 5, 3, 4)
  rc = andr_create(ctx, 0)
  if (rc /= ANDR_OK) stop "andr_create failed"
  write(*, '(a, i0, a, i0, a)') "mesh: ", nn, " nodes, ", ne, " triangles"

  ! ---- Laplace: assembly + LSQR through the F77 seam ----------------------------------------
  allocate(H(nsp, nn), G(nsp, ne))
  rc = andr_assemble_laplace(ctx, mesh, ANDR_COL_MAJOR, H, G, err)
  if (rc /= ANDR_OK) then
    print *, "andr_assemble_laplace: ", andr_error_message(ctx)
    stop 1
  end if
  write(*, '(a, 2es12.4)') "laplace self-check errors: ", err
  ! centroid rows: G q = H (-u) with u = x at the nodes
  nnz = int(ne, c_int64_t) * ne
  allocate(ia(ne + 1), ja(nnz), a(nnz), q(ne), b(ne), x(ne))
  k = 0
  do i = 1, ne
    ia(i) = k + 1
    do j = 1, ne
      k = k + 1
      ja(k) = j
      a(k) = G(nn + i, j)
    end do
    b(i) = -sum(H(nn + i, :) * xyz(1, :))
  end do
  ia(ne + 1) = nnz + 1
  m32 = ne
  n32 = ne
  maxit = 0
  eps = 1.0d-12
  x = 0.0d0
  call lsqr_precon(m32, n32, nnz, q, ia, ja, a)
  call lsqr_solve(m32, n32, nnz, maxit, eps, nit, ierr, q, ia, ja, a, b, x)
  worst = 0.0d0
  do i = 1, ne
    e1 = xyz(:, con(2, i)) - xyz(:, con(1, i))
    e2 = xyz(:, con(3, i)) - xyz(:, con(1, i))
    nrm = [e1(2) * e2(3) - e1(3) * e2(2), e1(3) * e2(1) - e1(1) * e2(3), e1(1) * e2(2) - e1(2) * e2(1)]
    nrm = nrm / norm2(nrm)
    worst = max(worst, abs(x(i) + nrm(1)))
  end do
  write(*, '(a, i0, a, i0, a, i0, a, es12.4)') "lsqr: maxit ", maxit, " nit ", nit, " ierr ", ierr, &
      "  max|q + n_x| ", worst

  ! ---- resistance tensor -------------------------------------------------------------------
  opts%form = ANDR_FORM_EXTERIOR
  opts%precision = ANDR_PREC_FP64
  if (trim(mode) == "mixed") opts%precision = ANDR_PREC_MIXED
  opts%mu = 1.0d0
  opts%eps = 0.0d0
  opts%maxit = 0
  opts%center = 0.0d0
  opts%restart = 0
  opts%problems = ANDR_PROBLEMS_RIGID
  rc = andr_resistance(ctx, mesh, opts, FT, rep)
  if (rc /= ANDR_OK) then
    print *, "andr_resistance: ", andr_error_message(ctx)
    stop 1
  end if
  write(*, '(a, 6i6)') "resistance iterations: ", rep%nit
  do k = 0, 5
    write(*, '(a, i0, a, 6es16.8)') "FT ", k, ": ", FT(6 * k + 1:6 * k + 6)
  end do
  ! 4. domain value at the centre of the mesh for the uniform flow u = (1, 0, 0), q = 0: the closure
  !    of the integrals returns the flow (GetStokesDomainValue, src/lib_Stokes.f90:3043-3390)
  allocate(ub(3, nn), qb(3, ne))
  ub = 0.0d0
  ub(1, :) = 1.0d0
  qb = 0.0d0
  do i = 1, 3
    pt(i) = sum(xyz(i, :)) / nn
  end do
  rc = andr_stokes_domain_values(ctx, mesh, 1_c_int32_t, pt, ub, qb, 1.0d0, vel, pd, ed)
  if (rc /= ANDR_OK) then
    print *, "andr_stokes_domain_values: ", andr_error_message(ctx)
    stop 1
  end if
  write(*, '(a, 3es16.8, a, es12.4)') "domain velocity: ", vel, " closure error: ", ed(1)
  rc = andr_destroy(ctx)

contains

  subroutine read_gmsh(fname, xyz, con, bcid)
    character(len=*), intent(in) :: fname
    real(c_double), allocatable, intent(out) :: xyz(:, :)
    integer(c_int32_t), allocatable, intent(out) :: con(:, :), bcid(:)
    character(len=256) :: line
    integer :: u, n, i, id, etype, ntag, nt, ios, vals(16)
    integer, allocatable :: remap(:)
    integer :: maxid
    real(c_double) :: p(3)
    integer(c_int32_t), allocatable :: tcon(:, :), tb(:)
    open(newunit=u, file=fname, status="old", action="read")
    do
      read(u, '(a)') line
      if (trim(line) == "$Nodes") exit
    end do
    read(u, *) n
    allocate(xyz(3, n), remap(0:0))
    maxid = 0
    block
      integer, allocatable :: ids(:)
      allocate(ids(n))
      do i = 1, n
        read(u, *) ids(i), p
        xyz(:, i) = p
      end do
      maxid = maxval(ids)
      deallocate(remap)
      allocate(remap(maxid))
      remap = 0
      do i = 1, n
        remap(ids(i)) = i
      end do
    end block
    do
      read(u, '(a)') line
      if (trim(line) == "$Elements") exit
    end do
    read(u, *) n
    allocate(tcon(3, n), tb(n))
    nt = 0
    do i = 1, n
      read(u, '(a)') line
      read(line, *, iostat=ios) id, etype, ntag
      if (etype /= 2) cycle
      read(line, *) vals(1:3 + ntag + 3)
      nt = nt + 1
      tb(nt) = max(1, vals(4))
      tcon(:, nt) = remap(vals(3 + ntag + 1:3 + ntag + 3))
    end do
    close(u)
    con = tcon(:, 1:nt)
    bcid = tb(1:nt)
  end subroutine

end program resistance_host
