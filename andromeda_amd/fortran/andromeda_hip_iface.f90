! andromeda_hip: ISO_C_BINDING interface of the MI355X BEM library (include/andromeda_hip.h).
!
! A Fortran host (the reference's and.inp/.msh/.bic drivers) uses this module to hand its mesh
! arrays (module mMesh, src/mod_mesh.f90:16-111) to the device assembly, solve and resistance
! entry points. The LSQR seam needs no interface: lsqr_precon_ and lsqr_solve_ are exported with
! the Fortran name mangling and the argument list of src/lib_lsqr.f:1-167, so the reference's
! implicit-interface calls (e.g. src/lib_CRS.f90:77-85) link against the library unchanged.
module andromeda_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ANDR_OK = 0
  integer(c_int), parameter :: ANDR_ROW_MAJOR = 0, ANDR_COL_MAJOR = 1
  integer(c_int), parameter :: ANDR_FORM_EXTERIOR = 0, ANDR_FORM_BOUNDED = 1
  integer(c_int), parameter :: ANDR_PREC_FP64 = 0, ANDR_PREC_MIXED = 1, ANDR_PREC_LEARNED = 2
  integer(c_int), parameter :: ANDR_PROBLEMS_RIGID = 0, ANDR_PROBLEMS_STRAIN = 1
  integer(c_int), parameter :: ANDR_MLP_NPARAM = 9288, ANDR_FAR_BINS = 16

  ! andr_mesh: pointers to the host's arrays (xyz(3,nnodes), con(3,nelem), bcid(nelem),
  ! normmul(nwalls)); index_base = 1 for Fortran node ids
  type, bind(C) :: andr_mesh
    integer(c_int32_t) :: nnodes, nelem, nwalls
    type(c_ptr) :: xyz, con, bcid, normmul
    integer(c_int32_t) :: index_base, itri
    integer(c_int32_t) :: nvert  ! entries of con per element: 3 (triangles) or 4 (quadrilaterals)
    integer(c_int32_t) :: iquad  ! Gauss-Legendre rule of quadrilaterals, IQUD 1..8 (0 -> 4)
  end type

  type, bind(C) :: andr_lsqr_opts
    integer(c_int32_t) :: maxit
    real(c_double) :: eps
    integer(c_int32_t) :: precon
  end type

  type, bind(C) :: andr_lsqr_stats
    integer(c_int32_t) :: nit(16), istop(16)
    real(c_double) :: seconds
  end type

  integer(c_int), parameter :: ANDR_BC_FUNCTION = 1, ANDR_BC_FLUX = 2, ANDR_BC_CONTACT = 3

  ! andr_subdomain: a subdomain's walls (loWalls), their NORMMUL factors, the diffusivity, and
  ! optional per-node wall ids of the EDGE / CORNER rules (nnodes entries, 0 = the default rule)
  type, bind(C) :: andr_subdomain
    integer(c_int32_t) :: nwalls
    type(c_ptr) :: walls, normmul
    real(c_double) :: diff
    type(c_ptr) :: node_wall
  end type

  type, bind(C) :: andr_laplace_report
    integer(c_int32_t) :: neq, nx, nb, nit, istop
    real(c_double) :: err(2), t_assemble, t_solve
  end type

  type, bind(C) :: andr_resistance_opts
    integer(c_int32_t) :: form, precision
    real(c_double) :: mu, eps
    integer(c_int32_t) :: maxit
    real(c_double) :: center(3)
    integer(c_int32_t) :: restart  ! fast forms: first GMRES cycle; 0 -> automatic
    integer(c_int32_t) :: problems ! 0: six rigid-body problems; 1: three strain fields (TorquePiCRS)
  end type

  type, bind(C) :: andr_resistance_report
    integer(c_int32_t) :: nit(6), status
    real(c_double) :: err(2), t_assemble, t_solve
    integer(c_int64_t) :: quad_points, pairs_learned, pairs_fallback
  end type

  interface
    integer(c_int) function andr_create(ctx, device) bind(C, name="andr_create")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(out) :: ctx
      integer(c_int32_t), value :: device
    end function

    integer(c_int) function andr_destroy(ctx) bind(C, name="andr_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function andr_last_error(ctx) bind(C, name="andr_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    ! sdFormIntegralMatrices (src/lib_int.f90:225-323): H(nsp, nnodes), G(nsp, nelem)
    integer(c_int) function andr_assemble_laplace(ctx, mesh, layout, H, G, err) &
        bind(C, name="andr_assemble_laplace")
      import :: c_int, c_ptr, c_int32_t, c_double, andr_mesh
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      integer(c_int32_t), value :: layout
      real(c_double), intent(out) :: H(*), G(*)
      real(c_double), intent(out) :: err(2)
    end function

    ! sdFormIntegralMatricesStokes (src/lib_Stokes.f90:1391-1555): T(6) and G(6) hold c_loc of
    ! the component matrices Txx..Tzz, Gxx..Gzz; prT/prG may be c_null_ptr
    integer(c_int) function andr_assemble_stokes(ctx, mesh, layout, extra_free_term, T, G, prT, prG, err) &
        bind(C, name="andr_assemble_stokes")
      import :: c_int, c_ptr, c_int32_t, c_double, andr_mesh
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      integer(c_int32_t), value :: layout
      real(c_double), value :: extra_free_term
      type(c_ptr), intent(in) :: T(6), G(6)
      type(c_ptr), value :: prT, prG
      real(c_double), intent(out) :: err(2)
    end function

    integer(c_int) function andr_lsqr_dense(ctx, m, n, A, nrhs, B, X, opts, stats) &
        bind(C, name="andr_lsqr_dense")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double, andr_lsqr_opts, andr_lsqr_stats
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: m, n
      real(c_double), intent(in) :: A(*)
      integer(c_int32_t), value :: nrhs
      real(c_double), intent(in) :: B(*)
      real(c_double), intent(inout) :: X(*)
      type(andr_lsqr_opts), intent(in) :: opts
      type(andr_lsqr_stats), intent(out) :: stats
    end function

    ! StokesInlet3sidesForceCRS / TorqueOmegaCRS (src/lib_particle.f90:176-476)
    integer(c_int) function andr_resistance(ctx, mesh, opts, FT, report) bind(C, name="andr_resistance")
      import :: c_int, c_ptr, c_double, andr_mesh, andr_resistance_opts, andr_resistance_report
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      type(andr_resistance_opts), intent(in) :: opts
      real(c_double), intent(out) :: FT(36)
      type(andr_resistance_report), intent(out) :: report
    end function

    integer(c_int) function andr_resistance_batch(ctx, nbodies, meshes, opts, FT, reports) &
        bind(C, name="andr_resistance_batch")
      import :: c_int, c_ptr, c_int32_t, c_double, andr_mesh, andr_resistance_opts, andr_resistance_report
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nbodies
      type(andr_mesh), intent(in) :: meshes(*)
      type(andr_resistance_opts), intent(in) :: opts
      real(c_double), intent(out) :: FT(*)
      type(andr_resistance_report), intent(out) :: reports(*)
    end function

    ! GetStokesDomainValue (src/lib_Stokes.f90:3043-3390) for npts points: pts(3,npts), u(3,nnodes),
    ! q(3,nelem) -> vel(3,npts), p(npts), err(npts)
    integer(c_int) function andr_stokes_domain_values(ctx, mesh, npts, pts, u, q, diff, vel, p, err) &
        bind(C, name="andr_stokes_domain_values")
      import :: c_int, c_ptr, c_int32_t, c_double, andr_mesh
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      integer(c_int32_t), value :: npts
      real(c_double), intent(in) :: pts(*), u(*), q(*)
      real(c_double), value :: diff
      real(c_double), intent(out) :: vel(*), p(*), err(*)
    end function

    ! pressureStokes (src/lib_Stokes.f90:4-94): p(nelem) at the element centres
    integer(c_int) function andr_stokes_surface_pressure(ctx, mesh, u, q, diff, p) &
        bind(C, name="andr_stokes_surface_pressure")
      import :: c_int, c_ptr, c_double, andr_mesh
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      real(c_double), intent(in) :: u(*), q(*)
      real(c_double), value :: diff
      real(c_double), intent(out) :: p(*)
    end function

    ! the LIS path's solve (parLISslvSet, src/mod_par.f90:288) on one GPU: BiCGSTAB + Jacobi
    integer(c_int) function andr_bicgstab_dense(ctx, n, A, nrhs, B, X, tol, maxit, stats) &
        bind(C, name="andr_bicgstab_dense")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double, andr_lsqr_stats
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: A(*)
      integer(c_int32_t), value :: nrhs
      real(c_double), intent(in) :: B(*)
      real(c_double), intent(inout) :: X(*)
      real(c_double), value :: tol
      integer(c_int32_t), value :: maxit
      type(andr_lsqr_stats), intent(out) :: stats
    end function

    ! WriteIntegralsToDisk / ReadIntegralsFile (src/lib_int.f90:136-221): H(nsp,nnodes) and
    ! G(nsp,nqnodes) passed row-major (ANDR_ROW_MAJOR), i.e. the transposes of Hmat / Gmat
    integer(c_int) function andr_integrals_write_laplace(path, nsp, nnodes, nqnodes, err, H, G) &
        bind(C, name="andr_integrals_write_laplace")
      import :: c_int, c_char, c_int32_t, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: nsp, nnodes, nqnodes
      real(c_double), intent(in) :: err(2), H(*), G(*)
    end function

    ! sdIntRowU / sdIntRowQ over a row range (0-based rows row_begin .. row_end - 1 of nsp)
    integer(c_int) function andr_assemble_laplace_rows(ctx, mesh, row_begin, row_end, layout, H, G, err) &
        bind(C, name="andr_assemble_laplace_rows")
      import :: c_int, c_ptr, c_int32_t, c_double, andr_mesh
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      integer(c_int32_t), value :: row_begin, row_end, layout
      real(c_double), intent(out) :: H(*), G(*), err(2)
    end function

    ! solveLaplace from the integrals on (src/laplace.f90): sdFormIntegralMatrices per subdomain,
    ! sdSetUpXandB, sdSetUpSysMrhsM_CRS, sdSolveCRS and CalFluxIntegralWalls for neqn equations;
    ! wall_bc(nwalls, neqn), u(nnodes, neqn), q(nelem, neqn), wall_flux(nwalls, neqn)
    integer(c_int) function andr_laplace_solve(ctx, mesh, nsd, sds, neqn, wall_bc, opts, u, q, wall_flux, &
        reports) bind(C, name="andr_laplace_solve")
      import :: c_int, c_ptr, c_int32_t, c_double, andr_mesh, andr_subdomain, andr_lsqr_opts, andr_laplace_report
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      integer(c_int32_t), value :: nsd
      type(andr_subdomain), intent(in) :: sds(*)
      integer(c_int32_t), value :: neqn
      integer(c_int32_t), intent(in) :: wall_bc(*)
      type(andr_lsqr_opts), intent(in) :: opts(*)
      real(c_double), intent(inout) :: u(*), q(*)
      real(c_double), intent(out) :: wall_flux(*)
      type(andr_laplace_report), intent(out) :: reports(*)
    end function

    ! the mixed-precision form of andr_laplace_solve (FP32 regular integrals and system storage)
    integer(c_int) function andr_laplace_solve_mixed(ctx, mesh, nsd, sds, neqn, wall_bc, opts, u, q, wall_flux, &
        reports) bind(C, name="andr_laplace_solve_mixed")
      import :: c_int, c_ptr, c_int32_t, c_double, andr_mesh, andr_subdomain, andr_lsqr_opts, andr_laplace_report
      type(c_ptr), value :: ctx
      type(andr_mesh), intent(in) :: mesh
      integer(c_int32_t), value :: nsd
      type(andr_subdomain), intent(in) :: sds(*)
      integer(c_int32_t), value :: neqn
      integer(c_int32_t), intent(in) :: wall_bc(*)
      type(andr_lsqr_opts), intent(in) :: opts(*)
      real(c_double), intent(inout) :: u(*), q(*)
      real(c_double), intent(out) :: wall_flux(*)
      type(andr_laplace_report), intent(out) :: reports(*)
    end function

    integer(c_int) function andr_set_quadrature_model(ctx, params, confidence, near_ratio, far_table) &
        bind(C, name="andr_set_quadrature_model")
      import :: c_int, c_ptr, c_float, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: params(*)
      real(c_float), value :: confidence
      real(c_double), value :: near_ratio
      integer(c_int32_t), intent(in) :: far_table(*)
    end function
  end interface

contains

  ! Message of the last failed call on ctx as a Fortran string.
  function andr_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    type(c_ptr) :: cp
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
      end function
    end interface
    cp = andr_last_error(ctx)
    if (.not. c_associated(cp)) then
      msg = ""
      return
    end if
    n = int(c_strlen(cp))
    call c_f_pointer(cp, p, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module andromeda_hip
