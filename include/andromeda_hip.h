/*
 * andromeda_hip.h -- C-ABI of the MI355X-native BEM hot path (assembly, solve, resistance).
 *
 * Plain C types only: pointers, sizes, 32/64-bit integers and doubles. Every entry point returns
 * an int status (0 ok; > 0 solver status compatible with the LSQR istop code; < 0 argument,
 * allocation or HIP error, message in andr_last_error()). Nothing aborts the host process.
 *
 * Host arrays stay owned by the caller and are never freed by the library. Device memory is owned
 * by the context and released by andr_destroy(). One context = one GPU and one HIP stream; a
 * context must not be used from two host threads at once.
 *
 * Each entry point names the reference routine (WildSmilodon/Andromeda v1.8) it replaces.
 */
#ifndef ANDROMEDA_HIP_H
#define ANDROMEDA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANDR_OK 0
#define ANDR_ERR_ARG (-1)
#define ANDR_ERR_HIP (-2)
#define ANDR_ERR_ALLOC (-3)
#define ANDR_ERR_STATE (-4)

/* storage order of matrices handed to / returned by the host */
#define ANDR_ROW_MAJOR 0
#define ANDR_COL_MAJOR 1 /* Fortran Hmat(nsp, nnodes) order, src/lib_int.f90:254-255 */

typedef struct andr_ctx andr_ctx;

/* Surface mesh of one body (flat 3-node triangles, or bilinear 4-node quadrilaterals mixed with
 * triangles), the data of module mMesh (src/mod_mesh.f90:16-111) passed explicitly. */
typedef struct {
  int32_t nnodes, nelem, nwalls;
  const double *xyz;     /* 3*nnodes: x,y,z of node 0, node 1, ...            */
  const int32_t *con;    /* nvert*nelem: vertex ids of every element            */
  const int32_t *bcid;   /* nelem: wall id 1..nwalls (NULL: every element wall 1) */
  const double *normmul; /* nwalls: NORMMUL factor per wall (NULL: all +1)    */
  int32_t index_base;    /* 0 for C ids, 1 for Fortran ids in con              */
  int32_t itri;          /* regular triangle rule, ITRI 1..5 (7..126 points); 0 -> 5 */
  int32_t nvert;         /* entries of con per element: 0 or 3 (triangles), or 4: 4-node
                            quadrilaterals, where a fourth entry below index_base marks a
                            triangle (src/mod_mesh.f90: element%type 2 / 3). Quadrilaterals are
                            accepted by the assembly entry points (seam 1); the resistance and
                            training entry points take triangle meshes */
  int32_t iquad;         /* Gauss-Legendre rule of quadrilaterals, IQUD 1..8 (N = 2,4,6,8,16,32,
                            40,48 per direction); 0 -> 4 (src/mod_par.f90:233-234) */
} andr_mesh;

/* Host-side checks of a mesh, the ones every entry point taking a mesh makes before it reads the
 * arrays: positive sizes, non-null xyz and con, nvert 0/3/4, every connectivity entry in range
 * (after index_base), every bcid in 1..nwalls, itri 0..5, iquad 0..8. Returns ANDR_OK or
 * ANDR_ERR_ARG with the reason in msg (msglen bytes, may be NULL). No context, no device (the
 * reference checks its input in its mesh readers, src/lib_meshRead.f90). */
int andr_mesh_validate(const andr_mesh *mesh, char *msg, int32_t msglen);

/* ---- context -------------------------------------------------------------------------- */
int andr_create(andr_ctx **ctx, int32_t device);
int andr_destroy(andr_ctx *ctx);
const char *andr_last_error(const andr_ctx *ctx);
/* Library version string; never NULL. */
const char *andr_version(void);

/* ---- seam 1: integral matrices ----------------------------------------------------------
 * Replaces sdFormIntegralMatrices (src/lib_int.f90:225-323; rows by sdIntRowU/sdIntRowQ,
 * src/lib_sdBem.f90:438-733). H is nsp x nnodes and G is nsp x nelem, nsp = nnodes + nelem,
 * node rows first, free term included. err[2] = {fspCerr, QspCerr} as logged by CoR_Integrals. */
int andr_assemble_laplace(andr_ctx *ctx, const andr_mesh *mesh, int32_t layout, double *H, double *G,
                          double err[2]);

/* Replaces sdFormIntegralMatricesStokes (src/lib_Stokes.f90:1391-1555; rows by sdIntRowUStokes
 * :1904-2245 and sdIntRowQStokes :1562-1897). T[6] = {Txx,Txy,Txz,Tyy,Tyz,Tzz} are nsp x nnodes,
 * G[6] the same components nsp x nelem, prT[3]/prG[3] the pressure matrices (pointers may be
 * NULL to skip). extra_free_term: 0 reproduces the reference; 8*pi gives the exterior form. */
int andr_assemble_stokes(andr_ctx *ctx, const andr_mesh *mesh, int32_t layout, double extra_free_term,
                         double *T[6], double *G[6], double *prT[3], double *prG[3], double err[2]);

/* Rows row_begin .. row_end - 1 (0-based, of the nsp = nnodes + nelem source points, node rows first)
 * of the same matrices: H and G hold (row_end - row_begin) rows (T, G, prT, prG likewise), in the
 * requested layout; err[2] covers those rows only. Replaces the per-row integration that the
 * reference's distributed path calls for the rows a process owns (sdIntRowU/Q, sdIntRowU/QStokes
 * from src/lib_StokesLis.f90:695,805). A null matrix pointer skips that matrix. */
int andr_assemble_laplace_rows(andr_ctx *ctx, const andr_mesh *mesh, int32_t row_begin, int32_t row_end,
                               int32_t layout, double *H, double *G, double err[2]);
int andr_assemble_stokes_rows(andr_ctx *ctx, const andr_mesh *mesh, int32_t row_begin, int32_t row_end,
                              int32_t layout, double extra_free_term, double *T[6], double *G[6], double *prT[3],
                              double *prG[3], double err[2]);

/* ---- seam 2: least squares --------------------------------------------------------------
 * Link-compatible with the Fortran routines of src/lib_lsqr.f:1-167 (gfortran/flang name mangling,
 * all arguments by reference, CSR with 1-based INTEGER(8) row pointers and INTEGER columns).
 * lsqr_precon_: q = 1/||column||_2, a <- a diag(q) in place (src/lib_lsqr.f:131-167).
 * lsqr_solve_: hot start from x, maxit == 0 -> 10(m+n+50) written back, ierr = istop if > 3
 * (src/lib_lsqr.f:33-127). Runs on the device of the process default context (ANDR_DEVICE env,
 * default 0). Duplicate (i, j) entries of the CSR add up, as in the reference's products, and q
 * sums their squares apart, as the reference's column norms do (computed on the host in row
 * order). A CSR with entries in under a third of the m x n positions stays sparse on the device
 * (A and its transpose in CSR, row-parallel products); a denser one (the reference's own systems
 * are dense matrices stored in CSR) becomes a dense copy for the matrix-core products. The
 * device copy of the matrix is reused by later lsqr_solve_ calls while the array pointers, the
 * sizes and a checksum of every byte of ia, ja and a are unchanged (else the matrix is uploaded
 * again); andr_lsqr_forget() drops it. The Fortran interface has no status argument for
 * lsqr_precon_: when it fails, q is set to NaN, a is left unscaled and every lsqr_solve_ returns
 * ierr = -1 until an lsqr_precon_ succeeds. */
void lsqr_precon_(const int32_t *m, const int32_t *n, const int64_t *nnz, double *q, const int64_t *ia,
                  const int32_t *ja, double *a);
void lsqr_solve_(const int32_t *m, const int32_t *n, const int64_t *nnz, int32_t *maxit, const double *eps,
                 int32_t *nit, int32_t *ierr, const double *q, const int64_t *ia, const int32_t *ja,
                 const double *a, const double *b, double *x);
void andr_lsqr_forget(void);

typedef struct {
  int32_t maxit;  /* 0 -> 10(m+n+50) */
  double eps;     /* atol = eps, btol = eps ||b|| / ||r0|| (src/lib_lsqr.f:84-100) */
  int32_t precon; /* 1: column scaling as lsqr_precon, 0: none */
} andr_lsqr_opts;
typedef struct {
  int32_t nit[16];
  int32_t istop[16];
  double seconds;
} andr_lsqr_stats;
/* Multi-RHS LSQR on a dense row-major m x n host matrix (no CSR), nrhs <= 16 right-hand sides
 * (B and X are m x nrhs and n x nrhs, column k contiguous). X is the hot start. Each column
 * follows its own LSQR recurrence; the products A V and A^T U run for all columns at once. */
int andr_lsqr_dense(andr_ctx *ctx, int64_t m, int64_t n, const double *A, int32_t nrhs, const double *B,
                    double *X, const andr_lsqr_opts *opts, andr_lsqr_stats *stats);

/* The Krylov alternative of the reference's LIS path for square systems (parLISslvSet
 * "-i bicgstab -p jacobi -tol 1.0e-8 -maxiter 1000", src/mod_par.f90:288; LIS itself is an external
 * library): Jacobi-preconditioned BiCGSTAB on a dense row-major n x n host matrix, nrhs <= 16
 * right-hand sides (B and X n x nrhs, column k contiguous, X the initial guess and the result).
 * tol <= 0: 1e-8 (||r||_2 <= tol ||b||_2); maxit <= 0: 1000. stats->istop per column: 0 converged,
 * 1 iteration limit, 2 breakdown; stats->nit the iterations. */
int andr_bicgstab_dense(andr_ctx *ctx, int64_t n, const double *A, int32_t nrhs, const double *B, double *X,
                        double tol, int32_t maxit, andr_lsqr_stats *stats);

/* andr_bicgstab_dense without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). A (n x n, row-major, leading dimension lda >= n),
 * B and X (n x nrhs, column k contiguous; X the initial guess and the result) are DEVICE pointers
 * owned by the caller; every work array lives in one fixed 64 MiB device array of the library
 * (one per device, held until the process ends), so n is limited to what fits it (about 27 000:
 * ANDR_ERR_ARG above). No hipMalloc, hipFree or memset; the call returns with X final. Calls that
 * overlap on one device share that array and race: serialize them. Results are those of
 * andr_bicgstab_dense on the same inputs (same kernels, same summation order). */
int andr_bicgstab_dense_noalloc(andr_ctx *ctx, int64_t n, const double *A, int64_t lda, int32_t nrhs, const double *B,
                                double *X, double tol, int32_t maxit, andr_lsqr_stats *stats);

/* ---- the Laplace boundary-value problem ---------------------------------------------------
 * A subdomain is a list of walls with their NORMMUL factors and a diffusivity (the SUBDOMAINS and
 * NORMMUL lines of the reference's BIC file, src/lib_sdBem.f90:319-430, src/lib_mesh.f90:1148-1160).
 * Its elements are those of its walls (wall by wall, in the order given), its nodes those of its
 * elements; a node takes its boundary condition from the wall of the last of its elements in that
 * order, unless node_wall (nnodes entries, may be NULL) gives a nonzero wall id for it (the EDGE and
 * CORNER rules of the BIC file, src/lib_sdBem.f90:112-300). */
typedef struct {
  int32_t nwalls;
  const int32_t *walls;     /* wall ids 1..mesh.nwalls */
  const double *normmul;    /* per listed wall; NULL: all +1 */
  double diff;              /* > 0 */
  const int32_t *node_wall; /* nnodes or NULL */
} andr_subdomain;

#define ANDR_BC_FUNCTION 1 /* u known (src/mod_eqns.f90 iFunction) */
#define ANDR_BC_FLUX 2     /* q known (iFlux) */
#define ANDR_BC_CONTACT 3  /* u and q unknown: a wall shared by two subdomains (iContact) */

typedef struct {
  int32_t neq, nx, nb;  /* equations (system rows), unknowns, knowns (sdSetUpXandB) */
  int32_t nit, istop;   /* LSQR iterations and stopping code (src/lib_lsqr.f:1125-1134) */
  double err[2];        /* largest fSP and qSP self-check errors of the subdomains' integrals */
  double t_assemble;    /* seconds: integral matrices of all subdomains (shared by the equations) */
  double t_solve;       /* seconds: this equation's system, solve and scatter */
} andr_laplace_report;

/* Replaces the reference's Laplace driver from the integrals on (solveLaplace, src/laplace.f90):
 * the integral matrices of every subdomain (sdFormIntegralMatrices, src/lib_int.f90:225-323), the
 * unknowns and knowns of every equation (sdSetUpXandB, src/lib_sdBem.f90:4-107), the system and
 * right-hand side (sdSetUpSysMrhsM_CRS and CRSxV, src/lib_CRS.f90:148-343,764-787), the column-scaled
 * LSQR solve with the hot start and the scatter back to u and q (sdSolveCRS, src/lib_CRS.f90:4-142),
 * and the flux integral of every wall (CalFluxIntegralWalls, src/lib_mesh.f90:851-882), all on the
 * device. neqn equations share the integrals: wall_bc[en * nwalls + w - 1] is the kind (ANDR_BC_*)
 * of wall w for equation en; u[en * nnodes + i] and q[en * nelem + e] hold the known values and the
 * initial values of the unknowns on entry and the solution on return; opts[en] (may be NULL) the LSQR
 * limits (maxit 0: 10(m+n+50); eps 0: 1e-15); wall_flux[en * nwalls + w - 1] (may be NULL) gets
 * sum_e A_e q_e over the elements of wall w; reports[en] (may be NULL). */
int andr_laplace_solve(andr_ctx *ctx, const andr_mesh *mesh, int32_t nsd, const andr_subdomain *sds, int32_t neqn,
                       const int32_t *wall_bc, const andr_lsqr_opts *opts, double *u, double *q, double *wall_flux,
                       andr_laplace_report *reports);
/* The mixed-precision form of andr_laplace_solve (no counterpart in the reference; same arguments and
 * results): the regular integrals in FP32 (singular ones, the free terms and the right-hand side in
 * FP64), the system stored in FP32, the LSQR products accumulated in FP64 and every LSQR vector and
 * the solution update in FP64. On the C1 sphere (1280 triangles, u = x) the fluxes stay within
 * 1e-6 (relative L2) of andr_laplace_solve's. */
int andr_laplace_solve_mixed(andr_ctx *ctx, const andr_mesh *mesh, int32_t nsd, const andr_subdomain *sds,
                             int32_t neqn, const int32_t *wall_bc, const andr_lsqr_opts *opts, double *u, double *q,
                             double *wall_flux, andr_laplace_report *reports);

/* The general Stokes problem of the reference's Stokes driver on one subdomain (its Stokes path has
 * one, src/lib_StokesFull.f90:183): the 12 integral matrices (sdFormIntegralMatricesStokes, bounded
 * form), the unknowns and knowns of the three velocity components (sdSetUpXandB per component,
 * stacked as StokesBigSystemXB, src/lib_Stokes.f90:1180-1276), the system rows of the unknown values
 * (stokesFormCRSsysMrhsM with formUrow / formQrow: -T and -(normmul/diff) G for the unknowns, +T and
 * +(normmul/diff) G times the knowns into b = rhsM b, src/lib_Stokes.f90:475-637), the column-scaled
 * LSQR solve with the hot start (StokesBigSystemSOLVEcrs, :417-469, with the solver limits of the first
 * equation, :626-629) and the scatter back (StokesDistributeUnknowns, src/lib_StokesFull.f90:379-501).
 * wall_bc[k * nwalls + w - 1] is the kind of wall w for velocity component k (x, y, z); u[k * nnodes + i]
 * and q[k * nelem + e] are the known and initial values on entry and the solution on return;
 * wall_force[k * nwalls + w - 1] (may be NULL) gets sum_e A_e q_k,e over wall w. Triangle meshes. */
int andr_stokes_solve(andr_ctx *ctx, const andr_mesh *mesh, const andr_subdomain *sd, const int32_t *wall_bc,
                      const andr_lsqr_opts *opts, double *u, double *q, double *wall_force,
                      andr_laplace_report *report);

/* The wall each node of each subdomain takes its condition from, as andr_laplace_solve uses it:
 * node_wall[s * nnodes + i] (0: node i is not in subdomain s). Host only: no context, no device.
 * ANDR_ERR_ARG with the reason in msg for an invalid mesh or subdomain. */
int andr_subdomain_node_walls(const andr_mesh *mesh, int32_t nsd, const andr_subdomain *sds, int32_t *node_wall,
                              char *msg, int32_t msglen);

/* ---- seam 3: resistance ------------------------------------------------------------------- */
#define ANDR_FORM_EXTERIOR 0 /* single surface, rigid-body velocity on the body                */
#define ANDR_FORM_BOUNDED 1  /* body (wall 1) at rest inside an outer wall 2 carrying the flow */

#define ANDR_PREC_FP64 0    /* plain path: FP64 assembly, FP64 storage, LSQR (drop-in)         */
#define ANDR_PREC_MIXED 1   /* FP32 regular quadrature, FP64 singular integrals, FP32 storage,
                               GMRES with FP64 vectors and FP64 restart residuals            */
#define ANDR_PREC_LEARNED 2 /* MIXED plus learned Gauss-order selection for near-field pairs */

typedef struct {
  int32_t form;      /* ANDR_FORM_* */
  int32_t precision; /* ANDR_PREC_* */
  double mu;         /* viscosity (subdomain diff), default 1 */
  double eps;        /* solver tolerance; 0 -> 1e-12 (LSQR eps, plain) or 1e-6 (GMRES relative
                        residual, mixed and learned) */
  int32_t maxit;     /* 0 -> default */
  double center[3];  /* torque and rotation reference point */
  int32_t restart;   /* fast forms: steps of the first GMRES cycle; 0 -> 128, longer (up to 192) while
                        the bodies of the call fit one wave of device memory */
  int32_t problems;  /* ANDR_PROBLEMS_RIGID (0): the six rigid-body problems; ANDR_PROBLEMS_STRAIN (1):
                        the three strain fields of StokesInlet3sidesTorquePiCRS (src/lib_particle.f90:
                        6-168, velocities :72-85): (0, z, y), (z, 0, x), (y, x, 0) about `center`, in
                        FT rows 0..2 (rows 3..5 zero) */
} andr_resistance_opts;
#define ANDR_PROBLEMS_RIGID 0
#define ANDR_PROBLEMS_STRAIN 1

typedef struct {
  int32_t nit[6];       /* iterations per rigid-body problem (LSQR iterations; GMRES: block rounds,
                           each reading the body's matrix once for all six problems) */
  int32_t status;       /* max solver status over the six problems: 0 converged; LSQR (plain path):
                           its istop when > 3 (4 ill-conditioned, 5 iteration limit); GMRES (mixed
                           and learned): 4 stagnated at a relative residual <= 1e-6 above eps (the
                           FP32 floor of the near-null direction, accepted), 5 iteration limit */
  double err[2];        /* integral self-checks (fspCerr-like, QspCerr-like) */
  double t_assemble;    /* seconds */
  double t_solve;       /* seconds */
  int64_t quad_points;  /* regular quadrature points evaluated */
  int64_t pairs_learned;  /* near-field pairs whose rule came from the network */
  int64_t pairs_fallback; /* near-field pairs sent to the fixed-order rule */
} andr_resistance_report;

/* Replaces StokesInlet3sidesForceCRS / TorqueOmegaCRS (src/lib_particle.f90:176-476): six
 * rigid-body problems (translations e_x, e_y, e_z, rotations e_k x (r - center)); FT[6*k + i] is
 * component i of (force, torque) for problem k, force = sum_e A_e q_e and torque =
 * sum_e A_e (c_e - center) x q_e over wall 1 (src/lib_particle.f90:1751-1831). With
 * opts->problems = ANDR_PROBLEMS_STRAIN it replaces StokesInlet3sidesTorquePiCRS
 * (src/lib_particle.f90:6-168): the three strain problems, whose torques are the Pi tensor. The
 * signs are those of the reference's bounded set-up (the body at rest in the flow): the drag of a
 * translation and the torque of a rotation are positive, and a strain field turns a body's long
 * axis toward its stretching direction (Jeffery's torque). */
int andr_resistance(andr_ctx *ctx, const andr_mesh *mesh, const andr_resistance_opts *opts, double FT[36],
                    andr_resistance_report *report);

/* Many bodies in one call: assembly of all bodies in one launch per wave of bodies that fits the
 * device memory, then all bodies solved together. FT is 36*nbodies; reports may be NULL. */
int andr_resistance_batch(andr_ctx *ctx, int32_t nbodies, const andr_mesh *meshes, const andr_resistance_opts *opts,
                          double *FT, andr_resistance_report *reports);

/* Replaces GetStokesDomainValue (src/lib_Stokes.f90:3043-3390) as calculateDomainValues
 * (src/lib_domain.f90:463-527) calls it for every domain point: velocity and pressure at npts points
 * pts[3*npts] of the flow domain from the boundary solution, u[3*nnodes] the nodal velocities (x, y,
 * z per node) and q[3*nelem] the element tractions, viscosity diff:
 * vel = -(sum T u + sum G q / diff) / 8 pi, p = -(diff sum prT.u - sum prG.q) / 8 pi, with every
 * element integrated by the mesh's rule (ITRI). err[npts] (may be NULL) is the reference's closure
 * error of the point's integrals; above 1e-2 the point is not inside the domain (getStokesDV,
 * src/lib_Stokes.f90:3020-3037). Triangle meshes. */
int andr_stokes_domain_values(andr_ctx *ctx, const andr_mesh *mesh, int32_t npts, const double *pts, const double *u,
                              const double *q, double diff, double *vel, double *p, double *err);

/* Replaces pressureStokes (src/lib_Stokes.f90:4-94): the surface pressure at the element centres,
 * p[nelem], from the centroid rows of the pressure integrals (prT with its row-sum correction, prG)
 * and the boundary solution u[3*nnodes], q[3*nelem], viscosity diff:
 * p = -2 (diff prT.(u - u_c) - prG.q) / 8 pi, u_c the mean velocity of the element's nodes.
 * Triangle meshes. */
int andr_stokes_surface_pressure(andr_ctx *ctx, const andr_mesh *mesh, const double *u, const double *q, double diff,
                                 double *p);

/* ---- integrals file (and.integrals.bin) -----------------------------------------------------
 * Replace the reference's writers and readers of its integrals file for one subdomain, so that
 * matrices assembled here are consumed by the unmodified host and the other way round:
 * WriteIntegralsToDiskStokes / ReadIntegralsFileStokes (src/lib_Stokes.f90:2742-2872) and
 * WriteIntegralsToDisk / ReadIntegralsFile (src/lib_int.f90:136-221). Fortran sequential unformatted
 * records with 4-byte length markers: (Stokes only) the id 1407; (nsp, nnodes, nqnodes); the two
 * self-check errors; then every matrix column by column, each column as records of 8192 values and a
 * last record with the rest (WrSMat / wrvec, src/lib_util.f90:201-287). Stokes matrices are passed
 * as andr_assemble_stokes returns them in ANDR_ROW_MAJOR layout: T and G in the order xx, xy, xz, yy,
 * yz, zz (the file holds xx, yy, zz, xy, xz, yz), prT and prG x, y, z. A reader checks the id and the
 * sizes as the reference does and returns ANDR_ERR_ARG on a mismatch or a short file; a NULL matrix
 * pointer skips that matrix. Host code: no context, no device. */
int andr_integrals_write_stokes(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes, const double err[2],
                                double *const T[6], double *const G[6], double *const prT[3], double *const prG[3]);
int andr_integrals_read_stokes(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes, double err[2],
                               double *T[6], double *G[6], double *prT[3], double *prG[3]);
int andr_integrals_write_laplace(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes, const double err[2],
                                 const double *H, const double *G);
int andr_integrals_read_laplace(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes, double err[2],
                                double *H, double *G);

/* Device-resident bodies (no reference counterpart: the reference reads a mesh and solves it once
 * per run). For callers that compute the tensors of the same bodies more than once, e.g. at several
 * precisions: andr_bodies_upload copies the geometry to the device and runs the closure tests of
 * the walls for opts->form and opts->mu (the other option fields are not used); the host meshes are
 * not referenced afterwards. andr_resistance_batch_resident is andr_resistance_batch on those bodies
 * (same form and mu, any precision), FT and reports in the order of the uploaded meshes.
 * andr_bodies_count returns the number of bodies; andr_bodies_free releases them and must come
 * before andr_destroy of the context. */
typedef struct andr_bodies andr_bodies;
int andr_bodies_upload(andr_ctx *ctx, int32_t nbodies, const andr_mesh *meshes, const andr_resistance_opts *opts,
                       andr_bodies **bodies);
int andr_bodies_count(const andr_bodies *bodies);
int andr_resistance_batch_resident(andr_ctx *ctx, const andr_bodies *bodies, const andr_resistance_opts *opts,
                                   double *FT, andr_resistance_report *reports);
int andr_bodies_free(andr_ctx *ctx, andr_bodies *bodies);

/* ---- learned quadrature (no reference counterpart: the reference uses one fixed rule) ----- */
/* Network: 6 inputs -> 3 hidden layers of 64 (ReLU) -> 8 classes; classes 0..6 are the Dunavant
 * rules with 3, 4, 6, 7, 12, 13, 19 points, class 7 the fixed-order rule of the mesh (ITRI).
 * Parameters are FP32 masters (ANDR_MLP_NPARAM floats: W1 6x64, b1, W2 64x64, b2, W3 64x64,
 * b3, W4 64x8, b4, weights stored input-major); the assembly kernel uses a BF16 copy. */
#define ANDR_MLP_NPARAM 9288
/* Features of near-field (source point, element) pairs of a body and the label of each pair:
 * the smallest rule whose G and T integrals agree with the fixed-order rule to rel_tol.
 * Pairs with distance/size < near_ratio give three records each (Laplace, Stokes-G, Stokes-H);
 * when there are more than max_pairs records, the pairs are sampled uniformly (deterministic hash
 * of the pair) to about 0.97 max_pairs. The number of records written is returned in *npairs. */
int andr_quad_training_pairs(andr_ctx *ctx, const andr_mesh *mesh, double near_ratio, double rel_tol,
                             int64_t max_pairs, float *features /* 6*max_pairs */, int32_t *labels,
                             int64_t *npairs);
/* Trains the network on the device (forward, backward and Adam kernels) for `epochs` passes of
 * minibatches; params in/out (ANDR_MLP_NPARAM floats). loss_out[epochs] gets the mean loss. */
int andr_mlp_train(andr_ctx *ctx, int64_t npairs, const float *features, const int32_t *labels, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, float *params, float *loss_out);
/* Forward pass on the device: probabilities 8*npairs. */
int andr_mlp_infer(andr_ctx *ctx, int64_t npairs, const float *features, const float *params, float *probs);
/* One forward+backward pass on a batch: loss and the gradient of the mean cross-entropy. */
int andr_mlp_loss_grad(andr_ctx *ctx, int64_t npairs, const float *features, const int32_t *labels,
                       const float *params, float *loss, float *grad);
/* Installs the rule selection used by ANDR_PREC_LEARNED. Pairs with distance/size below near_ratio
 * go through the network; each of its two kernels (Stokeslet, stresslet) takes the smallest class
 * whose cumulative softmax probability reaches `confidence` (a quantile of the predicted label),
 * the pair the larger of the two; class 7 is the fixed-order rule. Pairs farther out take far_table[b], the class
 * calibrated on training pairs for bin b = floor(2 log2(d / (h near_ratio))), b = 0..15
 * (beyond the last bin: far_table[15]). */
#define ANDR_FAR_BINS 16
int andr_set_quadrature_model(andr_ctx *ctx, const float *params, float confidence, double near_ratio,
                              const int32_t *far_table);
/* Far-field calibration: for pairs with distance/size in [near_ratio, near_ratio * 256) the
 * largest label per far bin (as in far_table), over the centroid rows of a mesh; -1 for a bin
 * that no pair of the mesh falls in. */
int andr_quad_far_calibration(andr_ctx *ctx, const andr_mesh *mesh, double near_ratio, double rel_tol,
                              int32_t *far_table);

/* ---- measurement (no reference counterpart: the reference only logs CPU timers) ------------
 * Per-kernel timing with HIP events recorded on the context stream around each launch of the hot
 * kernels, with the algorithmic bytes and flops of each launch. andr_profile(ctx, 1) resets and
 * enables, 0 disables; andr_profile_read returns 1 and fills entry idx (0..), 0 past the end. */
int andr_profile(andr_ctx *ctx, int32_t enable);
int andr_profile_read(andr_ctx *ctx, int32_t idx, char *name, int32_t namelen, int64_t *launches, double *ms,
                      double *bytes, double *flops);

/* ---- autotuner (no reference counterpart) ---------------------------------------------------
 * The plain solve's FP64 product A V picks its rows per wave (32, 64 or 128) by timing the
 * candidates on the first large system (>= 256 MB) of every (ceil log2 rows, ceil log2 columns)
 * bucket; the choice is kept per context and never changes a result bit (the column segments and
 * every summation order are fixed). andr_autotune(ctx, mode): mode 1 tunes (the default), 0 uses 64
 * rows per wave, 32 / 64 / 128 fix that many; andr_autotune_read returns 1 and fills entry idx (0..)
 * of the table, 0 past the end. */
int andr_autotune(andr_ctx *ctx, int32_t mode);
int andr_autotune_read(andr_ctx *ctx, int32_t idx, int32_t *log2_rows, int32_t *log2_cols, int32_t *rows_per_wave,
                       double *ms);
/* The fast forms' Krylov product splits the columns of the live systems into segments when few
 * rows are live (the slowest bodies of a batch), aiming at a number of waves (1024, 2048, 4096 or
 * 8192). With the autotuner on (mode 1), the first time a bucket of live rows (ceil log2 of the live
 * 32-row waves, below 4096 waves) occurs the candidates are timed on the live systems and the
 * fastest is kept in the context; otherwise 2048. andr_autotune_read_krylov returns 1 and fills
 * entry idx (0..) of that table, 0 past the end. */
int andr_autotune_read_krylov(andr_ctx *ctx, int32_t idx, int32_t *log2_row_waves, int32_t *target_waves,
                              double *ms);

#ifdef __cplusplus
}
#endif
#endif
