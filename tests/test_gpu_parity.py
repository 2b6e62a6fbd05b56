"""HIP path against the oracle on the same inputs (FP64 plain path: relative 1e-10 on every
matrix entry scale; solver outputs after tight convergence), and the mixed path against the plain
path (1e-6 on resistance tensors). Runs only on an MI355X."""
import numpy as np
import pytest

from andromeda_amd import mesh as M
from andromeda_amd import api
from andromeda_amd.metrics import tensor_error

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("sub", [2, 3])
def test_laplace_assembly(ctx, orc, sub):
    m = M.icosphere(sub)
    H, G, err = ctx.assemble_laplace(m)
    Ho, Go, erro = orc.assemble_laplace(m)
    assert _rel(H, Ho) < 1e-10 and _rel(G, Go) < 1e-10
    assert abs(err[0] - erro[0]) < 1e-12 and abs(err[1] - erro[1]) < 1e-9
    Hc, Gc, _ = ctx.assemble_laplace(m, layout=api.COL_MAJOR)
    assert np.array_equal(Hc, H) and np.array_equal(Gc, G)


def test_laplace_reference_mesh(ctx, orc):
    import os
    m = M.read_gmsh(os.path.join(os.path.dirname(__file__), "golden", "sphere-01.msh"))
    H, G, err = ctx.assemble_laplace(m)
    Ho, Go, _ = orc.assemble_laplace(m)
    assert _rel(H, Ho) < 1e-10 and _rel(G, Go) < 1e-10


def test_stokes_assembly(ctx, orc):
    m = M.geodesic_sphere(3)
    T, G, pT, pG, err = ctx.assemble_stokes(m, pressure=True)
    To, Go, pTo, pGo, erro = orc.assemble_stokes(m, pressure=True)
    for k in range(6):
        assert _rel(T[k], To[k]) < 1e-10, k
        assert _rel(G[k], Go[k]) < 1e-10, k
    for k in range(3):
        assert _rel(pT[k], pTo[k]) < 1e-10 and _rel(pG[k], pGo[k]) < 1e-10


def test_stokes_assembly_two_walls_exterior_term(ctx, orc):
    m = M.body_in_sphere(M.geodesic_sphere(2, 0.5), 8.0, 2)
    T, G, _, _, _ = ctx.assemble_stokes(m, extra=8 * np.pi, pressure=False)
    To, Go, _, _, _ = orc.assemble_stokes(m, extra=8 * np.pi)
    for k in range(6):
        assert _rel(T[k], To[k]) < 1e-10 and _rel(G[k], Go[k]) < 1e-10


def test_lsqr_dense_laplace_system(ctx, orc):
    m = M.icosphere(3)
    Ho, Go, _ = orc.assemble_laplace(m)
    sysM, rhsM = orc.laplace_dirichlet_system(m, Ho, Go)
    b = rhsM @ (-m.xyz[:, 0])
    ia, ja, a = orc.csr_from_dense(sysM)
    q = orc.lsqr_precon(ia, ja, a, m.nelem, m.nelem)
    xo, info = orc.lsqr_solve(ia, ja, a, m.nelem, m.nelem, q, b, eps=1e-12)
    X, st = ctx.lsqr_dense(sysM, b, eps=1e-12)
    assert abs(st["nit"][0] - info["nit"]) <= 2
    assert np.linalg.norm(X[:, 0] - xo) / np.linalg.norm(xo) < 1e-9


@pytest.mark.parametrize("m,n", [(301, 77), (1000, 999), (2050, 640), (64, 64)])
def test_lsqr_dense_shapes(ctx, m, n):
    """Least-squares solutions of six right-hand sides against numpy's lstsq. The shapes cover an odd
    leading dimension (scalar matrix loads), one and several column segments of the matrix-core
    product A V, a partial last column block (n % 16 != 0) and row counts that are not a multiple of
    the 64 rows a wave owns."""
    rng = np.random.default_rng(m * 7 + n)
    A = rng.normal(size=(m, n))
    A[:n, :n] += 4.0 * np.sqrt(m) * np.eye(n)
    B = rng.normal(size=(m, 6))
    X, st = ctx.lsqr_dense(A, B, eps=1e-14)
    Xo = np.linalg.lstsq(A, B, rcond=None)[0]
    assert all(s in (1, 2) for s in st["istop"]), st
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10


def test_lsqr_autotune_rows_per_wave_bitwise():
    """The autotuner's choice (rows per wave of the FP64 product A V) never changes a result bit:
    30 iterations on a 6000 x 6000 system (288 MB, 21 column segments) give identical iterates with
    32, 64 and 128 rows per wave and with the tuned choice, which the table then records."""
    c = api.Context(0)  # a fresh autotuner table
    try:
        rng = np.random.default_rng(5)
        n = 6000
        A = rng.normal(size=(n, n)) / np.sqrt(n) + np.eye(n)
        B = rng.normal(size=(n, 6))
        out = {}
        for mode in (32, 64, 128, 1):
            c.autotune(mode)
            out[mode] = c.lsqr_dense(A, B, maxit=30, eps=1e-30)[0]
        for mode in (64, 128, 1):
            assert np.array_equal(out[mode], out[32]), mode
        table = c.autotune_table()
        assert len(table) == 1 and table[0]["rows_per_wave"] in (32, 64, 128)
        assert (table[0]["log2_rows"], table[0]["log2_cols"]) == (13, 13)
        with pytest.raises(RuntimeError):
            c.autotune(3)
    finally:
        c.close()


def test_fortran_lsqr_shim(ctx, orc):
    rng = np.random.default_rng(11)
    n = 300
    A = rng.normal(size=(n, n)) + 10 * np.eye(n)
    b = rng.normal(size=n)
    ia, ja, a = orc.csr_from_dense(A)
    qo = orc.lsqr_precon(ia, ja, a, n, n)
    xo, info = orc.lsqr_solve(ia, ja, a, n, n, qo, b, eps=1e-14)
    # the Fortran host passes 1-based CSR
    ia1, ja1 = ia + 1, ja + 1
    a1 = A.reshape(-1).copy()
    q = api.lsqr_precon(n, n, ia1, ja1, a1)
    assert _rel(q, qo) < 1e-14 and _rel(a1, a) < 1e-14
    x = np.zeros(n)
    r = api.lsqr_solve(n, n, ia1, ja1, a1, q, b, x, maxit=0, eps=1e-14)
    assert r["ierr"] == 0 and r["maxit"] == 10 * (n + n + 50)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("kind", ["sphere", "superellipsoid"])
def test_resistance_plain_vs_oracle(ctx, orc, kind):
    if kind == "sphere":
        m = M.geodesic_sphere(3)
    else:
        m = M.random_superellipsoid(np.random.default_rng(3), 3)[0]
    FTo, nits, ierr = orc.resistance(m, exterior=True, eps=1e-12)
    r = ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12)
    assert r.status == 0
    assert _rel(r.FT, FTo) < 1e-9


def test_resistance_bounded_vs_oracle(ctx, orc):
    m = M.body_in_sphere(M.geodesic_sphere(2, 0.5), 16.0, 2)
    FTo, _, _ = orc.resistance(m, exterior=False, eps=1e-12)
    r = ctx.resistance(m, form=api.FORM_BOUNDED, precision=api.PREC_FP64, eps=1e-12)
    assert _rel(r.FT, FTo) < 1e-9


def test_mixed_vs_plain(ctx):
    m = M.random_superellipsoid(np.random.default_rng(7), 8)[0]
    p = ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12)
    x = ctx.resistance(m, precision=api.PREC_MIXED)
    assert tensor_error(x.FT, p.FT) < 1e-6


def test_krylov_product_precisions(ctx, monkeypatch):
    """At the default relative residual (1e-6) the Krylov products run on the FP32 matrix cores,
    at 1e-8 on the FP64 ones; ANDR_GMRES_FP64_PRODUCTS=1 forces FP64 products. On a superellipsoid
    and a fibre of the batched workload both product precisions give tensors within 1e-6 of the
    plain FP64 path and within 3e-7 of each other, in about the same number of Krylov steps."""
    allb = [m for m, _ in M.batch_bodies(512, 64)]
    bodies = [allb[42], allb[534]]
    plain = ctx.resistance_batch(bodies, precision=api.PREC_FP64, eps=1e-12)
    f32 = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    monkeypatch.setenv("ANDR_GMRES_FP64_PRODUCTS", "1")
    f64 = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    monkeypatch.delenv("ANDR_GMRES_FP64_PRODUCTS")
    tight = ctx.resistance_batch(bodies, precision=api.PREC_MIXED, eps=1e-8)
    for p, a, b, t in zip(plain, f32, f64, tight):
        print(f"F {a.FT.shape}: its f32 {a.nit.max()} f64 {b.nit.max()} (1e-8: {t.nit.max()}); "
              f"err f32 {tensor_error(a.FT, p.FT):.1e} f64 {tensor_error(b.FT, p.FT):.1e} "
              f"1e-8 {tensor_error(t.FT, p.FT):.1e}")
        assert a.status == 0 and b.status == 0 and t.status == 0
        assert tensor_error(a.FT, p.FT) < 1e-6 and tensor_error(b.FT, p.FT) < 1e-6
        assert tensor_error(t.FT, p.FT) < 1e-6
        assert tensor_error(a.FT, b.FT) < 3e-7
        assert a.nit.max() <= 1.15 * b.nit.max() + 2
        assert t.nit.max() > b.nit.max()


def test_converged_system_skip_is_exact(ctx, monkeypatch):
    """The Krylov product skips a system once every column has converged in the cycle (until the
    host's next map rebuild). The steps after a column's convergence are never used, so tensors and
    step counts with and without the skip are the same bit for bit (autotuner off: fixed splits)."""
    allb = [m for m, _ in M.batch_bodies(512, 64)]
    bodies = [allb[i] for i in (29, 111, 222, 370, 519, 534)]  # superellipsoids and fibres, 500..5120
    ctx.autotune(False)
    try:
        a = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
        monkeypatch.setenv("ANDR_GMRES_NO_SKIP", "1")
        b = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
        monkeypatch.delenv("ANDR_GMRES_NO_SKIP")
    finally:
        ctx.autotune(True)
    for x, y in zip(a, b):
        assert x.status == 0 and y.status == 0
        assert np.array_equal(x.nit, y.nit)
        assert np.array_equal(x.FT, y.FT)


def test_block_gmres_vs_independent(ctx, monkeypatch):
    """Block GMRES (one Krylov space for the six rigid-body problems of a body, the default) and
    one Krylov space per problem (ANDR_GMRES_INDEPENDENT=1) solve the same systems: on superellipsoids
    and fibres of the batched workload both stay within 1e-6 of the plain FP64 path and within 2e-7
    of each other (each stops at half the tolerance on its own least-squares estimate, ~1e-8 to 7e-8
    from the plain path), and the block form needs fewer rounds than the slowest column of the other
    needs steps (a round and a step each read the body's matrix once)."""
    allb = [m for m, _ in M.batch_bodies(512, 64)]
    bodies = [allb[i] for i in (29, 204, 523, 534)]  # superellipsoids of 500 and 2420, fibres of 496 and 2000
    plain = ctx.resistance_batch(bodies, precision=api.PREC_FP64, eps=1e-12)
    blk = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    monkeypatch.setenv("ANDR_GMRES_INDEPENDENT", "1")
    ind = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    monkeypatch.delenv("ANDR_GMRES_INDEPENDENT")
    for p, a, b in zip(plain, blk, ind):
        print(f"rounds {a.nit.max()} vs steps {b.nit.tolist()}; err block {tensor_error(a.FT, p.FT):.1e} "
              f"independent {tensor_error(b.FT, p.FT):.1e}")
        assert a.status == 0 and b.status == 0
        assert tensor_error(a.FT, p.FT) < 1e-6 and tensor_error(b.FT, p.FT) < 1e-6
        assert tensor_error(a.FT, b.FT) < 2e-7
        assert a.nit.max() < b.nit.max()


def test_krylov_product_autotuner(ctx):
    """A slow fibre next to a fast body: once the fast body is done the fibre alone is live, its
    columns are split into segments, and the autotuner times the candidate splits for that bucket
    and keeps the fastest. With the tuner off (fixed split) the tensors agree to rounding."""
    allb = [m for m, _ in M.batch_bodies(512, 64)]
    bodies = [allb[523], allb[29]]
    ctx.autotune(True)
    tuned = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    table = ctx.autotune_table_krylov()
    print("krylov autotuner:", table)
    assert table and all(e["target_waves"] in (1024, 2048, 4096, 8192) and e["ms"] > 0 for e in table)
    ctx.autotune(False)
    try:
        fixed = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    finally:
        ctx.autotune(True)
    for a, b in zip(tuned, fixed):
        assert a.status == 0 and b.status == 0
        assert tensor_error(a.FT, b.FT) < 1e-9
        assert abs(int(a.nit.max()) - int(b.nit.max())) <= 2


def test_slender_body_long_restarts(ctx):
    """A 496-element fibre of the batched workload that needs more Krylov rounds than a first cycle
    of 48 holds (about 60 block rounds): the cycles after the first run with a longer restart in
    the same buffers (about twice as long here), next to a body that converges in the first cycle;
    both stay within 1e-6 of the plain FP64 path. With the automatic first cycle (longer while the
    bodies fit one wave) the fibre converges without a restart, in fewer rounds, to the same tensor.
    (Much shorter cycles stall the fibre's hardest column, as restarted GMRES does on such
    operators: a first cycle of 32 leaves cycles of 64 after it, and its column reaches the
    iteration limit.)"""
    allb = [m for m, _ in M.batch_bodies(512, 64)]
    bodies = [allb[523], allb[29]]
    plain = ctx.resistance_batch(bodies, precision=api.PREC_FP64, eps=1e-12)
    fast = ctx.resistance_batch(bodies, precision=api.PREC_MIXED, restart=48)
    print("restarted rounds", fast[0].nit.tolist(), fast[1].nit.tolist())
    for f, p in zip(fast, plain):
        assert f.status == 0
        assert tensor_error(f.FT, p.FT) < 1e-6
    assert fast[0].nit.max() > 48 >= fast[1].nit.max()
    auto = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    assert auto[0].nit.max() <= 192 and auto[0].nit.max() < fast[0].nit.max()
    for f, p in zip(auto, plain):
        assert tensor_error(f.FT, p.FT) < 1e-6


def test_tiny_bodies_krylov_space_exhausted(ctx):
    """20- and 80-element spheres (60 and 240 unknowns): GMRES can span the whole space, so the
    Arnoldi process meets a (near) breakdown; the fast forms still agree with the plain path, and a
    batch with one tiny and one ordinary body gives the same tensors as the bodies alone."""
    tiny = [M.icosphere(0), M.icosphere(1)]
    for m in tiny:
        p = ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12)
        f = ctx.resistance(m, precision=api.PREC_MIXED)
        assert np.all(np.isfinite(f.FT)) and tensor_error(f.FT, p.FT) < 1e-6
    mixed = [tiny[0], M.icosphere(3)]
    rb = ctx.resistance_batch(mixed, precision=api.PREC_MIXED)
    for b, r in zip(mixed, rb):
        assert _rel(r.FT, ctx.resistance(b, precision=api.PREC_MIXED).FT) < 1e-12


def test_batch_equals_single(ctx):
    bodies = [M.random_superellipsoid(np.random.default_rng(s), 4)[0] for s in range(3)]
    for prec in (api.PREC_FP64, api.PREC_MIXED):
        rb = ctx.resistance_batch(bodies, precision=prec)
        for b, r in zip(bodies, rb):
            s = ctx.resistance(b, precision=prec)
            assert _rel(r.FT, s.FT) < 1e-12


def test_plain_batch_streams_bitwise(ctx):
    """Six plain FP64 solves of different sizes share four streams (two streams carry two solves):
    every tensor and iteration count is bit-for-bit that of the body solved alone."""
    bodies = [M.random_superellipsoid(np.random.default_rng(10 + i), nu)[0] for i, nu in enumerate((3, 5, 4, 6, 3, 4))]
    rb = ctx.resistance_batch(bodies, precision=api.PREC_FP64, eps=1e-12)
    for b, r in zip(bodies, rb):
        s = ctx.resistance(b, precision=api.PREC_FP64, eps=1e-12)
        assert np.array_equal(r.FT, s.FT) and np.array_equal(r.nit, s.nit)


def test_resident_bodies_match_host_meshes(ctx):
    """Bodies uploaded once give, at every form, the tensors and iteration counts of the call that
    takes host meshes, bit for bit; a viscosity other than the upload's is refused."""
    bodies = [M.random_superellipsoid(np.random.default_rng(20 + i), nu)[0] for i, nu in enumerate((3, 5, 4))]
    res = ctx.upload_bodies(bodies)
    try:
        assert len(res) == 3
        for prec in (api.PREC_FP64, api.PREC_MIXED):
            a = ctx.resistance_batch(bodies, precision=prec)
            b = ctx.resistance_batch_resident(res, precision=prec)
            for x, y in zip(a, b):
                assert np.array_equal(x.FT, y.FT) and np.array_equal(x.nit, y.nit)
        res.mu = 2.0
        with pytest.raises(RuntimeError):
            ctx.resistance_batch_resident(res)
    finally:
        res.free()


def test_sphere_20480_drag_and_fast_form(ctx):
    """The rigid-sphere configuration at full size: icosphere subdivision 5 (20 480 triangles, 61 440
    traction unknowns, a 30 GB FP64 system), unit radius and viscosity. The plain FP64 path gives the
    Stokes drag 6 pi mu a U and the rotational torque 8 pi mu a^3 to the mesh's discretisation error,
    and the fast form stays within 1e-6 of it."""
    m = M.icosphere(5)
    assert m.nelem == 20480
    p = ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12)
    assert p.status == 0
    K = np.diag(p.FT[:3, :3])
    Om = np.diag(p.FT[3:, 3:])
    assert np.all(np.abs(K / (6 * np.pi) - 1) < 5e-3), K
    assert np.all(np.abs(Om / (8 * np.pi) - 1) < 5e-3), Om
    f = ctx.resistance(m, precision=api.PREC_MIXED)
    assert tensor_error(f.FT, p.FT) < 1e-6


def _seam_system(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 6 * np.sqrt(n) * np.eye(n)
    return A, rng.normal(size=n), rng.normal(size=n)


def test_fortran_lsqr_shim_hot_start_and_iteration_limit(ctx, orc):
    """lsqr_solve_ as the reference's drivers call it: a hot start x0 != 0 (src/lib_lsqr.f:84-113:
    x <- x/q, u = b - A x, x = x0 + q dx) and maxit = 3 (istop = ierr = 5, nit = 3) give the
    oracle's iterates; maxit is not written back when nonzero."""
    n = 200
    A, b, x0 = _seam_system(n, 21)
    ia, ja, a = orc.csr_from_dense(A)
    qo = orc.lsqr_precon(ia, ja, a, n, n)
    ia1, ja1, a1 = ia + 1, ja + 1, A.reshape(-1).copy()
    q = api.lsqr_precon(n, n, ia1, ja1, a1)
    for maxit, eps in ((0, 1e-14), (3, 1e-15)):
        xo, info = orc.lsqr_solve(ia, ja, a, n, n, qo, b, x0=x0, maxit=maxit, eps=eps)
        x = x0.copy()
        r = api.lsqr_solve(n, n, ia1, ja1, a1, q, b, x, maxit=maxit, eps=eps)
        print(f"maxit {maxit}: nit {r['nit']} (oracle {info['nit']}) ierr {r['ierr']} (oracle {info['ierr']})")
        if maxit == 3:
            assert r["nit"] == info["nit"] == 3 and r["ierr"] == info["ierr"] == 5 and r["maxit"] == 3
            assert np.abs(x - xo).max() / np.abs(xo).max() < 1e-12
        else:
            assert r["ierr"] == 0 and abs(r["nit"] - info["nit"]) <= 2
            assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10


def _sparse_csr(m, n, per_row, seed, dup_rows=()):
    """1-based CSR: a dominant entry at (i, i mod n) plus per_row random columns per row; the rows in
    dup_rows store their dominant entry twice (two halves)."""
    rng = np.random.default_rng(seed)
    ia, ja, vals = [1], [], []
    for i in range(m):
        cols = rng.choice(n, size=per_row, replace=False)
        cols = cols[cols != i % n]
        ja += [c + 1 for c in cols]
        vals += list(rng.normal(size=cols.size))
        d = 4.0 + rng.uniform()
        if i in dup_rows:
            ja += [i % n + 1, i % n + 1]
            vals += [0.5 * d, 0.5 * d]
        else:
            ja.append(i % n + 1)
            vals.append(d)
        ia.append(len(ja) + 1)
    return np.array(ia, np.int64), np.array(ja, np.int32), np.array(vals)


def test_fortran_lsqr_shim_sparse_csr(ctx, orc):
    """A genuinely sparse CSR (under a third of the entries) stays sparse on the device: A and its
    transpose in CSR, products row by row. Against the oracle's CSR LSQR on the same matrix (with
    duplicate entries in some rows): the column scales bit for bit (both sum the squared stored
    entries in row order), the solution to 1e-10, iteration count and stop code alike. A 200000 x
    200000 system with 6 entries per row (a dense copy would take 320 GB) solves too."""
    m, n = 3000, 2000
    ia, ja, a = _sparse_csr(m, n, 7, 31, dup_rows={3, 50, 999})
    b = np.random.default_rng(32).normal(size=m)
    ao = a.copy()
    qo = orc.lsqr_precon(ia - 1, ja - 1, ao, m, n)
    q = api.lsqr_precon(m, n, ia, ja, a)
    assert np.array_equal(q, qo) and np.array_equal(a, ao)
    x = np.zeros(n)
    r = api.lsqr_solve(m, n, ia, ja, a, q, b, x, eps=1e-12)
    xo, info = orc.lsqr_solve(ia - 1, ja - 1, ao, m, n, qo, b, eps=1e-12)
    print(f"sparse {m}x{n}: its {r['nit']} (oracle {info['nit']}), ierr {r['ierr']} ({info['ierr']}), "
          f"x rel {np.abs(x - xo).max() / np.abs(xo).max():.1e}")
    assert r["ierr"] == info["ierr"] and abs(r["nit"] - info["nit"]) <= 1
    assert np.abs(x - xo).max() < 1e-10 * np.abs(xo).max()
    # large: only the sparse form fits
    M = 200000
    ia, ja, a = _sparse_csr(M, M, 5, 33)
    bl = np.random.default_rng(34).normal(size=M)
    ql = api.lsqr_precon(M, M, ia, ja, a)
    xl = np.zeros(M)
    r = api.lsqr_solve(M, M, ia, ja, a, ql, bl, xl, eps=1e-10)
    # residual of the unscaled system: A diag(q) (x / q) = b, a holds A diag(q)
    rows = np.repeat(np.arange(M), np.diff(ia))
    Ax = np.bincount(rows, weights=a * xl[ja - 1] / ql[ja - 1], minlength=M)
    res = np.linalg.norm(Ax - bl) / np.linalg.norm(bl)
    print(f"sparse {M}^2: its {r['nit']}, ierr {r['ierr']}, relative residual {res:.1e}")
    assert r["ierr"] == 0 and res < 1e-8
    api.lsqr_forget()


def test_fortran_lsqr_shim_duplicates_cache_and_failure(ctx, orc):
    """A CSR with duplicate (i, j) entries solves as their sum; a solve after the values change in
    place (same arrays) uploads the new matrix; a failed lsqr_precon_ sets q to NaN and the next
    lsqr_solve_ returns ierr = -1 until an lsqr_precon_ succeeds."""
    n = 120
    A, b, _ = _seam_system(n, 22)
    # duplicates: every diagonal entry split in two halves, stored twice in its row
    ia, ja, vals = [1], [], []
    for i in range(n):
        for j in range(n):
            if i == j:
                ja += [j + 1, j + 1]
                vals += [0.5 * A[i, j], 0.5 * A[i, j]]
            else:
                ja.append(j + 1)
                vals.append(A[i, j])
        ia.append(len(ja) + 1)
    ia1, ja1, a1 = np.array(ia, np.int64), np.array(ja, np.int32), np.array(vals)
    q = api.lsqr_precon(n, n, ia1, ja1, a1)
    x = np.zeros(n)
    r = api.lsqr_solve(n, n, ia1, ja1, a1, q, b, x, eps=1e-14)
    assert r["ierr"] == 0 and np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-10
    # same arrays, new values: the cached device copy must not be used
    a1 *= 2.0
    x2 = np.zeros(n)
    r = api.lsqr_solve(n, n, ia1, ja1, a1, q, b, x2, eps=1e-14)
    assert r["ierr"] == 0 and np.linalg.norm(x2 - 0.5 * x) / np.linalg.norm(x) < 1e-9
    # a failed precon (a column index out of range) is visible in q and in the next solve
    bad = ja1.copy()
    bad[5] = n + 3
    qb = api.lsqr_precon(n, n, ia1, bad, a1.copy())
    assert np.all(np.isnan(qb))
    r = api.lsqr_solve(n, n, ia1, ja1, a1, q, b, np.zeros(n), eps=1e-14)
    assert r["ierr"] == -1
    api.lsqr_precon(n, n, ia1, ja1, a1.copy())
    r = api.lsqr_solve(n, n, ia1, ja1, a1, q, b, np.zeros(n), eps=1e-14)
    assert r["ierr"] == 0
