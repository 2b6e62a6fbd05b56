"""Learned rule selection: the network's forward/backward kernels against the CPU oracle (itself
pinned by finite differences), the FP64 label generator against a numpy restatement, the
in-assembly BF16 classifier against the oracle's decisions, and what the trained network gains
over the fixed-order rule at the resistance tolerance."""
import math
import os
import re
import time

import numpy as np
import pytest

import mlp_oracle as MO
import quad_label_oracle as QL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "andromeda_amd", "quad_model.npz")


def dunavant_rules():
    src = open(os.path.join(ROOT, "andromeda_amd", "csrc", "andr_dunavant_rules.h")).read()
    rows = re.findall(r"\{([-0-9.e]+), ([-0-9.e]+), ([-0-9.e]+), ([-0-9.e]+)\}", src)
    P = np.array([[float(x) for x in r] for r in rows])
    offs = [0, 3, 7, 13, 20, 32, 45, 64]
    return [P[offs[i]:offs[i + 1]] for i in range(7)]


def fixed_rule(orc, itri=5):
    l1, l2, l3, w = orc.tri_rule(itri)
    return np.stack([l1, l2, l3, w], axis=1)


def synthetic_set(n, seed):
    """Features shaped like pair records, labels a fixed function of them."""
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 6), dtype=np.float32)
    X[:, 0] = rng.uniform(-1.0, 1.4, n)
    X[:, 1] = rng.uniform(0.0, 2.0, n)
    X[:, 2] = rng.uniform(-1.0, 1.0, n)
    k = rng.integers(0, 3, n)
    X[np.arange(n), 3 + k] = 1.0
    y = np.clip(np.floor((1.4 - X[:, 0]) * 3.0 + 0.5 * X[:, 1]), 0, 7).astype(np.int32)
    return X, y


# ---- CPU: the oracle itself ----------------------------------------------------------------------
def test_oracle_gradient_matches_finite_differences():
    X, y = synthetic_set(40, 1)
    p = MO.init_params(3).astype(np.float64)
    p += 0.05 * np.random.default_rng(4).standard_normal(p.size)  # nonzero biases too
    L, g = MO.loss_grad(p, X, y)
    assert abs(L - MO.loss(p, X, y)) < 1e-12
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(MO.NPARAM, 300, replace=False),
                          np.arange(MO.NPARAM - 8, MO.NPARAM)])  # all output biases
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (MO.loss(pp, X, y) - MO.loss(pm, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], dtype=np.float32)
    r = MO.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0          # halfway: ties to even
    assert r[2] == np.float32(1.0 + 2 ** -6)    # halfway up to the even neighbour
    assert r[3] == -2.5
    assert abs(r[4] - 1e-3) <= 2 ** -8 * 1e-3


def test_label_oracle_point_kernels(orc):
    """The numpy point kernels reproduce the C oracle's BemKernel and StokesKernel."""
    import ctypes as C
    rng = np.random.default_rng(0)
    lib = orc.lib()
    for _ in range(20):
        s, f, n = rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3)
        n /= np.linalg.norm(n)
        g, h = C.c_double(), C.c_double()
        lib.orc_bem_kernel(orc._p(s), orc._p(f), orc._p(n), C.byref(g), C.byref(h))
        G, H = QL.lap_point(s, f[None], n)
        assert abs(G[0] - g.value) < 1e-14 * abs(g.value) and abs(H[0] - h.value) < 1e-13 * abs(h.value) + 1e-15
        G9, T9, pg, pt = np.zeros(9), np.zeros(9), np.zeros(3), np.zeros(3)
        lib.orc_stokes_kernel(orc._p(s), orc._p(f), orc._p(n), orc._p(G9), orc._p(T9), orc._p(pg), orc._p(pt))
        Gs, Ts = QL.stokes_point(s, f[None], n)
        assert np.allclose(Gs[0].ravel(), G9, rtol=1e-13, atol=1e-14)
        assert np.allclose(Ts[0].ravel(), T9, rtol=1e-12, atol=1e-14)


def test_label_oracle_rules_are_nested(orc):
    """Far from the element every rule agrees (label 0); on the element's neighbour only the fixed
    rule does (label 7) at a tight tolerance."""
    fr, dr = fixed_rule(orc), dunavant_rules()
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    n = np.array([0.0, 0.0, 1.0])
    area = 0.5
    for k in range(3):
        assert QL.label(k, np.array([1600.0, 1200.0, 600.0]), x, n, area, fr, dr, 1e-7) == 0
        assert QL.label(k, np.array([0.5, 0.5, 0.05]), x, n, area, fr, dr, 1e-12) == 7


def test_far_table_fill_and_monotone():
    from andromeda_amd import quadrature_model as QM

    class Fake:
        def __init__(self, tabs):
            self.tabs = iter(tabs)

        def quad_far_calibration(self, m, near, tol):
            return np.array(next(self.tabs), dtype=np.int32)

    a = [5, 4, 4, 3, 3, 2, 1, 1, -1, -1, -1, -1, -1, -1, -1, -1]
    b = [6, 3, 4, 2, 3, 1, 2, 0, 0, -1, -1, -1, -1, -1, -1, -1]
    t = QM.far_table(Fake([a, b]), [None, None], 4.0, 1e-7)
    assert t.tolist() == [6, 4, 4, 3, 3, 2, 2, 1, 0, 0, 0, 0, 0, 0, 0, 0]
    t = QM.far_table(Fake([[-1] * 16]), [None], 4.0, 1e-7)
    assert t.tolist() == [7] * 16  # nothing seen: keep the fixed rule everywhere


def test_validation_stats():
    from andromeda_amd import quadrature_model as QM
    probs = np.array([[0.95, 0.05, 0, 0, 0, 0, 0, 0],   # 0.9-quantile class 0, label 1: unsafe
                      [0, 0, 0.6, 0.4, 0, 0, 0, 0],     # quantile reached at class 3 (7 points), label 2
                      [0, 0, 0, 0, 0, 0, 0.5, 0.5]])    # quantile at class 7: fixed rule
    s = QM.validation_stats(probs, np.array([1, 2, 7]), 0.9)
    assert s["taken"] == pytest.approx(2 / 3) and s["unsafe"] == pytest.approx(1 / 3)
    assert s["mean_points"] == pytest.approx((3 + 7 + 126) / 3)
    cls, taken = QM.decide(probs, 0.5)  # the median: classes 0, 2 and 6
    assert cls.tolist() == [0, 2, 6] and taken.tolist() == [True, True, True]


# ---- GPU: kernels against the oracle -------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_loss_grad_matches_oracle(ctx):
    X, y = synthetic_set(1000, 2)
    p = MO.init_params(7)
    p = (p + 0.05 * np.random.default_rng(8).standard_normal(p.size)).astype(np.float32)
    L, g = ctx.mlp_loss_grad(X, y, p)
    Lo, go = MO.loss_grad(p.astype(np.float64), X, y)
    assert abs(L - Lo) <= 1e-5 * abs(Lo)
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go)
    big = np.abs(go) > 1e-3 * np.abs(go).max()
    assert np.max(np.abs(g[big] - go[big]) / np.abs(go[big])) < 1e-2


@pytest.mark.gpu
def test_gpu_infer_matches_oracle(ctx):
    X, _ = synthetic_set(70000, 3)  # crosses the 65536-record chunk boundary
    p = MO.init_params(9)
    P = ctx.mlp_infer(X, p)
    Z, _ = MO.forward(p.astype(np.float64), X)
    Po = MO.softmax(Z)
    assert np.max(np.abs(P - Po)) < 2e-5


@pytest.mark.gpu
def test_gpu_training_learns(ctx):
    X, y = synthetic_set(20000, 4)
    p0 = MO.init_params(0)
    L0 = MO.loss(p0, X, y)
    t0 = time.perf_counter()
    p, losses = ctx.mlp_train(X, y, p0, epochs=20, batch=512, lr=3e-3, seed=1)
    dt = time.perf_counter() - t0
    print(f"training: {20 * (20000 // 512)} minibatches of 512 in {dt * 1e3:.1f} ms "
          f"({dt * 1e6 / (20 * (20000 // 512)):.1f} us per forward, backward and Adam step)")
    L1 = MO.loss(p.astype(np.float64), X, y)
    assert losses[-1] < losses[0] and L1 < 0.3 * L0
    Xv, yv = synthetic_set(5000, 5)
    acc = (ctx.mlp_infer(Xv, p).argmax(1) == yv).mean()
    assert acc > 0.9, acc


def _match_labels(Fg, Lg, Fo, Lo, radius=2e-5):
    """Fraction of device records whose label equals the label of an oracle record with the same
    features (within radius); pairs related by the mesh's symmetry share features."""
    from scipy.spatial import cKDTree
    tree = cKDTree(Fo)
    hits = tree.query_ball_point(Fg, r=radius)
    found = np.array([len(h) > 0 for h in hits])
    agree = np.array([len(h) > 0 and Lg[i] in set(Lo[h].tolist()) for i, h in enumerate(hits)])
    near = np.array([len(h) > 0 and np.min(np.abs(Lo[h] - Lg[i])) <= 1 for i, h in enumerate(hits)])
    return found.mean(), agree.mean(), near.mean()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["sphere", "superellipsoid"])
def test_gpu_labels_match_oracle(ctx, orc, shape):
    from andromeda_amd import mesh as M
    m = M.icosphere(1) if shape == "sphere" else M.superellipsoid([1.6, 1.0, 0.7], 4.0, 2)
    near, tol = 4.0, 1e-7
    Fg, Lg = ctx.quad_training_pairs(m, near_ratio=near, rel_tol=tol, max_pairs=200000)
    Fo, Lo = QL.near_records(m.xyz, m.con, fixed_rule(orc), dunavant_rules(), near, tol)
    assert len(Fg) == len(Fo)
    assert np.array_equal(np.bincount(Fg[:, 3:].argmax(1)), np.bincount(Fo[:, 3:].argmax(1)))
    found, agree, near1 = _match_labels(Fg.astype(np.float64), Lg, Fo, Lo)
    assert found == 1.0
    assert agree >= 0.995, agree
    assert near1 == 1.0


def _expected_classes(m, orc, model, mesh_nq=126):
    """Oracle decisions of the in-assembly classifier for every centroid-row pair of a body."""
    from andromeda_amd import quadrature_model as QM
    x, nrm, area, cen = QL.geometry(m.xyz, m.con)
    near, conf, far = float(model["near_ratio"]), float(model["confidence"]), model["far_table"]
    pts = learned = fallback = 0
    feats, where = [], []
    cls = {}
    for r in range(len(m.con)):
        for e in range(len(m.con)):
            if e == r:
                continue
            dh, asp, cs = QL.pair_geo(cen[r], nrm[r], x[e], nrm[e], area[e])
            if dh >= near:
                cls[(r, e)] = int(far[QL.far_bin(dh, near)])
            else:
                feats += [QL.features(dh, asp, cs, 1), QL.features(dh, asp, cs, 2)]
                where.append((r, e))
    if feats:
        # BF16 inputs, weights and activations as on the matrix cores; biases FP32
        Z = MO.forward_bf16(np.asarray(model["params"], np.float64), np.array(feats, dtype=np.float32))
        c, _ = QM.decide(MO.softmax(Z), conf)
        for i, (r, e) in enumerate(where):
            a, b = c[2 * i], c[2 * i + 1]
            if a == 7 or b == 7:
                cls[(r, e)] = 7
                fallback += 1
            else:
                cls[(r, e)] = int(max(a, b))
                learned += 1
    for v in cls.values():
        pts += mesh_nq if v == 7 else QL.DUN_POINTS[v]
    return pts, learned, fallback


def _model():
    if not os.path.exists(MODEL):
        pytest.skip("no trained model file (python -m andromeda_amd.quadrature_model --out ...)")
    z = np.load(MODEL)
    return {k: z[k] for k in z.files}


@pytest.mark.gpu
def test_gpu_classifier_matches_oracle(ctx, orc):
    from andromeda_amd import api
    from andromeda_amd import mesh as M
    model = _model()
    ctx.set_quadrature_model(model["params"], float(model["confidence"]), float(model["near_ratio"]),
                             model["far_table"])
    m = M.superellipsoid([1.5, 1.0, 0.8], 3.0, 3)
    r = ctx.resistance(m, precision=api.PREC_LEARNED)
    pts, nl, nf = _expected_classes(m, orc, model)
    assert r.pairs_learned + r.pairs_fallback == nl + nf
    assert abs(r.pairs_learned - nl) <= 0.005 * (nl + nf) + 1
    assert abs(r.quad_points - pts) <= 0.005 * pts


@pytest.mark.gpu
def test_gpu_learned_gain(ctx):
    """Held-out bodies (seeds 100-199 were not trained on): the learned form stays within 1e-6 of
    the plain FP64 path while integrating far fewer quadrature points than the fixed rule, and
    the network itself (not only the far table) removes points."""
    from andromeda_amd import api
    from andromeda_amd import quadrature_model as QM
    from andromeda_amd.metrics import tensor_error
    model = _model()
    bodies = [QM.seeded_body(s)[0] for s in (102, 103, 107)]
    plain = ctx.resistance_batch(bodies, precision=api.PREC_FP64, eps=1e-12)
    mixed = ctx.resistance_batch(bodies, precision=api.PREC_MIXED)
    # network switched off: every near pair keeps the fixed rule, far pairs use the table
    off = np.zeros(api.MLP_NPARAM, dtype=np.float32)
    off[-1] = 10.0  # output bias of class 7
    ctx.set_quadrature_model(off, float(model["confidence"]), float(model["near_ratio"]), model["far_table"])
    table_only = ctx.resistance_batch(bodies, precision=api.PREC_LEARNED)
    ctx.set_quadrature_model(model["params"], float(model["confidence"]), float(model["near_ratio"]),
                             model["far_table"])
    learned = ctx.resistance_batch(bodies, precision=api.PREC_LEARNED)
    for p, mx, t, l in zip(plain, mixed, table_only, learned):
        assert tensor_error(l.FT, p.FT) <= 1e-6
        assert tensor_error(t.FT, p.FT) <= 1e-6
        assert l.quad_points < t.quad_points < mx.quad_points
        assert l.pairs_learned > 0
    # over the three bodies: at most a quarter of the fixed rule's quadrature points
    assert sum(l.quad_points for l in learned) < 0.25 * sum(mx.quad_points for mx in mixed)
