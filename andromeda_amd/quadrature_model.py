"""Training of the learned rule selection (the ANDR_PREC_LEARNED form of the resistance solver).

The reference integrates every regular (collocation point, element) pair with the run's one
fixed rule (src/lib_Stokes.f90:2341-2446). The learned form chooses per pair the cheapest
Dunavant rule in {3, 4, 6, 7, 12, 13, 19} points whose G and T integrals agree with the fixed
rule to `rel_tol` (class 7: keep the fixed rule):

* near pairs (distance / longest edge < near_ratio) go through a 6 -> 3x64 -> 8 network;
  a pair whose softmax confidence is below `confidence` keeps the fixed rule;
* far pairs take the class of their distance bin from a table calibrated on the same meshes
  (largest label seen per bin, made monotone in distance).

Everything runs on the device: labels from FP64 integrals (andr_quad_training_pairs,
andr_quad_far_calibration), training by the library's forward/backward/Adam kernels
(andr_mlp_train), inference for the validation statistics (andr_mlp_infer).

    python -m andromeda_amd.quadrature_model --out andromeda_amd/quad_model.npz

trains on the meshes of seeds 0-99 and validates on seeds 100-199.
"""
import argparse
import math
import sys
import time

import numpy as np

from . import api
from . import mesh as M

NCLS = 8
DUN_POINTS = np.array([3, 4, 6, 7, 12, 13, 19], dtype=np.int64)


def seeded_body(seed: int):
    """Body of one seed: a fibre for seed % 9 == 8, else a superellipsoid; element count
    log-uniform in [500, 8000] (the batched workload's distribution)."""
    rng = np.random.default_rng(seed)
    f = math.exp(rng.uniform(math.log(500.0), math.log(8000.0)))
    nu = M.nu_for_elements(f)
    if seed % 9 == 8:
        return M.random_fibre(rng, 20 * nu * nu)
    return M.random_superellipsoid(rng, nu)


def large_body(seed: int):
    """8000-triangle superellipsoid of one seed (the single-body precision workload)."""
    rng = np.random.default_rng(10_000 + seed)
    return M.random_superellipsoid(rng, 20)


def bodies_for(seeds, n_large: int):
    out = [seeded_body(s)[0] for s in seeds]
    out += [large_body(s)[0] for s in list(seeds)[:n_large]]
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He initialisation of the 6 -> 64 -> 64 -> 64 -> 8 network (weights input-major)."""
    rng = np.random.default_rng(seed)
    parts = []
    for nin, nout in ((6, 64), (64, 64), (64, 64), (64, 8)):
        parts.append(rng.standard_normal(nin * nout) * math.sqrt(2.0 / nin))
        parts.append(np.zeros(nout))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == api.MLP_NPARAM
    return p


def collect_pairs(ctx, meshes, near_ratio, rel_tol, per_mesh, rng, log=print):
    """Near-pair records of every mesh, `per_mesh` of them drawn uniformly from each."""
    F, L = [], []
    for k, m in enumerate(meshes):
        # the device samples pairs uniformly down to about 2 per_mesh records
        f, lab = ctx.quad_training_pairs(m, near_ratio=near_ratio, rel_tol=rel_tol, max_pairs=2 * per_mesh)
        pick = rng.choice(len(lab), size=min(per_mesh, len(lab)), replace=False)
        F.append(f[pick])
        L.append(lab[pick])
        if k % 20 == 0:
            log(f"[quad] mesh {k}/{len(meshes)}: F={m.nelem}, {len(lab)} near records, "
                f"labels {np.bincount(lab, minlength=NCLS).tolist()}")
    return np.concatenate(F), np.concatenate(L)


def far_table(ctx, meshes, near_ratio, rel_tol):
    """Largest class per distance bin over all meshes; unseen bins take the nearer bin's class
    and the table is made non-increasing with distance."""
    t = np.full(api.FAR_BINS, -1, dtype=np.int64)
    for m in meshes:
        t = np.maximum(t, ctx.quad_far_calibration(m, near_ratio, rel_tol))
    last = NCLS - 1
    for b in range(api.FAR_BINS):
        if t[b] < 0:
            t[b] = last
        last = t[b]
    for b in range(api.FAR_BINS - 2, -1, -1):
        t[b] = max(t[b], t[b + 1])
    return t.astype(np.int32)


def decide(probs, conf):
    """Class chosen per record and whether it is taken (not class 7). The classes are ordered (a
    rule good enough for a pair stays good enough for every larger class), so the choice is the
    smallest class whose cumulative probability reaches conf: the conf-quantile of the predicted
    label."""
    cum = np.cumsum(probs, axis=1)
    reached = cum >= conf * cum[:, -1:]
    cls = np.where(reached.any(axis=1), reached.argmax(axis=1), NCLS - 1)
    taken = cls < NCLS - 1
    return cls, taken


def validation_stats(probs, labels, conf, nq_fixed=126):
    cls, taken = decide(probs, conf)
    unsafe = taken & (cls < labels)
    pts = np.where(taken, DUN_POINTS[np.minimum(cls, 6)], nq_fixed)
    best = np.where(labels < NCLS - 1, DUN_POINTS[np.minimum(labels, 6)], nq_fixed)
    return {"confidence": float(conf), "taken": float(taken.mean()), "unsafe": float(unsafe.mean()),
            "mean_points": float(pts.mean()), "oracle_points": float(best.mean()),
            "accuracy": float((cls == labels).mean())}


CONFIDENCES = (0.9, 0.95, 0.98, 0.99, 0.995, 0.998, 0.999, 0.9995)


def select_confidence(probs, labels, max_unsafe, log=print):
    """The quantile level with the fewest points whose validation unsafe rate (chosen class below the
    label) stays within max_unsafe."""
    stats = [validation_stats(probs, labels, c) for c in CONFIDENCES]
    ok = [s for s in stats if s["unsafe"] <= max_unsafe]
    chosen = min(ok, key=lambda s: s["mean_points"]) if ok else stats[-1]
    for s in stats:
        log(f"[quad] validation conf {s['confidence']:.4f}: taken {s['taken']:.3f} unsafe {s['unsafe']:.2e} "
            f"points {s['mean_points']:.1f} (labels {s['oracle_points']:.1f}) accuracy {s['accuracy']:.3f}")
    log(f"[quad] chosen confidence {chosen['confidence']}")
    return chosen


def reselect(ctx, path, max_unsafe=1e-4, log=print):
    """Chooses the decision level of a trained model again on its validation bodies (same seeds and
    sampling as training) and rewrites the model file; the network parameters are kept."""
    z = dict(np.load(path))
    near_ratio, rel_tol = float(z["near_ratio"]), float(z["rel_tol"])
    lo, hi = (int(v) for v in z["val_seeds"])
    rng = np.random.default_rng(0)
    va = bodies_for(range(lo, hi + 1), 10)
    Xva, Yva = collect_pairs(ctx, va, near_ratio, rel_tol, 20000 // 4, rng, log)
    chosen = select_confidence(ctx.mlp_infer(Xva, z["params"]), Yva, max_unsafe, log)
    z["confidence"] = np.float32(chosen["confidence"])
    z["val_unsafe"] = np.float64(chosen["unsafe"])
    z["val_points"] = np.float64(chosen["mean_points"])
    np.savez(path, **z)
    log(f"[quad] wrote {path}")
    return z


def train(ctx, near_ratio=4.0, rel_tol=1e-7, train_seeds=range(0, 100), val_seeds=range(100, 200),
          n_large=10, per_mesh=20000, epochs=30, batch=4096, lr=2e-3, max_unsafe=1e-4, seed=0, log=print):
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    tr = bodies_for(train_seeds, n_large)
    va = bodies_for(val_seeds, n_large)
    Xtr, Ytr = collect_pairs(ctx, tr, near_ratio, rel_tol, per_mesh, rng, log)
    Xva, Yva = collect_pairs(ctx, va, near_ratio, rel_tol, per_mesh // 4, rng, log)
    log(f"[quad] {len(Ytr)} training and {len(Yva)} validation records ({time.perf_counter() - t0:.1f} s); "
        f"training labels {np.bincount(Ytr, minlength=NCLS).tolist()}")
    t1 = time.perf_counter()
    far = far_table(ctx, tr, near_ratio, rel_tol)
    log(f"[quad] far table {far.tolist()} ({time.perf_counter() - t1:.1f} s)")
    t2 = time.perf_counter()
    params, losses = ctx.mlp_train(Xtr, Ytr, init_params(seed), epochs=epochs, batch=batch, lr=lr, seed=seed)
    log(f"[quad] trained {epochs} epochs in {time.perf_counter() - t2:.1f} s; loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    chosen = select_confidence(ctx.mlp_infer(Xva, params), Yva, max_unsafe, log)
    return {"params": params, "confidence": np.float32(chosen["confidence"]), "near_ratio": np.float64(near_ratio),
            "far_table": far, "rel_tol": np.float64(rel_tol), "losses": losses,
            "train_seeds": np.array([min(train_seeds), max(train_seeds)]),
            "val_seeds": np.array([min(val_seeds), max(val_seeds)]),
            "val_unsafe": np.float64(chosen["unsafe"]), "val_points": np.float64(chosen["mean_points"]),
            "n_train": np.int64(len(Ytr))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", help="train a model and write it here")
    ap.add_argument("--reselect", help="choose the decision level of this model file again")
    ap.add_argument("--near-ratio", type=float, default=4.0)
    ap.add_argument("--rel-tol", type=float, default=1e-7)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--per-mesh", type=int, default=20000)
    ap.add_argument("--train-seeds", type=int, default=100)
    args = ap.parse_args(argv)
    ctx = api.Context(0)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if args.reselect:
        reselect(ctx, args.reselect, log=log)
        return 0
    if not args.out:
        ap.error("--out or --reselect is required")
    model = train(ctx, near_ratio=args.near_ratio, rel_tol=args.rel_tol, epochs=args.epochs,
                  per_mesh=args.per_mesh, train_seeds=range(0, args.train_seeds),
                  val_seeds=range(100, 100 + args.train_seeds), log=log)
    np.savez(args.out, **model)
    log(f"[quad] wrote {args.out}")


if __name__ == "__main__":
    sys.exit(main())
