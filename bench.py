"""Resistance-tensor throughput on one MI355X.

Workload: the batched configuration of the benchmark -- 512 seeded superellipsoids
(|x/a|^p + |y/b|^p + |z/c|^p = 1, axes U[0.5, 2], p U[2, 8]) and 64 fibre-like bodies (aspect
U[10, 50]), element counts log-uniform in [500, 8000] rounded to geodesic sizes 20 nu^2, six
rigid-body problems per body (single-surface exterior Stokes). One step computes all 576 6x6
resistance tensors with the fast form (batched, FP32 quadrature and storage, FP64 solve; with the
learned Gauss-order selection when a trained model file is present). The plain FP64 path (FP64
assembly, FP64 LSQR, the drop-in) and the fast form are both timed on a size-stratified sample
of the same bodies, which also gives the fast form's error against FP64. The bodies are uploaded
to the device before the timed region (inputs resident in HBM); one more step from host meshes
gives the rate with the upload over PCIe. The CPU oracle is timed on a bounded sample on the host
cores.

Prints one JSON line (metric, value, roofline, cpu_baseline, the plain path's figures) plus a
human-readable summary on stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
METRIC = "resistance tensors/sec at ≤1e-6 rel err, batched mixed+learned vs plain FP64 GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_bodies(n_super, n_fibre):
    from andromeda_amd import mesh as M
    return [m for m, _ in M.batch_bodies(n_super, n_fibre)]


def _spread(idx, bodies, k):
    order = sorted(idx, key=lambda i: bodies[i].nelem)
    if k >= len(order):
        return order
    if k == 1:
        return [order[len(order) // 2]]
    return [order[int(round(j * (len(order) - 1) / (k - 1)))] for j in range(k)]


def stratified(bodies, k, n_fibre):
    """k bodies spread over the size range: a quarter of them fibres, the rest superellipsoids."""
    n_super = len(bodies) - n_fibre
    kf = min(n_fibre, k // 4)
    return sorted(_spread(range(n_super), bodies, k - kf) + _spread(range(n_super, len(bodies)), bodies, kf))


def traffic_ratio(kernel):
    """HBM bytes / algorithmic bytes of a kernel, measured with rocprofv3 FETCH_SIZE and WRITE_SIZE
    runs and recorded in profiles/*.md as 'traffic_ratio: <kernel> <ratio>' (None if absent)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_ratio:" and parts[1] == kernel:
                return float(parts[2])
    return None


def kernel_roofline(name, v):
    """Roofline entry of one profiled kernel: HBM-bound for the matrix products (algorithmic bytes
    over the mean launch time), VALU-bound otherwise (algorithmic flops)."""
    avg_ms = v["ms"] / v["launches"]
    if "gemv" in name:
        per_launch = v["bytes"] / v["launches"]
        ach = per_launch / (avg_ms * 1e-3) / 1e9
        ratio = traffic_ratio(name)
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": None if ratio is None else round(ratio * per_launch),
                "kernel": name, "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": round(per_launch)}
    ach = v["flops"] / v["launches"] / (avg_ms * 1e-3) / 1e12
    peak = FP32_PEAK_TFLOPS if "fp32" in name else FP64_PEAK_TFLOPS
    return {"bound": "valu", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
            "frac": round(ach / peak, 4), "traffic": None, "kernel": name, "avg_launch_ms": round(avg_ms, 4)}


def kernel_table(prof):
    """Per-kernel totals of a profiled region and the roofline of its dominant kernel."""
    kernels = {}
    for name, v in prof.items():
        sec = v["ms"] * 1e-3
        kernels[name] = {"launches": v["launches"], "ms_total": round(v["ms"], 3),
                         "GBps": round(v["bytes"] / sec / 1e9, 1) if sec > 0 else None,
                         "TFLOPs": round(v["flops"] / sec / 1e12, 2) if sec > 0 else None}
    if not prof:
        return kernels, None
    name, v = max(prof.items(), key=lambda kv: kv[1]["ms"])
    return kernels, kernel_roofline(name, v)


def cpu_baseline(bodies, sample_idx, plain_its):
    """The oracle (C, FP64: assembly on OpenMP threads, LSQR serial per right-hand side, as the
    reference runs it) on the stratified sample of the batch. Its two phases are timed on bounded
    pieces -- the assembly of 48 centroid rows of the sample's largest body (seconds per pair at
    the 126-point rule) and 8 LSQR iterations on a dense system of 4500 unknowns (seconds per
    iteration and matrix entry) -- and every sample body's time is extrapolated from its pair count,
    its 3F x 3F matrix and the LSQR iteration counts of the plain GPU path on the same body (the
    same algorithm). One body is run end to end to check the extrapolation."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    big = max(sample_idx, key=lambda i: bodies[i].nelem)
    mb = bodies[big]
    rows = np.arange(mb.nnodes, mb.nnodes + 48)
    pyoracle.assemble_stokes(mb, rows=rows[:4], extra=8 * np.pi)  # warm-up (thread pool)
    t0 = time.perf_counter()
    pyoracle.assemble_stokes(mb, rows=rows, extra=8 * np.pi)
    t_pair = (time.perf_counter() - t0) / (rows.size * mb.nelem)
    n = 4500
    rng = np.random.default_rng(0)
    A = rng.normal(size=(n, n)) + 8.0 * np.sqrt(n) * np.eye(n)
    ia, ja, a = pyoracle.csr_from_dense(A)
    q = pyoracle.lsqr_precon(ia, ja, a, n, n)
    bvec = rng.normal(size=n)
    t0 = time.perf_counter()
    pyoracle.lsqr_solve(ia, ja, a, n, n, q, bvec, maxit=8, eps=1e-30)
    t_entry = (time.perf_counter() - t0) / (8 * float(n) * n)
    del A, ia, ja, a

    def model(i, its):
        F = bodies[i].nelem
        return F * (F - 1) * t_pair + sum(its) * 9.0 * F * F * t_entry

    est = {i: model(i, plain_its[i]) for i in sample_idx}
    small = min(sample_idx, key=est.get)  # the cheapest body of the sample, run end to end
    t0 = time.perf_counter()
    pyoracle.resistance(bodies[small], exterior=True, eps=1e-12)
    t_small = time.perf_counter() - t0
    total = sum(est.values())
    return {"value": len(sample_idx) / total, "unit": "resistance tensors/s", "cores": threads, "kind": "port",
            "extrapolated": True,
            "seconds_per_body": {"min": round(min(est.values()), 2), "median": round(float(np.median(list(est.values()))), 2),
                                 "max": round(max(est.values()), 2)},
            "batch576_seconds_extrapolated": round(total * len(bodies) / len(sample_idx), 1),
            "phase_rates": {"assembly_s_per_pair_126pt": t_pair, "lsqr_s_per_iteration_entry": t_entry},
            "check": {"body": small, "F": bodies[small].nelem, "measured_s": round(t_small, 2),
                      "model_s": round(est[small], 2)},
            "sample": f"the {len(sample_idx)}-body stratified sample of the batch (F {min(bodies[i].nelem for i in sample_idx)}"
                      f"..{max(bodies[i].nelem for i in sample_idx)}), FP64 oracle, time per body extrapolated from "
                      f"measured phase rates (assembly: 48 rows of the {mb.nelem}-triangle body on {threads} OpenMP "
                      f"threads; LSQR: 8 serial iterations on a 4500^2 system) and the plain GPU path's LSQR "
                      f"iteration counts; the {bodies[small].nelem}-triangle body run end to end in "
                      f"{t_small:.1f} s (model {est[small]:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--bodies", type=int, default=576, help="total bodies (8/9 superellipsoids, 1/9 fibres)")
    ap.add_argument("--sample", type=int, default=96, help="bodies of the FP64-vs-fast comparison (stratified)")
    ap.add_argument("--precision", choices=["mixed", "learned", "auto"], default="auto")
    ap.add_argument("--model", default=os.path.join(ROOT, "andromeda_amd", "quad_model.npz"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ablation", action="store_true", help="skip the ablation steps (profiled runs)")
    ap.add_argument("--domain-only", type=int, default=0, metavar="NPTS",
                    help="only the domain evaluation: velocity and pressure at NPTS points inside a "
                         "20480-triangle sphere (the kernel's trace and roofline)")
    ap.add_argument("--plain-only", action="store_true",
                    help="only the plain FP64 path on the sample, one body at a time with per-kernel timing "
                         "(the workload of the plain path's kernel traces and counters)")
    ap.add_argument("--sphere20480", action="store_true",
                    help="only the single-system configuration: the plain FP64 path (assembly and LSQR of a "
                         "61440^2 system) on the 20480-triangle unit sphere, with its kernel roofline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 launches): the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 5 copies of {n * 4} bytes")
        return
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: runs on several GPUs are not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from andromeda_amd import api
    from andromeda_amd.metrics import entry_error, tensor_error

    if args.domain_only:
        from andromeda_amd import mesh as M
        m = M.icosphere(5)
        rng = np.random.default_rng(0)
        d = rng.normal(size=(args.domain_only, 3))
        pts = 0.8 * d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0, 1, size=(args.domain_only, 1)) ** (1 / 3)
        u, q = rng.normal(size=(m.nnodes, 3)), rng.normal(size=(m.nelem, 3))
        ctx = api.Context(0)
        ctx.stokes_domain_values(m, pts[:64], u, q)  # warm-up
        ctx.profile(True)
        t0 = time.perf_counter()
        ctx.stokes_domain_values(m, pts, u, q)
        el = time.perf_counter() - t0
        kernels, roofline = kernel_table(ctx.profile_read())
        ctx.profile(False)
        log(f"[bench] domain values: {args.domain_only} points x {m.nelem} triangles in {el:.3f} s "
            f"({args.domain_only / el:.0f} points/s); {kernels}")
        print(json.dumps({"domain_values": {"points": args.domain_only, "triangles": m.nelem,
                                            "points_per_s": round(args.domain_only / el, 1),
                                            "roofline": roofline, "kernels": kernels}}), flush=True)
        return

    if args.sphere20480:
        from andromeda_amd import mesh as M
        m = M.icosphere(5)
        ctx = api.Context(0)
        for w in range(args.warmup):
            ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12)
        ctx.profile(True)
        t0 = time.perf_counter()
        for s in range(args.steps):
            r = ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12)
        el = (time.perf_counter() - t0) / args.steps
        kernels, roofline = kernel_table(ctx.profile_read())
        ctx.profile(False)
        K, Om = np.diag(r.FT[:3, :3]), np.diag(r.FT[3:, 3:])
        drag = float(np.abs(K / (6 * np.pi) - 1).max())
        log(f"[bench] 20480-triangle sphere, plain FP64: {el:.3f} s per tensor (assembly {r.t_assemble:.3f} s, "
            f"LSQR {r.t_solve:.3f} s, its {r.nit.tolist()}); drag vs 6 pi: {drag:.2e}; {kernels}")
        print(json.dumps({"sphere_20480_plain_fp64": {
            "triangles": m.nelem, "unknowns": 3 * m.nelem, "system_bytes": 8 * (3 * m.nelem) ** 2,
            "tensors_per_s": round(1.0 / el, 4), "s_per_tensor": round(el, 3),
            "assembly_s": round(r.t_assemble, 3), "lsqr_s": round(r.t_solve, 3), "lsqr_its": r.nit.tolist(),
            "status": int(r.status), "drag_rel_err_vs_6pi": float(f"{drag:.3e}"),
            "torque_rel_err_vs_8pi": float(f"{np.abs(Om / (8 * np.pi) - 1).max():.3e}"),
            "roofline": roofline, "kernels": kernels}}), flush=True)
        return

    n_fibre = args.bodies // 9
    n_super = args.bodies - n_fibre
    t0 = time.perf_counter()
    bodies = make_bodies(n_super, n_fibre)
    log(f"[bench] {len(bodies)} bodies, {sum(b.nelem for b in bodies)} triangles, "
        f"sum F^2 = {sum(b.nelem ** 2 for b in bodies):.3e} (mesh generation {time.perf_counter() - t0:.1f} s)")

    ctx = api.Context(0)
    if args.plain_only:
        idx = stratified(bodies, args.sample, n_fibre)
        ctx.profile(True)
        tp = time.perf_counter()
        for i in idx:
            ctx.resistance_batch([bodies[i]], precision=api.PREC_FP64, eps=1e-12)
        t_plain = time.perf_counter() - tp
        kernels, roofline = kernel_table(ctx.profile_read())
        ctx.profile(False)
        log(f"[bench] plain FP64 only, bodies {idx} (F = {[bodies[i].nelem for i in idx]}): "
            f"{len(idx) / t_plain:.3f} tensors/s; autotuner {ctx.autotune_table()}")
        print(json.dumps({"plain_fp64_only": {"bodies": idx, "tensors_per_s": round(len(idx) / t_plain, 3),
                                              "roofline": roofline, "kernels": kernels}}), flush=True)
        return
    prec = api.PREC_MIXED
    prec_name = "mixed"
    if args.precision in ("learned", "auto") and os.path.exists(args.model):
        z = np.load(args.model)
        ctx.set_quadrature_model(z["params"], float(z["confidence"]), float(z["near_ratio"]), z["far_table"])
        prec, prec_name = api.PREC_LEARNED, "learned"
    elif args.precision == "learned":
        raise SystemExit(f"--precision learned needs {args.model}")

    # inputs resident in HBM before the timed region: the meshes are uploaded once
    tu = time.perf_counter()
    resident = ctx.upload_bodies(bodies)
    log(f"[bench] bodies uploaded: {time.perf_counter() - tu:.3f} s")
    for w in range(args.warmup):
        tw = time.perf_counter()
        ctx.resistance_batch_resident(resident, precision=prec)
        log(f"[bench] warmup {w}: {time.perf_counter() - tw:.2f} s")
    ctx.profile(True)
    t0 = time.perf_counter()  # the library calls return with their work finished
    results = None
    for s in range(args.steps):
        ts = time.perf_counter()
        results = ctx.resistance_batch_resident(resident, precision=prec)
        log(f"[bench] step {s}: {time.perf_counter() - ts:.2f} s (assembly "
            f"{sum(r.t_assemble for r in results):.2f} s, solve {sum(r.t_solve for r in results):.2f} s)")
    elapsed = time.perf_counter() - t0
    pairs = sum(b.nelem * (b.nelem - 1) for b in bodies)
    log(f"[bench] quadrature: {sum(r.quad_points for r in results) / pairs:.2f} points per pair over {pairs:.3e} "
        f"pairs; near pairs on a learned rule {sum(r.pairs_learned for r in results):.3e}, on the fixed rule "
        f"{sum(r.pairs_fallback for r in results):.3e}")
    worst = sorted(range(len(results)), key=lambda i: -int(results[i].nit.max()))[:4]
    log("[bench] most iterations: " + ", ".join(f"body {i} F={bodies[i].nelem} its={int(results[i].nit.max())}"
                                                for i in worst))
    prof = ctx.profile_read()
    ctx.profile(False)
    value = args.steps * len(bodies) / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"[bench] fast form ({prec_name}): {value:.2f} tensors/s, {ms_per_step:.0f} ms/step")
    # ablation on the same batch, one step each after a warm call: the fixed-order FP32 rule instead
    # of the learned rule order (mixed form), the learned form with FP64 Krylov products, and with
    # one Krylov space per right-hand side instead of block GMRES
    ablation = {"learned_step_ms": round(ms_per_step, 1)}
    if prec == api.PREC_LEARNED and not args.no_ablation:
        ctx.resistance_batch_resident(resident, precision=api.PREC_MIXED)
        ta = time.perf_counter()
        ctx.resistance_batch_resident(resident, precision=api.PREC_MIXED)
        ablation["mixed_step_ms"] = round(1e3 * (time.perf_counter() - ta), 1)
    if not args.no_ablation:
        os.environ["ANDR_GMRES_FP64_PRODUCTS"] = "1"
        try:
            ctx.resistance_batch_resident(resident, precision=prec)
            ta = time.perf_counter()
            ctx.resistance_batch_resident(resident, precision=prec)
            ablation[f"{prec_name}_fp64_krylov_products_step_ms"] = round(1e3 * (time.perf_counter() - ta), 1)
        finally:
            os.environ.pop("ANDR_GMRES_FP64_PRODUCTS", None)
        os.environ["ANDR_GMRES_INDEPENDENT"] = "1"
        try:
            ctx.resistance_batch_resident(resident, precision=prec)
            ta = time.perf_counter()
            ctx.resistance_batch_resident(resident, precision=prec)
            ablation[f"{prec_name}_independent_gmres_step_ms"] = round(1e3 * (time.perf_counter() - ta), 1)
        finally:
            os.environ.pop("ANDR_GMRES_INDEPENDENT", None)
    log(f"[bench] ablation: {ablation}")
    # the same step from host meshes (upload over PCIe inside the call), reported beside value
    th = time.perf_counter()
    ctx.resistance_batch(bodies, precision=prec)
    host_rate = len(bodies) / (time.perf_counter() - th)
    log(f"[bench] from host meshes (upload included): {host_rate:.2f} tensors/s")
    resident.free()

    # The accuracy condition of the metric, on the timed run's own tensors: the plain FP64 path on a
    # size-stratified sample (a quarter fibres), compared with the last timed step's results for the
    # same bodies; every tensor of the step is checked for the solver status.
    n_maxit = sum(1 for r in results if r.status == 5)
    n_floor = sum(1 for r in results if r.status == 4)
    idx = stratified(bodies, args.sample, n_fibre)
    sample = [bodies[i] for i in idx]
    sample_res = ctx.upload_bodies(sample)
    tp = time.perf_counter()
    plain = ctx.resistance_batch_resident(sample_res, precision=api.PREC_FP64, eps=1e-12)
    t_plain = time.perf_counter() - tp
    # the fast form on the same sample: one untimed call first (it re-fills the context's memory
    # cache, which still holds the plain path's FP64 systems), as the batch steps follow warmups
    ctx.resistance_batch_resident(sample_res, precision=prec)
    tf = time.perf_counter()
    fast_s = ctx.resistance_batch_resident(sample_res, precision=prec)
    t_fast = time.perf_counter() - tf
    sample_res.free()
    errs = [tensor_error(results[i].FT, p.FT) for i, p in zip(idx, plain)]
    entry_errs = [entry_error(results[i].FT, p.FT) for i, p in zip(idx, plain)]
    errs_alone = [tensor_error(f.FT, p.FT) for f, p in zip(fast_s, plain)]
    plain_rate = len(sample) / t_plain
    log(f"[bench] sample of {len(sample)}: plain FP64 {plain_rate:.2f} tensors/s, fast {len(sample) / t_fast:.2f}; "
        f"max err of the timed step {max(errs):.2e} (sample batch alone {max(errs_alone):.2e}); "
        f"plain LSQR its {[int(p.nit.max()) for p in plain]}")
    log("[bench] sample (body, F, fast its, err): " + ", ".join(
        f"({i}, {bodies[i].nelem}, {int(results[i].nit.max())}, {e:.1e})" for i, e in zip(idx, errs)))
    log(f"[bench] solver status of the timed step: {n_maxit} tensors at the iteration limit, {n_floor} accepted "
        f"at the FP32 floor")
    plain_status = sorted({int(p.status) for p in plain})
    log(f"[bench] plain FP64 LSQR status on the sample: {plain_status} (4: LSQR's condition limit, the reference "
        f"solver's own outcome on slender fibres)")
    accuracy_ok = max(errs) <= 1e-6 and n_maxit == 0 and all(p.status in (0, 4) for p in plain)

    kernels, roofline = kernel_table(prof)

    # kernel rooflines of the plain path on one large body of the sample (about 5000 triangles: a
    # system of ~2 GB that solves in about a second), profiled on its own
    big = max((b for b in sample if b.nelem <= 5200), key=lambda b: b.nelem,
              default=min(sample, key=lambda b: b.nelem))
    ctx.profile(True)
    ctx.resistance_batch([big], precision=api.PREC_FP64, eps=1e-12)
    plain_kernels, plain_roofline = kernel_table(ctx.profile_read())
    ctx.profile(False)
    log("[bench] plain FP64 kernels: " + ", ".join(
        f"{k} {v['GBps']} GB/s {v['TFLOPs']} TFLOP/s" for k, v in plain_kernels.items()))

    cpu = None if args.no_cpu else cpu_baseline(bodies, idx, {i: p.nit.tolist() for i, p in zip(idx, plain)})
    if cpu:
        log(f"[bench] cpu oracle: {cpu['value']:.4f} tensors/s ({cpu['sample']})")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "resistance tensors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "precision": {"quadrature": "f32" if prec_name == "mixed" else "f32, learned rule order",
                      "singular_integrals": "f64", "matrix_storage": "f32",
                      "krylov_product": "f32 matrix cores, f64 accumulation of 16-column blocks",
                      "krylov_basis_and_solve": "f64", "refinement_residual": "f64"},
        "data": "synthetic (seeded superellipsoids and fibres)",
        "config": {"workload": f"{len(bodies)} bodies ({n_super} superellipsoids + {n_fibre} fibres), "
                               f"F log-uniform [500, 8000], 6 rigid-body RHS, exterior Stokes",
                   "bodies": len(bodies), "triangles": int(sum(b.nelem for b in bodies)),
                   "form": prec_name},
        "plain_fp64": {"sample_bodies": len(sample), "tensors_per_s": round(plain_rate, 3),
                       "kernels_profiled_on": f"one body of {big.nelem} triangles",
                       "fast_tensors_per_s_same_sample": round(len(sample) / t_fast, 3),
                       "speedup_same_sample": round(t_plain / t_fast, 2),
                       "roofline": plain_roofline, "kernels": plain_kernels},
        "ablation": dict(ablation, plain_fp64_lsqr_tensors_per_s_on_sample=round(plain_rate, 3),
                         fast_tensors_per_s_on_sample=round(len(sample) / t_fast, 3),
                         note="steps of the whole batch: learned rule order (value), fixed-order FP32 rule "
                              "(mixed), FP64 instead of FP32 Krylov products, one Krylov space per column instead "
                              "of block GMRES; the plain FP64 LSQR path and the fast form on the stratified sample"),
        "host_meshes": {"tensors_per_s": round(host_rate, 3),
                        "note": "one step from host meshes, mesh upload over PCIe inside the call"},
        "accuracy": {"max_tensor_err_vs_fp64": float(f"{max(errs):.3e}"),
                     "max_entry_err_vs_fp64": float(f"{max(entry_errs):.3e}"),
                     "entry_err_of": "entries above 1e-3 of their block's norm", "bodies": len(sample),
                     "of": "the last timed step's tensors of a size-stratified sample, against the plain FP64 path",
                     "status_iteration_limit": n_maxit, "status_floor_accepted": n_floor,
                     "plain_lsqr_status": plain_status,
                     "within_1e-6": bool(accuracy_ok)},
        "roofline": roofline,
        "kernels": kernels,
        "autotuner": {"krylov_product_segments": ctx.autotune_table_krylov(),
                      "plain_gemv_n_rows_per_wave": ctx.autotune_table()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if not accuracy_ok:
        log("[bench] FAILED: a sampled tensor is more than 1e-6 from the plain FP64 path, or a solve stopped "
            "at its iteration limit")
        sys.exit(1)


if __name__ == "__main__":
    main()
