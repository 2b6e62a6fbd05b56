"""bench.py's contract pieces that need no GPU: more than one GPU is refused with a JSON message,
the roofline entry follows from the profiled totals (algorithmic bytes over the mean launch time,
HBM traffic from the ratio the committed profile records), and the accuracy sample is spread over
the size range with a quarter of it fibres."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_roofline_entry_from_profiled_totals():
    # 10 launches of 4.2e10 algorithmic bytes in 65 ms: 6.46 TB/s of the 8 TB/s peak
    v = {"launches": 10, "ms": 65.0, "bytes": 4.2e11, "flops": 2.9e12}
    r = bench.kernel_roofline("kb_gemv_sym", v)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(4.2e10 / 6.5e-3 / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, rel=1e-3)
    ratio = bench.traffic_ratio("kb_gemv_sym")  # from profiles/gmres_gemv_sym__batch576.md
    assert ratio is not None and 1.0 <= ratio < 1.1
    assert r["traffic"] == round(ratio * 4.2e10)
    # a VALU-bound kernel: algorithmic flops over the mean launch time against the FP32 peak
    a = bench.kernel_roofline("stokes_system_fp32", {"launches": 1, "ms": 150.0, "bytes": 1.8e11, "flops": 4.35e12})
    assert a["bound"] == "valu" and a["achieved"] == pytest.approx(29.0, rel=1e-3) and a["peak"] == 157.3
    # the dominant kernel is the one with the most time
    kernels, dom = bench.kernel_table({"kb_gemv_sym": v, "stokes_system_fp32": {"launches": 1, "ms": 50.0,
                                                                                "bytes": 1.8e11, "flops": 1.45e12}})
    assert dom["kernel"] == "kb_gemv_sym" and set(kernels) == {"kb_gemv_sym", "stokes_system_fp32"}
    assert kernels["kb_gemv_sym"]["GBps"] == pytest.approx(6461.5, rel=1e-3)


def test_stratified_sample():
    from andromeda_amd import mesh as M
    bodies = [m for m, _ in M.batch_bodies(64, 8)]
    idx = bench.stratified(bodies, 16, 8)
    assert len(idx) == 16 == len(set(idx))
    assert sum(1 for i in idx if i >= 64) == 4  # a quarter fibres
    sizes = [bodies[i].nelem for i in idx if i < 64]
    all_sizes = [b.nelem for b in bodies[:64]]
    assert min(sizes) == min(all_sizes) and max(sizes) == max(all_sizes)
