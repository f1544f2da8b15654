"""bench.py command-line contract on the CPU: one GPU only, a JSON line and
exit status 2 for any other GPU count."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_help_lists_the_modes():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--serve", "--otf", "--scf", "--peaks", "--measure-traffic",
                 "--profile-precision", "--scf-leg", "--serve-leg-requests"):
        assert flag in p.stdout


def test_traffic_table_is_used_only_for_the_kernel_sources_it_was_measured_on(tmp_path, monkeypatch):
    sys.path.insert(0, ROOT)
    import bench

    kern = {"f2jk::kx_bf16x3_kernel<10, 4, 8>": {"read_bytes": 1.5e10, "write_bytes": 3.0e9}}
    tab = {"kernel_sources_sha256": bench.kernel_sources_sha256(), "fetch_factor": 2.0, "write_factor": 1.0,
           "copy_bytes": 4 << 30, "entries": {"64:bf16x3": {"kernels": kern}}}
    path = tmp_path / "traffic.json"
    path.write_text(json.dumps(tab))
    monkeypatch.setattr(bench, "TRAFFIC_TABLE", str(path))
    t, note = bench.measured_traffic(64, "bf16x3", "f2jk::kx_bf16x3_kernel")
    assert t == 1.8e10 and "FETCH_SIZE x2" in note
    assert bench.measured_traffic(16, "bf16x3", "f2jk::kx_bf16x3_kernel")[0] is None
    tab["kernel_sources_sha256"] = "0" * 64  # measured on other kernel sources: not reported
    path.write_text(json.dumps(tab))
    t, note = bench.measured_traffic(64, "bf16x3", "f2jk::kx_bf16x3_kernel")
    assert t is None and "other kernel sources" in note
