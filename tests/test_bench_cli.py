"""bench.py's command-line contract on the CPU: a request for more than one
GPU ends with one JSON line saying this build measures one GPU, exit code 2,
before any device is touched."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_bench_refuses_hw_queues_beyond_32():
    """--hw-queues sets GPU_MAX_HW_QUEUES before HIP initialises; values
    outside [1, 32] end the run with a message before any device is touched."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--hw-queues", "40"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "--hw-queues must be in [1, 32]" in r.stderr


def _bench():
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_psi_gate_reports_dlogpsi_sign_and_zeros():
    b = _bench()
    g = b.psi_gate([1.0, -2.0, 0.0], [1.0, -2.0, 0.0])
    assert g["gate_met"] and g["max_abs_dlogpsi"] == 0.0 and g["nonzero"] == 2
    g = b.psi_gate([1.0, -2.0 * (1 + 1e-5)], [1.0, -2.0])
    assert not g["gate_met"] and abs(g["max_abs_dlogpsi"] - 1e-5) < 1e-9 and g["sign_agreement"] == 1.0
    g = b.psi_gate([1.0, 2.0], [1.0, -2.0])
    assert not g["gate_met"] and g["sign_agreement"] == 0.5
    g = b.psi_gate([1.0, 0.0], [1.0, 1e-30])
    assert not g["gate_met"] and not g["zeros_agree"]


def test_host_cores_bounded_by_nproc():
    usable, nproc = _bench().host_cores()
    assert 1 <= usable <= nproc
