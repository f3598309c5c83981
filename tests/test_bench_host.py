"""bench.py host-side contract (no GPU): the multi-GPU request is refused with
a JSON message and a non-zero status, the argument parser accepts the
documented flags, and the byte / flop models used for the roofline are
consistent with their definitions."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_argument_parser_accepts_documented_flags():
    sys.path.insert(0, ROOT)
    import bench
    old = sys.argv
    try:
        sys.argv = ["bench.py", "--steps", "5", "--warmup", "1", "--pe-only", "gauss", "--no-cpu"]
        a = bench._args()
        assert a.steps == 5 and a.warmup == 1 and a.pe_only == "gauss" and a.no_cpu
        sys.argv = ["bench.py", "--pe-only"]
        assert bench._args().pe_only == "all"
        sys.argv = ["bench.py"]
        a = bench._args()
        assert a.gpus == 1 and a.pe_only is None and a.precision is None
    finally:
        sys.argv = old


def test_pe_state_generator_is_band_limited_rest_plus_perturbation():
    import math
    import numpy as np
    sys.path.insert(0, ROOT)
    import bench
    y = bench._pe_state(21, 4, 7)
    assert y.shape == (13, 22, 22)
    cut = (2 * 21) // 3
    assert not y[:, cut + 1:, :].any()
    assert not np.triu(y[0], 1).any()
    assert y[0, 0, 0] == 0.0 and y[4, 0, 0] == 0.0
    assert abs(y[8, 0, 0].real / math.sqrt(4 * math.pi) - 260.0) < 5.0
    assert np.all(y[:, :, 0].imag == 0.0)
