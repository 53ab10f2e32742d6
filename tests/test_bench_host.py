"""Host-side parts of bench.py that need no GPU: the one-GPU contract and the
per-op schedule's byte count that `config.unfused_equivalent` reports."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_before_any_device_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_unfused_bytes_count_every_op_kind(bt):
    """2S per gate, 3S per circuit-placed channel and measurement, and 2S times
    the weight of the non-identity branches of each noisy gate's channel,
    here summed as sum_{k >= 1} ||K_k||_F^2 / d (the channel's own Kraus
    weights) instead of the first operator's diagonal that bench.py reads."""
    import bench
    from bluetangle_jl_amd import workloads as W
    cfg = W.config_c3(6, 4)  # holds gates, damping channels and two measurements
    S = 2 ** 6 * 8.0
    nm = cfg["noise"]

    def weight(ch):
        d = ch.kraus[0].shape[0]
        return sum(float(np.linalg.norm(k) ** 2) for k in ch.kraus[1:]) / d

    expect = 0.0
    kinds = {"gate": 0, "channel": 0, "measure": 0}
    for o in cfg["ops"]:
        if type(o).__name__ == "OpQC":
            kinds["channel"] += 1
            expect += 3 * S
        elif o.is_measurement:
            kinds["measure"] += 1
            expect += 3 * S
        else:
            kinds["gate"] += 1
            expect += 2 * S
            if o.noisy:
                expect += 2 * S * weight(nm.q1 if (o.q == 1 and o.control == -2) else nm.q2)
    assert kinds["gate"] > 0 and kinds["channel"] > 0 and kinds["measure"] > 0, kinds
    assert abs(bench.unfused_bytes_per_trajectory(cfg) - expect) <= 1e-9 * expect
