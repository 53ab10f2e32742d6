"""Writes the small golden fixtures of this directory with the FP64 oracle
(oracle/): inputs and expected outputs only.  tests/test_golden.py checks
them against an independent dense NumPy brute force (oracle/dense.py) on
the CPU and against the engine on the GPU.

    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from refcirc import brickwork_spec, random_spec, to_ref  # noqa: E402

NOISE = ("depolarizing", 1e-3, "depolarizing", 1e-2)


def main():
    # C1: N=12 brickwork, 20 layers, seed 2026, final probabilities (complex128)
    spec = brickwork_spec(12, 20, 2026, gate_noise=False)
    s, _, _, _, _ = O.run_sv(12, to_ref(spec), None, None)
    np.save(os.path.join(HERE, "c1_probs_n12.npy"), np.abs(s) ** 2)
    # C2-shaped density matrix at N=5: 8 layers, gate noise + amplitude damping
    spec = brickwork_spec(5, 8, 2026, ad_gamma=5e-3)
    rho = O.run_dm(5, to_ref(spec), O.ref_noise(*NOISE))
    np.save(os.path.join(HERE, "c2_rho_n5.npy"), rho)
    # trajectories: N=6 random noisy circuits with fixed uniforms u(seed, 0, .)
    cases = []
    kinds = ("gate1", "gate2", "ccx", "chan1", "chan2", "meas", "reset", "ifop")
    for seed in range(6):
        rng = np.random.default_rng(4400 + seed)
        spec = random_spec(6, 30, rng, kinds=kinds)
        nm = ("depolarizing", 0.05, "amplitude_damping", 0.1)
        u = O.uniforms(seed, 0, 16 * len(spec) + 16)
        st, outs, used, bnd, _ = O.run_sv(6, to_ref(spec), O.ref_noise(*nm), u, tol=1e-6)
        if bnd:
            continue
        cases.append({"seed": seed, "spec": spec, "noise": nm, "outcomes": outs, "draws": int(used),
                      "state_re": st.real.tolist(), "state_im": st.imag.tolist()})
    with open(os.path.join(HERE, "trajectories_n6.json"), "w") as f:
        json.dump(cases, f)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
