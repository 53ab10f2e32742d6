"""Seeded circuit generators for the benchmark configurations.

Brickwork: per layer RX(t1), RY(t2), RZ(t3) on every qubit (angles U[0, 2pi)
from a numpy Generator seeded with `seed`), then Op("CNOT", i, i+1) on the
even-start pairs (1,2),(3,4),... in even layers and the odd-start pairs
(2,3),(4,5),... in odd layers, then optionally OpQC("amplitude_damping",
gamma, q) on every qubit, then Op("MZ", q) on the qubits of a mask drawn with
probability `mz_prob` per qubit from the same generator.  Gate noise is a
NoiseModel(Noise1("depolarizing", p1), Noise2("depolarizing", p2)).
"""
from __future__ import annotations

import math

import numpy as np

from .noise import NoiseModel
from .ops import Op, OpQC


def brickwork_ops(N, layers, seed, ad_gamma=None, mz_prob=0.0):
    rng = np.random.default_rng(seed)
    ops = []
    for layer in range(layers):
        for q in range(1, N + 1):
            th = rng.uniform(0, 2 * math.pi, 3)
            ops.append(Op("RX(%r)" % float(th[0]), q))
            ops.append(Op("RY(%r)" % float(th[1]), q))
            ops.append(Op("RZ(%r)" % float(th[2]), q))
        start = 1 if layer % 2 == 0 else 2
        for i in range(start, N, 2):
            ops.append(Op("CNOT", i, i + 1))
        if ad_gamma is not None:
            for q in range(1, N + 1):
                ops.append(OpQC("amplitude_damping", ad_gamma, q))
        if mz_prob > 0:
            mask = rng.random(N) < mz_prob
            for q in range(1, N + 1):
                if mask[q - 1]:
                    ops.append(Op("MZ", q))
    return ops


def depolarizing_noise(p1=1e-3, p2=1e-2):
    return NoiseModel("depolarizing", p1, "depolarizing", p2)


# The configurations (sizes as stated for each workload)
def config_c1():
    """Plain statevector, N=12, 20 brickwork layers, final probabilities."""
    return dict(N=12, ops=brickwork_ops(12, 20, 2026), noise=False)


def config_c2():
    """Density matrix N=8, 16 layers, depolarizing 1e-3/1e-2 + amplitude damping 5e-3."""
    return dict(N=8, ops=brickwork_ops(8, 16, 2026, ad_gamma=5e-3), noise=depolarizing_noise())


def config_c3(N=22, layers=32):
    """Batched trajectories: N=22 brickwork, 32 layers, same noise, MZ with p=0.1 per qubit per layer."""
    return dict(N=N, ops=brickwork_ops(N, layers, 2026, ad_gamma=5e-3, mz_prob=0.1), noise=depolarizing_noise())
