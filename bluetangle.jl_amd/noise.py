"""Kraus tables and noise-model types (src/noise.jl:1-131, src/struct.jl:78-347)."""
from __future__ import annotations

import math
from typing import List, Sequence, Union

import numpy as np

from .gates import _RX, _RY, _RZ, _P, _is_it_measurement, gate


def noise_model(model: str, p: float, two_qubit: bool = False) -> List[np.ndarray]:
    """noise_model (noise.jl:52-131), quirks kept: ``phase_flip`` yields Y
    because noise.jl:80 compares the String with a Symbol, and the 2-qubit
    model is the product [kron(Ki,Kj) for Ki in ops for Kj in ops]."""
    m = model.lower()
    I, X, Y, Z = gate["I"], gate["X"], gate["Y"], gate["Z"]
    if m == "amplitude_damping":
        ops = [np.array([[1, 0], [0, math.sqrt(1 - p)]], dtype=complex),
               np.array([[0, math.sqrt(p)], [0, 0]], dtype=complex)]
    elif m == "phase_damping":
        ops = [np.array([[1, 0], [0, math.sqrt(1 - p)]], dtype=complex),
               np.array([[0, 0], [0, math.sqrt(p)]], dtype=complex)]
    elif m in ("phase_flip", "bit_flip", "bit_phase_flip"):
        ops = [math.sqrt(1 - p) * I, math.sqrt(p) * (X if m == "bit_flip" else Y)]
    elif m == "depolarizing_amp":
        ops = [math.sqrt(1 - 3 * p / 4) * I, math.sqrt(p / 4) * X, math.sqrt(p / 4) * Y, math.sqrt(p / 4) * Z]
    elif m == "depolarizing":
        ops = [math.sqrt(1 - p) * I, math.sqrt(p / 3) * X, math.sqrt(p / 3) * Y, math.sqrt(p / 3) * Z]
    elif m == "rot_z":
        ops = [_RZ(p)]
    elif m == "rot_y":
        ops = [_RY(p)]
    elif m == "rot_x":
        ops = [_RX(p)]
    elif m == "rot_p":
        ops = [_P(p)]
    elif m == "rot_xyz":
        ops = [_RX(p) / math.sqrt(3), _RY(p) / math.sqrt(3), _RZ(p) / math.sqrt(3)]
    elif _is_it_measurement(m):
        ops = [gate["P0"].copy(), gate["P1"].copy()]
    else:
        raise ValueError("Unknown quantum error model")
    ops = [np.asarray(o, dtype=np.complex128) for o in ops]
    if not two_qubit:
        return ops
    return [np.kron(a, b) for a in ops for b in ops]


def noise_model1(model, p):
    return noise_model(model, p, two_qubit=False)


def noise_model2(model, p):
    return noise_model(model, p, two_qubit=True)


def kraus_to_choi(kraus):
    d = kraus[0].shape[0]
    choi = np.zeros((d * d, d * d), dtype=complex)
    for K in kraus:
        v = np.asarray(K).reshape(-1, order="F")
        choi += np.outer(v, v.conj())
    return choi


def _isapprox(a, b) -> bool:
    """Julia's isapprox for arrays: norm(a-b) <= sqrt(eps) * max(norm(a), norm(b))."""
    rtol = math.sqrt(np.finfo(float).eps)
    return float(np.linalg.norm(a - b)) <= rtol * max(float(np.linalg.norm(a)), float(np.linalg.norm(b)))


def is_valid_quantum_channel(kraus: Sequence[np.ndarray]) -> bool:
    """struct.jl:291-309: trace preserving and completely positive."""
    n = kraus[0].shape[0]
    s = sum(np.asarray(K).conj().T @ np.asarray(K) for K in kraus)
    tp = _isapprox(s, np.eye(n))
    choi = kraus_to_choi(kraus)
    ev = np.linalg.eigvalsh((choi + choi.conj().T) / 2)
    cp = _isapprox(choi, choi.conj().T) and bool(np.all(np.round(ev, 10) >= 0))
    return bool(tp and cp)


class QuantumChannel:
    """QuantumChannel(q, model, p) / QuantumChannel(model, kraus, p) /
    QuantumChannel(model, p) (struct.jl:90-140)."""

    def __init__(self, *args):
        if len(args) == 3 and isinstance(args[0], int):
            q, model, p = args
            kraus = noise_model1(model, p) if q == 1 else noise_model2(model, p)
        elif len(args) == 3:
            model, kraus, p = args
        elif len(args) == 2:
            model, p = args
            kraus = noise_model1(model, p)
        else:
            raise TypeError("QuantumChannel(q, model, p) | (model, kraus, p) | (model, p)")
        kraus = [np.asarray(k, dtype=np.complex128) for k in kraus]
        if not is_valid_quantum_channel(kraus):
            raise ValueError("not valid kraus operators: not CPTP!")
        q = int(round(math.log2(kraus[0].shape[0])))
        if q not in (1, 2):
            raise ValueError("Noise models are available only for 1 and 2 qubits!")
        self.q = q
        self.name = str(model).lower()
        self.p = float(p)
        self.kraus = kraus

    def __repr__(self):
        return "QuantumChannel(q=%d, %s, p=%g)" % (self.q, self.name, self.p)


def Noise1(model, p):
    return QuantumChannel(1, model, p)


def Noise2(model, p):
    return QuantumChannel(2, model, p)


def custom_noise(q, name, kraus):
    if not is_valid_quantum_channel(kraus):
        raise ValueError("define valid kraus operators")
    return QuantumChannel(name, kraus, 0.0)


ChannelOrFalse = Union[QuantumChannel, bool]


class NoiseModel:
    """NoiseModel(n1, n2) and its shorthands (struct.jl:147-173); a channel
    slot may be ``False`` as in the reference."""

    def __init__(self, *args):
        if len(args) == 2 and isinstance(args[0], str):
            model, p = args
            n1, n2 = Noise1(model, p), Noise2(model, p)
        elif len(args) == 4:
            n1, n2 = Noise1(args[0], args[1]), Noise2(args[2], args[3])
        elif len(args) == 2:
            a, b = args
            n1 = Noise1(a[0], a[1]) if isinstance(a, (list, tuple)) else a
            n2 = Noise2(b[0], b[1]) if isinstance(b, (list, tuple)) else b
        else:
            raise TypeError("NoiseModel(n1, n2) | (model, p) | (m1, p1, m2, p2)")
        if isinstance(n1, QuantumChannel) and isinstance(n2, QuantumChannel):
            if n1.q != 1 or n2.q != 2:
                raise ValueError("size error in noise model (1|2)")
        elif isinstance(n1, QuantumChannel) and n1.q != 1:
            raise ValueError("size error in noise model (1)")
        elif isinstance(n2, QuantumChannel) and n2.q != 2:
            raise ValueError("size error in noise model (2)")
        self.q1 = n1
        self.q2 = n2

    def __repr__(self):
        return "NoiseModel(%r, %r)" % (self.q1, self.q2)
