"""Operator types: Op, OpQC, ifOp, OpF (src/struct.jl:6,182-249,365-453,666-744)."""
from __future__ import annotations

import math
from typing import Callable, List, Sequence, Union

import numpy as np

from .gates import _is_it_measurement, _name_with_arg_bool, gates, MEASUREMENTS
from .noise import is_valid_quantum_channel, noise_model1, noise_model2


class QuantumOps:
    """Abstract operation (struct.jl:6)."""

    def __mul__(self, state):
        # Op * state == apply(state, op) (all.jl:41)
        from .core import apply
        return apply(state, self)


def _get_op_num_qubits(qubit, target_qubit, control):
    """struct.jl:455-474"""
    if target_qubit == -1:
        if qubit == control:
            raise ValueError("`qubit` must differ from `control` qubit")
        return 1
    if qubit == target_qubit:
        raise ValueError("`qubit` and `target_qubit` must differ")
    if qubit == control and target_qubit == control:
        raise ValueError("either `qubit` or `target_qubit` must differ from `control` qubit")
    return 2


class Op(QuantumOps):
    """Op(name, qubit[, target]; noisy, control), Op(name, mat|f, qubit[, target]),
    Op(name, qubit, control, target) for CCZ/CCX/CCY/CSWAP, Op("RES", q) -> OpQC
    (struct.jl:365-453)."""

    def __new__(cls, name, *args, **kw):
        if len(args) == 1 and isinstance(args[0], (int, np.integer)) and name.upper() in ("RES", "RESET"):
            return OpQC("RES", [np.array([[1.0, 0], [0, 0]], dtype=complex),
                                np.array([[0, 1.0], [0, 0]], dtype=complex)], int(args[0]))
        return super().__new__(cls)

    def __init__(self, name, *args, type=None, noisy=True, control=-2):
        mat = None
        if len(args) >= 1 and (callable(args[0]) or isinstance(args[0], np.ndarray) or
                               isinstance(args[0], (list, tuple))):
            mat = args[0]
            args = args[1:]
        args = [int(a) for a in args]
        if len(args) == 1:
            qubit, target = args[0], -1
        elif len(args) == 2:
            qubit, target = args
        elif len(args) == 3 and mat is None:
            # three-qubit constructor (struct.jl:436-449)
            qubit, control_q, target = args
            base = {"CCZ": "CZ", "CCX": "CX", "CCY": "CY", "CSWAP": "SWAP"}.get(name)
            if base is None:
                raise ValueError("Unsupported three-qubit operation")
            name, control = base, control_q
        else:
            raise TypeError("Op(name, qubit[, target]) | Op(name, mat, qubit[, target])")
        self.name = name
        self.qubit = qubit
        self.target_qubit = target
        self.control = int(control)
        self.noisy = bool(noisy)
        self.q = _get_op_num_qubits(qubit, target, self.control)
        if mat is not None and callable(mat) and not isinstance(mat, np.ndarray):
            if type is not None:
                raise ValueError("setting `type` for gates constructed from functions is not supported")
            self.mat = mat
            self.type = "f"
            return
        m = gates(name) if mat is None else np.asarray(mat, dtype=np.complex128)
        if m.shape != (2 ** self.q, 2 ** self.q):
            raise ValueError("size of matrix %s not compatible with %d-qubit operation" % (m.shape, self.q))
        self.mat = m
        if self.q == 1 and _is_it_measurement(name):
            if self.control != -2:
                raise ValueError("measurement and control operations are incompatible.")
            self.type = "🔬"
            self.noisy = False
        else:
            self.type = "phase" if _name_with_arg_bool(name) else "op"

    @property
    def is_measurement(self):
        return self.type == "🔬"

    def __repr__(self):
        t = "" if self.target_qubit == -1 else ", %d" % self.target_qubit
        c = "" if self.control == -2 else "; control=%d" % self.control
        return 'Op("%s", %d%s%s)' % (self.name, self.qubit, t, c)


class OpQC(QuantumOps):
    """Circuit-placed quantum channel (struct.jl:182-249).  noisy == False, so
    no noise model channel follows it."""

    def __init__(self, name, *args):
        if len(args) < 2:
            raise TypeError("OpQC(model, p, qubit[, target]) | OpQC(name, kraus, qubit[, target])")
        if not isinstance(args[0], (list, tuple)):
            # OpQC(model, p, qubit, target=-1)
            p, qubit = float(args[0]), int(args[1])
            target = int(args[2]) if len(args) > 2 else -1
            q = 2 if target > 0 else 1
            kraus = noise_model1(name, p) if q == 1 else noise_model2(name, p)
            self._init(name, kraus, qubit, target, type=repr(p))
            self.p = p
            return
        kraus = list(args[0])
        qubit = int(args[1])
        target = int(args[2]) if len(args) > 2 else -1
        self._init(name, kraus, qubit, target, type="")
        self.p = None

    def _init(self, name, kraus, qubit, target, type=""):
        kraus = [np.asarray(k, dtype=np.complex128) for k in kraus]
        if not is_valid_quantum_channel(kraus):
            raise ValueError("not valid kraus operators: not CPTP!")
        q = int(round(math.log2(kraus[0].shape[0])))
        if q == 1 and target > 0:
            raise ValueError("for 1-qubit quantum channel, target_qubit should be -1")
        if q == 2 and target < 0:
            raise ValueError("for 2-qubit quantum channel, you must select the target_qubit")
        if q > 3:
            raise ValueError("Noise models are available only for up to 3 qubits!")
        self.q = q
        self.name = name.lower()
        self.qubit = qubit
        self.target_qubit = target if q == 2 else -1
        self.control = -2
        self.kraus = kraus
        self.noisy = False
        self.type = type

    def __repr__(self):
        return 'OpQC("%s", q=%d, %d, %d)' % (self.name, self.q, self.qubit, self.target_qubit)


class ifOp(QuantumOps):
    """Mid-circuit measurement with outcome-dependent op lists (struct.jl:666-696)."""

    def __init__(self, name, qubit, if0="I", if1="I"):
        if not _is_it_measurement(name):
            raise ValueError("select MX or MY or MZ or MR basis.")
        if isinstance(if0, str):
            if0 = [Op(if0, qubit)]
        if isinstance(if1, str):
            if1 = [Op(if1, qubit)]
        if isinstance(if0, Op):
            if0 = [if0]
        if isinstance(if1, Op):
            if1 = [if1]
        self.q = 1
        self.name = name
        self.mat = gates(name)
        self.qubit = int(qubit)
        self.target_qubit = -1
        self.control = -2
        self.if01 = (list(if0), list(if1))
        self.type = "🔬"

    def __repr__(self):
        return 'ifOp("%s", %d)' % (self.name, self.qubit)


class OpF(QuantumOps):
    """f(state) -> state (struct.jl:703-744): a function, a full-space matrix
    or an op list applied without noise."""

    def __init__(self, name, data):
        self.q = 1
        self.name = name
        self.data = data
        self.qubit = -1
        self.target_qubit = -1
        self.control = -2


def measurement_basis(name: str) -> int:
    return MEASUREMENTS[name.upper()]
