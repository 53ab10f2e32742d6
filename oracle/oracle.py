"""FP64 CPU oracle for BlueTangle.jl's apply path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  It is the checker, never the product:
the engine in ``bluetangle.jl_amd`` does not import it.

Layers:
  * ``libbt_oracle.so`` (bt_oracle.c): complex128 restatement of
    apply / apply_noise / born_measure_Z / _reset_Z / __ifOp_apply /
    __QuantumChannel_new_apply / to_rho, in a fast in-place mode and in the
    reference's own sparse full-space-operator mode.
  * this file: ctypes binding, the reference's gate table
    (src/gates.jl:22-59,369-510) and Kraus tables (src/noise.jl:52-131)
    restated in numpy, and a reference-level op description (``RefOp``) that
    mirrors the fields of Op / OpQC / ifOp (src/struct.jl:182-249,365-453,
    666-696).
"""
from __future__ import annotations

import ctypes
import math
import os
import re
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KIND_GATE, KIND_CHANNEL, KIND_MEASURE, KIND_RESET, KIND_IFOP = 0, 1, 2, 3, 4
MZ, MX, MY, MR = 0, 1, 2, 3


def build():
    """Compile libbt_oracle.so with make (gcc only)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libbt_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.bto_run_sv.argtypes = [ctypes.c_int, dp, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, dp,
                                 ctypes.c_int64, ctypes.c_double, ctypes.c_int, ip, ctypes.c_int64, i64p, i64p,
                                 i64p, i64p]
        L.bto_run_sv.restype = ctypes.c_int
        L.bto_run_traj.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                   ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_int, ip,
                                   ctypes.c_int64, i64p, dp, i64p]
        L.bto_run_traj.restype = ctypes.c_int
        L.bto_run_dm.argtypes = [ctypes.c_int, dp, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.bto_run_dm.restype = ctypes.c_int
        L.bto_last_error.restype = ctypes.c_char_p
        L.bto_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.bto_uniform.restype = ctypes.c_double
        L.bto_uniforms.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, dp]
        L.bto_full_operator.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, dp]
        L.bto_apply_gate.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.bto_partial_trace.argtypes = [ctypes.c_int, dp, ctypes.c_int, ip, dp]
        L.bto_calc_prob.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp]
        L.bto_weighted_sample.argtypes = [dp, ctypes.c_int, ctypes.c_double]
        L.bto_weighted_sample.restype = ctypes.c_int
        L.bto_num_threads.restype = ctypes.c_int
        L.bto_set_num_threads.argtypes = [ctypes.c_int]
        L.bto_set_num_threads.restype = None
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle error %d: %s" % (rc, lib().bto_last_error().decode()))


# ---------------------------------------------------------------------------
# Gate table (src/gates.jl:22-59) and parametrised gates (gates.jl:369-510)
# ---------------------------------------------------------------------------
_s2 = 1.0 / math.sqrt(2.0)
GATE = {
    "I": np.array([[1, 0], [0, 1]], dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "SX": 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]),
    "XSQRT": 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]),
    "Y": np.array([[0, -1j], [1j, 0]]),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
    "H": _s2 * np.array([[1, 1], [1, -1]], dtype=complex),
    "S": np.array([[1, 0], [0, 1j]]),
    "SD": np.array([[1, 0], [0, -1j]]),
    "T": np.array([[1, 0], [0, complex(round(math.cos(math.pi / 4), 10), round(math.sin(math.pi / 4), 10))]]),
    "TD": np.array([[1, 0], [0, complex(round(math.cos(math.pi / 4), 10), -round(math.sin(math.pi / 4), 10))]]),
    "HSP": _s2 * np.array([[1, -1j], [1, 1j]]),
    "HY": _s2 * np.array([[1, 1j], [1, -1j]]),
    "P0": np.array([[1, 0], [0, 0]], dtype=complex),
    "P1": np.array([[0, 0], [0, 1]], dtype=complex),
    "SP": np.array([[0, 1], [0, 0]], dtype=complex),
    "SM": np.array([[0, 0], [1, 0]], dtype=complex),
    "CX": np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex),
    "CNOT": np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex),
    "CY": np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]]),
    "CZ": np.diag([1, 1, 1, -1]).astype(complex),
    "CS": np.diag([1, 1, 1, 1j]),
    "CSD": np.diag([1, 1, 1, -1j]),
    "CI": np.eye(4, dtype=complex),
    "CH": np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, _s2, _s2], [0, 0, _s2, -_s2]], dtype=complex),
    "SWAP": np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex),
    "ISWAP": np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
    "FSWAP": np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, -1]], dtype=complex),
}
_t = complex(round(math.cos(math.pi / 4), 10), round(math.sin(math.pi / 4), 10))
GATE["CT"] = np.diag([1, 1, 1, _t])
GATE["CTD"] = np.diag([1, 1, 1, _t.conjugate()])
GATE["SYC"] = np.array([[1, 0, 0, 0], [0, 0, -1j, 0], [0, -1j, 0, 0], [0, 0, 0, np.exp(-1j * math.pi / 6)]])
GATE["ECR"] = _s2 * np.array([[0, 1, 0, 1j], [1, 0, -1j, 0], [0, 1j, 0, 1], [-1j, 0, 1, 0]])


def _rx(t):
    return np.array([[math.cos(t / 2), -1j * math.sin(t / 2)], [-1j * math.sin(t / 2), math.cos(t / 2)]])


def _ry(t):
    return np.array([[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]], dtype=complex)


def _rz(t):
    return np.array([[np.exp(-1j * t / 2), 0], [0, np.exp(1j * t / 2)]])


def _p(l):
    return np.array([[1, 0], [0, np.exp(1j * l)]])


PARAM = {
    "P": _p, "U1": _p, "RX": _rx, "RY": _ry, "RZ": _rz,
    "U2": lambda f, l: _s2 * np.array([[1, -np.exp(1j * l)], [np.exp(1j * f), np.exp(1j * (f + l))]]),
    "U3": lambda t, f, l: np.array([[math.cos(t / 2), -np.exp(1j * l) * math.sin(t / 2)],
                                    [np.exp(1j * f) * math.sin(t / 2), np.exp(1j * (f + l)) * math.cos(t / 2)]]),
    "CP": lambda l: np.diag([1, 1, 1, np.exp(1j * l)]),
    "GIVENS": lambda t: np.array([[1, 0, 0, 0], [0, math.cos(t), -math.sin(t), 0], [0, math.sin(t), math.cos(t), 0],
                                  [0, 0, 0, 1]], dtype=complex),
    "FSIM": lambda t, f: np.array([[1, 0, 0, 0], [0, math.cos(t), -1j * math.sin(t), 0],
                                   [0, -1j * math.sin(t), math.cos(t), 0], [0, 0, 0, np.exp(1j * f)]]),
    "SWAPA": lambda a: 0.5 * np.array([[2, 0, 0, 0], [0, 1 + np.exp(1j * math.pi * a), 1 - np.exp(1j * math.pi * a), 0],
                                       [0, 1 - np.exp(1j * math.pi * a), 1 + np.exp(1j * math.pi * a), 0],
                                       [0, 0, 0, 2]]),
    "RXX": lambda f: np.array([[math.cos(f), 0, 0, -1j * math.sin(f)], [0, math.cos(f), -1j * math.sin(f), 0],
                               [0, -1j * math.sin(f), math.cos(f), 0], [-1j * math.sin(f), 0, 0, math.cos(f)]]),
    "RYY": lambda f: np.array([[math.cos(f), 0, 0, 1j * math.sin(f)], [0, math.cos(f), -1j * math.sin(f), 0],
                               [0, -1j * math.sin(f), math.cos(f), 0], [1j * math.sin(f), 0, 0, math.cos(f)]]),
    "RZZ": lambda f: np.diag([np.exp(-1j * f / 2), np.exp(1j * f / 2), np.exp(1j * f / 2), np.exp(-1j * f / 2)]),
    "RXY": lambda f: np.array([[1, 0, 0, 0], [0, math.cos(f), -1j * math.sin(f), 0],
                               [0, -1j * math.sin(f), math.cos(f), 0], [0, 0, 0, 1]]),
}

MEAS_NAMES = {"MZ": MZ, "M(Z)": MZ, "MX": MX, "M(X)": MX, "MY": MY, "M(Y)": MY, "MR": MR, "M(R)": MR}


def _arith(text: str) -> float:
    """Evaluate a gate argument such as "0.3", ".1π", "pi/2" (the reference
    uses eval(Meta.parse(...)), gates.jl:348); only numbers, pi and + - * /."""
    import ast
    import operator as opm
    t = text.strip().replace("π", "*pi").replace("**pi", "*pi")
    if t.startswith("*pi"):
        t = "pi" + t[3:]
    t = re.sub(r"(?<![0-9a-z_.])\*pi", "pi", t)
    binops = {ast.Add: opm.add, ast.Sub: opm.sub, ast.Mult: opm.mul, ast.Div: opm.truediv}

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.Name) and n.id == "pi":
            return math.pi
        if isinstance(n, ast.BinOp) and type(n.op) in binops:
            return binops[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        raise ValueError("unsupported gate argument: %r" % text)

    return ev(ast.parse(t, mode="eval"))


def gate_matrix(name: str) -> np.ndarray:
    """gates(name) (gates.jl:369-453)."""
    up = name.upper()
    clean = up.split("(")[0]
    if clean in PARAM and "(" in name:
        args = [_arith(a) for a in re.sub(r"[()]", "", name.split("(", 1)[1]).split(",")]
        return np.asarray(PARAM[clean](*args), dtype=complex)
    if up in ("MZ", "M(Z)", "MR", "M(R)", "RES"):
        return GATE["I"].copy()
    if up in ("MX", "M(X)"):
        return GATE["H"].copy()
    if up in ("MY", "M(Y)"):
        return GATE["HSP"].copy()
    if up in GATE:
        return GATE[up].copy()
    raise KeyError("Gate %s not found" % name)


def noise_kraus(model: str, p: float, two_qubit: bool = False) -> List[np.ndarray]:
    """noise_model (noise.jl:52-131), including its quirks: `phase_flip`
    yields Y (noise.jl:80 compares a String with a Symbol) and the 2-qubit
    model is the product [kron(Ki,Kj) for Ki in ops for Kj in ops]
    (noise.jl:126)."""
    m = model.lower()
    I, X, Y, Z = GATE["I"], GATE["X"], GATE["Y"], GATE["Z"]
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
        ops = [_rz(p)]
    elif m == "rot_y":
        ops = [_ry(p)]
    elif m == "rot_x":
        ops = [_rx(p)]
    elif m == "rot_p":
        ops = [_p(p)]
    elif m == "rot_xyz":
        ops = [_rx(p) / math.sqrt(3), _ry(p) / math.sqrt(3), _rz(p) / math.sqrt(3)]
    elif m.upper() in MEAS_NAMES:
        ops = [GATE["P0"].copy(), GATE["P1"].copy()]
    else:
        raise ValueError("Unknown quantum error model")
    if not two_qubit:
        return [np.asarray(o, dtype=complex) for o in ops]
    return [np.kron(a, b) for a in ops for b in ops]


# ---------------------------------------------------------------------------
# Reference-level op description
# ---------------------------------------------------------------------------
@dataclass
class RefOp:
    kind: int
    q: int
    qubit: int
    target: int = -1
    control: int = -2
    noisy: bool = True
    mat: Optional[np.ndarray] = None
    kraus: Optional[List[np.ndarray]] = None
    basis: int = MZ
    if0: List["RefOp"] = field(default_factory=list)
    if1: List["RefOp"] = field(default_factory=list)
    name: str = ""


def ref_gate(name, qubit, target=-1, control=-2, noisy=True, mat=None):
    m = gate_matrix(name) if mat is None else np.asarray(mat, dtype=complex)
    up = name.upper()
    if up in MEAS_NAMES:
        return RefOp(KIND_MEASURE, 1, qubit, -1, -2, False, m, basis=MEAS_NAMES[up], name=up)
    q = 1 if target == -1 else 2
    return RefOp(KIND_GATE, q, qubit, target, control, noisy, m, name=name)


def ref_channel(model, p, qubit, target=-1):
    q = 2 if target > 0 else 1
    return RefOp(KIND_CHANNEL, q, qubit, target, -2, False, kraus=noise_kraus(model, p, q == 2), name=model)


def ref_channel_kraus(kraus, qubit, target=-1):
    kraus = [np.asarray(k, dtype=complex) for k in kraus]
    q = int(round(math.log2(kraus[0].shape[0])))
    return RefOp(KIND_CHANNEL, q, qubit, target if q == 2 else -1, -2, False, kraus=kraus, name="custom")


def ref_reset(qubit):
    return RefOp(KIND_RESET, 1, qubit, -1, -2, False,
                 kraus=[np.array([[1, 0], [0, 0]], dtype=complex), np.array([[0, 1], [0, 0]], dtype=complex)],
                 name="res")


def ref_ifop(name, qubit, if0, if1):
    return RefOp(KIND_IFOP, 1, qubit, -1, -2, False, gate_matrix(name), basis=MEAS_NAMES[name.upper()],
                 if0=list(if0), if1=list(if1), name=name.upper())


@dataclass
class RefNoise:
    k1: Optional[List[np.ndarray]]
    k2: Optional[List[np.ndarray]]


def ref_noise(model1, p1, model2=None, p2=None):
    """NoiseModel(Noise1(model1,p1), Noise2(model2,p2)) (struct.jl:147-173);
    model1/model2 = None stands for `false`."""
    if model2 is None and p2 is None and model1 is not None:
        model2, p2 = model1, p1
    k1 = noise_kraus(model1, p1, False) if model1 is not None else None
    k2 = noise_kraus(model2, p2, True) if model2 is not None else None
    return RefNoise(k1, k2)


# ---------------------------------------------------------------------------
# ctypes marshalling
# ---------------------------------------------------------------------------
class _COp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("q", ctypes.c_int32), ("qubit", ctypes.c_int32),
                ("target", ctypes.c_int32), ("control", ctypes.c_int32), ("noisy", ctypes.c_int32),
                ("nkraus", ctypes.c_int32), ("basis", ctypes.c_int32),
                ("mat", ctypes.POINTER(ctypes.c_double)), ("kraus", ctypes.POINTER(ctypes.c_double)),
                ("if0_begin", ctypes.c_int32), ("if0_end", ctypes.c_int32), ("if1_begin", ctypes.c_int32),
                ("if1_end", ctypes.c_int32), ("is_cx", ctypes.c_int32), ("is_cz", ctypes.c_int32)]


class _CNoise(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int32), ("k1", ctypes.POINTER(ctypes.c_double)),
                ("n2", ctypes.c_int32), ("k2", ctypes.POINTER(ctypes.c_double))]


def _colmajor(m):
    m = np.asarray(m, dtype=np.complex128)
    return np.ascontiguousarray(m.T).view(np.float64).reshape(-1).copy()


class _Packed:
    """Keeps numpy buffers alive for the C structs."""

    def __init__(self, ops: Sequence[RefOp], noise: Optional[RefNoise]):
        self.keep = []
        flat: List[RefOp] = []
        top = list(ops)
        # layout: top-level ops first, then branch ops appended
        arr_ops = list(top)
        branch_ranges = {}
        for i, op in enumerate(top):
            if op.kind == KIND_IFOP:
                b0 = len(arr_ops); arr_ops.extend(op.if0); e0 = len(arr_ops)
                b1 = len(arr_ops); arr_ops.extend(op.if1); e1 = len(arr_ops)
                branch_ranges[i] = (b0, e0, b1, e1)
        self.n = len(arr_ops)
        self.ntop = len(top)
        self.arr = (_COp * max(1, self.n))()
        for i, op in enumerate(arr_ops):
            c = self.arr[i]
            c.kind, c.q, c.qubit, c.target, c.control = op.kind, op.q, op.qubit, op.target, op.control
            c.noisy = 1 if op.noisy else 0
            c.basis = op.basis
            if op.mat is not None:
                mb = _colmajor(op.mat)
                self.keep.append(mb)
                c.mat = _dp(mb)
                c.is_cx = int(op.mat.shape == (4, 4) and np.array_equal(op.mat, GATE["CX"]))
                c.is_cz = int(op.mat.shape == (4, 4) and np.array_equal(op.mat, GATE["CZ"]))
            if op.kraus is not None:
                kb = np.concatenate([_colmajor(k) for k in op.kraus])
                self.keep.append(kb)
                c.kraus = _dp(kb)
                c.nkraus = len(op.kraus)
            if i in branch_ranges:
                c.if0_begin, c.if0_end, c.if1_begin, c.if1_end = branch_ranges[i]
        self.noise = None
        if noise is not None:
            self.noise = _CNoise()
            if noise.k1:
                b = np.concatenate([_colmajor(k) for k in noise.k1]); self.keep.append(b)
                self.noise.n1, self.noise.k1 = len(noise.k1), _dp(b)
            if noise.k2:
                b = np.concatenate([_colmajor(k) for k in noise.k2]); self.keep.append(b)
                self.noise.n2, self.noise.k2 = len(noise.k2), _dp(b)

    def noise_ptr(self):
        return ctypes.byref(self.noise) if self.noise is not None else None


def uniforms(seed: int, traj: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().bto_uniforms(seed, traj, n, _dp(out))
    return out


def run_sv(N, ops, noise=None, u=None, state=None, mode=0, tol=1e-9, outcap=4096):
    """Run ops on |0..0> (or `state`).  Returns (state, outcomes, used_draws, boundary, overflow)."""
    pk = _Packed(ops, noise)
    if state is None:
        s = np.zeros(2 ** N, dtype=np.complex128); s[0] = 1.0
    else:
        s = np.array(state, dtype=np.complex128, copy=True)
    sv = s.view(np.float64)
    if u is None:
        u = np.zeros(1)
    u = np.ascontiguousarray(u, dtype=np.float64)
    outs = np.zeros(outcap, dtype=np.int32)
    nout, used, bnd, ov = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    rc = lib().bto_run_sv(N, _dp(sv), ctypes.cast(pk.arr, ctypes.c_void_p), pk.ntop, pk.noise_ptr(), _dp(u), len(u),
                          tol, mode, outs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), outcap, ctypes.byref(nout),
                          ctypes.byref(used), ctypes.byref(bnd), ctypes.byref(ov))
    _check(rc)
    return s, outs[: nout.value].tolist(), used.value, bnd.value, ov.value


def run_traj(N, ops, noise, seed, traj0, ntraj, max_draws, mode=0, tol=1e-9, outcap=4096, keep_states=False):
    pk = _Packed(ops, noise)
    outs = np.zeros((ntraj, outcap), dtype=np.int32)
    nouts = np.zeros(ntraj, dtype=np.int64)
    states = np.zeros((ntraj, 2 ** N), dtype=np.complex128) if keep_states else None
    btot = ctypes.c_int64()
    rc = lib().bto_run_traj(N, ctypes.cast(pk.arr, ctypes.c_void_p), pk.ntop, pk.noise_ptr(), seed, traj0, ntraj,
                            max_draws, tol, mode, outs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), outcap,
                            nouts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                            _dp(states.view(np.float64)) if keep_states else None, ctypes.byref(btot))
    _check(rc)
    return [outs[t, : nouts[t]].tolist() for t in range(ntraj)], states, btot.value


def run_dm(N, ops, noise=None, rho=None, mode=0):
    """to_rho semantics: returns the column-major rho as a 2D numpy array (rho[r, c])."""
    pk = _Packed(ops, noise)
    D = 2 ** N
    if rho is None:
        r = np.zeros((D, D), dtype=np.complex128); r[0, 0] = 1.0
    else:
        r = np.array(rho, dtype=np.complex128)
    buf = np.asfortranarray(r)
    flat = buf.reshape(-1, order="F").copy()
    rc = lib().bto_run_dm(N, _dp(flat.view(np.float64)), ctypes.cast(pk.arr, ctypes.c_void_p), pk.ntop,
                          pk.noise_ptr(), mode)
    _check(rc)
    return flat.reshape((D, D), order="F")


def full_operator(N, mat, qubit, target=-1, control=-2):
    """Full-space operator built the reference's way (sparse kron fold)."""
    m = np.asarray(mat, dtype=complex)
    q = int(round(math.log2(m.shape[0])))
    D = 2 ** N
    out = np.zeros(2 * D * D)
    mb = _colmajor(m)
    is_cx = int(q == 2 and np.array_equal(m, GATE["CX"]))
    is_cz = int(q == 2 and np.array_equal(m, GATE["CZ"]))
    _check(lib().bto_full_operator(N, q, _dp(mb), qubit, target, control, is_cx, is_cz, _dp(out)))
    return out.view(np.complex128).reshape((D, D), order="F")


def apply_gate(N, state, mat, qubit, target=-1, control=-2, mode=0):
    m = np.asarray(mat, dtype=complex)
    q = int(round(math.log2(m.shape[0])))
    s = np.array(state, dtype=np.complex128, copy=True)
    mb = _colmajor(m)
    is_cx = int(q == 2 and np.array_equal(m, GATE["CX"]))
    is_cz = int(q == 2 and np.array_equal(m, GATE["CZ"]))
    _check(lib().bto_apply_gate(N, _dp(s.view(np.float64)), q, _dp(mb), qubit, target, control, is_cx, is_cz, mode))
    return s


def partial_trace(N, state, keep):
    keep = sorted(keep)
    k = len(keep)
    s = np.ascontiguousarray(state, dtype=np.complex128)
    out = np.zeros(2 * 4 ** k)
    kk = np.array(keep, dtype=np.int32)
    _check(lib().bto_partial_trace(N, _dp(s.view(np.float64)), k, kk.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   _dp(out)))
    return out.view(np.complex128).reshape((2 ** k, 2 ** k), order="F")


def calc_prob(N, state, kraus, qubit, target=-1):
    q = int(round(math.log2(kraus[0].shape[0])))
    s = np.ascontiguousarray(state, dtype=np.complex128)
    kb = np.concatenate([_colmajor(k) for k in kraus])
    probs = np.zeros(len(kraus))
    _check(lib().bto_calc_prob(N, _dp(s.view(np.float64)), q, len(kraus), _dp(kb), qubit, target, _dp(probs)))
    return probs


def weighted_sample(probs, u):
    p = np.ascontiguousarray(probs, dtype=np.float64)
    return lib().bto_weighted_sample(_dp(p), len(p), u)


def num_threads():
    return lib().bto_num_threads()


def set_num_threads(n):
    lib().bto_set_num_threads(int(n))
