"""Gate matrices and name parsing (src/gates.jl:22-59, 334-510).

The table has to match the reference's entries value for value, so it lists
the same gates; matrices are numpy complex128, row-major in Python and
handed to the engine column-major.
"""
from __future__ import annotations

import ast
import math
import operator
import re

import numpy as np

_r2 = 1.0 / math.sqrt(2.0)
_t4 = complex(round(math.cos(math.pi / 4), 10), round(math.sin(math.pi / 4), 10))  # round(..., sigdigits=10)


def _m(rows):
    return np.array(rows, dtype=np.complex128)


# gates.jl:22-59
gate = {
    "I": _m([[1, 0], [0, 1]]),
    "X": _m([[0, 1], [1, 0]]),
    "SX": 0.5 * _m([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]),
    "XSQRT": 0.5 * _m([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]),
    "Y": _m([[0, -1j], [1j, 0]]),
    "Z": _m([[1, 0], [0, -1]]),
    "H": _r2 * _m([[1, 1], [1, -1]]),
    "S": _m([[1, 0], [0, 1j]]),
    "SD": _m([[1, 0], [0, -1j]]),
    "T": _m([[1, 0], [0, _t4]]),
    "TD": _m([[1, 0], [0, _t4.conjugate()]]),
    "HSP": _r2 * _m([[1, -1j], [1, 1j]]),
    "HY": _r2 * _m([[1, 1j], [1, -1j]]),
    "H2": _m([[0.5, 0.5, 0.5, 0.5], [0.5, -0.5, 0.5, -0.5], [0.5, 0.5, -0.5, -0.5], [0.5, -0.5, -0.5, 0.5]]),
    "P0": _m([[1, 0], [0, 0]]),
    "P1": _m([[0, 0], [0, 1]]),
    "SP": _m([[0, 1], [0, 0]]),
    "SM": _m([[0, 0], [1, 0]]),
    "CX": _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]),
    "CNOT": _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]),
    "CY": _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]]),
    "CZ": _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]]),
    "CS": np.diag([1, 1, 1, 1j]).astype(np.complex128),
    "CT": np.diag([1, 1, 1, _t4]).astype(np.complex128),
    "CTD": np.diag([1, 1, 1, _t4.conjugate()]).astype(np.complex128),
    "CSD": np.diag([1, 1, 1, -1j]).astype(np.complex128),
    "CI": np.eye(4, dtype=np.complex128),
    "CH": _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, _r2, _r2], [0, 0, _r2, -_r2]]),
    "SWAP": _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]),
    "ISWAP": _m([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
    "FSWAP": _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, -1]]),
    "SYC": _m([[1, 0, 0, 0], [0, 0, -1j, 0], [0, -1j, 0, 0], [0, 0, 0, np.exp(-1j * math.pi / 6)]]),
    "ECR": _r2 * _m([[0, 1, 0, 1j], [1, 0, -1j, 0], [0, 1j, 0, 1], [-1j, 0, 1, 0]]),
}
gate["CCX"] = np.eye(8, dtype=np.complex128)
gate["CCX"][6:8, 6:8] = _m([[0, 1], [1, 0]])
gate["CCZ"] = np.diag([1, 1, 1, 1, 1, 1, 1, -1]).astype(np.complex128)


def _RX(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _m([[c, -1j * s], [-1j * s, c]])


def _RY(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _m([[c, -s], [s, c]])


def _RZ(t):
    return _m([[np.exp(-1j * t / 2), 0], [0, np.exp(1j * t / 2)]])


def _P(l):
    return _m([[1, 0], [0, np.exp(1j * l)]])


def _U2(f, l):
    return _r2 * _m([[1, -np.exp(1j * l)], [np.exp(1j * f), np.exp(1j * (f + l))]])


def _U3(t, f, l):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _m([[c, -np.exp(1j * l) * s], [np.exp(1j * f) * s, np.exp(1j * (f + l)) * c]])


def _CP(l):
    return np.diag([1, 1, 1, np.exp(1j * l)]).astype(np.complex128)


def _GIVENS(t):
    c, s = math.cos(t), math.sin(t)
    return _m([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]])


def _FSIM(t, f):
    c, s = math.cos(t), math.sin(t)
    return _m([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [0, 0, 0, np.exp(1j * f)]])


def _SWAPA(a):
    e = np.exp(1j * math.pi * a)
    return 0.5 * _m([[2, 0, 0, 0], [0, 1 + e, 1 - e, 0], [0, 1 - e, 1 + e, 0], [0, 0, 0, 2]])


def _RXX(f):
    c, s = math.cos(f), math.sin(f)
    return _m([[c, 0, 0, -1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [-1j * s, 0, 0, c]])


def _RYY(f):
    c, s = math.cos(f), math.sin(f)
    return _m([[c, 0, 0, 1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [1j * s, 0, 0, c]])


def _RZZ(f):
    return np.diag([np.exp(-1j * f / 2), np.exp(1j * f / 2), np.exp(1j * f / 2), np.exp(-1j * f / 2)])


def _RXY(f):
    c, s = math.cos(f), math.sin(f)
    return _m([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [0, 0, 0, 1]])


PARAM_GATES = {"P": _P, "RX": _RX, "RY": _RY, "RZ": _RZ, "U1": _P, "U2": _U2, "U3": _U3, "CP": _CP,
               "GIVENS": _GIVENS, "SWAPA": _SWAPA, "FSIM": _FSIM, "RXX": _RXX, "RYY": _RYY, "RZZ": _RZZ,
               "RXY": _RXY}
gates_with_phase = list(PARAM_GATES)  # gates.jl:361

MEASUREMENTS = {"MZ": 0, "M(Z)": 0, "MX": 1, "M(X)": 1, "MY": 2, "M(Y)": 2, "MR": 3, "M(R)": 3}


def _is_it_measurement(name: str) -> bool:
    """gates.jl:334-338"""
    return name.upper() in MEASUREMENTS


def _clean_name(name: str) -> str:
    """gates.jl:356"""
    return name.split("(")[0].upper()


def _name_with_arg_bool(name: str) -> bool:
    return _clean_name(name) in PARAM_GATES


_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.Pow: operator.pow}


def parse_argument(text: str) -> float:
    """Numeric gate argument ("0.3", ".1π", "pi/2", "2*pi/3").  The reference
    evaluates arbitrary Julia (eval(Meta.parse), gates.jl:348); here only
    numbers, pi/π and + - * / ^ are accepted."""
    t = text.strip().replace("π", " pi ").replace("^", "**")
    t = re.sub(r"(\d|\.)\s+pi", r"\1*pi", t)

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.Name) and n.id == "pi":
            return math.pi
        if isinstance(n, ast.BinOp) and type(n.op) in _BINOPS:
            return _BINOPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        raise ValueError("unsupported gate argument %r" % text)

    return ev(ast.parse(t, mode="eval"))


def _get_name_and_arguments(name: str):
    """gates.jl:340-353"""
    clean = _clean_name(name)
    if clean in PARAM_GATES and "(" in name:
        inner = name.split("(", 1)[1]
        args = [parse_argument(a) for a in re.sub(r"[()]", "", inner).split(",")]
        return clean, "(" + inner, args
    return clean


def gates(name: str) -> np.ndarray:
    """gates(op_name) (gates.jl:369-453)."""
    up = name.upper()
    clean = _clean_name(name)
    if clean in PARAM_GATES:
        _, _, args = _get_name_and_arguments(name)
        return np.asarray(PARAM_GATES[clean](*args), dtype=np.complex128)
    if up in ("M(Z)", "MZ", "M(R)", "MR", "RES"):
        return gate["I"].copy()
    if up in ("M(X)", "MX"):
        return gate["H"].copy()
    if up in ("M(Y)", "MY"):
        return gate["HSP"].copy()
    if up in gate:
        return gate[up].copy()
    raise KeyError("Gate %s not found" % name)
