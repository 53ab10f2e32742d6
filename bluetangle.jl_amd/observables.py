"""Expectation values and correlations on a device-resident state.

Mirrors expect / correlation of src/func.jl:17-23,91-151 and expand_multi_op /
string_to_matrix of src/ops.jl:928-944,1016-1019: the reference builds the
full-space operator and forms real(psi' O psi) or real(tr(rho O)); here the
operator is never built.  A product of one-qubit factors is split into a
Pauli string (each factor equal to s * I, X, Y or Z: bit masks and a scalar)
and at most three dense factors (a 2^k x 2^k block), and one FP64 reduction
over the state (bt_expect) evaluates it.  Factors beyond three dense ones are
expanded into Pauli sums (c_P = tr(P A) / 2), one reduction per term.
"""
from __future__ import annotations

import ctypes
import itertools
from typing import Dict, List, Sequence, Tuple, Union

import numpy as np

from . import _lib as L
from .core import DensityMatrix, DeviceState, StateVector, _colmajor, _ibuf
from .gates import gate, gates
from .ops import Op, QuantumOps

_PAULI_M = {"I": gate["I"], "X": gate["X"], "Y": gate["Y"], "Z": gate["Z"]}
_MAX_TERMS = 4096


def _pauli_form(m) -> Union[Tuple[str, complex], None]:
    """(P, s) with m == s * P exactly for P in I, X, Y, Z, else None."""
    m00, m01, m10, m11 = m[0, 0], m[0, 1], m[1, 0], m[1, 1]
    if m01 == 0 and m10 == 0:
        if m11 == m00:
            return "I", complex(m00)
        if m11 == -m00:
            return "Z", complex(m00)
    if m00 == 0 and m11 == 0:
        if m10 == m01:
            return "X", complex(m01)
        if m10 == -m01:
            return "Y", complex(-1j * m10)  # Y = [0 -i; i 0]
    return None


def _expect_raw(state: DeviceState, k: int, block, bits, xmask: int, zmask: int) -> complex:
    re, im = ctypes.c_double(), ctypes.c_double()
    if k:
        m = _colmajor(block)
        b = _ibuf(bits)
        rc = L.lib().bt_expect(state.handle, k, L.dptr(m), L.iptr(b), int(xmask), int(zmask), ctypes.byref(re),
                               ctypes.byref(im))
    else:
        rc = L.lib().bt_expect(state.handle, 0, None, None, int(xmask), int(zmask), ctypes.byref(re),
                               ctypes.byref(im))
    L.check(rc)
    return complex(re.value, im.value)


def _expect_product(state: DeviceState, factors: Dict[int, np.ndarray]) -> complex:
    """<prod_b A_b> for one-qubit factors A_b on bit positions b (complex)."""
    scalar = 1.0 + 0j
    xmask = zmask = 0
    dense: List[Tuple[int, np.ndarray]] = []
    for b in sorted(factors, reverse=True):  # MSB first
        A = np.asarray(factors[b], dtype=np.complex128)
        if A.shape != (2, 2):
            raise ValueError("one-qubit factors only")
        pf = _pauli_form(A)
        if pf is None:
            dense.append((b, A))
            continue
        p, s = pf
        scalar *= s
        if p == "X":
            xmask |= 1 << b
        elif p == "Z":
            zmask |= 1 << b
        elif p == "Y":
            xmask |= 1 << b
            zmask |= 1 << b
            scalar *= -1j  # Y[a, 1-a] = -i (-1)^a
    if scalar == 0:
        return 0j
    keep, extra = dense[:3], dense[3:]
    block = np.ones((1, 1), dtype=np.complex128)
    for _, A in keep:
        block = np.kron(block, A)
    kbits = [b for b, _ in keep]
    if not extra:
        return scalar * _expect_raw(state, len(keep), block, kbits, xmask, zmask)
    # expand the remaining dense factors into Pauli sums
    terms = []
    for b, A in extra:
        t = [(p, np.trace(_PAULI_M[p] @ A) / 2) for p in "IXYZ"]
        terms.append([(b, p, c) for p, c in t if c != 0])
    nterms = 1
    for t in terms:
        nterms *= len(t)
    if nterms > _MAX_TERMS:
        raise L.BTError(-1, "expectation: too many non-Pauli factors (%d Pauli terms)" % nterms)
    total = 0j
    for combo in itertools.product(*terms):
        x, z, c = xmask, zmask, scalar
        for b, p, coef in combo:
            c *= coef
            if p in ("X", "Y"):
                x |= 1 << b
            if p in ("Z", "Y"):
                z |= 1 << b
            if p == "Y":
                c *= -1j
        total += c * _expect_raw(state, len(keep), block, kbits, x, z)
    return total


def _bit(N: int, q: int) -> int:
    if q < 1 or q > N:
        raise ValueError("N must be larger than qubit")
    return N - q


def expand_multi_op(list_of_operators: str, qubits_applied: Sequence[int], N: int) -> Dict[int, np.ndarray]:
    """The factors of expand_multi_op (ops.jl:928-944): gate names split on ','
    placed on qubits_applied, identity elsewhere; returned as {bit: 2x2}
    (qubit q on bit N - q) instead of the 2^N x 2^N kron (ops.jl:1016-1019)."""
    ops = [s for s in str(list_of_operators).split(",")]
    if len(qubits_applied) != len(ops):
        raise ValueError("qubit number does not match with operators")
    out: Dict[int, np.ndarray] = {}
    for name, q in zip(ops, qubits_applied):
        m = gates(name)
        if m.shape != (2, 2):
            raise ValueError("expand_multi_op takes one-qubit operators (%s)" % name)
        out[_bit(N, int(q))] = m  # a later entry on the same qubit replaces the earlier one
    return out


def _op_block(op: Op, N: int, pars=None):
    """(bits MSB first, dense block) of a gate op, control as the block's MSB
    (the controlled operator P0 x I + P1 x U, hilbert.jl:156-157,52-54)."""
    mat = op.mat
    if callable(mat):
        mat = mat(*(pars or ()))
    U = np.asarray(mat, dtype=np.complex128)
    bits = [_bit(N, op.qubit)] if op.q == 1 else [_bit(N, op.qubit), _bit(N, op.target_qubit)]
    if op.control != -2:
        d = U.shape[0]
        G = np.eye(2 * d, dtype=np.complex128)
        G[d:, d:] = U
        return [_bit(N, op.control)] + bits, G
    return bits, U


def _expect_complex(state: DeviceState, what, pars=None) -> complex:
    N = state.N
    if isinstance(what, QuantumOps):
        if not isinstance(what, Op) or what.type == "🔬":
            raise ValueError("expect takes gate ops (not measurements or channels)")
        bits, G = _op_block(what, N, pars)
        if len(bits) == 1:
            return _expect_product(state, {bits[0]: G})
        return _expect_raw(state, len(bits), G, bits, 0, 0)
    if isinstance(what, np.ndarray) and what.ndim == 2:
        return _expect_full(state, what)
    raise TypeError("expect(state, op | op_str | matrix)")


def _expect_full(state: DeviceState, M) -> complex:
    """<psi|M|psi> / tr(rho M) for a full 2^N x 2^N matrix."""
    M = np.asarray(M, dtype=np.complex128)
    D = 2 ** state.N
    if M.shape != (D, D):
        raise ValueError("matrix must be 2^N x 2^N")
    if isinstance(state, DensityMatrix):
        # tr(rho M) = sum_rc conj(W_rc) rho_rc with W = M^dagger
        W = DensityMatrix(state.N, "c128").upload(M.conj().T)
        return W.inner(state)
    tmp = state.copy()
    L.check(L.lib().bt_apply_full(tmp.handle, L.dptr(_colmajor(M))))
    return state.inner(tmp)


def expect(state: DeviceState, what=None, qubit: Union[int, None] = None, pars=None):
    """expect(state|rho, op) / expect(state|rho, op_str) -> vector over qubits /
    expect(state|rho, matrix) (func.jl:91-100): the real part, as the reference.
    On a Measurement: <Z_q> from its sample (func.jl:39)."""
    from .measurement import Measurement, expect as _m_expect
    if isinstance(state, Measurement):
        return _m_expect(state, what)
    N = state.N
    if isinstance(what, str):
        m = gates(what)
        if qubit is not None:
            return _expect_product(state, {_bit(N, qubit): m}).real
        return np.array([_expect_product(state, {_bit(N, q): m}).real for q in range(1, N + 1)])
    return _expect_complex(state, what, pars).real


def correlation(state: DeviceState, list_of_operators, qubits_applied=None) -> float:
    """correlation(state|rho, "Z,Z", [2, 4]) (func.jl:142-151), or
    correlation(state, qubits) = <prod Z_q> from the exact probabilities
    (func.jl:17-23,166-182); correlation(m::Measurement, qubits) (func.jl:14)."""
    from .measurement import Measurement, correlation as _m_corr
    if isinstance(state, Measurement):
        return _m_corr(state, list_of_operators)
    N = state.N
    if qubits_applied is None:
        qubits = list(list_of_operators)
        z = 0
        for q in qubits:
            z ^= 1 << _bit(N, int(q))  # a repeated qubit cancels, as the parity sum does
        return _expect_raw(state, 0, None, [], 0, z).real
    return _expect_product(state, expand_multi_op(list_of_operators, list(qubits_applied), N)).real
