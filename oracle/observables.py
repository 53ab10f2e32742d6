"""CPU restatement of BlueTangle.jl's expectation values and correlations.

TEST INFRASTRUCTURE ONLY (checker, never the product).

Follows src/func.jl:91-101 (expect: real(state' * O * state), real(tr(rho * O))),
src/func.jl:97-98 (expect(state, op_str): one value per qubit 1..N),
src/func.jl:142-151 (correlation(state|rho, op_str, qubits)),
src/func.jl:17-23,166-182 (correlation(state, qubits): parity-signed sum of the
exact probabilities from sample_exact), src/ops.jl:928-944 (expand_multi_op)
and src/ops.jl:1016-1019 (string_to_matrix = foldl(kron, gates.(names))).

Two evaluations of the same quantity: `expand_multi_op` builds the dense
2^N x 2^N kron exactly as the reference (small N), and `apply_factors`
applies the one-qubit factors to the state tensor axis by axis (any N that
fits in memory).  Qubit q is tensor axis q - 1 of the row-major reshape
(2,)*N, i.e. bit N - q of the index (src/bit.jl:9-15).
"""
from __future__ import annotations

from typing import Dict, List, Sequence

import numpy as np

from oracle import gate_matrix


def expand_multi_op(list_of_operators: str, qubits_applied: Sequence[int], N: int) -> np.ndarray:
    """ops.jl:928-944 + ops.jl:1016-1019 (dense; N <= 12)."""
    ops = list_of_operators.split(",")
    if len(qubits_applied) != len(ops):
        raise ValueError("qubit number does not match with operators")
    names = ["I"] * N
    for op, q in zip(ops, qubits_applied):
        names[q - 1] = op
    out = np.ones((1, 1), dtype=complex)
    for n in names:
        out = np.kron(out, gate_matrix(n))
    return out


def expect_dense(x: np.ndarray, O: np.ndarray) -> float:
    """func.jl:91-92,100-101: real(psi' O psi) or real(tr(rho O))."""
    if x.ndim == 1:
        return float(np.real(np.vdot(x, O @ x)))
    return float(np.real(np.trace(x @ O)))


def apply_factors(psi: np.ndarray, factors: Dict[int, np.ndarray], N: int) -> np.ndarray:
    """(prod_q A_q) psi with A_q on qubit q, by tensor contraction per axis."""
    t = np.asarray(psi, dtype=complex).reshape((2,) * N)
    for q, A in factors.items():
        t = np.moveaxis(np.tensordot(np.asarray(A, dtype=complex), t, axes=([1], [q - 1])), 0, q - 1)
    return t.reshape(-1)


def expect_factors(x: np.ndarray, factors: Dict[int, np.ndarray], N: int) -> complex:
    """<psi|prod A_q|psi> or tr(rho prod A_q) (complex), without the 2^N matrix.
    rho: tr(rho O) = sum_a (O rho)[a, a]; O acts on the row index of each column."""
    if x.ndim == 1:
        return complex(np.vdot(x, apply_factors(x, factors, N)))
    D = x.shape[0]
    cols = np.stack([apply_factors(x[:, c], factors, N) for c in range(D)], axis=1)
    return complex(np.trace(cols))


def expect_str(x: np.ndarray, op_str: str, N: int) -> List[float]:
    """func.jl:97-98: [expect of op_str on qubit q for q in 1..N]."""
    m = gate_matrix(op_str)
    return [expect_factors(x, {q: m}, N).real for q in range(1, N + 1)]


def correlation_str(x: np.ndarray, list_of_operators: str, qubits: Sequence[int], N: int) -> float:
    """func.jl:142-151 via the factors of expand_multi_op."""
    ops = list_of_operators.split(",")
    if len(qubits) != len(ops):
        raise ValueError("qubit number does not match with operators")
    f: Dict[int, np.ndarray] = {}
    for op, q in zip(ops, qubits):
        f[q] = gate_matrix(op)
    return expect_factors(x, f, N).real


def correlation_z(state: np.ndarray, qubits: Sequence[int], N: int) -> float:
    """func.jl:17-23 + func.jl:166-182: sum over basis states of p_i, signed by
    the parity of the bits at `qubits` (fock_basis(i, N)[q] = bit N - q)."""
    p = np.abs(state) ** 2
    idx = np.arange(p.size, dtype=np.int64)
    par = np.zeros(p.size, dtype=np.int64)
    for q in qubits:
        par += (idx >> (N - q)) & 1
    return float(np.sum(np.where(par % 2 == 0, p, -p)))
