"""Brute-force dense restatement of BlueTangle.jl's apply path for tiny N.

TEST INFRASTRUCTURE ONLY (checker, never the product).

Every operator is a dense 2^N x 2^N numpy matrix built with np.kron exactly
as src/hilbert.jl builds its sparse ones (foldl(kron, ...), the
_swap_control_target permutation, the Pauli decomposition/reconstruction for
non-adjacent pairs, CCZX), and the trajectory rules are transcribed line by
line from src/hilbert.jl:469-553,669-696,752-759,810-819 and
src/struct.jl:9-76,565-594.  It exists to pin the C oracle (bt_oracle.c),
which uses strided in-place updates instead.  Pure-Python loops: N <= 8.
"""
from __future__ import annotations

import itertools
import math
from typing import List, Optional

import numpy as np

from oracle import (GATE, KIND_CHANNEL, KIND_GATE, KIND_IFOP, KIND_MEASURE, KIND_RESET, MR, MX, MY, RefNoise,
                    RefOp)

I2 = np.eye(2, dtype=complex)
PAULIS = [GATE["I"], GATE["X"], GATE["Y"], GATE["Z"]]


def kron_fold(ms):
    out = ms[0]
    for m in ms[1:]:
        out = np.kron(out, m)
    return out


def swap_control_target(m):
    """hilbert.jl:215-237 (invariant matrices returned unchanged)."""
    SW = GATE["SWAP"]
    if np.all(m @ SW - SW @ m == 0):
        return m
    perm = [0, 2, 1, 3]
    return m[np.ix_(perm, perm)]


def hilbert1(N, mat, qubit, control=-2):
    """hilbert.jl:143-159"""
    if N < qubit or N < control:
        raise ValueError("N must be larger than qubit")
    if control == -2:
        return kron_fold([mat if x == qubit else I2 for x in range(1, N + 1)])
    a = kron_fold([mat if x == qubit else (GATE["P1"] if x == control else I2) for x in range(1, N + 1)])
    b = kron_fold([GATE["P0"] if x == control else I2 for x in range(1, N + 1)])
    return a + b


def cczx(sym, N, i, j, k):
    """hilbert.jl:73-103"""
    i, j, k = N - i, N - j, N - k
    D = 2 ** N
    H = np.zeros((D, D), dtype=complex)
    for a in range(D):
        if sym == "CCZ":
            H[a, a] += -1 if ((a >> i) & (a >> j) & (a >> k) & 1) == 1 else 1
        else:
            if ((a >> i) & (a >> j) & 1) == 1:
                H[a, a ^ (1 << k)] += 1
            else:
                H[a, a] += 1
    return H


def hilbert2(N, mat, qubit, target, control=-2):
    """hilbert.jl:18-70 with the Pauli decomposition path (decompose.jl:5-100)."""
    distance = abs(qubit - target)
    if N < qubit or N < target or N < control:
        raise ValueError("N must be larger than qubits")
    final = swap_control_target(mat) if qubit > target else mat
    if control == -2:
        if distance == 1:
            return kron_fold([final if x == qubit else I2 for x in range(1, N + 1) if x != target])
        lo = min(qubit, target)
        # coefficients tr(A P)/4 in the reference's Iterators.product order
        total = None
        # Iterators.product(1:4, 1:4) runs the first index fastest
        for b in range(4):
            for a in range(4):
                coef = np.trace(final @ np.kron(PAULIS[a], PAULIS[b])) / 4
                factors = [I2] * (lo - 1) + [PAULIS[a]] + [I2] * (distance - 1) + [PAULIS[b]]
                term = coef * kron_fold(factors)
                total = term if total is None else total + term
        for _ in range(N - max(qubit, target)):
            total = np.kron(total, I2)
        return total
    if distance == 1:
        a = kron_fold([final if x == qubit else (GATE["P1"] if x == control else I2)
                       for x in range(1, N + 1) if x != target])
        b = kron_fold([GATE["P0"] if x == control else I2 for x in range(1, N + 1)])
        return a + b
    if np.array_equal(mat, GATE["CX"]):
        return cczx("CCX", N, qubit, control, target)
    if np.array_equal(mat, GATE["CZ"]):
        return cczx("CCZ", N, qubit, control, target)
    raise ValueError("Unsupported operation. Use 'CCZ' or 'CCX'.")


def hilbert3(N, mat, first):
    """hilbert.jl:106-128"""
    e = [I2] * N
    e.insert(first - 1, mat)
    for _ in range(3):
        e.pop(first)
    return kron_fold(e)


def expand(N, mat, qubit, target=-1, control=-2):
    m = np.asarray(mat, dtype=complex)
    if m.shape[0] == 2:
        return hilbert1(N, m, qubit, control)
    if m.shape[0] == 4:
        return hilbert2(N, m, qubit, target, control)
    return hilbert3(N, m, qubit)


def partial_trace(state, keep):
    """partial_trace(state, keep) via Julia's column-major reshape
    (linalg.jl:167-230 for 1 / adjacent 2 qubits; the general form
    linalg.jl:83-140 keeps ascending qubit order too)."""
    N = int(round(math.log2(len(state))))
    keep = sorted(keep)
    rho_full = np.outer(state, np.conj(state))
    # tensor with axes (q1, ..., qN) where q1 is the MSB
    t = rho_full.reshape([2] * (2 * N))
    tr_out = [q for q in range(1, N + 1) if q not in keep]
    # trace over tr_out: axes (q-1) and (N+q-1)
    letters = "abcdefghijklmnopqrstuvwxyz"
    row = list(letters[:N])
    col = list(letters[N:2 * N]) if 2 * N <= 26 else None
    for q in tr_out:
        col[q - 1] = row[q - 1]
    out_row = "".join(row[q - 1] for q in keep)
    out_col = "".join(col[q - 1] for q in keep)
    red = np.einsum("".join(row) + "".join(col) + "->" + out_row + out_col, t)
    k = len(keep)
    return red.reshape(2 ** k, 2 ** k)


def calc_prob(state, kraus, qubits):
    pA = partial_trace(state, qubits)
    return [float(np.real(np.trace(K @ pA @ K.conj().T))) for K in kraus]


class Draws:
    def __init__(self, u):
        self.u = list(u)
        self.pos = 0

    def next(self):
        v = self.u[self.pos]
        self.pos += 1
        return v


def weighted_sample(probs, u):
    """hilbert.jl:810-819 (0-based; past the total: last positive branch)."""
    c = 0.0
    for i, p in enumerate(probs):
        c += p
        if u <= c:
            return i
    i = len(probs) - 1
    while i > 0 and not probs[i] > 0:
        i -= 1
    return i


def channel_apply(state, kraus, qubit, target, draws):
    """struct.jl:31-41,51-55"""
    N = int(round(math.log2(len(state))))
    q = int(round(math.log2(kraus[0].shape[0])))
    if q == 1:
        probs = calc_prob(state, kraus, [qubit])
    elif q == 2:
        probs = calc_prob(state, kraus, [qubit, target])
    else:
        probs = calc_prob(state, kraus, [qubit, qubit + 1, qubit + 2])
    ind = weighted_sample(probs, draws.next())
    v = expand(N, kraus[ind], qubit, target if q == 2 else -1) @ state
    return v / np.linalg.norm(v)


def born_measure_z(state, qubit, draws):
    """hilbert.jl:682-696"""
    N = int(round(math.log2(len(state))))
    prob0 = float(np.real(partial_trace(state, [qubit])[0, 0]))
    ind = 0 if draws.next() < prob0 else 1
    v = hilbert1(N, GATE["P0"] if ind == 0 else GATE["P1"], qubit) @ state
    return v / np.linalg.norm(v), ind


def measure(state, basis, qubit, draws):
    """_born_measure / __ifOp_apply rotation (hilbert.jl:669-679, struct.jl:554-594)."""
    N = int(round(math.log2(len(state))))
    if basis == MR:
        basis = [MX, MY, 0][min(2, int(math.floor(draws.next() * 3)))]
    R = GATE["H"] if basis == MX else (GATE["HSP"] if basis == MY else GATE["I"])
    Rf = hilbert1(N, R, qubit)
    state = Rf @ state
    state, ind = born_measure_z(state, qubit, draws)
    return Rf.conj().T @ state, ind


def apply_noise(state, op: RefOp, nm: Optional[RefNoise], draws):
    """hilbert.jl:322-340"""
    if nm is None or not op.noisy:
        return state
    if op.q == 1:
        if op.control == -2:
            return channel_apply(state, nm.k1, op.qubit, -1, draws)
        return channel_apply(state, nm.k2, op.control, op.qubit, draws)
    if op.q == 2:
        return channel_apply(state, nm.k2, op.qubit, op.target, draws)
    return state


def apply(state, op: RefOp, nm, draws, outs: Optional[List[int]]):
    """apply(state, op; noise, track_measurements) (hilbert.jl:469-515)."""
    N = int(round(math.log2(len(state))))
    if op.kind == KIND_CHANNEL:
        return channel_apply(state, op.kraus, op.qubit, op.target, draws)
    if op.kind == KIND_RESET:
        state, ind = born_measure_z(state, op.qubit, draws)
        if ind == 1:
            state = hilbert1(N, GATE["X"], op.qubit) @ state
        return state
    if op.kind in (KIND_MEASURE, KIND_IFOP):
        state, ind = measure(state, op.basis, op.qubit, draws)
        if outs is not None:
            outs.append(ind)
        if op.kind == KIND_IFOP:
            for b in (op.if0 if ind == 0 else op.if1):
                state = apply(state, b, nm, draws, None)
        return state
    state = expand(N, op.mat, op.qubit, op.target, op.control) @ state
    return apply_noise(state, op, nm, draws)


def run_sv(N, ops, nm=None, u=(), state=None):
    if state is None:
        state = np.zeros(2 ** N, dtype=complex)
        state[0] = 1
    d = Draws(u)
    outs: List[int] = []
    for op in ops:
        state = apply(state, op, nm, d, outs)
    return state, outs, d.pos


def dm_channel(rho, kraus, qubit, target):
    """struct.jl:58-76"""
    N = int(round(math.log2(rho.shape[0])))
    new = np.zeros_like(rho)
    for K in kraus:
        e = expand(N, K, qubit, target if K.shape[0] == 4 else -1)
        new = new + e @ rho @ e.conj().T
    return new


def run_dm(N, ops, nm=None):
    """to_rho (ops.jl:806-844) with apply(rho, op; noise) (hilbert.jl:639-666)."""
    rho = np.zeros((2 ** N, 2 ** N), dtype=complex)
    rho[0, 0] = 1
    for op in ops:
        if op.kind in (KIND_CHANNEL, KIND_RESET):
            rho = dm_channel(rho, op.kraus, op.qubit, op.target)
            continue
        if op.kind in (KIND_MEASURE, KIND_IFOP):
            raise ValueError("fix this:")
        e = expand(N, op.mat, op.qubit, op.target, op.control)
        rho = e @ rho @ e.conj().T
        if nm is not None and op.noisy:
            if op.q == 1:
                rho = dm_channel(rho, nm.k1, op.qubit, -1) if op.control == -2 else \
                    dm_channel(rho, nm.k2, op.control, op.qubit)
            else:
                rho = dm_channel(rho, nm.k2, op.qubit, op.target)
    return rho
