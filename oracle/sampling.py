"""CPU restatement of BlueTangle.jl's final sampling and Measurement record.

TEST INFRASTRUCTURE ONLY (checker, never the product).

The reference draws shots with StatsBase.sample(0:2^N-1, Weights(|psi|^2),
shots) (src/ops.jl:46-60), whose RNG and algorithm are not pinned by any
reference test.  The engine replaces the RNG by explicit uniforms and fixes
the rule: shot s is the smallest i with u_s T < C_i, C the sequential FP64
cumulative sum and T its last entry.  `sample_rule` states that rule with a
sequential cumsum; a shot with u_s T within 1e-12 T of a C_i is flagged.

Also restated: get_probs_from_sample (ops.jl:72-84), _sample_to_expectation
(func.jl:166-182), mag_moments from a sample (func.jl:233-236), measure(state,
shots) / measure(sample, N) (ops.jl:488-521), and measure(circuit, shots)
(ops.jl:599-646) over the oracle's trajectories, with the final-basis rule
_final_measurement (ops.jl:856-905): X -> H, Y -> HSP on every qubit; R ->
per qubit one of H, HSP, I chosen by one draw u <= cumsum(1/3, 2/3, 1).
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

import oracle as O

TOL = 1e-12


def sample_rule(p: np.ndarray, u: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """(indices, boundary flags) for uniforms u over weights p (sequential cumsum)."""
    p = np.asarray(p, dtype=np.float64)
    C = np.cumsum(p)
    T = C[-1]
    t = np.asarray(u, dtype=np.float64) * T
    idx = np.searchsorted(C, t, side="right")  # smallest i with t < C_i
    idx = np.minimum(idx, len(p) - 1)
    lo = np.where(idx > 0, C[np.maximum(idx - 1, 0)], 0.0)
    hi = C[idx]
    flag = (t - lo < TOL * T) | (hi - t < TOL * T)
    return idx.astype(np.int64), flag


def get_probs_from_sample(sample: Sequence[int]) -> Tuple[List[int], List[float]]:
    """ops.jl:72-84 (a Dict of counts, then sorted by outcome)."""
    counts: Dict[int, int] = {}
    for o in sample:
        counts[int(o)] = counts.get(int(o), 0) + 1
    total = len(sample)
    outs = sorted(counts)
    return outs, [counts[o] / total for o in outs]


def fock(i: int, N: int) -> List[int]:
    """int2bin(i, N) (bit.jl:9-15): qubit 1 first (the MSB)."""
    return [(i >> (N - q)) & 1 for q in range(1, N + 1)]


def sample_to_expectation(bitstr, probs, qubits, N) -> float:
    """func.jl:166-182."""
    s = 0.0
    for b, p in zip(bitstr, probs):
        f = fock(int(b), N)
        if sum(f[q - 1] for q in qubits) % 2 == 0:
            s += p
        else:
            s -= p
    return s


def mag_moments(N, bitstr, probs, order) -> float:
    """func.jl:233-236: sum((sum(mag_basis(b, N)))^order .* sample)."""
    return float(sum(sum(1 - 2 * x for x in fock(int(b), N)) ** order * p for b, p in zip(bitstr, probs)))


def record(bitstr, probs, N):
    """The derived fields of a Measurement (ops.jl:500-503)."""
    expect = [sample_to_expectation(bitstr, probs, [q], N) for q in range(1, N + 1)]
    mm = [mag_moments(N, bitstr, probs, k) for k in range(1, 13)]
    return expect, mm


def final_basis_ref_ops(N: int, basis: str):
    b = basis.upper()
    if b == "Z":
        return []
    if b == "X":
        return [O.ref_gate("XBasis", q, mat=O.GATE["H"], noisy=False) for q in range(1, N + 1)]
    if b == "Y":
        return [O.ref_gate("YBasis", q, mat=O.GATE["HSP"], noisy=False) for q in range(1, N + 1)]
    if b == "R":
        ks = [O.GATE["H"] / np.sqrt(3.0), O.GATE["HSP"] / np.sqrt(3.0), O.GATE["I"] / np.sqrt(3.0)]
        return [O.ref_channel_kraus(ks, q) for q in range(1, N + 1)]
    raise ValueError("measurement_basis error!")


def measure_circuit(N, ops, shots, noise=None, basis="Z", seed=0, noisy_branch=None):
    """measure(circuit, shots) (ops.jl:599-646) on the oracle with the engine's
    uniform contract: the ops compiled (line-layout swaps, then layer order),
    then noiseless -> one state on stream (seed, 0), shots from its next
    draws; otherwise experiment t = trajectory on stream (seed, t) sampled once
    with its next draw.  Returns (indices, flags)."""
    import compile_ref as CR
    full = CR.layered(CR.line_swaps(list(ops))) + final_basis_ref_ops(N, basis)
    has_mid = any(o.kind in (O.KIND_MEASURE, O.KIND_IFOP) for o in ops)
    noisy = noise is not None or has_mid if noisy_branch is None else noisy_branch
    maxd = 8 * (len(full) + 4)
    if not noisy:
        u = O.uniforms(seed, 0, maxd + shots)
        s, _, used, _, _ = O.run_sv(N, full, noise, u=u)
        return sample_rule(np.abs(s) ** 2, u[used: used + shots])
    idx, flags = [], []
    for t in range(shots):
        u = O.uniforms(seed, t, maxd + 1)
        s, _, used, _, _ = O.run_sv(N, full, noise, u=u)
        i, f = sample_rule(np.abs(s) ** 2, u[used: used + 1])
        idx.append(int(i[0]))
        flags.append(bool(f[0]))
    return np.array(idx, dtype=np.int64), np.array(flags)
