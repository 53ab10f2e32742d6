"""Final sampling and the Measurement record.

Mirrors sample / sample_state / get_probs_from_sample (src/ops.jl:46-96),
measure(state, shots) and measure(sample, N) (src/ops.jl:488-524),
measure(circuit, shots) (src/ops.jl:599-646) with _final_measurement
(src/ops.jl:856-905), the Measurement struct (src/struct.jl:907-917) and the
sample-based expect / correlation / mag_moments (src/func.jl:14,39,166-182,
233-236).

Shots are drawn on the device (bt_sample, bt_run_batch_sample) from
explicit uniforms in place of StatsBase.sample's RNG: shot s takes the
smallest basis index i with u_s T < C_i (FP64 cumulative sum in index order).
The uniforms come from the engine's counter stream: measure(state, shots)
takes the next `shots` draws of the current stream; measure(circuit, shots)
gives experiment t the stream (seed, t), the final sample taking the draw
after the circuit's own.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence, Union

import numpy as np

from . import _lib as L
from .core import StateVector, current_stream, zero_state
from .gates import gate
from .noise import QuantumChannel
from .ops import Op, OpQC, QuantumOps

BASES = ("Z", "X", "Y", "R")


@dataclass
class Measurement:
    """struct Measurement (struct.jl:907-917).  bitstr: sorted 0-based basis
    indices that occurred; sample: their frequencies."""
    bitstr: np.ndarray
    sample: np.ndarray
    expect: np.ndarray
    mag_moments: np.ndarray
    measurement_basis: str
    number_of_experiment: int
    circuit_name: str
    number_of_qubits: int
    density_matrix: Optional[np.ndarray] = None
    boundary_draws: int = field(default=0, compare=False)


def _draws(n: int, rng=None) -> np.ndarray:
    rng = rng or current_stream()
    return np.array([rng() for _ in range(int(n))], dtype=np.float64)


def sample_with(state: StateVector, u) -> (np.ndarray, int):
    """Indices for the given uniforms (device), and how many lay within
    1e-12 of a cumulative boundary."""
    if not isinstance(state, StateVector):
        raise TypeError("sample takes a state vector (ops.jl:46)")
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
    out = np.zeros(u.size, dtype=np.uint64)
    nb = ctypes.c_int64()
    L.check(L.lib().bt_sample(state.handle, u.size, L.dptr(u), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                              ctypes.byref(nb)))
    return out.astype(np.int64), nb.value


def sample(state: StateVector, shots: int, rng=None) -> np.ndarray:
    """sample(state, shots) (ops.jl:46-60): `shots` basis indices drawn from
    |psi|^2."""
    return sample_with(state, _draws(shots, rng))[0]


def get_probs_from_sample(sample_: Sequence[int], N: int):
    """ops.jl:72-84: (sorted distinct outcomes, their frequencies)."""
    s = np.asarray(sample_, dtype=np.int64).reshape(-1)
    if s.size == 0:
        return np.zeros(0, dtype=np.int64), np.zeros(0)
    vals, counts = np.unique(s, return_counts=True)
    return vals, counts / s.size


def sample_state(state: StateVector, shots: int, rng=None):
    """sample_state (ops.jl:92): get_probs_from_sample of sample(state, shots)."""
    return get_probs_from_sample(sample(state, shots, rng), state.N)


def _fock_bits(bitstr, N):
    """fock_basis(i, N)[q] = bit N - q of i (bit.jl:33, 9-15): (len, N) array."""
    b = np.asarray(bitstr, dtype=np.int64).reshape(-1, 1)
    shifts = np.arange(N - 1, -1, -1, dtype=np.int64).reshape(1, -1)
    return (b >> shifts) & 1


def sample_to_expectation(bitstr, probs, qubits: Sequence[int], N: int) -> float:
    """_sample_to_expectation (func.jl:166-182): sum of p over outcomes with
    an even number of ones at `qubits`, minus the rest."""
    f = _fock_bits(bitstr, N)
    par = f[:, [q - 1 for q in qubits]].sum(axis=1) if len(qubits) else np.zeros(len(f), dtype=np.int64)
    p = np.asarray(probs, dtype=np.float64)
    return float(np.sum(np.where(par % 2 == 0, p, -p)))


def mag_moments(N: int, bitstr, probs, moment_order: int) -> float:
    """mag_moments(N, bitstr, sample, order) (func.jl:233-236): sum over
    outcomes of (sum_q (1 - 2 b_q))^order p."""
    mag = (1 - 2 * _fock_bits(bitstr, N)).sum(axis=1).astype(np.float64)
    return float(np.sum(mag ** moment_order * np.asarray(probs, dtype=np.float64)))


def _record(bitstr, probs, N, basis, nexp, name, boundary=0) -> Measurement:
    ex = np.array([sample_to_expectation(bitstr, probs, [q], N) for q in range(1, N + 1)])
    mm = np.array([mag_moments(N, bitstr, probs, k) for k in range(1, 13)])
    return Measurement(np.asarray(bitstr, dtype=np.int64), np.asarray(probs, dtype=np.float64), ex, mm, basis,
                       int(nexp), name, N, None, boundary)


def measure(x, number_of_experiment: int = -1, N: Optional[int] = None, label="state to measurement",
            rng=None) -> Measurement:
    """measure(state, number_of_experiment) (ops.jl:488-506): exact
    probabilities when number_of_experiment == -1, else sampled shots;
    measure(sample::Vector{Int}, N) (ops.jl:509-521)."""
    if isinstance(x, StateVector):
        n = x.N
        if number_of_experiment == -1:
            p = x.probabilities()
            nz = np.nonzero(p)[0]
            return _record(nz, p[nz], n, "0", -1, label)
        idx, nb = sample_with(x, _draws(number_of_experiment, rng))
        b, p = get_probs_from_sample(idx, n)
        return _record(b, p, n, "0", number_of_experiment, label, nb)
    if N is None:
        raise ValueError("measure(sample, N) needs N")
    s = np.asarray(x, dtype=np.int64).reshape(-1)
    b, p = get_probs_from_sample(s, N)
    return _record(b, p, N, "0", s.size, "sample to measurement")


def correlation(m: Measurement, qubits: Sequence[int]) -> float:
    """correlation(m::Measurement, qubits) (func.jl:14)."""
    return sample_to_expectation(m.bitstr, m.sample, list(qubits), m.number_of_qubits)


def expect(m: Measurement, qubit: Optional[int] = None):
    """expect(m::Measurement[, qubit]) (func.jl:39): <Z_q> from the sample."""
    if qubit is None:
        return m.expect.copy()
    return sample_to_expectation(m.bitstr, m.sample, [qubit], m.number_of_qubits)


# --- measure(circuit, shots) -------------------------------------------------
_R_CHANNEL = [gate["H"] / np.sqrt(3.0), gate["HSP"] / np.sqrt(3.0), gate["I"] / np.sqrt(3.0)]


def final_basis_ops(N: int, basis: str):
    """_final_measurement (ops.jl:856-905) as ops appended to the circuit:
    Z: identity (nothing); X: H on every qubit; Y: HSP on every qubit; R: per
    qubit one of H, HSP, I drawn uniformly, written as the channel
    {H, HSP, I} / sqrt(3) (one draw per qubit, u <= cumsum picks the branch;
    the reference's rand(1:3))."""
    b = basis.upper()
    if b not in BASES:
        raise ValueError("measurement_basis error!")
    if b == "Z":
        return []
    if b == "X":
        return [Op("XBasis", gate["H"], q, noisy=False) for q in range(1, N + 1)]
    if b == "Y":
        return [Op("YBasis", gate["HSP"], q, noisy=False) for q in range(1, N + 1)]
    return [OpQC("RBasis", _R_CHANNEL, q) for q in range(1, N + 1)]


def measure_circuit(circuit_or_ops, number_of_experiment: int, N: Optional[int] = None, noise=False,
                    measurement_basis: Optional[str] = None, readout_noise=False, seed: int = 0, prec="c64",
                    circuit_name: Optional[str] = None, id: int = 0, twirl: bool = False) -> Measurement:
    """measure(circuit, number_of_experiment, id) (ops.jl:599-646).  An op list
    is compiled first (swaps, layers) with Options(noise, measurement_basis,
    twirl); a compiled circuit brings its own options.  Without noise and
    mid-circuit measurements: one state (stream (seed, 0)), rotated to the
    final basis, sampled number_of_experiment times with the next draws of the
    same stream.  Otherwise: experiment t is a trajectory on stream (seed, t)
    sampled once with its next draw (bt_run_batch_sample; with twirling, each
    experiment runs its own op list), as the reference's per-experiment
    to_state + sample(state, 1)."""
    from .circuit import Circuit, _uniforms, compiled, run_shot
    from .compile import CompiledCircuit, Options
    if isinstance(circuit_or_ops, CompiledCircuit):
        cc = circuit_or_ops
    else:
        cc = compiled(circuit_or_ops, options=Options(
            circuit_name=circuit_name or "circuit", measurement_basis=measurement_basis or "Z", noise=noise,
            twirl=twirl, readout_noise=readout_noise))
    opts = cc.options
    if opts.readout_noise is not False and opts.readout_noise is not None:
        # the state-vector path calls apply_noise(state, op, ::QuantumChannel),
        # which has no method in the reference (ops.jl:869 vs hilbert.jl:322)
        raise L.BTError(-2, "readout noise on the state-vector path is not supported by the reference (ops.jl:869)")
    if number_of_experiment <= 0:
        raise ValueError("number_of_experiment must be positive")
    n = max(cc.N, N or 0)
    basis = opts.measurement_basis
    final = final_basis_ops(n, basis)
    noiseless = opts.noise is False and cc.stats["mid_measurement_count"] == 0
    if noiseless:
        st = zero_state(n, prec)
        _, used, _ = run_shot(cc, st, seed, 0, 0, final)
        idx, nb = sample_with(st, _uniforms(seed, 0, used, number_of_experiment))
    elif not opts.twirl:
        c = Circuit(cc.ops_for_shot(id) + final, n, opts.noise)
        idx, nb = c.run_batch_sample(number_of_experiment, seed=seed, prec=prec)
    else:
        idx = np.zeros(number_of_experiment, dtype=np.int64)
        nb = 0
        for t in range(number_of_experiment):
            st = zero_state(n, prec)
            _, used, _ = run_shot(cc, st, seed, t, id, final)
            i, b = sample_with(st, _uniforms(seed, t, used, 1))
            idx[t] = i[0]
            nb += b
    b, p = get_probs_from_sample(idx, n)
    return _record(b, p, n, basis, number_of_experiment, opts.circuit_name, nb)
