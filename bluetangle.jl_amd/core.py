"""Device-resident states and the per-op apply path
(src/hilbert.jl:322-364,469-557,639-696,752-759; src/struct.jl:9-76,554-594).

The reference returns a new SparseVector/SparseMatrixCSC from every apply;
here a state is a handle to HBM that apply() updates in place and returns.
Uniform draws follow Julia's global-RNG order from a counter stream
(``seed``), one per sampled channel, one per Born measurement and one extra
for the MR basis.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Union

import numpy as np

from . import _lib as L
from .gates import MEASUREMENTS, gate
from .noise import NoiseModel, QuantumChannel
from .ops import Op, OpF, OpQC, QuantumOps, ifOp

PRECISIONS = {"c64": L.BT_C64, "c128": L.BT_C128, "complex64": L.BT_C64, "complex128": L.BT_C128}

_initialised = False


def init(device: int = 0):
    global _initialised
    if not _initialised:
        L.check(L.lib().bt_init(device))
        _initialised = True


# ---------------------------------------------------------------------------
# uniform stream standing in for Julia's global rand()
# ---------------------------------------------------------------------------
class UniformStream:
    def __init__(self, seed=0, traj=0):
        self.seed_value = int(seed)
        self.traj = int(traj)
        self.draw = 0

    def __call__(self):
        u = L.lib().bt_uniform(self.seed_value, self.traj, self.draw)
        self.draw += 1
        return u


_stream = UniformStream()


def seed(s: int, traj: int = 0):
    """Random.seed!(s) for the engine's draws."""
    global _stream
    _stream = UniformStream(s, traj)
    return _stream


def current_stream() -> UniformStream:
    return _stream


def _colmajor(m) -> np.ndarray:
    m = np.asarray(m, dtype=np.complex128)
    return np.ascontiguousarray(m.T).reshape(-1).view(np.float64).copy()


def _ibuf(vals):
    return np.ascontiguousarray(np.array(vals, dtype=np.int32))


# ---------------------------------------------------------------------------
# states
# ---------------------------------------------------------------------------
class DeviceState:
    kind = L.BT_SV

    def __init__(self, N: int, prec="c64", _handle=None):
        init()
        self.N = int(N)
        self.prec = PRECISIONS[prec] if isinstance(prec, str) else int(prec)
        if _handle is None:
            h = ctypes.c_void_p()
            L.check(L.lib().bt_state_create(self.N, self.kind, self.prec, ctypes.byref(h)))
            _handle = h
        self.handle = _handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and L._lib is not None:
            try:
                L._lib.bt_state_destroy(h)
            except Exception:
                pass
            self.handle = None

    @property
    def length(self):
        return 2 ** self.N if self.kind == L.BT_SV else 4 ** self.N

    @property
    def precision(self):
        return "c64" if self.prec == L.BT_C64 else "c128"

    def copy(self):
        other = type(self)(self.N, self.prec)
        L.check(L.lib().bt_state_copy(other.handle, self.handle))
        return other

    def set_basis(self, index: int = 0):
        L.check(L.lib().bt_state_set_basis(self.handle, int(index)))
        return self

    def inner(self, other):
        """<self|other> in FP64 (either precision on either side)."""
        re, im = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib().bt_inner(self.handle, other.handle, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def sync(self):
        """Wait for the work queued on this state's stream."""
        L.check(L.lib().bt_state_sync(self.handle))
        return self

    def upload(self, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128).reshape(-1, order="F"))
        if a.size != self.length:
            raise ValueError("expected %d entries" % self.length)
        L.check(L.lib().bt_state_upload(self.handle, L.dptr(a.view(np.float64)), a.size))
        return self

    def _download(self):
        out = np.empty(self.length, dtype=np.complex128)
        L.check(L.lib().bt_state_download(self.handle, L.dptr(out.view(np.float64)), out.size))
        return out

    def probabilities(self) -> np.ndarray:
        out = np.empty(2 ** self.N, dtype=np.float64)
        L.check(L.lib().bt_probabilities(self.handle, L.dptr(out)))
        return out

    def norm2(self) -> float:
        v = ctypes.c_double()
        L.check(L.lib().bt_norm2(self.handle, ctypes.byref(v)))
        return v.value

    def gather(self, idx) -> np.ndarray:
        i = np.ascontiguousarray(np.asarray(idx, dtype=np.uint64))
        out = np.empty(i.size, dtype=np.complex128)
        L.check(L.lib().bt_gather(self.handle, i.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), i.size,
                                  L.dptr(out.view(np.float64))))
        return out


class StateVector(DeviceState):
    """Device state vector (replaces the reference's SparseVector{ComplexF64})."""
    kind = L.BT_SV

    def to_numpy(self) -> np.ndarray:
        return self._download()


class DensityMatrix(DeviceState):
    """Device density matrix, column-major (replaces SparseMatrixCSC)."""
    kind = L.BT_DM

    def to_numpy(self) -> np.ndarray:
        return self._download().reshape((2 ** self.N, 2 ** self.N), order="F")

    def trace_purity(self):
        t, p = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib().bt_trace_purity(self.handle, ctypes.byref(t), ctypes.byref(p)))
        return t.value, p.value


def zero_state(N: int, prec="c64") -> StateVector:
    """|0...0> (hilbert.jl:866)."""
    return StateVector(N, prec)


def zero_rho(N: int, prec="c64") -> DensityMatrix:
    """zero_state(N) * zero_state(N)' (ops.jl:809-810)."""
    return DensityMatrix(N, prec)


def from_numpy(arr, prec="c64"):
    a = np.asarray(arr, dtype=np.complex128)
    if a.ndim == 1:
        N = int(round(math.log2(a.size)))
        return StateVector(N, prec).upload(a)
    N = int(round(math.log2(a.shape[0])))
    return DensityMatrix(N, prec).upload(a)


def get_N(state) -> int:
    return state.N


# ---------------------------------------------------------------------------
# per-op kernels through the C ABI
# ---------------------------------------------------------------------------
def _bit(N, q):
    if q < 1 or q > N:
        raise ValueError("N must be larger than qubit")
    return N - q


def _gate(state, q, mat, qubit, target, control):
    N = state.N
    if N < qubit or N < target or N < control:
        raise ValueError("N must be larger than qubits")
    ctrl = 0
    if q == 1:
        bits = [_bit(N, qubit)]
    else:
        bits = [_bit(N, qubit), _bit(N, target)]
        if control != -2 and abs(qubit - target) != 1 and not (
                np.array_equal(mat, gate["CX"]) or np.array_equal(mat, gate["CZ"])):
            raise ValueError("Unsupported operation. Use 'CCZ' or 'CCX'.")
    if control != -2:
        ctrl = 1 << _bit(N, control)
    b = _ibuf(bits)
    L.check(L.lib().bt_apply_matrix(state.handle, len(bits), L.dptr(_colmajor(mat)), L.iptr(b), ctrl))


def channel_bits(N, q, qubit, target):
    """(apply bits, probability bits) of a channel: K's first factor on
    `qubit`; probabilities use the reduced rho in ascending qubit order
    (struct.jl:17-29, linalg.jl:83-230)."""
    if q == 1:
        a = [_bit(N, qubit)]
        return a, a
    if q == 2:
        a = [_bit(N, qubit), _bit(N, target)]
        p = [_bit(N, min(qubit, target)), _bit(N, max(qubit, target))]
        return a, p
    if qubit + 2 > N:
        raise ValueError("N must be larger than all three qubits")
    a = [N - qubit, N - qubit - 1, N - qubit - 2]
    return a, a


def _channel_sv(state, kraus, qubit, target, rng):
    q = int(round(math.log2(kraus[0].shape[0])))
    a, p = channel_bits(state.N, q, qubit, target)
    kb = np.concatenate([_colmajor(k) for k in kraus])
    ab, pb = _ibuf(a), _ibuf(p)
    chosen = ctypes.c_int()
    L.check(L.lib().bt_apply_kraus_traj(state.handle, q, len(kraus), L.dptr(kb), L.iptr(ab), L.iptr(pb), rng(),
                                        ctypes.byref(chosen), None))
    return chosen.value


def _channel_dm(state, kraus, qubit, target):
    q = int(round(math.log2(kraus[0].shape[0])))
    if q == 3:
        raise ValueError("3-qubit channels have no density-matrix method")
    a, _ = channel_bits(state.N, q, qubit, target)
    kb = np.concatenate([_colmajor(k) for k in kraus])
    L.check(L.lib().bt_apply_kraus_dm(state.handle, q, len(kraus), L.dptr(kb), L.iptr(_ibuf(a))))


def _measure(state, name, qubit, rng):
    basis = MEASUREMENTS[name.upper()]
    u = np.array([rng(), rng()] if basis == L.BT_BASIS_R else [rng(), 0.0])
    out = ctypes.c_int()
    L.check(L.lib().bt_measure(state.handle, _bit(state.N, qubit), basis, L.dptr(u), ctypes.byref(out), None))
    return out.value


def _reset(state, qubit, rng):
    out = ctypes.c_int()
    L.check(L.lib().bt_reset(state.handle, _bit(state.N, qubit), rng(), ctypes.byref(out)))
    return out.value


def apply_noise(state, op, noise: NoiseModel, rng=None):
    """apply_noise (hilbert.jl:322-364).  The state-vector method does not
    guard a `false` channel; the reference then fails, and so does this."""
    rng = rng or _stream
    if not getattr(op, "noisy", False) or isinstance(op, (OpQC, ifOp, OpF)):
        return state
    dm = isinstance(state, DensityMatrix)

    def ch(c, qubit, target):
        if not isinstance(c, QuantumChannel):
            raise ValueError("noise model has no channel for a %d-qubit operation" % op.q)
        if dm:
            _channel_dm(state, c.kraus, qubit, target)
        else:
            _channel_sv(state, c.kraus, qubit, target, rng)

    if op.q == 1:
        if op.control == -2:
            ch(noise.q1, op.qubit, -1)
        else:
            ch(noise.q2, op.control, op.qubit)
    elif op.q == 2:
        ch(noise.q2, op.qubit, op.target_qubit)
    return state


def _apply_one(state, op, noise, track, pars, rng, outs):
    dm = isinstance(state, DensityMatrix)
    if isinstance(op, OpF):
        if dm:
            raise TypeError("OpF has no density-matrix method")
        d = op.data
        if callable(d) and not isinstance(d, np.ndarray):
            res = d(state)
            return res if res is not None else state
        if isinstance(d, (list, tuple)):
            for o in d:
                _apply_one(state, o, False, False, pars, rng, None)
            return state
        m = _colmajor(np.asarray(d, dtype=np.complex128))
        L.check(L.lib().bt_apply_full(state.handle, L.dptr(m)))
        return state
    if isinstance(op, OpQC):
        if dm:
            _channel_dm(state, op.kraus, op.qubit, op.target_qubit)
        elif op.name.upper() in ("RES", "RESET"):
            _reset(state, op.qubit, rng)
        else:
            _channel_sv(state, op.kraus, op.qubit, op.target_qubit, rng)
        return state
    if op.type == "🔬":
        if dm:
            raise ValueError("fix this: density-matrix measurement is not implemented (hilbert.jl:772)")
        ind = _measure(state, op.name, op.qubit, rng)
        if isinstance(op, ifOp):
            for o in op.if01[0 if ind == 0 else 1]:
                _apply_one(state, o, noise, False, pars, rng, None)
        if track and outs is not None:
            outs.append(ind)
        return state
    if op.type == "f":
        if pars is None:
            raise ValueError("function-type gates need pars=")
        mat = np.asarray(op.mat(*pars), dtype=np.complex128)
    else:
        mat = op.mat
    if dm:
        _gate(state, op.q, mat, op.qubit, op.target_qubit, op.control)
    else:
        _gate(state, op.q, mat, op.qubit, op.target_qubit, op.control)
    if isinstance(noise, NoiseModel):
        apply_noise(state, op, noise, rng)
    return state


def _as_op(x):
    if isinstance(x, QuantumOps):
        return x
    if isinstance(x, tuple):
        return Op(*x)
    raise TypeError("ops are not valid!")


def apply(a, b, noise: Union[NoiseModel, bool] = False, track_measurements: bool = False, pars=None, rng=None):
    """apply(state, op), apply(state, ops), apply(op, state), apply(ops, state)
    with keyword `noise` and `track_measurements` (hilbert.jl:469-557).
    Returns the state, or (state, outcomes) when tracking."""
    rng = rng or _stream
    if isinstance(a, DeviceState):
        state, ops = a, b
    else:
        ops, state = a, b
    if not isinstance(state, DeviceState):
        raise TypeError("apply needs a device state (zero_state / from_numpy)")
    outs: List[int] = []
    if isinstance(ops, (list,)) or (isinstance(ops, tuple) and len(ops) and isinstance(ops[0], tuple)):
        for o in ops:
            _apply_one(state, _as_op(o), noise, track_measurements, pars, rng, outs)
    else:
        _apply_one(state, _as_op(ops), noise, track_measurements, pars, rng, outs)
    return (state, outs) if track_measurements else state


def sample_exact(state):
    """sample_exact (ops.jl:98-101,129-132): (0-based indices, probabilities)
    of the nonzero entries of |psi|^2 or diag(rho)."""
    p = state.probabilities()
    nz = np.nonzero(p)[0]
    return nz, p[nz]
