"""Circuit IR, and the circuit drivers run / to_state / to_rho (src/ops.jl:659-844).

``compile_ir`` lowers an op list plus a NoiseModel into the engine's flat IR
(bt_ir_op), expanding the noise routing of apply_noise (hilbert.jl:322-340)
into explicit channel ops right after each noisy gate, so that the engine
consumes uniforms in the reference's order.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Union

import numpy as np

from . import _lib as L
from .core import (DensityMatrix, StateVector, _colmajor, channel_bits, init, zero_rho, zero_state)
from .gates import MEASUREMENTS, gate
from .noise import NoiseModel, QuantumChannel
from .ops import Op, OpF, OpQC, QuantumOps, ifOp


class IR:
    """Flat op array + matrix pool (complex128, column-major)."""

    def __init__(self, nqubits: int):
        self.n = nqubits
        self.ops: List[L.IrOp] = []
        self.pool: List[np.ndarray] = []
        self.pool_len = 0

    def add_matrix(self, m) -> int:
        m = np.asarray(m, dtype=np.complex128)
        off = self.pool_len
        self.pool.append(_colmajor(m))
        self.pool_len += m.shape[0] * m.shape[1]
        return off

    def op(self, kind, bits, k=None, prob_bits=None, ctrl=0, mat=0, nkraus=0, basis=0, record=0):
        o = L.IrOp()
        o.kind = kind
        o.k = len(bits) if k is None else k
        for i, b in enumerate(bits):
            o.bits[i] = b
        for i, b in enumerate(prob_bits or []):
            o.prob_bits[i] = b
        o.ctrl_mask = ctrl
        o.mat = mat
        o.nkraus = nkraus
        o.basis = basis
        o.record = record
        return o

    def arrays(self):
        arr = (L.IrOp * max(1, len(self.ops)))(*self.ops)
        pool = np.concatenate(self.pool) if self.pool else np.zeros(2)
        return arr, np.ascontiguousarray(pool)


def _bit(N, q):
    if q < 1 or q > N:
        raise ValueError("N must be larger than qubit")
    return N - q


def _gate_record(ir, N, q, mat, qubit, target, control):
    if N < qubit or N < target or N < control:
        raise ValueError("N must be larger than qubits")
    ctrl = (1 << _bit(N, control)) if control != -2 else 0
    if q == 1:
        bits = [_bit(N, qubit)]
    else:
        bits = [_bit(N, qubit), _bit(N, target)]
        if control != -2 and abs(qubit - target) != 1 and not (
                np.array_equal(mat, gate["CX"]) or np.array_equal(mat, gate["CZ"])):
            raise ValueError("Unsupported operation. Use 'CCZ' or 'CCX'.")
    return ir.op(L.BT_OP_GATE, bits, ctrl=ctrl, mat=ir.add_matrix(mat))


def _channel_record(ir, N, kraus, qubit, target):
    q = int(round(math.log2(kraus[0].shape[0])))
    a, p = channel_bits(N, q, qubit, target)
    off = ir.add_matrix(kraus[0])
    for K in kraus[1:]:
        ir.add_matrix(K)
    return ir.op(L.BT_OP_CHANNEL, a, prob_bits=p, mat=off, nkraus=len(kraus))


def _noise_records(ir, N, op, noise):
    """apply_noise routing (hilbert.jl:322-340)."""
    if not isinstance(noise, NoiseModel) or not getattr(op, "noisy", False):
        return []

    def ch(c, qubit, target):
        if not isinstance(c, QuantumChannel):
            raise ValueError("noise model has no channel for a %d-qubit operation" % op.q)
        return _channel_record(ir, N, c.kraus, qubit, target)

    if op.q == 1:
        return [ch(noise.q1, op.qubit, -1)] if op.control == -2 else [ch(noise.q2, op.control, op.qubit)]
    return [ch(noise.q2, op.qubit, op.target_qubit)]


def _lower(ir, N, op, noise, pars, out: list, branches: list):
    if isinstance(op, OpF):
        if isinstance(op.data, (list, tuple)):
            for o in op.data:  # o * state: apply without noise
                _lower(ir, N, o, False, pars, out, branches)
            return
        raise TypeError("OpF with a function or full-space matrix cannot be lowered to the engine IR")
    if isinstance(op, OpQC):
        if op.name.upper() in ("RES", "RESET"):
            out.append(ir.op(L.BT_OP_RESET, [_bit(N, op.qubit)], basis=L.BT_BASIS_Z))
        else:
            out.append(_channel_record(ir, N, op.kraus, op.qubit, op.target_qubit))
        return
    if op.type == "🔬":
        basis = MEASUREMENTS[op.name.upper()]
        if isinstance(op, ifOp):
            rec = ir.op(L.BT_OP_IF, [_bit(N, op.qubit)], basis=basis, record=1)
            b0, b1 = [], []
            for o in op.if01[0]:
                _lower(ir, N, o, noise, pars, b0, branches)
            for o in op.if01[1]:
                _lower(ir, N, o, noise, pars, b1, branches)
            branches.append((rec, b0, b1))
            out.append(rec)
        else:
            out.append(ir.op(L.BT_OP_MEASURE, [_bit(N, op.qubit)], basis=basis, record=1))
        return
    if op.type == "f":
        if pars is None:
            raise ValueError("function-type gates need pars=")
        mat = np.asarray(op.mat(*pars), dtype=np.complex128)
    else:
        mat = op.mat
    out.append(_gate_record(ir, N, op.q, mat, op.qubit, op.target_qubit, op.control))
    out.extend(_noise_records(ir, N, op, noise))


def compile_ir(ops: Sequence[QuantumOps], N: int, noise=False, pars=None) -> IR:
    ir = IR(N)
    top: List[L.IrOp] = []
    branches = []
    for op in ops:
        _lower(ir, N, op, noise, pars, top, branches)
    ir.ops = list(top)
    for rec, b0, b1 in branches:
        rec.br0_begin = len(ir.ops)
        ir.ops.extend(b0)
        rec.br0_end = len(ir.ops)
        rec.br1_begin = len(ir.ops)
        ir.ops.extend(b1)
        rec.br1_end = len(ir.ops)
    ir.ntop = len(top)
    return ir


def get_N(ops: Sequence[QuantumOps]) -> int:
    """get_N(ops) (ops.jl:14-34)."""
    m = 0
    for op in ops:
        if isinstance(op, OpF):
            continue
        if isinstance(op, ifOp):
            m = max(m, op.qubit)
        elif isinstance(op, OpQC):
            m = max(m, op.qubit, op.target_qubit)
        else:
            m = max(m, op.qubit, op.target_qubit, op.control)
    return m


class Circuit:
    """A compiled op list on the device (bt_circuit) with its options.
    Stands in for the reference's Circuit + Options (struct.jl:835-873) on
    the apply path: the noise model is folded into the IR."""

    def __init__(self, ops: Sequence[QuantumOps], N: Optional[int] = None, noise=False, pars=None):
        init()
        self.N = get_N(ops) if N is None else N
        self.noise = noise
        self.ir = compile_ir(ops, self.N, noise, pars)
        self.nops = len(self.ir.ops)
        self.ntop = self.ir.ntop
        self._arr, self._pool = self.ir.arrays()
        h = ctypes.c_void_p()
        L.check(L.lib().bt_circuit_create(self.N, self._arr, self.nops, self.ntop, L.dptr(self._pool),
                                          self._pool.size // 2, ctypes.byref(h)))
        self.handle = h
        self.max_draws = int(L.lib().bt_circuit_max_draws(h))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and L._lib is not None:
            try:
                L._lib.bt_circuit_destroy(h)
            except Exception:
                pass
            self.handle = None

    # --- drivers ---------------------------------------------------------
    def run_on(self, state: StateVector, seed=0, traj=0, uniforms=None, outcap=None):
        """One trajectory on `state` (in place).  Returns (outcomes, stats)."""
        outcap = outcap or max(16, 2 * self.nops)
        outs = np.zeros(outcap, dtype=np.int32)
        nout = ctypes.c_int64()
        st = L.RunStats()
        if uniforms is not None:
            u = np.ascontiguousarray(uniforms, dtype=np.float64)
            up, nu = L.dptr(u), u.size
        else:
            up, nu = None, 0
        L.check(L.lib().bt_run_circuit(state.handle, self.handle, seed, traj, up, nu, L.iptr(outs), outcap,
                                       ctypes.byref(nout), ctypes.byref(st)))
        return outs[: nout.value].tolist(), st.as_dict()

    def run_fused(self, state: StateVector, max_k=5, norm_every=0):
        """The (noiseless, gate-only) circuit fused into dense blocks of at
        most max_k bits, one pass per block (complex64, max_k=5: FP32 MFMA).
        Returns (stats, FP64 norms after every norm_every blocks)."""
        nb = ctypes.c_int64()
        st = L.RunStats()
        cap = 4096 if norm_every > 0 else 0
        norms = np.zeros(max(cap, 1), dtype=np.float64)
        L.check(L.lib().bt_run_fused(state.handle, self.handle, max_k, norm_every, L.dptr(norms) if cap else None,
                                     cap, ctypes.byref(nb), ctypes.byref(st)))
        d = st.as_dict()
        d["blocks"] = nb.value
        nn = min(cap, nb.value // norm_every) if norm_every > 0 else 0
        return d, norms[:nn].tolist()

    def autotune_fused(self, state: StateVector):
        """Pick the fusion width for this noiseless circuit on states like
        `state` (max_k 3, 4, 5 or the register-tile executor, timed on a
        scratch state); run_fused(max_k=0) then uses it.  Returns (best k,
        device ms of k = 3, 4, 5 and the register-tile executor)."""
        k = ctypes.c_int()
        ms = np.zeros(4, dtype=np.float64)
        L.check(L.lib().bt_autotune_fused(state.handle, self.handle, ctypes.byref(k), L.dptr(ms)))
        return k.value, ms.tolist()

    def capture_dm(self, rho: DensityMatrix):
        """to_rho of this circuit captured as a HIP graph on `rho`
        (bt_graph_create_dm); .launch() replays the whole circuit (reset rho
        with set_basis(0) first)."""
        return FusedGraph(self, rho, dm=True)

    def capture(self, state: StateVector, max_k=5):
        """The fused form of this (noiseless, gate-only) circuit captured as a
        HIP graph bound to `state`; call .launch() to run it again."""
        return FusedGraph(self, state, max_k)

    def autotune(self, prec="c64", sample=64, seed=12345):
        """Pick the batched executor's tile bits, coalescing reserve and tiles
        per workgroup by timing `sample` trajectories of this circuit per
        candidate; the winner stays in effect (process-wide)."""
        from .core import PRECISIONS
        t = L.Tuning()
        L.check(L.lib().bt_autotune_batch(self.handle, PRECISIONS[prec], sample, seed, ctypes.byref(t)))
        return t.as_dict()

    def run_dm(self, rho: DensityMatrix):
        st = L.RunStats()
        L.check(L.lib().bt_run_circuit_dm(rho.handle, self.handle, ctypes.byref(st)))
        return st.as_dict()

    def run_batch(self, ntraj: int, seed=0, traj0=0, prec="c64", keep_states=False, outcap=None, mode="rounds"):
        """ntraj trajectories from |0..0>: mode "rounds" (lockstep launches over
        all resident trajectories) or "persistent" (one launch, device-side
        interpreters).  Returns (outcome lists, states or None, stats)."""
        from .core import PRECISIONS
        outcap = outcap or max(16, 2 * self.nops)
        outs = np.zeros((ntraj, outcap), dtype=np.int32)
        nouts = np.zeros(ntraj, dtype=np.int64)
        states = np.zeros((ntraj, 2 ** self.N), dtype=np.complex128) if keep_states else None
        st = L.RunStats()
        fn = L.lib().bt_run_batch if mode == "rounds" else L.lib().bt_run_batch_persistent
        L.check(fn(self.handle, PRECISIONS[prec], seed, traj0, ntraj, L.iptr(outs), outcap,
                                     nouts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     L.dptr(states.view(np.float64)) if keep_states else None,
                                     ctypes.byref(st)))
        return [outs[t, : nouts[t]].tolist() for t in range(ntraj)], states, st.as_dict()

    def run_batch_sample(self, ntraj: int, seed=0, traj0=0, prec="c64", outcap=None, with_outcomes=False):
        """run_batch, then one shot per trajectory from its final state (the
        bt_sample rule, uniform u(seed, traj, draws used)).  Returns (indices,
        boundary count[, outcome lists, stats])."""
        from .core import PRECISIONS
        outcap = outcap or max(16, 2 * self.nops)
        outs = np.zeros((ntraj, outcap), dtype=np.int32)
        nouts = np.zeros(ntraj, dtype=np.int64)
        samples = np.zeros(ntraj, dtype=np.uint64)
        nb = ctypes.c_int64()
        st = L.RunStats()
        L.check(L.lib().bt_run_batch_sample(self.handle, PRECISIONS[prec], seed, traj0, ntraj, L.iptr(outs), outcap,
                                            nouts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                            samples.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                            ctypes.byref(nb), ctypes.byref(st)))
        idx = samples.astype(np.int64)
        if with_outcomes:
            return idx, nb.value, [outs[t, : nouts[t]].tolist() for t in range(ntraj)], st.as_dict()
        return idx, nb.value


def _uniforms(seed: int, traj: int, start: int, n: int) -> np.ndarray:
    f = L.lib().bt_uniform
    return np.array([f(seed, traj, start + d) for d in range(int(n))], dtype=np.float64)


def compiled(ops_or_circuit, noise=False, options=None):
    """An op list compiled as the reference's run / to_state / measure do
    (compile: swap insertion on the default line layout, then layers), or
    a CompiledCircuit / device Circuit passed through."""
    from .compile import CompiledCircuit, Options, compile as _compile
    if isinstance(ops_or_circuit, (Circuit, CompiledCircuit)):
        return ops_or_circuit
    if isinstance(ops_or_circuit, str):  # OpenQASM text or file (run(qasm_or_path), ops.jl:720-727)
        from .qasm import from_qasm, qasm_source
        ops_or_circuit = from_qasm(qasm_source(ops_or_circuit))
    opts = options or Options(noise=noise)
    return _compile(list(ops_or_circuit), opts)


def shot_circuit(cc, seed: int, traj: int, id: int = 0, extra_ops=()):
    """(device circuit, index of the first apply draw) of one shot of a
    CompiledCircuit: the twirl rows take the first draws of the shot's stream
    (seed, traj), in op order."""
    from .compile import twirl_choice
    ntw = cc.twirl_count(id)
    choices = [twirl_choice(u) for u in _uniforms(seed, traj, 0, ntw)]
    ops = cc.ops_for_shot(id, choices) + list(extra_ops)
    return Circuit(ops, max(cc.N, 1), cc.options.noise), ntw


def run_shot(cc, state: StateVector, seed: int, traj: int, id: int = 0, extra_ops=()):
    """One shot of a CompiledCircuit on `state` (in place): (outcomes, draws used
    in total, stats)."""
    dc, start = shot_circuit(cc, seed, traj, id, extra_ops)
    if start == 0:
        outs, st = dc.run_on(state, seed=seed, traj=traj)
    else:
        outs, st = dc.run_on(state, uniforms=_uniforms(seed, traj, start, max(dc.max_draws, 1)))
    return outs, start + st["draws"], st


def to_state(ops_or_circuit, N=None, noise=False, seed=0, traj=0, prec="c64", id=0):
    """to_state (ops.jl:790-795): one trajectory from |0..0> of the compiled
    circuit's shot ops (stream (seed, traj))."""
    from .compile import CompiledCircuit
    c = compiled(ops_or_circuit, noise)
    if isinstance(c, CompiledCircuit):
        st = zero_state(max(c.N, N or 0), prec)
        run_shot(c, st, seed, traj, id)
        return st
    st = zero_state(c.N, prec)
    c.run_on(st, seed, traj)
    return st


def to_rho(ops_or_circuit, N=None, noise=False, prec="c64", id=0, seed=0):
    """to_rho (ops.jl:806-844): the exact density matrix of the layered ops
    (ZNE / twirling as to_rho expands them; twirl rows from stream (seed, 0))."""
    from .compile import CompiledCircuit, twirl_choice
    c = compiled(ops_or_circuit, noise)
    if isinstance(c, CompiledCircuit):
        ch = [twirl_choice(u) for u in _uniforms(seed, 0, 0, c.rho_twirl_count(id))]
        c = Circuit(c.rho_ops(id, ch), max(c.N, N or 0), c.options.noise)
    rho = zero_rho(c.N, prec)
    c.run_dm(rho)
    return rho


def run(ops_or_circuit, number_of_experiment: int = 1, N=None, noise=False, seed=0, prec="c64", batched=True,
        id=0, options=None):
    """run (ops.jl:659-692, and run(qasm_or_path) ops.jl:720-727):
    mid-circuit outcomes per shot.  An op list (or OpenQASM text / file) is
    compiled first (swaps, layers), so shots apply vcat(layers...) as the
    reference does.  Shot t uses the uniform stream (seed, t).  Batched: all
    shots in one bt_run_batch call (one launch per round); with twirling each
    shot has its own op list and runs on its own."""
    from .compile import CompiledCircuit
    if number_of_experiment <= 0:
        raise ValueError("`number_of_experiment` must be positive.")
    c = compiled(ops_or_circuit, noise, options)
    if isinstance(c, CompiledCircuit):
        if c.options.twirl:
            res = []
            for t in range(number_of_experiment):
                st = zero_state(max(c.N, N or 0), prec)
                o, _, _ = run_shot(c, st, seed, t, id)
                res.append(o)
            return res
        c = Circuit(c.ops_for_shot(id), max(c.N, N or 0), c.options.noise)
    if batched:
        outs, _, _ = c.run_batch(number_of_experiment, seed=seed, prec=prec)
        return outs
    res = []
    for t in range(number_of_experiment):
        st = zero_state(c.N, prec)
        o, _ = c.run_on(st, seed, t)
        res.append(o)
    return res


class FusedGraph:
    """A captured circuit (bt_graph_*): a fused noiseless circuit on a state
    vector, or a density-matrix circuit (dm=True); launch() enqueues every
    pass at once on the state's stream; state.sync() waits."""

    def __init__(self, circuit: Circuit, state, max_k=5, dm=False):
        h = ctypes.c_void_p()
        if dm:
            L.check(L.lib().bt_graph_create_dm(state.handle, circuit.handle, ctypes.byref(h)))
        else:
            L.check(L.lib().bt_graph_create_fused(state.handle, circuit.handle, max_k, ctypes.byref(h)))
        self.handle = h
        self.state = state  # the graph is bound to this state's memory
        self.passes = int(L.lib().bt_graph_passes(h))

    def launch(self):
        L.check(L.lib().bt_graph_launch(self.handle))
        return self.state

    def close(self):
        h = getattr(self, "handle", None)
        if h is not None and L._lib is not None:
            L._lib.bt_graph_destroy(h)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching executor (bt_server_*): requests (circuit, kind,
    precision, seed, traj) are admitted while their states fit in the HBM
    budget and advance together, one launch per round; finished requests
    leave and queued ones take their place.  A request's result does not
    depend on what else is resident."""

    def __init__(self, hbm_budget_bytes=0, max_slots=0):
        init()
        h = ctypes.c_void_p()
        L.check(L.lib().bt_server_create(int(hbm_budget_bytes), int(max_slots), ctypes.byref(h)))
        self.handle = h
        self._circ = {}  # keep submitted circuits alive until fetched

    def submit(self, circuit: Circuit, kind="sv", prec="c64", seed=0, traj=0) -> int:
        from .core import PRECISIONS
        kinds = {"sv": 0, "dm": 1}
        rid = ctypes.c_int64()
        L.check(L.lib().bt_server_submit(self.handle, circuit.handle, kinds[kind], PRECISIONS[prec], seed, traj,
                                         ctypes.byref(rid)))
        self._circ[rid.value] = circuit
        return rid.value

    def submit_qasm(self, qasm: str, kind="sv", prec="c64", seed=0, traj=0, noise=False) -> int:
        """A request in OpenQASM 2 (text or file): imported (qasm.jl:275-421),
        compiled to layer order, lowered with `noise`, and submitted."""
        cc = compiled(qasm, noise)
        return self.submit(Circuit(cc.ops(), cc.N, noise), kind, prec, seed, traj)

    def poll(self, wait=False, cap=256):
        """Finished requests: list of dicts (id, status, nout, latency_ms, queue_ms)."""
        res = (L.Result * cap)()
        n = ctypes.c_int()
        L.check(L.lib().bt_server_poll(self.handle, res, cap, 1 if wait else 0, ctypes.byref(n)))
        return [dict(id=r.id, status=r.status, nout=r.nout, latency_ms=r.latency_ms, queue_ms=r.queue_ms)
                for r in res[: n.value]]

    def fetch(self, rid: int, probs=True, outcap=None):
        """(outcomes, probabilities or None) of a finished request; frees its state."""
        c = self._circ.get(rid)
        N = c.N if c is not None else 0
        outcap = outcap or (max(16, 2 * c.nops) if c is not None else 4096)
        outs = np.zeros(outcap, dtype=np.int32)
        nout = ctypes.c_int64()
        p = np.zeros(2 ** N, dtype=np.float64) if probs else None
        L.check(L.lib().bt_server_fetch(self.handle, rid, L.iptr(outs), outcap, ctypes.byref(nout),
                                        L.dptr(p) if probs else None, p.size if probs else 0))
        self._circ.pop(rid, None)
        return outs[: min(outcap, nout.value)].tolist(), p

    def reserve(self, kind="sv", prec="c64", nqubits=1, count=1) -> int:
        """Pre-size the state pool on the calling thread (count buffers for
        requests of this kind, precision and qubit count, within the HBM
        budget), so admitting them allocates nothing on the serving thread.
        Returns the number of buffers added."""
        from .core import PRECISIONS
        kinds = {"sv": 0, "dm": 1}
        n = ctypes.c_int()
        L.check(L.lib().bt_server_reserve(self.handle, kinds[kind], PRECISIONS[prec], int(nqubits), int(count),
                                          ctypes.byref(n)))
        return n.value

    def stats(self):
        a, q = ctypes.c_int(), ctypes.c_int()
        u = ctypes.c_uint64()
        L.check(L.lib().bt_server_stats(self.handle, ctypes.byref(a), ctypes.byref(q), ctypes.byref(u)))
        return dict(active=a.value, queued=q.value, hbm_in_use=u.value)

    def close(self):
        h = getattr(self, "handle", None)
        if h is not None and L._lib is not None:
            L._lib.bt_server_destroy(h)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
