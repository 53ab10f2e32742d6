"""OpenQASM 2 import / export (qasm.jl:33-421) and run(qasm) (ops.jl:720-727).

CPU part: the import of a program that uses every statement form against the
op list the reference's rules give (known answer, written out by hand), the
imported circuit's operator against the OpenQASM meaning of the same
statements built independently with numpy (cx a,b = X on b controlled by a;
ccx a,b,c = X on c controlled by a and b; cswap a,b,c = swap of b and c
controlled by a; cu3(t,0,0) = controlled RY(t); qubit i = qasm index i - 1),
and export -> import round trips.  GPU part: run(qasm) against run(ops), and
a QASM request through the serving loop against the same circuit run alone."""
import math

import numpy as np
import pytest

import oracle as O

PROGRAM = """OPENQASM 2.0;
include "qelib1.inc";
qreg q[4];
creg c[4];
// a comment; with a semicolon
h q[0];
cx q[0],q[1];
rz(pi/2) q[2];
u3(0.1,0.2,0.3) q[3];
cu1(0.5) q[0],q[3];
cu3(0.7,0,0) q[1],q[2];
crx(0.25) q[2],q[3];
ccx q[0],q[1],q[2];
cswap q[0],q[1],q[3];
rxx(0.3) q[1],q[2];
fsim(0.1,0.2) q[0],q[1];
swap q[2],q[3];
sdg q[1];
sx q[0];
u2(0.1,0.2) q[1];
barrier q[0],q[1];
"""

EXPECTED = [("H", 1, -1, -2), ("X", 2, -1, 1), ("rz(pi/2)", 3, -1, -2), ("U3(0.1,0.2,0.3)", 4, -1, -2),
            ("U1(0.5)", 4, -1, 1), ("ry(0.7)", 3, -1, 2), ("rx(0.25)", 4, -1, 3), ("CX", 1, 3, 2),
            ("SWAP", 2, 4, 1), ("RXX(0.3)", 2, 3, -2), ("FSIM(0.1,0.2)", 1, 2, -2), ("SWAP", 3, 4, -2),
            ("SD", 2, -1, -2), ("XSQRT", 1, -1, -2), ("U2(0.1,0.2)", 2, -1, -2)]


@pytest.fixture(scope="module")
def qz(bt):
    import importlib
    return importlib.import_module(bt.__name__ + ".qasm")


def desc(ops):
    return [(o.name, o.qubit, o.target_qubit, o.control) for o in ops]


def to_ref_ops(ops):
    out = []
    for o in ops:
        if o.type == "🔬":
            out.append(O.ref_gate(o.name, o.qubit))
        else:
            out.append(O.ref_gate(o.name, o.qubit, o.target_qubit, o.control, mat=np.asarray(o.mat)))
    return out


# independent OpenQASM semantics on N qubits (qasm index k <-> qubit k + 1 <-> bit N - 1 - k)
def op_on(N, mats_by_qasm_index):
    out = np.ones((1, 1), dtype=complex)
    for k in range(N):
        out = np.kron(out, mats_by_qasm_index.get(k, np.eye(2)))
    return out


def controlled(N, ctrls, target_fn):
    """sum over control patterns: all controls 1 -> target_fn(), else identity."""
    D = 2 ** N
    U = np.zeros((D, D), dtype=complex)
    P0, P1 = np.diag([1, 0]), np.diag([0, 1])
    on = op_on(N, {c: P1 for c in ctrls})
    U += on @ target_fn()
    # identity on the complement of "all controls 1"
    U += np.eye(D) - on
    return U


def test_import_known_answer(bt, qz):
    ops = qz.from_qasm(PROGRAM)
    assert desc(ops) == EXPECTED
    ops = qz.from_qasm("measure q[2] -> c[2]; reset anc[0]; CX Q[1],Q[0];")
    assert ops[0].name == "MZ" and ops[0].qubit == 3
    assert isinstance(ops[1], bt.OpQC) and ops[1].qubit == 1
    assert (ops[2].name, ops[2].qubit, ops[2].control) == ("X", 1, 2)
    with pytest.raises(ValueError):
        qz.from_qasm("foo q[0];")
    with pytest.raises(ValueError):
        qz.from_qasm("h q;")


def test_import_semantics(bt, qz):
    """The imported ops, applied by the oracle, equal the OpenQASM meaning of
    the statements built with numpy.  (A controlled SWAP on non-adjacent qubits
    imports, but applying it raises, as hilbert.jl:22-26 does.)"""
    N = 4
    H, X = O.GATE["H"], O.GATE["X"]

    def ry(t):
        return np.array([[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]])

    def rx(t):
        return np.array([[math.cos(t / 2), -1j * math.sin(t / 2)], [-1j * math.sin(t / 2), math.cos(t / 2)]])

    SW = np.zeros((16, 16))
    for i in range(16):
        b = [(i >> (3 - k)) & 1 for k in range(4)]
        b[1], b[2] = b[2], b[1]
        SW[sum(v << (3 - k) for k, v in enumerate(b)), i] = 1
    prog = "h q[0]; cx q[0],q[1]; cu3(0.7,0,0) q[1],q[2]; crx(0.25) q[2],q[3]; ccx q[0],q[1],q[2]; " \
           "cswap q[0],q[1],q[2]; rz(pi/2) q[3];"
    steps = [op_on(N, {0: H}),
             controlled(N, [0], lambda: op_on(N, {1: X})),
             controlled(N, [1], lambda: op_on(N, {2: ry(0.7)})),
             controlled(N, [2], lambda: op_on(N, {3: rx(0.25)})),
             controlled(N, [0, 1], lambda: op_on(N, {2: X})),
             controlled(N, [0], lambda: SW),
             op_on(N, {3: np.diag([np.exp(-1j * math.pi / 4), np.exp(1j * math.pi / 4)])})]
    U = np.eye(16)
    for S in steps:
        U = S @ U
    rng = np.random.default_rng(1)
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    psi /= np.linalg.norm(psi)
    got = O.run_sv(N, to_ref_ops(qz.from_qasm(prog)), state=psi)[0]
    assert np.abs(got - U @ psi).max() < 1e-12


def test_round_trip(bt, qz):
    ops = qz.from_qasm(PROGRAM)
    text = qz.to_qasm(ops)
    assert text.startswith('OPENQASM 2.0;\ninclude "qelib1.inc";\nopaque crx(theta) a,b;')
    assert "qreg q[4];\ncreg c[4];" in text and text.endswith("u2(0.1,0.2) q[1];\n")
    again = qz.from_qasm(text)
    assert desc(again) == [(n if n != "U3(0.7,0,0)" else n, q, t, c) for n, q, t, c in desc(ops)]
    assert qz.to_qasm(again) == text
    # measurement / reset / controlled and plain forms
    ops = [bt.Op("MX", 1), bt.Op("MY", 2), bt.Op("RES", 3), bt.Op("Z", 2, control=1), bt.Op("T", 3),
           bt.Op("CZ", 1, 2, control=3), bt.Op("I", 1)]
    text = qz.to_qasm(ops, nqubits=3)
    body = text.split("creg c[3];\n")[1].strip().split("\n")
    assert body == ["h q[0];", "measure q[0] -> c[0];", "sdg q[1];", "h q[1];", "measure q[1] -> c[1];",
                    "reset q[2];", "cz q[0],q[1];", "t q[2];", "ccz q[0],q[2],q[1];"]
    assert "gate ccz a,b,c { h c; ccx a,b,c; h c; }" in text
    with pytest.raises(ValueError):
        qz.to_qasm([bt.OpQC("depolarizing", 0.1, 1)])
    with pytest.raises(ValueError):
        qz.to_qasm([])


# --- engine (GPU) ------------------------------------------------------------------
RUNNABLE = PROGRAM.replace("cswap q[0],q[1],q[3];", "cswap q[0],q[1],q[2];")


@pytest.mark.gpu
def test_gpu_run_qasm(bt, qz):
    text = RUNNABLE + "measure q[0] -> c[0]; measure q[2] -> c[2]; h q[3]; measure q[3] -> c[3];"
    nm = bt.NoiseModel(bt.Noise1("depolarizing", 0.02), bt.Noise2("depolarizing", 0.05))
    a = bt.run(text, 32, noise=nm, seed=4, prec="c128")
    b = bt.run(qz.from_qasm(text), 32, noise=nm, seed=4, prec="c128")
    assert a == b and len(a) == 32 and all(len(x) == 3 for x in a)


@pytest.mark.gpu
def test_gpu_server_qasm(bt, qz):
    from bluetangle_jl_amd.circuit import Circuit, Server, compiled
    text = RUNNABLE + "measure q[1] -> c[1];"
    nm = bt.NoiseModel(bt.Noise1("depolarizing", 0.02), bt.Noise2("depolarizing", 0.05))
    srv = Server()
    try:
        rid = srv.submit_qasm(text, "sv", "c128", seed=3, traj=5, noise=nm)
        done = []
        while not done:
            done = srv.poll(wait=True)
        outs, probs = srv.fetch(rid)
    finally:
        srv.close()
    cc = compiled(text, nm)
    st = bt.zero_state(cc.N, "c128")
    o2, _ = Circuit(cc.ops(), cc.N, nm).run_on(st, seed=3, traj=5)
    assert outs == o2
    assert np.abs(probs - st.probabilities()).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_gpu_qasm_vs_oracle(bt, qz, prec):
    """run(qasm) on the device against the oracle's run of the parsed program
    (qasm.jl:285-421 import, then vcat(layers...) as run applies it,
    ops.jl:659-692): the same outcome lists, and each shot's final state
    within 1e-12 (complex128) / 1e-5 (complex64), with a noise model."""
    from bluetangle_jl_amd.circuit import Circuit, compiled
    text = RUNNABLE + "measure q[0] -> c[0]; reset q[1]; measure q[2] -> c[2]; h q[3]; measure q[3] -> c[3];"
    nm = bt.NoiseModel(bt.Noise1("amplitude_damping", 0.05), bt.Noise2("depolarizing", 0.05))
    rnm = O.ref_noise("amplitude_damping", 0.05, "depolarizing", 0.05)
    shots, seed = 16, 21
    outs = bt.run(text, shots, noise=nm, seed=seed, prec=prec)
    cc = compiled(text, nm)
    ops = cc.ops()
    ref = []
    for o in ops:
        if isinstance(o, bt.OpQC):
            ref.append(O.ref_reset(o.qubit))
        elif o.type == "🔬":
            ref.append(O.ref_gate(o.name, o.qubit))
        else:
            # the SWAPs that compile inserts for routing are noiseless (noisy=False)
            ref.append(O.ref_gate(o.name, o.qubit, o.target_qubit, o.control, noisy=o.noisy, mat=np.asarray(o.mat)))
    _, states, _ = Circuit(ops, cc.N, nm).run_batch(shots, seed=seed, prec=prec, keep_states=True)
    tol = 1e-12 if prec == "c128" else 1e-5
    checked = 0
    for t in range(shots):
        r, routs, _, bnd, _ = O.run_sv(cc.N, ref, rnm, O.uniforms(seed, t, 256), tol=1e-6)
        if bnd:
            continue
        assert outs[t] == routs, t
        assert np.abs(states[t] - r).max() < tol, t
        checked += 1
    assert checked >= shots - 2
