"""Pins the FP64 oracle (oracle/bt_oracle.c) before anything is checked
against it.  Each check follows a test of the reference (test/runtests.jl) or
a closed form, and uses an independent implementation (dense np.kron
brute force, scipy.linalg.expm, textbook channel formulas), so a dropped term,
a wrong sign or a transposed factor in the oracle fails here."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import dense as Dn
import oracle as O
from refcirc import brickwork_spec, count_draws_upper, random_spec, to_ref

RNG = np.random.default_rng(20261202)


def rand_state(N, rng=RNG):
    v = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
    return v / np.linalg.norm(v)


# --- runtests.jl:8 -- int2bin(2,4) == [0,0,1,0]: qubit 1 is the MSB -------
def test_index_convention_kat():
    s, _, _, _, _ = O.run_sv(4, [O.ref_gate("X", 3)])
    assert np.argmax(np.abs(s)) == 2  # |0010>: qubit 3 set <-> index 2
    s, _, _, _, _ = O.run_sv(4, [O.ref_gate("X", 1)])
    assert np.argmax(np.abs(s)) == 8


# --- runtests.jl:10-17 -- non-local CNOT equals controlled X, bit-exact -----
@pytest.mark.parametrize("N,q,t", [(6, 2, 5), (5, 4, 1), (4, 1, 4), (6, 5, 2)])
def test_nonlocal_cnot_equals_controlled_x(N, q, t):
    a = O.full_operator(N, O.GATE["CNOT"], q, t)
    b = O.full_operator(N, O.GATE["X"], t, -1, q)
    assert np.array_equal(a, b)
    assert np.array_equal(a, Dn.hilbert2(N, O.GATE["CNOT"], q, t))
    assert np.array_equal(b, Dn.hilbert1(N, O.GATE["X"], t, q))


# --- every placement kind: C reference-mode build == dense np.kron ---------
@pytest.mark.parametrize("name,q,t,c", [
    ("H", 2, -1, -2), ("RX(0.7)", 1, -1, 3), ("ISWAP", 2, 3, -2), ("ISWAP", 3, 2, -2),
    ("ECR", 1, 4, -2), ("ECR", 4, 1, -2), ("SYC", 2, 3, 1), ("CX", 1, 4, 2), ("CZ", 4, 1, 3),
    ("RXX(0.3)", 3, 4, -2), ("CY", 4, 3, -2)])
def test_full_operator_matches_dense(name, q, t, c):
    N = 4
    m = O.gate_matrix(name)
    a = O.full_operator(N, m, q, t, c)
    b = Dn.expand(N, m, q, t, c)
    assert np.allclose(a, b, atol=1e-14)
    # fast in-place placement == operator product
    s = rand_state(N)
    assert np.allclose(O.apply_gate(N, s, m, q, t, c, mode=0), b @ s, atol=1e-14)
    assert np.allclose(O.apply_gate(N, s, m, q, t, c, mode=1), b @ s, atol=1e-14)


def test_first_factor_on_qubit():
    """A 2-qubit matrix's first tensor factor acts on `qubit`
    (hilbert.jl:30-36, _swap_control_target): kron(A,B) at (3,1) == A on 3, B on 1."""
    A, B = O.gate_matrix("RX(0.4)"), O.gate_matrix("RY(1.1)")
    N = 3
    m = np.kron(A, B)
    lhs = O.full_operator(N, m, 3, 1)
    rhs = Dn.hilbert1(N, A, 3) @ Dn.hilbert1(N, B, 1)
    assert np.allclose(lhs, rhs, atol=1e-14)


def test_hilbert3_order():
    m = np.kron(np.kron(O.gate_matrix("H"), O.gate_matrix("RX(0.2)")), O.gate_matrix("T"))
    N = 5
    full = Dn.hilbert3(N, m, 2)
    rhs = Dn.hilbert1(N, O.gate_matrix("H"), 2) @ Dn.hilbert1(N, O.gate_matrix("RX(0.2)"), 3) @ \
        Dn.hilbert1(N, O.gate_matrix("T"), 4)
    assert np.allclose(full, rhs, atol=1e-14)
    assert np.allclose(O.full_operator(N, m, 2), rhs, atol=1e-14)


# --- gate table vs closed forms ---------------------------------------------
def test_gate_table_closed_forms():
    X, Y, Z = O.GATE["X"], O.GATE["Y"], O.GATE["Z"]
    for th in (0.0, 0.3, 2.1, 5.9):
        assert np.allclose(O.gate_matrix("RX(%r)" % th), sla.expm(-0.5j * th * X), atol=1e-14)
        assert np.allclose(O.gate_matrix("RY(%r)" % th), sla.expm(-0.5j * th * Y), atol=1e-14)
        assert np.allclose(O.gate_matrix("RZ(%r)" % th), sla.expm(-0.5j * th * Z), atol=1e-14)
        assert np.allclose(O.gate_matrix("RZZ(%r)" % th), sla.expm(-0.5j * th * np.kron(Z, Z)), atol=1e-14)
    # SX == RX(pi/2) * exp(i pi/4) (gates.jl:25)
    assert np.allclose(O.gate_matrix("SX"), O.gate_matrix("RX(π/2)") * np.exp(1j * math.pi / 4), atol=1e-14)
    assert np.allclose(O.gate_matrix("RX(.5π)"), O.gate_matrix("RX(1.5707963267948966)"))
    # HSP = H * S^dagger ... the measurement rotation maps Y eigenstates to Z ones
    Hsp = O.gate_matrix("MY")
    ev = np.linalg.eigh(Y)[1]
    for k in range(2):
        v = Hsp @ ev[:, k]
        assert abs(abs(v[0]) ** 2 + abs(v[1]) ** 2 - 1) < 1e-14
        assert min(abs(v[0]), abs(v[1])) < 1e-14
    assert np.array_equal(O.gate_matrix("CNOT"), O.GATE["CX"])


# --- Kraus tables, including quirks Q1/Q2 ------------------------------------
def test_kraus_tables():
    for model in ("amplitude_damping", "phase_damping", "depolarizing", "bit_flip", "phase_flip",
                  "bit_phase_flip", "depolarizing_amp", "rot_x", "rot_xyz"):
        ks = O.noise_kraus(model, 0.23)
        s = sum(k.conj().T @ k for k in ks)
        assert np.allclose(s, np.eye(2), atol=1e-14), model
    # Q1: phase_flip yields Y (noise.jl:80 compares a String with a Symbol)
    assert np.allclose(O.noise_kraus("phase_flip", 0.1)[1], math.sqrt(0.1) * O.GATE["Y"])
    # Q2: two-qubit model is the 16-term product
    k2 = O.noise_kraus("depolarizing", 0.1, True)
    k1 = O.noise_kraus("depolarizing", 0.1)
    assert len(k2) == 16
    assert np.allclose(k2[4 * 1 + 3], np.kron(k1[1], k1[3]))
    ad = O.noise_kraus("amplitude_damping", 0.3)
    assert np.allclose(ad[0], [[1, 0], [0, math.sqrt(0.7)]]) and np.allclose(ad[1], [[0, math.sqrt(0.3)], [0, 0]])


# --- runtests.jl:19-35 -- Kraus probabilities three ways ---------------------
@pytest.mark.parametrize("q,t", [(2, 3), (2, 5), (5, 2), (1, 6)])
def test_kraus_probability_identity(q, t):
    N = 6
    s = rand_state(N)
    kraus = O.noise_kraus("depolarizing", 0.5, True)
    pA = O.partial_trace(N, s, [q, t])
    p1 = [float(np.real(np.trace(K @ pA @ K.conj().T))) for K in kraus]
    p3 = O.calc_prob(N, s, kraus, q, t)
    assert np.allclose(p1, p3, atol=1e-14)
    if q < t:  # Q3: the (min,max) reduced rho matches the application order only for q < t
        p2 = [np.sum(np.abs(O.full_operator(N, K, q, t) @ s) ** 2) for K in kraus]
        assert np.allclose(p1, p2, atol=1e-13)
    assert np.allclose(pA, Dn.partial_trace(s, [q, t]), atol=1e-14)
    # Pauli channel: p_k equals the Kraus weight (closed form)
    w = [np.real(np.trace(K.conj().T @ K)) / 4 for K in kraus]
    assert np.allclose(p3, w, atol=1e-13)


def test_amplitude_damping_probability_closed_form():
    N = 4
    s = rand_state(N)
    g = 0.37
    for q in range(1, N + 1):
        p = O.calc_prob(N, s, O.noise_kraus("amplitude_damping", g), q)
        mask = np.array([(i >> (N - q)) & 1 for i in range(2 ** N)], dtype=bool)
        p1 = g * np.sum(np.abs(s[mask]) ** 2)
        assert np.allclose(p, [1 - p1, p1], atol=1e-14)


# --- runtests.jl:54-73 -- sequential apply == product operator (1e-12) ------
def test_sequential_equals_product_operator():
    N, depth = 8, 50
    rng = np.random.default_rng(5)
    spec = random_spec(N, depth, rng, kinds=("gate1", "gate2", "ccx"))
    ops = to_ref(spec)
    s0 = rand_state(N, rng)
    seq, _, _, _, _ = O.run_sv(N, ops, state=s0)
    U = np.eye(2 ** N, dtype=complex)
    for op in ops:
        U = O.full_operator(N, op.mat, op.qubit, op.target, op.control) @ U
    assert np.allclose(seq, U @ s0, atol=1e-12)


# --- fast mode == reference-algorithm mode == dense, all op kinds -----------
@pytest.mark.parametrize("seed", range(6))
def test_modes_agree_on_random_noisy_circuits(seed):
    rng = np.random.default_rng(100 + seed)
    N = 4 + seed % 2
    spec = random_spec(N, 40, rng, kinds=("gate1", "gate2", "ccx", "chan1", "chan2", "meas", "reset", "ifop"))
    ops = to_ref(spec)
    nm = O.ref_noise("depolarizing", 0.05, "amplitude_damping", 0.1)
    u = O.uniforms(seed, 0, 4 * len(spec) + 10)
    sf, of, uf, _, _ = O.run_sv(N, ops, nm, u, mode=0)
    sr, orf, ur, _, _ = O.run_sv(N, ops, nm, u, mode=1)
    sd, od, ud = Dn.run_sv(N, ops, nm, u)
    assert of == orf == od
    assert uf == ur == ud
    assert uf <= count_draws_upper(spec, True) + 0
    assert np.allclose(sf, sr, atol=1e-12) and np.allclose(sf, sd, atol=1e-12)


def test_draw_order_contract():
    """One draw per noisy gate's channel, per OpQC, per Born measurement, two
    for MR (struct.jl:568), plus those of the chosen ifOp branch."""
    N = 3
    spec = [("gate", "H", 1, -1, -2, True), ("gate", "CX", 1, 2, -2, True), ("gate", "X", 3, -1, 2, True),
            ("gate", "Y", 2, -1, -2, False), ("chan", "amplitude_damping", 0.2, 3, -1), ("meas", "MR", 1),
            ("reset", 2), ("ifop", "MZ", 3, [("gate", "X", 3, -1, -2, True)], [("gate", "H", 3, -1, -2, True)])]
    ops = to_ref(spec)
    nm = O.ref_noise("depolarizing", 0.1)
    u = O.uniforms(9, 0, 64)
    _, outs, used, _, _ = O.run_sv(N, ops, nm, u)
    assert used == 3 + 1 + 2 + 1 + 1 + 1
    assert len(outs) == 2
    _, outs0, used0, _, _ = O.run_sv(N, ops, None, u)
    assert used0 == 1 + 2 + 1 + 1


def test_sampling_boundary_rules():
    """Q4: Kraus sampling takes u <= cumsum (hilbert.jl:815); Born takes
    u < prob0 (hilbert.jl:693)."""
    assert O.weighted_sample([0.25, 0.75], 0.25) == 0
    assert O.weighted_sample([0.25, 0.75], 0.2500000001) == 1
    # Q5: u beyond the FP64 total (the reference returns nothing,
    # hilbert.jl:814-818): the last branch with positive probability
    assert O.weighted_sample([0.25, 0.7], 0.99) == 1
    assert O.weighted_sample([0.5, 0.45, 0.0], 0.999) == 1
    # Born: |+> has prob0 = 1/2 up to rounding; choose a state with exact prob0
    s = np.array([math.sqrt(0.5), 0, math.sqrt(0.5), 0], dtype=complex)  # qubit 1: prob0 = 0.5 exactly? (0.5 + tiny)
    p0 = float(np.real(O.partial_trace(2, s, [1])[0, 0]))
    _, outs, _, _, _ = O.run_sv(2, [O.ref_gate("MZ", 1)], None, np.array([p0]), state=s)
    assert outs == [1]
    _, outs, _, _, _ = O.run_sv(2, [O.ref_gate("MZ", 1)], None, np.array([np.nextafter(p0, 0)]), state=s)
    assert outs == [0]


# --- runtests.jl:195-231 -- mid-circuit measurement / reset KATs ------------
def test_midcircuit_kats():
    ops_reset = [("gate", "X", 1, -1, -2, True), ("reset", 1), ("meas", "MZ", 1)]
    ops_ifop = [("gate", "X", 1, -1, -2, True),
                ("ifop", "MZ", 1, [("gate", "I", 1, -1, -2, True)], [("gate", "X", 1, -1, -2, True)])]
    ops_mid = [("gate", "X", 1, -1, -2, True), ("gate", "X", 2, -1, -2, True), ("gate", "CX", 2, 3, -2, True),
               ("gate", "CX", 1, 2, -2, True), ("gate", "X", 2, -1, -2, True), ("meas", "MZ", 2), ("reset", 2),
               ("gate", "CX", 2, 3, -2, True), ("gate", "CX", 1, 2, -2, True), ("gate", "CX", 2, 3, -2, True),
               ("gate", "X", 1, -1, -2, True)]
    for seed in range(5):
        u = O.uniforms(seed, 0, 16)
        assert O.run_sv(1, to_ref(ops_reset), None, u)[1] == [0]
        assert O.run_sv(1, to_ref(ops_ifop), None, u)[1] == [1]
        assert O.run_sv(3, to_ref(ops_mid), None, u)[1] == [1]


# --- runtests.jl:89-133 -- density matrix vs pure state and vs trajectories
def test_dm_noiseless_equals_pure():
    N = 4
    rng = np.random.default_rng(7)
    spec = random_spec(N, 20, rng, kinds=("gate1", "gate2", "ccx"))
    ops = to_ref(spec)
    s, _, _, _, _ = O.run_sv(N, ops)
    rho = O.run_dm(N, ops)
    assert np.allclose(rho, np.outer(s, s.conj()), atol=1e-10)
    assert np.allclose(O.run_dm(N, ops, mode=1), rho, atol=1e-12)
    assert np.allclose(Dn.run_dm(N, ops), rho, atol=1e-12)


def test_dm_vs_trajectory_mean():
    N = 4
    rng = np.random.default_rng(11)
    spec = random_spec(N, 20, rng, kinds=("gate1", "gate2"), nonadjacent=False, controls=False)
    ops = to_ref(spec)
    nm = O.ref_noise("depolarizing", 0.1)
    rho = O.run_dm(N, ops, nm)
    ntraj = 4000
    _, states, _ = O.run_traj(N, ops, nm, 31, 0, ntraj, 4 * len(spec), keep_states=True)
    # observable: sum_q <Z_q>; 5-sigma Monte-Carlo bound from the sample spread
    zsum = np.array([sum(1 - 2 * ((i >> (N - q)) & 1) for q in range(1, N + 1)) for i in range(2 ** N)])
    vals = np.abs(states) ** 2 @ zsum
    exact = float(np.real(np.diag(rho)) @ zsum)
    assert abs(exact - vals.mean()) < 5 * vals.std() / math.sqrt(ntraj) + 1e-12
    rho_mc = np.einsum("ti,tj->ij", states, states.conj()) / ntraj
    assert np.abs(rho_mc - rho).max() < 5 / math.sqrt(ntraj)


# --- runtests.jl:136-193 -- OpQC after each gate == NoiseModel, on rho -----
def test_opqc_equals_noisemodel_on_rho():
    N = 3
    rng = np.random.default_rng(3)
    spec = random_spec(N, 20, rng, kinds=("gate1", "gate2"), controls=False)
    p = 0.01
    nm = O.ref_noise("depolarizing", p)
    with_channels = []
    for s in spec:
        g = list(s)
        g[5] = True
        with_channels.append(tuple(g))
        with_channels.append(("chan", "depolarizing", p, s[2], s[3] if s[3] > 0 else -1))
    ops_plain = to_ref(with_channels)
    for op in ops_plain:
        op.noisy = False
    r1 = O.run_dm(N, ops_plain)
    r2 = O.run_dm(N, to_ref(spec), nm)
    assert np.linalg.norm(r1 - r2) < 1e-13


# --- closed forms of channels on rho ----------------------------------------
def test_dm_channel_closed_forms():
    N = 3
    rng = np.random.default_rng(21)
    v = rand_state(N, rng)
    rho0 = np.outer(v, v.conj())
    p, g, q = 0.2, 0.3, 2
    Xf, Yf, Zf = (Dn.hilbert1(N, O.GATE[n], q) for n in "XYZ")
    # depolarizing: (1-p) rho + p/3 sum sigma rho sigma
    r = O.run_dm(N, [O.ref_channel("depolarizing", p, q)], rho=rho0)
    cf = (1 - p) * rho0 + p / 3 * (Xf @ rho0 @ Xf + Yf @ rho0 @ Yf + Zf @ rho0 @ Zf)
    assert np.allclose(r, cf, atol=1e-14)
    # amplitude damping, written out on the (q, rest) blocks
    r = O.run_dm(N, [O.ref_channel("amplitude_damping", g, q)], rho=rho0)
    b = N - q
    cf = np.zeros_like(rho0)
    for i in range(2 ** N):
        for j in range(2 ** N):
            bi, bj = (i >> b) & 1, (j >> b) & 1
            if bi == 0 and bj == 0:
                cf[i, j] = rho0[i, j] + g * rho0[i | (1 << b), j | (1 << b)]
            elif bi == 1 and bj == 1:
                cf[i, j] = (1 - g) * rho0[i, j]
            else:
                cf[i, j] = math.sqrt(1 - g) * rho0[i, j]
    assert np.allclose(r, cf, atol=1e-14)
    assert abs(np.trace(r) - 1) < 1e-14
    assert np.allclose(r, r.conj().T, atol=1e-15)
    # reset on rho == amplitude damping with gamma = 1 (struct.jl:423)
    r = O.run_dm(N, [O.ref_reset(q)], rho=rho0)
    r2 = O.run_dm(N, [O.ref_channel("amplitude_damping", 1.0, q)], rho=rho0)
    assert np.allclose(r, r2, atol=1e-15)


def test_dm_measurement_raises():
    with pytest.raises(RuntimeError):
        O.run_dm(2, [O.ref_gate("MZ", 1)])


def test_missing_noise_channel_raises():
    """apply_noise on a state does not guard a `false` channel (hilbert.jl:328-335)."""
    nm = O.ref_noise(None, None, "depolarizing", 0.1)
    with pytest.raises(RuntimeError):
        O.run_sv(2, [O.ref_gate("H", 1)], nm, O.uniforms(0, 0, 4))


# --- the counter-based uniform stream --------------------------------------
def _splitmix(x):
    M = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & M
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
    return x ^ (x >> 31)


def test_uniform_stream_restated():
    # SplitMix64 of 0 is the published first output of seed 0 (Vigna's splitmix64.c)
    assert _splitmix(0) == 0xE220A8397B1DCDAF
    for seed, traj, draw in [(0, 0, 0), (1, 2, 3), (2 ** 63 + 5, 4095, 3199)]:
        z = _splitmix(seed ^ _splitmix(traj ^ _splitmix(draw)))
        assert O.lib().bto_uniform(seed, traj, draw) == (z >> 11) * 2.0 ** -53
    u = O.uniforms(42, 7, 100000)
    assert 0 <= u.min() and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.005


def test_brickwork_shape():
    for N, even, odd in ((12, 6, 5), (22, 11, 10), (33, 16, 16)):
        spec = brickwork_spec(N, 2, 0)
        cn = [s for s in spec if s[1] == "CNOT"]
        assert len(cn) == even + odd
        assert sum(1 for s in spec if s[1].startswith("R")) == 2 * 3 * N
