"""Expectation values and correlations (func.jl:17-23,91-151; ops.jl:928-944).

CPU part: pins the oracle's restatement (oracle/observables.py) with closed
forms, the reference's own density-matrix test (runtests.jl:89-133: the
T-expectation summed over qubits agrees between psi, rho = psi psi' and the
trajectory mean), and two independent evaluations of every quantity (dense
kron vs axis-by-axis factors; operator form vs probability-parity form).
GPU part: bt_expect through the C ABI against the oracle, complex64 within
1e-5 and complex128 within 1e-12 (absolute, on normalised states)."""
import math

import numpy as np
import pytest

import observables as OB
import oracle as O
from refcirc import random_spec, to_ref

NAMES = ["X", "Y", "Z", "H", "S", "T", "SX", "I", "RX(0.37)", "RY(1.9)", "P(0.4)"]


def rand_state(N, rng):
    v = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
    return v / np.linalg.norm(v)


def rand_rho(N, rng, rank=3):
    vs = [rand_state(N, rng) for _ in range(rank)]
    w = rng.random(rank)
    w /= w.sum()
    return sum(wi * np.outer(v, v.conj()) for wi, v in zip(w, vs))


# --- oracle pinning (CPU) ----------------------------------------------------
def test_closed_forms():
    N = 4
    # basis state |0101>: <Z_q> = +1 where the qubit is 0, -1 where it is 1
    s = np.zeros(16, complex); s[0b0101] = 1
    assert OB.expect_str(s, "Z", N) == [1.0, -1.0, 1.0, -1.0]
    plus = np.full(16, 0.25, complex)
    assert np.allclose(OB.expect_str(plus, "X", N), 1.0, atol=1e-15)
    assert np.allclose(OB.expect_str(plus, "Y", N), 0.0, atol=1e-15)
    assert np.allclose(OB.expect_str(plus, "Z", N), 0.0, atol=1e-15)
    pi = np.array([1, 1j]) / math.sqrt(2)  # |+i>
    s = pi
    for _ in range(N - 1):
        s = np.kron(s, pi)
    assert np.allclose(OB.expect_str(s, "Y", N), 1.0, atol=1e-14)
    # GHZ: <Z_i Z_j> = 1, <X...X> = 1, <Z_i> = 0, <Y Y X X> = -1
    g = np.zeros(16, complex); g[0] = g[15] = 1 / math.sqrt(2)
    assert abs(OB.correlation_str(g, "Z,Z", [1, 3], N) - 1) < 1e-15
    assert abs(OB.correlation_str(g, "X,X,X,X", [1, 2, 3, 4], N) - 1) < 1e-15
    assert abs(OB.correlation_str(g, "Y,Y,X,X", [1, 2, 3, 4], N) + 1) < 1e-15
    assert np.allclose(OB.expect_str(g, "Z", N), 0.0)
    assert abs(OB.correlation_z(g, [2, 4], N) - 1) < 1e-15


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_dense_kron_equals_factors(seed):
    rng = np.random.default_rng(seed)
    N = 5
    psi = rand_state(N, rng)
    rho = rand_rho(N, rng)
    for _ in range(6):
        m = int(rng.integers(1, N + 1))
        qs = [int(q) for q in rng.choice(np.arange(1, N + 1), m, replace=False)]
        ops = ",".join(NAMES[int(rng.integers(len(NAMES)))] for _ in qs)
        Od = OB.expand_multi_op(ops, qs, N)
        assert abs(OB.expect_dense(psi, Od) - OB.correlation_str(psi, ops, qs, N)) < 1e-13
        assert abs(OB.expect_dense(rho, Od) - OB.correlation_str(rho, ops, qs, N)) < 1e-13


def test_z_parity_equals_operator_form():
    rng = np.random.default_rng(7)
    N = 6
    psi = rand_state(N, rng)
    for qs in ([1], [2, 5], [1, 3, 6], [1, 2, 3, 4, 5, 6]):
        a = OB.correlation_z(psi, qs, N)
        b = OB.correlation_str(psi, ",".join("Z" for _ in qs), qs, N)
        assert abs(a - b) < 1e-14


def test_reference_density_matrix_test():
    """runtests.jl:89-133 with the oracle: sum of <T_q> agrees between psi and
    rho = psi psi' (b1, 1e-10), and the noisy trajectory mean approaches the
    noisy rho within 10 / exp_no (b3; here with the oracle's own 5-sigma)."""
    N = 4
    rng = np.random.default_rng(20)
    ops = to_ref(random_spec(N, 20, rng, kinds=("gate1", "gate2"), nonadjacent=False, controls=False))
    psi = O.run_sv(N, ops)[0]
    rho = O.run_dm(N, ops)
    assert abs(sum(OB.expect_str(psi, "T", N)) - sum(OB.expect_str(rho, "T", N))) < 1e-10
    nm = O.ref_noise("depolarizing", 0.1)
    rho_n = O.run_dm(N, ops, nm)
    ntraj = 2000
    _, states, _ = O.run_traj(N, ops, nm, 5, 0, ntraj, 4 * len(ops), keep_states=True)
    vals = np.array([sum(OB.expect_str(states[t], "T", N)) for t in range(ntraj)])
    exact = sum(OB.expect_str(rho_n, "T", N))
    assert abs(exact - vals.mean()) < max(5 * vals.std() / math.sqrt(ntraj), 1e-12)


def test_repeated_qubit_replaces():
    """expand_multi_op writes result[qubit] = op in order: the last one wins."""
    N = 3
    a = OB.expand_multi_op("X,Z", [2, 2], N)
    b = OB.expand_multi_op("Z", [2], N)
    assert np.array_equal(a, b)


def test_too_many_dense_factors_is_an_error(bt):
    """More than three non-Pauli one-qubit factors are expanded into Pauli
    sums; past 4096 terms the expectation is refused (argument error) before
    the state is touched."""
    import importlib
    obs = importlib.import_module(bt.__name__ + ".observables")
    from bluetangle_jl_amd import _lib as L
    A = np.array([[0.3, 0.2 - 0.1j], [0.5j, -0.7]])  # all four Pauli coefficients nonzero
    factors = {b: A for b in range(10)}  # 3 kept dense, 7 expanded: 4^7 terms
    with pytest.raises(L.BTError) as e:
        obs._expect_product(object(), factors)
    assert e.value.code == -1


# --- engine (GPU) ------------------------------------------------------------
TOL = {"c64": 1e-5, "c128": 1e-12}


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_gpu_expect_strings(bt, prec):
    rng = np.random.default_rng(101)
    N = 9
    psi = rand_state(N, rng)
    s = bt.from_numpy(psi, prec)
    for name in NAMES:
        got = bt.expect(s, name)
        want = OB.expect_str(psi, name, N)
        assert np.abs(np.asarray(got) - want).max() < TOL[prec], name
    for _ in range(12):
        m = int(rng.integers(1, 7))
        qs = [int(q) for q in rng.choice(np.arange(1, N + 1), m, replace=False)]
        ops = ",".join(NAMES[int(rng.integers(len(NAMES)))] for _ in qs)
        got = bt.correlation(s, ops, qs)
        want = OB.correlation_str(psi, ops, qs, N)
        assert abs(got - want) < TOL[prec], (ops, qs)
    for qs in ([1], [3, 7], [1, 2, 9], list(range(1, N + 1))):
        assert abs(bt.correlation(s, qs) - OB.correlation_z(psi, qs, N)) < TOL[prec]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_gpu_expect_rho(bt, prec):
    rng = np.random.default_rng(202)
    N = 5
    rho = rand_rho(N, rng)
    r = bt.from_numpy(rho, prec)
    for name in NAMES:
        got = bt.expect(r, name)
        want = OB.expect_str(rho, name, N)
        assert np.abs(np.asarray(got) - want).max() < TOL[prec], name
    for _ in range(10):
        m = int(rng.integers(1, N + 1))
        qs = [int(q) for q in rng.choice(np.arange(1, N + 1), m, replace=False)]
        ops = ",".join(NAMES[int(rng.integers(len(NAMES)))] for _ in qs)
        assert abs(bt.correlation(r, ops, qs) - OB.correlation_str(rho, ops, qs, N)) < TOL[prec], (ops, qs)
    M = rng.normal(size=(2 ** N, 2 ** N)) + 1j * rng.normal(size=(2 ** N, 2 ** N))
    assert abs(bt.expect(r, M) - OB.expect_dense(rho, M)) < TOL[prec] * 40


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_gpu_expect_ops_and_matrix(bt, prec):
    """expect(state, op) = real(psi' op.expand(N) psi) (func.jl:91), op.expand
    from the oracle's reference-algorithm operator builder."""
    rng = np.random.default_rng(303)
    N = 6
    psi = rand_state(N, rng)
    rho = np.outer(psi, psi.conj())
    s = bt.from_numpy(psi, prec)
    r = bt.from_numpy(rho, prec)
    cases = [("H", 2, -1, -2), ("T", 6, -1, -2), ("X", 3, -1, 1), ("CNOT", 2, 3, -2), ("CNOT", 5, 2, -2),
             ("SWAP", 1, 2, -2), ("RZZ(0.7)", 4, 5, -2), ("CZ", 2, 3, 6), ("RY(0.3)", 1, -1, 4)]
    for name, q, t, c in cases:
        op = bt.Op(name, q, control=c) if t == -1 else bt.Op(name, q, t, control=c)
        full = O.full_operator(N, O.gate_matrix(name), q, t, c)
        want = OB.expect_dense(psi, full)
        assert abs(bt.expect(s, op) - want) < TOL[prec], name
        assert abs(bt.expect(r, op) - want) < TOL[prec] * 4, name
    M = rng.normal(size=(2 ** N, 2 ** N)) + 1j * rng.normal(size=(2 ** N, 2 ** N))
    assert abs(bt.expect(s, M) - OB.expect_dense(psi, M)) < TOL[prec] * 40


@pytest.mark.gpu
def test_gpu_expect_many_dense_factors(bt):
    """More than three non-Pauli factors: expanded into Pauli sums."""
    rng = np.random.default_rng(404)
    N = 7
    psi = rand_state(N, rng)
    s = bt.from_numpy(psi, "c128")
    ops, qs = "H,T,RX(0.2),S,X,RY(1.1)", [1, 2, 4, 5, 6, 7]
    assert abs(bt.correlation(s, ops, qs) - OB.correlation_str(psi, ops, qs, N)) < 1e-12


@pytest.mark.gpu
def test_gpu_expect_large_state_z(bt):
    """N=24 complex64: <Z_q> for every q and a long Z string against the
    probability-parity form on the downloaded state."""
    rng = np.random.default_rng(505)
    N = 24
    psi = (rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)).astype(np.complex64)
    psi /= np.linalg.norm(psi.astype(np.complex128))
    s = bt.from_numpy(psi, "c64")
    ref = s.to_numpy()
    got = bt.expect(s, "Z")
    p = np.abs(ref) ** 2
    idx = np.arange(2 ** N)
    want = [float(np.sum(np.where((idx >> (N - q)) & 1, -p, p))) for q in range(1, N + 1)]
    assert np.abs(got - want).max() < 1e-9
    qs = [1, 5, 9, 17, 24]
    assert abs(bt.correlation(s, qs) - OB.correlation_z(ref, qs, N)) < 1e-9
