"""Final sampling and the Measurement record (ops.jl:46-96,488-524,599-646,
856-905; func.jl:14,39,166-182,233-236).

CPU part pins the oracle (oracle/sampling.py): sampling-rule known answers,
the empirical distribution against the probabilities, get_probs_from_sample
and mag_moments known answers, the sample-based <Z> against the exact
probability-parity form, and the final-basis rules against explicit
rotations.  GPU part: bt_sample and bt_run_batch_sample through the C ABI
against the oracle on the same uniforms -- indices identical except shots the
engine or the oracle flags as within 1e-12 of a cumulative boundary (counted,
and bounded)."""
import math

import numpy as np
import pytest

import observables as OB
import oracle as O
import sampling as SP
from refcirc import brickwork_spec, random_spec, to_bt, to_ref


def rand_state(N, rng):
    v = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
    return v / np.linalg.norm(v)


# --- oracle pinning (CPU) ----------------------------------------------------
def test_sample_rule_kats():
    p = np.array([0.0, 0.5, 0.0, 0.5])
    idx, flag = SP.sample_rule(p, np.array([0.0, 0.25, 0.4999, 0.5, 0.75, 0.999999]))
    assert idx.tolist() == [1, 1, 1, 3, 3, 3]
    assert flag.tolist() == [True, False, False, True, False, False]
    # one-hot state: every shot on that index; unnormalised weights scale out
    p = np.zeros(16); p[11] = 7.0
    idx, _ = SP.sample_rule(p, np.linspace(0, 0.999, 50))
    assert set(idx.tolist()) == {11}


def test_sample_rule_distribution():
    rng = np.random.default_rng(3)
    p = rng.random(32) ** 3
    p /= p.sum()
    u = O.uniforms(9, 0, 200000)
    idx, _ = SP.sample_rule(p, u)
    freq = np.bincount(idx, minlength=32) / u.size
    sigma = np.sqrt(p * (1 - p) / u.size)
    assert np.all(np.abs(freq - p) < 5 * sigma + 1e-12)


def test_probs_from_sample_and_moments_kats():
    outs, probs = SP.get_probs_from_sample([3, 1, 3, 3])
    assert outs == [1, 3] and probs == [0.25, 0.75]
    # |00>: magnetisation 2 -> moments 2^k; |01>, |10>: 0; |11>: -2
    assert [SP.mag_moments(2, [0], [1.0], k) for k in (1, 2, 3)] == [2.0, 4.0, 8.0]
    assert SP.mag_moments(2, [1, 2], [0.5, 0.5], 2) == 0.0
    assert SP.mag_moments(2, [3], [1.0], 3) == -8.0
    assert SP.fock(2, 4) == [0, 0, 1, 0]  # int2bin(2, 4), runtests.jl:8


def test_sample_expectation_equals_exact_parity():
    rng = np.random.default_rng(4)
    N = 6
    psi = rand_state(N, rng)
    p = np.abs(psi) ** 2
    b = np.nonzero(p)[0]
    for qs in ([1], [2, 5], [1, 3, 6]):
        assert abs(SP.sample_to_expectation(b, p[b], qs, N) - OB.correlation_z(psi, qs, N)) < 1e-13


def test_final_basis_rules():
    """X and Y bases rotate with H / HSP (ops.jl:874-896): <X_q> before equals
    <Z_q> after; R picks one of H, HSP, I per qubit with probability 1/3."""
    rng = np.random.default_rng(5)
    N = 4
    psi = rand_state(N, rng)
    for basis, op in (("X", "X"), ("Y", "Y")):
        rot, _, _, _, _ = O.run_sv(N, SP.final_basis_ref_ops(N, basis), state=psi)
        assert np.allclose(OB.expect_str(rot, "Z", N), OB.expect_str(psi, op, N), atol=1e-13)
    # R: the branch for qubit q is decided by draw q alone (state-independent weights)
    ops = SP.final_basis_ref_ops(N, "R")
    u = np.array([0.1, 0.5, 0.9, 0.2])
    rot, _, used, _, _ = O.run_sv(N, ops, state=psi, u=u)
    assert used == N
    mats = [O.GATE["H"], O.GATE["HSP"], O.GATE["I"]]
    want = psi
    for q, uq in enumerate(u, start=1):
        k = 0 if uq <= 1 / 3 else (1 if uq <= 2 / 3 else 2)
        want = OB.apply_factors(want, {q: mats[k]}, N)
    assert np.allclose(rot, want, atol=1e-13)


def test_readout_noise_on_state_vector_raises(bt):
    """measure(circuit) with readout noise calls apply_noise(state, op,
    ::QuantumChannel), which has no method in the reference (ops.jl:869 vs
    hilbert.jl:322): the engine raises an unsupported-operation error before
    any device work, and the readout options stay accepted by compile."""
    from bluetangle_jl_amd import _lib as L
    ops = [bt.Op("H", 1), bt.Op("CNOT", 1, 2)]
    with pytest.raises(L.BTError) as e:
        bt.measure_circuit(ops, 10, readout_noise=0.1)
    assert e.value.code == -2 and "ops.jl:869" in str(e.value)
    with pytest.raises(ValueError):
        bt.measure_circuit(ops, 0)


# --- engine (GPU) ------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("N,prec", [(10, "c128"), (16, "c64"), (22, "c64")])
def test_gpu_sample_rule(bt, N, prec):
    rng = np.random.default_rng(N)
    psi = rand_state(N, rng)
    psi[rng.integers(0, 2 ** N, size=2 ** N // 3)] = 0  # runs of zero probability
    s = bt.from_numpy(psi, prec)
    ref = s.to_numpy()
    u = O.uniforms(77, N, 20000)
    got, nb = bt.measurement.sample_with(s, u)
    want, flag = SP.sample_rule(np.abs(ref) ** 2, u)
    diff = got != want
    assert not np.any(diff & ~flag), np.nonzero(diff & ~flag)[0][:5]
    assert nb <= 5 and flag.sum() <= 5
    # zero-probability entries are never drawn
    assert np.all(np.abs(ref[got]) > 0)


@pytest.mark.gpu
def test_gpu_sample_edge_cases(bt):
    s = bt.zero_state(5, "c64")
    got, nb = bt.measurement.sample_with(s, [0.0, 0.5, 0.999999])
    assert got.tolist() == [0, 0, 0]
    s.set_basis(31)
    got, _ = bt.measurement.sample_with(s, [0.0, 0.3])
    assert got.tolist() == [31, 31]
    got, nb = bt.measurement.sample_with(s, [])
    assert got.size == 0 and nb == 0
    with pytest.raises(bt.BTError):
        bt.measurement.sample_with(s, [1.0])
    with pytest.raises(TypeError):
        bt.sample(bt.zero_rho(3), 4)


@pytest.mark.gpu
def test_gpu_measure_state_record(bt):
    rng = np.random.default_rng(8)
    N = 8
    psi = rand_state(N, rng)
    s = bt.from_numpy(psi, "c128")
    bt.seed(21)
    m = bt.measure(s, 5000)
    u = O.uniforms(21, 0, 5000)
    idx, flag = SP.sample_rule(np.abs(psi) ** 2, u)
    assert flag.sum() == 0
    outs, probs = SP.get_probs_from_sample(idx.tolist())
    assert m.bitstr.tolist() == outs and np.allclose(m.sample, probs, atol=0, rtol=0)
    ex, mm = SP.record(outs, probs, N)
    assert np.allclose(m.expect, ex, atol=1e-12) and np.allclose(m.mag_moments, mm, rtol=1e-12)
    assert abs(bt.expect(m, 3) - ex[2]) < 1e-12
    assert abs(bt.correlation(m, [1, 4]) - SP.sample_to_expectation(outs, probs, [1, 4], N)) < 1e-12
    # exact form (number_of_experiment = -1): sample_exact probabilities
    me = bt.measure(s)
    assert np.allclose(me.expect, OB.expect_str(psi, "Z", N), atol=1e-12)
    # measure(sample, N)
    m2 = bt.measure(idx.tolist(), N=N)
    assert m2.bitstr.tolist() == outs and m2.number_of_experiment == 5000


@pytest.mark.gpu
@pytest.mark.parametrize("basis", ["Z", "X", "Y", "R"])
def test_gpu_measure_circuit_noiseless(bt, basis):
    N = 7
    rng = np.random.default_rng(31)
    spec = random_spec(N, 30, rng, kinds=("gate1", "gate2"), nonadjacent=True, controls=True)
    m = bt.measure_circuit(to_bt(spec, bt), 3000, N=N, measurement_basis=basis, seed=5, prec="c128")
    idx, flag = SP.measure_circuit(N, to_ref(spec), 3000, basis=basis, seed=5)
    assert flag.sum() == 0 and m.boundary_draws == 0
    outs, probs = SP.get_probs_from_sample(idx.tolist())
    assert m.bitstr.tolist() == outs and np.array_equal(m.sample, probs)
    assert m.measurement_basis == basis


@pytest.mark.gpu
@pytest.mark.parametrize("basis", ["Z", "R"])
def test_gpu_measure_circuit_noisy(bt, basis):
    """Noisy branch: one trajectory per experiment, sampled once on the device
    (bt_run_batch_sample), against per-trajectory oracle runs."""
    N = 6
    spec = brickwork_spec(N, 6, seed=3, ad_gamma=0.05, mz_prob=0.2)
    nm_ref = O.ref_noise("depolarizing", 0.02, "depolarizing", 0.05)
    nm = bt.NoiseModel(bt.Noise1("depolarizing", 0.02), bt.Noise2("depolarizing", 0.05))
    shots = 300
    m = bt.measure_circuit(to_bt(spec, bt), shots, N=N, noise=nm, measurement_basis=basis, seed=9, prec="c128")
    idx, flag = SP.measure_circuit(N, to_ref(spec), shots, noise=nm_ref, basis=basis, seed=9)
    assert flag.sum() == 0 and m.boundary_draws == 0
    outs, probs = SP.get_probs_from_sample(idx.tolist())
    assert m.bitstr.tolist() == outs and np.array_equal(m.sample, probs)


@pytest.mark.gpu
def test_gpu_batch_sample_matches_batch(bt):
    """bt_run_batch_sample leaves the trajectories' outcomes as bt_run_batch
    has them, and its samples follow the rule on each final state."""
    N = 12
    spec = brickwork_spec(N, 6, seed=2, ad_gamma=5e-3, mz_prob=0.1)
    nm = bt.NoiseModel(bt.Noise1("depolarizing", 1e-3), bt.Noise2("depolarizing", 1e-2))
    c = bt.Circuit(to_bt(spec, bt), N, nm)
    ntraj = 64
    outs, states, _ = c.run_batch(ntraj, seed=4, prec="c64", keep_states=True)
    idx, nb, outs2, st = c.run_batch_sample(ntraj, seed=4, prec="c64", with_outcomes=True)
    assert outs2 == outs
    flags = 0
    for t in range(ntraj):
        st1 = bt.zero_state(N, "c64")
        _, stats = c.run_on(st1, seed=4, traj=t)
        u = O.uniforms(4, t, stats["draws"] + 1)[-1:]
        want, f = SP.sample_rule(np.abs(states[t]) ** 2, u)
        flags += int(f[0])
        if not f[0]:
            assert idx[t] == want[0], t
    assert flags <= 1 and nb <= 1
