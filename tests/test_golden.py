"""Golden fixtures (tests/golden, written by tests/golden/make_fixtures.py
with the oracle): pinned on the CPU against the reference-algorithm mode of
the oracle and an independent dense NumPy brute force (oracle/dense.py),
and checked on the GPU against the engine through the C ABI."""
import json
import os

import numpy as np
import pytest

import dense as Dn
import oracle as O
from refcirc import brickwork_spec, to_bt, to_ref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NOISE = ("depolarizing", 1e-3, "depolarizing", 1e-2)


def _traj_cases():
    with open(os.path.join(GOLD, "trajectories_n6.json")) as f:
        return json.load(f)


def _spec(c):
    """The fixture's op list with JSON lists turned back into spec tuples."""
    return [tuple(s) if s[0] != "ifop" else ("ifop", s[1], s[2], [tuple(a) for a in s[3]], [tuple(a) for a in s[4]])
            for s in c["spec"]]


def test_c1_probabilities_fixture():
    """C1 (N=12, 20 layers): the fixture equals the oracle's reference-
    algorithm mode (full-space sparse operators) to 1e-13."""
    spec = brickwork_spec(12, 20, 2026, gate_noise=False)
    ref = np.load(os.path.join(GOLD, "c1_probs_n12.npy"))
    s, _, _, _, _ = O.run_sv(12, to_ref(spec), None, None, mode=1)
    assert np.abs(np.abs(s) ** 2 - ref).max() < 1e-13
    assert abs(ref.sum() - 1.0) < 1e-12


def test_c1_generator_dense_bruteforce():
    """The same generator at N=7 (dense 128x128 operators): oracle == dense."""
    spec = brickwork_spec(7, 20, 2026, gate_noise=False)
    s, _, _, _, _ = O.run_sv(7, to_ref(spec), None, None)
    d, _, _ = Dn.run_sv(7, to_ref(spec))
    assert np.abs(s - d).max() < 1e-12


def test_c2_rho_fixture_dense():
    spec = brickwork_spec(5, 8, 2026, ad_gamma=5e-3)
    ref = np.load(os.path.join(GOLD, "c2_rho_n5.npy"))
    d = Dn.run_dm(5, to_ref(spec), O.ref_noise(*NOISE))
    assert np.abs(d - ref).max() < 1e-12
    assert abs(np.trace(ref).real - 1.0) < 1e-12


def test_trajectory_fixtures_dense():
    cases = _traj_cases()
    assert len(cases) >= 4
    for c in cases:
        spec = _spec(c)
        u = O.uniforms(c["seed"], 0, 16 * len(spec) + 16)
        d, outs, used = Dn.run_sv(6, to_ref(spec), O.ref_noise(*c["noise"]), u)
        assert outs == c["outcomes"] and used == c["draws"]
        assert np.abs(d - (np.array(c["state_re"]) + 1j * np.array(c["state_im"]))).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_engine_matches_fixtures(bt, prec):
    from bluetangle_jl_amd.circuit import Circuit
    tol = 1e-12 if prec == "c128" else 1e-5
    # C1 probabilities
    spec = brickwork_spec(12, 20, 2026, gate_noise=False)
    st = bt.zero_state(12, prec)
    Circuit(to_bt(spec, bt), 12).run_on(st)
    assert np.abs(st.probabilities() - np.load(os.path.join(GOLD, "c1_probs_n12.npy"))).max() < tol
    # C2-shaped rho
    spec = brickwork_spec(5, 8, 2026, ad_gamma=5e-3)
    rho = bt.zero_rho(5, prec)
    Circuit(to_bt(spec, bt), 5, bt.NoiseModel(*NOISE)).run_dm(rho)
    assert np.abs(rho.to_numpy() - np.load(os.path.join(GOLD, "c2_rho_n5.npy"))).max() < tol
    # trajectories with the fixtures' uniforms
    for c in _traj_cases():
        spec = _spec(c)
        cir = Circuit(to_bt(spec, bt), 6, bt.NoiseModel(*c["noise"]))
        st = bt.zero_state(6, prec)
        outs, _ = cir.run_on(st, seed=c["seed"], traj=0)
        assert outs == c["outcomes"]
        ref = np.array(c["state_re"]) + 1j * np.array(c["state_im"])
        assert np.allclose(st.to_numpy(), ref, rtol=0, atol=tol, equal_nan=True)
