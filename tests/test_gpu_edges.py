"""Edge cases of the engine's inputs: empty circuits on every executor, the
smallest states (N = 1, 2) through the circuit, batch and density-matrix
executors against the oracle, and the size limits (a state beyond HBM is an
out-of-memory error; more than 36 index bits, zero qubits or an empty batch
are argument errors, all before any kernel runs), and tile settings smaller
than an operator or than the state (density-matrix superoperators; a scale
pending across a reduction-only pass)."""
import gc

import numpy as np
import pytest

import oracle as O
from refcirc import random_spec, to_bt, to_ref

pytestmark = pytest.mark.gpu

TOL = {"c64": 1e-5, "c128": 1e-12}
NM = ("depolarizing", 0.05, "amplitude_damping", 0.1)


def test_empty_circuit_every_executor(bt):
    from bluetangle_jl_amd.circuit import Circuit
    N = 3
    e0 = np.zeros(2 ** N, dtype=np.complex128)
    e0[0] = 1
    c = Circuit([], N)
    for prec in ("c64", "c128"):
        st = bt.zero_state(N, prec)
        outs, _ = c.run_on(st)
        assert outs == [] and np.array_equal(st.to_numpy(), e0)
    outs, states, _ = c.run_batch(4, seed=1, prec="c128", keep_states=True)
    assert outs == [[]] * 4 and all(np.array_equal(s, e0) for s in states)
    st = bt.zero_state(N, "c64")
    stats, _ = c.run_fused(st)
    assert stats["blocks"] == 0 and np.array_equal(st.to_numpy(), e0)
    rho = bt.zero_rho(N, "c128")
    c.run_dm(rho)
    assert np.array_equal(rho.to_numpy(), np.outer(e0, e0))
    st = bt.zero_state(N, "c128")
    bt.apply(st, [])
    assert np.array_equal(st.to_numpy(), e0)


@pytest.mark.parametrize("N", [1, 2])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_smallest_states_every_executor(bt, N, prec):
    from bluetangle_jl_amd.circuit import Circuit
    kinds = ("gate1", "chan1", "meas", "reset", "ifop") if N == 1 else \
        ("gate1", "gate2", "chan1", "chan2", "meas", "reset", "ifop")
    checked = 0
    for seed in range(4):
        rng = np.random.default_rng(100 * N + seed)
        spec = random_spec(N, 30, rng, kinds=kinds)
        c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*NM))
        u = O.uniforms(seed, 0, 8 * len(spec) + 16)
        r, routs, _, bnd, _ = O.run_sv(N, to_ref(spec), O.ref_noise(*NM), u, tol=1e-6)
        if bnd:
            continue
        st = bt.zero_state(N, prec)
        outs, _ = c.run_on(st, seed=seed, traj=0)
        assert outs == routs
        assert np.allclose(st.to_numpy(), r, rtol=0, atol=TOL[prec], equal_nan=True)
        bouts, states, _ = c.run_batch(3, seed=seed, prec=prec, keep_states=True)
        assert bouts[0] == routs
        assert np.allclose(states[0], r, rtol=0, atol=TOL[prec], equal_nan=True)
        checked += 1
    assert checked >= 2
    # density matrix: the gate and channel ops of a circuit without measurements
    rng = np.random.default_rng(7 + N)
    spec = random_spec(N, 20, rng, kinds=("gate1", "gate2", "chan1", "chan2") if N == 2 else ("gate1", "chan1"))
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*NM))
    rho = bt.zero_rho(N, prec)
    c.run_dm(rho)
    want = O.run_dm(N, to_ref(spec), O.ref_noise(*NM))
    assert np.abs(rho.to_numpy() - want).max() < TOL[prec]


def test_size_limits(bt):
    from bluetangle_jl_amd import _lib as L
    from bluetangle_jl_amd.circuit import Circuit
    with pytest.raises(L.BTError) as e:
        bt.zero_state(36, "c64")  # 512 GiB: more than the device holds
    assert e.value.code == -4
    with pytest.raises(L.BTError) as e:
        bt.zero_state(37, "c64")
    assert e.value.code == -1
    with pytest.raises(L.BTError) as e:
        bt.zero_rho(19, "c64")  # 38 index bits
    assert e.value.code == -1
    with pytest.raises(L.BTError) as e:
        bt.zero_state(0, "c64")
    assert e.value.code == -1
    c = Circuit([bt.Op("H", 1)], 2)
    with pytest.raises(L.BTError) as e:
        c.run_batch(0, seed=1)
    assert e.value.code == -1
    # the device is still usable after the refused requests
    st = bt.zero_state(2, "c128")
    c.run_on(st)
    assert abs(st.to_numpy()[2] - 2 ** -0.5) < 1e-15
    del st
    gc.collect()


@pytest.fixture
def tile_bits(bt):
    from bluetangle_jl_amd import _lib
    L = _lib.lib()

    def set_tiles(T):
        _lib.check(L.bt_set_tile_bits(0, T))
        _lib.check(L.bt_set_tile_bits(1, T))
    yield set_tiles
    set_tiles(0)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_density_matrix_small_tile_setting(bt, tile_bits, prec):
    """A tile setting smaller than a superoperator (4 bits of vec(rho)) is
    raised by the density-matrix executors, alone and through the server."""
    from bluetangle_jl_amd.circuit import Circuit, Server
    tile_bits(3)
    rng = np.random.default_rng(77)
    N = 4
    spec = random_spec(N, 30, rng, kinds=("gate1", "gate2", "chan1", "chan2"))
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*NM))
    want = O.run_dm(N, to_ref(spec), O.ref_noise(*NM))
    rho = bt.zero_rho(N, prec)
    c.run_dm(rho)
    assert np.abs(rho.to_numpy() - want).max() < TOL[prec]
    srv = Server()
    try:
        rid = srv.submit(c, "dm", prec)
        done = []
        while not done:
            done = srv.poll(wait=True)
        _, probs = srv.fetch(rid)
    finally:
        srv.close()
    assert np.abs(probs - np.real(np.diag(want))).max() < TOL[prec]


@pytest.mark.parametrize("T", [3, 4])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_pending_scale_with_small_tiles(bt, tile_bits, T, prec):
    """A lazily decided measurement branch, then the reference's T gate (not
    exactly unitary): the scale pending across a reduction-only pass is
    applied (every executor, tiles smaller than the state)."""
    from bluetangle_jl_amd.circuit import Circuit
    from test_interp_emu import PENDING_SCALE_SPEC as spec
    tile_bits(T)
    N = 9
    nm = ("depolarizing", 0.05, "depolarizing", 0.1)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm))
    for mode in ("rounds", "persistent"):
        outs, states, _ = c.run_batch(6, seed=109, prec=prec, keep_states=True, mode=mode)
        for t in range(6):
            u = O.uniforms(109, t, 16 * len(spec) + 16)
            r, routs, _, bnd, _ = O.run_sv(N, to_ref(spec), O.ref_noise(*nm), u, tol=1e-6)
            assert not bnd
            assert outs[t] == routs, (mode, t)
            assert np.abs(states[t] - r).max() < TOL[prec], (mode, t)
    st = bt.zero_state(N, prec)
    o1, _ = c.run_on(st, seed=109, traj=0)
    u = O.uniforms(109, 0, 16 * len(spec) + 16)
    r, routs, _, _, _ = O.run_sv(N, to_ref(spec), O.ref_noise(*nm), u, tol=1e-6)
    assert o1 == routs and np.abs(st.to_numpy() - r).max() < TOL[prec]


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_annihilating_branch_gives_nan(bt, prec):
    """The qubit > target ordering quirk picks, for trajectories 4 and 6 of
    this circuit, a Kraus branch that annihilates the state: NaN everywhere,
    as the reference's normalize gives, through the per-op path and every
    executor; the other trajectories match the oracle."""
    from bluetangle_jl_amd.circuit import Circuit
    from test_interp_emu import ANNIHILATING_BRANCH_SPEC as spec
    N = 6
    nm = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm))
    refs = []
    for t in range(8):
        u = O.uniforms(158, t, 16 * len(spec) + 16)
        r, routs, _, bnd, _ = O.run_sv(N, to_ref(spec), O.ref_noise(*nm), u, tol=1e-6)
        assert not bnd
        refs.append((r, routs))
    assert [t for t in range(8) if np.isnan(refs[t][0]).all()] == [4, 6]
    for mode in ("rounds", "persistent"):
        outs, states, _ = c.run_batch(8, seed=158, prec=prec, keep_states=True, mode=mode)
        for t in range(8):
            assert outs[t] == refs[t][1], (mode, t)
            assert np.array_equal(np.isnan(states[t]), np.isnan(refs[t][0])), (mode, t)
            assert np.allclose(states[t], refs[t][0], rtol=0, atol=TOL[prec], equal_nan=True), (mode, t)
    for t in (4, 5, 6):
        st = bt.zero_state(N, prec)
        o1, _ = c.run_on(st, seed=158, traj=t)
        assert o1 == refs[t][1] and np.array_equal(np.isnan(st.to_numpy()), np.isnan(refs[t][0]))
        bt.seed(158, t)
        st2 = bt.zero_state(N, prec)
        st2, o2 = bt.apply(st2, to_bt(spec, bt), noise=bt.NoiseModel(*nm), track_measurements=True)
        assert o2 == refs[t][1] and np.array_equal(np.isnan(st2.to_numpy()), np.isnan(refs[t][0])), t
