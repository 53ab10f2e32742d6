"""GPU parity of the circuit executors against the FP64 oracle:
bt_run_circuit (one trajectory, grid tile passes), bt_run_batch (persistent
batched trajectories), bt_run_circuit_dm (to_rho).

Tolerances: complex128 1e-12, complex64 1e-5 (max amplitude / entry
error); outcomes identical unless the oracle counts a draw within 1e-6 of a
cumulative-probability boundary."""
import numpy as np
import pytest

import oracle as O
from refcirc import brickwork_spec, random_spec, to_bt, to_ref

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}
KINDS = ("gate1", "gate2", "ccx", "chan1", "chan2", "meas", "reset", "ifop")


@pytest.fixture
def tiles(bt):
    from bluetangle_jl_amd import _lib
    L = _lib.lib()

    def set_tiles(T):
        _lib.check(L.bt_set_tile_bits(0, T))
        _lib.check(L.bt_set_tile_bits(1, T))
    yield set_tiles
    set_tiles(0)


def oracle_traj(N, spec, nm_args, seed, traj):
    nm = O.ref_noise(*nm_args) if nm_args else None
    u = O.uniforms(seed, traj, 16 * len(spec) + 16)
    s, outs, used, bnd, _ = O.run_sv(N, to_ref(spec), nm, u, tol=1e-6)
    return s, outs, bnd


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("T", [0, 3])
def test_run_circuit_random(bt, tiles, prec, T):
    from bluetangle_jl_amd.circuit import Circuit
    tiles(T)
    for seed in range(6):
        rng = np.random.default_rng(300 + seed)
        N = 4 + seed % 4
        spec = random_spec(N, 50, rng, kinds=KINDS)
        nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
        c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
        st = bt.zero_state(N, prec)
        outs, stats = c.run_on(st, seed=seed, traj=0)
        r, routs, bnd = oracle_traj(N, spec, nm_args, seed, 0)
        if bnd:
            continue
        s = st.to_numpy()
        assert outs == routs, (seed, prec, T)
        assert np.array_equal(np.isnan(s), np.isnan(r))
        assert np.allclose(s, r, rtol=0, atol=TOL[prec], equal_nan=True), (seed, np.nanmax(np.abs(s - r)))


@pytest.mark.parametrize("mode", ["rounds", "persistent"])
@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("T", [0, 5])
def test_run_batch_matches_per_trajectory_oracle(bt, tiles, prec, T, mode):
    from bluetangle_jl_amd.circuit import Circuit
    tiles(T)
    N = 8
    spec = brickwork_spec(N, 6, 11, ad_gamma=5e-3, mz_prob=0.25)
    nm_args = ("depolarizing", 1e-3, "depolarizing", 1e-2)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
    ntraj = 48
    outs, states, stats = c.run_batch(ntraj, seed=21, prec=prec, keep_states=True, mode=mode)
    assert stats["passes"] > 0
    skipped = 0
    for t in range(ntraj):
        r, routs, bnd = oracle_traj(N, spec, nm_args, 21, t)
        if bnd:
            skipped += 1
            continue
        assert outs[t] == routs, t
        assert np.abs(states[t] - r).max() < TOL[prec], (t, np.abs(states[t] - r).max())
    assert skipped < ntraj // 4


@pytest.mark.parametrize("mode", ["rounds", "persistent"])
def test_run_batch_random_ops(bt, tiles, mode):
    """Every op kind through both batched executors (c128)."""
    from bluetangle_jl_amd.circuit import Circuit
    tiles(3)
    rng = np.random.default_rng(7)
    N = 6
    spec = random_spec(N, 60, rng, kinds=KINDS)
    nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
    outs, states, _ = c.run_batch(16, seed=3, prec="c128", keep_states=True, mode=mode)
    for t in range(16):
        r, routs, bnd = oracle_traj(N, spec, nm_args, 3, t)
        if bnd:
            continue
        assert outs[t] == routs
        assert np.allclose(states[t], r, rtol=0, atol=1e-12, equal_nan=True)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_run_circuit_dm_random(bt, tiles, prec):
    from bluetangle_jl_amd.circuit import Circuit
    for T, seed in [(0, 1), (4, 2), (0, 3)]:
        tiles(T)
        rng = np.random.default_rng(400 + seed)
        N = 3 + seed
        spec = random_spec(N, 40, rng, kinds=("gate1", "gate2", "ccx", "chan1", "chan2", "reset"))
        nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
        c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
        rho = bt.zero_rho(N, prec)
        c.run_dm(rho)
        ref = O.run_dm(N, to_ref(spec), O.ref_noise(*nm_args))
        assert np.abs(rho.to_numpy() - ref).max() < TOL[prec], (T, seed)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_kept_plans_follow_the_tiling(bt, tiles, prec):
    """A density-matrix run and a gate-only state-vector run keep their pass
    plans on the circuit (and the executor's staging on the state); changing
    the tiling plans them again: one circuit and one state run at tile
    settings 0, 4, 9 and 0 again match the oracle every time, and the pass
    count follows the tiling."""
    from bluetangle_jl_amd.circuit import Circuit
    rng = np.random.default_rng(480)
    N = 6
    spec = random_spec(N, 50, rng, kinds=("gate1", "gate2", "ccx", "chan1", "chan2"))
    nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
    ref = O.run_dm(N, to_ref(spec), O.ref_noise(*nm_args))
    rho = bt.zero_rho(N, prec)
    G = 9
    gspec = random_spec(G, 60, rng, kinds=("gate1", "gate2", "ccx"))
    g = Circuit(to_bt(gspec, bt), G, False)
    gref, _, _, _, _ = O.run_sv(G, to_ref(gspec), None, O.uniforms(0, 0, 16), tol=1e-6)
    psi = bt.zero_state(G, prec)
    passes = {}
    for T in (0, 4, 9, 0):
        tiles(T)
        st = c.run_dm(rho.set_basis(0))
        assert np.abs(rho.to_numpy() - ref).max() < TOL[prec], T
        _, sg = g.run_on(psi.set_basis(0))
        assert np.abs(psi.to_numpy() - gref).max() < TOL[prec], T
        passes.setdefault(T, []).append((st["passes"], sg["passes"]))
    assert passes[0][0] == passes[0][1]
    assert passes[4][0][0] > passes[9][0][0] and passes[4][0][1] > passes[9][0][1]


def test_c2_density_matrix_config(bt):
    """Config 2: rho N=8, brickwork 16 layers, depolarizing 1e-3 / 1e-2 and
    amplitude damping 5e-3 per qubit per layer; complex64 rho vs the
    ComplexF64 oracle, FP64 trace and purity."""
    from bluetangle_jl_amd.circuit import Circuit
    N = 8
    spec = brickwork_spec(N, 16, 2, ad_gamma=5e-3)
    nm_args = ("depolarizing", 1e-3, "depolarizing", 1e-2)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
    rho = bt.zero_rho(N, "c64")
    stats = c.run_dm(rho)
    ref = O.run_dm(N, to_ref(spec), O.ref_noise(*nm_args))
    got = rho.to_numpy()
    assert np.abs(got - ref).max() < 1e-5
    tr, pur = rho.trace_purity()
    assert abs(tr - np.trace(ref).real) < 1e-5
    assert abs(pur - np.sum(np.abs(ref) ** 2)) < 1e-5
    assert stats["passes"] > 0


def test_c1_to_state(bt):
    from bluetangle_jl_amd.circuit import Circuit, to_state
    N = 12
    spec = brickwork_spec(N, 20, 2026)
    st = to_state(Circuit(to_bt(spec, bt), N), prec="c128")
    r, _, _ = oracle_traj(N, spec, None, 0, 0)
    assert np.abs(st.to_numpy() - r).max() < 1e-12
    assert np.abs(st.probabilities() - np.abs(r) ** 2).max() < 1e-13


def test_run_driver_kats(bt):
    """run(ops, shots) returns the reference's KATs (runtests.jl:195-231)."""
    from bluetangle_jl_amd.circuit import run
    ops_reset = [bt.Op("X", 1), bt.Op("RES", 1), bt.Op("MZ", 1)]
    ops_ifop = [bt.Op("X", 1), bt.ifOp("MZ", 1, bt.Op("I", 1), bt.Op("X", 1))]
    ops_mid = [bt.Op("X", 1), bt.Op("X", 2), bt.Op("CX", 2, 3), bt.Op("CX", 1, 2), bt.Op("X", 2), bt.Op("MZ", 2),
               bt.Op("RES", 2), bt.Op("CX", 2, 3), bt.Op("CX", 1, 2), bt.Op("CX", 2, 3), bt.Op("X", 1)]
    for batched in (True, False):
        assert run(ops_reset, 3, batched=batched) == [[0]] * 3
        assert run(ops_ifop, 3, batched=batched) == [[1]] * 3
        assert run(ops_mid, 3, batched=batched) == [[1]] * 3


def test_autotuner_keeps_results(bt):
    """The autotuner picks a valid tuning; trajectories run under it still
    match the oracle (tuning changes the pass plan, not the results)."""
    from bluetangle_jl_amd import _lib
    from bluetangle_jl_amd.circuit import Circuit
    N = 12
    spec = brickwork_spec(N, 5, 3, ad_gamma=5e-3, mz_prob=0.2)
    nm_args = ("depolarizing", 1e-3, "depolarizing", 1e-2)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
    try:
        t = c.autotune("c64", sample=8)
        # N=12 is one tile per trajectory: every candidate runs one tile per workgroup
        assert 6 <= t["tile_bits"] <= 13 and t["coalesce_bits"] in (3, 4, 5) and t["tiles_per_block"] == 1
        assert t["ms_per_traj"] > 0
        outs, states, _ = c.run_batch(12, seed=5, prec="c64", keep_states=True)
        for tr in range(12):
            r, routs, bnd = oracle_traj(N, spec, nm_args, 5, tr)
            if bnd:
                continue
            assert outs[tr] == routs
            assert np.abs(states[tr] - r).max() < 1e-5
    finally:
        _lib.check(_lib.lib().bt_set_tuning(0, 0, 0, 0))


def test_autotuner_times_what_it_installs(bt):
    """At N=23 with 64 trajectories per stage-2 sample the engine's caps do
    not bind: the chosen tiles per workgroup is one of the timed candidates,
    installed as is."""
    from bluetangle_jl_amd import _lib
    from bluetangle_jl_amd.circuit import Circuit
    import ctypes
    N = 23
    spec = brickwork_spec(N, 2, 4, ad_gamma=5e-3, mz_prob=0.1)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel("depolarizing", 1e-3, "depolarizing", 1e-2))
    try:
        t = c.autotune("c64", sample=16)
        assert t["tiles_per_block"] in (4, 8, 16, 32), t
        got = [ctypes.c_int() for _ in range(3)]
        _lib.check(_lib.lib().bt_get_tuning(0, *[ctypes.byref(g) for g in got]))
        assert [g.value for g in got] == [t["tile_bits"], t["coalesce_bits"], t["tiles_per_block"]]
    finally:
        _lib.check(_lib.lib().bt_set_tuning(0, 0, 0, 0))


def test_autotuner_error_restores_tuning(bt):
    """A circuit that no pass can hold (a 4-bit gate with a control inside the
    tile: five operand bits) makes the first candidate fail: the autotuner
    returns that error and leaves the tuning as it was."""
    from bluetangle_jl_amd import _lib as L
    from bluetangle_jl_amd.circuit import IR
    import ctypes
    ir = IR(6)
    ir.ops = [ir.op(L.BT_OP_GATE, [0, 1, 2, 3], ctrl=1 << 4, mat=ir.add_matrix(np.eye(16)))]
    ir.ntop = 1
    arr, pool = ir.arrays()
    h = ctypes.c_void_p()
    L.check(L.lib().bt_circuit_create(6, arr, 1, 1, L.dptr(pool), pool.size // 2, ctypes.byref(h)))
    try:
        L.check(L.lib().bt_set_tuning(0, 12, 4, 16))
        t = L.Tuning()
        rc = L.lib().bt_autotune_batch(h, 0, 4, 1, ctypes.byref(t))
        assert rc != 0
        got = [ctypes.c_int() for _ in range(3)]
        L.check(L.lib().bt_get_tuning(0, *[ctypes.byref(g) for g in got]))
        assert [g.value for g in got] == [12, 4, 16]
    finally:
        L.check(L.lib().bt_set_tuning(0, 0, 0, 0))
        L.lib().bt_circuit_destroy(h)


@pytest.mark.parametrize("tpb", [3, 5])
def test_rounds_with_partial_workgroups(bt, tpb):
    """Tiles per workgroup that do not divide a trajectory's tiles (256 tiles
    of 2^6 amplitudes at N=14; 3 or 5 per workgroup, below both caps of the
    engine) leave each trajectory's last workgroup partly filled: outcomes
    and amplitudes still match the oracle."""
    from bluetangle_jl_amd import _lib
    from bluetangle_jl_amd.circuit import Circuit
    N = 14
    spec = brickwork_spec(N, 6, 8, ad_gamma=5e-3, mz_prob=0.2)
    nm_args = ("depolarizing", 1e-3, "depolarizing", 1e-2)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm_args))
    try:
        _lib.check(_lib.lib().bt_set_tuning(0, 6, 3, tpb))
        outs, states, _ = c.run_batch(64, seed=9, prec="c64", keep_states=True)
        for tr in range(0, 64, 8):
            r, routs, bnd = oracle_traj(N, spec, nm_args, 9, tr)
            if bnd:
                continue
            assert outs[tr] == routs
            assert np.abs(states[tr] - r).max() < 1e-5
    finally:
        _lib.check(_lib.lib().bt_set_tuning(0, 0, 0, 0))


def _haar(rng, d):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("N", [13, 16])
def test_tile_pass_every_pair(bt, N):
    """A random unitary on every qubit and on every ordered qubit pair, plus a
    controlled one per qubit, noiseless, through bt_run_circuit: fused
    register-tile passes (complex64: the matrix-core form; complex128: vector
    FMAs) against the oracle's application of the same matrices.  At N=13 the
    passes use all 15 complex64 register masks (5 one-bit, 10 two-bit), as
    the host emulation's dump of the same circuit's pass geometries shows."""
    from bluetangle_jl_amd.circuit import Circuit
    rng = np.random.default_rng(1000 + N)
    ops_bt, ops_ref = [], []
    for q in range(1, N + 1):
        u = _haar(rng, 2)
        ops_bt.append(bt.Op("U1", u, q))
        ops_ref.append(O.ref_gate("U1", q, mat=u))
    for a in range(1, N + 1):
        for b in range(a + 1, N + 1):
            u = _haar(rng, 4)
            x, y = (a, b) if rng.random() < 0.5 else (b, a)
            ops_bt.append(bt.Op("U2", u, x, y))
            ops_ref.append(O.ref_gate("U2", x, y, mat=u))
    for q in range(1, N + 1):
        u = _haar(rng, 2)
        c = 1 + (q + N // 2) % N
        ops_bt.append(bt.Op("UC", u, q, control=c))
        ops_ref.append(O.ref_gate("UC", q, control=c, mat=u))
    r, _, _, _, _ = O.run_sv(N, ops_ref, None, O.uniforms(0, 0, 16))
    circ = Circuit(ops_bt, N, None)
    for prec in ("c64", "c128"):
        st = bt.zero_state(N, prec)
        outs, stats = circ.run_on(st, seed=0, traj=0)
        s = st.to_numpy()
        err = float(np.abs(s - r).max())
        assert err <= TOL[prec], (prec, N, err)
