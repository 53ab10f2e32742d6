"""Random noisy circuits of every op kind (N = 3..12, six noise models, tile
settings 0 / 2..8) through every trajectory executor — batched rounds,
persistent batch, one-circuit executor, op-by-op apply — against the oracle
on the same uniforms, and their gate/channel part through the density-matrix
executor against the oracle's rho.  Outcomes identical (draws within 1e-6 of
a cumulative boundary are counted and skipped), amplitudes within 1e-12
(complex128) / 1e-5 (complex64), NaN where the reference's normalize gives
NaN."""
import numpy as np
import pytest

import oracle as O
from refcirc import random_spec, to_bt, to_ref

pytestmark = pytest.mark.gpu

KINDS = ("gate1", "gate2", "ccx", "chan1", "chan2", "meas", "reset", "ifop")
NOISES = [("depolarizing", 0.05, "amplitude_damping", 0.1), ("amplitude_damping", 0.2, "depolarizing", 0.1),
          ("bit_flip", 0.1, "phase_damping", 0.2), ("phase_flip", 0.1, "bit_flip", 0.05),
          ("depolarizing_amp", 0.1, "amplitude_damping", 0.3), ("rot_xyz", 0.2, "depolarizing", 0.02)]
TILES = [0, 2, 3, 4, 5, 6, 8]
TOL = {"c128": 1e-12, "c64": 1e-5}


@pytest.fixture
def tile_bits(bt):
    from bluetangle_jl_amd import _lib
    L = _lib.lib()

    def set_tiles(T):
        _lib.check(L.bt_set_tile_bits(0, T))
        _lib.check(L.bt_set_tile_bits(1, T))
    yield set_tiles
    set_tiles(0)


def close(a, b, prec):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.allclose(a, b, rtol=0, atol=TOL[prec], equal_nan=True)


@pytest.mark.parametrize("seed", range(36))
def test_random_circuit_every_executor(bt, tile_bits, seed):
    from bluetangle_jl_amd.circuit import Circuit
    rng = np.random.default_rng(50_000 + seed)
    N = int(rng.integers(3, 13))
    nm = NOISES[seed % len(NOISES)]
    spec = random_spec(N, int(rng.integers(20, 80)), rng, kinds=KINDS)
    T = TILES[seed % len(TILES)]
    tile_bits(min(T, N) if T else 0)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm))
    refs = []
    for t in range(4):
        u = O.uniforms(seed, t, 16 * len(spec) + 16)
        r, routs, _, bnd, _ = O.run_sv(N, to_ref(spec), O.ref_noise(*nm), u, tol=1e-6)
        refs.append((r, routs, bnd))
    for prec in ("c128", "c64"):
        for mode in ("rounds", "persistent"):
            outs, states, _ = c.run_batch(4, seed=seed, prec=prec, keep_states=True, mode=mode)
            for t, (r, routs, bnd) in enumerate(refs):
                if bnd:
                    continue
                assert outs[t] == routs, (prec, mode, t)
                assert close(states[t], r, prec), (prec, mode, t)
        r, routs, bnd = refs[0]
        if bnd:
            continue
        st = bt.zero_state(N, prec)
        o1, _ = c.run_on(st, seed=seed, traj=0)
        assert o1 == routs and close(st.to_numpy(), r, prec), (prec, "run_on")
        bt.seed(seed, 0)
        st2, o2 = bt.apply(bt.zero_state(N, prec), to_bt(spec, bt), noise=bt.NoiseModel(*nm),
                           track_measurements=True)
        assert o2 == routs and close(st2.to_numpy(), r, prec), (prec, "apply")
    dspec = [x for x in spec if x[0] in ("gate", "chan")]
    if dspec and N <= 7:
        cd = Circuit(to_bt(dspec, bt), N, bt.NoiseModel(*nm))
        want = O.run_dm(N, to_ref(dspec), O.ref_noise(*nm))
        for prec in ("c128", "c64"):
            rho = bt.zero_rho(N, prec)
            cd.run_dm(rho)
            assert np.abs(rho.to_numpy() - want).max() < 10 * TOL[prec], (prec, "dm")


KRAUS_KINDS = ("gate1", "gate2", "ccx", "kraus", "kraus", "meas", "reset", "ifop")
KRAUS_TILES = [0, 3, 4, 5, 8]


@pytest.mark.parametrize("seed", range(12))
def test_random_custom_kraus_every_executor(bt, tile_bits, seed):
    """OpQC with random custom Kraus lists on 1, 2 and 3 qubits (struct.jl:182-234:
    unitary mixtures, general CPTP lists whose branch needs the reduced rho
    over 1-3 qubits, tensored amplitude damping; 3 qubits act on
    qubit..qubit+2, hilbert.jl:106-128, struct.jl:44-55) through every
    trajectory executor against the oracle; the density-matrix executor runs
    the 1- and 2-qubit part and raises for a 3-qubit channel, which has no
    density-matrix method in the reference."""
    from bluetangle_jl_amd.circuit import Circuit
    rng = np.random.default_rng(60_000 + seed)
    N = int(rng.integers(3, 11))
    nm = NOISES[seed % len(NOISES)]
    spec = random_spec(N, int(rng.integers(20, 60)), rng, kinds=KRAUS_KINDS)
    if not any(s[0] == "kraus" and len(s[1][0]) == 8 for s in spec):
        spec.append(("kraus", __import__("refcirc").random_kraus(3, rng, "general"), 1, -1))
    T = KRAUS_TILES[seed % len(KRAUS_TILES)]
    tile_bits(min(T, N) if T else 0)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm))
    refs = []
    for t in range(4):
        u = O.uniforms(seed, t, 16 * len(spec) + 16)
        r, routs, _, bnd, _ = O.run_sv(N, to_ref(spec), O.ref_noise(*nm), u, tol=1e-6)
        refs.append((r, routs, bnd))
    for prec in ("c128", "c64"):
        for mode in ("rounds", "persistent"):
            outs, states, _ = c.run_batch(4, seed=seed, prec=prec, keep_states=True, mode=mode)
            for t, (r, routs, bnd) in enumerate(refs):
                if bnd:
                    continue
                assert outs[t] == routs, (prec, mode, t)
                assert close(states[t], r, prec), (prec, mode, t)
        r, routs, bnd = refs[0]
        if bnd:
            continue
        st = bt.zero_state(N, prec)
        o1, _ = c.run_on(st, seed=seed, traj=0)
        assert o1 == routs and close(st.to_numpy(), r, prec), (prec, "run_on")
        bt.seed(seed, 0)
        st2, o2 = bt.apply(bt.zero_state(N, prec), to_bt(spec, bt), noise=bt.NoiseModel(*nm),
                           track_measurements=True)
        assert o2 == routs and close(st2.to_numpy(), r, prec), (prec, "apply")
    dspec = [x for x in spec if x[0] == "gate" or (x[0] == "kraus" and len(x[1][0]) < 8)]
    if N <= 6:
        cd = Circuit(to_bt(dspec, bt), N, bt.NoiseModel(*nm))
        want = O.run_dm(N, to_ref(dspec), O.ref_noise(*nm))
        for prec in ("c128", "c64"):
            rho = bt.zero_rho(N, prec)
            cd.run_dm(rho)
            assert np.abs(rho.to_numpy() - want).max() < 10 * TOL[prec], (prec, "dm")
        c3 = Circuit(to_bt([x for x in spec if x[0] == "kraus" and len(x[1][0]) == 8][:1], bt), N)
        with pytest.raises(Exception):
            c3.run_dm(bt.zero_rho(N, "c64"))
        with pytest.raises(Exception):
            bt.apply(bt.zero_rho(N, "c64"), to_bt([x for x in spec if x[0] == "kraus" and len(x[1][0]) == 8][:1], bt))
