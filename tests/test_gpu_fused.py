"""GPU parity of the dense-block path: bt_apply_fused5 (5-bit blocks on FP32
MFMA) against a NumPy tensor contraction, and bt_run_fused (gate-only
circuits fused into blocks) against the FP64 oracle.  Tolerances: complex64
1e-5, complex128 1e-12 (max amplitude error)."""
import numpy as np
import pytest

import oracle as O
from refcirc import brickwork_spec, random_spec, to_bt, to_ref

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def apply_dense(psi, N, bits, m):
    """psi <- M psi, M acting on bits (bits[0] = MSB of M's index)."""
    k = len(bits)
    t = psi.reshape((2,) * N)
    axes = [N - 1 - b for b in bits]
    mt = m.reshape((2,) * (2 * k))
    out = np.tensordot(mt, t, axes=(list(range(k, 2 * k)), axes))
    out = np.moveaxis(out, list(range(k)), axes)
    return out.reshape(-1)


def rand_unitary(rng, D):
    q, _ = np.linalg.qr(rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D)))
    return q


@pytest.mark.parametrize("N,bits", [(12, [4, 3, 2, 1, 0]), (12, [11, 10, 9, 8, 7]), (13, [0, 12, 5, 7, 2]),
                                    (16, [15, 3, 9, 0, 6]), (16, [7, 8, 9, 10, 11]), (18, [17, 16, 0, 1, 9])])
def test_fused5_block(bt, N, bits):
    from bluetangle_jl_amd import _lib as L
    rng = np.random.default_rng(N * 100 + bits[0])
    m = rand_unitary(rng, 32)
    v = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
    v /= np.linalg.norm(v)
    st = bt.from_numpy(v, "c64")
    mm = np.ascontiguousarray(m.T).reshape(-1).view(np.float64).copy()  # column-major complex128
    b = np.array(bits, dtype=np.int32)
    L.check(L.lib().bt_apply_fused5(st.handle, L.dptr(mm), L.iptr(b)))
    ref = apply_dense(v, N, bits, m)
    assert np.abs(st.to_numpy() - ref).max() < 1e-5


@pytest.mark.parametrize("prec,max_k", [("c64", 5), ("c64", 4), ("c64", 2), ("c128", 5), ("c128", 3)])
def test_run_fused_random_circuits(bt, prec, max_k):
    from bluetangle_jl_amd.circuit import Circuit
    for seed in range(3):
        rng = np.random.default_rng(9100 + seed)
        N = 12 + seed
        spec = random_spec(N, 80, rng, kinds=("gate1", "gate2", "ccx") if max_k >= 3 else ("gate1", "gate2"))
        c = Circuit(to_bt(spec, bt), N)
        v = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
        v /= np.linalg.norm(v)
        st = bt.from_numpy(v, prec)
        stats, norms = c.run_fused(st, max_k=max_k, norm_every=4)
        r, _, _, _, _ = O.run_sv(N, to_ref(spec), None, None, state=v)
        assert np.abs(st.to_numpy() - r).max() < TOL[prec] * 10, (seed, np.abs(st.to_numpy() - r).max())
        assert 0 < stats["blocks"] < len(spec)
        assert len(norms) == stats["blocks"] // 4
        # the reference's T gate has 10-digit entries (gates.jl), so it is unitary only to ~1e-10
        assert all(abs(x - 1.0) < (1e-5 if prec == "c64" else 1e-8) for x in norms)


def test_run_fused_brickwork_matches_tile_engine(bt):
    """C5-shaped brickwork (RX/RY/RZ + CNOT) at N=16: fused MFMA blocks, the
    register-tile executor and the oracle agree."""
    from bluetangle_jl_amd.circuit import Circuit
    N = 16
    spec = brickwork_spec(N, 6, 5, gate_noise=False)
    c = Circuit(to_bt(spec, bt), N)
    a = bt.zero_state(N, "c64")
    stats, _ = c.run_fused(a, max_k=5)
    b = bt.zero_state(N, "c64")
    c.run_on(b)
    r, _, _, _, _ = O.run_sv(N, to_ref(spec), None, None)
    assert np.abs(a.to_numpy() - r).max() < 1e-5
    assert np.abs(b.to_numpy() - r).max() < 1e-5
    assert stats["blocks"] * 5 < len(spec)


def test_fused_rejects_noisy_circuits(bt):
    from bluetangle_jl_amd.circuit import Circuit
    c = Circuit([bt.Op("H", 1), bt.Op("MZ", 1)], 12)
    with pytest.raises(Exception):
        c.run_fused(bt.zero_state(12, "c64"))


def test_inner_across_precisions(bt):
    rng = np.random.default_rng(3)
    N = 14
    v = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
    w = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
    a64, a128 = bt.from_numpy(v, "c64"), bt.from_numpy(v, "c128")
    b128 = bt.from_numpy(w, "c128")
    assert abs(a128.inner(b128) - np.vdot(v, w)) < 1e-9 * np.abs(v).sum()
    assert abs(a64.inner(a128) - np.vdot(v, v)) < 1e-5 * np.vdot(v, v).real
    assert abs(a128.inner(a128).real - a128.norm2()) < 1e-9


@pytest.mark.parametrize("prec,max_k", [("c64", 5), ("c64", 3), ("c128", 3)])
def test_captured_graph_replays(bt, prec, max_k):
    """A captured fused circuit replayed several times, with eager work on the
    same stream between replays, matches the eager runs each time."""
    from bluetangle_jl_amd.circuit import Circuit
    N = 14
    spec = brickwork_spec(N, 6, 8, gate_noise=False)
    c = Circuit(to_bt(spec, bt), N)
    eager = bt.zero_state(N, prec)
    g_state = bt.zero_state(N, prec)
    g = c.capture(g_state, max_k=max_k)
    assert g.passes > 0
    tol = 1e-5 if prec == "c64" else 1e-12
    for rep in range(3):
        c.run_fused(eager, max_k=max_k)      # eager reference, applied once more each round
        g.launch()
        g_state.sync()
        n = g_state.norm2()                  # eager work between replays
        assert abs(n - eager.norm2()) < 10 * tol
        assert np.abs(g_state.to_numpy() - eager.to_numpy()).max() < 10 * tol, rep
    g.close()


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_fusion_width_autotuner(bt, prec):
    """bt_autotune_fused times max_k = 3, 4, 5 and the register-tile executor
    on a scratch state and records the fastest on the circuit; run_fused with
    max_k = 0 runs it, with the oracle's result (the caller's state is not
    touched by the tuning)."""
    from bluetangle_jl_amd.circuit import Circuit
    N = 14
    spec = brickwork_spec(N, 6, 31)
    c = Circuit(to_bt(spec, bt), N)
    st = bt.zero_state(N, prec)
    st.set_basis(5)
    k, ms = c.autotune_fused(st)
    assert k in (0, 3, 4, 5)
    assert ms[0] > 0 and ms[1] > 0 and ms[3] > 0
    assert (ms[2] > 0) == (prec == "c64")  # MFMA blocks: complex64 only
    assert np.abs(st.to_numpy() - np.eye(1, 2 ** N, 5)[0]).max() == 0
    st.set_basis(0)
    c.run_fused(st, max_k=0)
    r, _, _, _, _ = O.run_sv(N, to_ref(spec), None)
    assert np.abs(st.to_numpy() - r).max() < TOL[prec]


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_density_matrix_graph(bt, prec):
    """to_rho captured as one HIP graph (bt_graph_create_dm): every replay,
    with eager work and synchronisations between the replays, equals the
    eager density-matrix executor bit for bit and the oracle's rho; a circuit
    with a measurement is rejected."""
    from bluetangle_jl_amd.circuit import Circuit
    from refcirc import random_kraus
    rng = np.random.default_rng(77)
    N = 6
    spec = random_spec(N, 40, rng, kinds=("gate1", "gate2", "ccx", "chan1", "chan2"))
    spec += [("kraus", random_kraus(1, rng, "general"), 2, -1), ("kraus", random_kraus(2, rng, "general"), 3, 5)]
    spec += brickwork_spec(N, 3, 4, ad_gamma=0.05)
    nm = ("depolarizing", 0.02, "amplitude_damping", 0.05)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel(*nm))
    eager = bt.zero_rho(N, prec)
    c.run_dm(eager)
    want = eager.to_numpy()
    ref = O.run_dm(N, to_ref(spec), O.ref_noise(*nm))
    assert np.abs(want - ref).max() < 10 * TOL[prec]
    rho = bt.zero_rho(N, prec)
    g = c.capture_dm(rho)
    assert g.passes > 0
    other = bt.zero_state(N + 2, prec)
    for rep in range(4):
        rho.set_basis(0)
        g.launch()
        rho.sync()
        assert np.array_equal(rho.to_numpy(), want), rep
        bt.apply(other, bt.Op("H", 1 + rep % (N + 2)))  # eager work between replays
        other.sync()
    g.close()
    cm = Circuit(to_bt([("gate", "H", 1, -1, -2, True), ("meas", "MZ", 1)], bt), N)
    with pytest.raises(Exception):
        cm.capture_dm(bt.zero_rho(N, prec))
