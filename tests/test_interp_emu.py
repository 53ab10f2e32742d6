"""CPU checks of the engine's trajectory interpreter (bt_interp.hpp: operator
queue with fusion, tile planning, lazy/exact branch decisions, folded
renormalisation) through a host emulation of its passes, against the FP64
oracle on identical circuits and uniforms.  No GPU involved."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from refcirc import brickwork_spec, random_spec, to_bt, to_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_DIR = os.path.join(ROOT, "tests", "emu")
EMU_SO = os.path.join(EMU_DIR, "libbt_emu.so")
CSRC = os.path.join(ROOT, "bluetangle.jl_amd", "csrc")
SRCS = [os.path.join(EMU_DIR, "bt_emu.cpp"), os.path.join(CSRC, "bt_interp.hpp"), os.path.join(CSRC, "bt_circuit.hpp"),
        os.path.join(CSRC, "bt_tile.hpp"), os.path.join(CSRC, "bt_internal.hpp"), os.path.join(CSRC, "bt_fuse.hpp"),
        os.path.join(ROOT, "include", "bt.h")]
ROCM_INC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU_SO) or any(os.path.getmtime(s) > os.path.getmtime(EMU_SO) for s in SRCS):
        # host-only build: the HIP headers supply the vector types, no device code
        # built under a per-process name, then renamed: parallel test workers may build at once
        tmp = "%s.%d.tmp" % (EMU_SO, os.getpid())
        subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INC,
                        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-o", tmp, SRCS[0]], check=True)
        os.replace(tmp, EMU_SO)
    L = ctypes.CDLL(EMU_SO)
    L.emu_error.restype = ctypes.c_char_p
    L.emu_set_exec(0)
    yield L
    L.emu_set_exec(0)


def emu_run(emu, bt, spec, N, noise=False, seed=0, traj=0, T=3, cmin=1, state=None, fold=1, tol=1e-9):
    from bluetangle_jl_amd.circuit import compile_ir
    ir = compile_ir(to_bt(spec, bt), N, noise)
    arr, pool = ir.arrays()
    s = np.zeros(2 ** N, dtype=np.complex128)
    if state is None:
        s[0] = 1
    else:
        s[:] = state
    outs = np.zeros(4096, dtype=np.int32)
    nout = ctypes.c_int64()
    stats = np.zeros(5, dtype=np.int64)
    rc = emu.emu_run(N, arr, len(ir.ops), ir.ntop, pool.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(pool.size // 2),
                     ctypes.c_uint64(seed), ctypes.c_uint64(traj), T, cmin, fold,
                     s.view(np.float64).ctypes.data_as(ctypes.c_void_p), outs.ctypes.data_as(ctypes.c_void_p),
                     ctypes.c_int64(4096), ctypes.byref(nout), stats.ctypes.data_as(ctypes.c_void_p),
                     ctypes.c_double(tol))
    if rc != 0:
        raise RuntimeError("emu rc=%d: %s" % (rc, emu.emu_error().decode()))
    return s, outs[: nout.value].tolist(), stats


def oracle_run(spec, N, nm_args=None, seed=0, traj=0, state=None, tol=1e-9):
    nm = O.ref_noise(*nm_args) if nm_args else None
    u = O.uniforms(seed, traj, 16 * len(spec) + 16)
    s, outs, used, bnd, ov = O.run_sv(N, to_ref(spec), nm, u, state=state, tol=tol)
    return s, outs, used, bnd


KINDS = ("gate1", "gate2", "ccx", "chan1", "chan2", "meas", "reset", "ifop")


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("seed", range(24))
@pytest.mark.parametrize("T", [2, 3, 5])
def test_random_noisy_circuits(emu, bt, seed, T, mode):
    """Random circuits of every op kind against the oracle: mode 0 runs the
    interpreter's pass plans on the whole vector, modes 1 / 2 the device's
    register-tile passes thread by thread in complex128 / complex64."""
    rng = np.random.default_rng(1000 + seed)
    N = 3 + seed % 4
    spec = random_spec(N, 60, rng, kinds=KINDS)
    nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    tol = 1e-9 if mode < 2 else 1e-6
    r, routs, used, bnd = oracle_run(spec, N, nm_args, seed, tol=tol)
    if bnd:
        pytest.skip("boundary draw")
    emu.emu_set_exec(mode)
    try:
        s, outs, stats = emu_run(emu, bt, spec, N, bt.NoiseModel(*nm_args), seed, T=T, cmin=min(1, T), tol=tol)
    finally:
        emu.emu_set_exec(0)
    assert outs == routs
    assert stats[3] == used
    assert np.array_equal(np.isnan(s), np.isnan(r))
    assert np.allclose(s, r, rtol=0, atol=1e-11 if mode < 2 else 1e-5, equal_nan=True)


KRAUS_KINDS = ("gate1", "gate2", "kraus", "kraus", "meas", "chan1")


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("seed", range(16))
@pytest.mark.parametrize("T", [3, 5])
def test_random_custom_kraus_circuits(emu, bt, seed, T, mode):
    """OpQC with random custom Kraus lists on 1, 2 and 3 qubits (unitary
    mixtures, general CPTP lists whose branch probabilities need the reduced
    rho over 1-3 bits, tensored amplitude damping) against the oracle, through
    the pass plans and the register-tile passes (the 3-bit reduced rho is read
    from the tile image in LDS)."""
    rng = np.random.default_rng(7000 + seed)
    N = 3 + seed % 5
    spec = random_spec(N, 40, rng, kinds=KRAUS_KINDS)
    nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    tol = 1e-9 if mode < 2 else 1e-6
    r, routs, used, bnd = oracle_run(spec, N, nm_args, seed, tol=tol)
    if bnd:
        pytest.skip("boundary draw")
    emu.emu_set_exec(mode)
    try:
        s, outs, stats = emu_run(emu, bt, spec, N, bt.NoiseModel(*nm_args), seed, T=T, cmin=1, tol=tol)
    finally:
        emu.emu_set_exec(0)
    assert outs == routs
    assert stats[3] == used
    assert np.array_equal(np.isnan(s), np.isnan(r))
    assert np.allclose(s, r, rtol=0, atol=1e-11 if mode < 2 else 1e-5, equal_nan=True)


def test_three_qubit_channel_needs_exact_rho(emu, bt):
    """A general 3-qubit CPTP list (probabilities from the 3-bit reduced rho,
    struct.jl:44-55) placed where the state is entangled across the three
    qubits: the decision is exact, every branch index agrees with the oracle
    over many trajectories."""
    from refcirc import random_kraus
    rng = np.random.default_rng(99)
    N = 5
    ks = random_kraus(3, rng, "general")
    spec = [("gate", "H", 1, -1, -2, True), ("gate", "CNOT", 1, 2, -2, True), ("gate", "RY(0.7)", 3, -1, -2, True),
            ("gate", "CNOT", 2, 3, -2, True), ("gate", "RX(1.3)", 4, -1, -2, True),
            ("kraus", ks, 2, -1), ("meas", "MZ", 3), ("kraus", ks, 1, -1), ("meas", "MX", 2)]
    for mode in (0, 1, 2):
        for traj in range(12):
            tol = 1e-9 if mode < 2 else 1e-6
            r, routs, used, bnd = oracle_run(spec, N, None, 5, traj, tol=tol)
            if bnd:
                continue
            emu.emu_set_exec(mode)
            try:
                s, outs, stats = emu_run(emu, bt, spec, N, False, 5, traj, T=4, cmin=1, tol=tol)
            finally:
                emu.emu_set_exec(0)
            assert outs == routs and stats[3] == used
            assert np.allclose(s, r, rtol=0, atol=1e-11 if mode < 2 else 1e-5)


# A measurement branch whose probability is decided lazily (the state's
# scale c unknown), then a gate that is not exactly unitary (the reference's
# T, entries rounded to 10 digits): settling c needs a pass that only
# reduces, and the scale pending from before must survive it to the next
# pass that writes the state.
PENDING_SCALE_SPEC = [
    ("chan", "amplitude_damping", 0.1695030611856367, 4, 9),
    ("gate", "CX", 8, 7, 5, True),
    ("ifop", "MY", 1, [("gate", "X", 1, -1, -2, True)], [("gate", "H", 3, -1, -2, True)]),
    ("ifop", "MY", 5, [("gate", "X", 5, -1, -2, True)], [("gate", "H", 3, -1, -2, True)]),
    ("chan", "amplitude_damping", 0.15421939985210797, 2, 9),
    ("chan", "amplitude_damping", 0.08302095597305365, 3, 6),
    ("gate", "SX", 4, -1, -2, True),
    ("gate", "RX(2.432970)", 8, -1, -2, True),
    ("gate", "T", 4, -1, 1, True),
]


def emu_run_dm(emu, bt, spec, N, noise, T, cmin=1):
    from bluetangle_jl_amd.circuit import compile_ir
    ir = compile_ir(to_bt(spec, bt), N, noise)
    arr, pool = ir.arrays()
    D = 2 ** N
    rho = np.zeros(D * D, dtype=np.complex128)
    rho[0] = 1
    stats = np.zeros(5, dtype=np.int64)
    rc = emu.emu_run_dm(N, arr, len(ir.ops), ir.ntop, pool.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_int64(pool.size // 2), T, cmin, rho.view(np.float64).ctypes.data_as(ctypes.c_void_p),
                        stats.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("emu rc=%d: %s" % (rc, emu.emu_error().decode()))
    return rho.reshape((D, D), order="F")


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("T", [4, 6])
@pytest.mark.parametrize("seed", range(4))
def test_density_matrix_tile_passes(emu, bt, seed, T, mode):
    """to_rho as the density-matrix executor runs it (superoperators on the
    2N-bit vector of rho, no norm folding), with tiles smaller than the
    vector, against the oracle's rho."""
    rng = np.random.default_rng(600 + seed)
    N = 3 + seed % 3
    spec = random_spec(N, 40, rng, kinds=("gate1", "gate2", "ccx", "chan1", "chan2", "reset"))
    nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    want = O.run_dm(N, to_ref(spec), O.ref_noise(*nm_args))
    emu.emu_set_exec(mode)
    try:
        got = emu_run_dm(emu, bt, spec, N, bt.NoiseModel(*nm_args), T)
    finally:
        emu.emu_set_exec(0)
    assert np.abs(got - want).max() < (1e-12 if mode < 2 else 1e-5)


# Amplitude damping after a two-qubit gate whose qubit > target: the
# reference samples the branch from the reduced rho in (min, max) order and
# applies it with its first factor on the qubit, so for trajectories 4 and 6
# of this circuit (seed 158) it picks a branch that annihilates the state and
# normalize gives NaN everywhere.  The engine's arithmetic leaves a rounding
# residue there, which must not be renormalised into a state.
ANNIHILATING_BRANCH_SPEC = [
    ("ifop", "MY", 6, [("gate", "X", 6, -1, -2, True)], [("gate", "H", 6, -1, -2, True)]),
    ("gate", "P(4.998140)", 5, -1, -2, True),
    ("chan", "amplitude_damping", 0.08767670617724119, 3, 4),
    ("gate", "ECR", 6, 3, -2, True),
]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("T", [2, 4, 6])
def test_annihilating_branch_gives_nan(emu, bt, mode, T):
    N = 6
    nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    tol = 1e-9 if mode < 2 else 1e-6
    emu.emu_set_exec(mode)
    try:
        nan_trajs = []
        for traj in range(8):
            r, routs, used, bnd = oracle_run(ANNIHILATING_BRANCH_SPEC, N, nm_args, 158, traj, tol=tol)
            s, outs, _ = emu_run(emu, bt, ANNIHILATING_BRANCH_SPEC, N, bt.NoiseModel(*nm_args), 158, traj, T=T,
                                 cmin=1, tol=tol)
            assert outs == routs, traj
            assert np.array_equal(np.isnan(s), np.isnan(r)), traj
            assert np.allclose(s, r, rtol=0, atol=1e-11 if mode < 2 else 1e-5, equal_nan=True), traj
            if np.isnan(r).all():
                nan_trajs.append(traj)
        assert nan_trajs == [4, 6]
    finally:
        emu.emu_set_exec(0)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("T", [3, 4, 9])
def test_pending_scale_survives_reduction_only_pass(emu, bt, mode, T):
    N = 9
    nm_args = ("depolarizing", 0.05, "depolarizing", 0.1)
    tol = 1e-9 if mode < 2 else 1e-6
    emu.emu_set_exec(mode)
    try:
        for traj in range(6):
            r, routs, used, bnd = oracle_run(PENDING_SCALE_SPEC, N, nm_args, 109, traj, tol=tol)
            assert not bnd
            s, outs, _ = emu_run(emu, bt, PENDING_SCALE_SPEC, N, bt.NoiseModel(*nm_args), 109, traj, T=T, cmin=1,
                                 tol=tol)
            assert outs == routs, traj
            assert np.abs(s - r).max() < (1e-11 if mode < 2 else 1e-5), (traj, np.abs(s - r).max())
    finally:
        emu.emu_set_exec(0)


@pytest.mark.parametrize("seed", range(6))
def test_brickwork_trajectories(emu, bt, seed):
    """The trajectory configuration in miniature: brickwork, depolarizing
    1e-3 / 1e-2, amplitude damping 5e-3 per qubit per layer, MZ with p=0.1."""
    N = 8
    spec = brickwork_spec(N, 6, 77, ad_gamma=5e-3, mz_prob=0.3)
    nm_args = ("depolarizing", 1e-3, "depolarizing", 1e-2)
    r, routs, used, bnd = oracle_run(spec, N, nm_args, 5, seed)
    s, outs, stats = emu_run(emu, bt, spec, N, bt.NoiseModel(*nm_args), 5, seed, T=5, cmin=2)
    assert bnd == 0
    assert outs == routs
    assert np.allclose(s, r, rtol=0, atol=1e-12)


def test_non_unitary_gates_do_not_renormalise(emu, bt):
    """Plain gates never renormalise (hilbert.jl:505) even when the matrix is
    not unitary (P0, SP), while channels and measurements do."""
    N = 3
    spec = [("gate", "H", 1, -1, -2, True), ("gate", "H", 2, -1, -2, True), ("chan", "amplitude_damping", 0.3, 1, -1),
            ("gate", "P0", 2, -1, -2, True), ("gate", "RX(0.3)", 3, -1, -2, True),
            ("chan", "depolarizing", 0.2, 3, -1), ("gate", "SP", 1, -1, -2, True), ("gate", "H", 3, -1, -2, True)]
    for seed in range(10):
        r, routs, used, bnd = oracle_run(spec, N, None, seed)
        s, outs, stats = emu_run(emu, bt, spec, N, False, seed, T=2, cmin=1)
        assert np.allclose(s, r, atol=1e-12), seed
        assert abs(np.linalg.norm(s) - np.linalg.norm(r)) < 1e-12


def test_measure_groups_and_resets(emu, bt):
    N = 5
    spec = [("gate", "H", q, -1, -2, True) for q in range(1, 6)] + \
           [("gate", "CNOT", 1, 2, -2, True), ("gate", "CNOT", 3, 4, -2, True)] + \
           [("meas", "MZ", 1), ("meas", "MZ", 2), ("reset", 3), ("meas", "MZ", 4), ("meas", "MZ", 5),
            ("meas", "MZ", 3), ("meas", "MX", 2), ("meas", "MR", 4), ("meas", "MY", 1)]
    for seed in range(20):
        r, routs, used, bnd = oracle_run(spec, N, None, seed)
        s, outs, stats = emu_run(emu, bt, spec, N, False, seed, T=3, cmin=1)
        assert outs == routs, seed
        assert np.allclose(s, r, atol=1e-12)


def test_pass_count_per_layer(emu, bt):
    """Lazy decisions + fusion keep the trajectory brickwork near two
    read+write passes per layer."""
    N, layers = 14, 8
    spec = brickwork_spec(N, layers, 3, ad_gamma=5e-3, mz_prob=0.1)
    s, outs, stats = emu_run(emu, bt, spec, N, bt.NoiseModel("depolarizing", 1e-3, "depolarizing", 1e-2), 1, 0,
                             T=9, cmin=3)
    passes = stats[0] + 0.5 * stats[1]
    assert passes <= 3.0 * layers, stats
    r, routs, used, bnd = oracle_run(spec, N, ("depolarizing", 1e-3, "depolarizing", 1e-2), 1, 0)
    assert outs == routs and np.allclose(s, r, atol=1e-11)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("T", [2, 3, 5, 8])
def test_register_tile_pass_emulation(emu, bt, mode, seed, T):
    """The register-tile pass (bt_tile.hpp: geometry, per-thread load,
    relayout, in-register and LDS operators, FP64 reductions, store) run
    thread by thread on the CPU in complex128 (mode 1) and complex64 (mode 2)
    against the oracle."""
    rng = np.random.default_rng(5000 + seed)
    N = 4 + seed % 5
    spec = random_spec(N, 60, rng, kinds=KINDS)
    nm_args = ("depolarizing", 0.05, "amplitude_damping", 0.1)
    tol = 1e-9 if mode == 1 else 1e-6
    r, routs, used, bnd = oracle_run(spec, N, nm_args, seed, tol=tol)
    if bnd:
        pytest.skip("boundary draw")
    emu.emu_set_exec(mode)
    try:
        s, outs, stats = emu_run(emu, bt, spec, N, bt.NoiseModel(*nm_args), seed, T=T, cmin=min(1, T), tol=tol)
    finally:
        emu.emu_set_exec(0)
    assert outs == routs
    assert np.array_equal(np.isnan(s), np.isnan(r))
    assert np.allclose(s, r, rtol=0, atol=1e-11 if mode == 1 else 1e-5, equal_nan=True), np.nanmax(np.abs(s - r))


@pytest.mark.parametrize("mode", [1, 2])
def test_register_tile_pass_brickwork(emu, bt, mode):
    """Noisy brickwork with mid-circuit Z measurements at the default tile
    size (all 256 threads active, several tiles per pass)."""
    N = 14
    spec = brickwork_spec(N, 4, 3, ad_gamma=5e-3, mz_prob=0.25)
    nm_args = ("depolarizing", 1e-2, "depolarizing", 3e-2)
    tol = 1e-9 if mode == 1 else 1e-6
    r, routs, used, bnd = oracle_run(spec, N, nm_args, 4, tol=tol)
    assert not bnd
    emu.emu_set_exec(mode)
    emu.emu_template_phases.restype = ctypes.c_int64
    emu.emu_template_phases(1)
    try:
        s, outs, stats = emu_run(emu, bt, spec, N, bt.NoiseModel(*nm_args), 4, T=12 if mode == 1 else 13, cmin=1, tol=tol)
    finally:
        emu.emu_set_exec(0)
    assert outs == routs
    assert np.abs(s - r).max() < (1e-11 if mode == 1 else 1e-5)
    # the straight-line phase templates were exercised
    assert emu.emu_template_phases(1) > 0


def test_planner_reserves_a_deeper_run_when_free(emu, bt):
    """The pass planner reserves one low tile bit beyond the configured
    coalescing bits whenever the greedy pick still selects as many
    operators (128-byte runs for complex64 at 3 bits): on the benchmark's
    noisy brickwork at its size (N=22; 4 of its layers) at least 3/4 of the
    passes hold bit 3 (8 of 10; a reserve of 3 bits alone gives 6 of 10),
    without more passes than a plain 4-bit reserve needs, and the state
    still matches the oracle."""
    N = 22
    spec = brickwork_spec(N, 4, 3, ad_gamma=5e-3, mz_prob=0.1)
    nm_args = ("depolarizing", 1e-3, "depolarizing", 1e-2)
    r, routs, used, bnd = oracle_run(spec, N, nm_args, 2, tol=1e-6)
    assert not bnd
    emu.emu_deep_coalesced_passes.restype = ctypes.c_int64
    res = {}
    emu.emu_set_exec(2)  # the complex64 register-tile pass, as the device runs it
    try:
        for cmin in (3, 4):
            emu.emu_deep_coalesced_passes(1)
            s, outs, stats = emu_run(emu, bt, spec, N, bt.NoiseModel(*nm_args), 2, T=13, cmin=cmin, tol=1e-6)
            res[cmin] = (int(stats[0]), int(emu.emu_deep_coalesced_passes(1)))
            assert outs == routs
            assert np.abs(s - r).max() < 1e-5
    finally:
        emu.emu_set_exec(0)
    print("passes, passes holding the bit after the reserve:", res)
    passes3, deep3 = res[3]
    assert 4 * deep3 >= 3 * passes3, res
    assert passes3 <= res[4][0], res


def _apply_block_dense(state, N, k, bits, m):
    """psi <- M psi on bits (bits[0] = MSB of M's index), plain numpy."""
    psi = state.copy()
    D = 1 << k
    idx = np.arange(2 ** N)
    sub = np.zeros(2 ** N, dtype=np.int64)
    for i, b in enumerate(bits[:k]):
        sub |= ((idx >> b) & 1) << (k - 1 - i)
    rest = idx.copy()
    for b in bits[:k]:
        rest &= ~(1 << b)
    out = np.zeros_like(psi)
    for r in range(D):
        for c in range(D):
            if m[r, c] == 0:
                continue
            sel = sub == c
            tgt = rest[sel].copy()
            for i, b in enumerate(bits[:k]):
                tgt |= ((r >> (k - 1 - i)) & 1) << b
            np.add.at(out, tgt, m[r, c] * psi[sel])
    return out


@pytest.mark.parametrize("max_k,pad", [(2, 0), (3, 0), (4, 0), (5, 0), (5, 5)])
@pytest.mark.parametrize("seed", range(4))
def test_fusion_planner(emu, bt, max_k, pad, seed):
    """Fused blocks (bt_fuse.hpp) applied in order equal the gate-by-gate
    oracle run; every block stays within the fusion width."""
    from bluetangle_jl_amd.circuit import compile_ir
    rng = np.random.default_rng(7000 + seed)
    N = 6
    spec = random_spec(N, 40, rng, kinds=("gate1", "gate2", "ccx") if max_k >= 3 else ("gate1", "gate2"))
    ir = compile_ir(to_bt(spec, bt), N, False)
    arr, pool = ir.arrays()
    cap = 256
    ks = np.zeros(cap, dtype=np.int32)
    bits = np.zeros(5 * cap, dtype=np.int32)
    mats = np.zeros(2 * 1024 * cap, dtype=np.float64)
    emu.emu_fuse.restype = ctypes.c_int64
    nb = emu.emu_fuse(N, arr, len(ir.ops), ir.ntop, pool.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(pool.size // 2),
                      max_k, pad, ks.ctypes.data_as(ctypes.c_void_p), bits.ctypes.data_as(ctypes.c_void_p),
                      mats.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(cap))
    assert nb > 0, emu.emu_error()
    v = rng.normal(size=2 ** N) + 1j * rng.normal(size=2 ** N)
    psi = v.copy()
    cm = mats.view(np.complex128)
    for b in range(nb):
        k = int(ks[b])
        assert k <= max(max_k, pad)
        if pad:
            assert k == pad
        D = 1 << k
        m = cm[1024 * b: 1024 * b + D * D].reshape(D, D, order="F")
        psi = _apply_block_dense(psi, N, k, list(bits[5 * b: 5 * b + 5]), m)
    r, _, _, _ = oracle_run(spec, N, None, 0, state=v)
    assert np.abs(psi - r).max() < 1e-10
    assert nb < len(spec)  # fusion merged something
