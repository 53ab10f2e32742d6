"""The workloads at their full sizes, checked through properties that do not
depend on a CPU run of the same size (the oracle cannot hold or finish them):

* N=30 complex64 single-gate layers (8 GiB): RY(theta_q) on every qubit of
  |0..0> through the per-op kernel at every bit position gives the product
  state with <Z_q> = cos(theta_q) and <X_q> = sin(theta_q) (closed form); the
  adjoint layer in reverse order returns every <Z_q> to 1.
* N=22, 4096 noisy trajectories (the batched workload, 32 layers): every
  trajectory records one outcome per MZ; the second cohort run alone gives the
  same outcomes as inside the full batch; single trajectories through the
  one-circuit executor, in complex64 and complex128, give the batch's
  outcomes (draws within the boundary tolerance excepted) and states of
  fidelity 1 - 1e-6 with each other.  One trajectory of it (N=22 is within
  the oracle's reach: about a minute on 16 host threads) is also compared
  with the FP64 oracle directly: identical outcomes, amplitudes within 1e-5.
* N=33 complex64 brickwork (64 GiB): 4 layers as fused 5-qubit MFMA blocks,
  then the adjoint circuit through the register-tile executor, return the
  state to |0..0> (every <Z_q> = 1); the fused and the register-tile results
  of the forward circuit agree (fidelity) and stay normalised.
"""
import gc
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _adjoint_ops(bt, ops):
    """The adjoint circuit of rotations and CNOTs: reversed order,
    R(theta) -> R(-theta), CNOT self-inverse."""
    out = []
    for o in reversed(ops):
        name = o.name
        if name.startswith(("RX(", "RY(", "RZ(")):
            th = float(name[3:-1])
            out.append(bt.Op("%s(%r)" % (name[:2], -th), o.qubit))
        elif name == "CNOT":
            out.append(bt.Op("CNOT", o.qubit, o.target_qubit))
        else:
            raise AssertionError("unexpected op " + name)
    return out


def test_gate_layer_n30_closed_form(bt):
    N = 30
    rng = np.random.default_rng(30)
    th = rng.uniform(0.0, math.pi, N)
    st = bt.zero_state(N, "c64")
    try:
        for q in range(1, N + 1):
            bt.apply(st, bt.Op("RY(%r)" % float(th[q - 1]), q))
        z = np.asarray(bt.expect(st, "Z"))
        x = np.asarray(bt.expect(st, "X"))
        assert np.abs(z - np.cos(th)).max() < 1e-5, np.abs(z - np.cos(th)).max()
        assert np.abs(x - np.sin(th)).max() < 1e-5, np.abs(x - np.sin(th)).max()
        assert abs(st.norm2() - 1.0) < 1e-5
        for q in range(N, 0, -1):
            bt.apply(st, bt.Op("RY(%r)" % float(-th[q - 1]), q))
        z = np.asarray(bt.expect(st, "Z"))
        assert np.abs(z - 1.0).max() < 1e-5
    finally:
        del st
        gc.collect()


def test_c3_full_size_batch(bt):
    from bluetangle_jl_amd import workloads as W
    from bluetangle_jl_amd.circuit import Circuit
    cfg = W.config_c3()
    N = cfg["N"]
    c = Circuit(cfg["ops"], N, cfg["noise"])
    nmz = sum(1 for o in cfg["ops"] if getattr(o, "name", "") == "MZ")
    seed = 5
    outs, _, st = c.run_batch(4096, seed=seed, prec="c64")
    assert len(outs) == 4096 and all(len(o) == nmz for o in outs)
    assert st["passes"] > 0
    half, _, _ = c.run_batch(2048, seed=seed, traj0=2048, prec="c64")
    assert half == outs[2048:]
    compared = 0
    for t in (0, 1, 2047, 2048, 4095):
        a = bt.zero_state(N, "c64")
        b = bt.zero_state(N, "c128")
        try:
            oa, sa = c.run_on(a, seed=seed, traj=t)
            ob, sb = c.run_on(b, seed=seed, traj=t)
            if sa["boundary_draws"] or sb["boundary_draws"]:
                continue
            assert oa == outs[t], t
            assert ob == oa, t
            fa, fb = a.norm2(), b.norm2()
            assert abs(fa - 1.0) < 1e-5 and abs(fb - 1.0) < 1e-12
            assert abs(a.inner(b)) ** 2 / (fa * fb) > 1.0 - 1e-6
            compared += 1
        finally:
            del a, b
            gc.collect()
    assert compared >= 3


def test_c5_full_size_fused_and_adjoint(bt):
    from bluetangle_jl_amd import workloads as W
    from bluetangle_jl_amd.circuit import Circuit
    N = 33
    ops = W.brickwork_ops(N, 4, 2026)
    fwd, inv = Circuit(ops, N), Circuit(_adjoint_ops(bt, ops), N)
    a = bt.zero_state(N, "c64")
    try:
        stats, norms = fwd.run_fused(a, max_k=5, norm_every=1)
        assert stats["blocks"] > 0
        assert np.abs(np.asarray(norms) - 1.0).max() < 1e-5
        b = bt.zero_state(N, "c64")
        try:
            fwd.run_on(b)
            fa, fb = a.norm2(), b.norm2()
            assert abs(fa - 1.0) < 1e-5 and abs(fb - 1.0) < 1e-5
            assert abs(a.inner(b)) ** 2 / (fa * fb) > 1.0 - 1e-6
        finally:
            del b
            gc.collect()
        inv.run_on(a)
        z = np.asarray(bt.expect(a, "Z"))
        assert np.abs(z - 1.0).max() < 1e-5, np.abs(z - 1.0).max()
    finally:
        del a
        gc.collect()


def test_c3_trajectory_vs_oracle(bt):
    """One trajectory of the batched workload at its full size (N=22, 32
    layers, depolarizing 1e-3/1e-2, amplitude damping 5e-3, MZ p=0.1; complex64
    amplitudes) from bt_run_batch against the FP64 oracle's run of the same
    circuit on the same uniforms (ops.jl:659-692): identical outcomes, max
    amplitude error within 1e-5."""
    from bluetangle_jl_amd import workloads as W
    from bluetangle_jl_amd.circuit import Circuit
    from refcirc import brickwork_spec, count_draws_upper, to_ref
    cfg = W.config_c3()
    N = cfg["N"]
    c = Circuit(cfg["ops"], N, cfg["noise"])
    seed, ntraj = 5, 4
    outs, states, st = c.run_batch(ntraj, seed=seed, prec="c64", keep_states=True)
    spec = brickwork_spec(N, 32, 2026, ad_gamma=5e-3, mz_prob=0.1)
    assert len(spec) == len(cfg["ops"])
    nm = O.ref_noise("depolarizing", 1e-3, "depolarizing", 1e-2)
    ro, rs, bnd = O.run_traj(N, to_ref(spec), nm, seed, 0, 1, count_draws_upper(spec, True), mode=0, tol=1e-6,
                             keep_states=True)
    if bnd:
        pytest.skip("a draw within 1e-6 of a cumulative boundary")
    assert outs[0] == ro[0]
    err = float(np.abs(states[0] - rs[0]).max())
    assert err < 1e-5, err
    print("C3 trajectory 0 vs oracle: %d outcomes identical, max amplitude error %.3g" % (len(ro[0]), err))
