"""GPU tests of the serving executor (bt_server_*): continuous batching with
HBM-budget admission.  A request's outcomes and probabilities are the same
bit for bit whether it runs alone or in a mix, agree with the FP64 oracle,
and the budget bounds what is resident."""
import numpy as np
import pytest

import oracle as O
from refcirc import brickwork_spec, random_spec, to_bt, to_ref

pytestmark = pytest.mark.gpu

NM = ("depolarizing", 0.02, "amplitude_damping", 0.05)
KINDS = ("gate1", "gate2", "ccx", "chan1", "chan2", "meas", "reset", "ifop")


def _circuits(bt):
    from bluetangle_jl_amd.circuit import Circuit
    out = []
    for i, N in enumerate((5, 8, 11)):
        spec = random_spec(N, 40, np.random.default_rng(900 + i), kinds=KINDS)
        out.append((spec, N, Circuit(to_bt(spec, bt), N, bt.NoiseModel(*NM))))
    spec = brickwork_spec(6, 4, 9, ad_gamma=0.02)
    out.append((spec, 6, Circuit(to_bt(spec, bt), 6, bt.NoiseModel(*NM))))  # also run as a density matrix
    return out


def _drain(srv, n):
    done = {}
    while len(done) < n:
        for r in srv.poll(wait=True):
            done[r["id"]] = r
    return done


def test_mix_equals_alone_and_oracle(bt):
    from bluetangle_jl_amd.circuit import Server
    circs = _circuits(bt)
    reqs = []  # (circuit index, kind, prec, seed, traj)
    for t in range(6):
        for ci in range(3):
            reqs.append((ci, "sv", "c64" if t % 2 == 0 else "c128", 5, t))
    reqs.append((3, "dm", "c128", 0, 0))
    reqs.append((3, "dm", "c64", 0, 0))
    # the mix
    srv = Server()
    ids = [srv.submit(circs[ci][2], kind, prec, seed, traj) for ci, kind, prec, seed, traj in reqs]
    res = _drain(srv, len(ids))
    mixed = {}
    for rid, q in zip(ids, reqs):
        assert res[rid]["status"] == 0, res[rid]
        assert 0 <= res[rid]["queue_ms"] <= res[rid]["latency_ms"]
        mixed[rid] = srv.fetch(rid)
    assert srv.stats() == dict(active=0, queued=0, hbm_in_use=0)
    srv.close()
    # each alone
    for rid, (ci, kind, prec, seed, traj) in zip(ids, reqs):
        s1 = Server()
        r1 = s1.submit(circs[ci][2], kind, prec, seed, traj)
        _drain(s1, 1)
        outs, p = s1.fetch(r1)
        s1.close()
        assert outs == mixed[rid][0]
        assert np.array_equal(p, mixed[rid][1]), (ci, kind, prec, traj)
    # against the oracle
    for rid, (ci, kind, prec, seed, traj) in zip(ids, reqs):
        spec, N, _ = circs[ci]
        outs, p = mixed[rid]
        tol = 1e-12 if prec == "c128" else 1e-5
        if kind == "dm":
            ref = O.run_dm(N, to_ref(spec), O.ref_noise(*NM))
            assert np.abs(p - np.diag(ref).real).max() < tol
            continue
        u = O.uniforms(seed, traj, 16 * len(spec) + 16)
        r, routs, used, bnd, _ = O.run_sv(N, to_ref(spec), O.ref_noise(*NM), u, tol=1e-6)
        if bnd:
            continue
        assert outs == routs
        assert np.allclose(p, np.abs(r) ** 2, rtol=0, atol=tol, equal_nan=True)


def test_budget_admission_and_rejection(bt):
    from bluetangle_jl_amd.circuit import Circuit, Server
    N = 16
    spec = brickwork_spec(N, 6, 3, mz_prob=0.2)
    c = Circuit(to_bt(spec, bt), N, bt.NoiseModel("depolarizing", 1e-3, "depolarizing", 1e-2))
    state = (1 << N) * 8
    srv = Server(hbm_budget_bytes=2 * state)  # two complex64 states resident at most
    big = srv.submit(c, "sv", "c128", 1, 0)    # 2x the budget of one... fits alone (2 * 8 B * 2^N)
    ids = [srv.submit(c, "sv", "c64", 1, t) for t in range(8)]
    st = srv.stats()
    assert st["hbm_in_use"] <= 2 * state
    too_big = Server(hbm_budget_bytes=state // 2)
    rej = too_big.submit(c, "sv", "c64", 1, 0)
    r = too_big.poll(wait=True)
    assert r[0]["id"] == rej and r[0]["status"] < 0
    too_big.close()
    done = {}
    peak = 0
    while len(done) < len(ids) + 1:
        peak = max(peak, srv.stats()["hbm_in_use"])
        for x in srv.poll(wait=True):
            done[x["id"]] = x
            outs, _ = srv.fetch(x["id"], probs=False)
    assert peak <= 2 * state
    assert all(x["status"] == 0 for x in done.values())
    # the same trajectories through bt_run_batch give the same outcomes
    outs_b, _, _ = c.run_batch(8, seed=1, prec="c64")
    srv2 = Server()
    ids2 = [srv2.submit(c, "sv", "c64", 1, t) for t in range(8)]
    _drain(srv2, 8)
    for t, rid in enumerate(ids2):
        assert srv2.fetch(rid, probs=False)[0] == outs_b[t]
    srv2.close()
    srv.close()


def test_upper_sizes_mix_equals_alone(bt):
    """The serving configuration at the top of its size ranges: a state-vector
    request at N=28 (2 GiB complex64) and a density-matrix request at N=13
    (4^13 entries, 512 MiB complex64), with brickwork, depolarizing
    1e-3/1e-2, amplitude damping 5e-3 and MZ p=0.1 (state vector), beside
    small requests: each result is bit-identical alone and in the mix, the
    state vector's probabilities sum to 1 and the density matrix's diagonal
    (its trace) to 1."""
    from bluetangle_jl_amd.circuit import Circuit, Server
    from bluetangle_jl_amd import workloads as W
    nm = W.depolarizing_noise()
    big_sv = Circuit(W.brickwork_ops(28, 12, 41, ad_gamma=5e-3, mz_prob=0.1), 28, nm)
    big_dm = Circuit(W.brickwork_ops(13, 6, 42, ad_gamma=5e-3), 13, nm)
    small = Circuit(W.brickwork_ops(10, 20, 43, ad_gamma=5e-3, mz_prob=0.1), 10, nm)
    small_dm = Circuit(W.brickwork_ops(7, 10, 44, ad_gamma=5e-3), 7, nm)
    reqs = [(big_sv, "sv", "c64", 3, 0), (small, "sv", "c64", 3, 1), (big_dm, "dm", "c64", 0, 0),
            (small, "sv", "c128", 3, 2), (small_dm, "dm", "c64", 0, 0)]
    srv = Server()
    ids = [srv.submit(c, k, p, s, t) for c, k, p, s, t in reqs]
    res = _drain(srv, len(ids))
    mixed = {}
    for rid in ids:
        assert res[rid]["status"] == 0, res[rid]
        mixed[rid] = srv.fetch(rid)
    srv.close()
    o, p = mixed[ids[0]]
    assert p.size == 2 ** 28 and abs(p.sum() - 1.0) < 1e-5 and len(o) > 0
    assert abs(mixed[ids[2]][1].sum() - 1.0) < 1e-5
    for rid, (c, k, p, s, t) in zip(ids, reqs):
        s1 = Server()
        r1 = s1.submit(c, k, p, s, t)
        _drain(s1, 1)
        outs, probs = s1.fetch(r1)
        s1.close()
        assert outs == mixed[rid][0], (k, p, c.N)
        assert np.array_equal(probs, mixed[rid][1]), (k, p, c.N)
        del probs
    mixed.clear()


def test_arrivals_during_execution(bt):
    """Requests submitted while earlier ones are on the device (the server is
    mid-way through them) are admitted as rounds go, and every result is the
    same as the request alone."""
    import time
    from bluetangle_jl_amd.circuit import Circuit, Server
    from bluetangle_jl_amd import workloads as W
    nm = W.depolarizing_noise()
    long_c = Circuit(W.brickwork_ops(22, 60, 51, ad_gamma=5e-3, mz_prob=0.1), 22, nm)
    short = [Circuit(W.brickwork_ops(n, 10, 52 + n, ad_gamma=5e-3, mz_prob=0.1), n, nm) for n in (8, 12, 16)]
    srv = Server()
    first = [srv.submit(long_c, "sv", "c64", 7, t) for t in range(4)]
    t0 = time.time()
    while srv.stats()["active"] == 0 and time.time() - t0 < 30:
        time.sleep(0.001)
    assert srv.stats()["active"] > 0
    early = srv.poll()
    later = []
    for i in range(12):
        c = short[i % 3]
        later.append((srv.submit(c, "sv", "c64", 7, 100 + i), c, 100 + i))
    busy = srv.stats()
    done = {r["id"]: r for r in early}
    while len(done) < len(first) + len(later):
        for r in srv.poll(wait=True):
            done[r["id"]] = r
    # the long requests were still running when the short ones arrived
    assert len(early) < len(first) and busy["active"] + busy["queued"] > 0
    results = {rid: srv.fetch(rid) for rid in done}
    srv.close()
    for rid, c, traj in [(r, long_c, t) for t, r in enumerate(first)] + later:
        s1 = Server()
        r1 = s1.submit(c, "sv", "c64", 7, traj)
        _drain(s1, 1)
        outs, probs = s1.fetch(r1)
        s1.close()
        assert outs == results[rid][0] and np.array_equal(probs, results[rid][1]), (c.N, traj)


def test_mix_equals_alone_whatever_the_grouping(bt):
    """A request's tiles per workgroup (and with it the grouping of its FP64
    reductions) follows from its own size and the tuning alone, not from the
    other requests of a launch: an N=18 request alone and beside an N=26
    request (tiles per workgroup set to 32) gives the same outcomes and
    probabilities bit for bit."""
    from bluetangle_jl_amd import _lib
    from bluetangle_jl_amd.circuit import Circuit, Server
    from bluetangle_jl_amd import workloads as W
    nm = W.depolarizing_noise()
    mid = Circuit(W.brickwork_ops(18, 16, 61, ad_gamma=5e-3, mz_prob=0.2), 18, nm)
    big = Circuit(W.brickwork_ops(26, 4, 62, ad_gamma=5e-3, mz_prob=0.1), 26, nm)
    try:
        _lib.check(_lib.lib().bt_set_tuning(0, 0, 0, 32))
        res = {}
        for mix in (False, True):
            srv = Server()
            ids = [srv.submit(mid, "sv", "c64", 4, t) for t in range(3)]
            if mix:
                srv.submit(big, "sv", "c64", 4, 0)
            done = _drain(srv, len(ids) + (1 if mix else 0))
            assert all(done[i]["status"] == 0 for i in ids)
            res[mix] = [srv.fetch(i) for i in ids]
            srv.close()
        for (oa, pa), (ob, pb) in zip(res[False], res[True]):
            assert oa == ob and np.array_equal(pa, pb)
    finally:
        _lib.check(_lib.lib().bt_set_tuning(0, 0, 0, 0))


def test_reserve_presizes_the_pool(bt):
    """Server.reserve allocates state buffers on the caller's thread within
    the budget; requests of that size then run from the pool."""
    from bluetangle_jl_amd.circuit import Circuit, Server
    from bluetangle_jl_amd import workloads as W
    N = 16
    state = (1 << N) * 8
    c = Circuit(W.brickwork_ops(N, 6, 71, ad_gamma=5e-3, mz_prob=0.1), N, W.depolarizing_noise())
    srv = Server(hbm_budget_bytes=3 * state)
    assert srv.reserve("sv", "c64", N, 5) == 3  # the budget holds three
    ids = [srv.submit(c, "sv", "c64", 2, t) for t in range(6)]
    got = {}
    while len(got) < len(ids):  # fetch as they finish: unfetched states hold the budget
        for x in srv.poll(wait=True):
            assert x["status"] == 0
            got[x["id"]] = srv.fetch(x["id"], probs=False)[0]
    outs = [got[i] for i in ids]
    srv.close()
    ref, _, _ = c.run_batch(6, seed=2, prec="c64")
    assert outs == ref
