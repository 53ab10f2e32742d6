#!/usr/bin/env python3
"""Benchmark of the noisy-circuit apply path on one MI355X.

Workload (one "step"): 4096 seeded Kraus-sampled trajectories of an N=22
brickwork circuit, 32 layers (RX/RY/RZ on every qubit, CNOT pairs alternating
by layer), depolarizing noise p=1e-3 after 1-qubit gates and p=1e-2 after
CNOTs, amplitude damping gamma=5e-3 per qubit per layer, Z measurement of each
qubit with probability 0.1 after every layer (collapse + renormalise);
complex64 amplitudes, FP64 norms and probabilities, all trajectories in one
bt_run_batch call (lockstep rounds: one launch per round over every resident
trajectory).  Value: trajectories/s over the timed steps.

Beside it, on the same JSON line:
  roofline      the dominant kernel (the round pass, k_round<float>):
                algorithmic bytes per launch / average launch time (HIP events
                inside the library), against 8 TB/s;
  cpu_baseline  the FP64 CPU restatement of the reference (oracle/, fast
                in-place mode, OpenMP) on a bounded sample of the same
                circuit; with it, the same oracle in the reference's own
                algorithm (sparse full-space operators, one thread) on a
                sample of that circuit and on the two small configurations;
  plain_loop    the same circuit one trajectory per call (bt_run_circuit: one
                launch per pass, one circuit per launch);
  persistent    the same batch as one persistent launch (device-side
                interpreters), on a sample;
  per_op_loop   the plain form: apply(state, op; noise) op by op;
  gate_layer    single-gate layers on an N=30 complex64 state (per-op
                kernels), achieved GB/s against 8 TB/s; then expect(state, "Z")
                and expect(state, "X") over every qubit of that state;
  small_configs the N=12 complex128 brickwork (time per circuit) and the N=8
                noisy density matrix (complex64 vs complex128, trace, purity);
  serving       Poisson arrivals of mixed circuits through the continuous-
                batching server: p50/p99 latency, queue wait and
                throughput; then a second leg under a binding HBM budget;
  large_state   N=33 complex64 brickwork: 5-qubit blocks on FP32 MFMA vs the
                register-tile executor vs unfused kernels, error and fidelity
                against the complex128 run.
Before the timed steps the autotuner picks the tiling on a sample (untimed).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--ntraj 4096]
"""
import argparse
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "noisy trajectories/s + achieved HBM GB/s per gate layer; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# k_round's measured HBM traffic / algorithmic bytes on this circuit, by the
# tiling's coalescing bits (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated,
# tilings 13,3,16 and 13,4,16); a tiling not in the table reports no traffic
TRAFFIC_RATIO = {3: 1.125, 4: 1.000}
# k_round's operator flops per algorithmic byte on this circuit at 3 coalescing
# bits: the register operators run as v_mfma_f32_4x4x1f32 (512 flop each;
# rocprofv3 SQ_INSTS_VALU_MFMA_MOPS_F32 x 512 over the same launches'
# algorithmic bytes), against the FP32 matrix-core peak
MFMA_FLOP_PER_BYTE = {3: 11.505}
MFMA_FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def unfused_bytes_per_trajectory(cfg, elem=8):
    """The per-op schedule's algorithmic bytes for one trajectory of a circuit:
    every gate one read + write of the state (2S), every circuit-placed channel
    and measurement a reduction read plus a read + write (3S), every noise
    channel after a noisy gate 2S times the probability of a branch other than
    the first (a multiple of the identity for depolarizing noise); S = bytes of
    the state."""
    S = float(1 << cfg["N"]) * elem
    nm = cfg["noise"]
    miss = {q: 1.0 - abs(ch.kraus[0][0, 0]) ** 2 if ch else 0.0 for q, ch in ((1, nm.q1), (2, nm.q2))} if nm else {}
    total = 0.0
    for o in cfg["ops"]:
        if type(o).__name__ == "OpQC" or o.type == "\U0001f52c":
            total += 3 * S
        else:
            total += 2 * S
            if nm and o.noisy:
                total += 2 * S * miss[1 if (o.q == 1 and o.control == -2) else 2]
    return total


def c3_leg(bt, args):
    from bluetangle_jl_amd import workloads as W
    from bluetangle_jl_amd.circuit import Circuit
    cfg = W.config_c3(args.N, args.layers)
    c = Circuit(cfg["ops"], cfg["N"], cfg["noise"])
    tuning = None
    if args.tuning:  # explicit tile bits, coalescing bits, tiles per workgroup
        from bluetangle_jl_amd import _lib as L
        T, cb, tpb = (int(x) for x in args.tuning.split(","))
        L.check(L.lib().bt_set_tuning(0, T, cb, tpb))
        tuning = {"tile_bits": T, "coalesce_bits": cb, "tiles_per_block": tpb, "ms_per_traj": None}
    elif args.autotune > 0:  # untimed: pick tile bits / coalescing / tiles per workgroup on a sample
        tuning = c.autotune("c64", sample=args.autotune)
        log("autotune", tuning)
    else:  # the library's defaults
        import ctypes
        from bluetangle_jl_amd import _lib as L
        v = [ctypes.c_int(0) for _ in range(3)]
        L.check(L.lib().bt_get_tuning(0, *[ctypes.byref(x) for x in v]))
        tuning = {"tile_bits": v[0].value, "coalesce_bits": v[1].value, "tiles_per_block": v[2].value,
                  "ms_per_traj": None}
    stats = []
    for w in range(args.warmup):
        c.run_batch(args.ntraj, seed=1000 + w, traj0=0, prec="c64", mode=args.mode)
        log("warmup step %d done" % w)
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, _, st = c.run_batch(args.ntraj, seed=k + 1, traj0=0, prec="c64", mode=args.mode)
        stats.append(st)
        log("step %d: %.3f s device, %d launches" % (k, st["kernel_ms"] / 1e3, st["launches"]))
    dt = time.perf_counter() - t0
    return c, cfg, stats, dt, tuning


def precision_leg(bt, c, args):
    """Trajectory 0 of the benchmark circuit in complex64 against the same
    trajectory in complex128 (FP64 mode, validated against the oracle in the
    tests): max amplitude error, fidelity, identical outcomes."""
    import numpy as np
    a, b = bt.zero_state(c.N, "c64"), bt.zero_state(c.N, "c128")
    oa, _ = c.run_on(a, seed=1, traj=0)
    ob, _ = c.run_on(b, seed=1, traj=0)
    err = float(np.abs(a.to_numpy() - b.to_numpy()).max())
    fid = abs(a.inner(b)) ** 2 / (a.norm2() * b.norm2())
    return dict(max_amp_err_vs_fp64=err, fidelity_vs_fp64=fid, outcomes_identical=oa == ob,
                sample="trajectory 0 (seed 1) of the benchmark circuit")


def plain_loop_leg(bt, c, args):
    st = bt.zero_state(c.N, "c64")
    n = args.plain_traj
    c.run_on(st.set_basis(0), seed=7, traj=0)  # warm
    t0 = time.perf_counter()
    for t in range(n):
        c.run_on(st.set_basis(0), seed=7, traj=t)
    return n / (time.perf_counter() - t0)


def persistent_leg(bt, c, args):
    """The batching form as one persistent launch (bt_run_batch_persistent:
    each workgroup runs a trajectory's interpreter on its lane 0 and takes the
    next trajectory from an atomic counter), on a sample of the same batch."""
    n = args.persistent_traj
    c.run_batch(64, seed=3, prec="c64", mode="persistent")  # warm
    t0 = time.perf_counter()
    _, _, st = c.run_batch(n, seed=1, prec="c64", mode="persistent")
    dt = time.perf_counter() - t0
    return n / dt, st


def per_op_loop_leg(bt, cfg, args):
    """The plain form: apply(state, op; noise) op by op (one or two kernel
    launches per op, host draws), one trajectory at a time."""
    n = args.per_op_traj
    st = bt.zero_state(cfg["N"], "c64")
    t0 = time.perf_counter()
    for t in range(n):
        bt.seed(7, t)
        st.set_basis(0)
        bt.apply(st, cfg["ops"], noise=cfg["noise"], track_measurements=True)
    st.sync()
    return n / (time.perf_counter() - t0)


def gate_layer_leg(bt, args):
    N = args.gate_N
    st = bt.zero_state(N, "c64")
    ops = [bt.Op("H", q) for q in range(1, N + 1)]
    for op in ops[:2]:
        bt.apply(st, op)
    st.sync()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        for op in ops:
            bt.apply(st, op)
    st.sync()
    dt = time.perf_counter() - t0
    nlayers = reps * len(ops)
    gbs = nlayers * 2.0 * (1 << N) * 8 / dt / 1e9
    # every target bit position on its own: 4 back-to-back launches per bit
    per_bit = []
    for b in range(N):
        op = ops[N - 1 - b]  # qubit q acts on bit N - q
        bt.apply(st, op)
        st.sync()
        t1 = time.perf_counter()
        for _ in range(4):
            bt.apply(st, op)
        st.sync()
        per_bit.append(4 * 2.0 * (1 << N) * 8 / (time.perf_counter() - t1) / 1e9)
    # two-qubit gates: a random 4x4 unitary on each adjacent bit pair (b, b+1)
    import numpy as np
    rng = np.random.default_rng(2)
    U2, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
    per_pair = []
    for b in range(N - 1):
        op = bt.Op("U2", U2, N - b - 1, N - b)
        bt.apply(st, op)
        st.sync()
        t1 = time.perf_counter()
        for _ in range(4):
            bt.apply(st, op)
        st.sync()
        per_pair.append(4 * 2.0 * (1 << N) * 8 / (time.perf_counter() - t1) / 1e9)
    obs = observables_on(bt, st, N)
    del st
    two = dict(per_pair_gbs=[round(x, 1) for x in per_pair],
               mean_frac=round(sum(per_pair) / len(per_pair) / HBM_PEAK_GBS, 4),
               min_frac=round(min(per_pair) / HBM_PEAK_GBS, 4),
               min_pair=int(min(range(N - 1), key=lambda i: per_pair[i])),
               note="a random two-qubit unitary on bits (b, b+1), four launches per pair; per_pair_gbs[b]")
    return dict(N=N, dtype="complex64", layers=nlayers, ms_per_layer=round(dt / nlayers * 1e3, 4),
                achieved_gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4),
                per_bit_gbs=[round(x, 1) for x in per_bit], per_bit_min_frac=round(min(per_bit) / HBM_PEAK_GBS, 4),
                per_bit_mean_frac=round(sum(per_bit) / len(per_bit) / HBM_PEAK_GBS, 4),
                per_bit_min_bit=int(min(range(N), key=lambda i: per_bit[i])),
                note="one Hadamard per layer on each qubit in turn; read+write of the state per layer; wall time "
                     "of back-to-back launches; per_bit_gbs[b]: four launches on target bit b alone",
                two_qubit=two, observables=obs)


def observables_on(bt, st, N):
    """expect(state, "Z") and expect(state, "X") over every qubit of the
    resident state: one FP64 reduction (one read of the state) per qubit."""
    out = {}
    for name in ("Z", "X"):
        bt.expect(st, name)
        t0 = time.perf_counter()
        vals = bt.expect(st, name)
        dt = time.perf_counter() - t0
        gbs = N * (1 << N) * 8 / dt / 1e9
        out[name] = dict(ms_per_qubit=round(dt / N * 1e3, 4), achieved_gbs=round(gbs, 1),
                         frac=round(gbs / HBM_PEAK_GBS, 4), mean=round(float(sum(vals)) / N, 6))
    out["note"] = ("one read of the 2^N complex64 state per qubit (bt_expect), wall time including the host "
                   "round trip of each result")
    import numpy as np
    u = np.random.default_rng(1).random(1 << 20)
    bt.measurement.sample_with(st, u[:1024])
    t0 = time.perf_counter()
    idx, nb = bt.measurement.sample_with(st, u)
    dt = time.perf_counter() - t0
    out["sample"] = dict(shots=int(u.size), ms=round(dt * 1e3, 3), boundary=int(nb),
                         note="bt_sample: chunk sums (one read of the state), prefix, one wavefront per shot; "
                              "wall time including the upload of the uniforms and the download of the indices")
    return out


def small_configs_leg(bt, args):
    """The two small configurations: the N=12 complex128 brickwork (20 layers,
    final probabilities) and the N=8 noisy density matrix (16 layers,
    complex64 against complex128, FP64 trace and purity); time per circuit
    (state on the device, compiled once)."""
    import numpy as np
    from bluetangle_jl_amd import workloads as W
    from bluetangle_jl_amd.circuit import Circuit
    out = {}
    c1 = W.config_c1()
    cc = Circuit(c1["ops"], c1["N"], False)
    st = bt.zero_state(c1["N"], "c128")
    cc.run_on(st.set_basis(0))
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        cc.run_on(st.set_basis(0))
        p = st.probabilities()
    dt = (time.perf_counter() - t0) / reps
    # the same noiseless circuit as fused <= 3-bit blocks captured once in a
    # HIP graph on a resident state, replayed per circuit
    sg = bt.zero_state(c1["N"], "c128")
    graph = cc.capture(sg, max_k=3)
    sg.set_basis(0)
    graph.launch()
    pg = sg.probabilities()
    t0 = time.perf_counter()
    for _ in range(reps):
        sg.set_basis(0)
        graph.launch()
        pg = sg.probabilities()
    dtg = (time.perf_counter() - t0) / reps
    out["c1"] = dict(N=c1["N"], dtype="complex128", layers=20, ms_per_circuit=round(dt * 1e3, 3),
                     prob_sum_minus_1=float(p.sum() - 1.0),
                     note="bt_run_circuit (register-tile passes) + FP64 probabilities, per circuit",
                     graph_ms_per_circuit=round(dtg * 1e3, 3), graph_max_prob_diff=float(np.abs(pg - p).max()),
                     graph_note="fused <= 3-bit blocks captured once in a HIP graph, replayed per circuit "
                                "(reset + replay + FP64 probabilities)")
    del graph, sg
    c2 = W.config_c2()
    cd = Circuit(c2["ops"], c2["N"], c2["noise"])
    r64, r128 = bt.zero_rho(c2["N"], "c64"), bt.zero_rho(c2["N"], "c128")
    cd.run_dm(r128)
    t0 = time.perf_counter()
    for _ in range(reps):
        cd.run_dm(r64.set_basis(0))
    r64.sync()
    dt = (time.perf_counter() - t0) / reps
    tr, pu = r64.trace_purity()
    err = float(np.abs(r64.to_numpy() - r128.to_numpy()).max())
    # the same noisy density-matrix circuit captured once in a HIP graph on a
    # resident rho, replayed per circuit (reset + replay + FP64 trace/purity)
    rg = bt.zero_rho(c2["N"], "c64")
    graph = cd.capture_dm(rg)
    rg.set_basis(0)
    graph.launch()
    rg.trace_purity()
    t0 = time.perf_counter()
    for _ in range(reps):
        rg.set_basis(0)
        graph.launch()
        trg, pug = rg.trace_purity()
    dtg = (time.perf_counter() - t0) / reps
    gdiff = float(np.abs(rg.to_numpy() - r64.to_numpy()).max())
    out["c2"] = dict(N=c2["N"], dtype="complex64 (FP64 trace/purity)", layers=16, ms_per_circuit=round(dt * 1e3, 3),
                     max_entry_err_vs_fp64=err, trace=tr, purity=pu,
                     note="bt_run_circuit_dm: gates as E rho E^dagger, channels as one superoperator pass each",
                     graph_ms_per_circuit=round(dtg * 1e3, 3), graph_passes=graph.passes, graph_max_entry_diff=gdiff,
                     graph_note="the density-matrix passes captured once in a HIP graph (bt_graph_create_dm), "
                                "replayed per circuit (reset + replay + FP64 trace and purity)")
    del graph, rg
    return out


def serving_leg(bt, args, budget_gib=0.0, nreq=None, tag="serving"):
    """Poisson arrivals of the serving workload through the continuous-
    batching server.  Circuits: 128 distinct random ones (statevector N ~
    U{10..28} or density matrix N ~ U{6..13}, depth ~ U{10..100}, brickwork
    and noise as the batched workload), each request one of them with its own
    trajectory.  The server's state pool is pre-sized for the sizes present
    (Server.reserve, on this thread) before the clock starts.  budget_gib > 0:
    the HBM budget binds and requests wait for admission (queue_ms)."""
    import numpy as np
    from bluetangle_jl_amd import workloads as W
    from bluetangle_jl_amd.circuit import Circuit, Server
    rng = np.random.default_rng(args.serve_seed + (1 if budget_gib else 0))
    pool = []
    for i in range(128):
        dm = rng.random() < 0.25
        N = int(rng.integers(6, 14)) if dm else int(rng.integers(10, 29))
        depth = int(rng.integers(10, 101))
        ops = W.brickwork_ops(N, depth, int(rng.integers(1 << 30)), ad_gamma=5e-3,
                              mz_prob=0.0 if dm else 0.1)
        pool.append((Circuit(ops, N, W.depolarizing_noise()), "dm" if dm else "sv", N, depth))
    n = nreq or args.serve_requests
    pick = rng.integers(0, len(pool), size=n)
    gaps = rng.exponential(1.0 / args.serve_rate, size=n)
    budget = int(budget_gib * 2 ** 30)
    srv = Server(hbm_budget_bytes=budget)
    reserved = 0
    for kind, N in sorted({(k, N) for _, k, N, _ in pool}):
        reserved += srv.reserve(kind, "c64", N, 1 if budget else 2)
    ids = {}
    peak = [0]
    t_start = time.perf_counter()

    def producer():
        t = t_start
        for i in range(n):
            c, kind, N, depth = pool[pick[i]]
            t += gaps[i]
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids[srv.submit(c, kind, "c64", seed=11, traj=i)] = i
            peak[0] = max(peak[0], srv.stats()["hbm_in_use"])

    th = threading.Thread(target=producer)
    th.start()
    done = {}
    while len(done) < n:
        got = srv.poll(wait=True)
        if not got:
            time.sleep(0.001)
        for r in got:
            done[r["id"]] = r
            srv.fetch(r["id"], probs=False)
    t_end = time.perf_counter()
    th.join()
    srv.close()
    lat = np.array([r["latency_ms"] for r in done.values()])
    qms = np.array([r["queue_ms"] for r in done.values()])
    bad = sum(1 for r in done.values() if r["status"] != 0)
    out = dict(requests=n, failed=bad, rate_per_s=args.serve_rate,
               p50_ms=round(float(np.percentile(lat, 50)), 2), p99_ms=round(float(np.percentile(lat, 99)), 2),
               queue_p50_ms=round(float(np.percentile(qms, 50)), 2),
               queue_p99_ms=round(float(np.percentile(qms, 99)), 2),
               throughput_rps=round(n / (t_end - t_start), 2), pool_buffers_reserved=int(reserved),
               workload="Poisson arrivals; statevector N~U{10..28} (75%) or density matrix N~U{6..13} (25%), "
                        "brickwork depth~U{10..100}, depolarizing 1e-3/1e-2 + amplitude damping 5e-3, "
                        "MZ p=0.1 per qubit per layer (statevector); complex64; 128 distinct circuits, one "
                        "trajectory per request")
    if budget:
        out["hbm_budget_gib"] = budget_gib
        out["peak_hbm_in_use_gib"] = round(peak[0] / 2 ** 30, 3)
    return out


def c5_leg(bt, args):
    """Large single state: N=33 complex64 (64 GiB), brickwork without noise,
    gates fused into 5-qubit 32x32 blocks on FP32 MFMA; FP64 norm every 4
    blocks; error against the same circuit run in complex128 (FP64 mode,
    128 GiB) on 2^20 sampled amplitudes plus the global fidelity; beside it
    the register-tile executor on the complex64 state and the unfused
    one-gate-per-pass kernels on the first layer."""
    import numpy as np
    from bluetangle_jl_amd import workloads as W
    from bluetangle_jl_amd.circuit import Circuit
    N, layers = args.c5_N, args.c5_layers
    ops = W.brickwork_ops(N, layers, 2026)
    c = Circuit(ops, N)
    amps = float(1 << N)
    a = bt.zero_state(N, "c64")
    st, norms = c.run_fused(a, max_k=5, norm_every=4)
    blocks = st["blocks"]
    fused_gbs = st["bytes"] / (st["kernel_ms"] * 1e-3) / 1e9
    # algorithmic: 8 * 2^k flop per amplitude for a k-bit block; executed on
    # the MFMA pipes: three real 32 x 32 products (Gauss form), 192 per amplitude
    tflops = blocks * amps * 256.0 / (st["kernel_ms"] * 1e-3) / 1e12
    mfma_tflops = tflops * 192.0 / 256.0
    log("c5 fused: %d blocks, %.1f ms, %.0f GB/s" % (blocks, st["kernel_ms"], fused_gbs))
    r = bt.zero_state(N, "c128")
    t0 = time.perf_counter()
    _, sr = c.run_on(r)
    ref_s = time.perf_counter() - t0
    rng = np.random.default_rng(33)
    idx = rng.integers(0, 1 << N, size=1 << 20, dtype=np.uint64)
    err = float(np.abs(a.gather(idx) - r.gather(idx)).max())
    ov = a.inner(r)
    fid = abs(ov) ** 2 / (a.norm2() * r.norm2())
    del r
    a.set_basis(0)
    t0 = time.perf_counter()
    _, stile = c.run_on(a)
    tile_s = time.perf_counter() - t0
    fusion = None
    if args.c5_autotune:  # fusion width: max_k 3, 4, 5 (MFMA) or the register-tile executor, on a scratch state
        k, ms = c.autotune_fused(a)
        fusion = dict(best=("register-tile executor" if k == 0 else "fused blocks, max_k=%d" % k),
                      ms={"k3": round(ms[0], 1), "k4": round(ms[1], 1), "k5_mfma": round(ms[2], 1),
                          "register_tile": round(ms[3], 1)})
        log("c5 fusion autotune", fusion)
    per_layer = len(ops) // layers
    a.set_basis(0)
    a.sync()
    t0 = time.perf_counter()
    for op in ops[:per_layer]:
        bt.apply(a, op)
    a.sync()
    unf_s = time.perf_counter() - t0
    unf_gbs = per_layer * 2.0 * amps * 8 / unf_s / 1e9
    del a
    return dict(workload="N=%d complex64 brickwork, %d layers, %d gates, no noise" % (N, layers, len(ops)),
                fused_blocks=blocks, fused_s=round(st["kernel_ms"] / 1e3, 3), fused_gbs=round(fused_gbs, 1),
                fused_frac=round(fused_gbs / HBM_PEAK_GBS, 4), block_tflops=round(tflops, 1),
                mfma_tflops=round(mfma_tflops, 1), mfma_frac=round(mfma_tflops / 157.3, 4), norm_checks=len(norms),
                max_norm_dev=round(max(abs(x - 1.0) for x in norms), 9) if norms else None,
                max_amp_err_vs_fp64=err, fidelity_vs_fp64=fid, fp64_reference_s=round(ref_s, 3),
                tile_executor_s=round(tile_s, 3), tile_executor_passes=int(stile["passes"]),
                unfused_layer_s=round(unf_s, 3), unfused_gbs=round(unf_gbs, 1),
                unfused_s_extrapolated=round(unf_s * layers, 2), fusion_autotune=fusion)


def cpu_leg(args):
    """The oracle (FP64 CPU restatement of the reference, fast in-place mode,
    OpenMP) on the same circuit.  Bounded sample: the first L layers of one
    trajectory (the circuit's layers are identical in cost), L chosen from a
    one-layer probe so the sample takes about args.cpu_seconds; the rate is
    scaled to whole trajectories."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from refcirc import brickwork_spec, count_draws_upper, to_ref
    nm = O.ref_noise("depolarizing", 1e-3, "depolarizing", 1e-2)

    def run_layers(L):
        spec = brickwork_spec(args.N, L, 2026, ad_gamma=5e-3, mz_prob=0.1)
        ops = to_ref(spec)
        md = count_draws_upper(spec, True)
        t0 = time.perf_counter()
        O.run_traj(args.N, ops, nm, 1, 0, 1, md, mode=0, tol=1e-6)
        return time.perf_counter() - t0

    t1 = run_layers(1)
    L = int(max(1, min(args.layers, args.cpu_seconds // max(t1, 1e-3))))
    dt = run_layers(L)
    per_traj = dt * args.layers / L
    out = dict(value=round(1.0 / per_traj, 5), unit="trajectories/s", cores=int(O.num_threads()), kind="port",
               sample="first %d of %d layers of one trajectory of the same N=%d circuit (%.1f s), scaled to a "
                      "whole trajectory; FP64 in-place strided kernels, OpenMP" % (L, args.layers, args.N, dt))
    if args.cpu_reference:
        out["reference_algorithm"] = cpu_reference_leg(args, O, nm, brickwork_spec, count_draws_upper, to_ref)
    out["host"] = host_info()
    return out


def host_info():
    """The host the CPU legs ran on: CPU model, logical CPUs (and those this
    process may use), RAM, OpenMP threads, the oracle's compiler flags."""
    model, ram = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    ram = round(int(line.split()[1]) / 2 ** 20, 1)
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return dict(cpu=model, logical_cpus=os.cpu_count(), usable_cpus=usable, ram_gib=ram,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                oracle_build="gcc -O2 -std=c11 -fopenmp (oracle/Makefile)",
                julia_threads="none: no Julia in the image; the reference-algorithm leg runs on one thread, "
                              "as the reference (which starts no threads)")


def cpu_reference_leg(args, O, nm, brickwork_spec, count_draws_upper, to_ref):
    """The oracle in the reference's own algorithm (every gate, Kraus operator
    and projector built as a full 2^N x 2^N sparse matrix and applied by a
    sparse product; hilbert.jl:18-159, struct.jl:31-41), on one thread as the
    reference runs.  The N=22 trajectory: every 4th op of the second layer (the
    op mix of a layer, on a dense state), scaled by the circuit's op count.  The two small
    configurations are timed whole, in this mode and in the in-place mode."""
    import numpy as np
    nthreads = int(O.num_threads())
    O.set_num_threads(1)
    try:
        # layer 2 from the state after layer 1 (dense, as in every later
        # layer: the sparse product skips zero amplitudes, so the first layer
        # from |0..0> would understate the cost)
        n1 = len(to_ref(brickwork_spec(args.N, 1, 2026, ad_gamma=5e-3, mz_prob=0.1)))
        spec2 = brickwork_spec(args.N, 2, 2026, ad_gamma=5e-3, mz_prob=0.1)
        ops2 = to_ref(spec2)
        u = O.uniforms(1, 0, count_draws_upper(spec2, True) + 16)
        s1, _, used, _, _ = O.run_sv(args.N, ops2[:n1], nm, u=u, mode=0, tol=1e-6)
        sub = ops2[n1:][::4]
        total = len(to_ref(brickwork_spec(args.N, args.layers, 2026, ad_gamma=5e-3, mz_prob=0.1)))
        t0 = time.perf_counter()
        O.run_sv(args.N, sub, nm, u=u[used:], state=s1, mode=1, tol=1e-6)
        dt = time.perf_counter() - t0
        del s1
        per_traj = dt * total / len(sub)
        res = dict(value=float("%.4g" % (1.0 / per_traj)), unit="trajectories/s", cores=1,
                   sample="every 4th op of the second layer of the same N=%d circuit, from the state after the "
                          "first (%d ops, %.1f s), scaled by the circuit's %d ops; sparse full-space operators, one "
                          "thread" % (args.N, len(sub), dt, total))

        def per_circuit(fn, min_s=0.5):
            fn()
            n, t0 = 0, time.perf_counter()
            while True:
                fn()
                n += 1
                el = time.perf_counter() - t0
                if el >= min_s:
                    return el / n * 1e3

        c1 = to_ref(brickwork_spec(12, 20, 7, gate_noise=False))
        c2 = to_ref(brickwork_spec(8, 16, 7, gate_noise=True, ad_gamma=5e-3))
        small = {}
        for mode, key in ((1, "reference_ms"), (0, "in_place_ms")):
            small.setdefault("c1", {})[key] = round(per_circuit(
                lambda: np.abs(O.run_sv(12, c1, None, mode=mode)[0]) ** 2), 3)
            small.setdefault("c2", {})[key] = round(per_circuit(lambda: O.run_dm(8, c2, nm, mode=mode)), 3)
        small["note"] = ("time per circuit on one thread: C1 N=12 noiseless brickwork, 20 layers, with |psi|^2; "
                         "C2 N=8 noisy rho, 16 layers; FP64")
        res["small_configs"] = small
        return res
    finally:
        O.set_num_threads(nthreads)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ntraj", type=int, default=4096)
    ap.add_argument("--N", type=int, default=22)
    ap.add_argument("--layers", type=int, default=32)
    ap.add_argument("--mode", default="rounds", choices=["rounds", "persistent"])
    ap.add_argument("--autotune", type=int, default=256, help="trajectories per autotuner candidate (0: off)")
    ap.add_argument("--tuning", default="", help="T,C,B: fixed tile bits, coalescing bits, tiles per workgroup")
    ap.add_argument("--plain-traj", type=int, default=64)
    ap.add_argument("--per-op-traj", type=int, default=4)
    ap.add_argument("--persistent-traj", type=int, default=1024)
    ap.add_argument("--gate-N", type=int, default=30)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-budget-gib", type=float, default=16.0, help="HBM budget of the second serving leg (0: off)")
    ap.add_argument("--serve-budget-requests", type=int, default=300)
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-reference", type=int, default=1,
                    help="also time the oracle in the reference's sparse full-space-operator algorithm (one thread)")
    ap.add_argument("--skip", default="", help="comma list of legs to skip: cpu,plain,gate,small,serve,c5")
    ap.add_argument("--no-c5", dest="c5", action="store_false", help="skip the N=33 large-state leg")
    ap.add_argument("--c5-N", type=int, default=33)
    ap.add_argument("--c5-layers", type=int, default=40)
    ap.add_argument("--c5-autotune", type=int, default=1, help="run the fusion-width autotuner on the C5 circuit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported (the value is one GPU's "
              "throughput)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    skip = set(s for s in args.skip.split(",") if s)
    import __graft_entry__ as G
    bt = G.load_package()
    bt.init()
    c, cfg, stats, dt, tuning = c3_leg(bt, args)
    value = args.ntraj * args.steps / dt
    kms = sum(s["kernel_ms"] for s in stats)
    nbytes = sum(s["bytes"] for s in stats)
    launches = sum(s["launches"] for s in stats)
    achieved = nbytes / (kms * 1e-3) / 1e9
    passes = sum(s["passes"] for s in stats) / (args.ntraj * args.steps)
    cb = (tuning or {}).get("coalesce_bits")
    unfused = unfused_bytes_per_trajectory(cfg)
    fused = nbytes / (args.ntraj * args.steps)
    ratio, fpb = TRAFFIC_RATIO.get(cb), MFMA_FLOP_PER_BYTE.get(cb)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 (fp32 amplitudes; fp64 norms, probabilities, sampling sums)",
        "data": "synthetic",
        "config": {"workload": "batched noisy trajectories: %d trajectories of an N=%d brickwork, %d layers, "
                               "depolarizing 1e-3/1e-2, amplitude damping 5e-3, MZ p=0.1 per qubit per layer"
                               % (args.ntraj, args.N, args.layers),
                   "ntraj": args.ntraj, "N": args.N, "layers": args.layers, "executor": args.mode,
                   "passes_per_trajectory": round(passes, 3),
                   "boundary_draws": int(sum(s["boundary_draws"] for s in stats)),
                   "tuning": tuning,
                   "unfused_equivalent": {
                       "gbs": round(value * unfused / 1e9, 1),
                       "bytes_per_trajectory": round(unfused, 1),
                       "fused_bytes_per_trajectory": round(fused, 1),
                       "reduction": round(unfused / fused, 1),
                       "note": "an equivalent, not a measured rate: the per-op schedule's algorithmic bytes "
                               "(gate 2S, amplitude damping and MZ 3S, noise branches 2S x their probability; "
                               "S = the complex64 state's bytes) per trajectory x trajectories/s"}},
        "roofline": {"bound": "hbm", "kernel": "k_round<float>" if args.mode == "rounds" else "k_batch<float>",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(nbytes / max(1, launches) * ratio, 1) if ratio else None,
                     "traffic_source": ("HBM read+write per launch = algorithmic x %.3f, the FETCH_SIZE/WRITE_SIZE "
                                        "ratio measured for this kernel, circuit and %d coalescing bits "
                                        "(profiles/k_round_traffic_pmc.txt)" % (ratio, cb)) if ratio else
                                       "not measured for %s coalescing bits" % cb,
                     "launches": int(launches), "avg_launch_ms": round(kms / max(1, launches), 4),
                     "bytes_per_launch": round(nbytes / max(1, launches), 1),
                     "matrix": {"achieved_tflops": round(achieved * fpb / 1e3, 1) if fpb else None,
                                "peak_tflops": MFMA_FP32_PEAK_TFLOPS,
                                "frac": round(achieved * fpb / 1e3 / MFMA_FP32_PEAK_TFLOPS, 4) if fpb else None,
                                "flop_per_byte": fpb,
                                "note": "the register operators' FP32 matrix-core work per algorithmic byte "
                                        "(rocprofv3 SQ_INSTS_VALU_MFMA_MOPS_F32 x 512, "
                                        "profiles/k_round_c3_n22_4096traj.txt) against the FP32 MFMA peak"}},
    }
    out["precision"] = precision_leg(bt, c, args)
    log("precision", out["precision"])
    if "plain" not in skip:
        pl = plain_loop_leg(bt, c, args)
        out["plain_loop"] = {"value": round(pl, 3), "unit": "trajectories/s", "sample": args.plain_traj,
                             "batched_speedup": round(value / pl, 2),
                             "executor": "bt_run_circuit per trajectory (fused tile passes, one launch per pass)"}
        log("plain loop %.2f traj/s" % pl)
        pv, pst = persistent_leg(bt, c, args)
        out["persistent"] = {"value": round(pv, 2), "unit": "trajectories/s", "sample": args.persistent_traj,
                             "rounds_over_persistent": round(value / pv, 2),
                             "passes_per_trajectory": round(pst["passes"] / args.persistent_traj, 3),
                             "executor": "bt_run_batch_persistent: one launch, device-side interpreters"}
        log("persistent %.2f traj/s" % pv)
        po = per_op_loop_leg(bt, cfg, args)
        out["per_op_loop"] = {"value": round(po, 4), "unit": "trajectories/s", "sample": args.per_op_traj,
                              "batched_speedup": round(value / po, 1),
                              "executor": "apply(state, op; noise) per op, one trajectory at a time"}
        log("per-op loop %.3f traj/s" % po)
    if "gate" not in skip:
        out["gate_layer"] = gate_layer_leg(bt, args)
        log("gate layer", out["gate_layer"])
    if "small" not in skip:
        out["small_configs"] = small_configs_leg(bt, args)
        log("small configs", out["small_configs"])
    if "serve" not in skip:
        out["serving"] = serving_leg(bt, args)
        log("serving", out["serving"])
        if args.serve_budget_gib > 0:
            out["serving"]["budget_leg"] = serving_leg(bt, args, budget_gib=args.serve_budget_gib,
                                                       nreq=args.serve_budget_requests)
            log("serving (budget)", out["serving"]["budget_leg"])
    if args.c5 and "c5" not in skip:
        out["large_state"] = c5_leg(bt, args)
        log("c5", out["large_state"])
    if "cpu" not in skip:
        out["cpu_baseline"] = cpu_leg(args)
        log("cpu", out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
