"""Neutral circuit descriptions shared by the oracle and the engine in tests.

A spec is a list of tuples, turned into oracle RefOps with ``to_ref`` (the
oracle's own gate/Kraus tables) or into engine ops with ``to_bt`` (the
engine's tables), so a mistake in one table cannot hide in both.

  ("gate", name, qubit, target, control, noisy)
  ("chan", model, p, qubit, target)            # OpQC(model, p, qubit[, target])
  ("kraus", [K...], qubit, target)             # OpQC("custom", kraus, qubit[, target]); 1-3 qubits
  ("meas", name, qubit)                        # Op("MZ"/"MX"/"MY"/"MR", qubit)
  ("reset", qubit)                             # Op("RES", qubit)
  ("ifop", name, qubit, spec0, spec1)          # ifOp(name, qubit, if0, if1)
"""
import math

import numpy as np

import oracle as O

ONE_Q = ["X", "Y", "Z", "H", "S", "T", "SD", "TD", "SX", "HY", "I"]
ONE_Q_PARAM = ["RX", "RY", "RZ", "P"]
TWO_Q = ["CNOT", "CZ", "SWAP", "ISWAP", "FSWAP", "SYC", "ECR", "CY", "CH", "CS"]
TWO_Q_PARAM = ["CP", "RXX", "RYY", "RZZ", "RXY", "GIVENS"]


def to_ref(spec):
    out = []
    for s in spec:
        kind = s[0]
        if kind == "gate":
            _, name, q, t, c, noisy = s
            out.append(O.ref_gate(name, q, t, c, noisy))
        elif kind == "chan":
            _, model, p, q, t = s
            out.append(O.ref_channel(model, p, q, t))
        elif kind == "kraus":
            _, ks, q, t = s
            out.append(O.ref_channel_kraus(ks, q, t))
        elif kind == "meas":
            out.append(O.ref_gate(s[1], s[2]))
        elif kind == "reset":
            out.append(O.ref_reset(s[1]))
        elif kind == "ifop":
            _, name, q, s0, s1 = s
            out.append(O.ref_ifop(name, q, to_ref(s0), to_ref(s1)))
        else:
            raise ValueError(kind)
    return out


def to_bt(spec, bt):
    out = []
    for s in spec:
        kind = s[0]
        if kind == "gate":
            _, name, q, t, c, noisy = s
            if t == -1:
                out.append(bt.Op(name, q, control=c, noisy=noisy))
            else:
                out.append(bt.Op(name, q, t, control=c, noisy=noisy))
        elif kind == "chan":
            _, model, p, q, t = s
            out.append(bt.OpQC(model, p, q, t))
        elif kind == "kraus":
            _, ks, q, t = s
            out.append(bt.OpQC("custom", [np.asarray(k) for k in ks], q, t))
        elif kind == "meas":
            out.append(bt.Op(s[1], s[2]))
        elif kind == "reset":
            out.append(bt.Op("RES", s[1]))
        elif kind == "ifop":
            _, name, q, s0, s1 = s
            out.append(bt.ifOp(name, q, to_bt(s0, bt), to_bt(s1, bt)))
        else:
            raise ValueError(kind)
    return out


def count_draws_upper(spec, noisy_model):
    n = 0
    for s in spec:
        if s[0] == "gate":
            n += 1 if (noisy_model and s[5]) else 0
        elif s[0] in ("chan", "kraus", "reset"):
            n += 1
        elif s[0] == "meas":
            n += 2 if s[1].upper() in ("MR", "M(R)") else 1
        elif s[0] == "ifop":
            n += (2 if s[1].upper() in ("MR", "M(R)") else 1) + count_draws_upper(s[3], noisy_model) + \
                count_draws_upper(s[4], noisy_model)
    return n


def random_unitary(D, rng):
    z = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_kraus(q, rng, style=None):
    """A random CPTP Kraus list on q qubits (2^q x 2^q), one of three kinds:
    'mix' (sqrt(w_i) U_i: state-independent probabilities), 'general'
    (A_i S^-1/2 with S = sum A_i'A_i: probabilities need the reduced rho) or
    'damping' (amplitude damping on each qubit, tensored: decided from
    eigenvalue bounds unless the draw is close to 1)."""
    D = 2 ** q
    style = style or ["mix", "general", "damping"][rng.integers(3)]
    if style == "mix":
        n = int(rng.integers(2, 4))
        w = rng.dirichlet(np.ones(n))
        return [math.sqrt(wi) * random_unitary(D, rng) for wi in w]
    if style == "general":
        n = int(rng.integers(1, 4))
        A = [rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D)) for _ in range(n)]
        S = sum(a.conj().T @ a for a in A)
        ev, V = np.linalg.eigh(S)
        Sm = (V * ev ** -0.5) @ V.conj().T
        return [a @ Sm for a in A]
    g = float(rng.uniform(0.02, 0.3))
    k1 = [np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=complex),
          np.array([[0, math.sqrt(g)], [0, 0]], dtype=complex)]
    ks = k1
    for _ in range(q - 1):
        ks = [np.kron(a, b) for a in ks for b in k1]
    return ks


def random_spec(N, depth, rng, kinds=("gate1", "gate2"), nonadjacent=True, controls=True):
    """Random gate lists in the spirit of random_ops (gates.jl:154-214),
    drawn from our own seeded numpy generator."""
    spec = []
    for _ in range(depth):
        k = kinds[rng.integers(len(kinds))]
        if k == "gate1" or N == 1:
            q = int(rng.integers(1, N + 1))
            if rng.random() < 0.5:
                name = ONE_Q[rng.integers(len(ONE_Q))]
            else:
                name = "%s(%.6f)" % (ONE_Q_PARAM[rng.integers(len(ONE_Q_PARAM))], rng.uniform(0, 2 * math.pi))
            c = -2
            if controls and N >= 2 and rng.random() < 0.2:
                c = int(rng.choice([x for x in range(1, N + 1) if x != q]))
            spec.append(("gate", name, q, -1, c, True))
        elif k == "gate2":
            if nonadjacent:
                q, t = [int(x) for x in rng.choice(np.arange(1, N + 1), 2, replace=False)]
            else:
                q = int(rng.integers(1, N))
                t = q + 1
                if rng.random() < 0.5:
                    q, t = t, q
            if rng.random() < 0.5:
                name = TWO_Q[rng.integers(len(TWO_Q))]
            else:
                name = "%s(%.6f)" % (TWO_Q_PARAM[rng.integers(len(TWO_Q_PARAM))], rng.uniform(0, 2 * math.pi))
            spec.append(("gate", name, q, t, -2, True))
        elif k == "ccx" and N >= 3:
            q, c, t = [int(x) for x in rng.choice(np.arange(1, N + 1), 3, replace=False)]
            spec.append(("gate", "CX", q, t, c, True))
        elif k == "chan1":
            model = ["amplitude_damping", "phase_damping", "depolarizing", "bit_flip", "phase_flip",
                     "depolarizing_amp", "rot_x", "rot_xyz"][rng.integers(8)]
            spec.append(("chan", model, float(rng.uniform(0.05, 0.4)), int(rng.integers(1, N + 1)), -1))
        elif k == "chan2" and N >= 2:
            q, t = [int(x) for x in rng.choice(np.arange(1, N + 1), 2, replace=False)]
            model = ["amplitude_damping", "depolarizing", "bit_flip"][rng.integers(3)]
            if model == "amplitude_damping" and q > t:
                # Q3: for q > t the reference samples with (min,max)-ordered
                # probabilities but applies K with its first factor on q, which can
                # pick a zero-norm branch (NaN in the reference too); keep q < t here.
                q, t = t, q
            spec.append(("chan", model, float(rng.uniform(0.05, 0.4)), q, t))
        elif k == "kraus":
            # OpQC(name, kraus, qubit[, target]) with a random CPTP list on 1, 2
            # or 3 qubits (struct.jl:182-234; 3 qubits act on qubit..qubit+2)
            q = int(rng.integers(1, min(N, 3) + 1))
            if q == 1:
                spec.append(("kraus", random_kraus(1, rng), int(rng.integers(1, N + 1)), -1))
            elif q == 2:
                a, b = [int(x) for x in rng.choice(np.arange(1, N + 1), 2, replace=False)]
                spec.append(("kraus", random_kraus(2, rng), a, b))
            else:
                spec.append(("kraus", random_kraus(3, rng), int(rng.integers(1, N - 1)), -1))
        elif k == "meas":
            spec.append(("meas", ["MZ", "MX", "MY", "MR"][rng.integers(4)], int(rng.integers(1, N + 1))))
        elif k == "reset":
            spec.append(("reset", int(rng.integers(1, N + 1))))
        elif k == "ifop":
            q = int(rng.integers(1, N + 1))
            spec.append(("ifop", ["MZ", "MX", "MY", "MR"][rng.integers(4)], q,
                         [("gate", "X", q, -1, -2, True)],
                         [("gate", "H", int(rng.integers(1, N + 1)), -1, -2, True)]))
    return spec


def brickwork_spec(N, layers, seed, gate_noise=True, ad_gamma=None, mz_prob=0.0):
    """The brickwork of the benchmark configurations: per layer RX, RY, RZ on
    every qubit (angles U[0,2pi) from `seed`), CNOT(i,i+1) on even-start
    pairs in even layers and odd-start pairs in odd layers, then optional
    OpQC amplitude damping on every qubit, then Op("MZ", q) with probability
    `mz_prob` per qubit (mask drawn once from `seed`)."""
    rng = np.random.default_rng(seed)
    spec = []
    for layer in range(layers):
        for q in range(1, N + 1):
            th = rng.uniform(0, 2 * math.pi, 3)
            spec.append(("gate", "RX(%r)" % float(th[0]), q, -1, -2, True))
            spec.append(("gate", "RY(%r)" % float(th[1]), q, -1, -2, True))
            spec.append(("gate", "RZ(%r)" % float(th[2]), q, -1, -2, True))
        start = 1 if layer % 2 == 0 else 2
        for i in range(start, N, 2):
            spec.append(("gate", "CNOT", i, i + 1, -2, True))
        if ad_gamma is not None:
            for q in range(1, N + 1):
                spec.append(("chan", "amplitude_damping", ad_gamma, q, -1))
        if mz_prob > 0:
            mask = rng.random(N) < mz_prob
            for q in range(1, N + 1):
                if mask[q - 1]:
                    spec.append(("meas", "MZ", q))
    return spec
