"""CPU restatement of BlueTangle.jl's circuit compilation on RefOp lists.

TEST INFRASTRUCTURE ONLY (checker, never the product).

Follows compile (src/ops.jl:421-445) with the default layout, a line of N
qubits (src/struct.jl:781-801, src/layout.jl:16-50), whose swap insertion
(_make_swaps, src/layout.jl:69-90) puts SWAP pairs around a non-adjacent
2-qubit op and rebuilds the op from its name next to the target;
get_layers (src/ops.jl:283-329); get_stats and _calculate_circuit_depth
(src/ops.jl:191-278); the shot expansion for ZNE (src/ops.jl:555-573) and
Pauli twirling (src/noise.jl:136-255) with explicit twirl rows.

Each op's layer is computed from scratch here by a brute-force rule (the
first layer after the last layer, among all earlier ops, that touched any of
its qubits), a different formulation from the reference's running
dictionary, and the two are cross-checked in the tests.
"""
from __future__ import annotations

from typing import Dict, List, Sequence

import oracle as O

IFOP = O.KIND_IFOP


def touched(op) -> List[int]:
    """Qubits whose layer fixes op's layer (ops.jl:301-309): qubit, target for
    2-qubit ops, control for gates (not for ifOp)."""
    qs = [op.qubit]
    if op.q == 2:
        qs.append(op.target)
    if op.kind == O.KIND_GATE and op.control != -2:
        qs.append(op.control)
    return qs


def marks(op) -> List[int]:
    """Qubits whose layer the op then sets (ops.jl:316-326): qubit, target for
    2-qubit ops, control for Op gates."""
    qs = [op.qubit]
    if op.q == 2:
        qs.append(op.target)
    if op.kind in (O.KIND_GATE, O.KIND_MEASURE) and op.control != -2:
        qs.append(op.control)
    return qs


def layer_index(ops: Sequence) -> List[int]:
    """1-based layer of every op, brute force over the earlier ops."""
    lay: List[int] = []
    for i, op in enumerate(ops):
        need = set(touched(op))
        best = 0
        for j in range(i):
            if need & set(marks(ops[j])):
                best = max(best, lay[j])
        lay.append(best + 1)
    return lay


def layered(ops: Sequence) -> List:
    """vcat(get_layers(ops)...): stable sort by layer."""
    lay = layer_index(ops)
    return [op for _, _, op in sorted(zip(lay, range(len(ops)), ops), key=lambda t: (t[0], t[1]))]


def line_swaps(ops: Sequence) -> List:
    """compile's swap insertion on the default line layout."""
    out = []
    for op in ops:
        if op.kind == O.KIND_IFOP or op.q == 1 or abs(op.qubit - op.target) == 1:
            out.append(op)
            continue
        q, t = op.qubit, op.target
        step = 1 if t > q else -1
        path = list(range(q, t + step, step))
        sw = [(path[i], path[i + 1]) for i in range(len(path) - 2)]
        new_q = path[-2] if q < t else t + 1
        new = O.ref_gate(op.name, new_q, t, noisy=op.noisy)
        out += [O.ref_gate("SWAP", a, b, noisy=False) for a, b in sw] + [new] + \
               [O.ref_gate("SWAP", a, b, noisy=False) for a, b in reversed(sw)]
    return out


def depth(ops: Sequence) -> int:
    """_calculate_circuit_depth (ops.jl:250-278)."""
    layer: Dict[int, int] = {}
    d = 0
    for op in ops:
        m = layer.get(op.qubit, 0)
        if op.q == 2:
            m = max(m, layer.get(op.target, 0))
        layer[op.qubit] = m + 1
        if op.q == 2:
            layer[op.target] = m + 1
        d = max(d, m + 1)
    return d


TWIRL_CNOT = "YYXZ YZXY YIYX YXYI ZYIY ZZIZ ZIZI ZXZX IYZY IZZZ IIII IXIX XYYZ XZYY XIXX XXXI".split()
TWIRL_CNOT_REV = "YYZX YZYI YIYZ YXZY ZYYX ZZZI ZIZZ ZXYY IYXY IZIZ IIII IXXX XYIY XZXZ XIXI XXIX".split()


def shot_ops(ops: Sequence, id: int = 0, twirl: bool = False, choices: Sequence[int] = ()) -> List:
    """_expanded_ops_for_shot (ops.jl:555-573) on CNOT/CX ops, with apply_twirl
    (noise.jl:232-255) taking rows from `choices`."""
    it = iter(choices)
    out = []

    def tw(op):
        table = TWIRL_CNOT if op.qubit < op.target else TWIRL_CNOT_REV
        row = table[next(it)]
        least = min(op.qubit, op.target)
        r = [O.ref_gate(row[0], least), O.ref_gate(row[1], least + 1), op, O.ref_gate(row[2], least),
             O.ref_gate(row[3], least + 1)]
        return r

    for op in ops:
        cnot = op.kind == O.KIND_GATE and op.q == 2 and op.name.upper() in ("CNOT", "CX")
        if id > 0 and cnot:
            for _ in range(2 * id + 1):
                out.extend(tw(op) if twirl else [op])
        elif twirl and cnot:
            out.extend(tw(op))
        else:
            out.append(op)
    return out
