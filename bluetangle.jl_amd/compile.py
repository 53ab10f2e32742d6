"""Circuit compilation: layout swaps, layering, statistics, and the per-shot
op expansion (ZNE, Pauli twirling) that feeds the engine's IR.

Mirrors compile (src/ops.jl:421-445), get_stats / _calculate_circuit_depth /
get_layers (src/ops.jl:191-329), optimize_simple (src/ops.jl:335-402), the
Layout struct and its swap insertion (src/struct.jl:770-825,
src/layout.jl:1-110), Options and Circuit (src/struct.jl:835-873),
_expanded_ops_for_shot (src/ops.jl:555-573) and apply_twirl / _twirl_set
(src/noise.jl:136-255).

The reference's drivers apply the ops of vcat(circuit.layers...), not the
order they were written in, so a shot consumes its uniforms in layer order;
run / to_state / to_rho / measure_circuit here do the same.  Twirl choices
(rand(list) in the reference) take the first draws of a shot's stream, one
per twirled gate in op order (index floor(16 u)), because
_expanded_ops_for_shot draws them before the shot's apply; the apply then
consumes the stream from there on.
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from .noise import NoiseModel, QuantumChannel
from .ops import Op, OpF, OpQC, QuantumOps, ifOp


# --- Options / Layout ----------------------------------------------------------
@dataclass
class Options:
    """Options (struct.jl:835-854); measurement_basis is upper-cased."""
    circuit_name: str = "circuit"
    measurement_basis: str = "Z"
    noise: Union[NoiseModel, bool] = False
    twirl: bool = False
    readout_noise: Union[QuantumChannel, bool] = False
    measurement_mitigate: bool = False
    density_matrix: bool = False

    def __post_init__(self):
        self.measurement_basis = self.measurement_basis.upper()


class Layout:
    """Layout (struct.jl:770-825, layout.jl:1-110): physical qubits are the
    nonzero cells of a grid, numbered in column-major order; neighbours are
    the 4-connected cells.  swap(op) inserts SWAP (or FSWAP) pairs around a
    non-local 2-qubit op along a shortest path.  The reference walks a
    Julia Set of neighbours, whose order decides between equally short paths
    on a 2-D grid; here neighbours are taken in ascending order (a line,
    the default layout, has one shortest path, so there the ops agree)."""

    def __init__(self, layout, noisy_swap: bool = False, fswap: bool = False):
        g = np.asarray(layout)
        if g.ndim == 1:
            g = g.reshape(1, -1)
        self.layout = g
        cells = [(r, c) for c in range(g.shape[1]) for r in range(g.shape[0]) if g[r, c] > 0]
        self.num_to_pos = {i + 1: p for i, p in enumerate(cells)}
        pos_to_num = {p: i for i, p in self.num_to_pos.items()}
        self.N = len(cells)
        self.neighbors: Dict[int, List[int]] = {}
        for q, (r, c) in self.num_to_pos.items():
            nb = [pos_to_num[(r + dr, c + dc)] for dr, dc in ((-1, 0), (1, 0), (0, -1), (0, 1))
                  if (r + dr, c + dc) in pos_to_num]
            self.neighbors[q] = sorted(nb)
        self.connectivity = (sum(len(v) for v in self.neighbors.values()) / self.N) if self.N else 0.0
        self.noisy_swap = noisy_swap
        self.fswap = fswap

    def check(self, q1: int, q2: int) -> bool:
        return q1 in self.neighbors[q2]

    def shortest_path(self, start: int, end: int) -> List[int]:
        """_find_shortest_path (layout.jl:92-110): breadth-first."""
        visited = set()
        queue = deque([(start, [start])])
        while queue:
            cur, path = queue.popleft()
            if cur == end:
                return path
            for nb in self.neighbors[cur]:
                if nb not in visited:
                    visited.add(nb)
                    queue.append((nb, path + [nb]))
        return []

    def swap(self, op: QuantumOps) -> List[QuantumOps]:
        """_make_swaps (layout.jl:69-90)."""
        if op.q == 1 or op.qubit in self.neighbors[op.target_qubit]:
            return [op]
        name = "FSWAP" if self.fswap else "SWAP"
        path = self.shortest_path(op.qubit, op.target_qubit)
        swaps = [(path[i], path[i + 1]) for i in range(len(path) - 2)]
        # the op is rebuilt from its name next to the target (layout.jl:86)
        new_q = path[-2] if op.qubit < op.target_qubit else op.target_qubit + 1
        new_op = Op(op.name, new_q, op.target_qubit, noisy=op.noisy)
        pre = [Op(name, a, b, noisy=self.noisy_swap) for a, b in swaps]
        post = [Op(name, a, b, noisy=self.noisy_swap) for a, b in reversed(swaps)]
        return pre + [new_op] + post


# --- statistics and layers -------------------------------------------------------
def _target_find(op) -> int:
    """struct.jl:746-753."""
    return -1 if isinstance(op, (ifOp, OpF)) else op.target_qubit


def _control_find(op) -> int:
    """struct.jl:755-761."""
    return -2 if isinstance(op, (ifOp, OpQC, OpF)) else op.control


def calculate_circuit_depth(ops: Sequence[QuantumOps]) -> int:
    """_calculate_circuit_depth (ops.jl:250-278): controls do not count."""
    layer: Dict[int, int] = {}
    depth = 0
    for op in ops:
        m = layer.get(op.qubit, 0)
        if op.q == 2:
            m = max(m, layer.get(op.target_qubit, 0))
        cur = m + 1
        depth = max(depth, cur)
        layer[op.qubit] = cur
        if op.q == 2:
            layer[op.target_qubit] = cur
    return depth


def get_stats(ops: Sequence[QuantumOps]) -> Dict[str, int]:
    """get_stats (ops.jl:191-240)."""
    ops = [o for o in ops if not isinstance(o, OpF)]
    N = length = cx = sw = twoq = mid = 0
    for op in ops:
        length += 1
        name = op.name.upper()
        if isinstance(op, ifOp):
            cur = op.qubit
        elif isinstance(op, OpQC):
            cur = max(op.qubit, op.target_qubit)
        else:
            cur = max(op.qubit, op.target_qubit, op.control)
        N = max(N, cur)
        if name in ("CX", "CNOT"):
            cx += 1
        if name == "SWAP":
            sw += 1
        if op.q == 2:
            twoq += 1
        if op.type == "🔬":
            mid += 1
    return dict(N=N, len=length, depth=calculate_circuit_depth(ops), cx_count=cx, swap_count=sw,
                two_qubit_count=twoq, mid_measurement_count=mid)


def get_layers(ops: Sequence[QuantumOps]) -> List[List[QuantumOps]]:
    """get_layers (ops.jl:283-329): each op goes one layer past the last
    layer that touched its qubits (target and control included); an OpF is
    a layer of its own that fences every qubit."""
    layers: List[List[QuantumOps]] = []
    ql: Dict[int, int] = {}
    nonf = [o for o in ops if not isinstance(o, OpF)]
    max_q = max(max(o.qubit, _target_find(o), _control_find(o)) for o in nonf) if nonf else 0
    for op in ops:
        if isinstance(op, OpF):
            layers.append([op])
            for q in range(1, max_q + 1):
                ql[q] = len(layers)
            continue
        lay = 0
        if op.q == 2:
            lay = max(ql.get(op.qubit, 0), ql.get(op.target_qubit, 0), ql.get(op.control, 0)) + 1
        elif op.q == 1:
            if isinstance(op, ifOp):
                lay = ql.get(op.qubit, 0) + 1
            else:
                lay = max(ql.get(op.qubit, 0), ql.get(op.control, 0)) + 1
        while len(layers) < lay:
            layers.append([])
        layers[lay - 1].append(op)
        ql[op.qubit] = lay
        if op.q == 2:
            ql[op.target_qubit] = lay
        if isinstance(op, Op) and op.control != -2:
            ql[op.control] = lay
    return layers


def _remove_duplicate(layers, l, i, num_layers) -> bool:
    """ops.jl:370-388 (0-based l, i)."""
    op = layers[l][i]

    def find(layer):
        for k, x in enumerate(layer):
            if x.name == op.name and x.qubit == op.qubit and x.control == -2:
                return k
        return None

    for j in range(1, num_layers + 1):
        if l + j >= len(layers) or find(layers[l + j]) is None:
            return False
    del layers[l][i]
    for j in range(1, num_layers + 1):
        del layers[l + j][find(layers[l + j])]
    return True


def _remove_two_qubit_duplicate(layers, l, i) -> bool:
    """ops.jl:390-402."""
    op = layers[l][i]
    if l + 1 >= len(layers):
        return False
    for k, x in enumerate(layers[l + 1]):
        if x.name == op.name and x.qubit == op.qubit and x.target_qubit == op.target_qubit and x.control == -2:
            del layers[l][i]
            del layers[l + 1][k]
            return True
    return False


def optimize_simple(layers: List[List[QuantumOps]]) -> List[List[QuantumOps]]:
    """optimize_simple (ops.jl:335-368): removes an X/Y/Z/H followed in the
    next layer by the same gate on the same qubit, S/SX/XSQRT repeated over
    the next three layers, and adjacent repeated CX/CNOT/CY/CZ/ECR/SWAP/FSWAP
    pairs, re-layering until the layer count stops changing."""
    layers = [list(l) for l in layers]
    n = -1
    while n != len(layers):
        n = len(layers)
        for l in range(len(layers)):
            i = 0
            while i < len(layers[l]):
                op = layers[l][i]
                removed = False
                if op.q == 1 and op.control == -2 and l < len(layers) - 1:
                    if op.name in ("X", "Y", "Z", "H"):
                        removed = _remove_duplicate(layers, l, i, 1)
                    elif op.name in ("S", "SX", "XSQRT") and l + 3 <= len(layers) - 1:
                        removed = _remove_duplicate(layers, l, i, 3)
                elif (op.q == 2 and op.control == -2 and l < len(layers) - 1 and
                      op.name in ("CX", "CNOT", "CY", "CZ", "ECR", "SWAP", "FSWAP")):
                    removed = _remove_two_qubit_duplicate(layers, l, i)
                if not removed:
                    i += 1
        layers = get_layers([o for lay in layers for o in lay])
    return layers


def get_optimized_layers(ops: Sequence[QuantumOps]) -> List[List[QuantumOps]]:
    """ops.jl:407."""
    return optimize_simple(get_layers(ops))


# --- twirling and ZNE ------------------------------------------------------------
def _sets(rows: str) -> List[Tuple[str, str, str, str]]:
    return [tuple(r) for r in rows.split()]


# Pauli frames (P1 P2 before, P3 P4 after, on (least, least + 1)) that leave
# the gate unchanged: noise.jl:138-218, the same 16 rows in the same order
_TWIRL = {
    "ecr": _sets("YYYY YZZI YIZZ YXYX ZYZY ZZYI ZIYZ ZXZX IYXX IZIZ IIII IXXY XYIX XZXZ XIXI XXIY"),
    "ecr_rev": _sets("YYYY YZYZ YIXX YXXI ZYIZ ZZIY ZIZI ZXZX IYZZ IZZY IIII IXIX XYXY XZXZ XIYX XXYI"),
    "cnot": _sets("YYXZ YZXY YIYX YXYI ZYIY ZZIZ ZIZI ZXZX IYZY IZZZ IIII IXIX XYYZ XZYY XIXX XXXI"),
    "cnot_rev": _sets("YYZX YZYI YIYZ YXZY ZYYX ZZZI ZIZZ ZXYY IYXY IZIZ IIII IXXX XYIY XZXZ XIXI XXIX"),
}


def twirl_set(name: str, qubit: int, target_qubit: int, choice: int) -> Tuple[str, str, str, str]:
    """_twirl_set (noise.jl:136-222) with the row given by `choice` (0..15)."""
    ecr = name.upper() == "ECR"
    key = ("ecr" if ecr else "cnot") + ("" if qubit < target_qubit else "_rev")
    return _TWIRL[key][choice]


def is_twirlable(op) -> bool:
    return op.q == 2 and op.name.upper() in ("CNOT", "CX", "ECR")


def apply_twirl(op: QuantumOps, choice: int) -> List[QuantumOps]:
    """apply_twirl (noise.jl:232-255): P1(least), P2(least+1), op, P3(least),
    P4(least+1), the Paulis as ordinary (noisy) gates."""
    if not is_twirlable(op):
        return [op]
    least = min(op.qubit, op.target_qubit)
    out: List[QuantumOps] = []
    for i, p in enumerate(twirl_set(op.name, op.qubit, op.target_qubit, choice), start=1):
        fq = least if i % 2 == 1 else least + 1
        if i == 3:
            out.append(op)
        out.append(Op(p, fq))
    return out


def _is_cnot_like(op) -> bool:
    """ops.jl:526."""
    return op.q == 2 and op.name.upper() in ("CNOT", "CX")


def twirl_choice(u: float) -> int:
    return min(int(u * 16.0), 15)


# --- Circuit ---------------------------------------------------------------------
@dataclass
class CompiledCircuit:
    """Circuit (struct.jl:867-873): stats after swap insertion, stats of the
    ops as given, options, layout, layers."""
    stats: Dict[str, int]
    stats_init: Dict[str, int]
    options: Options
    layout: Layout
    layers: List[List[QuantumOps]]

    @property
    def N(self) -> int:
        return self.stats["N"]

    def ops(self) -> List[QuantumOps]:
        return [o for lay in self.layers for o in lay]

    def twirl_count(self, id: int = 0) -> int:
        """Twirled gates of one shot (= twirl draws at the head of its stream)."""
        if not self.options.twirl:
            return 0
        n = 0
        for op in self.ops():
            if _is_cnot_like(op):
                n += (2 * id + 1) if id > 0 else 1
        return n

    def ops_for_shot(self, id: int = 0, choices: Sequence[int] = ()) -> List[QuantumOps]:
        """_expanded_ops_for_shot (ops.jl:555-573): with id > 0 every CNOT/CX
        becomes 2 id + 1 copies (ZNE), each twirled when options.twirl, taking
        the twirl rows from `choices` in order."""
        it = iter(choices)
        out: List[QuantumOps] = []
        tw = self.options.twirl
        for op in self.ops():
            if id > 0 and _is_cnot_like(op):
                for _ in range(2 * id + 1):
                    out.extend(apply_twirl(op, next(it)) if tw else [op])
            elif tw and _is_cnot_like(op):
                out.extend(apply_twirl(op, next(it)))
            else:
                out.append(op)
        return out

    def rho_ops(self, id: int = 0, choices: Sequence[int] = ()) -> List[QuantumOps]:
        """The op sequence of to_rho (ops.jl:806-844): ZNE repeats ops named
        exactly "CNOT" or "CX" (case-sensitive there); with options.twirl every
        2-qubit op goes through apply_twirl (which twirls CNOT/CX/ECR)."""
        it = iter(choices)
        out: List[QuantumOps] = []
        tw = self.options.twirl
        for op in self.ops():
            if op.q == 2 and op.name in ("CNOT", "CX") and id > 0:
                for _ in range(2 * id + 1):
                    out.extend(apply_twirl(op, next(it)) if tw else [op])
            elif op.q == 2 and tw:
                out.extend(apply_twirl(op, next(it)) if is_twirlable(op) else [op])
            else:
                out.append(op)
        return out

    def rho_twirl_count(self, id: int = 0) -> int:
        if not self.options.twirl:
            return 0
        n = 0
        for op in self.ops():
            if op.q == 2 and op.name in ("CNOT", "CX") and id > 0:
                n += 2 * id + 1
            elif op.q == 2 and is_twirlable(op):
                n += 1
        return n


def compile(ops: Sequence[QuantumOps], options: Optional[Options] = None,
            layout: Union[Layout, int] = -1) -> CompiledCircuit:
    """compile (ops.jl:421-445)."""
    options = options or Options()
    stats_init = get_stats(ops)
    N = stats_init["N"]
    if isinstance(layout, int) and layout == -1:
        layout = Layout(np.ones((1, N), dtype=int))
    elif N > layout.N:
        raise ValueError("The number of physical qubits in the layout does not match the requirements of the "
                         "operations.")
    local: List[QuantumOps] = []
    for op in ops:
        local.extend(layout.swap(op) if not isinstance(op, (OpF, ifOp)) else [op])
    return CompiledCircuit(get_stats(local), stats_init, options, layout, get_layers(local))
