"""OpenQASM 2 import / export: the wire format of circuits handed to run and
to the serving loop.

Mirrors from_qasm and to_qasm of src/qasm.jl:33-421 and run(qasm_or_path)
(src/ops.jl:694-727).  Import lower-cases each statement, skips the
boilerplate (OPENQASM, include, qreg, creg, opaque, gate, barrier), and maps
statements onto ops with 1-based qubits (qasm index + 1, any register name):
cx/cy/cz become controlled one-qubit X/Y/Z (so their noise follows the
controlled-1q route of apply_noise), cu1/cp a controlled U1, cu3 with zero
phi and lambda a controlled RY, ccx a b c the three-qubit Op("CCX", a, b, c),
cswap a b c Op("CSWAP", b, a, c).  Export writes the header, the sorted
opaque / gate declarations the body needs, the registers and one or more
statements per op.
"""
from __future__ import annotations

import os
import re
from typing import List, Optional, Sequence, Tuple

from .ops import Op, OpQC, QuantumOps


def _qubit(tok: str) -> int:
    m = re.search(r"\[(\d+)\]\s*[;,]?\s*$", tok.strip())
    if m is None:
        raise ValueError("Expected qubit like q[<int>]; got: %s" % tok)
    return int(m.group(1)) + 1


_Q = r"([^\s,;]+)"
_A = r"\s*\(\s*(.+?)\s*\)"
_RULES: List[Tuple[re.Pattern, str]] = [(re.compile(p), k) for p, k in [
    (r"^measure\s+([^\s]+)\s*->\s*([^\s]+)\s*;", "measure"),
    (r"^reset\s+([^\s;]+)\s*;", "reset"),
    (r"^(x|y|z|h|s|sdg|t|tdg|sx)\s+([^\s;]+)\s*;", "unary"),
    (r"^(rx|ry|rz)" + _A + r"\s+([^\s;]+)\s*;", "rot"),
    (r"^u1" + _A + r"\s+([^\s;]+)\s*;", "u1"),
    (r"^u2\s*\(\s*(.+?)\s*,\s*(.+?)\s*\)\s+([^\s;]+)\s*;", "u2"),
    (r"^u3\s*\(\s*(.+?)\s*,\s*(.+?)\s*,\s*(.+?)\s*\)\s+([^\s;]+)\s*;", "u3"),
    (r"^(?:cu1|cp)" + _A + r"\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "cu1"),
    (r"^cu3\s*\(\s*(.+?)\s*,\s*(.+?)\s*,\s*(.+?)\s*\)\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "cu3"),
    (r"^(crx|cry|crz)" + _A + r"\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "crot"),
    (r"^(cx|cnot|cy|cz|swap|iswap|fswap|syc|ecr)\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "binary"),
    (r"^(ccx|ccy|ccz|cswap)\s+" + _Q + r"\s*,\s*" + _Q + r"\s*,\s*([^\s;]+)\s*;", "ternary"),
    (r"^(rxx|ryy|rzz|rxy)" + _A + r"\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "binrot"),
    (r"^givens" + _A + r"\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "givens"),
    (r"^fsim\s*\(\s*(.+?)\s*,\s*(.+?)\s*\)\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "fsim"),
    (r"^swapa" + _A + r"\s+" + _Q + r"\s*,\s*([^\s;]+)\s*;", "swapa"),
]]
_SKIP = ("openqasm", "include", "qreg", "creg", "opaque", "gate", "barrier")
_UNARY = {"sdg": "SD", "tdg": "TD", "sx": "XSQRT"}


def _statement_ops(stmt: str) -> List[QuantumOps]:
    s = stmt.lower()
    if s.startswith(_SKIP):
        return []
    for pat, kind in _RULES:
        m = pat.match(s)
        if m is None:
            continue
        g = [x.strip() for x in m.groups()]
        if kind == "measure":
            return [Op("MZ", _qubit(g[0]))]
        if kind == "reset":
            return [Op("RES", _qubit(g[0]))]
        if kind == "unary":
            return [Op(_UNARY.get(g[0], g[0].upper()), _qubit(g[1]))]
        if kind == "rot":
            return [Op("%s(%s)" % (g[0], g[1]), _qubit(g[2]))]
        if kind == "u1":
            return [Op("U1(%s)" % g[0], _qubit(g[1]))]
        if kind == "u2":
            return [Op("U2(%s,%s)" % (g[0], g[1]), _qubit(g[2]))]
        if kind == "u3":
            return [Op("U3(%s,%s,%s)" % (g[0], g[1], g[2]), _qubit(g[3]))]
        if kind == "cu1":
            return [Op("U1(%s)" % g[0], _qubit(g[2]), control=_qubit(g[1]))]
        if kind == "cu3":
            th, ph, la, c, t = g[0], g[1], g[2], _qubit(g[3]), _qubit(g[4])
            if ph == "0" and la == "0":
                return [Op("ry(%s)" % th, t, control=c)]
            return [Op("U3(%s,%s,%s)" % (th, ph, la), t, control=c)]
        if kind == "crot":
            return [Op("%s(%s)" % (g[0][1:], g[1]), _qubit(g[3]), control=_qubit(g[2]))]
        if kind == "binary":
            a, b = _qubit(g[1]), _qubit(g[2])
            single = {"cx": "X", "cnot": "X", "cy": "Y", "cz": "Z"}.get(g[0])
            if single:
                return [Op(single, b, control=a)]
            return [Op(g[0].upper(), a, b)]
        if kind == "ternary":
            a, b, c = _qubit(g[1]), _qubit(g[2]), _qubit(g[3])
            if g[0] == "cswap":
                return [Op("CSWAP", b, a, c)]
            return [Op(g[0].upper(), a, b, c)]
        if kind == "binrot":
            return [Op("%s(%s)" % (g[0].upper(), g[1]), _qubit(g[2]), _qubit(g[3]))]
        if kind == "givens":
            return [Op("GIVENS(%s)" % g[0], _qubit(g[1]), _qubit(g[2]))]
        if kind == "fsim":
            return [Op("FSIM(%s,%s)" % (g[0], g[1]), _qubit(g[2]), _qubit(g[3]))]
        if kind == "swapa":
            return [Op("SWAPA(%s)" % g[0], _qubit(g[1]), _qubit(g[2]))]
    raise ValueError("Unsupported or unrecognized statement: %s" % stmt)


def qasm_source(qasm_or_path: str, is_file: Optional[bool] = None) -> str:
    """_qasm_to_source (ops.jl:694-712)."""
    if is_file is True:
        with open(qasm_or_path) as f:
            return f.read()
    if is_file is False:
        return str(qasm_or_path)
    if os.path.isfile(qasm_or_path):
        with open(qasm_or_path) as f:
            return f.read()
    t = str(qasm_or_path)
    if ";" in t or "\n" in t or "openqasm" in t.lower():
        return t
    raise ValueError("Input is neither an existing file nor recognizable OpenQASM text: `%s`" % qasm_or_path)


def from_qasm(qasm: str) -> List[QuantumOps]:
    """from_qasm (qasm.jl:275-421): a file path or QASM text -> ops."""
    text = qasm
    if "\n" not in qasm and ";" not in qasm and os.path.isfile(qasm):
        with open(qasm) as f:
            text = f.read()
    src = re.sub(r"//.*", "", text)
    src = "\n".join(line.strip() for line in src.split("\n"))
    ops: List[QuantumOps] = []
    for part in src.split(";"):
        if part.strip():
            ops.extend(_statement_ops(part.strip() + ";"))
    return ops


# --- export ---------------------------------------------------------------------
def _split(name: str):
    if "(" in name:
        head, rest = name.split("(", 1)
        return head.upper(), rest.replace(")", "").rstrip()
    return name.upper(), None


def to_qasm(ops_or_circuit, nqubits: Optional[int] = None, qreg: str = "q", creg: str = "c",
            filename: Optional[str] = None) -> str:
    """to_qasm (qasm.jl:33-255); a compiled circuit is exported as
    vcat(layers...)."""
    ops = ops_or_circuit.ops() if hasattr(ops_or_circuit, "layers") else list(ops_or_circuit)
    if not ops:
        raise ValueError("ops vector is empty")

    def maxidx(op):
        m = op.qubit
        if op.target_qubit > 0:
            m = max(m, op.target_qubit)
        if op.control > 0:
            m = max(m, op.control)
        return m

    N = max(maxidx(o) for o in ops) if nqubits is None else nqubits
    if N < 1:
        raise ValueError("nqubits must be >= 1")
    opaques = set()

    def lines(op) -> List[str]:
        if isinstance(op, OpQC):
            if op.name.split("(")[0].upper() in ("RES", "RESET"):
                return ["reset %s[%d];" % (qreg, op.qubit - 1)]
            raise ValueError("Cannot export quantum channel '%s' to OpenQASM; only RES/RESET (gamma=1) is "
                             "supported." % op.name)
        if not isinstance(op, Op):
            raise ValueError("Unsupported operation for export: %s." % type(op).__name__)
        q, t, c = op.qubit, op.target_qubit, op.control
        Q = "%s[%d]" % (qreg, q - 1)
        T = "%s[%d]" % (qreg, t - 1) if t >= 1 else None
        Cq = "%s[%d]" % (qreg, c - 1) if c >= 1 else None
        nameU = op.name.split("(")[0].upper()
        base, args = _split(op.name)
        # measurements are matched on the name before any "(" (qasm.jl:96-103),
        # so only the MZ / MX / MY spellings export
        if nameU == "MZ":
            return ["measure %s -> %s[%d];" % (Q, creg, q - 1)]
        if nameU == "MX":
            return ["h %s;" % Q, "measure %s -> %s[%d];" % (Q, creg, q - 1)]
        if nameU == "MY":
            return ["sdg %s;" % Q, "h %s;" % Q, "measure %s -> %s[%d];" % (Q, creg, q - 1)]
        if nameU in ("RES", "RESET"):
            return ["reset %s;" % Q]
        if t < 0:
            if c >= 1:
                if base in ("X", "Y", "Z"):
                    return ["c%s %s,%s;" % (base.lower(), Cq, Q)]
                if base in ("RX", "RY", "RZ"):
                    g = "c" + base.lower()
                    opaques.add("opaque %s(theta) a,b;" % g)
                    return ["%s(%s) %s,%s;" % (g, args, Cq, Q)]
                if base in ("P", "U1"):
                    return ["cu1(%s) %s,%s;" % (args, Cq, Q)]
                if base == "U2":
                    return ["cu3(pi/2,%s) %s,%s;" % (args, Cq, Q)]
                if base == "U3":
                    return ["cu3(%s) %s,%s;" % (args, Cq, Q)]
                raise ValueError("Controlled %s not supported" % op.name)
            if base in ("P", "U1"):
                return ["u1(%s) %s;" % (args, Q)]
            if base in ("U2", "U3"):
                return ["%s(%s) %s;" % (base.lower(), args, Q)]
            if base in ("RX", "RY", "RZ"):
                return ["%s(%s) %s;" % (base.lower(), args if args is not None else "", Q)]
            simple = {"X": "x", "Y": "y", "Z": "z", "H": "h", "S": "s", "SD": "sdg", "T": "t", "TD": "tdg",
                      "SX": "sx", "XSQRT": "sx"}
            if nameU == "I":
                return []
            if nameU in simple:
                return ["%s %s;" % (simple[nameU], Q)]
            raise ValueError("Unsupported 1-qubit op: %s" % op.name)
        if base in ("CX", "CNOT"):
            return ["ccx %s,%s,%s;" % (Q, Cq, T)] if c >= 1 else ["cx %s,%s;" % (Q, T)]
        if base in ("CY", "CZ"):
            if c >= 1:
                body = "s c; ccx a,b,c; sdg c;" if base == "CY" else "h c; ccx a,b,c; h c;"
                opaques.add("gate c%s a,b,c { %s }" % (base.lower(), body))
                return ["c%s %s,%s,%s;" % (base.lower(), Q, Cq, T)]
            return ["%s %s,%s;" % (base.lower(), Q, T)]
        if base == "SWAP":
            return ["cswap %s,%s,%s;" % (Cq, Q, T)] if c >= 1 else ["swap %s,%s;" % (Q, T)]
        if base in ("ISWAP", "FSWAP", "SYC", "ECR"):
            opaques.add("opaque %s a,b;" % base.lower())
            return ["%s %s,%s;" % (base.lower(), Q, T)]
        if base == "CP":
            return ["cu1(%s) %s,%s;" % (args, Q, T)]
        params = {"RXX": "theta", "RYY": "theta", "RZZ": "theta", "RXY": "theta", "GIVENS": "theta",
                  "FSIM": "theta,phi", "SWAPA": "alpha"}
        if base in params:
            opaques.add("opaque %s(%s) a,b;" % (base.lower(), params[base]))
            return ["%s(%s) %s,%s;" % (base.lower(), args, Q, T)]
        raise ValueError("Unsupported operation for export: %s" % op.name)

    body: List[str] = []
    for op in ops:
        body.extend(lines(op))
    out = "\n".join(['OPENQASM 2.0;', 'include "qelib1.inc";'] + sorted(opaques) +
                    ["qreg %s[%d];" % (qreg, N), "creg %s[%d];" % (creg, N)] + body) + "\n"
    if filename is not None:
        if not filename:
            raise ValueError("filename must be a non-empty string")
        with open(filename, "w") as f:
            f.write(out)
    return out
