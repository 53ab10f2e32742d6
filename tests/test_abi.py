"""CPU checks of the C-ABI boundary: the library loads and exports every
entry point include/bt.h declares; host-side logic of the Python mirror
(gate table vs the oracle's, op construction and validation)."""
import math
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "bt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(bt):
    import ctypes
    from bluetangle_jl_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "libbt.so not built"
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding covers the same set
    assert sorted(_lib.SYMBOLS) == syms


def test_uniform_matches_oracle(bt):
    from bluetangle_jl_amd import _lib
    L = _lib.lib()
    for seed, traj, draw in [(0, 0, 0), (5, 9, 1234), (2 ** 63 + 1, 4095, 99)]:
        assert L.bt_uniform(seed, traj, draw) == O.lib().bto_uniform(seed, traj, draw)


def test_gate_table_matches_oracle_table(bt):
    for name in list(O.GATE):
        assert np.allclose(bt.gates(name), O.gate_matrix(name), atol=0, rtol=0) or \
            np.allclose(bt.gates(name), O.gate_matrix(name), atol=1e-15), name
    for name in ["RX(0.3)", "RY(.1π)", "RZ(2*pi/3)", "P(1.2)", "U2(0.1,0.2)", "U3(0.1,0.2,0.3)", "CP(0.4)",
                 "GIVENS(0.5)", "FSIM(0.2,0.9)", "SWAPA(0.3)", "RXX(0.7)", "RYY(0.7)", "RZZ(0.7)", "RXY(0.7)"]:
        assert np.allclose(bt.gates(name), O.gate_matrix(name), atol=1e-15), name


def test_noise_tables_match_oracle(bt):
    for model in ("amplitude_damping", "phase_damping", "depolarizing", "bit_flip", "phase_flip",
                  "bit_phase_flip", "depolarizing_amp", "rot_x", "rot_y", "rot_z", "rot_p", "rot_xyz"):
        for two in (False, True):
            a, b = bt.noise_model(model, 0.17, two), O.noise_kraus(model, 0.17, two)
            assert len(a) == len(b)
            assert all(np.allclose(x, y, atol=1e-15) for x, y in zip(a, b)), model


def test_op_construction_rules(bt):
    """struct.jl:365-474 validation and constructors."""
    op = bt.Op("CNOT", 1, 2)
    assert op.q == 2 and op.type == "op"
    assert bt.Op("RX(0.2)", 3).type == "phase"
    m = bt.Op("MZ", 2)
    assert m.type == "🔬" and m.noisy is False
    with pytest.raises(ValueError):
        bt.Op("MZ", 2, control=1)
    with pytest.raises(ValueError):
        bt.Op("X", 2, control=2)
    with pytest.raises(ValueError):
        bt.Op("CNOT", 2, 2)
    ccx = bt.Op("CCX", 1, 2, 3)
    assert ccx.name == "CX" and ccx.control == 2 and ccx.qubit == 1 and ccx.target_qubit == 3
    res = bt.Op("RES", 1)
    assert isinstance(res, bt.OpQC) and res.noisy is False
    qc = bt.OpQC("amplitude_damping", 0.1, 2)
    assert qc.q == 1 and qc.type == repr(0.1)
    with pytest.raises(ValueError):
        bt.OpQC("depolarizing", 0.1, 1, 2).__class__ and bt.OpQC("custom", [np.eye(2) * 2], 1)
    with pytest.raises(ValueError):
        bt.NoiseModel(bt.Noise2("depolarizing", 0.1), bt.Noise1("depolarizing", 0.1))
    nm = bt.NoiseModel("depolarizing", 0.1)
    assert nm.q1.q == 1 and nm.q2.q == 2 and len(nm.q2.kraus) == 16
    f = bt.Op("f", lambda t: bt.gates("RX(%r)" % t), 1)
    assert f.type == "f"


def test_channel_bits_convention(bt):
    from bluetangle_jl_amd.core import channel_bits
    # N=5: qubit q <-> bit 5-q; 2q probabilities in ascending qubit order
    assert channel_bits(5, 1, 2, -1) == ([3], [3])
    assert channel_bits(5, 2, 4, 2) == ([1, 3], [3, 1])
    assert channel_bits(5, 3, 2, -1) == ([3, 2, 1], [3, 2, 1])
