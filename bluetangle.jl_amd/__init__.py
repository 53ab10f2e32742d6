"""MI355X-native engine for BlueTangle.jl's noisy-circuit apply path.

Mirrors the reference's Julia API (Op, OpQC, ifOp, OpF, QuantumChannel,
NoiseModel, apply, to_state, to_rho, run, sample_exact, expect, correlation) over the C ABI of
libbt.so (include/bt.h).  There is no CPU fallback.
"""
from ._lib import BTError
from .gates import gate, gates
from .noise import (Noise1, Noise2, NoiseModel, QuantumChannel, custom_noise, is_valid_quantum_channel,
                    noise_model, noise_model1, noise_model2)
from .ops import Op, OpF, OpQC, QuantumOps, ifOp
from .core import (DensityMatrix, DeviceState, StateVector, UniformStream, apply, apply_noise, current_stream,
                   from_numpy, get_N, init, sample_exact, seed, zero_rho, zero_state)
from .circuit import Circuit, Server, compile_ir, run, to_rho, to_state
from .observables import correlation, expand_multi_op, expect
from . import measurement
from .qasm import from_qasm, to_qasm
from .compile import (CompiledCircuit, Layout, Options, apply_twirl, compile, get_layers, get_optimized_layers,
                      get_stats, optimize_simple)
from .measurement import (Measurement, get_probs_from_sample, measure, measure_circuit, sample, sample_state)

__all__ = [
    "BTError", "gate", "gates", "Noise1", "Noise2", "NoiseModel", "QuantumChannel", "custom_noise",
    "is_valid_quantum_channel", "noise_model", "noise_model1", "noise_model2", "Op", "OpF", "OpQC", "QuantumOps",
    "ifOp", "DensityMatrix", "DeviceState", "StateVector", "UniformStream", "apply", "apply_noise",
    "current_stream", "correlation", "Measurement", "get_probs_from_sample", "measure", "measure_circuit",
    "measurement", "sample", "sample_state", "CompiledCircuit", "Layout", "Options", "apply_twirl", "compile",
    "get_layers", "get_optimized_layers", "get_stats", "optimize_simple", "from_qasm", "to_qasm", "expand_multi_op", "expect", "from_numpy", "get_N", "init", "sample_exact", "seed", "zero_rho", "zero_state",
    "Circuit", "Server", "compile_ir", "run", "to_rho", "to_state",
]
