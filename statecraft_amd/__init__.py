"""statecraft_amd -- MI355X (gfx950) statevector engine for StateCraft.

The product is libstatecraft_hip.so (C-ABI in include/statecraft_hip.h, HIP
kernels in csrc/).  This package holds the in-tree build recipe and the
host-side mirror of StateCraft's engine API (circuit, engine, noise, rng) used
by tests and the benchmark.
"""
from . import capi
from .circuit import QuantumCircuit, lower

__all__ = ["capi", "QuantumCircuit", "lower"]
