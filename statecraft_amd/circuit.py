"""QuantumCircuit mirror and its lowering to the C-ABI op buffer.

Names, argument order and validation follow the reference's immutable circuit
record (core/src/main/java/com/omaarr90/statecraft/quantum/QuantumCircuit.java:9-574)
and gate records (Hadamard.java, PauliX/Y/Z.java, SGate.java, SdgGate.java,
MatrixSingleQubitGate.java).  Rotation matrices follow
OpenQasmCircuitParser.rotationX/Y/Z/phaseGate/uGate
(core/src/main/java/com/omaarr90/statecraft/core/parse/OpenQasmCircuitParser.java:646-684);
zz_phase and rx follow AlgorithmSuiteSpecs.zzPhase/rx
(app/src/main/java/com/omaarr90/statecraft/AlgorithmSuiteSpecs.java:283-296).

Parameterised gates optionally carry gradient metadata (param index, Pauli
generator, coefficient) so the adjoint engine can differentiate them; this
has no reference counterpart.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import capi
from .capi import IllegalArgumentException, NullPointerException, UnsupportedOperationException


# ---------------------------------------------------------------- gates
@dataclass(frozen=True)
class SingleQubitGate:
    name: str
    matrix: tuple  # (m00, m01, m10, m11) complex
    param: int = -1
    gen: int = capi.GEN_NONE
    coeff: float = 0.0

    def element(self, row: int, col: int) -> complex:
        if row not in (0, 1):
            raise IllegalArgumentException(f"row must be 0 or 1: {row}")
        if col not in (0, 1):
            raise IllegalArgumentException(f"col must be 0 or 1: {col}")
        return self.matrix[2 * row + col]


_S = 1.0 / math.sqrt(2.0)


def Hadamard() -> SingleQubitGate:
    return SingleQubitGate("H", (complex(_S, 0.0), complex(_S, 0.0), complex(_S, 0.0), complex(-_S, 0.0)))


def PauliX() -> SingleQubitGate:
    return SingleQubitGate("X", (0j, 1 + 0j, 1 + 0j, 0j))


def PauliY() -> SingleQubitGate:
    return SingleQubitGate("Y", (0j, complex(0.0, -1.0), complex(0.0, 1.0), 0j))


def PauliZ() -> SingleQubitGate:
    return SingleQubitGate("Z", (1 + 0j, 0j, 0j, complex(-1.0, 0.0)))


def SGate() -> SingleQubitGate:
    return SingleQubitGate("S", (1 + 0j, 0j, 0j, complex(0.0, 1.0)))


def SdgGate() -> SingleQubitGate:
    return SingleQubitGate("Sdg", (1 + 0j, 0j, 0j, complex(0.0, -1.0)))


def MatrixSingleQubitGate(name: str, m00, m01, m10, m11) -> SingleQubitGate:
    return SingleQubitGate(name, (complex(m00), complex(m01), complex(m10), complex(m11)))


def _cis(a: float) -> complex:
    return complex(math.cos(a), math.sin(a))


def rotation_x(angle: float, param: int = -1) -> SingleQubitGate:
    h = angle / 2.0
    c, s = math.cos(h), math.sin(h)
    return SingleQubitGate("rx", (complex(c, 0.0), complex(0.0, -s), complex(0.0, -s), complex(c, 0.0)), param,
                           capi.GEN_X, 0.5)


def rotation_y(angle: float, param: int = -1) -> SingleQubitGate:
    h = angle / 2.0
    c, s = math.cos(h), math.sin(h)
    return SingleQubitGate("ry", (complex(c, 0.0), complex(-s, 0.0), complex(s, 0.0), complex(c, 0.0)), param,
                           capi.GEN_Y, 0.5)


def rotation_z(angle: float, param: int = -1) -> SingleQubitGate:
    h = angle / 2.0
    return SingleQubitGate("rz", (_cis(-h), 0j, 0j, _cis(h)), param, capi.GEN_Z, 0.5)


def phase_gate(angle: float) -> SingleQubitGate:
    return SingleQubitGate("p", (1 + 0j, 0j, 0j, _cis(angle)))


def u_gate(theta: float, phi: float, lam: float) -> SingleQubitGate:
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return SingleQubitGate("u", (complex(c, 0.0), _cis(lam) * -s, _cis(phi) * s, _cis(phi + lam) * c))


def rx_suite(angle: float, param: int = -1, coeff: float = 0.5) -> SingleQubitGate:
    """AlgorithmSuiteSpecs.rx: exp(-i angle X / 2); coeff = d(angle/2)/d(param)."""
    h = angle / 2.0
    c, s = math.cos(h), math.sin(h)
    return SingleQubitGate(f"Rx({angle:.3f})", (complex(c, 0.0), complex(0.0, -s), complex(0.0, -s), complex(c, 0.0)),
                           param, capi.GEN_X, coeff)


def zz_phase(gamma: float) -> list:
    """AlgorithmSuiteSpecs.zzPhase: diag(e^{-i g}, e^{i g}, e^{i g}, e^{-i g}) = exp(-i g Z Z)."""
    a, b = _cis(-gamma), _cis(gamma)
    return [a, b, b, a]


# ---------------------------------------------------------------- operations
@dataclass(frozen=True)
class Operation:
    kind: int
    q0: int = -1
    q1: int = -1
    controls: tuple = ()
    matrix: tuple = ()
    gate: SingleQubitGate | None = None
    param: int = -1
    gen: int = capi.GEN_NONE
    coeff: float = 0.0

    def qubits(self) -> tuple:
        if self.kind == capi.OP_MEASURE:
            return tuple(self.controls)
        qs = [self.q0] + ([self.q1] if self.q1 >= 0 else []) + list(self.controls)
        return tuple(qs)


class QuantumCircuit:
    """Immutable circuit; every append returns a new circuit (QuantumCircuit.java:29-106)."""

    def __init__(self, qubit_count: int, operations: Sequence[Operation] = ()):
        if qubit_count <= 0:
            raise IllegalArgumentException("qubitCount must be positive")
        self.qubit_count = int(qubit_count)
        self._ops = tuple(operations)
        for op in self._ops:
            self._validate(op)

    @property
    def operations(self) -> tuple:
        return self._ops

    def qubitCount(self) -> int:  # noqa: N802 (reference name)
        return self.qubit_count

    # --- validation (Operation.validateTargets, QuantumCircuit.java:362-574)
    def _validate(self, op: Operation) -> None:
        n = self.qubit_count
        k = op.kind
        if k == capi.OP_SINGLE:
            if op.q0 < 0:
                raise IllegalArgumentException("qubit index must be non-negative")
            if op.q0 >= n:
                raise IllegalArgumentException(f"target qubit out of range: {op.q0}")
        elif k == capi.OP_CNOT:
            if op.q0 < 0:
                raise IllegalArgumentException("control qubit index must be non-negative")
            if op.q1 < 0:
                raise IllegalArgumentException("target qubit index must be non-negative")
            if op.q0 == op.q1:
                raise IllegalArgumentException("control and target qubits must differ")
            if op.q0 >= n:
                raise IllegalArgumentException(f"control qubit out of range: {op.q0}")
            if op.q1 >= n:
                raise IllegalArgumentException(f"target qubit out of range: {op.q1}")
        elif k in (capi.OP_TWO, capi.OP_DIAG2, capi.OP_SWAP):
            if op.q0 < 0 or op.q1 < 0:
                raise IllegalArgumentException("qubit indices must be non-negative")
            if op.q0 == op.q1:
                raise IllegalArgumentException("cannot swap qubit with itself" if k == capi.OP_SWAP
                                               else "qubits must be distinct")
            if op.q0 >= n:
                raise IllegalArgumentException(f"first qubit out of range: {op.q0}")
            if op.q1 >= n:
                raise IllegalArgumentException(f"second qubit out of range: {op.q1}")
        elif k == capi.OP_MCU:
            if op.q0 < 0:
                raise IllegalArgumentException("target qubit index must be non-negative")
            if not op.controls:
                raise IllegalArgumentException("at least one control qubit is required")
            for c in op.controls:
                if c < 0:
                    raise IllegalArgumentException("control qubit index must be non-negative")
                if c == op.q0:
                    raise IllegalArgumentException("control and target qubits must differ")
            for a, b in zip(op.controls, op.controls[1:]):
                if a == b:
                    raise IllegalArgumentException(f"duplicate control qubit: {a}")
            if op.q0 >= n:
                raise IllegalArgumentException(f"target qubit out of range: {op.q0}")
            for c in op.controls:
                if c >= n:
                    raise IllegalArgumentException(f"control qubit out of range: {c}")
        elif k == capi.OP_MEASURE:
            if not op.controls:
                raise IllegalArgumentException("at least one qubit must be measured")
            for q in op.controls:
                if q < 0:
                    raise IllegalArgumentException("measured qubit index must be non-negative")
            for a, b in zip(op.controls, op.controls[1:]):
                if a == b:
                    raise IllegalArgumentException(f"duplicate measured qubit: {a}")
            for q in op.controls:
                if q >= n:
                    raise IllegalArgumentException(f"measured qubit out of range: {q}")
        else:
            raise IllegalArgumentException(f"unknown operation kind {k}")

    def _with(self, op: Operation) -> "QuantumCircuit":
        self._validate(op)
        c = QuantumCircuit.__new__(QuantumCircuit)
        c.qubit_count = self.qubit_count
        c._ops = self._ops + (op,)
        return c

    # --- builders (QuantumCircuit.java:29-106)
    def append(self, gate: SingleQubitGate, target: int) -> "QuantumCircuit":
        if gate is None:
            raise NullPointerException("gate")
        return self._with(Operation(capi.OP_SINGLE, int(target), gate=gate, param=gate.param, gen=gate.gen,
                                    coeff=gate.coeff))

    def append_cnot(self, control: int, target: int) -> "QuantumCircuit":
        return self._with(Operation(capi.OP_CNOT, int(control), int(target)))

    def append_two_qubit_unitary(self, matrix, first: int, second: int) -> "QuantumCircuit":
        m = tuple(complex(x) for x in np.asarray(matrix, dtype=np.complex128).reshape(-1))
        if len(m) != 16:
            raise IllegalArgumentException("two-qubit matrix must contain 16 elements")
        return self._with(Operation(capi.OP_TWO, int(first), int(second), matrix=m))

    def append_diagonal_two_qubit(self, diagonal, first: int, second: int, param: int = -1,
                                  coeff: float = 0.0) -> "QuantumCircuit":
        d = tuple(complex(x) for x in np.asarray(diagonal, dtype=np.complex128).reshape(-1))
        if len(d) != 4:
            raise IllegalArgumentException("two-qubit diagonal must contain 4 elements")
        return self._with(Operation(capi.OP_DIAG2, int(first), int(second), matrix=d, param=param,
                                    gen=capi.GEN_ZZ if param >= 0 else capi.GEN_NONE, coeff=coeff))

    def append_controlled_phase(self, angle: float, control: int, target: int) -> "QuantumCircuit":
        phase = complex(math.cos(angle), math.sin(angle))
        return self.append_diagonal_two_qubit([1 + 0j, 1 + 0j, 1 + 0j, phase], control, target)

    def append_controlled(self, gate: SingleQubitGate, control: int, target: int) -> "QuantumCircuit":
        return self.append_multi_control(gate, target, control)

    def append_controlled_x(self, control: int, target: int) -> "QuantumCircuit":
        return self.append_controlled(PauliX(), control, target)

    def append_controlled_y(self, control: int, target: int) -> "QuantumCircuit":
        return self.append_controlled(PauliY(), control, target)

    def append_controlled_z(self, control: int, target: int) -> "QuantumCircuit":
        return self.append_controlled(PauliZ(), control, target)

    def append_toffoli(self, c1: int, c2: int, target: int) -> "QuantumCircuit":
        return self.append_multi_control(PauliX(), target, c1, c2)

    def append_swap(self, first: int, second: int) -> "QuantumCircuit":
        return self._with(Operation(capi.OP_SWAP, int(first), int(second)))

    def append_multi_control(self, gate: SingleQubitGate, target: int, *controls: int) -> "QuantumCircuit":
        if gate is None:
            raise NullPointerException("gate")
        if len(controls) == 0:
            raise IllegalArgumentException("at least one control qubit is required")
        cs = tuple(sorted(int(c) for c in controls))
        return self._with(Operation(capi.OP_MCU, int(target), controls=cs, gate=gate, param=gate.param, gen=gate.gen,
                                    coeff=gate.coeff))

    def measure(self, *qubits: int) -> "QuantumCircuit":
        if len(qubits) == 0:
            raise IllegalArgumentException("at least one qubit must be measured")
        return self._with(Operation(capi.OP_MEASURE, controls=tuple(sorted(int(q) for q in qubits))))


def lower(circuit: QuantumCircuit):
    """Lower a circuit to (ops structured array, complex matrix pool) for sc_apply_ops."""
    ops = circuit.operations
    arr = np.zeros(len(ops), dtype=capi.OP_DTYPE)
    mats: list = []
    for i, op in enumerate(ops):
        r = arr[i]
        r["kind"], r["q0"], r["q1"], r["mat"] = op.kind, op.q0, op.q1, -1
        r["param"], r["gen"], r["coeff"] = op.param, op.gen, op.coeff
        if op.kind in (capi.OP_SINGLE, capi.OP_MCU):
            r["mat"] = len(mats)
            mats.extend(op.gate.matrix)
        elif op.kind in (capi.OP_TWO, capi.OP_DIAG2):
            r["mat"] = len(mats)
            mats.extend(op.matrix)
        if op.kind == capi.OP_MCU or op.kind == capi.OP_MEASURE:
            m = 0
            for c in op.controls:
                m |= 1 << c
            r["controls"] = m
        elif op.kind == capi.OP_CNOT:
            r["controls"] = 1 << op.q0
    return arr, np.asarray(mats, dtype=np.complex128)
