"""HipStatevectorEngine: the SimulatorEngine mirror over the C-ABI.

Request / result types and the simulate() control flow follow
core/src/main/java/com/omaarr90/statecraft/core/engine/{SimulatorEngine,SimulationRequest,
MeasurementInstruction,MeasurementResult,SimulationResult}.java and
StatevectorEngine.simulate (engines/.../statevector/StatevectorEngine.java:63-167).
All amplitude work runs in libstatecraft_hip.so; this module only validates,
lowers, draws SplittableRandom variates in reference order and assembles
results.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import capi
from .capi import (IllegalArgumentException, IllegalStateException, NullPointerException,
                   UnsupportedOperationException)
from .circuit import QuantumCircuit, lower
from .noise import CompositeChannel, NoiseModel
from .rng import SplittableRandom


# ---------------------------------------------------------------- request / result types
@dataclass(frozen=True)
class MeasurementInstruction:
    shots: int
    measured_qubits: Optional[tuple] = None
    seed: Optional[int] = None
    mode: str = "COUNTS"

    def __post_init__(self):
        if self.shots <= 0:
            raise IllegalArgumentException("shots must be positive")
        if self.measured_qubits is not None:
            qs = list(self.measured_qubits)
            if not qs:
                raise IllegalArgumentException("measured qubit list must not be empty")
            for q in qs:
                if q < 0:
                    raise IllegalArgumentException(f"qubit index must be non-negative: {q}")
            qs.sort()
            for a, b in zip(qs, qs[1:]):
                if a == b:
                    raise IllegalArgumentException(f"duplicate qubit index: {a}")
            object.__setattr__(self, "measured_qubits", tuple(qs))
        if self.mode not in ("COUNTS", "SAMPLES"):
            raise IllegalArgumentException("mode must be COUNTS or SAMPLES")

    @staticmethod
    def counts_all(shots: int) -> "MeasurementInstruction":
        return MeasurementInstruction(shots)

    @staticmethod
    def counts(shots: int, *qubits: int) -> "MeasurementInstruction":
        if not qubits:
            raise IllegalArgumentException("at least one qubit must be provided")
        return MeasurementInstruction(shots, tuple(qubits))

    @staticmethod
    def samples_all(shots: int) -> "MeasurementInstruction":
        return MeasurementInstruction(shots, None, None, "SAMPLES")

    @staticmethod
    def samples(shots: int, *qubits: int) -> "MeasurementInstruction":
        if not qubits:
            raise IllegalArgumentException("at least one qubit must be provided")
        return MeasurementInstruction(shots, tuple(qubits), None, "SAMPLES")

    def with_seed(self, seed: int) -> "MeasurementInstruction":
        return MeasurementInstruction(self.shots, self.measured_qubits, int(seed), self.mode)


@dataclass(frozen=True)
class SimulationRequest:
    circuit: QuantumCircuit
    initial_state: Optional[np.ndarray] = None
    basis_state_qubits: Optional[tuple] = None
    measurement: Optional[MeasurementInstruction] = None
    noise_model: Optional[NoiseModel] = None
    noise_seed: Optional[int] = None
    return_final_state: bool = True

    def __post_init__(self):
        c = self.circuit
        if c is None:
            raise NullPointerException("circuit")
        n = c.qubit_count
        if self.initial_state is not None and len(self.initial_state) != (1 << n):
            raise IllegalArgumentException(
                f"Initial state qubit count {int(np.log2(len(self.initial_state)))} does not match circuit qubit "
                f"count {n}")
        if self.basis_state_qubits is not None:
            qs = sorted(int(q) for q in self.basis_state_qubits)
            for q in qs:
                if q < 0 or q >= n:
                    raise IllegalArgumentException(f"basis-state qubit out of range: {q}")
            for a, b in zip(qs, qs[1:]):
                if a == b:
                    raise IllegalArgumentException(f"duplicate basis-state qubit: {a}")
            object.__setattr__(self, "basis_state_qubits", tuple(qs))
        if self.initial_state is not None and self.basis_state_qubits is not None:
            raise IllegalArgumentException("initialState and basisStateQubits are mutually exclusive")
        if self.measurement is not None and self.measurement.measured_qubits is not None:
            for q in self.measurement.measured_qubits:
                if q < 0 or q >= n:
                    raise IllegalArgumentException(f"measurement qubit out of range: {q}")
        if not self.return_final_state and self.measurement is None:
            raise IllegalArgumentException(
                "simulation must request the final state or provide a measurement instruction")

    @staticmethod
    def zero_state(circuit: QuantumCircuit) -> "SimulationRequest":
        return SimulationRequest(circuit)

    @staticmethod
    def with_initial_state(circuit: QuantumCircuit, state) -> "SimulationRequest":
        if state is None:
            raise NullPointerException("state")
        return SimulationRequest(circuit, np.asarray(state, dtype=np.complex128))

    def _replace(self, **kw) -> "SimulationRequest":
        d = dict(circuit=self.circuit, initial_state=self.initial_state, basis_state_qubits=self.basis_state_qubits,
                 measurement=self.measurement, noise_model=self.noise_model, noise_seed=self.noise_seed,
                 return_final_state=self.return_final_state)
        d.update(kw)
        return SimulationRequest(**d)

    def with_measurement(self, instruction: MeasurementInstruction, include_final_state: Optional[bool] = None):
        if instruction is None:
            raise NullPointerException("instruction")
        kw = {"measurement": instruction}
        if include_final_state is not None:
            kw["return_final_state"] = include_final_state
        return self._replace(**kw)

    def without_final_state(self) -> "SimulationRequest":
        if self.measurement is None:
            raise IllegalStateException("cannot drop final state without measurement instruction")
        return self._replace(return_final_state=False)

    def with_noise_model(self, model: NoiseModel) -> "SimulationRequest":
        if model is None:
            raise NullPointerException("model")
        return self._replace(noise_model=model)

    def with_noise_seed(self, seed: int) -> "SimulationRequest":
        return self._replace(noise_seed=int(seed))

    def with_basis_state(self, *qubits: int) -> "SimulationRequest":
        return self._replace(initial_state=None, basis_state_qubits=tuple(qubits))


@dataclass(frozen=True)
class Histogram:
    measured_qubits: tuple
    shots: int
    counts: dict

    def __post_init__(self):
        if sum(self.counts.values()) != self.shots:
            raise IllegalArgumentException(
                f"histogram counts sum {sum(self.counts.values())} does not match shots {self.shots}")


@dataclass(frozen=True)
class Samples:
    measured_qubits: tuple
    shots: int
    outcomes: list


@dataclass(frozen=True)
class BitstringHistogram:
    """Outcomes wider than 31 bits as MSB-left bit strings (TensorNetworkEngine.java:425-431 convention)."""
    measured_qubits: tuple
    shots: int
    counts: dict


@dataclass(frozen=True)
class BitstringSamples:
    measured_qubits: tuple
    shots: int
    outcomes: list


@dataclass(frozen=True)
class SimulationResult:
    final_state: Optional[np.ndarray]
    measurement: Optional[object]


# ---------------------------------------------------------------- engine
class HipStatevectorEngine:
    """SimulatorEngine implementation backed by the gfx950 engine (id "statevector-hip")."""

    ID = "statevector-hip"

    def __init__(self, device: int = 0, exec_mode: int = capi.EXEC_SWEEP):
        self.ctx = capi.Context(device, exec_mode)

    def id(self) -> str:
        return self.ID

    def close(self) -> None:
        """Releases the device context (its stream, scratch and cached plans)."""
        self.ctx.close()

    def simulate(self, request: SimulationRequest) -> SimulationResult:
        if request is None:
            raise NullPointerException("request")
        circuit = request.circuit
        n = circuit.qubit_count
        st = capi.State(self.ctx, n, 1)
        try:
            if request.initial_state is not None:
                st.upload(request.initial_state)
            elif request.basis_state_qubits is not None:
                idx = 0
                for q in request.basis_state_qubits:
                    idx |= 1 << q
                st.set_basis(idx)
            ops, mats = lower(circuit)
            noise = request.noise_model
            measure_ops = [op for op in circuit.operations if op.kind == capi.OP_MEASURE]
            if noise is not None and noise.has_noise():
                self._run_noisy(st, circuit, ops, mats, noise, request.noise_seed)
            else:
                st.apply(ops, mats)
            final = st.download() if request.return_final_state else None
            meas = None
            if request.measurement is not None:
                meas = self._measure(st, n, request.measurement, measure_ops)
            return SimulationResult(final, meas)
        finally:
            st.close()

    # StatevectorEngine.java:79-151 (noisy branch) and :371-407
    def _run_noisy(self, st, circuit, ops, mats, noise: NoiseModel, seed):
        """The noisy circuit in one sc_run_noisy call.  Every scheduled channel and target becomes an event; its
        uniform (the nextDouble() behind rng.nextDouble(total), one per event in circuit order) is drawn here and
        the branch is chosen on the device from the live weights.  A validation error ends the event list where
        the reference would raise it; the events before it still run, so a failing channel among them is
        reported first, as in the reference's order."""
        rng = SplittableRandom(seed if seed is not None else int.from_bytes(os.urandom(8), "little"))
        n = circuit.qubit_count
        events, kraus, where = [], [], []  # where[e]: (channel, target) for messages
        seen_measure = False
        pending_error, end = None, len(ops)
        for i, op in enumerate(circuit.operations):
            if op.kind == capi.OP_MEASURE:
                seen_measure = True
                continue
            if seen_measure:
                pending_error = UnsupportedOperationException(
                    "Unitary operations cannot follow measurement operations in the circuit")
                end = i
                break
            try:
                for sch in noise.scheduled_channels_after(op):
                    self._collect_channel(sch.channel, sch.target_qubits, n, i, events, kraus, where)
            except (IllegalArgumentException, IllegalStateException) as e:
                pending_error, end = e, i + 1
                break
        u = rng.next_doubles(len(events)) if events else np.zeros(0)
        try:
            st.run_noisy(ops[:end], mats, events, kraus, u)
        except capi.NoisyRunFailed as f:
            channel, t = where[f.event]
            w = f.weights[:len(channel.kraus)]
            if f.kind == 1:
                raise IllegalStateException(f"invalid Kraus branch weight for channel {channel} on targets [{t}]: "
                                            f"weights={w}, branch={w[f.branch]}") from None
            if f.kind == 2:
                raise IllegalStateException(f"invalid Kraus branch weights for channel {channel} on targets [{t}]: "
                                            f"total={sum(w)}, weights={w}") from None
            raise IllegalStateException(
                f"state collapsed to invalid norm after Kraus application: normSq={f.norm_sq}") from None
        if pending_error is not None:
            raise pending_error

    def _collect_channel(self, channel, mapped, n, after, events, kraus, where):
        # applyErrorChannel (StatevectorEngine.java:382-407) and resolveTargets (:409-424), composite recursion
        if isinstance(channel, CompositeChannel):
            comp_map = mapped if mapped is not None and len(mapped) == 1 else None
            for c in channel.channels:
                self._collect_channel(c, comp_map, n, after, events, kraus, where)
            return
        targets = tuple(mapped) if mapped is not None else channel.affected_qubits()
        if len(targets) == 0:
            raise IllegalArgumentException("noise channel must target at least one qubit")
        seen = set()
        for t in targets:
            if t < 0 or t >= n:
                raise IllegalArgumentException(f"noise channel qubit out of range: {t}")
            if t in seen:
                raise IllegalArgumentException(f"noise channel contains duplicate qubit: {t}")
            seen.add(t)
        first = len(kraus)
        kraus.extend(np.asarray(k, np.complex128).reshape(2, 2) for k in channel.kraus)
        for t in targets:  # every built-in channel decomposes to single-qubit operators
            events.append((after, t, first, len(channel.kraus)))
            where.append((channel, t))

    # StatevectorEngine.java:229-272
    def _measure(self, st, n, instr: MeasurementInstruction, measure_ops):
        if instr.measured_qubits is not None:
            measured = tuple(instr.measured_qubits)
        elif measure_ops:
            measured = tuple(sorted({q for op in measure_ops for q in op.controls}))
        else:
            measured = tuple(range(n))
        if len(measured) == 0:
            raise IllegalStateException("No qubits available for measurement")
        seed = instr.seed if instr.seed is not None else int.from_bytes(os.urandom(8), "little")
        u = SplittableRandom(seed).next_doubles(instr.shots)
        outcomes, _gap = st.sample(measured, u)
        m = len(measured)
        if instr.mode == "COUNTS":
            vals, cnts = np.unique(outcomes, return_counts=True)
            if m <= 31:
                return Histogram(measured, instr.shots, {int(v): int(c) for v, c in zip(vals, cnts)})
            return BitstringHistogram(measured, instr.shots,
                                      {format(int(v), f"0{m}b"): int(c) for v, c in zip(vals, cnts)})
        if m <= 31:
            return Samples(measured, instr.shots, [int(v) for v in outcomes])
        return BitstringSamples(measured, instr.shots, [format(int(v), f"0{m}b") for v in outcomes])


def engines() -> dict:
    """SimulatorEngines.loadAll analogue for this package: id -> factory."""
    return {HipStatevectorEngine.ID: HipStatevectorEngine}
