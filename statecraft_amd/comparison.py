"""Comparison-runner front end: the command line and JSON output of the reference's
StatecraftComparisonRunner (engines/src/test/java/com/omaarr90/statecraft/engines/comparison/
StatecraftComparisonRunner.java), served by the HIP engine, so the external benchmark harness
(scripts/benchmark/comparison/run_comparison.py) can run this engine as one more runner.

    python -m statecraft_amd.comparison --circuit-json circuit.json [--fixture-id ID] [--category C]
        [--warmup-runs 1] [--timed-runs 5] [--parallelism 1] [--noise-type T --noise-qubits 0,1
        --noise-trajectories 512 --noise-seed S --noise-probability p | --noise-gamma g | --noise-lambda l |
        --noise-t1 t1 --noise-t2 t2 --noise-gate-time t]

Circuits use the reference's JSON circuit format (core/src/main/java/com/omaarr90/statecraft/core/parse/
JsonCircuitParser.java:52-229).  One JSON object is printed: the reference's fields in its order
(StatecraftComparisonRunner.java:119-147), numbers formatted with %.17g.
"""
from __future__ import annotations

import json
import math
import sys
import time
from typing import List, Optional

import numpy as np

from .capi import IllegalArgumentException
from .circuit import (Hadamard, PauliX, PauliY, PauliZ, QuantumCircuit, SdgGate, SGate)
from .engine import HipStatevectorEngine, SimulationRequest
from .noise import ErrorChannel, NoiseModel

DEFAULT_NOISE_SEED = 0x51A7EC0FFEEA


class CircuitParseException(ValueError):
    pass


def _int_field(node: dict, field: str, context: str, required: bool = False) -> Optional[int]:
    v = node.get(field)
    if v is None:
        if required:
            raise CircuitParseException(f"Missing required field '{field}' in {context}")
        return None
    if isinstance(v, bool) or not isinstance(v, int):
        raise CircuitParseException(f"Field '{field}' in {context} must be an integer")
    return v


def _require(v, message: str):
    if v is None:
        raise CircuitParseException(message)
    return v


def _index(q: int, n: int, name: str) -> None:
    if q < 0 or q >= n:
        raise CircuitParseException(f"{name} index {q} out of range for {n} qubits")


def parse_circuit_json(text: str) -> QuantumCircuit:
    """JsonCircuitParser.parse (JsonCircuitParser.java:52-229): gates h x y z s sdg cx cy cz swap cp ccx mcx mcy
    mcz measure barrier; a unitary after a measurement is rejected."""
    try:
        root = json.loads(text)
    except json.JSONDecodeError as e:
        raise CircuitParseException(f"Invalid JSON: {e.msg}") from e
    if root is None:
        raise CircuitParseException("JSON source is empty")
    if not isinstance(root, dict):
        raise CircuitParseException("Root JSON value must be an object")
    n = _int_field(root, "qubits", "root", required=True)
    ops = root.get("operations")
    if ops is not None and not isinstance(ops, list):
        raise CircuitParseException("'operations' must be an array")
    if n <= 0:
        raise CircuitParseException("'qubits' must be positive")
    circ = QuantumCircuit(n)
    measured_seen = False
    singles = {"h": Hadamard, "x": PauliX, "y": PauliY, "z": PauliZ, "s": SGate, "sdg": SdgGate}
    for i, op in enumerate(ops or []):
        ctx = f"operations[{i}]"
        if not isinstance(op, dict):
            raise CircuitParseException(f"{ctx} must be an object")
        gate = op.get("gate")
        if not isinstance(gate, str):
            raise CircuitParseException(f"Missing required field 'gate' in {ctx}")
        gate = gate.strip().lower()
        target = _int_field(op, "target", ctx)
        control = _int_field(op, "control", ctx)
        first = _int_field(op, "first", ctx)
        second = _int_field(op, "second", ctx)
        if measured_seen and gate not in ("measure", "barrier"):
            raise CircuitParseException(f"Unitary gate '{gate}' cannot follow a measurement operation")
        if gate in singles:
            t = _require(target, f"gate '{gate}' requires 'target'")
            _index(t, n, "target")
            circ = circ.append(singles[gate](), t)
        elif gate in ("cx", "cy", "cz"):
            c = _require(control, f"gate '{gate}' requires 'control'")
            t = _require(target, f"gate '{gate}' requires 'target'")
            if c == t:
                raise CircuitParseException(f"gate '{gate}' requires distinct control and target qubits")
            _index(c, n, "control")
            _index(t, n, "target")
            circ = {"cx": circ.append_cnot, "cy": circ.append_controlled_y, "cz": circ.append_controlled_z}[gate](c, t)
        elif gate == "swap":
            a = _require(first, "gate 'swap' requires 'first'")
            b = _require(second, "gate 'swap' requires 'second'")
            _index(a, n, "first")
            _index(b, n, "second")
            if a == b:
                raise CircuitParseException("gate 'swap' requires distinct qubits")
            circ = circ.append_swap(a, b)
        elif gate == "cp":
            c = _require(control, "gate 'cp' requires 'control'")
            t = _require(target, "gate 'cp' requires 'target'")
            angle = op.get("angle")
            if angle is None or isinstance(angle, bool) or not isinstance(angle, (int, float)):
                raise CircuitParseException("gate 'cp' requires 'angle'")
            if not math.isfinite(float(angle)):
                raise CircuitParseException("gate 'cp' requires a finite angle")
            _index(c, n, "control")
            _index(t, n, "target")
            if c == t:
                raise CircuitParseException("gate 'cp' requires distinct control and target qubits")
            circ = circ.append_controlled_phase(float(angle), c, t)
        elif gate in ("ccx", "mcx", "mcy", "mcz"):
            t = _require(target, f"gate '{gate}' requires 'target'")
            controls = op.get("controls")
            if not isinstance(controls, list) or not controls:
                raise CircuitParseException(f"gate '{gate}' requires 'controls'")
            cs = sorted(set(int(c) for c in controls))
            for c in cs:
                _index(c, n, "control")
            if gate == "ccx" and len(cs) != 2:
                raise CircuitParseException("gate 'ccx' requires exactly two control qubits")
            if t in cs:
                raise CircuitParseException(f"gate '{gate}' requires distinct controls and target")
            _index(t, n, "target")
            g = {"ccx": PauliX, "mcx": PauliX, "mcy": PauliY, "mcz": PauliZ}[gate]()
            circ = circ.append_multi_control(g, t, *cs)
        elif gate == "measure":
            measured_seen = True
            qs = op.get("qubits")
            qs = list(range(n)) if qs is None else sorted(set(int(q) for q in qs))
            for q in qs:
                _index(q, n, "measured")
            circ = circ.measure(*qs)
        elif gate == "barrier":
            pass
        else:
            raise CircuitParseException(f"Unknown gate: '{op.get('gate')}'")
    return circ


def fmt(v: float) -> str:
    """Java String.format(Locale.US, "%.17g", v).  Java's Formatter starts from the shortest decimal that
    round-trips (Double.toString's digits, as Python's repr) and pads it with zeros to 17 significant digits;
    decimal notation for exponents -4..16, else d.dddde[+-]XX."""
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    if v == 0.0:
        return sign + "0." + "0" * 16
    from decimal import Decimal

    t = Decimal(repr(abs(v))).normalize().as_tuple()
    digits = "".join(str(d) for d in t.digits)
    x = len(digits) + t.exponent - 1  # decimal exponent of the leading digit
    digits = (digits + "0" * 17)[:17]
    if -4 <= x < 17:
        if x >= 0:
            frac = digits[x + 1:]
            return sign + digits[:x + 1] + ("." + frac if frac else "")
        return sign + "0." + "0" * (-x - 1) + digits
    return sign + digits[0] + "." + digits[1:] + ("e-%02d" % -x if x < 0 else "e+%02d" % x)


class Options:
    def __init__(self):
        self.fixture_id = ""
        self.category = ""
        self.circuit_json = None
        self.warmup_runs = 1
        self.timed_runs = 5
        self.parallelism = 1
        self.noise_type = None
        self.noise_qubits: List[int] = []
        self.noise_trajectories = 512
        self.noise_seed = DEFAULT_NOISE_SEED
        self.noise_probability = 0.0
        self.noise_gamma = 0.0
        self.noise_lambda = 0.0
        self.noise_t1 = 0.0
        self.noise_t2 = 0.0
        self.noise_gate_time = 0.0

    @staticmethod
    def parse(argv: List[str]) -> "Options":
        """Options.parse (StatecraftComparisonRunner.java:228-255)."""
        o = Options()
        setters = {
            "--fixture-id": lambda v: setattr(o, "fixture_id", v),
            "--category": lambda v: setattr(o, "category", v),
            "--circuit-json": lambda v: setattr(o, "circuit_json", v),
            "--warmup-runs": lambda v: setattr(o, "warmup_runs", int(v)),
            "--timed-runs": lambda v: setattr(o, "timed_runs", int(v)),
            "--parallelism": lambda v: setattr(o, "parallelism", int(v)),
            "--noise-type": lambda v: setattr(o, "noise_type", v),
            "--noise-qubits": lambda v: setattr(o, "noise_qubits",
                                                [int(x.strip()) for x in v.split(",")] if v.strip() else []),
            "--noise-trajectories": lambda v: setattr(o, "noise_trajectories", int(v)),
            "--noise-seed": lambda v: setattr(o, "noise_seed", int(v)),
            "--noise-probability": lambda v: setattr(o, "noise_probability", float(v)),
            "--noise-gamma": lambda v: setattr(o, "noise_gamma", float(v)),
            "--noise-lambda": lambda v: setattr(o, "noise_lambda", float(v)),
            "--noise-t1": lambda v: setattr(o, "noise_t1", float(v)),
            "--noise-t2": lambda v: setattr(o, "noise_t2", float(v)),
            "--noise-gate-time": lambda v: setattr(o, "noise_gate_time", float(v)),
        }
        i = 0
        while i < len(argv):
            arg = argv[i]
            if arg not in setters:
                raise IllegalArgumentException(f"unknown argument: {arg}")
            if i + 1 >= len(argv):
                raise IllegalArgumentException(f"missing value for {arg}")
            setters[arg](argv[i + 1])
            i += 2
        if o.circuit_json is None:
            raise IllegalArgumentException("--circuit-json is required")
        if o.noise_type is not None and not o.noise_qubits:
            raise IllegalArgumentException("--noise-qubits is required when --noise-type is provided")
        return o

    def channel(self, qubit: int) -> ErrorChannel:
        """createChannel (StatecraftComparisonRunner.java:96-106)."""
        t = self.noise_type
        if t == "depolarizing":
            return ErrorChannel.depolarizing(self.noise_probability, qubit)
        if t == "phase_flip":
            return ErrorChannel.phase_flip(self.noise_probability, qubit)
        if t == "amplitude_damping":
            return ErrorChannel.amplitude_damping(self.noise_gamma, qubit)
        if t == "phase_damping":
            return ErrorChannel.phase_damping(self.noise_lambda, qubit)
        if t == "thermal_relaxation":
            return ErrorChannel.thermal_relaxation(self.noise_t1, self.noise_t2, self.noise_gate_time, qubit)
        raise IllegalArgumentException(f"unknown noise type: {t}")

    def noise_model(self) -> Optional[NoiseModel]:
        if self.noise_type is None:
            return None
        model = NoiseModel()
        for q in self.noise_qubits:
            model.after_all_gates(self.channel(q))
        return model


def _run_once(engine: HipStatevectorEngine, circuit: QuantumCircuit, noise: Optional[NoiseModel], seed: int):
    req = SimulationRequest.zero_state(circuit)
    if noise is not None:
        req = req.with_noise_model(noise).with_noise_seed(seed)
    return engine.simulate(req).final_state


def _probabilities(state) -> np.ndarray:
    a = np.asarray(state, dtype=np.complex128)
    return a.real * a.real + a.imag * a.imag


def run(argv: List[str], engine: Optional[HipStatevectorEngine] = None) -> str:
    """StatecraftComparisonRunner.main (:27-64): warm-up runs, timed runs (System.nanoTime around each
    simulate), probabilities of the last run, or trajectory-averaged probabilities under noise."""
    o = Options.parse(argv)
    with open(o.circuit_json, "r", encoding="utf-8") as f:
        circuit = parse_circuit_json(f.read())
    own = engine is None
    engine = engine or HipStatevectorEngine()
    try:
        noise = o.noise_model()
        for i in range(o.warmup_runs):
            _run_once(engine, circuit, noise, o.noise_seed + i)
        timed: List[float] = []
        final = None
        if noise is None:
            for i in range(o.timed_runs):
                t0 = time.perf_counter_ns()
                final = _run_once(engine, circuit, None, o.noise_seed + i)
                timed.append((time.perf_counter_ns() - t0) / 1e6)
            probs = _probabilities(final)
        else:
            probs = None
            for _ in range(o.timed_runs):
                t0 = time.perf_counter_ns()
                agg = np.zeros(1 << circuit.qubit_count)
                for t in range(o.noise_trajectories):
                    agg += _probabilities(_run_once(engine, circuit, noise, o.noise_seed + t))
                probs = agg * (1.0 / o.noise_trajectories)
                timed.append((time.perf_counter_ns() - t0) / 1e6)
        return render(o, circuit, timed, final, probs)
    finally:
        if own:
            engine.close()


def render(o: Options, circuit: QuantumCircuit, timed: List[float], final, probs) -> str:
    """renderResult (StatecraftComparisonRunner.java:119-147): field order and %.17g numbers."""
    def s(v: str) -> str:
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'

    unitary_ops = sum(1 for op in circuit.operations if op.kind != 7)
    parts = [
        '"runner":' + s("statecraft"),
        '"fixture_id":' + s(o.fixture_id),
        '"category":' + s(o.category),
        '"status":' + s("ok"),
        f'"qubits":{circuit.qubit_count}',
        f'"operations":{unitary_ops}',
        f'"threads":{o.parallelism}',
        f'"warmup_runs":{o.warmup_runs}',
        f'"timed_runs":{o.timed_runs}',
        f'"noise_trajectories":{0 if o.noise_type is None else o.noise_trajectories}',
        '"timed_runs_ms":[' + ",".join(fmt(v) for v in timed) + "]",
        '"mean_ms":' + fmt(sum(timed) / len(timed) if timed else 0.0),
        '"probabilities":[' + ",".join(fmt(float(v)) for v in (probs if probs is not None else [])) + "]",
    ]
    if final is not None:
        amps = np.asarray(final, dtype=np.complex128)
        parts.append('"statevector":[' + ",".join("[" + fmt(a.real) + "," + fmt(a.imag) + "]" for a in amps) + "]")
    return "{" + ",".join(parts) + "}"


def main(argv=None) -> int:
    print(run(list(sys.argv[1:] if argv is None else argv)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
