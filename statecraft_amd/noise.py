"""Noise channels and the NoiseModel schedule (host side).

Kraus sets follow the reference channel classes
(core/src/main/java/com/omaarr90/statecraft/core/noise/DepolarizingChannel.java:49-76,
AmplitudeDampingChannel.java:47-62, PhaseFlipChannel.java:50-64,
PhaseDampingChannel.java:51-64, ThermalRelaxationChannel.java:61-132,
CompositeChannel.java:19-79).  NoiseModel.scheduled_channels_after follows
NoiseModel.scheduledChannelsAfter (NoiseModel.java:211-241): gate-class
channels, then idle-qubit channels in java.util.HashMap iteration order, then
global channels.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .capi import IllegalArgumentException, IllegalStateException, UnsupportedOperationException

_BOUNDED_TOL = 1e-12
_RATE_TOL = 1e-12


def _op(m00, m01, m10, m11) -> np.ndarray:
    return np.array([m00, m01, m10, m11], dtype=np.complex128)


class ErrorChannel:
    kind = "channel"

    def __init__(self, qubits):
        qubits = tuple(int(q) for q in qubits)
        if len(qubits) != 1:
            raise IllegalArgumentException(f"{self.kind} channel currently only supports single qubits")
        self.qubits = qubits
        self.kraus = self._build()

    def affected_qubits(self) -> tuple:
        return self.qubits

    def num_qubits(self) -> int:
        return 1

    # ---- factories (ErrorChannel.java:42-131)
    @staticmethod
    def depolarizing(p: float, *qubits: int) -> "ErrorChannel":
        return DepolarizingChannel(p, qubits)

    @staticmethod
    def amplitude_damping(gamma: float, *qubits: int) -> "ErrorChannel":
        return AmplitudeDampingChannel(gamma, qubits)

    @staticmethod
    def phase_flip(p: float, *qubits: int) -> "ErrorChannel":
        return PhaseFlipChannel(p, qubits)

    @staticmethod
    def phase_damping(lam: float, *qubits: int) -> "ErrorChannel":
        return PhaseDampingChannel(lam, qubits)

    @staticmethod
    def thermal_relaxation(t1: float, t2: float, gate_time: float, *qubits: int) -> "ErrorChannel":
        return ThermalRelaxationChannel(t1, t2, gate_time, qubits)

    @staticmethod
    def compose(*channels: "ErrorChannel") -> "CompositeChannel":
        return CompositeChannel(channels)


def _prob(name, v):
    if v < 0.0 or v > 1.0:
        raise IllegalArgumentException(f"{name} must be in [0,1], got {v}")
    return float(v)


class DepolarizingChannel(ErrorChannel):
    kind = "depolarizing"

    def __init__(self, p, qubits):
        self.p = _prob("probability", p)
        super().__init__(qubits)

    def _build(self):
        a, s = math.sqrt(1.0 - self.p), math.sqrt(self.p / 3.0)
        return np.stack([_op(a, 0, 0, a), _op(0, s, s, 0), _op(0, complex(0, -s), complex(0, s), 0),
                         _op(s, 0, 0, -s)])


class AmplitudeDampingChannel(ErrorChannel):
    kind = "amplitude damping"

    def __init__(self, gamma, qubits):
        self.gamma = _prob("gamma", gamma)
        super().__init__(qubits)

    def _build(self):
        return np.stack([_op(1, 0, 0, math.sqrt(1.0 - self.gamma)), _op(0, math.sqrt(self.gamma), 0, 0)])


class PhaseFlipChannel(ErrorChannel):
    kind = "phase flip"

    def __init__(self, p, qubits):
        self.p = _prob("probability", p)
        super().__init__(qubits)

    def _build(self):
        a, s = math.sqrt(1.0 - self.p), math.sqrt(self.p)
        return np.stack([_op(a, 0, 0, a), _op(s, 0, 0, -s)])


class PhaseDampingChannel(ErrorChannel):
    kind = "phase damping"

    def __init__(self, lam, qubits):
        self.lam = _prob("lambda", lam)
        super().__init__(qubits)

    def _build(self):
        return np.stack([_op(1, 0, 0, math.sqrt(1.0 - self.lam)), _op(0, 0, 0, math.sqrt(self.lam))])


def _bounded(name, v):
    if not math.isfinite(v):
        raise IllegalStateException(f"{name} must be finite, got {v}")
    if v < 0.0:
        if v >= -_BOUNDED_TOL:
            return 0.0
        raise IllegalStateException(f"{name} must be in [0,1], got {v}")
    if v > 1.0:
        if v <= 1.0 + _BOUNDED_TOL:
            return 1.0
        raise IllegalStateException(f"{name} must be in [0,1], got {v}")
    return v


class ThermalRelaxationChannel(ErrorChannel):
    kind = "thermal relaxation"

    def __init__(self, t1, t2, gate_time, qubits):
        if t1 <= 0.0 or not math.isfinite(t1):
            raise IllegalArgumentException(f"T1 must be positive and finite, got {t1}")
        if t2 <= 0.0 or not math.isfinite(t2):
            raise IllegalArgumentException(f"T2 must be positive and finite, got {t2}")
        if t2 > 2.0 * t1:
            raise IllegalArgumentException(f"T2 must satisfy T2 <= 2*T1 (physical constraint), got T1={t1}, T2={t2}")
        if gate_time < 0.0 or not math.isfinite(gate_time):
            raise IllegalArgumentException(f"gate time must be non-negative and finite, got {gate_time}")
        self.t1, self.t2, self.gate_time = float(t1), float(t2), float(gate_time)
        super().__init__(qubits)

    def _build(self):
        if self.gate_time == 0.0:
            surv, deph = 1.0, 1.0
        else:
            surv = _bounded("relaxation survival", math.exp(-self.gate_time / self.t1))
            inv_t2, half_inv_t1 = 1.0 / self.t2, 1.0 / (2.0 * self.t1)
            rate = inv_t2 - half_inv_t1
            if rate < 0.0:
                tol = _RATE_TOL * max(1.0, abs(inv_t2) + abs(half_inv_t1))
                if rate >= -tol:
                    rate = 0.0
                else:
                    raise IllegalStateException(f"pure dephasing rate must be non-negative, got {rate}")
            deph = _bounded("pure dephasing factor", math.exp(-self.gate_time * rate))
        decay = _bounded("decay probability", 1.0 - surv)
        ss = math.sqrt(surv)
        plus = _bounded("thermal plus factor", 0.5 * (1.0 + deph))
        minus = _bounded("thermal minus factor", 0.5 * (1.0 - deph))
        k0d0 = math.sqrt(plus)
        k1d0 = math.sqrt(minus)
        return np.stack([_op(k0d0, 0, 0, ss * k0d0), _op(k1d0, 0, 0, -(ss * k1d0)), _op(0, math.sqrt(decay), 0, 0)])


class CompositeChannel(ErrorChannel):
    kind = "composite"

    def __init__(self, channels):
        if len(channels) == 0:
            raise IllegalArgumentException("composite channel requires at least one channel")
        self.channels = tuple(channels)
        qs = set()
        for c in channels:
            qs.update(c.affected_qubits())
        self.qubits = tuple(sorted(qs))
        self.kraus = None

    def num_qubits(self):
        raise UnsupportedOperationException("CompositeChannel does not provide a single Kraus decomposition. "
                                            "Use getChannels() for sequential application.")


def _is_single_qubit(ch) -> bool:
    if isinstance(ch, CompositeChannel):
        return all(_is_single_qubit(c) for c in ch.channels)
    return True


def java_hashmap_order(keys_in_insertion_order) -> list:
    """Iteration order of a java.util.HashMap<Integer, ...> (default capacity 16, load factor 0.75) filled in
    the given insertion order: buckets by (h ^ h>>>16) & (cap-1), insertion order within a bucket."""
    keys = list(dict.fromkeys(keys_in_insertion_order))
    cap = 16
    while len(keys) > cap * 0.75:
        cap *= 2

    def bucket(k):
        h = k & 0xFFFFFFFF
        return (h ^ (h >> 16)) & (cap - 1)

    return sorted(keys, key=lambda k: (bucket(k), keys.index(k)))


@dataclass
class ScheduledChannel:
    channel: ErrorChannel
    target_qubits: tuple


class NoiseModel:
    """Gate-class, idle-qubit and global channels (NoiseModel.java:46-288)."""

    def __init__(self):
        self.gate_noise: dict = {}   # op kind -> [channels]
        self.qubit_noise: dict = {}  # qubit -> [channels], insertion ordered
        self.global_noise: list = []

    def after_gate(self, op_kind: int, channel: ErrorChannel) -> "NoiseModel":
        self.gate_noise.setdefault(op_kind, []).append(channel)
        return self

    def on_qubits(self, channel: ErrorChannel, *qubits: int) -> "NoiseModel":
        if len(qubits) == 0:
            raise IllegalArgumentException("must specify at least one qubit")
        if len(set(qubits)) != len(qubits):
            raise IllegalArgumentException("duplicate qubit index")
        if not _is_single_qubit(channel):
            raise IllegalArgumentException("onQubits requires channels that decompose to single-qubit operators")
        for q in qubits:
            self.qubit_noise.setdefault(int(q), []).append(channel)
        return self

    def after_all_gates(self, channel: ErrorChannel) -> "NoiseModel":
        self.global_noise.append(channel)
        return self

    def has_noise(self) -> bool:
        return bool(self.gate_noise or self.qubit_noise or self.global_noise)

    def scheduled_channels_after(self, op) -> list:
        out = []
        for ch in self.gate_noise.get(op.kind, []):
            out.append(ScheduledChannel(ch, ch.affected_qubits()))
        affected = set(op.qubits())
        for q in java_hashmap_order(self.qubit_noise.keys()):
            if q in affected:
                continue
            for ch in self.qubit_noise[q]:
                out.append(ScheduledChannel(ch, (q,)))
        for ch in self.global_noise:
            out.append(ScheduledChannel(ch, ch.affected_qubits()))
        return out
