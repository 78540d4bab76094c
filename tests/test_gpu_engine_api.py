"""The drop-in engine API (HipStatevectorEngine.simulate) at configuration sizes, with default options.

A caller of the reference's SimulatorEngine gets the engine with its defaults; these tests run the same path:
config 1 (20 qubits, Ry layer + QFT, 4096 shots, seed 42) through simulate() against the oracle, and a
measurement of 32 qubits, which the reference's Integer-keyed Histogram cannot hold, as a BitstringHistogram
with MSB-left strings (TensorNetworkEngine.java:425-431).
"""
import numpy as np
import pytest

import oracle
from statecraft_amd import capi, workloads
from statecraft_amd.circuit import PauliX, QuantumCircuit, lower
from statecraft_amd.engine import (BitstringHistogram, BitstringSamples, Histogram, HipStatevectorEngine,
                                   MeasurementInstruction, SimulationRequest)
from statecraft_amd.rng import SplittableRandom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    e = HipStatevectorEngine()
    yield e
    e.close()


def test_config1_through_simulate_with_defaults(engine):
    n = 20
    circ, _ = workloads.ry_qft(n, 42)
    req = SimulationRequest.zero_state(circ).with_measurement(MeasurementInstruction.counts_all(4096).with_seed(42))
    res = engine.simulate(req)
    ops, mats = lower(circ)
    want = oracle.run_ops(n, ops, mats)
    assert np.max(np.abs(res.final_state - want)) <= 1e-12
    assert isinstance(res.measurement, Histogram) and res.measurement.shots == 4096
    u = SplittableRandom(42).next_doubles(4096)
    out = oracle.sample(res.final_state, list(range(n)), u)
    vals, cnts = np.unique(out, return_counts=True)
    assert res.measurement.counts == {int(v): int(c) for v, c in zip(vals, cnts)}


def test_wide_measurement_is_msb_left_bitstrings(engine):
    n = 32
    ones = (0, 3, 17, 30, 31)
    c = QuantumCircuit(n)
    for q in ones:
        c = c.append(PauliX(), q)
    want = "".join("1" if (n - 1 - k) in ones else "0" for k in range(n))
    res = engine.simulate(SimulationRequest.zero_state(c).with_measurement(
        MeasurementInstruction.counts_all(64).with_seed(9), False))
    assert res.final_state is None
    assert isinstance(res.measurement, BitstringHistogram)
    assert res.measurement.counts == {want: 64}
    res = engine.simulate(SimulationRequest.zero_state(c).with_measurement(
        MeasurementInstruction.samples_all(8).with_seed(9), False))
    assert isinstance(res.measurement, BitstringSamples) and res.measurement.outcomes == [want] * 8
