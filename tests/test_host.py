"""CPU tests: the C-ABI library loads and exports every declared symbol; host-side logic (circuit
validation and lowering, SplittableRandom, NoiseModel scheduling, workload definitions)."""
import math
import re
from pathlib import Path

import numpy as np
import pytest

from statecraft_amd import capi, workloads
from statecraft_amd.capi import IllegalArgumentException
from statecraft_amd.circuit import Hadamard, PauliX, PauliZ, QuantumCircuit, lower, rotation_y
from statecraft_amd.noise import ErrorChannel, NoiseModel, java_hashmap_order

HEADER = Path(__file__).resolve().parent.parent / "include" / "statecraft_hip.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s*(sc_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    lib = capi.lib()
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(capi.EXPORTED)
    assert lib.sc_version() == 10000


def test_op_record_layout_matches_header():
    assert capi.OP_DTYPE.itemsize == 40
    assert capi.OP_DTYPE.fields["controls"][1] == 16
    assert capi.OP_DTYPE.fields["coeff"][1] == 32


def test_circuit_validation_messages():
    c = QuantumCircuit(3)
    with pytest.raises(IllegalArgumentException, match="target qubit out of range: 3"):
        c.append(Hadamard(), 3)
    with pytest.raises(IllegalArgumentException, match="control and target qubits must differ"):
        c.append_cnot(1, 1)
    with pytest.raises(IllegalArgumentException, match="cannot swap qubit with itself"):
        c.append_swap(2, 2)
    with pytest.raises(IllegalArgumentException, match="duplicate control qubit"):
        c.append_multi_control(PauliX(), 2, 0, 0)
    with pytest.raises(IllegalArgumentException, match="qubitCount must be positive"):
        QuantumCircuit(0)


def test_lowering_encodes_reference_fields():
    c = (QuantumCircuit(4).append(Hadamard(), 1).append_cnot(2, 0).append_multi_control(PauliZ(), 3, 2, 0)
         .append_controlled_phase(0.5, 1, 3).append_swap(0, 3).measure(2, 1))
    ops, mats = lower(c)
    assert list(ops["kind"]) == [capi.OP_SINGLE, capi.OP_CNOT, capi.OP_MCU, capi.OP_DIAG2, capi.OP_SWAP,
                                 capi.OP_MEASURE]
    assert ops[1]["q0"] == 2 and ops[1]["q1"] == 0
    assert ops[2]["controls"] == 0b101 and ops[2]["q0"] == 3
    assert ops[5]["controls"] == 0b110
    d = mats[ops[3]["mat"]:ops[3]["mat"] + 4]
    assert d[3] == complex(math.cos(0.5), math.sin(0.5)) and d[0] == 1


def test_java_hashmap_iteration_order():
    # keys below the capacity iterate in ascending order; 20 shares bucket 4 with 4 in a 16-bucket table
    assert java_hashmap_order([3, 1, 2]) == [1, 2, 3]
    assert java_hashmap_order([5, 20]) == [20, 5]
    assert java_hashmap_order([20, 4]) == [20, 4]
    assert java_hashmap_order([4, 20]) == [4, 20]


def test_noise_schedule_order():
    dep = ErrorChannel.depolarizing(0.1, 0)
    amp = ErrorChannel.amplitude_damping(0.2, 1)
    model = NoiseModel().after_gate(capi.OP_CNOT, dep).on_qubits(amp, 2, 1).after_all_gates(dep)
    c = QuantumCircuit(3).append_cnot(0, 1)
    sched = model.scheduled_channels_after(c.operations[0])
    assert [s.target_qubits for s in sched] == [(0,), (2,), (0,)]


def test_workload_sizes_and_credit():
    c2 = workloads.random_layered(30, 20, 7)
    kinds = [op.kind for op in c2.operations]
    assert kinds.count(capi.OP_SINGLE) == 1200 and kinds.count(capi.OP_CNOT) == 290
    assert workloads.amp_update_credit(c2) == 1200 * (1 << 30) + 290 * (1 << 29)
    c1, th = workloads.ry_qft(20, 42)
    assert len(c1.operations) == 20 + 20 + 190 + 10
    assert np.all((th >= 0) & (th < 2 * math.pi))


def test_grover_closed_form_small():
    import oracle

    circ, marked = workloads.grover(8, 3, seed=5)
    ops, mats = lower(circ)
    s = oracle.run_ops(8, ops, mats)
    assert abs(abs(s[marked]) ** 2 - workloads.grover_marked_probability(8, 3)) <= 1e-12
