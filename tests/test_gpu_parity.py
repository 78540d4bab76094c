"""GPU parity: the HIP engine (through the C-ABI) against the pinned CPU oracle on the same seeded inputs.

Bar: amplitudes and probabilities within 1e-12 max-abs; sampled outcomes bit-identical given the same
uniforms.  Every execution mode (one kernel per op, dense fused blocks, LDS-blocked sweeps) is covered.
"""
import math

import numpy as np
import pytest

import bruteforce as bf
import oracle
from circuits import random_circuit, random_state
from statecraft_amd import capi, workloads
from statecraft_amd.circuit import Hadamard, PauliX, PauliZ, QuantumCircuit, lower
from statecraft_amd.rng import SplittableRandom

pytestmark = pytest.mark.gpu
EPS = 1e-12
MODES = [capi.EXEC_DIRECT, capi.EXEC_FUSED, capi.EXEC_SWEEP]


def run_gpu(ctx, n, ops, mats, init=None, mode=capi.EXEC_SWEEP):
    ctx.set_option(capi.OPT_EXEC_MODE, mode)
    st = capi.State(ctx, n)
    if init is not None:
        st.upload(init)
    st.apply(ops, mats)
    out = st.download()
    st.close()
    return out


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 13, 14])
def test_random_circuits(hip_ctx, mode, n):
    rng = np.random.default_rng(1000 * n + mode)
    ops, mats = random_circuit(n, 80, rng)
    init = random_state(n, rng)
    got = run_gpu(hip_ctx, n, ops, mats, init, mode)
    want = oracle.run_ops(n, ops, mats, init)
    assert np.max(np.abs(got - want)) <= EPS


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_fused_block_sizes(hip_ctx, k):
    rng = np.random.default_rng(77 + k)
    n = 9
    ops, mats = random_circuit(n, 120, rng, max_controls=2)
    init = random_state(n, rng)
    hip_ctx.set_option(capi.OPT_MAX_FUSED, k)
    try:
        got = run_gpu(hip_ctx, n, ops, mats, init, capi.EXEC_FUSED)
    finally:
        hip_ctx.set_option(capi.OPT_MAX_FUSED, 4)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


@pytest.mark.parametrize("low", [0, 2, 4, 6, 8])
@pytest.mark.parametrize("n", [12, 15, 17])
def test_sweep_layouts(hip_ctx, low, n):
    rng = np.random.default_rng(31 * n + low)
    ops, mats = random_circuit(n, 400, rng)
    init = random_state(n, rng)
    hip_ctx.set_option(capi.OPT_LOW_QUBITS, low)
    try:
        got = run_gpu(hip_ctx, n, ops, mats, init, capi.EXEC_SWEEP)
    finally:
        hip_ctx.set_option(capi.OPT_LOW_QUBITS, 4)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


def test_sweep_dense_fusion_path(hip_ctx):
    # long runs of 1q gates on four qubits trigger the fused 16x16 register block
    rng = np.random.default_rng(5)
    n = 13
    ops, mats = random_circuit(n, 600, rng, kinds=[capi.OP_SINGLE, capi.OP_TWO])
    ops["q0"] = ops["q0"] % 4
    ops["q1"] = np.where(ops["kind"] == capi.OP_TWO, (ops["q0"] + 1 + (ops["q1"] % 3)) % 4, -1)
    init = random_state(n, rng)
    got = run_gpu(hip_ctx, n, ops, mats, init, capi.EXEC_SWEEP)
    assert hip_ctx.plan_stats()["fused_blocks"] > 0
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


def test_many_controls_outside_chunk(hip_ctx):
    rng = np.random.default_rng(9)
    n = 18
    ops, mats = random_circuit(n, 60, rng, kinds=[capi.OP_MCU, capi.OP_SINGLE, capi.OP_CNOT])
    init = random_state(n, rng)
    for mode in MODES:
        got = run_gpu(hip_ctx, n, ops, mats, init, mode)
        assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS, mode


def test_batched_per_element_matrices(hip_ctx):
    rng = np.random.default_rng(12)
    n, B = 13, 3
    ops, mats0 = random_circuit(n, 50, rng)
    # same structure, element-specific matrices (a global phase per element keeps every matrix unitary)
    mats = np.stack([mats0 * np.exp(1j * 0.1 * b) for b in range(B)])
    init = np.stack([random_state(n, rng) for _ in range(B)])
    for mode in MODES:
        hip_ctx.set_option(capi.OPT_EXEC_MODE, mode)
        st = capi.State(hip_ctx, n, B)
        st.upload(init)
        st.apply_batched(ops, mats)
        got = st.download()
        st.close()
        for b in range(B):
            want = oracle.run_ops(n, ops, mats[b], init[b])
            assert np.max(np.abs(got[b] - want)) <= EPS, (mode, b)
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)


def test_shared_circuit_batch(hip_ctx):
    rng = np.random.default_rng(13)
    n, B = 12, 4
    ops, mats = random_circuit(n, 70, rng)
    init = np.stack([random_state(n, rng) for _ in range(B)])
    st = capi.State(hip_ctx, n, B)
    st.upload(init)
    st.apply(ops, mats)
    got = st.download()
    for b in range(B):
        assert np.max(np.abs(got[b] - oracle.run_ops(n, ops, mats, init[b]))) <= EPS


def test_ry_qft_20q_config(hip_ctx):
    n = 20
    circ, thetas = workloads.ry_qft(n, 42)
    ops, mats = lower(circ)
    for mode in MODES:
        got = run_gpu(hip_ctx, n, ops, mats, None, mode)
        c = np.ones(1 << n)
        idx = np.arange(1 << n)
        for q, t in enumerate(thetas):
            c *= np.where((idx >> q) & 1, math.sin(t / 2), math.cos(t / 2))
        assert np.max(np.abs(got - bf.qft_closed_form(c.astype(complex)))) <= EPS, mode
    want = oracle.run_ops(n, ops, mats)
    assert np.max(np.abs(got - want)) <= EPS
    # 4096 shots, seed 42, all qubits: bit-identical outcomes
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    st = capi.State(hip_ctx, n)
    st.apply(ops, mats)
    u = SplittableRandom(42).next_doubles(4096)
    out, gap = st.sample(list(range(n)), u)
    assert np.array_equal(out, oracle.sample(want, list(range(n)), u))
    p = st.probabilities()
    assert np.max(np.abs(p - oracle.probabilities(want, list(range(n))))) <= EPS
    st.close()


def test_random_layered_family_20q(hip_ctx):
    n = 20
    circ = workloads.random_layered(n, 20, 7)
    ops, mats = lower(circ)
    want = oracle.run_ops(n, ops, mats)
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    st = capi.State(hip_ctx, n)
    st.apply(ops, mats)
    got = st.download()
    assert np.max(np.abs(got - want)) <= EPS
    u = SplittableRandom(8192).next_doubles(8192)
    out, gap = st.sample(list(range(n)), u)
    assert np.array_equal(out, oracle.sample(want, list(range(n)), u))
    cs = st.checksum()
    ocs = oracle.checksum(want)
    assert abs(cs[0] - ocs[0]) <= EPS and abs(cs[1] - ocs[1]) <= EPS
    st.close()


@pytest.mark.parametrize("measured", [[0], [19], [3, 11], [0, 1, 2, 17], list(range(0, 20, 2)), list(range(1, 20))])
def test_marginals_and_samples(hip_ctx, measured):
    rng = np.random.default_rng(len(measured) * 7 + measured[0])
    n = 20
    s = random_state(n, rng)
    st = capi.State(hip_ctx, n)
    st.upload(s)
    p = st.probabilities(measured)
    assert np.max(np.abs(p - oracle.probabilities(s, measured))) <= EPS
    u = SplittableRandom(3).next_doubles(5000)
    out, _ = st.sample(measured, u)
    assert np.array_equal(out, oracle.sample(s, measured, u))
    st.close()


def test_unnormalised_state_sampling(hip_ctx):
    # sum of probabilities far from 1 triggers normalize's rescale branch (StatevectorEngine.java:306-311)
    rng = np.random.default_rng(4)
    s = random_state(14, rng) * 1.7
    st = capi.State(hip_ctx, 14)
    st.upload(s)
    u = SplittableRandom(5).next_doubles(3000)
    out, _ = st.sample(list(range(14)), u)
    assert np.array_equal(out, oracle.sample(s, list(range(14)), u))
    st.close()


def test_zero_state_sampling_raises(hip_ctx):
    st = capi.State(hip_ctx, 4)
    st.upload(np.zeros(16, complex))
    with pytest.raises(capi.IllegalStateException, match="zero probability"):
        st.sample([0, 1], [0.5])
    st.close()


def test_error_paths(hip_ctx):
    st = capi.State(hip_ctx, 3)
    ops, mats = lower(QuantumCircuit(5).append(Hadamard(), 4))
    with pytest.raises(capi.IllegalArgumentException, match="target qubit out of range: 4"):
        st.apply(ops, mats)
    ops, mats = lower(QuantumCircuit(3).measure(0).append(Hadamard(), 1))
    with pytest.raises(capi.UnsupportedOperationException, match="cannot follow measurement"):
        st.apply(ops, mats)
    with pytest.raises(capi.IllegalArgumentException):
        st.upload(np.zeros(4, complex))
    st.close()
    with pytest.raises(capi.UnsupportedOperationException):
        capi.State(hip_ctx, 35)


def test_determinism(hip_ctx):
    rng = np.random.default_rng(21)
    n = 16
    ops, mats = random_circuit(n, 200, rng)
    a = run_gpu(hip_ctx, n, ops, mats, None, capi.EXEC_SWEEP)
    b = run_gpu(hip_ctx, n, ops, mats, None, capi.EXEC_SWEEP)
    assert np.array_equal(a, b)


def test_plan_cache_reuse_and_classification(hip_ctx):
    """Plans are cached by circuit structure: circuits that differ only in parameters reuse a plan and must
    read their own matrices; a matrix that changes class (diagonal / real / exchange) at the same op must not
    reuse the plan made for the other class."""
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    n = 16
    for seed in (3, 4, 5):
        ops, mats = lower(workloads.random_layered(n, 6, seed))
        got = run_gpu(hip_ctx, n, ops, mats, None, capi.EXEC_SWEEP)
        assert np.max(np.abs(got - oracle.run_ops(n, ops, mats))) <= EPS, seed
    rng = np.random.default_rng(3)
    ops, mats = random_circuit(n, 120, rng, kinds=[capi.OP_SINGLE, capi.OP_CNOT])
    init = random_state(n, rng)
    singles = np.flatnonzero(ops["kind"] == capi.OP_SINGLE)
    for variant in range(3):
        m = np.array(mats, copy=True)
        for i in singles[::3]:
            o = int(ops[i]["mat"])
            m[o:o + 4] = [np.array([1, 0, 0, 1j]), np.array([0, 1, 1, 0]), np.array([1, 1, 1, -1]) / np.sqrt(2)][variant]
        got = run_gpu(hip_ctx, n, ops, m, init, capi.EXEC_SWEEP)
        assert np.max(np.abs(got - oracle.run_ops(n, ops, m, init))) <= EPS, variant


def test_basis_state_is_written_when_first_needed(hip_ctx):
    """sc_state_set_basis only records the basis state: a runtime-compiled sweep generates it in its first
    pass, every other operation writes it first.  Each reader sees |k> in every batch element."""
    n, B, k = 13, 2, 4097
    want = np.zeros(1 << n, complex)
    want[k] = 1.0
    st = capi.State(hip_ctx, n, B)
    st.set_basis(k)
    assert np.array_equal(st.download().reshape(B, -1), np.stack([want] * B))
    st.set_basis(k)
    assert st.read(k, 1, elem=1)[0] == 1.0 and st.norm_sq(1) == 1.0
    st.set_basis(k)
    assert st.probabilities()[k] == 1.0
    st.set_basis(k)
    assert np.all(st.expectation([0], [1], [1.0]) == -1.0)  # <Z_0> on |4097>: bit 0 is set
    st.close()


@pytest.mark.parametrize("jit", [0, 2])
def test_circuit_from_a_basis_state(hip_ctx, jit):
    """A circuit applied right after set_basis: generated in the first pass (runtime-compiled sweeps) or
    written before the first op (other paths); both equal the oracle from the basis state."""
    rng = np.random.default_rng(31 + jit)
    n, k = 15, 12345
    ops, mats = random_circuit(n, 120, rng)
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    hip_ctx.set_option(capi.OPT_JIT, jit)
    init = np.zeros(1 << n, complex)
    init[k] = 1.0
    st = capi.State(hip_ctx, n)
    st.set_basis(k)
    st.apply(ops, mats)
    got = st.download()
    st.close()
    hip_ctx.set_option(capi.OPT_JIT, 0)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS
