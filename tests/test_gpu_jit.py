"""GPU parity of the runtime-compiled sweep passes (SC_OPT_JIT 1 and 2): each pass as its own kernel with the
pass constant-folded, against the oracle (1e-12 max-abs) and against the compiled sweep kernel."""
import math

import numpy as np
import pytest

import bruteforce as bf
import oracle
from circuits import random_circuit, random_state, unit_pivot_circuit
from statecraft_amd import capi, workloads
from statecraft_amd.circuit import lower
from test_plan_host import _diag_real_circuit

pytestmark = pytest.mark.gpu
EPS = 1e-12


@pytest.fixture(params=[1, 2], ids=["jit", "jit-prefetch"])
def jit_ctx(hip_ctx, request):
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    hip_ctx.set_option(capi.OPT_JIT, request.param)
    yield hip_ctx
    hip_ctx.set_option(capi.OPT_JIT, 0)
    hip_ctx.set_option(capi.OPT_LOW_QUBITS, 4)
    hip_ctx.set_option(capi.OPT_REMAP, 1)


def run(ctx, n, ops, mats, init=None, batch=1, batched=False):
    st = capi.State(ctx, n, batch)
    if init is not None:
        st.upload(init)
    if batched:
        st.apply_batched(ops, mats)
    else:
        st.apply(ops, mats)
    out = st.download()
    st.close()
    return out


@pytest.mark.parametrize("n,low,remap", [(12, 0, 1), (13, 4, 1), (15, 4, 0), (17, 8, 1), (16, 2, 1)])
def test_jit_random_circuits(jit_ctx, n, low, remap):
    rng = np.random.default_rng(71 * n + low)
    ops, mats = random_circuit(n, 300, rng)
    init = random_state(n, rng)
    jit_ctx.set_option(capi.OPT_LOW_QUBITS, low)
    jit_ctx.set_option(capi.OPT_REMAP, remap)
    got = run(jit_ctx, n, ops, mats, init)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS
    jit_ctx.set_option(capi.OPT_JIT, 0)
    interp = run(jit_ctx, n, ops, mats, init)
    assert np.max(np.abs(got - interp)) <= EPS


def test_jit_dense_fusion(jit_ctx):
    rng = np.random.default_rng(5)
    n = 13
    ops, mats = random_circuit(n, 600, rng, kinds=[capi.OP_SINGLE, capi.OP_TWO])
    ops["q0"] = ops["q0"] % 4
    ops["q1"] = np.where(ops["kind"] == capi.OP_TWO, (ops["q0"] + 1 + (ops["q1"] % 3)) % 4, -1)
    init = random_state(n, rng)
    got = run(jit_ctx, n, ops, mats, init)
    assert jit_ctx.plan_stats()["fused_blocks"] > 0
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


def test_jit_kernels_follow_parameters(jit_ctx):
    """Circuits of one structure share their pass kernels; every run must read its own matrices."""
    n = 16
    for seed in (3, 4, 5):
        ops, mats = lower(workloads.random_layered(n, 8, seed))
        got = run(jit_ctx, n, ops, mats)
        assert np.max(np.abs(got - oracle.run_ops(n, ops, mats))) <= EPS, seed


def test_jit_outer_controls(jit_ctx):
    rng = np.random.default_rng(9)
    n = 18
    ops, mats = random_circuit(n, 60, rng, kinds=[capi.OP_MCU, capi.OP_SINGLE, capi.OP_CNOT])
    init = random_state(n, rng)
    got = run(jit_ctx, n, ops, mats, init)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


@pytest.mark.parametrize("n,low", [(13, 4), (16, 8)])
def test_jit_diagonal_tables_and_real_blocks(jit_ctx, n, low):
    rng = np.random.default_rng(17 + n)
    ops, mats = lower(_diag_real_circuit(n, rng))
    init = random_state(n, rng)
    jit_ctx.set_option(capi.OPT_LOW_QUBITS, low)
    got = run(jit_ctx, n, ops, mats, init)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


def test_jit_batched_per_element_matrices(jit_ctx):
    rng = np.random.default_rng(12)
    n, B = 13, 3
    ops, mats0 = random_circuit(n, 50, rng)
    mats = np.stack([np.asarray(mats0) * np.exp(1j * 0.3 * b) for b in range(B)])
    init = np.stack([random_state(n, rng) for _ in range(B)])
    got = run(jit_ctx, n, ops, mats, init, batch=B, batched=True)
    for b in range(B):
        assert np.max(np.abs(got[b] - oracle.run_ops(n, ops, mats[b], init[b]))) <= EPS


def test_jit_ry_qft_20q(jit_ctx):
    n = 20
    circ, thetas = workloads.ry_qft(n, 42)
    ops, mats = lower(circ)
    got = run(jit_ctx, n, ops, mats)
    c = np.ones(1 << n)
    idx = np.arange(1 << n)
    for q, t in enumerate(thetas):
        c *= np.where((idx >> q) & 1, math.sin(t / 2), math.cos(t / 2))
    assert np.max(np.abs(got - bf.qft_closed_form(c.astype(complex)))) <= EPS


def test_jit_random_layered_20q(jit_ctx):
    n = 20
    ops, mats = lower(workloads.random_layered(n, 20, 7))
    got = run(jit_ctx, n, ops, mats)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats))) <= EPS


@pytest.mark.parametrize("case", ["layered20", "random15", "batched13"])
def test_jit_in_register_transitions(jit_ctx, case):
    """SC_OPT_PERMLANE 1: stages entered through v_permlane32_swap / v_permlane16_swap exchanges (the 20-qubit
    layered circuit's plan has such transitions in most passes)."""
    jit_ctx.set_option(capi.OPT_LOW_QUBITS, 4)
    jit_ctx.set_option(capi.OPT_REMAP, 0)
    jit_ctx.set_option(capi.OPT_PERMLANE, 1)
    try:
        if case == "layered20":
            n = 20
            ops, mats = lower(workloads.random_layered(n, 20, 11))
            got = run(jit_ctx, n, ops, mats)
            want = oracle.run_ops(n, ops, mats)
        elif case == "random15":
            n = 15
            rng = np.random.default_rng(515)
            ops, mats = random_circuit(n, 400, rng)
            init = random_state(n, rng)
            got = run(jit_ctx, n, ops, mats, init)
            want = oracle.run_ops(n, ops, mats, init)
        else:
            n, batch = 13, 3
            ops, _ = lower(workloads.random_layered(n, 8, 3))
            per = [lower(workloads.random_layered(n, 8, 3 + 10 * b))[1] for b in range(batch)]
            got = run(jit_ctx, n, ops, np.stack(per), batch=batch, batched=True)
            want = np.stack([oracle.run_ops(n, ops, per[b]) for b in range(batch)])
    finally:
        jit_ctx.set_option(capi.OPT_PERMLANE, 0)
    assert np.max(np.abs(got - want)) <= EPS


@pytest.mark.parametrize("n,general,diagonals", [(13, False, False), (15, True, True)])
def test_jit_unit_pivot_forms(jit_ctx, n, general, diagonals):
    """Ry / Rx / H as unit-pivot ops with the pass's common factor in a table, a dense op or a scaling op."""
    rng = np.random.default_rng(40 + n)
    ops, mats = unit_pivot_circuit(n, 150, rng, general, diagonals)
    init = random_state(n, rng)
    got = run(jit_ctx, n, ops, mats, init)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


def test_jit_batched_unit_pivots_per_element(jit_ctx):
    """Per-element matrices: the unit-pivot form is used only where it holds for every element.  Op 0 is Y in
    element 1 (zero pivots) and Ry elsewhere, so it keeps the general form; the others differ in angle only."""
    rng = np.random.default_rng(77)
    n, B = 14, 3
    ops, mats0 = unit_pivot_circuit(n, 120, rng, False, True)
    ops[0]["kind"], ops[0]["q0"], ops[0]["q1"], ops[0]["mat"] = capi.OP_SINGLE, 3, -1, len(mats0)
    mats = []
    for b in range(B):
        m = np.array(mats0, dtype=np.complex128)
        single = np.nonzero(ops["kind"] == capi.OP_SINGLE)[0]
        for i in single[1:]:
            o = int(ops[i]["mat"])
            g = m[o:o + 4]
            if g[1] != 0 and g[0] == g[3] and g[0].imag == 0:  # Ry / Rx: another angle per element
                t = 0.37 * (b + 1)
                c, s_ = np.cos(t), np.sin(t)
                rot = np.array([[c, -s_], [s_, c]])
                m[o:o + 4] = (g.reshape(2, 2) @ rot if g[1].imag == 0 else g.reshape(2, 2)).reshape(-1)
        extra = np.array([0, -1j, 1j, 0]) if b == 1 else np.array([0.6, -0.8, 0.8, 0.6])
        mats.append(np.concatenate([m, extra]))
    mats = np.stack(mats)
    init = np.stack([random_state(n, rng) for _ in range(B)])
    got = run(jit_ctx, n, ops, mats, init, batch=B, batched=True)
    for b in range(B):
        assert np.max(np.abs(got[b] - oracle.run_ops(n, ops, mats[b], init[b]))) <= EPS, b


@pytest.fixture(params=[1, 2], ids=["jit-13", "jit-prefetch-13"])
def jit13_ctx(hip_ctx, request):
    """Runtime-compiled passes over 13-qubit chunks (512 threads, 128 KiB of LDS, one workgroup per CU)."""
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    hip_ctx.set_option(capi.OPT_JIT, request.param)
    hip_ctx.set_option(capi.OPT_CHUNK_BITS, 13)
    yield hip_ctx
    hip_ctx.set_option(capi.OPT_CHUNK_BITS, 12)
    hip_ctx.set_option(capi.OPT_JIT, 0)
    hip_ctx.set_option(capi.OPT_LOW_QUBITS, 4)
    hip_ctx.set_option(capi.OPT_REMAP, 1)


@pytest.mark.parametrize("n,low,remap", [(13, 0, 1), (14, 3, 0), (16, 4, 1), (18, 3, 0), (17, 8, 1)])
def test_jit13_random_circuits(jit13_ctx, n, low, remap):
    rng = np.random.default_rng(53 * n + low)
    ops, mats = random_circuit(n, 300, rng)
    init = random_state(n, rng)
    jit13_ctx.set_option(capi.OPT_LOW_QUBITS, low)
    jit13_ctx.set_option(capi.OPT_REMAP, remap)
    got = run(jit13_ctx, n, ops, mats, init)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS


def test_jit13_small_states_fall_back_to_twelve(jit13_ctx):
    """A 12-qubit state has no 13-qubit chunk: the plan uses 12."""
    rng = np.random.default_rng(4)
    ops, mats = random_circuit(12, 80, rng)
    init = random_state(12, rng)
    got = run(jit13_ctx, 12, ops, mats, init)
    assert np.max(np.abs(got - oracle.run_ops(12, ops, mats, init))) <= EPS


def test_jit13_batched_unit_pivots_and_layered(jit13_ctx):
    rng = np.random.default_rng(19)
    n, B = 15, 3
    ops, mats0 = unit_pivot_circuit(n, 150, rng, True, True)
    mats = np.stack([np.asarray(mats0) * np.exp(1j * 0.2 * b) for b in range(B)])
    init = np.stack([random_state(n, rng) for _ in range(B)])
    got = run(jit13_ctx, n, ops, mats, init, batch=B, batched=True)
    for b in range(B):
        assert np.max(np.abs(got[b] - oracle.run_ops(n, ops, mats[b], init[b]))) <= EPS, b
    jit13_ctx.set_option(capi.OPT_REMAP, 0)
    n = 20
    ops, mats = lower(workloads.random_layered(n, 12, 5))
    got = run(jit13_ctx, n, ops, mats)
    assert np.max(np.abs(got - oracle.run_ops(n, ops, mats))) <= EPS


def test_jit13_grover_closed_form(jit13_ctx):
    """Grover on 20 qubits with 3 low positions: the marked probability follows the closed form."""
    jit13_ctx.set_option(capi.OPT_LOW_QUBITS, 3)
    jit13_ctx.set_option(capi.OPT_REMAP, 0)
    n, it = 20, 3
    circ, marked = workloads.grover(n, it, 7)
    ops, mats = lower(circ)
    st = capi.State(jit13_ctx, n)
    st.apply(ops, mats)
    p = abs(st.read(marked)[0]) ** 2
    st.close()
    assert abs(p - workloads.grover_marked_probability(n, it)) <= 1e-12
