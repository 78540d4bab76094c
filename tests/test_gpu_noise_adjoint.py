"""GPU parity for noisy trajectories, Pauli expectations and adjoint gradients.

Noise: the engine's per-channel path and the batched trajectory kernels against the oracle's restatement of
StatevectorEngine.applyErrorChannel with the same SplittableRandom streams (1e-12).
Expectations: the oracle's Pauli-sum evaluation (1e-12).
Gradients: parameter-shift on the oracle (1e-10) and central finite differences h=1e-6 (1e-7, the accuracy
a double-precision central difference can reach).
"""
import math

import numpy as np
import pytest

import oracle
from circuits import random_circuit, random_state
from statecraft_amd import capi, workloads
from statecraft_amd.circuit import Hadamard, PauliX, QuantumCircuit, lower
from statecraft_amd.engine import HipStatevectorEngine, MeasurementInstruction, SimulationRequest
from statecraft_amd.noise import ErrorChannel, NoiseModel
from statecraft_amd.rng import SplittableRandom

pytestmark = pytest.mark.gpu
EPS = 1e-12


def oracle_noisy(circuit, model, seed):
    """Oracle run of a noisy circuit with the schedule NoiseModel.scheduledChannelsAfter yields."""
    n = circuit.qubit_count
    ops, mats = lower(circuit)
    s = oracle.zero_state(n)
    holder = [seed]
    for i, op in enumerate(circuit.operations):
        if op.kind == capi.OP_MEASURE:
            continue
        s = oracle.run_ops(n, ops[i:i + 1], mats, s)
        for sch in model.scheduled_channels_after(op):
            chans = sch.channel.channels if hasattr(sch.channel, "channels") else [sch.channel]
            for ch in chans:
                targets = sch.target_qubits if not hasattr(sch.channel, "channels") or len(sch.target_qubits) == 1 \
                    else ch.affected_qubits()
                for t in targets:
                    oracle.apply_channel_1q(s, t, ch.kraus, holder)
    return s


def test_engine_noisy_simulation_matches_oracle():
    circuit = (QuantumCircuit(5).append(Hadamard(), 0).append(Hadamard(), 3).append_cnot(3, 1).append_swap(0, 4)
               .append_controlled_phase(math.pi / 5.0, 4, 2).append_multi_control(PauliX(), 2, 0, 1, 3))
    model = (NoiseModel().after_all_gates(ErrorChannel.phase_flip(0.35, 0))
             .after_all_gates(ErrorChannel.amplitude_damping(0.2, 3))
             .on_qubits(ErrorChannel.depolarizing(0.3, 0), 4, 1))
    eng = HipStatevectorEngine()
    for seed in (99, 100, 101, 7):
        res = eng.simulate(SimulationRequest.zero_state(circuit).with_noise_model(model).with_noise_seed(seed))
        want = oracle_noisy(circuit, model, seed)
        assert np.max(np.abs(res.final_state - want)) <= EPS, seed


def test_engine_golden_noise_histogram():
    # StatevectorNoiseTest golden: X then amplitude damping gamma=1 always relaxes to |0>: counts {0: 16}
    circuit = QuantumCircuit(1).append(PauliX(), 0)
    model = NoiseModel().after_all_gates(ErrorChannel.amplitude_damping(1.0, 0))
    req = (SimulationRequest.zero_state(circuit).with_noise_model(model).with_noise_seed(3)
           .with_measurement(MeasurementInstruction.counts_all(16).with_seed(1)))
    res = HipStatevectorEngine().simulate(req)
    assert res.measurement.counts == {0: 16}


def test_engine_measurement_counts_bit_identical():
    circ, _ = workloads.ry_qft(12, 42)
    req = SimulationRequest.zero_state(circ).with_measurement(MeasurementInstruction.counts_all(4096).with_seed(42))
    res = HipStatevectorEngine().simulate(req)
    ops, mats = lower(circ)
    s = oracle.run_ops(12, ops, mats)
    out = oracle.sample(s, list(range(12)), SplittableRandom(42).next_doubles(4096))
    vals, cnts = np.unique(out, return_counts=True)
    assert res.measurement.counts == {int(v): int(c) for v, c in zip(vals, cnts)}


def _noise_schedule(n, n_ops, rng, every):
    """Channels after every `every`-th op (one or two each, on random qubits): sc_run_noisy events, the stacked
    Kraus operators, and (op, target, operators) per event for the oracle."""
    chans = [ErrorChannel.amplitude_damping(0.3, 0).kraus, ErrorChannel.phase_damping(0.4, 0).kraus,
             ErrorChannel.depolarizing(0.2, 0).kraus, ErrorChannel.thermal_relaxation(50.0, 60.0, 10.0, 0).kraus,
             ErrorChannel.phase_flip(0.25, 0).kraus]
    events, kraus, per = [], [], []
    for i in range(0, n_ops, every):
        for _ in range(int(rng.integers(1, 3))):
            k = chans[int(rng.integers(0, len(chans)))]
            t = int(rng.integers(0, n))
            events.append((i, t, len(kraus), len(k)))
            per.append((i, t, k))
            kraus.extend(k)
    return events, kraus, per


def _oracle_schedule(n, ops, mats, per, seed, init):
    s = np.array(init, np.complex128, copy=True)
    holder, j = [seed], 0
    for i in range(len(ops)):
        s = oracle.run_ops(n, ops[i:i + 1], mats, s)
        while j < len(per) and per[j][0] == i:
            oracle.apply_channel_1q(s, per[j][1], per[j][2], holder)
            j += 1
    return s


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("n,B,every", [(6, 1, 1), (13, 1, 1), (13, 1, 7), (16, 3, 5), (12, 4, 12)])
def test_run_noisy_matches_oracle(n, B, every, jit):
    """sc_run_noisy: state-dependent channels (amplitude and phase damping, depolarising, thermal relaxation,
    phase flip) with the branch chosen on the device, after single ops (per-op kernels between channels) and
    after longer segments (the context's execution mode), every batch element with its own uniforms, against
    the oracle's applyErrorChannel with the same SplittableRandom streams (1e-12)."""
    ctx = capi.Context(0)
    ctx.set_option(capi.OPT_JIT, jit)
    try:
        rng = np.random.default_rng(9000 + 100 * n + 10 * B + every)
        ops, mats = random_circuit(n, 36, rng, max_controls=2)
        events, kraus, per = _noise_schedule(n, len(ops), rng, every)
        seeds = [4000 + 17 * b for b in range(B)]
        u = np.stack([SplittableRandom(sd).next_doubles(len(events)) for sd in seeds])
        init = np.stack([random_state(n, rng) for _ in range(B)])
        st = capi.State(ctx, n, B)
        st.upload(init.reshape(-1))
        st.run_noisy(ops, mats, events, kraus, u)
        got = st.download().reshape(B, -1)
        st.close()
        for b in range(B):
            want = _oracle_schedule(n, ops, mats, per, seeds[b], init[b])
            assert np.max(np.abs(got[b] - want)) <= EPS, (b, np.max(np.abs(got[b] - want)))
    finally:
        ctx.close()


@pytest.mark.parametrize("n", [5, 13])
def test_run_noisy_channels_after_one_op(hip_ctx, n):
    """Several channels after one op share sweeps (the apply of one with the weights of the next): on the same
    qubit, on another qubit below and above it, then a channel after a later op; against the oracle (1e-12)."""
    rng = np.random.default_rng(9500 + n)
    ops, mats = random_circuit(n, 12, rng, max_controls=1)
    ad = list(ErrorChannel.amplitude_damping(0.35, 0).kraus)
    pd = list(ErrorChannel.phase_damping(0.5, 0).kraus)
    dep = list(ErrorChannel.depolarizing(0.4, 0).kraus)
    kraus = ad + pd + dep
    sched = [(2, 3, ad), (2, 3, pd), (2, n - 1, dep), (2, 0, ad), (2, 0, ad), (7, 1, pd), (7, n - 2, ad)]
    first = {id(ad): 0, id(pd): 2, id(dep): 4}
    events = [(i, t, first[id(k)], len(k)) for i, t, k in sched]
    per = [(i, t, k) for i, t, k in sched]
    for seed in (31, 32, 33):
        u = SplittableRandom(seed).next_doubles(len(events)).reshape(1, -1)
        init = random_state(n, rng)
        st = capi.State(hip_ctx, n, 1)
        st.upload(init)
        st.run_noisy(ops, mats, events, kraus, u)
        got = st.download()
        st.close()
        want = _oracle_schedule(n, ops, mats, per, seed, init)
        assert np.max(np.abs(got - want)) <= EPS, seed


def test_run_noisy_reports_the_first_failing_channel(hip_ctx):
    """A channel whose weights are all zero (total <= 1e-15) or not a number fails the run with the event, its
    weights and the kind (the reference's IllegalStateException); later calls on the context still work."""
    n = 12
    circuit = QuantumCircuit(n)
    for q in range(n):
        circuit = circuit.append(Hadamard(), q)
    ops, mats = lower(circuit)
    damp = ErrorChannel.amplitude_damping(0.3, 0).kraus
    zero = [np.zeros((2, 2)), np.zeros((2, 2))]
    nan = [np.eye(2), np.array([[np.nan, 0], [0, 0]])]
    kraus = list(damp) + zero + nan
    events = [(3, 1, 0, 2), (5, 2, 2, 2), (6, 0, 4, 2)]
    st = capi.State(hip_ctx, n, 1)
    st.set_basis(0)
    with pytest.raises(capi.NoisyRunFailed) as e:
        st.run_noisy(ops, mats, events, kraus, np.full((1, 3), 0.5))
    assert (e.value.event, e.value.kind, e.value.weights[:2]) == (1, 2, [0.0, 0.0])
    assert "total=0" in str(e.value)
    st.set_basis(0)
    with pytest.raises(capi.NoisyRunFailed) as e:
        st.run_noisy(ops, mats, events[2:], kraus, np.full((1, 1), 0.5))
    assert (e.value.event, e.value.kind, e.value.branch) == (0, 1, 1)
    assert abs(e.value.weights[0] - 1.0) <= 1e-12 and np.isnan(e.value.weights[1])
    # argument checks (before any kernel)
    with pytest.raises(capi.IllegalArgumentException):
        st.run_noisy(ops, mats, [(5, 0, 0, 2), (3, 0, 0, 2)], kraus, np.full((1, 2), 0.5))
    with pytest.raises(capi.IllegalArgumentException):
        st.run_noisy(ops, mats, [(3, n, 0, 2)], kraus, np.full((1, 1), 0.5))
    with pytest.raises(capi.IllegalArgumentException):
        st.run_noisy(ops, mats, [(3, 0, 5, 2)], kraus, np.full((1, 1), 0.5))
    with pytest.raises(capi.IllegalArgumentException):
        st.run_noisy(ops, mats, [(3, 0, 0, 2)], kraus, np.full((1, 1), 1.0))
    # the context still runs: H on every qubit, damping on qubit 1 with weights (0.85, 0.15) and r = 0.75 picks
    # K_0 = diag(1, sqrt(0.7)); the state is K_0 |+..+> / sqrt(0.85)
    st.set_basis(0)
    st.run_noisy(ops, mats, events[:1], kraus, np.full((1, 1), 0.75))
    bit1 = (np.arange(1 << n) >> 1) & 1
    want = np.where(bit1 == 1, np.sqrt(0.7), 1.0) * 2.0 ** (-n / 2) / np.sqrt(0.85)
    assert np.max(np.abs(st.download() - want)) <= EPS
    st.close()


@pytest.mark.parametrize("n,B,p", [(12, 8, 0.1), (14, 5, 0.3)])
def test_batched_trajectories(hip_ctx, n, B, p):
    circuit = workloads.ghz(n)
    ops, mats = lower(circuit)
    noise_after = (ops["kind"] == capi.OP_CNOT).astype(np.uint8)
    n_draws = 2 * int(noise_after.sum())
    seeds = [1000 + b for b in range(B)]
    u = np.stack([SplittableRandom(s).next_doubles(n_draws) for s in seeds])
    st = capi.State(hip_ctx, n, B)
    st.run_depolarizing_trajectories(ops, mats, noise_after, p, u)
    got = st.download()
    dep = ErrorChannel.depolarizing(p, 0).kraus
    for b in range(B):
        s = oracle.zero_state(n)
        holder = [seeds[b]]
        for i in range(len(ops)):
            s = oracle.run_ops(n, ops[i:i + 1], mats, s)
            if noise_after[i]:
                oracle.apply_channel_1q(s, int(ops[i]["q0"]), dep, holder)
                oracle.apply_channel_1q(s, int(ops[i]["q1"]), dep, holder)
        assert np.max(np.abs(got[b] - s)) <= EPS, b
    # trajectory-averaged GHZ parity <Z^n>
    par = st.expectation([0], [(1 << n) - 1], [1.0])
    for b in range(B):
        assert abs(par[b] - oracle.expect_pauli(got[b], [0], [(1 << n) - 1], [1.0])) <= EPS
    st.close()


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("n,B", [(14, 48), (20, 24)])
def test_trajectories_with_z_terms(n, B, fused):
    """sc_run_depolarizing_trajectories_z: the GHZ parity and other Z strings of every trajectory, summed by the
    noisy circuit's last pass (SC_OPT_FUSED_TERMS 1) or by the term kernel afterwards (0), against the plain
    call followed by the term kernel (1e-14) and the oracle (1e-12)."""
    ctx = capi.Context(0)  # library defaults: runtime-compiled passes
    ctx.set_option(capi.OPT_FUSED_TERMS, fused)
    try:
        ops, mats, noise_after, u = workloads.ghz_trajectory_uniforms(B, n, 77)
        z = np.array([(1 << n) - 1, 1, (1 << (n - 1)) | 1, 0b1011 << 3, 0], np.uint64)
        st = capi.State(ctx, n, B)
        st.set_basis(0)
        got = st.run_depolarizing_trajectories_z(ops, mats, noise_after, 0.05, u, z)
        states = st.download()
        st.set_basis(0)
        st.run_depolarizing_trajectories(ops, mats, noise_after, 0.05, u)
        want = st.expectation_terms(np.zeros(len(z), np.uint64), z)
        assert np.max(np.abs(got - want)) <= 1e-14
        assert np.array_equal(states, st.download())
        for b in (0, B - 1):
            ref = [oracle.expect_pauli(states[b], [0], [int(zz)], [1.0]) for zz in z]
            assert np.max(np.abs(got[b] - ref)) <= EPS
        st.close()
    finally:
        ctx.close()


def test_pauli_expectations(hip_ctx):
    rng = np.random.default_rng(17)
    n, B = 13, 3
    states = np.stack([random_state(n, rng) for _ in range(B)])
    xs = [int(x) for x in rng.integers(0, 1 << n, 12)]
    zs = [int(z) for z in rng.integers(0, 1 << n, 12)]
    cs = rng.normal(size=12)
    st = capi.State(hip_ctx, n, B)
    st.upload(states)
    got = st.expectation(xs, zs, cs)
    for b in range(B):
        assert abs(got[b] - oracle.expect_pauli(states[b], xs, zs, cs)) <= EPS
    st.close()


@pytest.mark.parametrize("n,B", [(13, 3), (20, 2)])
def test_pauli_terms_each_term(hip_ctx, n, B):
    """Per-term expectations: 35 diagonal terms (two launches of one x-group), repeated x masks, Y factors."""
    rng = np.random.default_rng(n)
    states = np.stack([random_state(n, rng) for _ in range(B)])
    full = (1 << n) - 1
    xs = [0] * 35 + [int(x) for x in rng.integers(1, 1 << n, 6)] * 2 + [1 << (n - 1), 1]
    zs = [int(z) for z in rng.integers(0, 1 << n, len(xs))]
    zs[0], zs[1], zs[-1] = 0, full, full
    st = capi.State(hip_ctx, n, B)
    st.upload(states)
    got = st.expectation_terms(xs, zs)
    assert got.shape == (B, len(xs))
    for b in range(B):
        want = [oracle.expect_pauli(states[b], [x], [z], [1.0]) for x, z in zip(xs, zs)]
        assert np.max(np.abs(got[b] - want)) <= EPS, b
    assert np.allclose(got[:, 0], 1.0, atol=EPS)  # identity term: the norm
    # the weighted sum agrees with the per-term values
    cs = rng.normal(size=len(xs))
    assert np.max(np.abs(st.expectation(xs, zs, cs) - got @ cs)) <= EPS
    st.close()


@pytest.fixture(params=[0, 2], ids=["per-op", "adjoint-passes"])
def adj_ctx(hip_ctx, request):
    """The reverse sweep per op (SC_OPT_JIT 0), or as runtime-compiled adjoint passes (n >= 12)."""
    hip_ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    hip_ctx.set_option(capi.OPT_JIT, request.param)
    yield hip_ctx
    hip_ctx.set_option(capi.OPT_JIT, 0)


def hea_energy(n, layers, thetas, obs):
    ops, mats = lower(workloads.hea(n, layers, thetas))
    return oracle.expect_pauli(oracle.run_ops(n, ops, mats), *obs)


@pytest.mark.parametrize("n,layers,ckpt", [(6, 2, 0), (12, 2, 10), (13, 3, 25)])
def test_adjoint_hea_vs_parameter_shift(adj_ctx, n, layers, ckpt):
    hip_ctx = adj_ctx
    rng = np.random.default_rng(n)
    P = 2 * n * layers
    B = 2
    thetas = rng.uniform(0, 2 * math.pi, size=(B, P))
    obs = workloads.hea_observable(n, rng.normal(size=n), rng.normal(size=n))
    ops, _ = lower(workloads.hea(n, layers, thetas[0]))
    mats = np.stack([lower(workloads.hea(n, layers, thetas[b]))[1] for b in range(B)])
    st = capi.State(hip_ctx, n, B)
    grads, energies = st.adjoint_gradient(ops, mats, P, *obs, checkpoint_every=ckpt, per_element=True)
    st.close()
    for b in range(B):
        e0 = hea_energy(n, layers, thetas[b], obs)
        assert abs(energies[b] - e0) <= EPS
        for p in range(0, P, max(1, P // 12)):
            tp, tm = thetas[b].copy(), thetas[b].copy()
            tp[p] += math.pi / 2
            tm[p] -= math.pi / 2
            ps = 0.5 * (hea_energy(n, layers, tp, obs) - hea_energy(n, layers, tm, obs))
            assert abs(grads[b, p] - ps) <= 1e-10, (b, p)
            h = 1e-6
            tp, tm = thetas[b].copy(), thetas[b].copy()
            tp[p] += h
            tm[p] -= h
            fd = (hea_energy(n, layers, tp, obs) - hea_energy(n, layers, tm, obs)) / (2 * h)
            assert abs(grads[b, p] - fd) <= 1e-7, (b, p)


def test_adjoint_checkpoint_reanchor(adj_ctx):
    hip_ctx = adj_ctx
    n, layers = 12, 2
    rng = np.random.default_rng(3)
    P = 2 * n * layers
    thetas = rng.uniform(0, 2 * math.pi, size=P)
    obs = workloads.hea_observable(n, rng.normal(size=n), rng.normal(size=n))
    ops, mats = lower(workloads.hea(n, layers, thetas))
    st = capi.State(hip_ctx, n, 1)
    g0, _ = st.adjoint_gradient(ops, mats, P, *obs, checkpoint_every=16)
    assert hip_ctx.adjoint_stats()["reanchors"] == 0
    hip_ctx.set_option(capi.OPT_TEST_DRIFT, 1000)  # inject 1e-9 drift before every checkpoint check
    try:
        st.set_basis(0)
        g1, _ = st.adjoint_gradient(ops, mats, P, *obs, checkpoint_every=16)
    finally:
        hip_ctx.set_option(capi.OPT_TEST_DRIFT, 0)
    stats = hip_ctx.adjoint_stats()
    assert stats["reanchors"] > 0 and stats["max_drift"] > 1e-10
    assert np.max(np.abs(g1 - g0)) <= 1e-10
    st.close()


def test_adjoint_qaoa_vs_finite_differences(adj_ctx):
    hip_ctx = adj_ctx
    n, p = 12, 2
    rng = np.random.default_rng(44)
    gam, bet = rng.uniform(0, math.pi, p), rng.uniform(0, math.pi, p)
    w = rng.uniform(0.5, 1.5, n)
    const, xm, zm, cf = workloads.maxcut_observable(n, w)
    circ = workloads.qaoa_ring(n, p, gam, bet, w)
    ops, mats = lower(circ)
    st = capi.State(hip_ctx, n, 1)
    grads, energies = st.adjoint_gradient(ops, mats, 2 * p, xm, zm, cf)
    st.close()

    def energy(g, b):
        o, m = lower(workloads.qaoa_ring(n, p, g, b, w))
        return oracle.expect_pauli(oracle.run_ops(n, o, m), xm, zm, cf)

    assert abs(energies[0] - energy(gam, bet)) <= EPS
    h = 1e-6
    for l in range(p):
        for which in (0, 1):
            gp, gm, bp, bm = gam.copy(), gam.copy(), bet.copy(), bet.copy()
            if which == 0:
                gp[l] += h
                gm[l] -= h
            else:
                bp[l] += h
                bm[l] -= h
            fd = (energy(gp, bp) - energy(gm, bm)) / (2 * h)
            assert abs(grads[0, 2 * l + which] - fd) <= 1e-7


def test_adjoint_passes_equal_per_op_reverse_sweep(hip_ctx):
    """Runtime-compiled adjoint passes and the per-op reverse sweep give the same gradients and energies
    (batch of 3 parameter sets, checkpoints every 40 ops, 14 qubits)."""
    n, layers, B = 14, 3, 3
    rng = np.random.default_rng(99)
    P = 2 * n * layers
    thetas = rng.uniform(0, 2 * math.pi, size=(B, P))
    obs = workloads.hea_observable(n, rng.normal(size=n), rng.normal(size=n))
    ops, _ = lower(workloads.hea(n, layers, thetas[0]))
    mats = np.stack([lower(workloads.hea(n, layers, thetas[b]))[1] for b in range(B)])
    out = []
    for jit in (0, 2):
        hip_ctx.set_option(capi.OPT_JIT, jit)
        try:
            st = capi.State(hip_ctx, n, B)
            out.append(st.adjoint_gradient(ops, mats, P, *obs, checkpoint_every=40, per_element=True))
            final = st.download()
            st.close()
        finally:
            hip_ctx.set_option(capi.OPT_JIT, 0)
        for b in range(B):  # the state is left at the circuit's output
            o, m = lower(workloads.hea(n, layers, thetas[b]))
            assert np.max(np.abs(final[b] - oracle.run_ops(n, o, m))) <= 1e-12
    (g0, e0), (g1, e1) = out
    assert np.max(np.abs(g1 - g0)) <= 1e-12
    assert np.max(np.abs(e1 - e0)) <= 1e-12
