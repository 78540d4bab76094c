"""GPU parity at the full sizes of the five benchmark configurations.

The reference's own near-limit test simulates a 22-qubit basis state and checks it and its norm
(StatevectorEngineTest.java:366-380); these tests do the same at every configuration's real size, where the
64-bit addressing, the deposit of many chunk-fixed bits, 32-control masks and the exact sampler over 2^30 and
2^33 outcomes are exercised.  Checks that do not need the whole state on the host run on the device
(sc_state_max_deviation):

* 30 qubits, random layered circuit of depth 20 (1490 ops): the probability checksum is 1 within 1e-12; 8192
  shots over all 30 qubits equal the oracle's sequential sampler on the downloaded 16 GiB state bit for bit;
  the circuit followed by its inverse returns |0...0> within 1e-12 in every amplitude.
* 33 qubits, Grover with 32-control Z gates, 8 iterations: every one of the 2^33 amplitudes equals the closed
  form a|m> + b sum_{x != m} |x> within 1e-12; 1024 shots land where the near-uniform CDF puts them.
* 4096 batched QAOA rings of 12-22 qubits: <C> of the first instance of every size against the oracle
  (1e-12); 1024 GHZ-20 depolarising trajectories, 8 of them against the oracle's channel restatement with the
  same SplittableRandom streams (1e-12).
* 24-qubit, 8-layer HEA, 384 parameters, batch 64: adjoint gradients against parameter shift on the GPU
  engine (1e-10) for 16 parameters of two sets and against central differences h = 1e-6 (1e-7); energies
  against the oracle (1e-12).
"""
import math
import sys

import numpy as np
import pytest

import oracle
from statecraft_amd import capi, workloads
from statecraft_amd.circuit import QuantumCircuit, lower, rotation_y, rotation_z
from statecraft_amd.noise import ErrorChannel
from statecraft_amd.rng import SplittableRandom

pytestmark = pytest.mark.gpu
EPS = 1e-12


def say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def big_ctx():
    # library defaults: runtime-compiled passes, chunk and low qubits chosen per circuit structure
    ctx = capi.Context(0)
    yield ctx
    ctx.close()


def random_layered_inverse(n, depth, seed):
    """The inverse of workloads.random_layered(n, depth, seed): reversed ops, Ry(-theta), Rz(-phi)."""
    u = SplittableRandom(seed).next_doubles(2 * n * depth) * 2.0 * math.pi
    c = QuantumCircuit(n)
    for d in reversed(range(depth)):
        start = 0 if d % 2 == 0 else 1
        for a in reversed(range(start, n - 1, 2)):
            c = c.append_cnot(a, a + 1)
        for q in reversed(range(n)):
            k = 2 * (d * n + q)
            c = c.append(rotation_z(-float(u[k + 1])), q)
            c = c.append(rotation_y(-float(u[k])), q)
    return c


def test_c2_random_layered_30q(big_ctx):
    n, depth, seed = 30, 20, 7
    ops, mats = lower(workloads.random_layered(n, depth, seed))
    st = capi.State(big_ctx, n)
    try:
        say("c2: 30 qubits, depth 20, 1490 ops")
        st.apply(ops, mats)
        say(f"c2: geometry {big_ctx.plan_geometry()}")
        cs = st.checksum()
        assert abs(cs[0] - 1.0) <= EPS, cs
        u = SplittableRandom(seed).next_doubles(8192)
        got, gap = st.sample(list(range(n)), u)
        stats = big_ctx.sample_stats()
        say(f"c2: checksum {cs}, sample stats {stats}, boundary gap {gap:.3e}; downloading 16 GiB")
        host = st.download()
        want = oracle.sample(host, list(range(n)), u)
        mism = np.nonzero(got != want)[0]
        assert len(mism) == 0, f"{len(mism)} of 8192 shots differ, first at shot {mism[0]}"
        # the oracle's checksum on the same amplitudes (different summation order: 1e-12)
        ocs = oracle.checksum(host)
        assert abs(ocs[0] - cs[0]) <= EPS and abs(ocs[1] - cs[1]) <= EPS
        del host
        say("c2: inverse circuit")
        iops, imats = lower(random_layered_inverse(n, depth, seed))
        st.apply(iops, imats)
        dev = st.max_deviation(0, 1.0, 0.0)
        say(f"c2: max |psi - |0>| after the inverse = {dev:.3e}")
        assert dev <= EPS
    finally:
        st.close()


def test_c3_grover_33q(big_ctx):
    n, iters, seed = 33, 8, 33
    circ, marked = workloads.grover(n, iters, seed)
    ops, mats = lower(circ)
    assert max(bin(int(c)).count("1") for c in ops["controls"]) == 32
    st = capi.State(big_ctx, n)
    try:
        say(f"c3: 33 qubits (128 GiB), Grover, 8 iterations, marked {marked}")
        st.apply(ops, mats)
        say(f"c3: geometry {big_ctx.plan_geometry()}")
        # the diffusion H X MCZ X H is -(2|s><s| - I): each iteration adds a factor -1, (-1)^8 = 1
        th = math.asin(2.0 ** (-n / 2.0))
        a = math.sin((2 * iters + 1) * th)
        b = math.cos((2 * iters + 1) * th) / math.sqrt(2.0 ** n - 1.0)
        dev = st.max_deviation(marked, a, b)
        say(f"c3: max |psi - closed form| over 2^33 amplitudes = {dev:.3e}")
        assert dev <= EPS
        pm = abs(st.read(marked)[0]) ** 2
        assert abs(pm - workloads.grover_marked_probability(n, iters)) <= 1e-12 * 1e-4 + 1e-20
        u = SplittableRandom(seed).next_doubles(1024)
        got, _ = st.sample(list(range(n)), u)
        say(f"c3: sample stats {big_ctx.sample_stats()}")
        # near-uniform CDF: cum[i] ~ (i + 1) b^2 (+ a^2 - b^2 past the marked index); the sequential running sum
        # drifts from it by at most 2^33 half-ulps of 1, about 2^13 outcomes
        b2, a2 = b * b, a * a
        pred = np.where(u <= marked * b2, np.ceil(u / b2) - 1, np.ceil((u - a2 + b2) / b2) - 1)
        assert np.all(np.abs(got - pred) <= 2 ** 14), np.max(np.abs(got - pred))
        assert np.all((got >= 0) & (got < 2 ** n))
    finally:
        st.close()


def test_c4_qaoa_batch_and_ghz_trajectories(big_ctx):
    batches = workloads.qaoa_batches(4096, 4, 4)
    assert sum(len(idx) for _, idx, _, _, _ in batches) == 4096
    say(f"c4: {len(batches)} QAOA size batches")
    for n, idx, ops, mats, w in batches:
        B = len(idx)
        st = capi.State(big_ctx, n, B)
        try:
            st.apply_batched(ops, mats)
            zz = st.expectation_terms(np.zeros(n, np.uint64), workloads.ring_zz_masks(n))
            mean_c = 0.5 * np.sum(w * (1.0 - zz), axis=1)
            assert np.all((mean_c >= -EPS) & (mean_c <= np.sum(w, axis=1) + EPS))
            s = oracle.run_ops(n, ops, mats[0])
            want = 0.5 * sum(w[0, i] * (1.0 - oracle.expect_pauli(s, [0], [int(workloads.ring_zz_masks(n)[i])], [1.0]))
                             for i in range(n))
            assert abs(mean_c[0] - want) <= EPS, (n, mean_c[0], want)
        finally:
            st.close()
    nt, nq, p = 1024, 20, 0.01
    ops, mats, noise_after, u = workloads.ghz_trajectory_uniforms(nt, nq)
    st = capi.State(big_ctx, nq, nt)
    try:
        say("c4: 1024 GHZ-20 trajectories")
        st.run_depolarizing_trajectories(ops, mats, noise_after, p, u)
        parity = st.expectation([0], [(1 << nq) - 1], [1.0])
        assert np.all(np.abs(parity) <= 1.0 + EPS)
        dep = ErrorChannel.depolarizing(p, 0).kraus
        for t in (0, 1, 2, 3, 511, 700, 1022, 1023):
            s = oracle.zero_state(nq)
            holder = [1000 + t]
            for i in range(len(ops)):
                s = oracle.run_ops(nq, ops[i:i + 1], mats, s)
                if noise_after[i]:
                    oracle.apply_channel_1q(s, int(ops[i]["q0"]), dep, holder)
                    oracle.apply_channel_1q(s, int(ops[i]["q1"]), dep, holder)
            got = st.read(0, 1 << nq, elem=t)
            assert np.max(np.abs(got - s)) <= EPS, t
            assert abs(parity[t] - oracle.expect_pauli(s, [0], [(1 << nq) - 1], [1.0])) <= EPS
    finally:
        st.close()


def test_c5_adjoint_hea_24q(big_ctx):
    n, layers, batch = 24, 8, 64
    P = 2 * n * layers
    rng = np.random.default_rng(5)
    thetas = SplittableRandom(5).next_doubles(batch * P).reshape(batch, P) * 2 * math.pi
    obs = workloads.hea_observable(n, rng.normal(size=n), rng.normal(size=n))
    ops, _ = lower(workloads.hea(n, layers, thetas[0]))
    mats = np.stack([lower(workloads.hea(n, layers, thetas[k]))[1] for k in range(batch)])
    st = capi.State(big_ctx, n, batch)
    try:
        say("c5: adjoint, 24 qubits, 384 parameters, batch 64")
        grads, energies = st.adjoint_gradient(ops, mats, P, *obs, checkpoint_every=144, per_element=True)
    finally:
        st.close()
    one = capi.State(big_ctx, n, 1)

    def energy(th):
        o, m = lower(workloads.hea(n, layers, th))
        one.set_basis(0)
        one.apply(o, m)
        return float(one.expectation(*obs)[0])

    try:
        params = list(range(0, P, P // 16))[:16]
        for k in (0, 37):
            assert abs(energy(thetas[k]) - energies[k]) <= EPS
            for p in params:
                tp, tm = thetas[k].copy(), thetas[k].copy()
                tp[p] += math.pi / 2
                tm[p] -= math.pi / 2
                ps = 0.5 * (energy(tp) - energy(tm))
                assert abs(ps - grads[k, p]) <= 1e-10, (k, p, ps, grads[k, p])
        say("c5: parameter shift agrees; central differences")
        h = 1e-6
        for p in params[::2]:
            tp, tm = thetas[0].copy(), thetas[0].copy()
            tp[p] += h
            tm[p] -= h
            fd = (energy(tp) - energy(tm)) / (2 * h)
            assert abs(fd - grads[0, p]) <= 1e-7, (p, fd, grads[0, p])
    finally:
        one.close()
    say("c5: energy of set 0 on the oracle (24 qubits, 576 ops)")
    s = oracle.run_ops(n, ops, mats[0])
    assert abs(oracle.expect_pauli(s, *obs) - energies[0]) <= EPS
