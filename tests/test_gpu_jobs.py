"""sc_run_jobs: batched circuits and their per-term expectations recorded once as a HIP graph and replayed.

Every replay must give what the separate entry points give (sc_state_set_basis, sc_apply_ops_batched,
sc_expectation_pauli_terms): bit for bit with SC_OPT_FUSED_TERMS 0 (the same plans, kernels and partial sums),
within 1e-14 when Z-string terms are summed in the last pass (another summation order); and the first instance
of each job must match the oracle within 1e-12.  A replay with new parameters reuses the graph.
"""
import numpy as np
import pytest

import oracle
from statecraft_amd import capi, workloads
from statecraft_amd.circuit import lower

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jit_ctx():
    ctx = capi.Context(0, capi.EXEC_SWEEP)
    ctx.set_option(capi.OPT_JIT, 2)
    ctx.set_option(capi.OPT_REMAP, 0)
    yield ctx
    ctx.close()


def _jobs(ctx, batches, states):
    return [(st, 0, ops, mats, np.zeros(n, np.uint64), workloads.ring_zz_masks(n))
            for st, (n, idx, ops, mats, w) in zip(states, batches)]


def _separate(states, batches):
    outs = []
    for st, (n, idx, ops, mats, w) in zip(states, batches):
        st.set_basis(0)
        st.apply_batched(ops, mats)
        outs.append(st.expectation_terms(np.zeros(n, np.uint64), workloads.ring_zz_masks(n)))
    return outs


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("chunk_bits", [12, 13])
def test_graph_jobs_equal_separate_calls_and_oracle(jit_ctx, chunk_bits, fused):
    jit_ctx.set_option(capi.OPT_CHUNK_BITS, chunk_bits)
    jit_ctx.set_option(capi.OPT_FUSED_TERMS, fused)
    batches = [b for b in workloads.qaoa_batches(300, 4, 11) if b[0] <= 16]
    assert len(batches) >= 3
    states = [capi.State(jit_ctx, n, len(idx)) for n, idx, _, _, _ in batches]
    try:
        for rnd in range(3):
            if rnd == 2:  # new parameters, same structure: the recorded graph is replayed
                rng = np.random.default_rng(3)
                new = []
                for n, idx, ops, mats, w in batches:
                    g, b = rng.random((len(idx), 4)) * np.pi, rng.random((len(idx), 4)) * np.pi
                    m = np.stack([lower(workloads.qaoa_ring(n, 4, g[k], b[k], w[k]))[1] for k in range(len(idx))])
                    new.append((n, idx, ops, m, w))
                batches = new
            got, used = capi.run_jobs(jit_ctx, _jobs(jit_ctx, batches, states))
            assert used
            want = _separate(states, batches)
            for g, w in zip(got, want):
                assert np.array_equal(g, w) if not fused else np.max(np.abs(g - w)) <= 1e-14
            for (n, idx, ops, mats, w), g in zip(batches, got):
                s = oracle.run_ops(n, ops, mats[0])
                zz = [oracle.expect_pauli(s, [0], [int(z)], [1.0]) for z in workloads.ring_zz_masks(n)]
                assert np.max(np.abs(g[0] - zz)) <= 1e-12
    finally:
        for st in states:
            st.close()
        jit_ctx.set_option(capi.OPT_CHUNK_BITS, 12)
        jit_ctx.set_option(capi.OPT_FUSED_TERMS, 1)


def test_fused_and_kernel_terms_in_one_call(jit_ctx):
    """A job with an X term keeps the term kernel, an all-Z job sums its terms in its last pass; 18- and
    20-qubit batches, whose last passes cover several elements per workgroup and several workgroups per
    element."""
    batches = [b for b in workloads.qaoa_batches(200, 4, 17) if b[0] in (18, 20)]
    assert len(batches) == 2
    states = [capi.State(jit_ctx, n, len(idx)) for n, idx, _, _, _ in batches]
    try:
        jobs = []
        for k, (st, (n, idx, ops, mats, w)) in enumerate(zip(states, batches)):
            z = workloads.ring_zz_masks(n)
            x = np.zeros(n, np.uint64)
            if k == 0:
                x[1] = 1 << 3  # X_3 Z_1 Z_2: the x-paired term kernel
            jobs.append((st, 0, ops, mats, x, z))
        got, used = capi.run_jobs(jit_ctx, jobs)
        assert used
        for st, (n, idx, ops, mats, w), (_, _, _, _, x, z), g in zip(states, batches, jobs, got):
            st.set_basis(0)
            st.apply_batched(ops, mats)
            want = st.expectation_terms(x, z)
            assert np.max(np.abs(g - want)) <= 1e-14
            s0 = oracle.run_ops(n, ops, mats[0])
            ref = [oracle.expect_pauli(s0, [int(xx)], [int(zz)], [1.0]) for xx, zz in zip(x, z)]
            assert np.max(np.abs(g[0] - ref)) <= 1e-12
    finally:
        for st in states:
            st.close()


def test_jobs_without_runtime_compiled_passes_run_one_by_one(hip_ctx):
    batches = [b for b in workloads.qaoa_batches(120, 4, 13) if b[0] <= 13]
    states = [capi.State(hip_ctx, n, len(idx)) for n, idx, _, _, _ in batches]
    try:
        got, used = capi.run_jobs(hip_ctx, _jobs(hip_ctx, batches, states))
        assert not used
        for g, w in zip(got, _separate(states, batches)):
            assert np.array_equal(g, w)
    finally:
        for st in states:
            st.close()


def test_jobs_reject_bad_input(jit_ctx):
    st = capi.State(jit_ctx, 12, 2)
    try:
        ops, mats = workloads.qaoa_batches(40, 4, 14)[0][2:4]
        with pytest.raises(capi.IllegalArgumentException):
            capi.run_jobs(jit_ctx, [(st, 1 << 12, ops[:0], np.zeros((2, 1), complex), [0], [1])])
        with pytest.raises(capi.IllegalArgumentException):
            capi.run_jobs(jit_ctx, [(st, 0, ops[:0], np.zeros((2, 1), complex), [0], [1 << 12])])
    finally:
        st.close()
