"""sc_run_jobs: batched circuits and their per-term expectations recorded once as a HIP graph and replayed.

Every replay must give what the separate entry points give (sc_state_set_basis, sc_apply_ops_batched,
sc_expectation_pauli_terms: the same plans, kernels and partial sums, so bit for bit), and the first
instance of each job must match the oracle within 1e-12.  A replay with new parameters reuses the graph.
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


@pytest.mark.parametrize("chunk_bits", [12, 13])
def test_graph_jobs_equal_separate_calls_and_oracle(jit_ctx, chunk_bits):
    jit_ctx.set_option(capi.OPT_CHUNK_BITS, chunk_bits)
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
                assert np.array_equal(g, w)
            for (n, idx, ops, mats, w), g in zip(batches, got):
                s = oracle.run_ops(n, ops, mats[0])
                zz = [oracle.expect_pauli(s, [0], [int(z)], [1.0]) for z in workloads.ring_zz_masks(n)]
                assert np.max(np.abs(g[0] - zz)) <= 1e-12
    finally:
        for st in states:
            st.close()
        jit_ctx.set_option(capi.OPT_CHUNK_BITS, 12)


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
