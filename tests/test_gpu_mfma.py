"""GPU parity of the matrix-core dense blocks (SC_OPT_MFMA 2: fused-mode 4- and 5-qubit blocks computed as
complex GEMMs with v_mfma_f64_16x16x4_f64), against the pinned CPU oracle and against the FP64 VALU block
kernel (SC_OPT_MFMA 0) on the same circuits.

Bar: amplitudes within 1e-12 max-abs of the oracle (the GEMM sums a row in another order than the VALU
kernel, so the two GPU paths agree to rounding, not bit for bit).
"""
import numpy as np
import pytest

import oracle
from circuits import random_circuit, random_state
from statecraft_amd import capi

pytestmark = pytest.mark.gpu
EPS = 1e-12


def run_fused(ctx, n, ops, mats, init, k, mfma, batch=1):
    ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_FUSED)
    ctx.set_option(capi.OPT_MAX_FUSED, k)
    ctx.set_option(capi.OPT_MFMA, 2 if mfma else 0)
    try:
        st = capi.State(ctx, n, batch=batch)
        st.upload(init)
        st.apply(ops, mats)
        out = st.download()
        st.close()
    finally:
        ctx.set_option(capi.OPT_MFMA, 1)
        ctx.set_option(capi.OPT_MAX_FUSED, 4)
        ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_SWEEP)
    return out


@pytest.mark.parametrize("k", [4, 5])
@pytest.mark.parametrize("n", [8, 9, 10, 13, 16])
def test_mfma_blocks_match_oracle(hip_ctx, k, n):
    # n - k < 4 leaves fewer than 16 groups per block: those blocks take the VALU kernel
    rng = np.random.default_rng(4100 + 10 * n + k)
    ops, mats = random_circuit(n, 150, rng, max_controls=2)
    init = random_state(n, rng)
    want = oracle.run_ops(n, ops, mats, init)
    got = run_fused(hip_ctx, n, ops, mats, init, k, mfma=True)
    assert np.max(np.abs(got - want)) <= EPS
    valu = run_fused(hip_ctx, n, ops, mats, init, k, mfma=False)
    assert np.max(np.abs(got - valu)) <= EPS


def test_mfma_blocks_batched(hip_ctx):
    # one circuit on three states: the groups of every element form one GEMM
    n, batch = 11, 3
    rng = np.random.default_rng(4242)
    ops, mats = random_circuit(n, 120, rng, max_controls=1)
    init = np.stack([random_state(n, rng) for _ in range(batch)])
    got = run_fused(hip_ctx, n, ops, mats, init.reshape(-1), 5, mfma=True, batch=batch).reshape(batch, -1)
    for b in range(batch):
        assert np.max(np.abs(got[b] - oracle.run_ops(n, ops, mats, init[b]))) <= EPS


def test_mfma_option_roundtrip(hip_ctx):
    assert hip_ctx.get_option(capi.OPT_MFMA) == 1  # library default: 5-qubit blocks on the matrix cores
    hip_ctx.set_option(capi.OPT_MFMA, 2)
    assert hip_ctx.get_option(capi.OPT_MFMA) == 2
    with pytest.raises(capi.IllegalArgumentException):
        hip_ctx.set_option(capi.OPT_MFMA, 3)
    hip_ctx.set_option(capi.OPT_MFMA, 1)


@pytest.mark.parametrize("k", [4, 5])
def test_mfma_tile_shapes_agree(hip_ctx, k):
    """Every tile shape of SC_OPT_MFMA_TILE gives the same amplitudes bit for bit (one product per group in
    the same order), within 1e-12 of the oracle; a batch of three 9-qubit states (an odd tile count with
    5-qubit blocks) runs the two-tile shapes as one-tile steps."""
    for n, batch in ((13, 1), (9, 3)):
        rng = np.random.default_rng(4300 + 10 * n + k)
        ops, mats = random_circuit(n, 120, rng, max_controls=1)
        init = np.stack([random_state(n, rng) for _ in range(batch)])
        outs = []
        try:
            for shape in range(6):
                hip_ctx.set_option(capi.OPT_MFMA_TILE, shape)
                outs.append(run_fused(hip_ctx, n, ops, mats, init.reshape(-1), k, mfma=True, batch=batch))
        finally:
            hip_ctx.set_option(capi.OPT_MFMA_TILE, -1)
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])
        got = outs[0].reshape(batch, -1)
        for b in range(batch):
            assert np.max(np.abs(got[b] - oracle.run_ops(n, ops, mats, init[b]))) <= EPS


def test_mfma_tile_autotuner():
    """With SC_OPT_MFMA_TILE -1 (the default) the first matrix-core block of a size times the six shapes and
    keeps one for the context; the choice is readable and read-only, and the result matches the oracle."""
    ctx = capi.Context(0)
    try:
        assert ctx.get_option(capi.OPT_MFMA_TILE) == -1
        assert (ctx.get_option(capi.OPT_MFMA_PICK4), ctx.get_option(capi.OPT_MFMA_PICK5)) == (-1, -1)
        n = 12
        rng = np.random.default_rng(4400)
        ops, mats = random_circuit(n, 80, rng, max_controls=1)
        init = random_state(n, rng)
        # library default SC_OPT_MFMA 1: only the 5-qubit blocks run on the matrix cores
        ctx.set_option(capi.OPT_EXEC_MODE, capi.EXEC_FUSED)
        ctx.set_option(capi.OPT_MAX_FUSED, 5)
        st = capi.State(ctx, n)
        st.upload(init)
        st.apply(ops, mats)
        got = st.download()
        st.close()
        assert np.max(np.abs(got - oracle.run_ops(n, ops, mats, init))) <= EPS
        assert 0 <= ctx.get_option(capi.OPT_MFMA_PICK5) < 6 and ctx.get_option(capi.OPT_MFMA_PICK4) == -1
        with pytest.raises(capi.IllegalArgumentException):
            ctx.set_option(capi.OPT_MFMA_PICK5, 0)
        with pytest.raises(capi.IllegalArgumentException):
            ctx.set_option(capi.OPT_MFMA_TILE, 6)
    finally:
        ctx.close()
