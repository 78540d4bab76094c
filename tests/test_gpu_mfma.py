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
