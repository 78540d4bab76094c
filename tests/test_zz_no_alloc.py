"""sc_expectation_pauli_terms_noalloc: the per-term expectation entry point that allocates no device memory.

It shares k_pauli_terms with sc_expectation_pauli_terms and keeps its partial sums in a fixed device array
instead of the context's scratch.  Two calls in a row, the second on a smaller input, must equal the
allocating entry point's results (the same kernels and the same number of partial sums at these sizes, so
bit for bit; the bar for expectations is 1e-12), and a request beyond the array is refused before any work.
"""
import numpy as np
import pytest

from circuits import random_state
from statecraft_amd import capi

pytestmark = pytest.mark.gpu


def _terms(n, count, rng):
    xm = rng.integers(0, 1 << n, size=count, dtype=np.uint64)
    xm[: count // 3] = 0  # diagonal terms: the x = 0 group
    xm[count // 3: count // 3 + 5] = xm[count // 3]  # a group of several terms sharing one x mask
    zm = rng.integers(0, 1 << n, size=count, dtype=np.uint64)
    return xm, zm


def _states(ctx, n, batch, rng):
    st = capi.State(ctx, n, batch)
    st.upload(np.concatenate([random_state(n, rng) for _ in range(batch)]))
    return st


def test_noalloc_matches_allocating_entry_point_twice():
    rng = np.random.default_rng(0x0A11)
    ctx = capi.Context(0)
    try:
        for n, batch, count in ((15, 3, 40), (12, 2, 9)):
            st = _states(ctx, n, batch, rng)
            xm, zm = _terms(n, count, rng)
            got = st.expectation_terms_noalloc(xm, zm)
            want = st.expectation_terms(xm, zm)
            assert got.shape == (batch, count)
            assert np.max(np.abs(got - want)) <= 1e-12
            assert np.array_equal(got, want)
            st.close()
    finally:
        ctx.close()


def test_noalloc_refuses_more_than_its_array():
    ctx = capi.Context(0)
    try:
        st = capi.State(ctx, 2, 1)
        count = (1 << 21) + 1  # term rows x batch beyond the 32 MiB array
        z = np.zeros(count, np.uint64)
        with pytest.raises(capi.IllegalArgumentException, match="scratch"):
            st.expectation_terms_noalloc(z, z)
        st.close()
    finally:
        ctx.close()
