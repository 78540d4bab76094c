"""Independent brute-force statevector semantics (test infrastructure).

Formulated as tensor contractions on a (2,)*n array (axis n-1-q <-> qubit q),
a different construction from both the oracle's index loops and the kernels,
so that a transposed operand, a wrong bit order or a dropped term in either
shows up as a mismatch.  Semantics follow QuantumCircuit.apply
(core/src/main/java/com/omaarr90/statecraft/quantum/QuantumCircuit.java:144-360).
"""
from __future__ import annotations

import numpy as np

OP_SINGLE, OP_CNOT, OP_TWO, OP_DIAG2, OP_SWAP, OP_MCU, OP_MEASURE = range(1, 8)


def _apply_dense(state: np.ndarray, n: int, qubits: list, u: np.ndarray) -> np.ndarray:
    """u acts on `qubits` with qubits[0] as the MOST significant bit of its local index."""
    k = len(qubits)
    t = state.reshape((2,) * n)
    axes = [n - 1 - q for q in qubits]
    ut = u.reshape((2,) * (2 * k))
    out = np.tensordot(ut, t, axes=(list(range(k, 2 * k)), axes))
    # tensordot puts the k output axes first; move them back into place
    out = np.moveaxis(out, list(range(k)), axes)
    return out.reshape(-1)


def controlled(g: np.ndarray, c: int) -> np.ndarray:
    """Dense matrix on (controls..., target) with target least significant: I - P + P (x) g."""
    d = 1 << (c + 1)
    u = np.eye(d, dtype=np.complex128)
    u[d - 2:, d - 2:] = g
    return u


def apply_op(state: np.ndarray, n: int, kind: int, q0: int, q1: int, controls: list, m: np.ndarray) -> np.ndarray:
    if kind == OP_SINGLE:
        return _apply_dense(state, n, [q0], m.reshape(2, 2))
    if kind == OP_CNOT:
        x = np.array([[0, 1], [1, 0]], dtype=np.complex128)
        return _apply_dense(state, n, [q0, q1], controlled(x, 1))
    if kind == OP_TWO:
        return _apply_dense(state, n, [q0, q1], m.reshape(4, 4))
    if kind == OP_DIAG2:
        return _apply_dense(state, n, [q0, q1], np.diag(m.reshape(4)))
    if kind == OP_SWAP:
        s = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
        return _apply_dense(state, n, [q0, q1], s)
    if kind == OP_MCU:
        return _apply_dense(state, n, list(controls) + [q0], controlled(m.reshape(2, 2), len(controls)))
    raise ValueError(kind)


def run(n: int, ops, mats, init=None) -> np.ndarray:
    st = np.zeros(1 << n, dtype=np.complex128)
    st[0] = 1.0
    if init is not None:
        st = np.array(init, dtype=np.complex128, copy=True)
    for r in ops:
        kind = int(r["kind"])
        if kind == OP_MEASURE:
            continue
        mat = int(r["mat"])
        cnt = {OP_SINGLE: 4, OP_MCU: 4, OP_DIAG2: 4, OP_TWO: 16}.get(kind, 0)
        m = np.asarray(mats[mat:mat + cnt]) if cnt else None
        cm = int(r["controls"])
        controls = [q for q in range(64) if (cm >> q) & 1] if kind == OP_MCU else []
        st = apply_op(st, n, kind, int(r["q0"]), int(r["q1"]), controls, m)
    return st


def pauli_matrix(n: int, xmask: int, zmask: int) -> np.ndarray:
    """Dense 2^n matrix of the Pauli string with X on xmask, Z on zmask, Y where both."""
    I = np.eye(2, dtype=np.complex128)
    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
    Z = np.diag([1.0, -1.0]).astype(np.complex128)
    m = np.ones((1, 1), dtype=np.complex128)
    for q in reversed(range(n)):  # kron order: qubit n-1 is the most significant factor
        x, z = (xmask >> q) & 1, (zmask >> q) & 1
        p = Y if (x and z) else X if x else Z if z else I
        m = np.kron(m, p)
    return m


def qft_closed_form(c: np.ndarray) -> np.ndarray:
    """Output of the reference QFT construction (AlgorithmSuiteSpecs.buildQftCircuit, H + cp + final swaps) on
    input amplitudes c: out = sqrt(N) * ifft(c[rev]) [rev], rev = n-bit reversal."""
    N = len(c)
    n = N.bit_length() - 1
    idx = np.arange(N)
    rev = np.zeros(N, dtype=np.int64)
    for b in range(n):
        rev |= ((idx >> b) & 1) << (n - 1 - b)
    y = np.sqrt(N) * np.fft.ifft(c[rev])
    return y[rev]
