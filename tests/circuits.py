"""Random circuit generators for parity tests (every op kind, both qubit orders, special gates)."""
from __future__ import annotations

import numpy as np

from statecraft_amd import capi


def random_unitary(d: int, rng) -> np.ndarray:
    z = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph


def random_state(n: int, rng) -> np.ndarray:
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


X = np.array([0, 1, 1, 0], dtype=np.complex128)
Z = np.array([1, 0, 0, -1], dtype=np.complex128)


def random_circuit(n: int, count: int, rng, kinds=None, max_controls=None):
    """Returns (ops structured array, complex mats)."""
    if kinds is None:
        kinds = [capi.OP_SINGLE, capi.OP_CNOT, capi.OP_TWO, capi.OP_DIAG2, capi.OP_SWAP, capi.OP_MCU]
    if n < 2:
        kinds = [capi.OP_SINGLE]
    ops = np.zeros(count, dtype=capi.OP_DTYPE)
    mats = []
    for i in range(count):
        k = int(rng.choice(kinds))
        if k == capi.OP_MCU and n < 2:
            k = capi.OP_SINGLE
        r = ops[i]
        r["kind"], r["mat"], r["q1"], r["param"] = k, -1, -1, -1
        if k in (capi.OP_SINGLE, capi.OP_MCU):
            q = int(rng.integers(0, n))
            r["q0"] = q
            flavour = rng.integers(0, 4)
            if flavour == 0:
                g = X
            elif flavour == 1:
                g = np.array([np.exp(1j * rng.uniform(0, 6.3)), 0, 0, np.exp(1j * rng.uniform(0, 6.3))])
            else:
                g = random_unitary(2, rng).reshape(-1)
            r["mat"] = len(mats)
            mats.extend(g)
            if k == capi.OP_MCU:
                others = [x for x in range(n) if x != q]
                cmax = len(others) if max_controls is None else min(max_controls, len(others))
                c = int(rng.integers(1, cmax + 1))
                cs = rng.choice(others, size=c, replace=False)
                r["controls"] = int(sum(1 << int(x) for x in cs))
        else:
            a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
            r["q0"], r["q1"] = a, b
            if k == capi.OP_CNOT:
                r["controls"] = 1 << a
            elif k == capi.OP_TWO:
                r["mat"] = len(mats)
                mats.extend(random_unitary(4, rng).reshape(-1))
            elif k == capi.OP_DIAG2:
                r["mat"] = len(mats)
                mats.extend(np.exp(1j * rng.uniform(0, 6.3, size=4)))
    return ops, np.asarray(mats if mats else [0j], dtype=np.complex128)


def structured_1q(rng, general: bool = True) -> np.ndarray:
    """A 2x2 gate (row-major) of the structured kinds circuits use: Ry, Rx, H, and with `general` also the
    ones without equal real pivots (Y: zero pivots, a phase times Ry: complex pivots, a random unitary)."""
    t = rng.uniform(0, 2 * np.pi)
    c, s = np.cos(t / 2), np.sin(t / 2)
    kinds = 3 + (3 if general else 0)
    k = int(rng.integers(0, kinds))
    if k == 0:
        return np.array([c, -s, s, c], dtype=np.complex128)
    if k == 1:
        return np.array([c, -1j * s, -1j * s, c], dtype=np.complex128)
    if k == 2:
        return np.array([1, 1, 1, -1], dtype=np.complex128) / np.sqrt(2)
    if k == 3:
        return np.array([0, -1j, 1j, 0], dtype=np.complex128)
    if k == 4:
        return np.exp(1j * t) * np.array([c, -s, s, c], dtype=np.complex128)
    return random_unitary(2, rng).reshape(-1)


def unit_pivot_circuit(n: int, count: int, rng, general: bool = True, diagonals: bool = True):
    """Structured single-qubit gates mixed with CNOTs (and Rz / two-qubit diagonals with `diagonals`)."""
    ops = np.zeros(count, dtype=capi.OP_DTYPE)
    mats = []
    for i in range(count):
        r = ops[i]
        r["mat"], r["q1"], r["param"] = -1, -1, -1
        u = rng.uniform()
        if u < 0.6:
            r["kind"], r["q0"], r["mat"] = capi.OP_SINGLE, int(rng.integers(0, n)), len(mats)
            mats.extend(structured_1q(rng, general))
        elif u < 0.8 or not diagonals:
            a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
            r["kind"], r["q0"], r["q1"] = capi.OP_CNOT, a, b
        elif u < 0.9:
            ph = rng.uniform(0, 6.3)
            r["kind"], r["q0"], r["mat"] = capi.OP_SINGLE, int(rng.integers(0, n)), len(mats)
            mats.extend([np.exp(-0.5j * ph), 0, 0, np.exp(0.5j * ph)])
        else:
            a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
            g = rng.uniform(0, 6.3)
            r["kind"], r["q0"], r["q1"], r["mat"] = capi.OP_DIAG2, a, b, len(mats)
            mats.extend(np.exp(1j * g * np.array([-1, 1, 1, -1])))
    return ops, np.array(mats, dtype=np.complex128)
