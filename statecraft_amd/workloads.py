"""Circuit families of the benchmark configurations, built with the circuit mirror.

Angles are drawn from SplittableRandom(seed).nextDouble() in the order the
builders below state, so every consumer (GPU engine, CPU oracle, tests)
sees identical circuits.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi
from .circuit import Hadamard, PauliX, PauliZ, QuantumCircuit, rotation_y, rotation_z, rx_suite, zz_phase
from .rng import SplittableRandom

TWO_PI = 2.0 * math.pi


def qft(circuit: QuantumCircuit, n: int) -> QuantumCircuit:
    """AlgorithmSuiteSpecs.buildQftCircuit (app/.../AlgorithmSuiteSpecs.java:257-270)."""
    for target in range(n):
        circuit = circuit.append(Hadamard(), target)
        for control in range(target + 1, n):
            circuit = circuit.append_controlled_phase(math.pi / (1 << (control - target)), control, target)
    for i in range(n // 2):
        circuit = circuit.append_swap(i, n - i - 1)
    return circuit


def ry_qft(n: int = 20, seed: int = 42):
    """Config 1: Ry(theta_i) on every qubit (theta_i = 2*pi*u_i, u from SplittableRandom(seed), qubit order),
    then the textbook QFT.  Returns (circuit, thetas)."""
    thetas = SplittableRandom(seed).next_doubles(n) * TWO_PI
    c = QuantumCircuit(n)
    for q in range(n):
        c = c.append(rotation_y(float(thetas[q])), q)
    return qft(c, n), thetas


def random_layered(n: int = 30, depth: int = 20, seed: int = 7) -> QuantumCircuit:
    """Config 2: per layer Ry(theta) Rz(phi) on every qubit (for each qubit: theta then phi, each 2*pi*u), then
    CNOT bricks on (0,1),(2,3).. for even layers and (1,2),(3,4).. for odd layers."""
    rng = SplittableRandom(seed)
    u = rng.next_doubles(2 * n * depth) * TWO_PI
    c = QuantumCircuit(n)
    k = 0
    for d in range(depth):
        for q in range(n):
            c = c.append(rotation_y(float(u[k])), q)
            c = c.append(rotation_z(float(u[k + 1])), q)
            k += 2
        start = 0 if d % 2 == 0 else 1
        for a in range(start, n - 1, 2):
            c = c.append_cnot(a, a + 1)
    return c


def amp_update_credit(circuit: QuantumCircuit) -> int:
    """Amplitudes each op of the unfused circuit must change: 1q / 2q dense / diagonal: 2^n; CNOT and
    multi-controlled U with c controls: 2^(n-c); SWAP: 2^(n-1)."""
    n = circuit.qubit_count
    total = 0
    for op in circuit.operations:
        if op.kind in (capi.OP_SINGLE, capi.OP_TWO, capi.OP_DIAG2):
            total += 1 << n
        elif op.kind == capi.OP_CNOT:
            total += 1 << (n - 1)
        elif op.kind == capi.OP_MCU:
            total += 1 << (n - len(op.controls))
        elif op.kind == capi.OP_SWAP:
            total += 1 << (n - 1)
    return total


def grover(n: int = 33, iterations: int = 8, seed: int = 33):
    """Config 3: Grover search for a marked n-bit string m = nextLong() & (2^n - 1) of SplittableRandom(seed).
    Oracle: X on the zero bits of m, Z on qubit n-1 controlled by all others, X again.  Diffusion:
    H^n X^n (multi-controlled Z) X^n H^n.  Returns (circuit, marked)."""
    marked = int(SplittableRandom(seed).next_longs(1)[0]) & ((1 << n) - 1)
    controls = tuple(range(n - 1))
    c = QuantumCircuit(n)
    for q in range(n):
        c = c.append(Hadamard(), q)
    for _ in range(iterations):
        zeros = [q for q in range(n) if not (marked >> q) & 1]
        for q in zeros:
            c = c.append(PauliX(), q)
        c = c.append_multi_control(PauliZ(), n - 1, *controls)
        for q in zeros:
            c = c.append(PauliX(), q)
        for q in range(n):
            c = c.append(Hadamard(), q)
        for q in range(n):
            c = c.append(PauliX(), q)
        c = c.append_multi_control(PauliZ(), n - 1, *controls)
        for q in range(n):
            c = c.append(PauliX(), q)
        for q in range(n):
            c = c.append(Hadamard(), q)
    return c, marked


def grover_marked_probability(n: int, iterations: int) -> float:
    """Closed form sin^2((2k+1) asin(2^{-n/2})) of the standard Grover iteration."""
    th = math.asin(2.0 ** (-n / 2.0))
    return math.sin((2 * iterations + 1) * th) ** 2


def ghz(n: int = 20) -> QuantumCircuit:
    c = QuantumCircuit(n).append(Hadamard(), 0)
    for q in range(n - 1):
        c = c.append_cnot(q, q + 1)
    return c


def qaoa_ring(n: int, p: int, gammas, betas, weights) -> QuantumCircuit:
    """QAOA MaxCut on a ring: H on all, then p layers of zzPhase(gamma*w_e) on ring edges (i, i+1 mod n) and
    rx(2*beta) on every qubit (AlgorithmSuiteSpecs.zzPhase / rx).  Parameter indices: gamma_l = 2l,
    beta_l = 2l+1."""
    c = QuantumCircuit(n)
    for q in range(n):
        c = c.append(Hadamard(), q)
    for layer in range(p):
        g, b = float(gammas[layer]), float(betas[layer])
        for i in range(n):
            j = (i + 1) % n
            c = c.append_diagonal_two_qubit(zz_phase(g * float(weights[i])), i, j, param=2 * layer,
                                            coeff=float(weights[i]))
        for q in range(n):
            c = c.append(rx_suite(2.0 * b, param=2 * layer + 1, coeff=1.0), q)
    return c


def maxcut_observable(n: int, weights):
    """<C> = sum_e w_e (1 - <Z_i Z_j>)/2: returns (const, xmask, zmask, coeff) with C = const + sum coeff*P."""
    xm = np.zeros(n, np.uint64)
    zm = np.array([(1 << i) | (1 << ((i + 1) % n)) for i in range(n)], dtype=np.uint64)
    coeff = -0.5 * np.asarray(weights, dtype=np.float64)
    return 0.5 * float(np.sum(weights)), xm, zm, coeff


def hea(n: int, layers: int, thetas) -> QuantumCircuit:
    """Hardware-efficient ansatz: per layer Ry(theta) then Rz(theta') on every qubit, then a CNOT ring
    (i, i+1 mod n).  Parameter order: layer, qubit, (ry, rz)."""
    c = QuantumCircuit(n)
    k = 0
    for _ in range(layers):
        for q in range(n):
            c = c.append(rotation_y(float(thetas[k]), param=k), q)
            c = c.append(rotation_z(float(thetas[k + 1]), param=k + 1), q)
            k += 2
        for q in range(n):
            c = c.append_cnot(q, (q + 1) % n)
    return c


def hea_observable(n: int, a, b):
    """H = sum_i a_i Z_i Z_{i+1 mod n} + sum_i b_i X_i."""
    xm = [0] * n + [1 << i for i in range(n)]
    zm = [(1 << i) | (1 << ((i + 1) % n)) for i in range(n)] + [0] * n
    coeff = list(a) + list(b)
    return np.array(xm, np.uint64), np.array(zm, np.uint64), np.array(coeff, np.float64)


def with_params(circuit_builder, thetas):
    return circuit_builder(thetas)


def qaoa_batches(instances: int = 4096, p: int = 4, seed: int = 4):
    """Config 4's QAOA instances grouped by size: sizes n ~ U{12..22} (nextLong() % 11 + 12 per instance), then
    per size in increasing order gamma, beta ~ U[0, pi) (B x p each) and ring weights ~ U[0.5, 1.5) (B x n), all
    from one SplittableRandom(seed).  Returns [(n, instance indices, ops, per-instance matrices, weights)]."""
    from .circuit import lower

    rng = SplittableRandom(seed)
    sizes = (rng.next_longs(instances) % np.uint64(11)).astype(np.int64) + 12
    out = []
    for n in range(12, 23):
        idx = np.nonzero(sizes == n)[0]
        if len(idx) == 0:
            continue
        B = len(idx)
        g = rng.next_doubles(B * p).reshape(B, p) * math.pi
        b = rng.next_doubles(B * p).reshape(B, p) * math.pi
        w = 0.5 + rng.next_doubles(B * n).reshape(B, n)
        ops, _ = lower(qaoa_ring(n, p, g[0], b[0], w[0]))
        mats = np.stack([lower(qaoa_ring(n, p, g[k], b[k], w[k]))[1] for k in range(B)])
        out.append((n, idx, ops, mats, w))
    return out


def ring_zz_masks(n: int) -> np.ndarray:
    """Z-masks of the ring edges (i, i+1 mod n)."""
    return np.array([(1 << i) | (1 << ((i + 1) % n)) for i in range(n)], dtype=np.uint64)


def ghz_trajectory_uniforms(trajectories: int = 1024, n: int = 20, first_seed: int = 1000):
    """Config 4's noisy GHZ-n: the circuit's ops, matrices, the noise flag per op (depolarising after every
    CNOT, on control then target) and, per trajectory t, the nextDouble() stream of SplittableRandom(first_seed
    + t), one value per scheduled channel."""
    from .circuit import lower

    ops, mats = lower(ghz(n))
    noise_after = (ops["kind"] == capi.OP_CNOT).astype(np.uint8)
    draws = 2 * int(noise_after.sum())
    u = np.stack([SplittableRandom(first_seed + t).next_doubles(draws) for t in range(trajectories)])
    return ops, mats, noise_after, u

