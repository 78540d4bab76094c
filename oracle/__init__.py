"""CPU oracle for the statevector hot path -- TEST INFRASTRUCTURE ONLY.

A plain-C restatement (sv_oracle.c) of StateCraft's StatevectorEngine /
StatevectorOps / SplittableRandom, loaded with ctypes.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only
as the checker or the timed CPU baseline; the product never calls it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libsv_oracle.so"

_lib = None
DP = ctypes.POINTER(ctypes.c_double)
I32P = ctypes.POINTER(ctypes.c_int32)
I64P = ctypes.POINTER(ctypes.c_int64)
U64P = ctypes.POINTER(ctypes.c_uint64)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = ctypes.CDLL(str(LIB))
        c, i, u64, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double
        sig = {
            "orc_sr_next_long": (u64, [U64P]),
            "orc_sr_next_double": (d, [U64P]),
            "orc_sr_next_double_bounded": (d, [U64P, d]),
            "orc_sr_doubles": (None, [u64, DP, i64]),
            "orc_apply_single": (None, [DP, i, i, DP]),
            "orc_apply_cnot": (None, [DP, i, i, i]),
            "orc_apply_two": (None, [DP, i, i, i, DP]),
            "orc_apply_diag2": (None, [DP, i, i, i, DP]),
            "orc_apply_swap": (None, [DP, i, i, i]),
            "orc_apply_mcu": (None, [DP, i, i, u64, DP]),
            "orc_run_ops": (i, [DP, i, c, i64, DP]),
            "orc_probabilities": (None, [DP, i, I32P, i, DP]),
            "orc_normalize": (i, [DP, u64]),
            "orc_cumulative": (None, [DP, DP, u64]),
            "orc_sample_outcome": (i64, [DP, i64, d]),
            "orc_sample": (i, [DP, i, I32P, i, DP, i64, I64P]),
            "orc_norm2sq": (d, [DP, u64]),
            "orc_renormalize": (i, [DP, i, DP]),
            "orc_kraus_weight": (d, [DP, i, i, DP]),
            "orc_sample_weighted_index": (i, [DP, i, d]),
            "orc_apply_channel_1q": (i, [DP, i, i, DP, i, U64P, ctypes.POINTER(ctypes.c_int)]),
            "orc_run_noisy": (i, [DP, i, c, i64, DP, I64P, I32P, I32P, DP, I32P, I32P, u64]),
            "orc_expect_pauli_term": (d, [DP, i, u64, u64]),
            "orc_expect_pauli": (d, [DP, i, U64P, U64P, DP, i]),
            "orc_checksum_weight": (d, [u64]),
            "orc_probability_checksum": (None, [DP, i, DP]),
            "orc_run_ops_parallel": (i, [DP, i, c, i64, DP, i, i64]),
            "orc_apply_single_parallel": (None, [DP, i, i, DP, i, i64]),
            "orc_gate_loop": (None, [DP, i, I32P, DP, i, i, i, i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _dp(a):
    return a.ctypes.data_as(DP)


def _f64(state):
    return np.ascontiguousarray(state, dtype=np.complex128).view(np.float64)


def zero_state(n: int) -> np.ndarray:
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1.0
    return s


def run_ops(n: int, ops, mats, init=None) -> np.ndarray:
    """StatevectorEngine op loop on a copy of `init` (|0..0> by default)."""
    st = zero_state(n) if init is None else np.array(init, dtype=np.complex128, copy=True)
    ops = np.ascontiguousarray(ops)
    m = np.ascontiguousarray(mats if len(mats) else np.zeros(1), dtype=np.complex128)
    rc = lib().orc_run_ops(_dp(st.view(np.float64)), n, ops.ctypes.data_as(ctypes.c_void_p), len(ops),
                           _dp(m.view(np.float64)))
    if rc != 0:
        raise RuntimeError(f"oracle run_ops failed with status {rc}")
    return st


def run_ops_inplace(state: np.ndarray, n: int, ops, mats) -> None:
    """Applies ops to `state` (complex128, contiguous) without copying it."""
    ops = np.ascontiguousarray(ops)
    m = np.ascontiguousarray(mats if len(mats) else np.zeros(1), dtype=np.complex128)
    rc = lib().orc_run_ops(_dp(state.view(np.float64)), n, ops.ctypes.data_as(ctypes.c_void_p), len(ops),
                           _dp(m.view(np.float64)))
    if rc != 0:
        raise RuntimeError(f"oracle run_ops failed with status {rc}")


def probabilities(state, measured) -> np.ndarray:
    n = int(np.log2(len(state)))
    meas = np.ascontiguousarray(measured, np.int32)
    out = np.empty(1 << len(meas))
    s = _f64(state)
    lib().orc_probabilities(_dp(s), n, meas.ctypes.data_as(I32P), len(meas), _dp(out))
    return out


def sample(state, measured, uniforms) -> np.ndarray:
    n = int(np.log2(len(state)))
    meas = np.ascontiguousarray(measured, np.int32)
    u = np.ascontiguousarray(uniforms, np.float64)
    out = np.empty(len(u), np.int64)
    s = _f64(state)
    rc = lib().orc_sample(_dp(s), n, meas.ctypes.data_as(I32P), len(meas), _dp(u), len(u), out.ctypes.data_as(I64P))
    if rc != 0:
        raise RuntimeError(f"oracle sample failed with status {rc}")
    return out


def cumulative(p) -> np.ndarray:
    p = np.ascontiguousarray(p, np.float64)
    out = np.empty_like(p)
    lib().orc_cumulative(_dp(p), _dp(out), len(p))
    return out


def sample_outcome(cum, u: float) -> int:
    cum = np.ascontiguousarray(cum, np.float64)
    return int(lib().orc_sample_outcome(_dp(cum), len(cum), float(u)))


def sr_doubles(seed: int, count: int) -> np.ndarray:
    out = np.empty(count)
    lib().orc_sr_doubles(ctypes.c_uint64(seed & ((1 << 64) - 1)), _dp(out), count)
    return out


def sr_longs(seed: int, count: int) -> list:
    s = ctypes.c_uint64(seed & ((1 << 64) - 1))
    return [int(lib().orc_sr_next_long(ctypes.byref(s))) for _ in range(count)]


def norm2sq(state) -> float:
    s = _f64(state)
    return float(lib().orc_norm2sq(_dp(s), len(s) // 2))


def kraus_weight(state, target: int, k) -> float:
    n = int(np.log2(len(state)))
    kk = np.ascontiguousarray(k, np.complex128).reshape(4)
    return float(lib().orc_kraus_weight(_dp(_f64(state)), n, target, _dp(kk.view(np.float64))))


def apply_channel_1q(state, target: int, kraus, seed_holder: list):
    """Mutates `state` in place; seed_holder[0] is the SplittableRandom state (advanced)."""
    n = int(np.log2(len(state)))
    k = np.ascontiguousarray(kraus, np.complex128).reshape(-1, 4)
    s = ctypes.c_uint64(seed_holder[0])
    chosen = ctypes.c_int()
    rc = lib().orc_apply_channel_1q(_dp(state.view(np.float64)), n, target, _dp(k.reshape(-1).view(np.float64)),
                                    len(k), ctypes.byref(s), ctypes.byref(chosen))
    seed_holder[0] = s.value
    if rc != 0:
        raise RuntimeError(f"oracle channel failed with status {rc}")
    return chosen.value


def expect_pauli(state, xmask, zmask, coeff) -> float:
    n = int(np.log2(len(state)))
    xm = np.ascontiguousarray(xmask, np.uint64)
    zm = np.ascontiguousarray(zmask, np.uint64)
    c = np.ascontiguousarray(coeff, np.float64)
    return float(lib().orc_expect_pauli(_dp(_f64(state)), n, xm.ctypes.data_as(U64P), zm.ctypes.data_as(U64P), _dp(c),
                                        len(c)))


def checksum(state):
    n = int(np.log2(len(state)))
    out = np.empty(2)
    lib().orc_probability_checksum(_dp(_f64(state)), n, _dp(out))
    return float(out[0]), float(out[1])


def run_ops_parallel(state: np.ndarray, n: int, ops, mats, threads: int, min_work: int = 4096) -> None:
    """Single-qubit and CNOT ops on `threads` threads (StatevectorParallelExecutor's split), in place."""
    ops = np.ascontiguousarray(ops)
    m = np.ascontiguousarray(mats if len(mats) else np.zeros(1), dtype=np.complex128)
    rc = lib().orc_run_ops_parallel(_dp(state.view(np.float64)), n, ops.ctypes.data_as(ctypes.c_void_p), len(ops),
                                    _dp(m.view(np.float64)), int(threads), int(min_work))
    if rc != 0:
        raise RuntimeError(f"oracle run_ops_parallel failed with status {rc}")


def apply_single_parallel(state: np.ndarray, n: int, target: int, g, threads: int, min_work: int = 4096) -> None:
    gg = np.ascontiguousarray(g, np.complex128).reshape(4)
    lib().orc_apply_single_parallel(_dp(state.view(np.float64)), n, int(target), _dp(gg.view(np.float64)),
                                    int(threads), int(min_work))


def gate_loop(state: np.ndarray, n: int, targets, gates, iterations: int, threads: int, min_work: int) -> None:
    """`iterations` passes over single-qubit gates (targets, 4 complex each) on `threads` threads, in place."""
    t = np.ascontiguousarray(targets, np.int32)
    g = np.ascontiguousarray(gates, np.complex128).reshape(-1, 4)
    lib().orc_gate_loop(_dp(state.view(np.float64)), n, t.ctypes.data_as(I32P), _dp(g.view(np.float64)), len(t),
                        int(iterations), int(threads), int(min_work))

