"""ctypes binding of libstatecraft_hip.so (include/statecraft_hip.h).

The shared library is the product; this module only marshals numpy arrays
into the C-ABI and maps status codes onto the reference's exception classes
(StatevectorEngine throws IllegalArgumentException, UnsupportedOperationException
and IllegalStateException; see engines/.../statevector/StatevectorEngine.java).
There is no fallback: if the library is missing, loading fails loudly.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libstatecraft_hip.so"

# ---- constants (statecraft_hip.h) -------------------------------------------
SC_OK, SC_E_NULL, SC_E_ARG, SC_E_UNSUPPORTED, SC_E_STATE, SC_E_OOM, SC_E_HIP = range(7)
OP_SINGLE, OP_CNOT, OP_TWO, OP_DIAG2, OP_SWAP, OP_MCU, OP_MEASURE = range(1, 8)
GEN_NONE, GEN_X, GEN_Y, GEN_Z, GEN_ZZ = range(5)
(OPT_EXEC_MODE, OPT_MAX_FUSED, OPT_PROFILE, OPT_LOW_QUBITS, OPT_TEST_DRIFT, OPT_REMAP,
 OPT_SWEEP_BLOCKS_PER_CU, OPT_JIT, OPT_SWEEP_PASS_CAP, OPT_CHUNK_BITS) = range(1, 11)
OPT_DIRECT_IO = 11
OPT_MFMA = 12
OPT_PASS_SCHED = 13
OPT_PERMLANE = 14
OPT_FUSED_TERMS = 15
OPT_MFMA_TILE = 16
OPT_MFMA_PICK4 = 17
OPT_MFMA_PICK5 = 18
EXEC_DIRECT, EXEC_FUSED, EXEC_SWEEP = range(3)
KFAM_GATE, KFAM_FUSED, KFAM_SWEEP, KFAM_MEASURE, KFAM_NOISE, KFAM_ADJOINT, KFAM_COPY = range(7)

OP_DTYPE = np.dtype(
    [
        ("kind", np.int32),
        ("q0", np.int32),
        ("q1", np.int32),
        ("mat", np.int32),
        ("controls", np.uint64),
        ("param", np.int32),
        ("gen", np.int32),
        ("coeff", np.float64),
    ],
    align=True,
)
assert OP_DTYPE.itemsize == 40


class NullPointerException(TypeError):
    pass


class IllegalArgumentException(ValueError):
    pass


class UnsupportedOperationException(NotImplementedError):
    pass


class IllegalStateException(RuntimeError):
    pass


class HipError(RuntimeError):
    pass


_EXC = {
    SC_E_NULL: NullPointerException,
    SC_E_ARG: IllegalArgumentException,
    SC_E_UNSUPPORTED: UnsupportedOperationException,
    SC_E_STATE: IllegalStateException,
    SC_E_OOM: UnsupportedOperationException,
    SC_E_HIP: HipError,
}

EXPORTED = [
    "sc_last_error", "sc_version", "sc_ctx_create", "sc_ctx_destroy", "sc_ctx_set_option", "sc_ctx_get_option",
    "sc_ctx_synchronize", "sc_ctx_device_info", "sc_state_create", "sc_state_destroy", "sc_state_qubits",
    "sc_state_set_basis", "sc_state_upload", "sc_state_download", "sc_state_copy", "sc_apply_ops",
    "sc_apply_ops_batched", "sc_probabilities", "sc_sample", "sc_norm_sq", "sc_probability_checksum",
    "sc_kraus_weights", "sc_apply_kraus", "sc_run_noisy", "sc_run_depolarizing_trajectories",
    "sc_run_depolarizing_trajectories_z",
    "sc_expectation_pauli",
    "sc_expectation_pauli_terms",
    "sc_adjoint_gradient", "sc_adjoint_stats", "sc_profile_reset", "sc_profile_read", "sc_last_plan_stats",
    "sc_plan_sweeps_host", "sc_state_read_range", "sc_jit_compile_host", "sc_adjoint_plan_host",
    "sc_last_sample_stats", "sc_cumulative_host", "sc_sample_host", "sc_state_max_deviation",
    "sc_expectation_pauli_terms_noalloc", "sc_plan_pass_stats_host", "sc_run_jobs", "sc_last_plan_geometry",
    "sc_plan_geometry_host", "sc_plan_product_host",
]

_lib = None

P = ctypes.c_void_p
DP = ctypes.POINTER(ctypes.c_double)
I32P = ctypes.POINTER(ctypes.c_int32)
I64P = ctypes.POINTER(ctypes.c_int64)
U64P = ctypes.POINTER(ctypes.c_uint64)
U8P = ctypes.POINTER(ctypes.c_uint8)
u64 = ctypes.c_uint64
i32 = ctypes.c_int
i64 = ctypes.c_int64


def _sig(lib):
    s = {
        "sc_last_error": (ctypes.c_char_p, []),
        "sc_version": (i32, []),
        "sc_ctx_create": (i32, [i32, ctypes.POINTER(P)]),
        "sc_ctx_destroy": (i32, [P]),
        "sc_ctx_set_option": (i32, [P, i32, i64]),
        "sc_ctx_get_option": (i32, [P, i32, I64P]),
        "sc_ctx_synchronize": (i32, [P]),
        "sc_ctx_device_info": (i32, [P, U64P, U64P, ctypes.POINTER(i32)]),
        "sc_state_create": (i32, [P, i32, i32, ctypes.POINTER(P)]),
        "sc_state_destroy": (i32, [P]),
        "sc_state_qubits": (i32, [P, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "sc_state_set_basis": (i32, [P, u64]),
        "sc_state_upload": (i32, [P, DP, u64]),
        "sc_state_download": (i32, [P, DP, u64]),
        "sc_state_copy": (i32, [P, P]),
        "sc_state_read_range": (i32, [P, i32, u64, u64, DP]),
        "sc_apply_ops": (i32, [P, P, u64, DP, u64]),
        "sc_apply_ops_batched": (i32, [P, P, u64, DP, u64]),
        "sc_probabilities": (i32, [P, i32, I32P, i32, DP]),
        "sc_sample": (i32, [P, i32, I32P, i32, DP, i64, I64P, DP]),
        "sc_norm_sq": (i32, [P, i32, DP]),
        "sc_probability_checksum": (i32, [P, i32, DP]),
        "sc_kraus_weights": (i32, [P, i32, i32, DP, i32, DP]),
        "sc_apply_kraus": (i32, [P, i32, i32, DP, DP]),
        "sc_run_depolarizing_trajectories": (i32, [P, P, u64, DP, u64, U8P, ctypes.c_double, DP, u64]),
        "sc_run_depolarizing_trajectories_z": (i32, [P, P, u64, DP, u64, U8P, ctypes.c_double, DP, u64, U64P, i32,
                                                     DP]),
        "sc_run_noisy": (i32, [P, P, u64, DP, u64, P, u64, DP, u64, DP, P]),
        "sc_expectation_pauli": (i32, [P, U64P, U64P, DP, i32, DP]),
        "sc_expectation_pauli_terms": (i32, [P, U64P, U64P, i32, DP]),
        "sc_adjoint_gradient": (i32, [P, P, u64, DP, u64, i32, U64P, U64P, DP, i32, i32, i32, DP, DP]),
        "sc_adjoint_stats": (i32, [P, U64P, DP]),
        "sc_profile_reset": (i32, [P]),
        "sc_profile_read": (i32, [P, i32, U64P, DP, DP]),
        "sc_last_plan_stats": (i32, [P, U64P, U64P, U64P]),
        "sc_plan_sweeps_host": (i32, [i32, P, u64, DP, u64, i32, i32, i32, i32, i32, U64P, U64P, U64P, DP, DP]),
        "sc_jit_compile_host": (i32, [i32, P, u64, DP, u64, i32, i32, i32, i32, i32, U64P, DP, U64P,
                                      ctypes.c_char_p]),
        "sc_adjoint_plan_host": (i32, [i32, P, u64, DP, u64, i32, i32, DP, DP, DP]),
        "sc_last_sample_stats": (i32, [P, U64P]),
        "sc_state_max_deviation": (i32, [P, i32, u64, DP, DP, DP]),
        "sc_expectation_pauli_terms_noalloc": (i32, [P, U64P, U64P, i32, DP]),
        "sc_run_jobs": (i32, [P, P, i32, ctypes.POINTER(i32)]),
        "sc_plan_pass_stats_host": (i32, [i32, P, u64, DP, u64, i32, i32, i32, i32, i32, I32P, I32P, DP, I32P,
                                          ctypes.POINTER(i32)]),
        "sc_last_plan_geometry": (i32, [P, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32), DP]),
        "sc_plan_geometry_host": (i32, [i32, P, u64, DP, u64, i32, i32, i32, i32, ctypes.POINTER(i32),
                                        ctypes.POINTER(i32), ctypes.POINTER(i32), DP]),
        "sc_plan_product_host": (i32, [i32, P, u64, DP, u64, u64, i32, i32, ctypes.POINTER(i32), DP]),
        "sc_cumulative_host": (i32, [DP, u64, DP, U64P]),
        "sc_sample_host": (i32, [i32, DP, I32P, i32, DP, i64, I64P, DP, U64P]),
    }
    for name, (res, args) in s.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def lib():
    """Load the HIP engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("STATECRAFT_HIP_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(
                f"{path} is missing: build the HIP engine with `python -m statecraft_amd.build` "
                "(there is no CPU fallback)"
            )
        l = ctypes.CDLL(str(path))
        _sig(l)
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc != SC_OK:
        msg = lib().sc_last_error().decode("utf-8", "replace")
        raise _EXC.get(rc, RuntimeError)(msg)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(DP)


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def complex_to_f64(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


def plan_sweeps_host(n: int, ops, mats, low_qubits: int = 4, state=None, remap: bool = True, jit: bool = False,
                     pass_cap: int = -1, chunk_bits: int = 12, permlane: bool = False):
    """Host-only sweep plan statistics; with `state` (complex 2^n) also executes the plan on the host.  jit: the
    runtime-compiled passes' plan; permlane (with jit): with in-register stage transitions."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
    a, b, c, f = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
    out = None
    ptr = None
    if state is not None:
        out = np.array(state, dtype=np.complex128, copy=True)
        ptr = dptr(out.view(np.float64))
    check(lib().sc_plan_sweeps_host(int(n), ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2, int(low_qubits),
                                    1 if remap else 0, (2 if permlane else 1) if jit else 0, int(pass_cap),
                                    int(chunk_bits),
                                    ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(f), ptr))
    return {"passes": a.value, "stages": b.value, "fused_blocks": c.value, "fp64_ops_per_amp": f.value}, out


def plan_pass_stats_host(n: int, ops, mats, low_qubits: int = 4, remap: bool = False, jit: bool = True,
                         chunk_bits: int = 12):
    """Host-only: per sweep pass of a circuit's plan, (stages, barrier exchanges, FP64 operations per amplitude,
    direct-HBM flags)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
    cap = 4096
    st, br, fp, dr = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap), np.zeros(cap, np.int32)
    cnt = ctypes.c_int()
    check(lib().sc_plan_pass_stats_host(int(n), ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2,
                                        int(low_qubits), 1 if remap else 0, 1 if jit else 0, int(chunk_bits), cap,
                                        st.ctypes.data_as(I32P), br.ctypes.data_as(I32P), dptr(fp),
                                        dr.ctypes.data_as(I32P), ctypes.byref(cnt)))
    k = min(cnt.value, cap)
    return [(int(st[i]), int(br[i]), float(fp[i]), int(dr[i])) for i in range(k)]


def plan_geometry_host(n: int, ops, mats, low_qubits: int = -1, remap: bool = False, jit: int = 2,
                       chunk_bits: int = 0):
    """Host-only: the sweep geometry a context with these options picks for the circuit: (chunk bits, low
    qubits, geometries planned, predicted ns per amplitude)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
    cb, lq, nc, ns = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    check(lib().sc_plan_geometry_host(int(n), ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2,
                                      int(low_qubits), 1 if remap else 0, int(jit), int(chunk_bits), ctypes.byref(cb),
                                      ctypes.byref(lq), ctypes.byref(nc), ctypes.byref(ns)))
    return cb.value, lq.value, nc.value, ns.value


def plan_product_host(n: int, ops, mats, basis: int = 0, low_qubits: int = -1, chunk_bits: int = 0):
    """Host-only: the circuit run from |basis> with its leading single-qubit ops as the first pass's
    product-state input; returns (state, ops absorbed)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
    out = np.zeros(1 << n, dtype=np.complex128)
    k = ctypes.c_int()
    check(lib().sc_plan_product_host(int(n), ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2, int(basis),
                                     int(low_qubits), int(chunk_bits), ctypes.byref(k), dptr(out.view(np.float64))))
    return out, k.value


def adjoint_plan_host(n: int, ops, mats, n_params: int, psi, lam, low_qubits: int = 4):
    """Host-only reverse sweep of an adjoint gradient as sweep passes: returns (grads, psi, lambda) after it."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
    p = np.array(psi, dtype=np.complex128, copy=True)
    lm = np.array(lam, dtype=np.complex128, copy=True)
    g = np.zeros(max(n_params, 1), dtype=np.float64)
    check(lib().sc_adjoint_plan_host(int(n), ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2, int(n_params),
                                     int(low_qubits), dptr(p.view(np.float64)), dptr(lm.view(np.float64)), dptr(g)))
    return g[:n_params], p, lm


def jit_compile_host(n: int, ops, mats, low_qubits: int = 4, remap: bool = True, threads: int = 1,
                     dump_dir: str | None = None, variant: int = 1, chunk_bits: int = 12):
    """Host-only: compile the runtime-compiled pass kernels of a circuit's sweep plan (no device needed)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
    k, sec, size = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64()
    check(lib().sc_jit_compile_host(int(n), ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2, int(low_qubits),
                                    1 if remap else 0, int(variant), int(chunk_bits), int(threads), ctypes.byref(k),
                                    ctypes.byref(sec),
                                    ctypes.byref(size), dump_dir.encode() if dump_dir else None))
    return {"kernels": k.value, "seconds": sec.value, "code_bytes": size.value}


class NoiseEvent(ctypes.Structure):
    """sc_noise_event: after op `after_op`, `n_kraus` single-qubit Kraus operators (from complex offset `kraus`)
    act on `target`."""
    _fields_ = [("after_op", ctypes.c_uint64), ("target", ctypes.c_int32), ("n_kraus", ctypes.c_int32),
                ("kraus", ctypes.c_uint64)]


class NoiseFailure(ctypes.Structure):
    """sc_noise_failure: the first failing channel of sc_run_noisy (event -1: none)."""
    _fields_ = [("event", ctypes.c_int64), ("elem", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("branch", ctypes.c_int32), ("pad", ctypes.c_int32), ("weights", ctypes.c_double * 4),
                ("norm_sq", ctypes.c_double)]


class NoisyRunFailed(IllegalStateException):
    """sc_run_noisy's SC_E_STATE, with the failing channel's record."""

    def __init__(self, msg: str, failure: NoiseFailure):
        super().__init__(msg)
        self.event = int(failure.event)
        self.elem = int(failure.elem)
        self.kind = int(failure.kind)
        self.branch = int(failure.branch)
        self.weights = [float(x) for x in failure.weights]
        self.norm_sq = float(failure.norm_sq)


class _Job(ctypes.Structure):
    _fields_ = [("state", P), ("basis", ctypes.c_uint64), ("ops", P), ("n_ops", ctypes.c_uint64), ("mats", DP),
                ("mat_stride", ctypes.c_uint64), ("xmask", U64P), ("zmask", U64P), ("n_terms", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("out", DP)]


def run_jobs(ctx: "Context", jobs):
    """sc_run_jobs: each job (state, basis, ops, per-element matrices [batch, m], xmask, zmask) runs its circuit on
    every element from |basis> and yields the per-term expectations [batch, n_terms].  Returns (outs, whether
    one graph launch ran them)."""
    keep, arr = [], (_Job * len(jobs))()
    outs = []
    for k, (st, basis, ops, mats, xm, zm) in enumerate(jobs):
        o = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        m = np.ascontiguousarray(mats, dtype=np.complex128)
        assert m.ndim == 2 and m.shape[0] == st.batch
        mv = m.view(np.float64)
        x = np.ascontiguousarray(xm, np.uint64)
        z = np.ascontiguousarray(zm, np.uint64)
        out = np.empty((st.batch, len(z)))
        keep += [o, m, x, z]
        outs.append(out)
        arr[k] = _Job(st.h, int(basis), o.ctypes.data_as(P), len(o), dptr(mv), m.shape[1], x.ctypes.data_as(U64P),
                      z.ctypes.data_as(U64P), len(z), 0, dptr(out))
    used = ctypes.c_int()
    check(lib().sc_run_jobs(ctx.h, ctypes.cast(arr, P), len(jobs), ctypes.byref(used)))
    return outs, bool(used.value)


SAMPLE_STATS = ("fast_segments", "serial_segments", "zero_segments", "gather_rounds", "rescaled")


def cumulative_host(values):
    """Host-only: the exact segment-parallel running sums sc_sample uses (returns cum, stats)."""
    v = as_f64(values)
    cum = np.empty_like(v)
    st = np.zeros(5, dtype=np.uint64)
    check(lib().sc_cumulative_host(dptr(v), len(v), dptr(cum), st.ctypes.data_as(U64P)))
    return cum, dict(zip(SAMPLE_STATS[:4], (int(x) for x in st[:4])))


def sample_host(n: int, state, measured, uniforms):
    """Host-only: sc_sample's exact flow on a host state (returns outcomes, boundary gap, stats)."""
    a = complex_to_f64(state)
    meas = np.ascontiguousarray(measured, np.int32)
    u = as_f64(uniforms)
    out = np.empty(len(u), dtype=np.int64)
    gap = ctypes.c_double()
    st = np.zeros(5, dtype=np.uint64)
    check(lib().sc_sample_host(int(n), dptr(a), meas.ctypes.data_as(I32P), len(meas), dptr(u), len(u),
                               out.ctypes.data_as(I64P), ctypes.byref(gap), st.ctypes.data_as(U64P)))
    return out, gap.value, dict(zip(SAMPLE_STATS, (int(x) for x in st)))


class Context:
    """One device + one HIP stream (StatevectorEngine(StatevectorExecutionConfig) analogue)."""

    def __init__(self, device: int = 0, mode: int | None = None):
        h = P()
        check(lib().sc_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self._states = weakref.WeakSet()  # states to release before the context
        _live_contexts.add(self)
        if mode is not None:
            self.set_option(OPT_EXEC_MODE, mode)

    def set_option(self, opt: int, value: int) -> None:
        check(lib().sc_ctx_set_option(self.h, opt, int(value)))

    def get_option(self, opt: int) -> int:
        v = ctypes.c_int64()
        check(lib().sc_ctx_get_option(self.h, opt, ctypes.byref(v)))
        return v.value

    def synchronize(self) -> None:
        check(lib().sc_ctx_synchronize(self.h))

    def device_info(self):
        f, t, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        check(lib().sc_ctx_device_info(self.h, ctypes.byref(f), ctypes.byref(t), ctypes.byref(c)))
        return f.value, t.value, c.value

    def profile_reset(self) -> None:
        check(lib().sc_profile_reset(self.h))

    def profile_read(self, family: int):
        n, ms, by = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
        check(lib().sc_profile_read(self.h, family, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)))
        return n.value, ms.value, by.value

    def plan_stats(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().sc_last_plan_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"sweeps": a.value, "launches": b.value, "fused_blocks": c.value}

    def plan_geometry(self):
        """Geometry of the last sweep plan this context used."""
        cb, lq, nc, ns = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().sc_last_plan_geometry(self.h, ctypes.byref(cb), ctypes.byref(lq), ctypes.byref(nc),
                                          ctypes.byref(ns)))
        return {"chunk_bits": cb.value, "low_qubits": lq.value, "candidates": nc.value, "ns_per_amp": ns.value}

    def sample_stats(self):
        """Exact-sum statistics of the last sample() on this context."""
        st = np.zeros(5, dtype=np.uint64)
        check(lib().sc_last_sample_stats(self.h, st.ctypes.data_as(U64P)))
        return dict(zip(SAMPLE_STATS, (int(x) for x in st)))

    def adjoint_stats(self):
        a, b = ctypes.c_uint64(), ctypes.c_double()
        check(lib().sc_adjoint_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return {"reanchors": a.value, "max_drift": b.value}

    def close(self) -> None:
        """Release the context; its states that are still open are released first."""
        if getattr(self, "h", None):
            for s in list(self._states):
                s.close()
            lib().sc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_live_contexts: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def _release_all() -> None:
    # Device memory and streams are released while the HIP runtime is still up, whatever order the
    # interpreter would finalise module globals in.
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


class State:
    """A resident batch of 2^n complex128 states in HBM."""

    def __init__(self, ctx: Context, n: int, batch: int = 1):
        h = P()
        check(lib().sc_state_create(ctx.h, int(n), int(batch), ctypes.byref(h)))
        self.h, self.ctx, self.n, self.batch = h, ctx, int(n), int(batch)
        ctx._states.add(self)

    @property
    def size(self) -> int:
        return self.batch << self.n

    def set_basis(self, index: int) -> None:
        check(lib().sc_state_set_basis(self.h, int(index)))

    def upload(self, amps) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128).reshape(-1)
        check(lib().sc_state_upload(self.h, dptr(a.view(np.float64)), a.size))

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.complex128)
        check(lib().sc_state_download(self.h, dptr(out.view(np.float64)), out.size))
        return out if self.batch == 1 else out.reshape(self.batch, -1)

    def copy_from(self, other: "State") -> None:
        check(lib().sc_state_copy(self.h, other.h))

    def read(self, offset: int, count: int = 1, elem: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=np.complex128)
        check(lib().sc_state_read_range(self.h, elem, int(offset), int(count), dptr(out.view(np.float64))))
        return out

    def apply(self, ops: np.ndarray, mats) -> None:
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
        check(lib().sc_apply_ops(self.h, ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2))

    def apply_batched(self, ops: np.ndarray, mats_per_elem) -> None:
        """mats_per_elem: complex array [batch, n_mat]."""
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        m = np.ascontiguousarray(mats_per_elem, dtype=np.complex128)
        assert m.ndim == 2 and m.shape[0] == self.batch
        check(lib().sc_apply_ops_batched(self.h, ops.ctypes.data_as(P), len(ops), dptr(m.view(np.float64)),
                                         m.shape[1]))

    def probabilities(self, measured=None, elem: int = 0) -> np.ndarray:
        meas = np.arange(self.n, dtype=np.int32) if measured is None else np.ascontiguousarray(measured, np.int32)
        out = np.empty(1 << len(meas), dtype=np.float64)
        check(lib().sc_probabilities(self.h, elem, meas.ctypes.data_as(I32P), len(meas), dptr(out)))
        return out

    def sample(self, measured, uniforms, elem: int = 0):
        meas = np.ascontiguousarray(measured, np.int32)
        u = as_f64(uniforms)
        out = np.empty(len(u), dtype=np.int64)
        gap = ctypes.c_double()
        check(lib().sc_sample(self.h, elem, meas.ctypes.data_as(I32P), len(meas), dptr(u), len(u),
                              out.ctypes.data_as(I64P), ctypes.byref(gap)))
        return out, gap.value

    def max_deviation(self, index: int, a: complex, b: complex = 0.0, elem: int = 0) -> float:
        """max_i |psi_i - r_i| with r_index = a, r_i = b elsewhere, reduced on the device."""
        av = np.array([complex(a).real, complex(a).imag])
        bv = np.array([complex(b).real, complex(b).imag])
        v = ctypes.c_double()
        check(lib().sc_state_max_deviation(self.h, elem, int(index), dptr(av), dptr(bv), ctypes.byref(v)))
        return v.value

    def norm_sq(self, elem: int = 0) -> float:
        v = ctypes.c_double()
        check(lib().sc_norm_sq(self.h, elem, ctypes.byref(v)))
        return v.value

    def checksum(self, elem: int = 0):
        out = np.empty(2)
        check(lib().sc_probability_checksum(self.h, elem, dptr(out)))
        return float(out[0]), float(out[1])

    def kraus_weights(self, target: int, kraus, elem: int = 0) -> np.ndarray:
        k = complex_to_f64(np.asarray(kraus, np.complex128).reshape(-1, 4))
        w = np.empty(k.size // 8)
        check(lib().sc_kraus_weights(self.h, elem, target, dptr(k), len(w), dptr(w)))
        return w

    def apply_kraus(self, target: int, op, elem: int = 0) -> float:
        k = complex_to_f64(np.asarray(op, np.complex128).reshape(4))
        ns = ctypes.c_double()
        check(lib().sc_apply_kraus(self.h, elem, target, dptr(k), ctypes.byref(ns)))
        return ns.value

    def expectation(self, xmask, zmask, coeff) -> np.ndarray:
        xm = np.ascontiguousarray(xmask, np.uint64)
        zm = np.ascontiguousarray(zmask, np.uint64)
        c = as_f64(coeff)
        out = np.empty(self.batch)
        check(lib().sc_expectation_pauli(self.h, xm.ctypes.data_as(U64P), zm.ctypes.data_as(U64P), dptr(c), len(c),
                                         dptr(out)))
        return out

    def expectation_terms(self, xmask, zmask) -> np.ndarray:
        """Per-term expectations <P_t> of every batch element: array [batch, n_terms]."""
        xm = np.ascontiguousarray(xmask, np.uint64)
        zm = np.ascontiguousarray(zmask, np.uint64)
        assert xm.shape == zm.shape
        out = np.empty((self.batch, len(zm)))
        check(lib().sc_expectation_pauli_terms(self.h, xm.ctypes.data_as(U64P), zm.ctypes.data_as(U64P), len(zm),
                                               dptr(out)))
        return out

    def expectation_terms_noalloc(self, xmask, zmask) -> np.ndarray:
        """expectation_terms through the entry point that allocates no device memory (fixed 32 MiB scratch)."""
        xm = np.ascontiguousarray(xmask, np.uint64)
        zm = np.ascontiguousarray(zmask, np.uint64)
        assert xm.shape == zm.shape
        out = np.empty((self.batch, len(zm)))
        check(lib().sc_expectation_pauli_terms_noalloc(self.h, xm.ctypes.data_as(U64P), zm.ctypes.data_as(U64P),
                                                       len(zm), dptr(out)))
        return out

    def adjoint_gradient(self, ops, mats, n_params: int, xmask, zmask, coeff, checkpoint_every: int = 0,
                         per_element: bool = False):
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        m = np.ascontiguousarray(mats, dtype=np.complex128)
        n_mat = m.shape[-1] if m.ndim == 2 else m.size
        xm = np.ascontiguousarray(xmask, np.uint64)
        zm = np.ascontiguousarray(zmask, np.uint64)
        c = as_f64(coeff)
        grads = np.empty(self.batch * n_params)
        energies = np.empty(self.batch)
        mf = m.reshape(-1).view(np.float64) if m.size else np.zeros(2)
        check(lib().sc_adjoint_gradient(self.h, ops.ctypes.data_as(P), len(ops), dptr(mf), n_mat,
                                        1 if per_element else 0, xm.ctypes.data_as(U64P), zm.ctypes.data_as(U64P),
                                        dptr(c), len(c), n_params, checkpoint_every, dptr(grads), dptr(energies)))
        return grads.reshape(self.batch, n_params), energies

    def run_noisy(self, ops, mats, events, kraus, uniforms) -> None:
        """sc_run_noisy: events are (after_op, target, first Kraus operator index, operator count) with the
        operators [count][2][2] complex stacked in `kraus`; uniforms [batch, n_events] of nextDouble() values.
        A failing channel raises NoisyRunFailed."""
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
        k = complex_to_f64(np.concatenate([np.asarray(op, np.complex128).reshape(4) for op in kraus])
                           if len(kraus) else np.zeros(4, np.complex128))
        ev = (NoiseEvent * max(1, len(events)))()
        for i, (after, target, first, count) in enumerate(events):
            ev[i] = NoiseEvent(int(after), int(target), int(count), 4 * int(first))
        u = np.ascontiguousarray(uniforms, np.float64).reshape(self.batch, len(events))
        if not u.size:
            u = np.zeros(1)  # not read without events; the library wants a pointer
        fail = NoiseFailure()
        rc = lib().sc_run_noisy(self.h, ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2,
                                ctypes.cast(ev, P), len(events), dptr(k), k.size // 2, dptr(u), ctypes.byref(fail))
        if rc == SC_E_STATE and fail.event >= 0:
            raise NoisyRunFailed(lib().sc_last_error().decode("utf-8", "replace"), fail)
        check(rc)

    def run_depolarizing_trajectories(self, ops, mats, noise_after, p: float, uniforms) -> None:
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
        na = np.ascontiguousarray(noise_after, np.uint8)
        u = np.ascontiguousarray(uniforms, np.float64)
        assert u.ndim == 2 and u.shape[0] == self.batch
        check(lib().sc_run_depolarizing_trajectories(self.h, ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2,
                                                     na.ctypes.data_as(U8P), float(p), dptr(u), u.shape[1]))

    def run_depolarizing_trajectories_z(self, ops, mats, noise_after, p: float, uniforms, zmask) -> np.ndarray:
        """run_depolarizing_trajectories, then <Z^zmask[t]> of every trajectory: array [batch, n_terms]."""
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        m = complex_to_f64(mats if len(mats) else np.zeros(1, np.complex128))
        na = np.ascontiguousarray(noise_after, np.uint8)
        u = np.ascontiguousarray(uniforms, np.float64)
        zm = np.ascontiguousarray(zmask, np.uint64)
        assert u.ndim == 2 and u.shape[0] == self.batch
        out = np.empty((self.batch, len(zm)))
        check(lib().sc_run_depolarizing_trajectories_z(self.h, ops.ctypes.data_as(P), len(ops), dptr(m), m.size // 2,
                                                       na.ctypes.data_as(U8P), float(p), dptr(u), u.shape[1],
                                                       zm.ctypes.data_as(U64P), len(zm), dptr(out)))
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().sc_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
