"""Benchmark of the statevector hot path on one MI355X.

Default workload (one step): a 30-qubit complex128 state reset to |0...0>, the random layered circuit of
depth 20 (per layer Ry Rz on every qubit, then CNOT bricks; 1490 ops), the probability-vector checksum and
8192 shots over all 30 qubits.  Metric: amplitude updates per second, crediting each op of the unfused
circuit with the amplitudes it must change (1q / diagonal: 2^n, CNOT: 2^(n-1)).

Also reported: the sweep kernel's achieved HBM bandwidth (HIP events on the engine's stream, bytes =
32 B x 2^n per sweep) against the MI355X peak, and the CPU oracle (plain C port of the reference
StatevectorOps, 1 thread) timed on a bounded sample of the same workload.

  python bench.py [--steps K] [--warmup W] [--workload c2|c1|c3|c5] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from statecraft_amd import capi, workloads  # noqa: E402
from statecraft_amd.circuit import lower  # noqa: E402
from statecraft_amd.rng import SplittableRandom  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "30q amp-updates/s & %HBM peak; adjoint grads/s vs param-shift, 1 GPU"
MODES = {"direct": capi.EXEC_DIRECT, "fused": capi.EXEC_FUSED, "sweep": capi.EXEC_SWEEP}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n, ops, mats, seconds: float):
    """The oracle (C port of StatevectorOps, 1 thread) on the first ops of the same circuit until `seconds`."""
    import oracle

    st = np.zeros(1 << n, dtype=np.complex128)
    st[0] = 1.0
    credit = 0
    done = 0
    t0 = time.perf_counter()
    for i in range(len(ops)):
        oracle.run_ops_inplace(st, n, ops[i:i + 1], mats)
        k = int(ops[i]["kind"])
        credit += (1 << (n - 1)) if k == capi.OP_CNOT else (1 << n)
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    del st
    return credit / dt, done, dt


def bench_c2(args):
    n, depth = args.qubits, args.depth
    circ = workloads.random_layered(n, depth, args.seed)
    ops, mats = lower(circ)
    credit = workloads.amp_update_credit(circ)
    ctx = capi.Context(0, MODES[args.mode])
    ctx.set_option(capi.OPT_LOW_QUBITS, args.low_qubits)
    ctx.set_option(capi.OPT_REMAP, 0 if args.no_remap else 1)
    free, total, cus = ctx.device_info()
    st = capi.State(ctx, n)
    measured = list(range(n))
    u = SplittableRandom(args.seed).next_doubles(args.shots)

    def step():
        st.set_basis(0)
        st.apply(ops, mats)
        cs = st.checksum()
        out, gap = st.sample(measured, u)
        return cs, out, gap

    for i in range(args.warmup):
        step()
        log(f"warmup step {i + 1}/{args.warmup} done")
    plan = ctx.plan_stats()
    ctx.synchronize()
    ctx.profile_reset()
    ctx.set_option(capi.OPT_PROFILE, 1)
    t0 = time.perf_counter()
    for i in range(args.steps):
        cs, out, gap = step()
        log(f"timed step {i + 1}/{args.steps} done")
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.set_option(capi.OPT_PROFILE, 0)
    launches, ms, nbytes = ctx.profile_read(capi.KFAM_SWEEP if args.mode == "sweep" else
                                            capi.KFAM_FUSED if args.mode == "fused" else capi.KFAM_GATE)
    g_launches, g_ms, _ = ctx.profile_read(capi.KFAM_GATE)
    m_launches, m_ms, _ = ctx.profile_read(capi.KFAM_MEASURE)
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = credit * args.steps / elapsed
    avg_ms = ms / max(launches, 1)
    bytes_per_launch = nbytes / max(launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    sweeps = plan["sweeps"]
    circuit_ms = ms / args.steps
    # measured device-copy bandwidth (plain copy kernel, same byte count per launch as one sweep)
    copy_gbs = None
    if not args.no_copy_reference:
        dst = capi.State(ctx, n)
        ctx.set_option(capi.OPT_PROFILE, 1)
        for _ in range(3):
            dst.copy_from(st)
        ctx.set_option(capi.OPT_PROFILE, 0)
        c_n, c_ms, c_bytes = ctx.profile_read(capi.KFAM_COPY)
        copy_gbs = (c_bytes / c_n) / (c_ms / c_n * 1e-3) / 1e9 if c_n else None
        dst.close()
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"random-layered-{n}q-depth{depth}",
            "qubits": n, "depth": depth, "ops": int(len(ops)), "shots": args.shots, "exec_mode": args.mode,
            "sweeps_per_step": sweeps, "kernel_launches_per_step": plan["launches"],
            "fused_16x16_blocks": plan["fused_blocks"], "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": {"sweep": "k_sweep", "fused": "k_dense", "direct": "k_single"}[args.mode],
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
            "launches": launches,
            "copy_kernel_gbs": copy_gbs,
            "frac_of_copy": (achieved / copy_gbs) if copy_gbs else None,
            "fp64_ops_per_amp_per_step": capi.plan_sweeps_host(n, ops, mats, args.low_qubits, None,
                                                               not args.no_remap)[0]["fp64_ops_per_amp"]
            if args.mode == "sweep" else None,
        },
        "amp_updates_per_step": credit,
        "gate_ms_per_step": circuit_ms,
        "measure_ms_per_step": m_ms / args.steps,
        "checksum": list(cs),
        "sample_boundary_gap": gap,
        "device": {"hbm_total_gib": total / 2**30, "cus": cus},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread) ...")
        rate, done, dt = cpu_baseline(n, ops, mats, args.cpu_seconds)
        result["cpu_baseline"] = {
            "value": rate, "unit": "amp-updates/s", "cores": 1, "kind": "port",
            "sample": f"first {done} ops of the same {n}-qubit circuit in {dt:.1f} s (C port of StatevectorOps, "
                      f"scalar, 1 thread; the Java reference needs a JDK, absent on this box)",
        }
    st.close()
    ctx.close()
    return result


def bench_c1(args):
    n = 20
    circ, _ = workloads.ry_qft(n, 42)
    ops, mats = lower(circ)
    credit = workloads.amp_update_credit(circ)
    ctx = capi.Context(0, MODES[args.mode])
    st = capi.State(ctx, n)
    u = SplittableRandom(42).next_doubles(4096)
    amps = np.empty(1 << n, dtype=np.complex128)

    def step():
        st.set_basis(0)
        st.apply(ops, mats)
        st.sample(list(range(n)), u)
        return st.download()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    el = time.perf_counter() - t0
    return {"metric": "c1 ry+qft 20q amp-updates/s (incl. 4096 shots + amplitude download)",
            "value": credit * args.steps / el, "unit": "amp-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "ry-qft-20q", "plan": ctx.plan_stats()}}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--shots", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--mode", default="sweep", choices=list(MODES))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-copy-reference", action="store_true")
    ap.add_argument("--low-qubits", type=int, default=4)
    ap.add_argument("--no-remap", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is designed, not built",
              file=sys.stderr)
        return 2
    res = bench_c2(args) if args.workload == "c2" else bench_c1(args)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
