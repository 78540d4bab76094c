"""Benchmark of the statevector hot path on one MI355X.

Default workload (one step): a 30-qubit complex128 state reset to |0...0>, the random layered circuit of
depth 20 (per layer Ry Rz on every qubit, then CNOT bricks; 1490 ops), the probability-vector checksum and
8192 shots over all 30 qubits.  Metric: amplitude updates per second, crediting each op of the unfused
circuit with the amplitudes it must change (1q / diagonal: 2^n, CNOT: 2^(n-1)).

Also reported: the sweep kernel's achieved HBM bandwidth (HIP events on the engine's stream, bytes =
32 B x 2^n per sweep) against the MI355X peak; the adjoint-gradient half of the metric (24-qubit HEA, 384
parameters, batch 64, against parameter shift); short runs of the other configurations (20-qubit Ry+QFT,
33-qubit Grover, batched QAOA + GHZ-20 trajectories); and the CPU oracle (plain C port of the reference
StatevectorOps split over the host's threads like StatevectorParallelExecutor) timed on a bounded sample of
the same circuit family at 29 qubits, the reference's dense limit.

  python bench.py [--steps K] [--warmup W] [--workload c1|c2|c3|c4|c5] [--no-cpu-baseline] [--no-adjoint]
                  [--no-other-configs]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

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


def make_context(args) -> capi.Context:
    """A context with the library's defaults (runtime-compiled passes, chunk and low qubits chosen per circuit
    structure), changed only by the options given on the command line."""
    ctx = capi.Context(0, MODES[args.mode])
    if args.low_qubits is not None:
        ctx.set_option(capi.OPT_LOW_QUBITS, args.low_qubits)
    if args.remap:
        ctx.set_option(capi.OPT_REMAP, 1)
    ctx.set_option(capi.OPT_SWEEP_BLOCKS_PER_CU, args.blocks_per_cu)
    if args.jit is not None:
        ctx.set_option(capi.OPT_JIT, args.jit)
    ctx.set_option(capi.OPT_SWEEP_PASS_CAP, args.pass_cap)
    ctx.set_option(capi.OPT_DIRECT_IO, 0 if args.no_direct_io else 1)
    if args.chunk_bits is not None:
        ctx.set_option(capi.OPT_CHUNK_BITS, args.chunk_bits)
    if args.mfma is not None:
        ctx.set_option(capi.OPT_MFMA, args.mfma)
    if args.mfma_tile is not None:
        ctx.set_option(capi.OPT_MFMA_TILE, args.mfma_tile)
    if args.max_fused is not None:
        ctx.set_option(capi.OPT_MAX_FUSED, args.max_fused)
    if args.pass_sched is not None:
        ctx.set_option(capi.OPT_PASS_SCHED, args.pass_sched)
    if args.permlane is not None:
        ctx.set_option(capi.OPT_PERMLANE, args.permlane)
    if args.fused_terms is not None:
        ctx.set_option(capi.OPT_FUSED_TERMS, args.fused_terms)
    return ctx


def context_options(ctx) -> dict:
    return {"jit": ctx.get_option(capi.OPT_JIT), "remap": bool(ctx.get_option(capi.OPT_REMAP)),
            "low_qubits_option": ctx.get_option(capi.OPT_LOW_QUBITS),
            "chunk_bits_option": ctx.get_option(capi.OPT_CHUNK_BITS),
            "pass_sched": ctx.get_option(capi.OPT_PASS_SCHED), "mfma": ctx.get_option(capi.OPT_MFMA),
            "permlane": ctx.get_option(capi.OPT_PERMLANE), "fused_terms": ctx.get_option(capi.OPT_FUSED_TERMS),
            "mfma_tile": ctx.get_option(capi.OPT_MFMA_TILE),
            "mfma_tile_autotuned": [ctx.get_option(capi.OPT_MFMA_PICK4), ctx.get_option(capi.OPT_MFMA_PICK5)]}


def host_threads() -> int:
    """Host threads for the CPU baseline: the process's CPU share (OMP_NUM_THREADS caps it when set)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(env))) if env and env.isdigit() else n


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _next_int_pow2(rng: SplittableRandom, bound: int) -> int:
    """SplittableRandom.nextInt(bound) for a power-of-two bound: the low bits of mix32(nextSeed()) (the JDK's
    published algorithm; not checked against a JDK here)."""
    m64 = (1 << 64) - 1
    rng.seed = (rng.seed + 0x9E3779B97F4A7C15) & m64
    z = rng.seed
    z = ((z ^ (z >> 33)) * 0x62A9D9ED799705F5) & m64
    r = ((((z ^ (z >> 28)) * 0xCB24D0A5C88C35B3) & m64) >> 32) & 0xFFFFFFFF
    return r & (bound - 1)


def cpu_baseline(depth: int, seed: int, seconds: float):
    """The reference's CPU path restated (oracle/sv_parallel.c: StatevectorParallelExecutor's leaf split on P
    threads, the scalar pair kernels), timed on this host: the same random layered family at 29 qubits, the
    reference's dense limit, op by op for about `seconds`; then the stock StatevectorKernelMicrobenchmark shape
    (16 qubits, 384 random X/Y/Z/H gates from SplittableRandom(0xBEEF), 64 iterations, serial and parallel)."""
    import oracle

    P = host_threads()
    n = 29
    ops, mats = lower(workloads.random_layered(n, depth, seed))
    st = np.zeros(1 << n, dtype=np.complex128)
    st[:] = 0.0  # touch every page before timing
    st[0] = 1.0
    credit, done = 0, 0
    t0 = time.perf_counter()
    for i in range(len(ops)):
        oracle.run_ops_parallel(st, n, ops[i:i + 1], mats, P, 4096)
        credit += (1 << (n - 1)) if int(ops[i]["kind"]) == capi.OP_CNOT else (1 << n)
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    del st
    # StatevectorKernelMicrobenchmark (engines/src/test/.../StatevectorKernelMicrobenchmark.java:27-107)
    from statecraft_amd.circuit import Hadamard, PauliX, PauliY, PauliZ

    rng = SplittableRandom(0xBEEF)
    cands = [PauliX(), PauliY(), PauliZ(), Hadamard()]
    gates = []
    for _ in range(384):
        g = cands[_next_int_pow2(rng, 4)]
        t = _next_int_pow2(rng, 16)
        gates.append((t, np.array([g.element(0, 0), g.element(0, 1), g.element(1, 0), g.element(1, 1)])))
    u = rng.next_doubles(2 << 16) - 0.5
    base = (u[0::2] + 1j * u[1::2]).astype(np.complex128)
    base *= 1.0 / math.sqrt(float(np.sum(base.real * base.real + base.imag * base.imag)))

    targets = np.array([t for t, _ in gates], np.int32)
    mats16 = np.stack([g for _, g in gates])

    def run(threads, iters):
        s = base.copy()
        t1 = time.perf_counter()
        oracle.gate_loop(s, 16, targets, mats16, iters, threads, 1024)
        return time.perf_counter() - t1

    run(1, 16)
    serial_s, par_s = run(1, 64), run(P, 64)
    micro_updates = (1 << 16) * 384 * 64
    return {
        "value": credit / dt, "unit": "amp-updates/s", "cores": P, "kind": "port",
        "sample": f"first {done} of the 1490 ops of the same random layered family at 29 qubits (the reference's "
                  f"dense limit; 8 GiB state) in {dt:.1f} s on {P} threads of {cpu_model()}: C restatement of "
                  f"StatevectorOps' pair kernels split like StatevectorParallelExecutor (the Java reference needs "
                  f"a JDK, absent on this host)",
        "microbenchmark": {
            "shape": "16 qubits, 384 random X/Y/Z/H gates x 64 iterations (StatevectorKernelMicrobenchmark)",
            "serial_ms": 1e3 * serial_s, "parallel_ms": 1e3 * par_s, "threads": P,
            "serial_amp_updates_per_s": micro_updates / serial_s,
            "parallel_amp_updates_per_s": micro_updates / par_s,
        },
    }


def pmc_traffic(fetch_csv: str, write_csv: str, copy_bytes: float):
    """HBM bytes per sweep launch from two rocprofv3 --pmc runs of this script (FETCH_SIZE, then WRITE_SIZE;
    KiB per dispatch), each scaled by what the same counter reads for k_copy, whose traffic is known
    (copy_bytes / 2 read and copy_bytes / 2 written per launch)."""
    import csv
    import re

    def per_dispatch(path, counter):
        acc = {}
        for row in csv.DictReader(open(path)):
            if row["Counter_Name"] != counter:
                continue
            k = row["Kernel_Name"].replace("sc::(anonymous namespace)::", "").replace("void ", "")
            k = re.sub(r"sc_pass_\d+", "sc_pass_*", re.sub(r"\(.*", "", k))
            s, d = acc.get(k, (0.0, set()))
            acc[k] = (s + float(row["Counter_Value"]), d | {row["Dispatch_Id"]})
        return {k: s / len(d) for k, (s, d) in acc.items()}

    f, w = per_dispatch(fetch_csv, "FETCH_SIZE"), per_dispatch(write_csv, "WRITE_SIZE")
    kern = "sc_pass_*" if "sc_pass_*" in f else "k_sweep"
    if kern not in f or kern not in w or "k_copy" not in f or "k_copy" not in w:
        return None
    f_scale = (copy_bytes / 2) / f["k_copy"]
    w_scale = (copy_bytes / 2) / w["k_copy"]
    return f[kern] * f_scale + w[kern] * w_scale


def profiled_traffic(workload: str, chunk_bits: int, low_qubits: int):
    """HBM bytes per sweep launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled against k_copy)
    for this workload and geometry, from the summary committed under profiles/ (None when none matches)."""
    import re

    path = Path(ROOT) / "profiles" / "sc_pass_random_layered_30q.md"
    try:
        text = path.read_text()
    except OSError:
        return None, None
    for m in re.finditer(r"^traffic: workload=(\S+) chunk_bits=(\d+) low_qubits=(\d+) bytes_per_launch=(\d+)$", text,
                         re.M):
        if (m.group(1), int(m.group(2)), int(m.group(3))) == (workload, chunk_bits, low_qubits):
            return float(m.group(4)), "profiles/" + path.name
    return None, None


def bench_c2(args):
    n, depth = args.qubits, args.depth
    circ = workloads.random_layered(n, depth, args.seed)
    ops, mats = lower(circ)
    credit = workloads.amp_update_credit(circ)
    ctx = make_context(args)
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

    # time to first result: the first step of a fresh context plans the circuit under each candidate geometry
    # and compiles the chosen plan's pass kernels (hipRTC) before running it
    first_step_ms = None
    for i in range(args.warmup):
        tw = time.perf_counter()
        step()
        ctx.synchronize()
        if i == 0:
            first_step_ms = 1e3 * (time.perf_counter() - tw)
        log(f"warmup step {i + 1}/{args.warmup} done")
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
    plan = ctx.plan_stats()
    geom = ctx.plan_geometry()  # the geometry the library picked for this circuit
    opts = context_options(ctx)
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
        dst.copy_from(st)  # first touch of the new allocation, untimed
        ctx.synchronize()
        ctx.profile_reset()
        ctx.set_option(capi.OPT_PROFILE, 1)
        for _ in range(3):
            dst.copy_from(st)
        ctx.set_option(capi.OPT_PROFILE, 0)
        c_n, c_ms, c_bytes = ctx.profile_read(capi.KFAM_COPY)
        copy_gbs = (c_bytes / c_n) / (c_ms / c_n * 1e-3) / 1e9 if c_n else None
        dst.close()
    fp64_per_step = (capi.plan_sweeps_host(n, ops, mats, geom["low_qubits"], None, opts["remap"], opts["jit"] != 0,
                                           args.pass_cap, geom["chunk_bits"])[0]["fp64_ops_per_amp"]
                     if args.mode == "sweep" else None)
    traffic, traffic_source = None, None
    if args.traffic_csv:
        traffic = pmc_traffic(args.traffic_csv[0], args.traffic_csv[1], 32.0 * (1 << n))
        traffic_source = "rocprofv3 --pmc csv files given with --traffic-csv"
    elif args.mode == "sweep" and opts["jit"]:
        traffic, traffic_source = profiled_traffic(f"random-layered-{n}q-depth{depth}", geom["chunk_bits"],
                                                   geom["low_qubits"])
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
            "blocks_per_cu": args.blocks_per_cu, "pass_cap": args.pass_cap, **opts,
            # chosen by the library's geometry autotuner unless fixed by the options above
            "chunk_bits": geom["chunk_bits"], "low_qubits": geom["low_qubits"],
            "geometries_planned": geom["candidates"], "predicted_gate_ms": geom["ns_per_amp"] * (1 << n) * 1e-6,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "kernel": {"sweep": "sc_pass_*" if opts["jit"] else "k_sweep", "fused": "k_dense, k_dense_mfma",
                       "direct": "k_single"}[args.mode],
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
            "launches": launches,
            "copy_kernel_gbs": copy_gbs,
            "frac_of_copy": (achieved / copy_gbs) if copy_gbs else None,
            # the same launches against the FP64 vector peak (an FMA counts 2 flops)
            "fp64_tflops": (2.0 * fp64_per_step * (1 << n) / (circuit_ms * 1e-3) / 1e12) if fp64_per_step else None,
            "fp64_frac": (2.0 * fp64_per_step * (1 << n) / (circuit_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS)
            if fp64_per_step else None,
            "fp64_ops_per_amp_per_step": fp64_per_step,
        },
        "amp_updates_per_step": credit,
        "first_step_ms": first_step_ms,
        "gate_ms_per_step": circuit_ms,
        "measure_ms_per_step": m_ms / args.steps,
        "checksum": list(cs),
        "sample_boundary_gap": gap,
        # the exact sequential CDF: 4096-value segments summed as integers / serially / all zero
        "sample_exact_sums": ctx.sample_stats(),
        "device": {"hbm_total_gib": total / 2**30, "cus": cus},
    }
    st.close()
    ctx.close()
    return result


def bench_c1(args):
    n = 20
    circ, _ = workloads.ry_qft(n, 42)
    ops, mats = lower(circ)
    credit = workloads.amp_update_credit(circ)
    ctx = make_context(args)
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


def _line(metric, value, unit, args, ms, config, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
         "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
         "data": "synthetic", "config": config}
    d.update(extra)
    return d


def bench_c3(args):
    """Grover search, 33 qubits (128 GiB state), 8 iterations; marked probability vs the closed form."""
    n = args.qubits if args.qubits != 30 else 33
    circ, marked = workloads.grover(n, 8, 33)
    ops, mats = lower(circ)
    credit = workloads.amp_update_credit(circ)
    ctx = make_context(args)
    st = capi.State(ctx, n)
    want = workloads.grover_marked_probability(n, 8)

    def step():
        st.set_basis(0)
        st.apply(ops, mats)
        return abs(st.read(marked)[0]) ** 2

    for _ in range(args.warmup):
        step()
    ctx.profile_reset()
    ctx.set_option(capi.OPT_PROFILE, 1)
    t0 = time.perf_counter()
    for i in range(args.steps):
        p = step()
        log(f"c3 step {i + 1}/{args.steps}: p(marked) = {p:.17g} (closed form {want:.17g})")
    ctx.synchronize()
    el = time.perf_counter() - t0
    ctx.set_option(capi.OPT_PROFILE, 0)
    geom = ctx.plan_geometry()
    launches, ms, nbytes = ctx.profile_read(capi.KFAM_SWEEP)
    achieved = (nbytes / launches) / (ms / launches * 1e-3) / 1e9 if launches else 0.0
    u = SplittableRandom(33).next_doubles(1024)
    out, gap = st.sample(list(range(n)), u)
    res = _line(f"{n}q grover amp-updates/s", credit * args.steps / el, "amp-updates/s", args,
                1e3 * el / args.steps,
                {"workload": f"grover-{n}q-8-iterations", "qubits": n, "ops": len(ops), "marked": marked,
                 "plan": ctx.plan_stats(), **context_options(ctx), "chunk_bits": geom["chunk_bits"],
                 "low_qubits": geom["low_qubits"], "geometries_planned": geom["candidates"]},
                marked_probability=p, closed_form=want, rel_error=abs(p - want) / want,
                roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS,
                          "kernel": "sc_pass_*" if ctx.get_option(capi.OPT_JIT) else "k_sweep"},
                shots_1024_distinct=int(len(np.unique(out))))
    st.close()
    ctx.close()
    return res


def bench_c4(args):
    """Batched QAOA MaxCut rings (instances of 12..22 qubits, p = 4) and depolarising GHZ-20 trajectories."""
    n_inst = args.instances
    ctx = make_context(args)
    batches = [(n, len(idx), ops, mats, w, workloads.ring_zz_masks(n))
               for n, idx, ops, mats, w in workloads.qaoa_batches(n_inst, 4, 4)]

    # the instance states stay allocated in HBM across steps (47.7 GiB for 4096 instances); a step starts
    # every instance from |0..0>
    states = [capi.State(ctx, n, B) for n, B, _, _, _, _ in batches]

    graph_used = []

    def qaoa():
        # every size batch from |0..0>, its circuit and every edge's <Z_i Z_j> (one read of the states per batch),
        # all recorded once as one HIP graph (sc_run_jobs) and replayed with this step's matrices
        zz_all, used = capi.run_jobs(ctx, [(st, 0, ops, mats, np.zeros(n, np.uint64), zm)
                                           for st, (n, B, ops, mats, w, zm) in zip(states, batches)])
        graph_used.append(used)
        mean_c = []
        for zz, (n, B, ops, mats, w, zm) in zip(zz_all, batches):
            mean_c.extend(list(0.5 * np.sum(w * (1.0 - zz), axis=1)))  # <C> = sum_e w_e (1 - <Z_i Z_j>_e) / 2
        return mean_c

    # GHZ-20 trajectories with depolarising p = 0.01 on both operands after every CNOT
    nt, nq = args.trajectories, 20
    gops, gmats, noise_after, u = workloads.ghz_trajectory_uniforms(nt, nq)

    tstate = capi.State(ctx, nq, nt)

    def trajectories():
        # the noisy circuit and every trajectory's parity <Z^20> (summed by the circuit's last pass)
        tstate.set_basis(0)
        parity = tstate.run_depolarizing_trajectories_z(gops, gmats, noise_after, 0.01, u,
                                                        np.array([(1 << nq) - 1], np.uint64))[:, 0]
        ctx.synchronize()
        return parity

    for i in range(args.warmup):  # plans and pass kernels of every instance size and noise segment
        qaoa()
        trajectories()
        log(f"c4 warmup {i + 1}/{args.warmup} done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mean_c = qaoa()
    t_qaoa = (time.perf_counter() - t0) / args.steps
    t1 = time.perf_counter()
    for _ in range(args.steps):
        parity = trajectories()
    t_traj = (time.perf_counter() - t1) / args.steps
    for st in states + [tstate]:
        st.close()
    opts = context_options(ctx)
    geoms = {str(n): list(capi.plan_geometry_host(n, ops, mats[0], opts["low_qubits_option"], opts["remap"],
                                                  opts["jit"], opts["chunk_bits_option"])[:2])
             for n, B, ops, mats, w, zm in batches} if opts["jit"] else None
    return _line("batched QAOA instances/s (12-22 q, p=4) + GHZ-20 depolarising trajectories/s",
                 n_inst / t_qaoa, "instances/s", args, 1e3 * (t_qaoa + t_traj),
                 {"workload": "qaoa-ring-batched+ghz20-trajectories", "instances": n_inst,
                  "trajectories": nt, **opts, "chunk_and_low_qubits_by_size": geoms},
                 qaoa_seconds=t_qaoa, mean_cut_expectation=float(np.mean(mean_c)),
                 qaoa_one_graph_launch=bool(graph_used and all(graph_used)),
                 trajectories_per_s=nt / t_traj, trajectory_seconds=t_traj,
                 ghz_parity_mean=float(np.mean(parity)))


def bench_c5(args):
    """Adjoint gradients of a 24-qubit, 8-layer hardware-efficient ansatz (384 parameters) for a batch of
    parameter sets, against parameter shift on the same engine (timed on a sample, extrapolated)."""
    n, layers, batch = 24, 8, args.param_sets
    P = 2 * n * layers
    rng = np.random.default_rng(5)
    thetas = SplittableRandom(5).next_doubles(batch * P).reshape(batch, P) * 2 * math.pi
    obs = workloads.hea_observable(n, rng.normal(size=n), rng.normal(size=n))
    ops, _ = lower(workloads.hea(n, layers, thetas[0]))
    mats = np.stack([lower(workloads.hea(n, layers, thetas[k]))[1] for k in range(batch)])
    ctx = make_context(args)
    st = capi.State(ctx, n, batch)
    # warm-up with the timed call's segmentation: plans and pass kernels of every checkpoint segment
    st.adjoint_gradient(ops, mats[:1].repeat(batch, 0), P, *obs, checkpoint_every=args.checkpoint, per_element=True)
    geom = ctx.plan_geometry()  # of the forward run (the reverse sweep's passes use 12-qubit chunks)
    st.set_basis(0)
    t0 = time.perf_counter()
    grads, energies = st.adjoint_gradient(ops, mats, P, *obs, checkpoint_every=args.checkpoint, per_element=True)
    ctx.synchronize()
    t_adj = time.perf_counter() - t0
    st.close()
    # parameter shift on the plain engine: E(theta +- pi/2) per parameter, one state per evaluation
    one = capi.State(ctx, n, 1)
    sample = list(range(0, P, max(1, P // args.shift_sample)))[:args.shift_sample]
    # the shifted circuits are lowered before the timed region: only engine work is timed
    shifted = []
    for p in sample:
        for sgn in (1, -1):
            th = thetas[0].copy()
            th[p] += sgn * math.pi / 2
            shifted.append(lower(workloads.hea(n, layers, th)))
    one.set_basis(0)
    one.apply(*shifted[0])  # plan (and kernels) of the structure, as for the adjoint warm-up
    ps = []
    t1 = time.perf_counter()
    for k in range(len(sample)):
        vals = []
        for o, m in shifted[2 * k:2 * k + 2]:
            one.set_basis(0)
            one.apply(o, m)
            vals.append(one.expectation(*obs)[0])
        ps.append(0.5 * (vals[0] - vals[1]))
    ctx.synchronize()
    t_ps = time.perf_counter() - t1
    one.close()
    per_grad_ps = t_ps / len(sample)
    adj_rate = batch * P / t_adj
    ps_rate = 1.0 / per_grad_ps
    err = float(np.max(np.abs(np.array(ps) - grads[0, sample])))
    return _line("adjoint gradients/s, 24q HEA 384 params", adj_rate, "gradients/s", args, 1e3 * t_adj,
                 {"workload": "hea-24q-8-layers-384-params", "param_sets": batch,
                  "checkpoint_every": args.checkpoint, **context_options(ctx),
                  "forward_chunk_bits": geom["chunk_bits"], "forward_low_qubits": geom["low_qubits"]},
                 param_shift_gradients_per_s=ps_rate, speedup_vs_param_shift=adj_rate / ps_rate,
                 param_shift_sample=len(sample), max_abs_diff_vs_param_shift=err,
                 adjoint_stats=ctx.adjoint_stats(), energy_0=float(energies[0]))


def bench_noisy(args):
    """State-dependent noise in one call (sc_run_noisy): GHZ-n with amplitude damping (gamma 0.02) and phase
    damping (lambda 0.03) on both qubits after every CNOT, one trajectory per run, for n = 12, 20, 26; beside it
    the per-channel loop the engine ran before (sc_kraus_weights, the branch chosen on the host,
    sc_apply_kraus per channel and target) on the same uniforms, and the largest difference of the two."""
    from statecraft_amd.noise import ErrorChannel
    ad = list(ErrorChannel.amplitude_damping(0.02, 0).kraus)
    pd = list(ErrorChannel.phase_damping(0.03, 0).kraus)
    kraus = ad + pd
    ctx = make_context(args)
    sizes = {}
    for n in (12, 20, 26):
        ops, mats = lower(workloads.ghz(n))
        events = [(i, q, f, 2) for i in range(len(ops)) if ops[i]["kind"] == capi.OP_CNOT
                  for q in (int(ops[i]["q0"]), int(ops[i]["q1"])) for f in (0, 2)]
        u = SplittableRandom(5).next_doubles(len(events)).reshape(1, -1)
        st = capi.State(ctx, n, 1)

        def device():
            st.set_basis(0)
            st.run_noisy(ops, mats, events, kraus, u)

        def per_channel():
            st.set_basis(0)
            done = 0
            for j, (i, t, f, c) in enumerate(events):
                if i + 1 > done:
                    st.apply(ops[done:i + 1], mats)
                    done = i + 1
                w = st.kraus_weights(t, kraus[f:f + c])
                r, cum, idx = u[0, j] * float(sum(w)), 0.0, c - 1
                for q, x in enumerate(w):
                    cum += float(x)
                    if r < cum:
                        idx = q
                        break
                st.apply_kraus(t, kraus[f + idx])
            if done < len(ops):
                st.apply(ops[done:], mats)

        times = {}
        for name, fn in (("run_noisy", device), ("per_channel", per_channel)):
            for _ in range(max(1, args.warmup)):
                fn()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn()
            ctx.synchronize()
            times[name] = (time.perf_counter() - t0) / args.steps
            if name == "run_noisy":
                a = st.download()
        diff = float(np.max(np.abs(st.download() - a)))
        st.close()
        sizes[str(n)] = {"channels": len(events), "run_noisy_ms": 1e3 * times["run_noisy"],
                         "per_channel_ms": 1e3 * times["per_channel"],
                         "speedup": times["per_channel"] / times["run_noisy"], "max_abs_diff": diff}
        log(f"noisy {n} q: {sizes[str(n)]}")
    ctx.close()
    t = sizes["20"]["run_noisy_ms"] * 1e-3
    return _line("noisy GHZ-20 runs/s (amplitude + phase damping after every CNOT, sc_run_noisy)", 1.0 / t, "runs/s",
                 args, 1e3 * t, {"workload": "ghz-noisy-damping", "gamma": 0.02, "lambda": 0.03}, sizes=sizes)


def other_configs(args):
    """Short runs of the other benchmark configurations at their full sizes, reported beside the headline:
    20-qubit Ry+QFT (3 steps), 33-qubit Grover (1 step), batched QAOA + GHZ-20 trajectories (2 steps); one
    warm-up step each (planning and kernel compilation)."""
    out = {}
    for name, fn, steps, keys in (
            ("c1_ry_qft_20q", bench_c1, 3, ()),
            ("c3_grover_33q", bench_c3, 1, ("marked_probability", "rel_error", "roofline")),
            ("c4_qaoa_batched_and_ghz20", bench_c4, 2, ("trajectories_per_s", "qaoa_seconds",
                                                        "qaoa_one_graph_launch", "mean_cut_expectation"))):
        log(f"{name} ...")
        a = argparse.Namespace(**vars(args))
        a.steps, a.warmup, a.qubits = steps, 1, 30
        r = fn(a)
        out[name] = {"metric": r["metric"], "value": r["value"], "unit": r["unit"], "steps": steps,
                     "ms_per_step": r["ms_per_step"], **{k: r[k] for k in keys}}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "noisy"])
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--trajectories", type=int, default=1024)
    ap.add_argument("--param-sets", type=int, default=64)
    ap.add_argument("--checkpoint", type=int, default=144)
    ap.add_argument("--shift-sample", type=int, default=8)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--shots", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--mode", default="sweep", choices=list(MODES))
    ap.add_argument("--max-fused", type=int, default=None, choices=[0, 1, 2, 3, 4, 5],
                    help="most qubits of a dense fused block (SC_OPT_MAX_FUSED; fused mode: 1..5)")
    ap.add_argument("--pass-sched", type=int, default=None, choices=[0, 1, 2, 3, 4, 5, 6],
                    help="SC_OPT_PASS_SCHED: wave-priority variant of the runtime-compiled passes (library default 4)")
    ap.add_argument("--permlane", type=int, default=None, choices=[0, 1],
                    help="SC_OPT_PERMLANE: stage transitions inside the registers where the plan allows (library "
                         "default 0)")
    ap.add_argument("--fused-terms", type=int, default=None, choices=[0, 1],
                    help="SC_OPT_FUSED_TERMS: c4's Z-string terms summed in each job's last pass (library default 1)")
    ap.add_argument("--mfma", type=int, default=None, choices=[0, 1, 2],
                    help="SC_OPT_MFMA, fused mode: 0 VALU blocks only, 1 5-qubit blocks on the FP64 matrix cores "
                         "(library default), 2 also 4-qubit blocks")
    ap.add_argument("--mfma-tile", type=int, default=None, choices=[-1, 0, 1, 2, 3, 4, 5],
                    help="SC_OPT_MFMA_TILE, fused mode: the matrix-core blocks' tile shape (-1: autotuned, the "
                         "library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="c2: skip the adjoint-gradient block of the metric")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="c2: skip the short runs of the 20-qubit QFT, 33-qubit Grover and batched configurations")
    ap.add_argument("--no-direct-io", action="store_true",
                    help="SC_OPT_DIRECT_IO 0: every pass goes through LDS at both ends")
    ap.add_argument("--no-copy-reference", action="store_true")
    ap.add_argument("--low-qubits", type=int, default=None,
                    help="SC_OPT_LOW_QUBITS: coalesced low positions every pass keeps local (default: the library "
                         "picks per circuit)")
    ap.add_argument("--chunk-bits", type=int, default=None, choices=[0, 12, 13],
                    help="SC_OPT_CHUNK_BITS: local qubits per chunk of the runtime-compiled passes (default: the "
                         "library picks per circuit from its measured pass-time model)")
    ap.add_argument("--remap", action="store_true", help="sweep passes relabel qubits in place (SC_OPT_REMAP)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="resident sweep workgroups per CU (0: all that fit)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--traffic-csv", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="rocprofv3 --pmc counter_collection.csv files of FETCH_SIZE and WRITE_SIZE runs of this "
                         "workload: fills roofline.traffic")
    ap.add_argument("--pass-cap", type=int, default=-1,
                    help="SC_OPT_SWEEP_PASS_CAP: FP64 operations per amplitude one pass takes on (-1 default, 0 none)")
    ap.add_argument("--jit", type=int, default=None, choices=[0, 1, 2],
                    help="SC_OPT_JIT: 0 compiled sweep kernel, 1 runtime-compiled pass kernels, 2 the same with "
                         "next-chunk prefetch (the library default; kernels are compiled during warmup)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is designed, not built",
              file=sys.stderr)
        return 2
    res = {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5,
           "noisy": bench_noisy}[args.workload](args)
    if args.workload == "c2" and not args.no_adjoint:
        # the metric's second half: adjoint gradients of the 24-qubit, 384-parameter HEA against parameter shift
        log("adjoint block (24-qubit HEA, 384 parameters, batch 64) ...")
        c5 = bench_c5(args)
        res["adjoint"] = {k: c5[k] for k in ("value", "unit", "ms_per_step", "param_shift_gradients_per_s",
                                             "speedup_vs_param_shift", "max_abs_diff_vs_param_shift",
                                             "param_shift_sample")}
        res["adjoint"]["config"] = c5["config"]
    if args.workload == "c2" and not args.no_other_configs:
        res["other_configs"] = other_configs(args)
    if args.workload == "c2" and not args.no_cpu_baseline:
        log("cpu baseline (reference path restated in C, all host threads, 29 qubits) ...")
        res["cpu_baseline"] = cpu_baseline(args.depth, args.seed, args.cpu_seconds)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
