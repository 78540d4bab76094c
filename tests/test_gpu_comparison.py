"""The comparison-runner front end on the GPU: the reference's benchmark-harness fixtures through
`python -m statecraft_amd.comparison`'s code path.  Ideal fixtures must give their textbook probabilities
(1e-12) and the oracle's final state; noisy fixtures must match an exact density-matrix evolution of the same
circuit and channels within Monte-Carlo error, and repeat exactly for the same seeds."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from statecraft_amd import comparison
from statecraft_amd.circuit import lower
from test_comparison_host import BY_ID, EXPECTED, circuit_text
from test_gpu_noise_adjoint import oracle_noisy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from statecraft_amd.engine import HipStatevectorEngine

    e = HipStatevectorEngine()
    yield e
    e.close()


def run_fixture(engine, tmp_path, fid, extra=()):
    f = BY_ID[fid]
    path = tmp_path / f"{fid}.json"
    path.write_text(circuit_text(f))
    args = ["--circuit-json", str(path), "--fixture-id", fid, "--category", f["category"], "--timed-runs", "2"]
    return json.loads(comparison.run(args + list(extra), engine=engine))


@pytest.mark.parametrize("fid", sorted(EXPECTED))
def test_ideal_fixtures(engine, tmp_path, fid):
    d = run_fixture(engine, tmp_path, fid)
    assert d["runner"] == "statecraft" and d["status"] == "ok" and d["fixture_id"] == fid
    assert len(d["timed_runs_ms"]) == 2 and d["noise_trajectories"] == 0
    p = np.array(d["probabilities"])
    want = np.zeros_like(p)
    for k, v in EXPECTED[fid].items():
        want[k] = v
    assert np.max(np.abs(p - want)) <= 1e-12
    circ = comparison.parse_circuit_json(circuit_text(BY_ID[fid]))
    ops, mats = lower(circ)
    sv = np.array([complex(re, im) for re, im in d["statevector"]])
    assert np.max(np.abs(sv - oracle.run_ops(circ.qubit_count, ops, mats))) <= 1e-12


# Kraus sets from the channel definitions (textbook forms)
def _kraus(noise):
    t = noise["type"]
    I2, X = np.eye(2), np.array([[0, 1], [1, 0]])
    Y, Z = np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0])
    if t == "depolarizing":
        p = noise["probability"]
        return [np.sqrt(1 - p) * I2, np.sqrt(p / 3) * X, np.sqrt(p / 3) * Y, np.sqrt(p / 3) * Z]
    if t == "phase_flip":
        p = noise["probability"]
        return [np.sqrt(1 - p) * I2, np.sqrt(p) * Z]
    if t == "amplitude_damping":
        g = noise["gamma"]
        return [np.array([[1, 0], [0, np.sqrt(1 - g)]]), np.array([[0, np.sqrt(g)], [0, 0]])]
    if t == "phase_damping":
        lam = noise["lambda"]
        return [np.array([[1, 0], [0, np.sqrt(1 - lam)]]), np.array([[0, 0], [0, np.sqrt(lam)]])]
    return None


def _on(n, q, m):
    """2x2 matrix m on qubit q of n (bit q of the index = qubit q)."""
    full = np.array([[1.0]])
    for k in reversed(range(n)):
        full = np.kron(full, m if k == q else np.eye(2))
    return full


def _unitary(n, ops, mats):
    """Dense matrix of a lowered circuit, column by column through the oracle."""
    d = 1 << n
    cols = []
    for j in range(d):
        e = np.zeros(d, dtype=complex)
        e[j] = 1.0
        cols.append(oracle.run_ops(n, ops, mats, e))
    return np.array(cols).T


@pytest.mark.parametrize("fid", ["noise_depolarizing_1q", "noise_phase_flip_2q", "noise_amplitude_damping_2q",
                                 "noise_phase_damping_3q"])
def test_noisy_fixtures_match_density_matrix(engine, tmp_path, fid):
    f = BY_ID[fid]
    noise = f["noise"]
    flags = ["--noise-type", noise["type"], "--noise-qubits", ",".join(str(q) for q in noise["qubits"]),
             "--noise-trajectories", "512", "--noise-seed", str(noise["seed"]), "--timed-runs", "1",
             "--warmup-runs", "0"]
    for key, flag in (("probability", "--noise-probability"), ("gamma", "--noise-gamma"),
                      ("lambda", "--noise-lambda")):
        if key in noise:
            flags += [flag, str(noise[key])]
    d = run_fixture(engine, tmp_path, fid, flags)
    p = np.array(d["probabilities"])
    assert d["noise_trajectories"] == 512 and "statevector" not in d
    assert abs(p.sum() - 1.0) <= 1e-12
    # exact: every gate, then every noise qubit's channel
    n = f["qubits"]
    circ = comparison.parse_circuit_json(circuit_text(f))
    rho = np.zeros((1 << n, 1 << n), dtype=complex)
    rho[0, 0] = 1.0
    ks = _kraus(noise)
    for op in circ.operations:
        o, m = lower(type(circ)(n, [op]))
        U = _unitary(n, o, m)
        rho = U @ rho @ U.conj().T
        for q in noise["qubits"]:
            rho = sum(_on(n, q, k) @ rho @ _on(n, q, k).conj().T for k in ks)
    exact = np.real(np.diag(rho))
    assert np.max(np.abs(p - exact)) <= 0.1  # 512 trajectories: a few standard errors
    # the pin: trajectory t runs with seed S + t, so the oracle's applyErrorChannel with the same seeds gives
    # the same 512 trajectories and the same average
    o = comparison.Options.parse(["--circuit-json", str(tmp_path / f"{fid}.json"), "--fixture-id", fid,
                                  "--category", f["category"]] + flags)
    model = o.noise_model()
    avg = np.zeros(1 << n)
    for t in range(512):
        s = oracle_noisy(circ, model, noise["seed"] + t)
        avg += s.real * s.real + s.imag * s.imag
    assert np.max(np.abs(p - avg / 512)) <= 1e-12
    # the same seeds give the same trajectories
    again = run_fixture(engine, tmp_path, fid, flags)
    assert again["probabilities"] == d["probabilities"]


def test_thermal_relaxation_fixture(engine, tmp_path):
    f = BY_ID["noise_thermal_relaxation_3q"]
    noise = f["noise"]
    flags = ["--noise-type", "thermal_relaxation", "--noise-qubits", "0,1,2", "--noise-t1", str(noise["t1"]),
             "--noise-t2", str(noise["t2"]), "--noise-gate-time", str(noise["gate_time"]),
             "--noise-trajectories", "128", "--noise-seed", str(noise["seed"]), "--timed-runs", "1"]
    d = run_fixture(engine, tmp_path, "noise_thermal_relaxation_3q", flags)
    p = np.array(d["probabilities"])
    assert abs(p.sum() - 1.0) <= 1e-12 and p.min() >= 0.0
    # the same 128 trajectories (seed S + t) through the oracle's applyErrorChannel
    circ = comparison.parse_circuit_json(circuit_text(f))
    o = comparison.Options.parse(["--circuit-json", str(tmp_path / "noise_thermal_relaxation_3q.json"),
                                  "--fixture-id", "noise_thermal_relaxation_3q", "--category", f["category"]] + flags)
    model = o.noise_model()
    avg = np.zeros(1 << f["qubits"])
    for t in range(128):
        s = oracle_noisy(circ, model, noise["seed"] + t)
        avg += s.real * s.real + s.imag * s.imag
    assert np.max(np.abs(p - avg / 128)) <= 1e-12
