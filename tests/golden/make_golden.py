"""Writes tests/golden/random_circuits.npz: random circuits over every op kind with their outputs from the
independent tensor-contraction brute force (tests/bruteforce.py).  Run from the repo root:
    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import bruteforce as bf  # noqa: E402
from circuits import random_circuit, random_state  # noqa: E402


def main():
    rng = np.random.default_rng(20240917)
    out = {}
    cases = [(1, 12), (2, 20), (3, 30), (4, 40), (5, 40), (6, 60), (7, 60), (8, 60)]
    for c, (n, count) in enumerate(cases):
        ops, mats = random_circuit(n, count, rng)
        init = random_state(n, rng)
        out[f"n{c}"] = np.array(n)
        out[f"ops{c}"] = np.frombuffer(ops.tobytes(), dtype=np.uint8)
        out[f"mats{c}"] = mats
        out[f"init{c}"] = init
        out[f"out{c}"] = bf.run(n, ops, mats, init)
    out["count"] = np.array(len(cases))
    np.savez_compressed(HERE / "random_circuits.npz", **out)
    print("wrote", HERE / "random_circuits.npz")


if __name__ == "__main__":
    main()
