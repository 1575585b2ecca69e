"""Writes tests/golden/golden.json: oracle outputs of the full-size
configurations, as hashes and sampled voxels, so that GPU tests can check the
512^3 result without running the oracle on the GPU machine.

The inputs come from itk-rs_amd/workloads.py (seeded, identical bits on any
device); the outputs from the oracle's threaded restatement
(oracle/gauss_oracle.c, orc_ref_faithful_f32), which tests/test_oracle.py
pins as bit-identical to the plain restatement.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from __graft_entry__ import load_package  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    load_package()
    from itk_rs_amd import workloads as wl

    out = {}
    threads = os.cpu_count() or 1

    t = time.time()
    x, sigma, sp = wl.c1("cpu")
    a = x.numpy()
    y = oracle.generate_region(a, [0, 0], sp, sigma, [0, 0], [512, 512])
    out["c1"] = {"input_sha256": sha(a), "output_sha256": sha(y), "sigma": sigma}
    print(f"c1 done {time.time() - t:.1f}s", flush=True)

    t = time.time()
    x, sigma, sp = wl.c2("cpu")
    a = x.numpy()
    print(f"c2 input {time.time() - t:.1f}s", flush=True)
    y = oracle.ref_faithful(a, [0, 0, 0], sp, sigma, [0, 0, 0], [512, 512, 512], 1, threads)
    rng = np.random.default_rng(2024)
    idx = np.sort(rng.choice(y.size, 2048, replace=False))
    out["c2"] = {
        "input_sha256": sha(a), "output_sha256": sha(y), "sigma": sigma,
        "sample_flat_index": idx.tolist(),
        "sample_value_bits": y.reshape(-1)[idx].view(np.uint32).tolist(),
    }
    print(f"c2 done {time.time() - t:.1f}s", flush=True)

    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json"), "w") as fh:
        json.dump(out, fh)


if __name__ == "__main__":
    main()
