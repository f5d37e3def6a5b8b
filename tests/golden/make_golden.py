"""Write golden vectors from the reference's own FSM core.

The reference has no tests or fixtures (its only sample input is missing),
so the known-answer vectors come from running its runFSM
(mpi_openacc/src/phi3D_fsm.c, compiled from source by oracle/Makefile into
oracle/_ref/libpdfs_ref.so) on small seeded inputs:

  case_*: random init (sources 0, far 1e6, the rest U(0,20)), ghosts +inf;
          after each single ordering sn = 1..8 (each from the same input),
          after one pass (sn = 1..8 in sequence) and converged (1e-12).
  c0:     64^3 point source (config c0), converged; stored as sha256 of the
          padded array bytes, the pass count and a sample of values.

Usage: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_ctypes as O  # noqa: E402

CASES = [
    # (name, X, Y, Z, (dx, dy, dz), seed)
    ("a", 4, 5, 6, (1.0, 1.0, 1.0), 11),
    ("b", 8, 8, 8, (1.0, 1.0, 1.0), 12),
    ("c", 7, 3, 9, (0.5, 0.7, 1.3), 13),
    ("d", 12, 10, 6, (2.0, 1.0, 0.25), 14),
]


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle)"
    out = {}
    for name, X, Y, Z, h, seed in CASES:
        a = O.random_field(X, Y, Z, seed)
        out[f"{name}_dims"] = np.array([X, Y, Z])
        out[f"{name}_h"] = np.array(h)
        out[f"{name}_input"] = a.copy()
        for sn in range(1, 9):
            b = a.copy()
            O.ref_run_sn(b, sn, h)
            out[f"{name}_sn{sn}"] = b
        b = a.copy()
        for sn in range(1, 9):
            O.ref_run_sn(b, sn, h)
        out[f"{name}_pass1"] = b
        b = a.copy()
        passes, _, md = O.ref_solve(b, h, tol=1e-12, max_passes=200)
        out[f"{name}_converged"] = b
        out[f"{name}_passes"] = np.array(passes)
    c0 = O.gen_point(64, 64, 64, 12345)
    passes, _, md = O.ref_solve(c0, tol=1e-12, max_passes=200)
    out["c0_passes"] = np.array(passes)
    out["c0_sha256"] = np.frombuffer(hashlib.sha256(c0.tobytes()).digest(), dtype=np.uint8)
    rng = np.random.default_rng(0)
    idx = rng.integers(0, c0.size, 256)
    out["c0_sample_idx"] = idx
    out["c0_sample_val"] = c0.reshape(-1)[idx]
    path = os.path.join(HERE, "ref_fsm_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes; c0 passes", passes)


if __name__ == "__main__":
    main()
