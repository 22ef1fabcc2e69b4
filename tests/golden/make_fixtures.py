"""Writes the small regression fixtures in tests/golden/ from the oracle.

Fixtures (all .npy, allow_pickle=False):
  pattern_N16_NITER2_{A,B,C}.npy   reference init (apps/sgemm/main.cpp:46-52) and
                                   host_mm golden (main.cpp:15-26) for N_16__NITER_2
  uniform_N32_{A,B}.npy            seeded uniform[-1,1) inputs (seeds 101, 102)
  uniform_N32_host_mm.npy          host_mm golden of those inputs (float32)
  uniform_N32_fmaf.npy             the device-order fmaf chain (kernel.cpp:186-275)
  uniform_N32_f64.npy              FP64 product
MANIFEST.txt lists the SHA-256 of every file.  Run: python tests/golden/make_fixtures.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def fixtures():
    out = {}
    N, NITER = 16, 2
    a, b = oracle.pattern_init(NITER * N * N)
    A = a.reshape(NITER, N, N); B = b.reshape(NITER, N, N)
    out["pattern_N16_NITER2_A"] = A
    out["pattern_N16_NITER2_B"] = B
    out["pattern_N16_NITER2_C"] = np.stack([oracle.host_mm(A[i], B[i]) for i in range(NITER)])
    n = 32
    A = oracle.gen("uniform", 101, n * n).reshape(n, n)
    B = oracle.gen("uniform", 102, n * n).reshape(n, n)
    out["uniform_N32_A"] = A
    out["uniform_N32_B"] = B
    out["uniform_N32_host_mm"] = oracle.host_mm(A, B)
    out["uniform_N32_fmaf"] = oracle.fmaf_chain_mm(A, B)
    out["uniform_N32_f64"] = oracle.gemm_f64(A, B)[0]
    return out


def main():
    lines = []
    for name, arr in fixtures().items():
        path = os.path.join(HERE, name + ".npy")
        np.save(path, arr, allow_pickle=False)
        lines.append(f"{hashlib.sha256(open(path, 'rb').read()).hexdigest()}  {name}.npy")
    with open(os.path.join(HERE, "MANIFEST.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
