"""Regression against the committed fixtures in tests/golden/ (oracle output)."""
import hashlib
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name + ".npy"), allow_pickle=False)


def test_manifest_hashes():
    for line in open(os.path.join(HERE, "MANIFEST.txt")):
        digest, name = line.split()
        assert hashlib.sha256(open(os.path.join(HERE, name), "rb").read()).hexdigest() == digest


def test_oracle_reproduces_fixtures():
    A = load("pattern_N16_NITER2_A"); B = load("pattern_N16_NITER2_B")
    C = load("pattern_N16_NITER2_C")
    for it in range(2):
        assert np.array_equal(oracle.host_mm(A[it], B[it]), C[it])
        assert np.array_equal(oracle.pattern_slice(16, it), C[it])
    A = load("uniform_N32_A"); B = load("uniform_N32_B")
    assert np.array_equal(oracle.gen("uniform", 101, 32 * 32).reshape(32, 32), A)
    assert np.array_equal(oracle.host_mm(A, B), load("uniform_N32_host_mm"))
    assert np.array_equal(oracle.fmaf_chain_mm(A, B), load("uniform_N32_fmaf"))
    assert np.array_equal(oracle.gemm_f64(A, B)[0], load("uniform_N32_f64"))


@pytest.mark.gpu
def test_gpu_kernels_on_fixtures():
    torch = pytest.importorskip("torch")
    import hb_hammerbench_amd as hb
    A = load("pattern_N16_NITER2_A"); B = load("pattern_N16_NITER2_B")
    C = load("pattern_N16_NITER2_C")
    a = torch.from_numpy(A.ravel()).cuda(); b = torch.from_numpy(B.ravel()).cuda()
    for fn in (hb.sgemm, hb.sgemm_bf16x3):
        c = torch.empty_like(a)
        fn(a, b, c, 16, 2)
        torch.cuda.synchronize()
        assert np.array_equal(c.cpu().numpy().reshape(2, 16, 16), C)
    A = load("uniform_N32_A"); B = load("uniform_N32_B"); R = load("uniform_N32_f64")
    S = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
    a = torch.from_numpy(A.ravel()).cuda(); b = torch.from_numpy(B.ravel()).cuda()
    c = torch.empty_like(a)
    hb.sgemm(a, b, c, 32, 1)
    torch.cuda.synchronize()
    g = 34 * 2.0 ** -24 / (1 - 34 * 2.0 ** -24)
    assert np.all(np.abs(c.cpu().numpy().reshape(32, 32) - R) <= g * S)
    sse, _ = oracle.sse(load("uniform_N32_host_mm"), c.cpu().numpy().reshape(32, 32))
    assert sse < 0.01
