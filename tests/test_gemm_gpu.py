"""GPU tests of the update GEMM (rocblas_amd_dgemm / _sgemm, the kernels the
blocked solve's trailing updates run on), exact on small-integer data: every
product and partial sum is an integer below 2^24, so FP64 and FP32 results
must equal numpy's bit for bit whatever the summation order.  Shapes cover the
LDS-DMA FP64 kernel (even M, K a multiple of 8, 16-byte aligned operands) at
tile-aligned and ragged M / N, with beta = 0 and beta != 0, an operand offset
by one element (misaligned: the register-staged kernel), odd M and K, the
transposed forms (the DMA kernels store a k-contiguous operand row-major with
a swizzle, an x-contiguous one k-major) and FP32 (its DMA kernel loads quads:
shapes with extents and leading dimensions that are or are not multiples of
4)."""
import numpy as np
import pytest
import torch

import rocblas_amd as rb
from tests.gpu_util import GpuHandle, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available()
    hh = GpuHandle()
    yield hh
    hh.close()


def _mat(rng, rows, cols, ld, dtype):
    M = np.zeros((ld, cols), dtype=dtype, order="F")
    M[:rows] = rng.integers(-4, 5, (rows, cols))
    return M


CASES = [
    # m, n, k, transA, transB, alpha, beta, lda pad, ldb pad, ldc pad, shift
    (2048, 2048, 256, "N", "N", 1.0, 0.0, 0, 0, 0, 0),
    (2050, 2100, 72, "N", "N", -1.0, 1.0, 2, 4, 1, 0),
    (1664, 2304, 8, "N", "N", 1.0, 1.0, 0, 0, 0, 0),
    (2176, 2048, 512, "N", "N", -1.0, 1.0, 6, 2, 3, 0),
    (2048, 2048, 256, "N", "N", 1.0, 1.0, 0, 0, 0, 1),   # misaligned A and B
    (2047, 2048, 96, "N", "N", 1.0, 0.0, 1, 0, 0, 0),    # odd M
    (2048, 2048, 100, "N", "N", 1.0, 0.0, 0, 0, 0, 0),   # K not a multiple of 8
    (2048, 2056, 64, "T", "N", -1.0, 1.0, 0, 0, 0, 0),
    (2048, 2048, 64, "N", "T", 1.0, 1.0, 0, 0, 0, 0),
    (2050, 2098, 40, "T", "N", 1.0, 1.0, 2, 0, 0, 0),    # K only a multiple of 8
    (2049, 2048, 64, "T", "N", -1.0, 0.0, 0, 2, 1, 0),   # odd M, A transposed
    (2048, 2050, 48, "N", "T", 1.0, 1.0, 0, 4, 0, 0),
    (2052, 2046, 96, "T", "T", -1.0, 1.0, 2, 2, 0, 0),
    (2048, 2047, 64, "N", "T", 1.0, 0.0, 0, 1, 0, 0),    # odd N, B transposed
    (2052, 2100, 96, "N", "N", -1.0, 1.0, 4, 4, 0, 0),   # ragged tiles, FP32 quads
    (2048, 2052, 48, "N", "T", 1.0, 1.0, 0, 4, 0, 0),    # K a multiple of 16 only
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}{c[3]}{c[4]}s{c[10]}"
                                             for c in CASES])
@pytest.mark.parametrize("prec", ["d", "s"])
def test_gemm_exact(h, case, prec):
    m, n, k, ta, tb, alpha, beta, pa, pb, pc, shift = case
    dt = np.float64 if prec == "d" else np.float32
    rng = np.random.default_rng(m * 7 + n + k)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + pa, br + pb, m + pc
    A = _mat(rng, ar, ac, lda, dt)
    B = _mat(rng, br, bc, ldb, dt)
    C = _mat(rng, m, n, ldc, dt)
    opA = A[:ar] if ta == "N" else A[:ar].T
    opB = B[:br] if tb == "N" else B[:br].T
    want = C.copy()
    want[:m] = alpha * (opA.astype(np.float64) @ opB.astype(np.float64)) + beta * C[:m]
    tdt = torch.float64 if prec == "d" else torch.float32
    # `shift` puts A and B one element into their buffers (8 / 4 bytes off 16)
    dA = torch.zeros(lda * ac + shift, dtype=tdt, device="cuda")
    dB = torch.zeros(ldb * bc + shift, dtype=tdt, device="cuda")
    dA[shift:] = torch.from_numpy(A.ravel(order="F"))
    dB[shift:] = torch.from_numpy(B.ravel(order="F"))
    dC = torch.from_numpy(C.ravel(order="F").copy()).cuda()
    esz = dA.element_size()
    say("gemm", prec, case)
    st = rb.gemm(h, ta, tb, m, n, k, alpha, dA.data_ptr() + shift * esz, lda,
                 dB.data_ptr() + shift * esz, ldb, beta, dC.data_ptr(), ldc, precision=prec)
    assert st == rb.rocblas_status.success
    torch.cuda.synchronize()
    got = dC.cpu().numpy().reshape(n, ldc).T
    assert np.array_equal(got[:m], want[:m].astype(dt)), np.abs(got[:m] - want[:m]).max()
    assert np.array_equal(got[m:], C[m:])  # rows past m untouched
