"""GPU tests of the ILP64 entry points with leading dimensions above INT32_MAX.

The cases are the reference's own stress cases trsm_left_64 / trsm_right_64
(clients/gtest/trsm_gtest.yaml:1236-1269; the right-side one has ldb < M and
checks the size error, so it is also run with ldb = M): single precision,
device pointer mode, lda or ldb = 2^31 + 1 (the buffer spans ~26 GB of HBM; only the used
entries are written), in every uplo / transA combination.  The reference
solves them with its 64-bit substitution fallback
(library/src/src64/blas3/rocblas_trsm_kernels_64.cpp:150-215); this library
gathers the operand with the huge leading dimension into a compact copy, so
each result must equal, bit for bit, the same solve through the 32-bit entry
point on compactly stored operands, and agree with the oracle.  One FP64 case
(lda = 2^31 + 1, ~52 GB) covers the double form.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from tests.gpu_util import GpuHandle, dev, host, say

pytestmark = pytest.mark.gpu
S = rb.rocblas_status
HUGE = (1 << 31) + 1
SIDE = {"L": 141, "R": 142}
FILL = {"U": 121, "L": 122}
OP = {"N": 111, "T": 112}
DIAG = {"N": 131, "U": 132}


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available()
    oracle.build()
    hd = GpuHandle()
    rb.check(hd.set_pointer_mode(1), "set_pointer_mode")  # device alpha, as the cases ask
    yield hd
    rb.check(hd.set_pointer_mode(0), "set_pointer_mode")
    hd.close()
    torch.cuda.empty_cache()


def _strided(buf, rows, cols, ld):
    """(rows x cols) column-major view of a flat device buffer with leading dimension ld."""
    return buf.as_strided((rows, cols), (1, ld))


def _run(h, dtype, side, uplo, trans, diag, m, n, lda, ldb, seed):
    np_t = np.float32 if dtype == torch.float32 else np.float64
    rng = np.random.default_rng(seed)
    k = m if side == "L" else n
    A = oracle.diag_dominant_triangular(k, uplo, rng, dtype=np_t, poison=0.0)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np_t))
    alpha = torch.tensor([0.75], dtype=dtype, device="cuda")
    fn64 = rb.lib().rocblas_strsm_64 if dtype == torch.float32 else rb.lib().rocblas_dtrsm_64
    fn32 = rb.lib().rocblas_strsm if dtype == torch.float32 else rb.lib().rocblas_dtrsm
    args = (SIDE[side], FILL[uplo], OP[trans], DIAG[diag])

    # compact reference through the 32-bit entry point
    dAc, dBc = dev(A), dev(B)
    assert fn32(h.ptr, *args, m, n, alpha.data_ptr(), dAc.data_ptr(), k, dBc.data_ptr(),
                m) == S.success

    # the same operands with the given leading dimensions
    dA = torch.empty((k - 1) * lda + k, dtype=dtype, device="cuda")
    dB = torch.empty((n - 1) * ldb + m, dtype=dtype, device="cuda")
    _strided(dA, k, k, lda).copy_(torch.from_numpy(A).cuda())
    _strided(dB, m, n, ldb).copy_(torch.from_numpy(B).cuda())
    say("ilp64 huge ld", dtype, side, uplo, trans, diag, m, n, lda, ldb)
    assert fn64(h.ptr, *args, m, n, alpha.data_ptr(), dA.data_ptr(), lda, dB.data_ptr(),
                ldb) == S.success
    torch.cuda.synchronize()
    got = _strided(dB, m, n, ldb).cpu().numpy()
    want = host(dBc, (m, n))
    assert np.array_equal(got, want), (side, uplo, trans)
    # the oracle agrees
    ref = B.copy(order="F")
    oracle.trsm(side, uplo, trans, diag, m, n, 0.75, np.asfortranarray(A), ref)
    tol = (5e-4 if dtype == torch.float32 else 1e-11) * max(1.0, np.abs(ref).max())
    assert np.abs(got - ref).max() <= tol
    del dA, dB
    torch.cuda.empty_cache()


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_left_huge_lda(h, uplo, trans):
    _run(h, torch.float32, "L", uplo, trans, "N", 4, 8000, HUGE, 10, 1)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_left_huge_ldb(h, uplo, trans):
    _run(h, torch.float32, "L", uplo, trans, "N", 8000, 4, 8001, HUGE, 2)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_right_huge_lda(h, uplo, trans):
    """The reference's right-side case {M: 512, N: 4, lda: 2^31 + 1, ldb: 10}
    has ldb < M, so it checks the size error (returned before any memory is
    touched); the same case with ldb = M is solved."""
    one = torch.tensor([1.0], dtype=torch.float32, device="cuda")
    dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
    st = rb.lib().rocblas_strsm_64(h.ptr, SIDE["R"], FILL[uplo], OP[trans], DIAG["U"], 512, 4,
                                   one.data_ptr(), dummy.data_ptr(), HUGE, dummy.data_ptr(), 10)
    assert st == S.invalid_size
    _run(h, torch.float32, "R", uplo, trans, "U", 512, 4, HUGE, 512, 3)


def test_double_huge_lda_and_strided_batch(h):
    """FP64 with lda = 2^31 + 1, and a strided batch of two instances whose B
    has ldb = 2^31 + 1 (each instance gathered and solved on its own)."""
    _run(h, torch.float64, "L", "U", "T", "N", 3, 500, HUGE, 5, 4)
    rng = np.random.default_rng(5)
    m, n, batch = 200, 2, 2
    A = oracle.diag_dominant_triangular(m, "L", rng, poison=0.0)
    Bs = [np.asfortranarray(rng.uniform(-1, 1, (m, n))) for _ in range(batch)]
    alpha = torch.tensor([1.0], dtype=torch.float64, device="cuda")
    dA = dev(A)
    wants = []
    for B in Bs:
        dBc = dev(B)
        assert rb.lib().rocblas_dtrsm(h.ptr, 141, 122, 111, 131, m, n, alpha.data_ptr(),
                                      dA.data_ptr(), m, dBc.data_ptr(), m) == S.success
        wants.append(dBc)
    stride_b = (n - 1) * HUGE + m
    dB = torch.empty(batch * stride_b, dtype=torch.float64, device="cuda")
    for b, B in enumerate(Bs):
        _strided(dB[b * stride_b:], m, n, HUGE).copy_(torch.from_numpy(B).cuda())
    say("ilp64 huge ldb strided batched")
    assert rb.lib().rocblas_dtrsm_strided_batched_64(
        h.ptr, 141, 122, 111, 131, m, n, alpha.data_ptr(), dA.data_ptr(), m, 0, dB.data_ptr(),
        HUGE, stride_b, batch) == S.success
    torch.cuda.synchronize()
    for b in range(batch):
        got = _strided(dB[b * stride_b:], m, n, HUGE).cpu().numpy()
        assert np.array_equal(got, host(wants[b], (m, n))), b
    del dB
    torch.cuda.empty_cache()
