"""GPU tests of rocblas_amd_dtrsm_mixed (FP32 or split-BF16x3 trailing
updates, FP64 diagonal-block inverses, FP64 residual and iterative
refinement) through the C ABI.

The mixed form answers the same problem as rocblas_dtrsm
(library/src/blas3/rocblas_trsm.cpp); results are checked against the oracle
(the netlib loop nest of the clients' host cblas_dtrsm) and by the normwise
backward error the workload demands (<= 1e-12); the call's own measurement
of it must meet the tolerance it was given.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from tests.gpu_util import GpuHandle, dev, host, say
from tests.helpers import ALL_CASES, integer_kat, op_dense

pytestmark = pytest.mark.gpu
S = rb.rocblas_status


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available()
    oracle.build()
    hd = GpuHandle()
    yield hd
    hd.close()


def eta_np(side, uplo, trans, diag, alpha, A, B0, X):
    k = A.shape[1]
    op = op_dense(A, uplo, trans, diag, k)
    R = alpha * B0 - (op @ X if side == "L" else X @ op)
    n1 = lambda M: np.abs(M).sum(axis=0).max()
    return n1(R) / (n1(op) * n1(X) + abs(alpha) * n1(B0))


@pytest.mark.parametrize("form", ["fp32", "bf16x3"])
@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
@pytest.mark.parametrize("diag", ["N", "U"])
def test_against_oracle_and_backward_error(h, side, uplo, trans, diag, form):
    """Sizes with full and partial diagonal blocks (block 128): the refined result
    meets eta <= 1e-12, agrees with the oracle, and the reported eta is the
    measured one."""
    rng = np.random.default_rng(200 + 2 * ALL_CASES.index((side, uplo, trans)) + (diag == "U"))
    for m, n in [(300, 70), (256, 129), (90, 333)]:
        say("mixed", form, side, uplo, trans, diag, m, n)
        k = m if side == "L" else n
        A = oracle.diag_dominant_triangular(k, uplo, rng, lda=k + 1)
        if diag == "U":
            oracle.make_unit_diagonal(A, uplo, k)
        B = np.zeros((m + 2, n), order="F")
        B[:m] = rng.uniform(-1, 1, (m, n))
        alpha = -0.75
        dA, dB = dev(A), dev(B)
        st, info = rb.trsm_mixed(h, side, uplo, trans, diag, m, n, alpha, dA.data_ptr(), k + 1,
                                 dB.data_ptr(), m + 2, block=128, max_refine=8, tol=1e-14,
                                 form=form)
        assert st == S.success
        torch.cuda.synchronize()
        got = host(dB, B.shape)
        assert np.all(got[m:] == 0.0)  # padding rows untouched
        X = got[:m]
        ref = B.copy(order="F")
        oracle.trsm(side, uplo, trans, diag, m, n, alpha, A, ref)
        e = eta_np(side, uplo, trans, diag, alpha, A, B[:m], X)
        assert e <= 1e-12, e
        # one leaf (k <= 128) is a pure FP64 product: no refinement needed
        assert info.converged == 1 and info.iterations >= (1 if k > 128 else 0)
        assert 0.0 <= info.backward_error <= 1e-14
        assert np.abs(X - ref[:m]).max() <= 1e-11 * np.abs(ref[:m]).max()


def test_integer_problem(h):
    """Integer solution of an integer system (NaN in the unreferenced triangle):
    the refined X rounds to the exact X and is within 1e-12 of it."""
    rng = np.random.default_rng(5)
    m, n = 260, 33
    A, B, X = integer_kat("L", "U", "N", "N", m, n, rng)
    dA, dB = dev(A), dev(B)
    st, info = rb.trsm_mixed(h, "L", "U", "N", "N", m, n, 1.0, dA.data_ptr(), m, dB.data_ptr(),
                             m, block=64)
    assert st == S.success and info.converged == 1
    torch.cuda.synchronize()
    got = host(dB, B.shape)
    assert np.array_equal(np.rint(got), X)
    assert np.abs(got - X).max() <= 1e-12 * np.abs(X).max()


def test_fixed_steps_and_block_sizes(h):
    """tol <= 0 runs exactly max_refine steps (no measurement); the answer does
    not depend on the block size beyond the refinement's accuracy."""
    rng = np.random.default_rng(9)
    m, n = 700, 40
    A = oracle.diag_dominant_triangular(m, "L", rng, poison=0.0)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    outs = []
    for nb in (64, 128, 256, 700, 0):
        say("mixed block", nb)
        dA, dB = dev(A), dev(B)
        st, info = rb.trsm_mixed(h, "L", "L", "T", "N", m, n, 2.0, dA.data_ptr(), m,
                                 dB.data_ptr(), m, block=nb, max_refine=2, tol=0.0)
        assert st == S.success
        assert info.iterations == 2 and info.backward_error == -1.0
        torch.cuda.synchronize()
        X = host(dB, B.shape)
        assert eta_np("L", "L", "T", "N", 2.0, A, B, X) <= 1e-12
        outs.append(X)
    for X in outs[1:]:
        assert np.abs(X - outs[0]).max() <= 1e-12 * np.abs(outs[0]).max()


def test_argument_checks_and_alpha_zero(h):
    m = n = 8
    dA = torch.zeros(m * m, dtype=torch.float64, device="cuda")
    dB = torch.ones(m * n, dtype=torch.float64, device="cuda")
    L = rb.lib()
    one = ctypes.c_double(1.0)
    ap = ctypes.addressof(one)
    f = lambda *a: L.rocblas_amd_dtrsm_mixed(*a)
    assert f(None, 141, 121, 111, 131, m, n, ap, dA.data_ptr(), m, dB.data_ptr(), m, 0, 0, 3,
             1e-13, None) == S.invalid_handle
    assert f(h.ptr, 140, 121, 111, 131, m, n, ap, dA.data_ptr(), m, dB.data_ptr(), m, 0, 0, 3,
             1e-13, None) == S.invalid_value
    assert f(h.ptr, 141, 121, 111, 131, m, n, ap, dA.data_ptr(), m, dB.data_ptr(), m, 7, 0, 3,
             1e-13, None) == S.invalid_value
    assert f(h.ptr, 141, 121, 111, 131, m, n, ap, dA.data_ptr(), m - 1, dB.data_ptr(), m, 0, 0,
             3, 1e-13, None) == S.invalid_size
    assert f(h.ptr, 141, 121, 111, 131, m, n, ap, dA.data_ptr(), m, dB.data_ptr(), m, 0, -1, 3,
             1e-13, None) == S.invalid_size
    assert f(h.ptr, 141, 121, 111, 131, m, n, ap, None, m, dB.data_ptr(), m, 0, 0, 3, 1e-13,
             None) == S.invalid_pointer
    assert f(h.ptr, 141, 121, 111, 131, 0, n, ap, None, m, None, m, 0, 0, 3, 1e-13,
             None) == S.success
    st, info = rb.trsm_mixed(h, "L", "U", "N", "N", m, n, 0.0, 0, m, dB.data_ptr(), m)
    assert st == S.success
    torch.cuda.synchronize()
    assert torch.count_nonzero(dB).item() == 0


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
@pytest.mark.parametrize("diag", ["N", "U"])
def test_reported_eta_through_the_dma_residual(h, side, uplo, trans, diag):
    """At 2056 x 2048 the FP64 residual alpha B - op(A) X runs on the DMA GEMM
    (the triangle's dense part) plus the masked kernel (its 128-wide diagonal
    blocks, the last one 8 wide).  With no refinement the reported eta is the
    first low-precision solve's, far above rounding level, so it must equal a
    plain PyTorch float64 residual's to many digits: a missing or doubled
    block, or a wrong mask, would change it at O(1)."""
    from tests.gpu_util import eta_torch, logical
    m, n = (2056, 2048) if side == "L" else (2048, 2056)
    k = m if side == "L" else n
    say("mixed dma residual", side, uplo, trans, diag)
    g = torch.Generator(device="cuda").manual_seed(17 + ALL_CASES.index((side, uplo, trans)))
    dA = torch.randn(k * k, dtype=torch.float64, device="cuda", generator=g) / k ** 0.5
    dA.view(k, k).diagonal().uniform_(1.0, 2.0, generator=g)
    dB0 = torch.rand(m * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dB = dB0.clone()
    st, info = rb.trsm_mixed(h, side, uplo, trans, diag, m, n, 1.5, dA.data_ptr(), k,
                             dB.data_ptr(), m, block=256, max_refine=0, tol=1e-30)
    assert st == S.success and info.iterations == 0
    torch.cuda.synchronize()
    eta = eta_torch(side, uplo, trans, diag, 1.5, logical(dA, k, k, k), logical(dB0, m, m, n),
                    logical(dB, m, m, n))
    assert 1e-12 < eta < 1e-3, eta
    assert abs(info.backward_error - eta) <= 1e-6 * eta, (info.backward_error, eta)


@pytest.mark.parametrize("side,uplo,trans", [("L", "U", "N"), ("R", "L", "T")])
def test_first_solve_accuracy_per_form(h, side, uplo, trans):
    """The unrefined first solve (max_refine 0) carries the update form's
    accuracy: FP32 products keep the backward error below 2e-6, split-BF16x3
    products (three BF16 terms, about 16 bits) below 1e-4 and measurably above
    FP32's -- a single BF16 product (8 bits) would give about 1e-3, so a wrong
    split or a dropped cross term fails the bound.  Refinement then meets
    1e-12 within two steps."""
    rng = np.random.default_rng(31)
    m, n = (1024, 96) if side == "L" else (96, 1024)
    k = m if side == "L" else n
    # diagonal U[1,2], off-diagonal U[-1,1]/k: values that neither FP32 nor
    # BF16 halves hold exactly (small integers would split without error)
    A = np.asfortranarray(np.triu(rng.uniform(-1, 1, (k, k)) / k, 1) if uplo == "U"
                          else np.tril(rng.uniform(-1, 1, (k, k)) / k, -1))
    A[np.arange(k), np.arange(k)] = rng.uniform(1, 2, k)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    bounds = {"fp32": 2e-6, "bf16x3": 1e-4}
    first = {}
    for form, hi in bounds.items():
        say("mixed first solve", form, side, uplo, trans)
        dA, dB = dev(A), dev(B)
        st, info = rb.trsm_mixed(h, side, uplo, trans, "N", m, n, 1.0, dA.data_ptr(), k,
                                 dB.data_ptr(), m, block=128, max_refine=0, tol=0.0, form=form)
        assert st == S.success
        torch.cuda.synchronize()
        e0 = eta_np(side, uplo, trans, "N", 1.0, A, B, host(dB, B.shape))
        assert 0.0 < e0 <= hi, (form, e0)
        first[form] = e0
        dB = dev(B)
        st, info = rb.trsm_mixed(h, side, uplo, trans, "N", m, n, 1.0, dA.data_ptr(), k,
                                 dB.data_ptr(), m, block=128, max_refine=5, tol=1e-13, form=form)
        assert st == S.success and info.converged == 1 and info.iterations <= 2
        torch.cuda.synchronize()
        assert eta_np(side, uplo, trans, "N", 1.0, A, B, host(dB, B.shape)) <= 1e-12
    # 16-bit products are measurably coarser than FP32 ones
    assert first["bf16x3"] > 2.0 * first["fp32"], first
