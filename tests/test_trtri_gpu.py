"""GPU parity tests of rocblas_{s,d}trtri{,_batched,_strided_batched} through the C ABI.

Mirrors clients/include/blas3/testing_trtri.hpp: the bad-argument table, the
comparison against the host xTRTI2 (here the oracle's restatement) at the
reference's tolerance eps * max(10000, N) for double (clients/include/unit.hpp:
631-638), plus exact integer inverses, the cleared opposite triangle, in-place
calls and batched forms.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from tests.gpu_util import GpuHandle, dev, host, ptr_array, say
from tests.helpers import op_dense

pytestmark = pytest.mark.gpu
S = rb.rocblas_status
EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available()
    oracle.build()
    hd = GpuHandle()
    yield hd
    hd.close()


def tri_of(X, uplo, n, unit):
    T = np.triu(X[:n, :n]) if uplo == "U" else np.tril(X[:n, :n])
    if unit:
        np.fill_diagonal(T, 1.0)
    return T


def test_bad_args(h):
    """clients/include/blas3/testing_trtri.hpp:42-90"""
    say("trtri bad args")
    N = lda = 100
    dA = torch.zeros(N * lda, dtype=torch.float64, device="cuda")
    dI = torch.zeros(N * lda, dtype=torch.float64, device="cuda")
    L = rb.lib()
    A, I = dA.data_ptr(), dI.data_ptr()
    assert L.rocblas_dtrtri(h.ptr, 123, 131, N, A, lda, I, lda) == S.invalid_value
    assert L.rocblas_dtrtri(h.ptr, 121, 143, N, A, lda, I, lda) == S.invalid_value
    assert L.rocblas_dtrtri(h.ptr, 121, 131, -1, A, lda, I, lda) == S.invalid_size
    assert L.rocblas_dtrtri(h.ptr, 121, 131, N, A, lda - 1, I, lda) == S.invalid_size
    assert L.rocblas_dtrtri(h.ptr, 121, 131, N, A, lda, I, lda - 1) == S.invalid_size
    assert L.rocblas_dtrtri(None, 121, 131, N, A, lda, I, lda) == S.invalid_handle
    assert L.rocblas_dtrtri(h.ptr, 121, 131, N, None, lda, I, lda) == S.invalid_pointer
    assert L.rocblas_dtrtri(h.ptr, 121, 131, N, A, lda, None, lda) == S.invalid_pointer
    assert L.rocblas_dtrtri(h.ptr, 121, 131, 0, None, lda, None, lda) == S.success
    assert L.rocblas_dtrtri_strided_batched(h.ptr, 121, 131, N, A, lda, 0, I, lda, 0,
                                            -1) == S.invalid_size
    assert L.rocblas_dtrtri_batched(h.ptr, 121, 131, N, None, lda, None, lda, 0) == S.success
    torch.cuda.synchronize()


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_against_oracle(h, uplo, diag):
    rng = np.random.default_rng(41)
    for n in (1, 5, 32, 63, 64, 65, 130, 333):
        say("trtri", uplo, diag, n)
        A = oracle.diag_dominant_triangular(n, uplo, rng, lda=n + 3)  # NaN elsewhere
        if diag == "U":
            oracle.make_unit_diagonal(A, uplo, n)
        ref = A.copy(order="F")
        oracle.trtri(uplo, diag, n, ref)
        dA = dev(A)
        dI = torch.full(((n + 1) * n,), 7.0, dtype=torch.float64, device="cuda")
        st = rb.trtri(h, uplo, diag, n, dA.data_ptr(), n + 3, dI.data_ptr(), n + 1)
        assert st == S.success
        torch.cuda.synchronize()
        got = host(dI, (n + 1, n))
        R = tri_of(ref, uplo, n, diag == "U")
        G = got[:n, :n]
        tol = EPS64 * max(10000.0, n)
        assert np.abs(G - R).max() / max(1.0, np.abs(R).max()) <= tol
        # opposite triangle cleared, padding rows untouched, input unchanged
        off = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
        assert np.all(G[off] == 0.0)
        assert np.all(got[n:, :] == 7.0)
        assert np.array_equal(host(dA, A.shape), A, equal_nan=True)
        # invA * A = I
        T = op_dense(A, uplo, "N", diag, n)
        assert np.abs(G @ T - np.eye(n)).max() <= 1e-12


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_exact_integer_inverse(h, uplo):
    """I minus the unit sub/super-diagonal has the all-ones triangle as inverse:
    every value on the way is a small integer, so the result is exact."""
    n = 200
    A = np.zeros((n, n), order="F")
    np.fill_diagonal(A, 1.0)
    for i in range(n - 1):
        if uplo == "L":
            A[i + 1, i] = -1.0
        else:
            A[i, i + 1] = -1.0
    dA = dev(A)
    dI = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    assert rb.trtri(h, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
    torch.cuda.synchronize()
    want = np.tril(np.ones((n, n))) if uplo == "L" else np.triu(np.ones((n, n)))
    assert np.array_equal(host(dI, (n, n)), want)


def test_in_place_and_batched_forms(h):
    rng = np.random.default_rng(43)
    n, batch = 70, 5
    for uplo in "UL":
        say("trtri batched", uplo)
        As = [oracle.diag_dominant_triangular(n, uplo, rng, poison=0.0) for _ in range(batch)]
        singles = []
        for A in As:
            dA = dev(A)
            dI = torch.zeros(n * n, dtype=torch.float64, device="cuda")
            assert rb.trtri(h, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
            singles.append(dI)
        torch.cuda.synchronize()
        # in place: invA == A
        dA = dev(As[0])
        assert rb.trtri(h, uplo, "N", n, dA.data_ptr(), n, dA.data_ptr(), n) == S.success
        torch.cuda.synchronize()
        assert torch.equal(dA, singles[0])
        # strided batched
        stride = n * n + 11
        flatA = np.zeros(stride * batch)
        for b, A in enumerate(As):
            flatA[b * stride:b * stride + n * n] = A.ravel(order="F")
        dAs = torch.from_numpy(flatA).cuda()
        dIs = torch.zeros(stride * batch, dtype=torch.float64, device="cuda")
        assert rb.trtri_strided_batched(h, uplo, "N", n, dAs.data_ptr(), n, stride, dIs.data_ptr(),
                                        n, stride, batch) == S.success
        # pointer-array batched
        dAl = [dev(A) for A in As]
        dIl = [torch.zeros(n * n, dtype=torch.float64, device="cuda") for _ in As]
        pa, pi = ptr_array(dAl), ptr_array(dIl)
        assert rb.trtri_batched(h, uplo, "N", n, pa.data_ptr(), n, pi.data_ptr(), n,
                                batch) == S.success
        torch.cuda.synchronize()
        for b in range(batch):
            assert torch.equal(dIs[b * stride:b * stride + n * n], singles[b])
            assert torch.equal(dIl[b], singles[b])


def test_float32(h):
    rng = np.random.default_rng(47)
    for uplo in "UL":
        n = 90
        A = oracle.diag_dominant_triangular(n, uplo, rng, poison=0.0).astype(np.float32)
        A = np.asfortranarray(A)
        ref = A.copy(order="F")
        oracle.trtri(uplo, "N", n, ref)
        dA = dev(A)
        dI = torch.zeros(n * n, dtype=torch.float32, device="cuda")
        assert rb.trtri(h, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n,
                        precision="s") == S.success
        torch.cuda.synchronize()
        G = host(dI, (n, n))
        R = tri_of(ref, uplo, n, False)
        assert np.abs(G - R).max() / np.abs(R).max() <= EPS32 * max(1000.0, n)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_blocked_doubling_sizes(h, uplo, diag):
    """Orders that exercise the doubling steps of the blocked inversion: equal
    groups batched in one launch (n = 1024), a shorter last group at several
    levels (777, 2050), odd leading dimensions; NaN in the opposite triangle
    of A, a sentinel in invA's padding rows.  Checked against the oracle's
    xTRTI2 (777) and by ||invA T - I|| on the device (all)."""
    rng = np.random.default_rng(53)
    for n in (777, 1024, 2050):
        say("trtri blocked", uplo, diag, n)
        A = oracle.diag_dominant_triangular(n, uplo, rng, lda=n + 1)
        if diag == "U":
            oracle.make_unit_diagonal(A, uplo, n)
        dA = dev(A)
        dI = torch.full(((n + 3) * n,), 7.0, dtype=torch.float64, device="cuda")
        st = rb.trtri(h, uplo, diag, n, dA.data_ptr(), n + 1, dI.data_ptr(), n + 3)
        assert st == S.success
        torch.cuda.synchronize()
        got = host(dI, (n + 3, n))
        G = got[:n, :n]
        off = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
        assert np.all(G[off] == 0.0)
        assert np.all(got[n:, :] == 7.0)
        T = op_dense(A, uplo, "N", diag, n)
        Gt = torch.from_numpy(np.ascontiguousarray(G)).cuda()
        Tt = torch.from_numpy(np.ascontiguousarray(T)).cuda()
        err = float((Gt @ Tt - torch.eye(n, dtype=torch.float64, device="cuda")).abs().max())
        assert err <= 1e-12, (n, err)
        if n == 777:
            ref = A.copy(order="F")
            oracle.trtri(uplo, diag, n, ref)
            R = tri_of(ref, uplo, n, diag == "U")
            assert np.abs(G - R).max() / max(1.0, np.abs(R).max()) <= EPS64 * max(10000.0, n)


def test_blocked_batched_matches_single(h):
    """Strided and pointer-array batches of order 300 (two doubling levels per
    instance, instances as the GEMM batch) equal the single-instance calls."""
    rng = np.random.default_rng(59)
    n, batch = 300, 3
    for uplo in "LU":
        As = [oracle.diag_dominant_triangular(n, uplo, rng, poison=0.0) for _ in range(batch)]
        singles = []
        for A in As:
            dA = dev(A)
            dI = torch.zeros(n * n, dtype=torch.float64, device="cuda")
            assert rb.trtri(h, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
            singles.append(dI)
        say("trtri blocked batched", uplo)
        flatA = np.concatenate([A.ravel(order="F") for A in As])
        dAs = torch.from_numpy(flatA).cuda()
        dIs = torch.zeros(n * n * batch, dtype=torch.float64, device="cuda")
        assert rb.trtri_strided_batched(h, uplo, "N", n, dAs.data_ptr(), n, n * n, dIs.data_ptr(),
                                        n, n * n, batch) == S.success
        torch.cuda.synchronize()
        for b in range(batch):
            got = dIs[b * n * n:(b + 1) * n * n]
            assert torch.allclose(got, singles[b], rtol=0, atol=1e-13), b


def _bidiagonal(n, uplo, diag_value, poison):
    """Column-major I minus the unit sub (super) diagonal on the device, with
    `diag_value` on the diagonal and `poison` in the opposite triangle; its
    inverse is the all-ones triangle (with the diagonal taken as one)."""
    V = torch.eye(n, dtype=torch.float64, device="cuda") * diag_value  # V[col, row]
    i = torch.arange(n - 1, device="cuda")
    if uplo == "L":
        V[i, i + 1] = -1.0
        V.masked_fill_(torch.ones(n, n, dtype=torch.bool, device="cuda").tril(-1), poison)
    else:
        V[i + 1, i] = -1.0
        V.masked_fill_(torch.ones(n, n, dtype=torch.bool, device="cuda").triu(1), poison)
    return V.reshape(-1).contiguous()


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_split_products_exact_and_repeatable(h, uplo, diag):
    """Orders whose doubling products run on the DMA kernel with the diagonal
    blocks of the triangular factors masked in the same kernel: with the K
    steps of the top levels balanced over the workgroups (n = 2048, 4096; at
    n = 3000 the uneven top group of the upper case, a transposed operand and
    tiles cut by the edges), and on the masked kernels below (n = 1024).  The inverse of the bidiagonal
    I - J is the all-ones triangle: every partial sum is a small integer, so
    the result is exact whatever the split, and NaN in A's opposite triangle
    (and, for a unit diagonal, 3 on A's diagonal) must never enter.  A second
    call gives the same bits."""
    for n in (1024, 2048, 3000, 4096):
        say("trtri split", uplo, diag, n)
        dA = _bidiagonal(n, uplo, 3.0 if diag == "U" else 1.0, float("nan"))
        dI = torch.full((n * n,), 5.0, dtype=torch.float64, device="cuda")
        assert rb.trtri(h, uplo, diag, n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
        torch.cuda.synchronize()
        ones = torch.ones(n, n, dtype=torch.float64, device="cuda")
        want = (ones.triu() if uplo == "L" else ones.tril()).reshape(-1)  # [col, row] view
        assert torch.equal(dI, want), n
        dI2 = torch.full((n * n,), 5.0, dtype=torch.float64, device="cuda")
        assert rb.trtri(h, uplo, diag, n, dA.data_ptr(), n, dI2.data_ptr(), n) == S.success
        torch.cuda.synchronize()
        assert torch.equal(dI, dI2), n
        assert h.get_async_status() == S.success


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_split_products_accuracy(h, uplo):
    """Random well-conditioned triangles at n = 2048 and 4096 (split top level):
    ||invA T - I||_max on the device, and bitwise equal results on a repeat."""
    for n in (2048, 3000, 4096):
        say("trtri split accuracy", uplo, n)
        dA = torch.empty(n * n, dtype=torch.float64, device="cuda")
        assert rb.synth_tri(h, uplo, "uniform", n, 7, dA.data_ptr(), n, scale=1.0 / n) \
            == S.success
        dI = torch.empty_like(dA)
        dI2 = torch.empty_like(dA)
        assert rb.trtri(h, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
        assert rb.trtri(h, uplo, "N", n, dA.data_ptr(), n, dI2.data_ptr(), n) == S.success
        torch.cuda.synchronize()
        assert torch.equal(dI, dI2)
        T = (torch.tril if uplo == "L" else torch.triu)(dA.view(n, n).T)
        X = dI.view(n, n).T
        err = float((X @ T - torch.eye(n, dtype=torch.float64, device="cuda")).abs().max())
        assert err <= 1e-12, (n, err)
        del dA, dI, dI2, T, X
        torch.cuda.empty_cache()


def _small_workspace_handle():
    """A handle whose user-managed workspace (1 MiB) cannot hold the balanced
    products' partial tiles, so trtri's top level at n = 4096 splits its tiles'
    K ranges with in-order waits instead."""
    ht = GpuHandle()
    ws = torch.empty(1 << 17, dtype=torch.float64, device="cuda")
    assert rb.lib().rocblas_set_workspace(ht.ptr, ws.data_ptr(), ws.numel() * 8) == S.success
    return ht, ws


def test_trtri_workspace_query_and_fallback(h):
    """The size query asks for the balanced products' partial tiles (n = 4096:
    2 x 512 tiles of 128 x 128 doubles with 256 CUs); with a workspace too small
    for them the call still succeeds (the other split), exact on the bidiagonal
    and accurate on a random triangle."""
    n = 4096
    sz = ctypes.c_size_t(0)
    L = rb.lib()
    assert L.rocblas_start_device_memory_size_query(h.ptr) == S.success
    st = rb.trtri(h, "L", "N", n, 8, n, 16, n)
    assert st in (S.size_unchanged, S.size_increased)
    assert L.rocblas_stop_device_memory_size_query(h.ptr, ctypes.byref(sz)) == S.success
    # the top level: 16 x 16 tiles whose K steps of 8 sum to 16 x 16 x 136
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    workgroups = min(2 * cus, 16 * 16 * 136 // 16)
    assert sz.value >= 2 * workgroups * 128 * 128 * 8, sz.value
    ht, ws = _small_workspace_handle()
    try:
        for uplo in "LU":
            say("trtri small workspace", uplo)
            dA = _bidiagonal(n, uplo, 1.0, 0.0)
            dI = torch.full((n * n,), 5.0, dtype=torch.float64, device="cuda")
            assert rb.trtri(ht, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
            torch.cuda.synchronize()
            ones = torch.ones(n, n, dtype=torch.float64, device="cuda")
            assert torch.equal(dI, (ones.triu() if uplo == "L" else ones.tril()).reshape(-1))
            dR = torch.empty(n * n, dtype=torch.float64, device="cuda")
            assert rb.synth_tri(ht, uplo, "uniform", n, 11, dR.data_ptr(), n, scale=1.0 / n) \
                == S.success
            assert rb.trtri(ht, uplo, "N", n, dR.data_ptr(), n, dI.data_ptr(), n) == S.success
            torch.cuda.synchronize()
            T = (torch.tril if uplo == "L" else torch.triu)(dR.view(n, n).T)
            err = float((dI.view(n, n).T @ T - torch.eye(n, dtype=torch.float64,
                                                          device="cuda")).abs().max())
            assert err <= 1e-12, err
            assert ht.get_async_status() == S.success
    finally:
        ht.close()
        del ws


def test_split_wait_timeout_reported_and_healed(h):
    """The split products' bounded waits (the in-order form, taken when the
    workspace cannot hold the balanced products' partial tiles), forced to
    give up through the handle's test hook (every kernel still runs to its
    end): the handle reports the error once, the next trtri call returns it
    without work, and after the hook is cleared the tile counters are back at
    zero, so the result is the exact one again."""
    n, uplo = 4096, "L"
    dA = _bidiagonal(n, uplo, 1.0, 0.0)
    ones = torch.ones(n, n, dtype=torch.float64, device="cuda")
    want = ones.triu().reshape(-1)
    ht, ws = _small_workspace_handle()
    try:
        assert ht.set_chain_test_timeout(1) == S.success
        dI = torch.zeros(n * n, dtype=torch.float64, device="cuda")
        say("trtri split timeout forced")
        assert rb.trtri(ht, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
        torch.cuda.synchronize()
        assert ht.get_async_status() == S.internal_error
        assert ht.get_async_status() == S.success
        assert rb.trtri(ht, uplo, "N", n, dA.data_ptr(), n, dI.data_ptr(), n) == S.success
        torch.cuda.synchronize()
        assert ht.set_chain_test_timeout(0) == S.success
        dJ = torch.full((n * n,), 5.0, dtype=torch.float64, device="cuda")
        assert rb.trtri(ht, uplo, "N", n, dA.data_ptr(), n, dJ.data_ptr(), n) \
            == S.internal_error
        torch.cuda.synchronize()
        assert bool((dJ == 5.0).all())
        say("trtri split after timeout")
        assert rb.trtri(ht, uplo, "N", n, dA.data_ptr(), n, dJ.data_ptr(), n) == S.success
        torch.cuda.synchronize()
        assert torch.equal(dJ, want)
        assert ht.get_async_status() == S.success
    finally:
        ht.close()
        del ws
