"""GPU parity tests of rocblas_{s,d}trsm{,_batched,_strided_batched} through the C ABI.

Mirrors the reference's test strategy (clients/include/blas3/testing_trsm*.hpp,
clients/gtest/trsm_gtest.yaml): the bad-argument status table, quick returns,
alpha == 0, the known-X forward-error/residual check at 40*eps*M, batched and
strided-batched forms, repeatability and graph capture -- plus exact checks the
reference lacks: exact-rational golden vectors and bit-exact integer problems.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from rocblas_amd import synth
from tests.gpu_util import (GpuHandle, dev, eta_torch, gpu_trsm, host, logical, ptr_array, say)
from tests.helpers import ALL_CASES, integer_kat, load_rational_cases

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps
S = rb.rocblas_status


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    rb.lib()
    oracle.build()
    hd = GpuHandle()
    yield hd
    hd.close()


# ---------------------------------------------------------------- statuses --

@pytest.mark.parametrize("mode", ["host", "device"])
def test_bad_arg_table(h, mode):
    """clients/include/blas3/testing_trsm.hpp:33-186, both pointer modes."""
    say("bad args", mode)
    M = N = lda = ldb = 100
    dA = torch.zeros(lda * M, dtype=torch.float64, device="cuda")
    dB = torch.zeros(ldb * N, dtype=torch.float64, device="cuda")
    one_h, zero_h = ctypes.c_double(1.0), ctypes.c_double(0.0)
    one_d = torch.ones(1, dtype=torch.float64, device="cuda")
    zero_d = torch.zeros(1, dtype=torch.float64, device="cuda")
    h.set_pointer_mode(rb.rocblas_pointer_mode.device if mode == "device"
                       else rb.rocblas_pointer_mode.host)
    try:
        alpha = one_d.data_ptr() if mode == "device" else ctypes.addressof(one_h)
        zero = zero_d.data_ptr() if mode == "device" else ctypes.addressof(zero_h)
        L = rb.lib()
        A, B = dA.data_ptr(), dB.data_ptr()
        s, u, t, d = 141, 121, 111, 131

        def call(*a):
            r = L.rocblas_dtrsm(*a)
            torch.cuda.synchronize()
            return r

        assert call(h.ptr, 143, u, t, d, M, N, alpha, A, lda, B, ldb) == S.invalid_value
        assert call(h.ptr, s, 143, t, d, M, N, alpha, A, lda, B, ldb) == S.invalid_value
        assert call(h.ptr, s, u, 143, d, M, N, alpha, A, lda, B, ldb) == S.invalid_value
        assert call(h.ptr, s, u, t, 143, M, N, alpha, A, lda, B, ldb) == S.invalid_value
        assert call(h.ptr, s, u, t, d, -1, N, alpha, A, lda, B, ldb) == S.invalid_size
        assert call(h.ptr, s, u, t, d, M, -1, alpha, A, lda, B, ldb) == S.invalid_size
        assert call(h.ptr, s, u, t, d, M, N, alpha, A, lda, B, M - 1) == S.invalid_size
        assert call(h.ptr, 141, u, t, d, M, N, alpha, A, M - 1, B, ldb) == S.invalid_size
        assert call(h.ptr, 142, u, t, d, M, N, alpha, A, N - 1, B, ldb) == S.invalid_size
        assert call(None, s, u, t, d, M, N, alpha, A, lda, B, ldb) == S.invalid_handle
        assert call(h.ptr, s, u, t, d, M, N, None, A, lda, B, ldb) == S.invalid_pointer
        if mode == "host":
            assert call(h.ptr, s, u, t, d, M, N, alpha, None, lda, B, ldb) == S.invalid_pointer
        assert call(h.ptr, s, u, t, d, M, N, alpha, A, lda, None, ldb) == S.invalid_pointer
        # alpha == 0: A may be null, B := 0
        dB.fill_(7.0)
        assert call(h.ptr, s, u, t, d, M, N, zero, None, lda, B, ldb) == S.success
        assert float(dB.abs().max()) == 0.0
        # m == 0 / n == 0: every pointer may be null
        assert call(h.ptr, s, u, t, d, 0, N, None, None, lda, None, ldb) == S.success
        assert call(h.ptr, s, u, t, d, M, 0, None, None, lda, None, ldb) == S.success
        # batched forms: batch_count < 0 is invalid, == 0 is a quick return
        assert L.rocblas_dtrsm_strided_batched(h.ptr, s, u, t, d, M, N, alpha, A, lda, 0, B,
                                               ldb, 0, -1) == S.invalid_size
        assert L.rocblas_dtrsm_strided_batched(h.ptr, s, u, t, d, M, N, None, None, lda, 0,
                                               None, ldb, 0, 0) == S.success
        assert L.rocblas_dtrsm_batched(h.ptr, s, u, t, d, M, N, alpha, None, lda, None, ldb,
                                       0) == S.success
        assert L.rocblas_dtrsm_batched(h.ptr, s, u, t, d, M, N, alpha, A, lda, None, ldb,
                                       1) == S.invalid_pointer
    finally:
        h.set_pointer_mode(rb.rocblas_pointer_mode.host)


def test_handle_services(h):
    L = rb.lib()
    sz = ctypes.c_size_t()
    assert L.rocblas_get_device_memory_size(h.ptr, ctypes.byref(sz)) == S.success
    assert sz.value == 128 << 20
    assert L.rocblas_is_managing_device_memory(h.ptr)
    mode = ctypes.c_int32(-1)
    assert L.rocblas_get_pointer_mode(h.ptr, ctypes.byref(mode)) == S.success
    assert mode.value == 0
    st = ctypes.c_void_p()
    assert L.rocblas_get_stream(h.ptr, ctypes.byref(st)) == S.success
    assert (st.value or 0) == torch.cuda.current_stream().cuda_stream
    # size-query protocol (library/src/handle.cpp:449-483): trsm reports, does no work
    assert L.rocblas_stop_device_memory_size_query(h.ptr, ctypes.byref(sz)) \
        == S.size_query_mismatch
    assert L.rocblas_start_device_memory_size_query(h.ptr) == S.success
    assert L.rocblas_start_device_memory_size_query(h.ptr) == S.size_query_mismatch
    assert L.rocblas_is_device_memory_size_query(h.ptr)
    dB = torch.full((64 * 64,), 3.0, dtype=torch.float64, device="cuda")
    st2 = rb.trsm(h, "L", "U", "N", "N", 64, 64, 1.0, None, 64, dB.data_ptr(), 64)
    assert st2 in (S.size_unchanged, S.size_increased)
    st2 = rb.trsm(h, "L", "U", "N", "N", 4096, 4096, 1.0, None, 4096, None, 4096)
    assert st2 in (S.size_unchanged, S.size_increased)
    assert L.rocblas_stop_device_memory_size_query(h.ptr, ctypes.byref(sz)) == S.success
    torch.cuda.synchronize()
    assert float(dB.min()) == 3.0 and float(dB.max()) == 3.0  # untouched
    # user-managed / user-owned workspace and back
    h2 = GpuHandle()
    assert L.rocblas_set_device_memory_size(h2.ptr, 1 << 20) == S.success
    assert L.rocblas_is_user_managing_device_memory(h2.ptr)
    buf = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    assert L.rocblas_set_workspace(h2.ptr, buf.data_ptr(), buf.numel()) == S.success
    assert not L.rocblas_is_managing_device_memory(h2.ptr)
    assert L.rocblas_set_workspace(h2.ptr, None, 0) == S.success
    assert L.rocblas_is_managing_device_memory(h2.ptr)
    h2.close()


# ------------------------------------------------------------ exact checks --

def test_golden_rational_vectors(h):
    cases = load_rational_cases()
    say("golden rational", len(cases))
    for c in cases:
        m, n = c["m"], c["n"]
        st, X = gpu_trsm(h, c["side"], c["uplo"], c["trans"], c["diag"], m, n, c["alpha"],
                         c["A"], c["B"])
        assert st == S.success
        got, want = X[:m, :n], c["X"][:m, :n]
        assert np.all(np.isfinite(got)), c
        scale = max(1.0, np.abs(want).max())
        assert np.abs(got - want).max() <= 64 * EPS64 * scale, (c["side"], c["uplo"],
                                                               c["trans"], c["diag"], m, n)
        if c["ldb"] > m:
            assert np.all(np.isnan(X[m:, :])), "rows past m must be untouched"


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_integer_kat_bitwise(h, side, uplo, trans):
    """Exactly solvable problems: every intermediate of substitution and of the
    GEMM updates is an exact small integer or dyadic, so X is reproduced bit for bit
    on both the small kernel (k <= 64) and the recursive blocked path."""
    rng = np.random.default_rng(101)
    for diag in "NU":
        for (m, n) in [(1, 1), (5, 3), (17, 9), (33, 40), (64, 16), (65, 7), (130, 33),
                       (200, 150), (300, 64)]:
            for alpha in (1.0, -0.5):
                say("kat", side, uplo, trans, diag, m, n, alpha)
                k = m if side == "L" else n
                A, B, X = integer_kat(side, uplo, trans, diag, m, n, rng, lda=k + 2, ldb=m + 3,
                                      alpha=alpha)
                st, got = gpu_trsm(h, side, uplo, trans, diag, m, n, alpha, A, B)
                assert st == S.success
                assert np.array_equal(got[:m, :n], X[:m, :n]), (diag, m, n, alpha)


# ------------------------------------------------- reference's acceptance --

@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_reference_known_x_procedure(h, side, uplo, trans):
    """clients/include/blas3/testing_trsm.hpp:297-315, 433-448: diagonally dominant A
    (unit-diagonal variant via make_unit_diagonal), random X, B := op(A) X / alpha,
    forward error and residual <= 40*eps*M."""
    rng = np.random.default_rng(202)
    for (m, n) in [(7, 5), (32, 16), (64, 64), (65, 130), (257, 100), (600, 300)]:
        for diag in "NU":
            say("known-x", side, uplo, trans, diag, m, n)
            alpha = 2.5
            k = m if side == "L" else n
            A = oracle.diag_dominant_triangular(k, uplo, rng, lda=k + 1)
            if diag == "U":
                oracle.make_unit_diagonal(A, uplo, k)
            X = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
            B = X.copy(order="F")
            oracle.trmm(side, uplo, trans, diag, m, n, 1.0 / alpha, A, B)
            st, got = gpu_trsm(h, side, uplo, trans, diag, m, n, alpha, A, B)
            assert st == S.success
            fwd = oracle.norm1_rel_diff(m, n, got, X)
            assert fwd <= 40 * EPS64 * m, fwd
            R = np.asfortranarray(got[:m, :n])
            oracle.trmm(side, uplo, trans, diag, m, n, 1.0 / alpha, A, R)
            assert oracle.norm1_rel_diff(m, n, R, B) <= 40 * EPS64 * m


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_float32(h, side, uplo, trans):
    rng = np.random.default_rng(303)
    for (m, n) in [(16, 16), (48, 20), (100, 70), (300, 129)]:
        say("strsm", side, uplo, trans, m, n)
        k = m if side == "L" else n
        A = oracle.diag_dominant_triangular(k, uplo, rng)
        X = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        B = X.copy(order="F")
        oracle.trmm(side, uplo, trans, "N", m, n, 1.0, A, B)
        A32 = np.asfortranarray(A.astype(np.float32))
        B32 = np.asfortranarray(B.astype(np.float32))
        st, got = gpu_trsm(h, side, uplo, trans, "N", m, n, 1.0, A32, B32, precision="s")
        assert st == S.success
        got64 = np.asfortranarray(got.astype(np.float64))
        assert oracle.norm1_rel_diff(m, n, got64, X) <= 40 * EPS32 * m
        eta = oracle.backward_error(side, uplo, trans, "N", m, n, 1.0, A32, B32, got)
        assert eta <= 1e-5


# --------------------------------------------------------- batched forms --

@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("L", "U", "T"), ("R", "L", "T"),
                                              ("R", "U", "N")])
def test_small_to_panel_handover_batched_matches_single(h, side, uplo, trans):
    """Triangles of 33..64 rows go to the small LDS kernel or to the panel
    solve by k and the vector count (48 < k, or 32 < k with at most 16
    vectors): shapes on both sides of each boundary, a strided batch of three
    against the single-instance calls bit for bit, and against the oracle."""
    rng = np.random.default_rng(4040 + ALL_CASES.index((side, uplo, trans)))
    for (k, nv) in [(33, 16), (33, 17), (48, 16), (48, 17), (49, 64), (64, 1), (64, 33),
                    (40, 70)]:
        m, n = (k, nv) if side == "L" else (nv, k)
        say("handover", side, uplo, trans, m, n)
        batch, alpha = 3, -0.75
        As = [oracle.diag_dominant_triangular(k, uplo, rng) for _ in range(batch)]
        Bs = [np.asfortranarray(rng.uniform(-1, 1, (m, n))) for _ in range(batch)]
        dA = torch.from_numpy(np.concatenate([A.ravel(order="F") for A in As])).cuda()
        dB = torch.from_numpy(np.concatenate([B.ravel(order="F") for B in Bs])).cuda()
        st = rb.trsm_strided_batched(h, side, uplo, trans, "N", m, n, alpha, dA.data_ptr(), k,
                                     k * k, dB.data_ptr(), m, m * n, batch)
        assert st == S.success
        torch.cuda.synchronize()
        out = dB.cpu().numpy()
        for b in range(batch):
            st1, single = gpu_trsm(h, side, uplo, trans, "N", m, n, alpha, As[b], Bs[b])
            assert st1 == S.success
            got = np.asfortranarray(out[b * m * n:(b + 1) * m * n].reshape(n, m).T)
            assert np.array_equal(got, single), (k, nv, b)
            ref = Bs[b].copy(order="F")
            oracle.trsm(side, uplo, trans, "N", m, n, alpha, As[b], ref)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (k, nv, b)


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_strided_and_pointer_batched_match_single(h, side, uplo, trans):
    rng = np.random.default_rng(404)
    for (m, n, batch) in [(32, 16, 7), (17, 33, 3), (100, 20, 5), (1, 9, 4)]:
        say("batched", side, uplo, trans, m, n, batch)
        k = m if side == "L" else n
        lda, ldb = k + 1, m + 2
        sa, sb = lda * k + 3, ldb * n + 5
        A = np.full(sa * batch, np.nan)
        B = rng.uniform(-1, 1, sb * batch)
        for b in range(batch):
            Ab = oracle.diag_dominant_triangular(k, uplo, rng, lda=lda)
            A[b * sa:b * sa + lda * k] = Ab.ravel(order="F")
        dA = torch.from_numpy(A).cuda()
        dB = torch.from_numpy(B.copy()).cuda()
        st = rb.trsm_strided_batched(h, side, uplo, trans, "N", m, n, 0.5, dA.data_ptr(), lda, sa,
                                     dB.data_ptr(), ldb, sb, batch)
        assert st == S.success
        # pointer-array form on separate copies
        As = [torch.from_numpy(A[b * sa:b * sa + lda * k].copy()).cuda() for b in range(batch)]
        Bs = [torch.from_numpy(B[b * sb:b * sb + ldb * n].copy()).cuda() for b in range(batch)]
        pa, pb = ptr_array(As), ptr_array(Bs)
        st = rb.trsm_batched(h, side, uplo, trans, "N", m, n, 0.5, pa.data_ptr(), lda,
                             pb.data_ptr(), ldb, batch)
        assert st == S.success
        torch.cuda.synchronize()
        Bout = dB.cpu().numpy()
        for b in range(batch):
            Ab = np.asfortranarray(A[b * sa:b * sa + lda * k].reshape(k, lda).T)
            Bb = np.asfortranarray(B[b * sb:b * sb + ldb * n].reshape(n, ldb).T)
            st1, single = gpu_trsm(h, side, uplo, trans, "N", m, n, 0.5, Ab, Bb)
            got = np.asfortranarray(Bout[b * sb:b * sb + ldb * n].reshape(n, ldb).T)
            assert np.array_equal(got, single), b
            assert np.array_equal(host(Bs[b], (ldb, n)), single), b
            # gaps between instances untouched
            assert np.array_equal(Bout[b * sb + ldb * n:(b + 1) * sb],
                                  B[b * sb + ldb * n:(b + 1) * sb])
            ref = Bb.copy(order="F")
            oracle.trsm(side, uplo, trans, "N", m, n, 0.5, Ab, ref)
            assert oracle.norm1_rel_diff(m, n, got, ref) <= 40 * EPS64 * max(m, n)


# ---------------------------------------------------- pointer mode / alpha --

@pytest.mark.parametrize("m,n", [(40, 24), (200, 90)])
def test_device_pointer_mode_matches_host(h, m, n):
    rng = np.random.default_rng(505)
    for side, uplo, trans in ALL_CASES:
        say("device alpha", side, uplo, trans, m, n)
        k = m if side == "L" else n
        A = oracle.diag_dominant_triangular(k, uplo, rng)
        B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        _, want = gpu_trsm(h, side, uplo, trans, "N", m, n, -1.5, A, B)
        dA, dB = dev(A), dev(B)
        alpha_d = torch.tensor([-1.5], dtype=torch.float64, device="cuda")
        h.set_pointer_mode(rb.rocblas_pointer_mode.device)
        try:
            st = rb.trsm(h, side, uplo, trans, "N", m, n, None, dA.data_ptr(), k, dB.data_ptr(), m,
                         alpha_ptr=alpha_d.data_ptr())
            assert st == S.success
            torch.cuda.synchronize()
            assert np.array_equal(host(dB, (m, n)), want)
            # device alpha == 0 with A == nullptr: B := 0, A never read
            alpha_d.zero_()
            st = rb.trsm(h, side, uplo, trans, "N", m, n, None, None, k, dB.data_ptr(), m,
                         alpha_ptr=alpha_d.data_ptr())
            assert st == S.success
            torch.cuda.synchronize()
            assert float(dB.abs().max()) == 0.0
        finally:
            h.set_pointer_mode(rb.rocblas_pointer_mode.host)


def test_unreferenced_memory_untouched_and_unread(h):
    """NaN in the other triangle, on a unit diagonal and in B's ldb padding never
    reaches the solution, and the padding is left as it was."""
    rng = np.random.default_rng(606)
    for side, uplo, trans in ALL_CASES:
        for (m, n) in [(20, 12), (150, 70)]:
            k = m if side == "L" else n
            A = oracle.diag_dominant_triangular(k, uplo, rng, lda=k + 3)  # NaN elsewhere
            oracle.make_unit_diagonal(A, uplo, k)  # NaN diagonal, diag=U
            B = np.full((m + 4, n), np.nan, order="F")
            B[:m] = rng.uniform(-1, 1, (m, n))
            st, got = gpu_trsm(h, side, uplo, trans, "U", m, n, 1.0, A, B)
            assert st == S.success
            assert np.all(np.isfinite(got[:m])) and np.all(np.isnan(got[m:]))


def test_repeatability_bitwise(h):
    """clients/gtest/trsm_gtest.yaml:1310-1334: identical results run to run."""
    rng = np.random.default_rng(707)
    for (m, n) in [(32, 16), (500, 300)]:
        A = synth.tri_uniform(m, "U", 1)
        B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        outs = [gpu_trsm(h, "L", "U", "T", "N", m, n, 1.0, A, B)[1] for _ in range(3)]
        assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_graph_capture_replays_match_eager(h):
    """Host-pointer-mode calls captured into a HIP graph replay to the eager result
    (clients/common/client_utility.cpp:354-385), with eager work between replays."""
    rng = np.random.default_rng(808)
    m, n = 300, 100
    A = synth.tri_uniform(m, "L", 3)
    B0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    _, want = gpu_trsm(h, "L", "L", "N", "N", m, n, 2.0, A, B0)
    dA, dB0 = dev(A), dev(B0)
    dB = dB0.clone()
    s = torch.cuda.Stream()
    hg = GpuHandle()
    rb.check(hg.set_stream(s.cuda_stream))
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        st = rb.trsm(hg, "L", "L", "N", "N", m, n, 2.0, dA.data_ptr(), m, dB.data_ptr(), m)
    assert st == S.success
    for rep in range(4):
        dB.copy_(dB0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(host(dB, (m, n)), want), rep
        _ = (dB0 * 2).sum().item()  # eager work between replays
    hg.close()


@pytest.mark.parametrize("case", ["chained", "recursive", "device_alpha"])
def test_graph_capture_multi_launch_chained_device_mode(h, case):
    """Graph capture of the paths the one-launch test above does not reach: the
    chained solve (its counters must be back at zero for every replay), the
    recursive path (scale, GEMM updates and leaf launches), and device
    pointer mode (alpha read on the device, no synchronisation).  Every replay,
    with eager work between replays, equals the eager result bit for bit."""
    rng = np.random.default_rng(811)
    m, n, uplo, trans = {"chained": (512, 256, "L", "N"), "recursive": (1500, 700, "U", "T"),
                         "device_alpha": (700, 40, "L", "T")}[case]
    say("graph", case, m, n)
    A = synth.tri_uniform(m, uplo, 5)
    B0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    _, want = gpu_trsm(h, "L", uplo, trans, "N", m, n, -1.5, A, B0)
    dA, dB0 = dev(A), dev(B0)
    dB = dB0.clone()
    s = torch.cuda.Stream()
    hg = GpuHandle()
    rb.check(hg.set_stream(s.cuda_stream))
    alpha_d = torch.full((1,), -1.5, dtype=torch.float64, device="cuda")
    if case == "device_alpha":
        rb.check(hg.set_pointer_mode(rb.rocblas_pointer_mode.device))
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        if case == "device_alpha":
            st = rb.trsm(hg, "L", uplo, trans, "N", m, n, None, dA.data_ptr(), m, dB.data_ptr(),
                         m, alpha_ptr=alpha_d.data_ptr())
        else:
            st = rb.trsm(hg, "L", uplo, trans, "N", m, n, -1.5, dA.data_ptr(), m, dB.data_ptr(),
                         m)
    assert st == S.success
    for rep in range(4):
        dB.copy_(dB0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(host(dB, (m, n)), want), rep
        _ = (dB0 * 2).abs().sum().item()
        if case == "chained":
            # an eager chained solve on the same handle between replays
            dE = dB0.clone()
            assert rb.trsm(hg, "L", uplo, trans, "N", m, n, -1.5, dA.data_ptr(), m,
                           dE.data_ptr(), m) == S.success
            torch.cuda.synchronize()
            assert np.array_equal(host(dE, (m, n)), want)
    assert hg.get_async_status() == S.success
    hg.close()


def test_chained_wait_timeout_is_reported_and_heals(h):
    """A chained launch whose waits give up (forced from the host: a few polls
    and an unreachable target, so every kernel still runs to its end) leaves a
    wrong result; the handle reports it through rocblas_amd_get_async_status
    and through the next launching call, which does no work; the counters are
    back at zero, so the following calls -- with more vector panels than the
    failed launch used -- give the normal result."""
    rng = np.random.default_rng(812)
    m, n = 512, 256
    A = synth.tri_uniform(m, "L", 8)
    B0 = np.asfortranarray(rng.uniform(-1, 1, (m, n + 512)))
    _, want = gpu_trsm(h, "L", "L", "N", "N", m, n, 1.0, A, B0[:, :n].copy(order="F"))
    _, want_wide = gpu_trsm(h, "L", "L", "N", "N", m, n + 512, 1.0, A, B0)
    dA = dev(A)
    ht = GpuHandle()
    try:
        # 1: reported by the query, once
        assert ht.set_chain_test_timeout(1) == S.success
        dB = dev(B0[:, :n].copy(order="F"))
        say("chain timeout forced (query)")
        assert rb.trsm(ht, "L", "L", "N", "N", m, n, 1.0, dA.data_ptr(), m, dB.data_ptr(), m)             == S.success
        assert ht.get_async_status() == S.internal_error
        assert ht.get_async_status() == S.success
        # 2: reported by the next launching call, which leaves B alone
        dB = dev(B0[:, :n].copy(order="F"))
        say("chain timeout forced (next call)")
        assert rb.trsm(ht, "L", "L", "N", "N", m, n, 1.0, dA.data_ptr(), m, dB.data_ptr(), m)             == S.success
        torch.cuda.synchronize()
        assert ht.set_chain_test_timeout(0) == S.success
        dC = dev(B0[:, :n].copy(order="F"))
        assert rb.trsm(ht, "L", "L", "N", "N", m, n, 1.0, dA.data_ptr(), m, dC.data_ptr(), m)             == S.internal_error
        torch.cuda.synchronize()
        assert np.array_equal(host(dC, (m, n)), B0[:, :n])
        # 3: healed -- same width, then more panels than the failed launch had
        say("chain after timeout")
        assert rb.trsm(ht, "L", "L", "N", "N", m, n, 1.0, dA.data_ptr(), m, dC.data_ptr(), m)             == S.success
        torch.cuda.synchronize()
        assert np.array_equal(host(dC, (m, n)), want)
        dW = dev(B0)
        assert rb.trsm(ht, "L", "L", "N", "N", m, n + 512, 1.0, dA.data_ptr(), m, dW.data_ptr(),
                       m) == S.success
        torch.cuda.synchronize()
        assert np.array_equal(host(dW, (m, n + 512)), want_wide)
        assert ht.get_async_status() == S.success
    finally:
        ht.close()


# --------------------------------------------------------- workload sizes --

def test_batched_workload_size_all_cases(h):
    """The strided-batched workload: 20000 instances, m=32, n=16, every side/uplo/trans
    case, A = (diag U[1,2], off-diagonal U[-1,1]/m) / m, B = U[-1,1]; per-instance
    backward error and agreement with the CPU oracle."""
    batch, m, n = 20000, 32, 16
    for side, uplo, trans in ALL_CASES:
        say("workload batch", side, uplo, trans)
        k = m if side == "L" else n
        A = synth.tri_uniform(k, uplo, 0, batch=batch, scale=1.0 / k)
        Bm = synth.rhs_uniform(m, n, 0, batch=batch)
        Af = np.ascontiguousarray(np.swapaxes(A, 1, 2)).ravel()
        Bf = np.ascontiguousarray(np.swapaxes(Bm, 1, 2)).ravel()
        dA = torch.from_numpy(Af).cuda()
        dB = torch.from_numpy(Bf.copy()).cuda()
        st = rb.trsm_strided_batched(h, side, uplo, trans, "N", m, n, 1.0, dA.data_ptr(), k, k * k,
                                     dB.data_ptr(), m, m * n, batch)
        assert st == S.success
        got = dB.cpu().numpy()
        ref = Bf.copy()
        oracle.trsm_batched_omp(side, uplo, trans, "N", m, n, 1.0, Af, k, k * k, ref, m, m * n,
                                batch)
        Xg = got.reshape(batch, n, m).transpose(0, 2, 1)
        Xr = ref.reshape(batch, n, m).transpose(0, 2, 1)
        rel = np.abs(Xg - Xr).max(axis=(1, 2)) / np.abs(Xr).max(axis=(1, 2))
        assert rel.max() <= 1e-12
        # per-instance backward error, vectorised float64 (residual is O(m^2 n))
        Ad = A[:, :k, :k].copy()
        Ad = np.where(np.isnan(Ad), 0.0, Ad)
        Ad = np.triu(Ad) if uplo == "U" else np.tril(Ad)
        op = np.swapaxes(Ad, 1, 2) if trans != "N" else Ad
        B0 = Bm[:, :m, :n]
        R = B0 - (op @ Xg if side == "L" else Xg @ op)
        n1 = lambda M: np.abs(M).sum(axis=1).max(axis=1)
        eta = n1(R) / (n1(op) * n1(Xg) + n1(B0))
        assert eta.max() <= 1e-14


def test_large_single_solve_backward_error(h):
    """m = n = 4096, upper, N(0,1)/sqrt(m) off-diagonal (condition number ~1e3):
    backward error from a plain PyTorch float64 residual on the device."""
    m = n = 4096
    say("large", m)
    A = synth.tri_normal(m, "U", 0)
    Bm = synth.rhs_uniform(m, n, 0)
    dA, dB0 = dev(A), dev(Bm)
    dB = dB0.clone()
    st = rb.trsm(h, "L", "U", "N", "N", m, n, 1.0, dA.data_ptr(), m, dB.data_ptr(), m)
    assert st == S.success
    torch.cuda.synchronize()
    eta = eta_torch("L", "U", "N", "N", 1.0, logical(dA, m, m, m), logical(dB0, m, m, n),
                    logical(dB, m, m, n))
    say("large eta", eta)
    assert eta <= 1e-14


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_chained_single_solve(h, side, uplo, trans):
    """One instance with a long triangle and a narrow right-hand side runs as
    the chained launch (a workgroup per 64-row block and 16-vector panel,
    blocks chained through completion counters): one vector (the trsv case),
    ragged panels, k not a multiple of 64, unit and non-unit diagonals.  The
    result meets the oracle at the reference's 40*eps*k bound and equals, bit
    for bit, the same instance solved inside a batch of two (which takes the
    one-workgroup-per-panel kernel for k <= 512)."""
    rng = np.random.default_rng(909)
    for (k, nv, diag) in [(65, 1, "N"), (130, 17, "U"), (512, 256, "N"), (777, 3, "N"),
                          (2049, 40, "U"), (4096, 1, "N")]:
        m, n = (k, nv) if side == "L" else (nv, k)
        say("chain", side, uplo, trans, diag, m, n)
        A = oracle.diag_dominant_triangular(k, uplo, rng, lda=k + 1)
        if diag == "U":
            oracle.make_unit_diagonal(A, uplo, k)
        B = np.asfortranarray(rng.uniform(-1, 1, (m + 2, n)))
        st, got = gpu_trsm(h, side, uplo, trans, diag, m, n, 0.75, A, B)
        assert st == S.success
        ref = B.copy(order="F")
        oracle.trsm(side, uplo, trans, diag, m, n, 0.75, A, ref)
        assert oracle.norm1_rel_diff(m, n, got, ref) <= 40 * EPS64 * k
        assert np.array_equal(got[m:], B[m:]), "rows past m must be untouched"
        if k <= 512:
            lda, ldb = A.shape[0], B.shape[0]
            sa, sb = lda * k, ldb * n
            dA = torch.from_numpy(np.concatenate([A.ravel(order="F")] * 2)).cuda()
            dB = torch.from_numpy(np.concatenate([B.ravel(order="F")] * 2)).cuda()
            st = rb.trsm_strided_batched(h, side, uplo, trans, diag, m, n, 0.75, dA.data_ptr(),
                                         lda, sa, dB.data_ptr(), ldb, sb, 2)
            assert st == S.success
            both = dB.cpu().numpy()
            for b in range(2):
                assert np.array_equal(np.asfortranarray(both[b * sb:(b + 1) * sb]
                                                        .reshape(n, ldb).T), got), b


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_recursive_solve_leaves_match_batched(h, side, uplo, trans):
    """A single instance too large for one chained launch recurses; its
    leaves (k <= 512) run as chained launches when they have at most two
    vector panels per 64-row block (12288 x 192: 512-row leaves, 8 x 12
    workgroups each), else one workgroup per panel (1024 x 1024).  The result
    meets the oracle's bound (1024) or the backward-error bound (12288, data
    generated on the device) and equals, bit for bit, the same instance
    inside a batch of two, whose leaves always take the
    one-workgroup-per-panel kernel."""
    rng = np.random.default_rng(910)
    for (k, nv) in [(1024, 1024), (12288, 192)]:
        m, n = (k, nv) if side == "L" else (nv, k)
        say("recursive leaves", side, uplo, trans, m, n)
        if k <= 2048:
            A = oracle.diag_dominant_triangular(k, uplo, rng, lda=k)
            B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
            st, got = gpu_trsm(h, side, uplo, trans, "N", m, n, -0.5, A, B)
            assert st == S.success
            ref = B.copy(order="F")
            oracle.trsm(side, uplo, trans, "N", m, n, -0.5, A, ref)
            assert oracle.norm1_rel_diff(m, n, got, ref) <= 40 * EPS64 * k
            dA1 = torch.from_numpy(A.ravel(order="F").copy()).cuda()
            dB1 = torch.from_numpy(B.ravel(order="F").copy()).cuda()
            X1 = torch.from_numpy(np.ascontiguousarray(got.ravel(order="F"))).cuda()
        else:
            g = torch.Generator(device="cuda").manual_seed(k + nv)
            dA1 = torch.randn(k * k, dtype=torch.float64, device="cuda", generator=g) / k ** 0.5
            dA1.view(k, k).diagonal().add_(2.0)
            dB1 = torch.rand(m * n, dtype=torch.float64, device="cuda", generator=g) - 0.5
            X1 = dB1.clone()
            st = rb.trsm(h, side, uplo, trans, "N", m, n, -0.5, dA1.data_ptr(), k,
                         X1.data_ptr(), m)
            assert st == S.success
            torch.cuda.synchronize()
            Al = dA1.view(k, k).T
            eta = eta_torch(side, uplo, trans, "N", -0.5, Al, dB1.view(n, m).T, X1.view(n, m).T)
            assert eta <= 1e-14, eta
        sa, sb = k * k, m * n
        dA = torch.cat([dA1, dA1])
        dB = torch.cat([dB1, dB1])
        st = rb.trsm_strided_batched(h, side, uplo, trans, "N", m, n, -0.5, dA.data_ptr(), k, sa,
                                     dB.data_ptr(), m, sb, 2)
        assert st == S.success
        torch.cuda.synchronize()
        for b in range(2):
            assert torch.equal(dB[b * sb:(b + 1) * sb], X1), b
        del dA, dB, dA1, dB1, X1


# ------------------------------------------------------------ trsm_ex --

def test_trsm_ex_forms_match_typed_entry_points(h):
    """rocblas_trsm{,_batched,_strided_batched}_ex (library/src/blas3/rocblas_trsm.cpp:28-73):
    f64/f32 by compute_type give the typed entry points' results bit for bit, with or
    without an invA hint; other types are not implemented; argument checks are
    trsm's."""
    rng = np.random.default_rng(77)
    L = rb.lib()
    m, n, batch = 200, 37, 3
    for prec, dt, code in (("d", np.float64, 152), ("s", np.float32, 151)):
        A = oracle.diag_dominant_triangular(m, "U", rng, poison=0.0).astype(dt)
        B = rng.uniform(-1, 1, (m, n)).astype(dt)
        A, B = np.asfortranarray(A), np.asfortranarray(B)
        say("trsm_ex", prec)
        dA = dev(A)
        want = dev(B)
        assert rb.trsm(h, "L", "U", "T", "N", m, n, 1.5, dA.data_ptr(), m, want.data_ptr(), m,
                       precision=prec) == S.success
        alpha = (ctypes.c_double if prec == "d" else ctypes.c_float)(1.5)
        junk = torch.full((128 * m,), float("nan"), dtype=dA.dtype, device="cuda")
        for invA, size in ((None, 0), (junk.data_ptr(), 128 * m)):
            got = dev(B)
            assert L.rocblas_trsm_ex(h.ptr, 141, 121, 112, 131, m, n, ctypes.addressof(alpha),
                                     dA.data_ptr(), m, got.data_ptr(), m, invA, size,
                                     code) == S.success
            torch.cuda.synchronize()
            assert torch.equal(got, want)
        # strided and pointer-array batched
        dAs = torch.cat([dA] * batch)
        dBs = torch.cat([dev(B)] * batch)
        assert L.rocblas_trsm_strided_batched_ex(h.ptr, 141, 121, 112, 131, m, n,
                                                 ctypes.addressof(alpha), dAs.data_ptr(), m, m * m,
                                                 dBs.data_ptr(), m, m * n, batch, None, 0, 0,
                                                 code) == S.success
        Bl = [dev(B) for _ in range(batch)]
        pa, pb = ptr_array([dA] * batch), ptr_array(Bl)
        assert L.rocblas_trsm_batched_ex(h.ptr, 141, 121, 112, 131, m, n, ctypes.addressof(alpha),
                                         pa.data_ptr(), m, pb.data_ptr(), m, batch, None, 0,
                                         code) == S.success
        torch.cuda.synchronize()
        for b in range(batch):
            assert torch.equal(dBs[b * m * n:(b + 1) * m * n], want)
            assert torch.equal(Bl[b], want)
    one = ctypes.c_double(1.0)
    assert L.rocblas_trsm_ex(h.ptr, 141, 121, 111, 131, 4, 4, ctypes.addressof(one), None, 4, None,
                             4, None, 0, 154) == S.not_implemented
    assert L.rocblas_trsm_ex(h.ptr, 141, 121, 111, 131, 4, 4, ctypes.addressof(one), None, 3, None,
                             4, None, 0, 152) == S.invalid_size
    assert L.rocblas_trsm_ex(None, 141, 121, 111, 131, 4, 4, ctypes.addressof(one), None, 4, None,
                             4, None, 0, 152) == S.invalid_handle


# ------------------------------------------------------------- ILP64 --

def test_ilp64_forms(h):
    """rocblas_{s,d}trsm{,_batched,_strided_batched}_64
    (library/src/src64/blas3/rocblas_trsm_64.cpp): 64-bit sizes give the 32-bit
    results bit for bit; a triangle above INT32_MAX is invalid_size, checks are
    trsm's."""
    rng = np.random.default_rng(64)
    L = rb.lib()
    m, n, batch = 150, 40, 2
    A = oracle.diag_dominant_triangular(n, "L", rng, poison=0.0)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    dA = dev(A)
    want = dev(B)
    assert rb.trsm(h, "R", "L", "N", "U", m, n, -2.0, dA.data_ptr(), n, want.data_ptr(),
                   m) == S.success
    alpha = ctypes.c_double(-2.0)
    got = dev(B)
    assert L.rocblas_dtrsm_64(h.ptr, 142, 122, 111, 132, m, n, ctypes.addressof(alpha),
                              dA.data_ptr(), n, got.data_ptr(), m) == S.success
    dAs = torch.cat([dA] * batch)
    dBs = torch.cat([dev(B)] * batch)
    assert L.rocblas_dtrsm_strided_batched_64(h.ptr, 142, 122, 111, 132, m, n,
                                              ctypes.addressof(alpha), dAs.data_ptr(), n, n * n,
                                              dBs.data_ptr(), m, m * n, batch) == S.success
    Bl = [dev(B) for _ in range(batch)]
    pa, pb = ptr_array([dA] * batch), ptr_array(Bl)
    assert L.rocblas_dtrsm_batched_64(h.ptr, 142, 122, 111, 132, m, n, ctypes.addressof(alpha),
                                      pa.data_ptr(), n, pb.data_ptr(), m, batch) == S.success
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    for b in range(batch):
        assert torch.equal(dBs[b * m * n:(b + 1) * m * n], want)
        assert torch.equal(Bl[b], want)
    big = 1 << 31
    assert L.rocblas_dtrsm_64(h.ptr, 141, 122, 111, 131, big, 1, ctypes.addressof(alpha),
                              dA.data_ptr(), big, got.data_ptr(), big) == S.invalid_size
    assert L.rocblas_dtrsm_64(h.ptr, 141, 122, 111, 131, 4, 1, ctypes.addressof(alpha),
                              dA.data_ptr(), 3, got.data_ptr(), 4) == S.invalid_size
    assert L.rocblas_dtrsm_64(None, 141, 122, 111, 131, 4, 1, ctypes.addressof(alpha),
                              dA.data_ptr(), 4, got.data_ptr(), 4) == S.invalid_handle
    assert L.rocblas_dtrsm_64(h.ptr, 141, 122, 111, 131, 0, 1, ctypes.addressof(alpha),
                              None, 4, None, 4) == S.success


@pytest.mark.parametrize("precision", ["d", "s"])
def test_dma_gemm_batch_loop_matches_single(h, precision):
    """2048 x 2048 solves with batch 2: the recursion's first update is
    1024 x 2048 x 1024 over two instances, which runs the LDS-DMA GEMM's batch
    loop (grid y, per-instance strides, LDS reuse across instances).  Each
    instance of the strided and the pointer-array forms equals its
    single-instance solve bit for bit."""
    rng = np.random.default_rng(2048)
    n, batch = 2048, 2
    dt = np.float64 if precision == "d" else np.float32
    tdt = torch.float64 if precision == "d" else torch.float32
    As = [oracle.diag_dominant_triangular(n, "U", rng, dtype=dt, poison=0.0) for _ in range(batch)]
    Bs = [np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(dt)) for _ in range(batch)]
    singles = []
    for A, B in zip(As, Bs):
        dA, dB = dev(A), dev(B)
        assert rb.trsm(h, "L", "U", "N", "N", n, n, 1.0, dA.data_ptr(), n, dB.data_ptr(), n,
                       precision) == S.success
        singles.append(dB)
    say("dma batch loop", precision)
    dAs = torch.cat([dev(A) for A in As])
    dBs = torch.cat([dev(B) for B in Bs])
    assert rb.trsm_strided_batched(h, "L", "U", "N", "N", n, n, 1.0, dAs.data_ptr(), n, n * n,
                                   dBs.data_ptr(), n, n * n, batch, precision) == S.success
    dAl = [dev(A) for A in As]
    dBl = [dev(B) for B in Bs]
    pa, pb = ptr_array(dAl), ptr_array(dBl)
    assert rb.trsm_batched(h, "L", "U", "N", "N", n, n, 1.0, pa.data_ptr(), n, pb.data_ptr(), n,
                           batch, precision) == S.success
    torch.cuda.synchronize()
    for b in range(batch):
        assert torch.equal(dBs[b * n * n:(b + 1) * n * n], singles[b]), ("strided", b)
        assert torch.equal(dBl[b], singles[b]), ("pointer array", b)
    assert singles[0].dtype == tdt
