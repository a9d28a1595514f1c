"""GPU parity tests of rocblas_amd_{s,d}trsm_vbatched (variable-size batched
solve in one persistent launch) through the C ABI.

There is no reference counterpart; each instance must equal what the
reference's rocblas_dtrsm gives for it (library/src/blas3/rocblas_trsm.cpp),
which these tests check against the oracle (the netlib loop nest the clients'
host cblas_dtrsm runs): bit for bit on exactly-solvable integer problems,
within a backward-error bound on random ones, and against the single-instance
entry point.
"""
import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from tests.gpu_util import GpuHandle, dev, host, ptr_array, say
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


class VbProblem:
    """Device-side argument arrays of one vbatched call over a list of
    (A, B) Fortran arrays; B buffers are kept so results can be read back."""

    def __init__(self, side, ms, ns, As, Bs, dtype=torch.float64):
        self.ms, self.ns = list(ms), list(ns)
        self.dA = [torch.from_numpy(np.ascontiguousarray(A.ravel(order="F"))).to("cuda", dtype)
                   for A in As]
        self.dB = [torch.from_numpy(np.ascontiguousarray(B.ravel(order="F"))).to("cuda", dtype)
                   for B in Bs]
        self.shapes = [B.shape for B in Bs]
        i32 = lambda v: torch.tensor(v, dtype=torch.int32, device="cuda")
        self.m = i32(self.ms)
        self.n = i32(self.ns)
        self.lda = i32([A.shape[0] for A in As])
        self.ldb = i32([B.shape[0] for B in Bs])
        self.pa = ptr_array(self.dA)
        self.pb = ptr_array(self.dB)
        self.batch = len(As)

    def run(self, h, side, uplo, trans, diag, alpha, precision="d", batch=None):
        return rb.trsm_vbatched(h, side, uplo, trans, diag, self.m.data_ptr(), self.n.data_ptr(),
                                alpha, self.pa.data_ptr(), self.lda.data_ptr(),
                                self.pb.data_ptr(), self.ldb.data_ptr(),
                                self.batch if batch is None else batch, precision)

    def result(self, i):
        return host(self.dB[i], self.shapes[i])


def test_bad_args(h):
    say("vbatched bad args")
    L = rb.lib()
    P = VbProblem("L", [4], [4], [np.eye(4, order="F")], [np.ones((4, 4), order="F")])
    assert L.rocblas_amd_dtrsm_vbatched(h.ptr, 140, 121, 111, 131, P.m.data_ptr(), P.n.data_ptr(),
                                        None, P.pa.data_ptr(), P.lda.data_ptr(), P.pb.data_ptr(),
                                        P.ldb.data_ptr(), 1) == S.invalid_value
    assert L.rocblas_amd_dtrsm_vbatched(h.ptr, 141, 120, 111, 131, P.m.data_ptr(), P.n.data_ptr(),
                                        None, P.pa.data_ptr(), P.lda.data_ptr(), P.pb.data_ptr(),
                                        P.ldb.data_ptr(), 1) == S.invalid_value
    assert L.rocblas_amd_dtrsm_vbatched(h.ptr, 141, 121, 111, 133, P.m.data_ptr(), P.n.data_ptr(),
                                        None, P.pa.data_ptr(), P.lda.data_ptr(), P.pb.data_ptr(),
                                        P.ldb.data_ptr(), 1) == S.invalid_value
    assert P.run(h, "L", "L", "N", "N", 1.0, batch=-1) == S.invalid_size
    assert P.run(h, "L", "L", "N", "N", 1.0, batch=0) == S.success
    assert L.rocblas_amd_dtrsm_vbatched(None, 141, 121, 111, 131, P.m.data_ptr(), P.n.data_ptr(),
                                        None, P.pa.data_ptr(), P.lda.data_ptr(), P.pb.data_ptr(),
                                        P.ldb.data_ptr(), 1) == S.invalid_handle
    assert P.run(h, "L", "L", "N", "N", None) == S.invalid_pointer
    assert L.rocblas_amd_dtrsm_vbatched(h.ptr, 141, 121, 111, 131, None, P.n.data_ptr(),
                                        None, P.pa.data_ptr(), P.lda.data_ptr(), P.pb.data_ptr(),
                                        P.ldb.data_ptr(), 1) == S.invalid_pointer
    torch.cuda.synchronize()


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_integer_kat_bitwise_mixed_sizes(h, side, uplo, trans):
    """Exactly solvable instances of uneven size (k from 1 to 150, padded leading
    dimensions, NaN in every unreferenced entry of A): every result is X bit for bit."""
    rng = np.random.default_rng(1000 + ALL_CASES.index((side, uplo, trans)))
    sizes = [(1, 1), (3, 70), (17, 9), (64, 64), (65, 1), (33, 100), (150, 20), (8, 130), (31, 31),
             (100, 65), (2, 2), (47, 48)]
    for diag in "NU":
        say("vbatched kat", side, uplo, trans, diag)
        As, Bs, Xs = [], [], []
        for (m, n) in sizes:
            k = m if side == "L" else n
            A, B, X = integer_kat(side, uplo, trans, diag, m, n, rng, lda=k + 2, ldb=m + 1)
            As.append(A)
            Bs.append(B)
            Xs.append(X)
        P = VbProblem(side, [s[0] for s in sizes], [s[1] for s in sizes], As, Bs)
        assert P.run(h, side, uplo, trans, diag, 1.0) == S.success
        torch.cuda.synchronize()
        for i, (m, n) in enumerate(sizes):
            got = P.result(i)
            assert np.array_equal(got[:m, :n], Xs[i][:m, :n]), (i, m, n)
            assert np.all(got[m:, :] == 0.0)  # padding rows of B untouched


@pytest.mark.parametrize("side,uplo,trans", ALL_CASES)
def test_random_against_oracle(h, side, uplo, trans):
    """Workload-like instances (m log-uniform in [8, 512], n uniform in [1, 64]) in
    every case, alpha = -1.5: agreement with the oracle and backward error."""
    rng = np.random.default_rng(7 + 3 * ALL_CASES.index((side, uplo, trans)))
    batch = 40
    ms = np.exp(rng.uniform(np.log(8), np.log(512), batch)).astype(int)
    ns = rng.integers(1, 65, batch)
    if side == "R":
        ms, ns = ns, ms
    As, Bs = [], []
    for m, n in zip(ms, ns):
        k = m if side == "L" else n
        As.append(oracle.diag_dominant_triangular(int(k), uplo, rng))
        Bs.append(np.asfortranarray(rng.uniform(-1, 1, (m, n))))
    say("vbatched random", side, uplo, trans)
    P = VbProblem(side, ms, ns, As, Bs)
    alpha = -1.5
    assert P.run(h, side, uplo, trans, "N", alpha) == S.success
    torch.cuda.synchronize()
    for i, (m, n) in enumerate(zip(ms, ns)):
        k = m if side == "L" else n
        ref = Bs[i].copy(order="F")
        oracle.trsm(side, uplo, trans, "N", int(m), int(n), alpha, As[i], ref)
        got = P.result(i)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (i, m, n)
        op = op_dense(As[i], uplo, trans, "N", k)
        R = alpha * Bs[i] - (op @ got if side == "L" else got @ op)
        n1 = lambda M: np.abs(M).sum(axis=0).max()
        eta = n1(R) / (n1(op) * n1(got) + abs(alpha) * n1(Bs[i]))
        assert eta <= 1e-14, (i, m, n, eta)


@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("R", "U", "T")])
def test_large_triangles_against_oracle(h, side, uplo, trans):
    """Triangles beyond 512 (up to 1100) in the same batch as small ones take
    the same panel path: agreement with the oracle and backward error."""
    rng = np.random.default_rng(77 + ALL_CASES.index((side, uplo, trans)))
    kv = [(600, 20), (1100, 5), (700, 70), (513, 1), (40, 9), (1024, 16)]  # (k, vectors)
    ms = [k if side == "L" else v for k, v in kv]
    ns = [v if side == "L" else k for k, v in kv]
    As = [oracle.diag_dominant_triangular(k, uplo, rng) for k, _ in kv]
    Bs = [np.asfortranarray(rng.uniform(-1, 1, (m, n))) for m, n in zip(ms, ns)]
    say("vbatched large", side, uplo, trans)
    P = VbProblem(side, ms, ns, As, Bs)
    alpha = 0.75
    assert P.run(h, side, uplo, trans, "N", alpha) == S.success
    torch.cuda.synchronize()
    for i, ((k, _), m, n) in enumerate(zip(kv, ms, ns)):
        ref = Bs[i].copy(order="F")
        oracle.trsm(side, uplo, trans, "N", m, n, alpha, As[i], ref)
        got = P.result(i)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (i, m, n)
        op = op_dense(As[i], uplo, trans, "N", k)
        R = alpha * Bs[i] - (op @ got if side == "L" else got @ op)
        n1 = lambda M: np.abs(M).sum(axis=0).max()
        eta = n1(R) / (n1(op) * n1(got) + abs(alpha) * n1(Bs[i]))
        assert eta <= 1e-14, (i, m, n, eta)


def test_matches_single_instance_entry_and_skips_malformed(h):
    """Each instance matches rocblas_dtrsm on it to 1e-13; instances with an
    empty size or a too-small leading dimension are left untouched; alpha = 0
    zeroes B without reading A; order of the instances does not matter."""
    rng = np.random.default_rng(99)
    sizes = [(200, 7), (12, 64), (0, 5), (40, 0), (90, 33), (5, 5), (256, 1), (20, 20)]
    As, Bs = [], []
    for m, n in sizes:
        As.append(oracle.diag_dominant_triangular(max(m, 1), "L", rng, poison=0.0))
        Bs.append(np.asfortranarray(rng.uniform(-1, 1, (max(m, 1), max(n, 1)))))
    say("vbatched vs single")
    P = VbProblem("L", [s[0] for s in sizes], [s[1] for s in sizes], As, Bs)
    P.lda[5] = 4  # lda < m: malformed, skipped
    assert P.run(h, "L", "L", "T", "N", 2.0) == S.success
    torch.cuda.synchronize()
    for i, (m, n) in enumerate(sizes):
        got = P.result(i)
        if i in (2, 3, 5):
            assert np.array_equal(got, Bs[i]), i
            continue
        dA, dB = dev(As[i]), dev(Bs[i])
        assert rb.trsm(h, "L", "L", "T", "N", m, n, 2.0, dA.data_ptr(), m, dB.data_ptr(),
                       m) == S.success
        torch.cuda.synchronize()
        want = host(dB, Bs[i].shape)
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), i
    # reversed instance order gives the same results bit for bit
    first = [P.result(i) for i in range(len(sizes))]
    Q = VbProblem("L", [s[0] for s in sizes][::-1], [s[1] for s in sizes][::-1], As[::-1],
                  Bs[::-1])
    Q.lda[len(sizes) - 1 - 5] = 4
    assert Q.run(h, "L", "L", "T", "N", 2.0) == S.success
    torch.cuda.synchronize()
    for i in range(len(sizes)):
        assert np.array_equal(Q.result(len(sizes) - 1 - i), first[i]), i
    # alpha = 0: B := 0 and A is not read (pointer array entries set to NULL)
    Z = VbProblem("L", [s[0] for s in sizes], [s[1] for s in sizes], As, Bs)
    Z.pa.zero_()
    assert Z.run(h, "L", "L", "N", "N", 0.0) == S.success
    torch.cuda.synchronize()
    for i, (m, n) in enumerate(sizes):
        if m > 0 and n > 0:
            assert np.all(Z.result(i)[:m, :n] == 0.0), i


def test_device_pointer_mode_and_float32(h):
    rng = np.random.default_rng(5)
    sizes = [(60, 10), (9, 3), (130, 64)]
    As = [oracle.diag_dominant_triangular(m, "U", rng, poison=0.0) for m, _ in sizes]
    Bs = [np.asfortranarray(rng.uniform(-1, 1, (m, n))) for m, n in sizes]
    say("vbatched device alpha")
    P = VbProblem("L", [s[0] for s in sizes], [s[1] for s in sizes], As, Bs)
    assert P.run(h, "L", "U", "N", "N", 3.0) == S.success
    torch.cuda.synchronize()
    host_res = [P.result(i) for i in range(len(sizes))]
    D = VbProblem("L", [s[0] for s in sizes], [s[1] for s in sizes], As, Bs)
    alpha = torch.tensor([3.0], dtype=torch.float64, device="cuda")
    rb.check(h.set_pointer_mode(rb.rocblas_pointer_mode.device))
    try:
        st = rb.trsm_vbatched(h, "L", "U", "N", "N", D.m.data_ptr(), D.n.data_ptr(), None,
                              D.pa.data_ptr(), D.lda.data_ptr(), D.pb.data_ptr(),
                              D.ldb.data_ptr(), D.batch, alpha_ptr=alpha.data_ptr())
    finally:
        rb.check(h.set_pointer_mode(rb.rocblas_pointer_mode.host))
    assert st == S.success
    torch.cuda.synchronize()
    for i in range(len(sizes)):
        assert np.array_equal(D.result(i), host_res[i]), i
    say("vbatched float32")
    F = VbProblem("L", [s[0] for s in sizes], [s[1] for s in sizes], As, Bs, dtype=torch.float32)
    assert F.run(h, "L", "U", "N", "N", 3.0, precision="s") == S.success
    torch.cuda.synchronize()
    for i, (m, n) in enumerate(sizes):
        ref = Bs[i].astype(np.float32, order="F")
        oracle.trsm("L", "U", "N", "N", m, n, 3.0, np.asfortranarray(As[i].astype(np.float32)), ref)
        got = F.result(i)
        assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max(), i


def test_graph_capture_with_new_sizes(h):
    """One captured call replays correctly after the size and pointer arrays are
    rewritten on the device (the plan kernel resets the work counter itself)."""
    rng = np.random.default_rng(11)
    cap = 30
    mmax, nmax = 300, 64
    As = [oracle.diag_dominant_triangular(mmax, "L", rng, poison=0.0) for _ in range(cap)]
    B0s = [np.asfortranarray(rng.uniform(-1, 1, (mmax, nmax))) for _ in range(cap)]
    P = VbProblem("L", [mmax] * cap, [nmax] * cap, As, B0s)
    s = torch.cuda.Stream()
    hg = GpuHandle()
    rb.check(hg.set_stream(s.cuda_stream))
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        st = P.run(hg, "L", "L", "N", "N", 1.0)
    assert st == S.success
    for rep in range(4):
        say("vbatched graph replay", rep)
        ms = rng.integers(1, mmax + 1, cap)
        ns = rng.integers(1, nmax + 1, cap)
        P.m.copy_(torch.from_numpy(ms.astype(np.int32)))
        P.n.copy_(torch.from_numpy(ns.astype(np.int32)))
        for i in range(cap):
            P.dB[i].copy_(torch.from_numpy(B0s[i].ravel(order="F")))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i in range(cap):
            m, n = int(ms[i]), int(ns[i])
            ref = B0s[i].copy(order="F")
            oracle.trsm("L", "L", "N", "N", m, n, 1.0, As[i], ref)
            got = P.result(i)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (rep, i, m, n)
        _ = (P.dB[0] * 2).sum().item()
    hg.close()
