"""GPU test of rocblas_amd_dtrsm_vbatched_noalloc through the C ABI (collected last).

The entry point is rocblas_amd_dtrsm_vbatched with its planning scratch in a
library-owned device array instead of the handle workspace.  Both run the
same kernels, so their results must be equal bit for bit; two calls in a row
(the second smaller, another case) check that the shared array carries no
state from one call to the next.  No reference counterpart.
"""
import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from tests.gpu_util import GpuHandle, host, ptr_array, say

pytestmark = pytest.mark.gpu
S = rb.rocblas_status
MAX_BATCH = 65536


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available()
    oracle.build()
    hd = GpuHandle()
    yield hd
    hd.close()


def _problem(side, uplo, batch, kmax, seed):
    """Random instances (k log-uniform in [8, kmax], 1..64 vectors) with padded
    leading dimensions; returns the host arrays and sizes."""
    rng = np.random.default_rng(seed)
    ks = np.exp(rng.uniform(np.log(8), np.log(kmax), batch)).astype(int)
    nv = rng.integers(1, 65, batch)
    ms, ns = (ks, nv) if side == "L" else (nv, ks)
    As, Bs = [], []
    for m, n, k in zip(ms, ns, ks):
        A = oracle.diag_dominant_triangular(int(k), uplo, rng, lda=int(k) + 3)
        B = np.asfortranarray(np.zeros((m + 1, n)))
        B[:m, :] = rng.uniform(-1, 1, (m, n))
        As.append(A)
        Bs.append(B)
    return [int(v) for v in ms], [int(v) for v in ns], As, Bs


class _Args:
    """Device arrays of one variable-size batched call; fresh copies of B."""

    def __init__(self, ms, ns, As, Bs):
        i32 = lambda v: torch.tensor(v, dtype=torch.int32, device="cuda")
        self.dA = [torch.from_numpy(np.ascontiguousarray(A.ravel(order="F"))).cuda() for A in As]
        self.dB = [torch.from_numpy(np.ascontiguousarray(B.ravel(order="F"))).cuda() for B in Bs]
        self.shapes = [B.shape for B in Bs]
        self.m, self.n = i32(ms), i32(ns)
        self.lda = i32([A.shape[0] for A in As])
        self.ldb = i32([B.shape[0] for B in Bs])
        self.pa, self.pb = ptr_array(self.dA), ptr_array(self.dB)
        self.batch = len(As)

    def call(self, fn, h, side, uplo, trans, diag, alpha, batch=None):
        return fn(h, side, uplo, trans, diag, self.m.data_ptr(), self.n.data_ptr(), alpha,
                  self.pa.data_ptr(), self.lda.data_ptr(), self.pb.data_ptr(),
                  self.ldb.data_ptr(), self.batch if batch is None else batch)

    def results(self):
        return [host(b, s) for b, s in zip(self.dB, self.shapes)]


def _vbatched(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, batch):
    return rb.trsm_vbatched(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, batch, "d")


CALLS = [
    # (side, uplo, trans, diag, batch, kmax, alpha, seed): the second call is smaller
    ("L", "U", "T", "N", 300, 700, -1.25, 11),
    ("R", "L", "N", "U", 37, 200, 0.5, 12),
]


def test_two_calls_match_workspace_entry_bitwise(h):
    """Two calls in a row, the second with fewer and smaller instances in
    another case: each equals rocblas_amd_dtrsm_vbatched's result bit for bit
    (and the oracle's to 1e-12), padding rows and NaN-poisoned unreferenced
    entries untouched."""
    for side, uplo, trans, diag, batch, kmax, alpha, seed in CALLS:
        ms, ns, As, Bs = _problem(side, uplo, batch, kmax, seed)
        say("noalloc vs workspace entry", side, uplo, trans, diag, "batch", batch, "kmax", kmax)
        ref = _Args(ms, ns, As, Bs)
        got = _Args(ms, ns, As, Bs)
        assert ref.call(_vbatched, h, side, uplo, trans, diag, alpha) == S.success
        assert got.call(rb.trsm_vbatched_noalloc, h, side, uplo, trans, diag, alpha) == S.success
        torch.cuda.synchronize()
        R, G = ref.results(), got.results()
        for i, (m, n) in enumerate(zip(ms, ns)):
            assert np.array_equal(G[i], R[i]), (side, uplo, trans, i, m, n)
            assert np.all(G[i][m:, :] == 0.0)
        for i in range(0, batch, 7):
            m, n = ms[i], ns[i]
            x = Bs[i].copy(order="F")
            oracle.trsm(side, uplo, trans, diag, m, n, alpha, As[i], x)
            assert np.abs(G[i] - x).max() <= 1e-12 * max(1.0, np.abs(x).max()), (i, m, n)


def test_size_limit(h):
    """batch_count above the static array's capacity is rejected before any
    work; at the limit the arguments pass the host checks (an empty problem of
    65536 instances whose sizes are all 0 is skipped on the device)."""
    ms = [0] * MAX_BATCH
    ns = [0] * MAX_BATCH
    i32 = lambda v: torch.tensor(v, dtype=torch.int32, device="cuda")
    m, n = i32(ms), i32(ns)
    ld = i32([1] * MAX_BATCH)
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    pa = ptr_array([dummy] * MAX_BATCH)
    say("noalloc size limit")
    st = rb.trsm_vbatched_noalloc(h, "L", "L", "N", "N", m.data_ptr(), n.data_ptr(), 1.0,
                                  pa.data_ptr(), ld.data_ptr(), pa.data_ptr(), ld.data_ptr(),
                                  MAX_BATCH + 1)
    assert st == S.invalid_size
    st = rb.trsm_vbatched_noalloc(h, "L", "L", "N", "N", m.data_ptr(), n.data_ptr(), 1.0,
                                  pa.data_ptr(), ld.data_ptr(), pa.data_ptr(), ld.data_ptr(),
                                  MAX_BATCH)
    assert st == S.success
    torch.cuda.synchronize()
    assert float(dummy.item()) == 0.0
