"""GPU tests at the full sizes of the measured workloads, through the C ABI.

* single solve, m=512 n=256, side L lower: against the oracle;
* m = n = 16384, side L upper, off-diagonal N(0,1)/sqrt(m): the plain FP64
  solve and the mixed-precision form, each with its normwise backward error
  from a plain PyTorch float64 residual on the device (and the mixed form's
  own reported value checked against it);
* 4096 variable-size instances (m log-uniform in [8, 512], n in 1..64) solved
  by one graph-captured _vbatched launch, replayed with eager work between
  replays, every instance checked against the oracle loop nest and its own
  backward error.

The reference's procedure for a batch checks every instance
(clients/include/blas3/testing_trsm_strided_batched.hpp:755-800); so do these.
"""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from rocblas_amd import synth, workloads
from tests.gpu_util import GpuHandle, dev, eta_torch, host, logical, say

pytestmark = pytest.mark.gpu
S = rb.rocblas_status
EPS64 = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    oracle.build()
    hd = GpuHandle()
    yield hd
    hd.close()


def test_single_512x256_against_oracle(h):
    m, n = 512, 256
    say("workload single", m, n)
    A = synth.tri_uniform(m, "L", 0)
    B = synth.rhs_uniform(m, n, 0)
    dA, dB = dev(A), dev(B)
    assert rb.trsm(h, "L", "L", "N", "N", m, n, 1.0, dA.data_ptr(), m, dB.data_ptr(), m) \
        == S.success
    torch.cuda.synchronize()
    got = host(dB, (m, n))
    ref = B.copy(order="F")
    oracle.trsm("L", "L", "N", "N", m, n, 1.0, A, ref)
    assert oracle.norm1_rel_diff(m, n, got, ref) <= 40 * EPS64 * m
    assert oracle.backward_error("L", "L", "N", "N", m, n, 1.0, A, B, got) <= 1e-15
    assert h.get_async_status() == S.success


@pytest.fixture(scope="module")
def large_problem(h):
    m = 16384
    dA = torch.empty(m * m, dtype=torch.float64, device="cuda")
    dB0 = torch.empty(m * m, dtype=torch.float64, device="cuda")
    rb.check(rb.synth_tri(h, "U", "normal", m, 0, dA.data_ptr(), m))
    rb.check(rb.synth_rhs(h, m, m, 0, dB0.data_ptr(), m))
    torch.cuda.synchronize()
    yield m, dA, dB0
    del dA, dB0
    torch.cuda.empty_cache()


def test_large_16384_plain_fp64(h, large_problem):
    m, dA, dB0 = large_problem
    say("workload large plain", m)
    dB = dB0.clone()
    assert rb.trsm(h, "L", "U", "N", "N", m, m, 1.0, dA.data_ptr(), m, dB.data_ptr(), m) \
        == S.success
    torch.cuda.synchronize()
    eta = eta_torch("L", "U", "N", "N", 1.0, logical(dA, m, m, m), logical(dB0, m, m, m),
                    logical(dB, m, m, m))
    say("large plain eta", eta)
    assert eta <= 1e-14
    del dB


def test_large_16384_mixed(h, large_problem):
    m, dA, dB0 = large_problem
    say("workload large mixed", m)
    dB = dB0.clone()
    st, info = rb.trsm_mixed(h, "L", "U", "N", "N", m, m, 1.0, dA.data_ptr(), m, dB.data_ptr(),
                             m, max_refine=10, tol=5e-13)
    assert st == S.success
    torch.cuda.synchronize()
    eta = eta_torch("L", "U", "N", "N", 1.0, logical(dA, m, m, m), logical(dB0, m, m, m),
                    logical(dB, m, m, m))
    say("large mixed eta", eta, "reported", info.backward_error, "steps", info.iterations)
    assert info.converged == 1
    assert eta <= 1e-12
    # the call's own residual accepted X: same value up to the rounding of two
    # different FP64 residual computations
    assert abs(info.backward_error - eta) <= 0.05 * eta + 4 * EPS64
    del dB


def _instance_check(args):
    m, n, A, B0, X = args
    ref = B0.copy(order="F")
    oracle.trsm("L", "L", "N", "N", m, n, 1.0, A, ref)
    diff = oracle.norm1_rel_diff(m, n, X, ref)
    L = np.tril(A)
    R = B0 - L @ X
    n1 = lambda Z: np.abs(Z).sum(axis=0).max()
    eta = n1(R) / (n1(L) * n1(X) + n1(B0))
    return diff, eta


def test_variable_batch_4096_graph(h):
    """4096 instances in one captured graph, replayed three times with eager
    work in between: every replay gives the same bits, and every instance
    meets the oracle at 40*eps*m and a backward error of 1e-14."""
    say("workload vbatched 4096")
    w = workloads.variable_batch(h, 4096, 8, 512, 1, 64, 0)
    ms, ns, offA, offB = w["ms"], w["ns"], w["offA"], w["offB"]
    dA, dB0, dB = w["dA"], w["dB0"], w["dB"]
    gs = torch.cuda.Stream()
    hg = GpuHandle()
    rb.check(hg.set_stream(gs.cuda_stream))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=gs):
        st = workloads.call_vbatched(hg, w)
    assert st == S.success
    first = None
    for rep in range(3):
        dB.copy_(dB0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        if first is None:
            first = dB.clone()
        else:
            assert torch.equal(dB, first), rep
        _ = (dB0 * 3).abs().sum().item()  # eager work between replays
    hg.close()
    Ah = dA.cpu().numpy()
    B0h = dB0.cpu().numpy()
    Xh = first.cpu().numpy()
    jobs = []
    for i in range(len(ms)):
        m, n = int(ms[i]), int(ns[i])
        A = np.asfortranarray(Ah[offA[i]:offA[i + 1]].reshape(m, m).T)
        B0 = np.asfortranarray(B0h[offB[i]:offB[i + 1]].reshape(n, m).T)
        X = np.asfortranarray(Xh[offB[i]:offB[i + 1]].reshape(n, m).T)
        jobs.append((m, n, A, B0, X))
    with cf.ThreadPoolExecutor(8) as ex:
        res = list(ex.map(_instance_check, jobs))
    diffs = np.array([r[0] for r in res])
    etas = np.array([r[1] for r in res])
    bound = 40 * EPS64 * ms.astype(float)
    say("vbatched 4096: worst diff/bound", float((diffs / bound).max()), "worst eta",
        float(etas.max()))
    assert np.all(diffs <= bound)
    assert etas.max() <= 1e-14
    del w, dA, dB0, dB, first
    torch.cuda.empty_cache()
