"""GPU tests of the continuous-batching serving loop (rocblas_amd_server_*).

Every request is the solve L X = B with a resident synthetic factor; results are
regenerated on the host (bit-identical uniform streams) and checked against the
oracle and by backward error (<= 1e-12 FP64, <= 1e-5 FP32), alone and in a
mix, under an admission budget small enough to force evictions.
"""
import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from rocblas_amd import serve
from tests.gpu_util import say
from tests.helpers import op_dense

pytestmark = pytest.mark.gpu
S = rb.rocblas_status


@pytest.fixture(scope="module", autouse=True)
def gpu():
    assert torch.cuda.is_available()
    torch.cuda.init()
    oracle.build()


def eta(A, B, X):
    k = A.shape[0]
    L = op_dense(A.astype(np.float64), "L", "N", "N", k)
    Xd, Bd = X.astype(np.float64), B.astype(np.float64)
    R = Bd - L @ Xd
    n1 = lambda M: np.abs(M).sum(axis=0).max()
    return n1(R) / (n1(L) * n1(Xd) + n1(Bd))


def mixed_rows(seed, count, arrival_step):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(count):
        if i >= 4 and rng.uniform() < 0.25:
            j = int(rng.integers(0, i))
            base = rows[j]
            m, prec, fid, fseed = base["m"], base["precision"], base["factor_id"], base["factor_seed"]
        else:
            m = int(np.exp(rng.uniform(np.log(16), np.log(900))))
            prec = int(rng.integers(0, 2))
            fid, fseed = 1000 + i, 77 + i
        rows.append(dict(factor_id=fid, factor_seed=fseed, rhs_seed=5000 + i,
                         arrival=i * arrival_step, m=m, n=int(rng.integers(1, 48)),
                         precision=prec, keep=1))
    return rows


def check_result(req, X):
    A, B = serve.host_problem(req)
    tol = 1e-12 if req.precision == 0 else 1e-5
    e = eta(A, B, X)
    assert e <= tol, (req.m, req.n, req.precision, e)
    ref = B.copy(order="F")
    oracle.trsm("L", "L", "N", "N", req.m, req.n, 1.0, A, ref)
    rel = np.abs(X.astype(np.float64) - ref).max() / max(np.abs(ref).max(), 1e-300)
    assert rel <= (1e-10 if req.precision == 0 else 1e-3), (req.m, req.n, rel)


def test_mix_matches_alone_and_oracle():
    """40 requests of mixed size and precision arriving 0.2 ms apart (25% reuse
    a resident factor): every result meets the accuracy target and equals, bit
    for bit, the same request served alone."""
    rows = mixed_rows(1, 40, 2e-4)
    reqs = serve.make_requests(rows)
    srv = serve.Server(budget=1 << 30, streams=4)
    try:
        say("serve mix")
        lat, st = srv.run(reqs)
        assert st.completed == len(rows) and st.rejected == 0
        assert np.all(lat >= 0)
        assert st.factor_hits >= 1 and st.batched_launches >= 1
        mix = [srv.result(i) for i in range(len(rows))]
        for i in range(len(rows)):
            check_result(reqs[i], mix[i])
        say("serve alone")
        for i in range(len(rows)):
            one = dict(rows[i], arrival=0.0)
            lat1, st1 = srv.run(serve.make_requests([one]))
            assert st1.completed == 1
            alone = srv.result(0)
            assert np.array_equal(alone, mix[i]), i
    finally:
        srv.close()


def test_small_budget_evicts_and_stays_correct():
    """An arena of 8 MiB holds two or three factors (m in [300, 600]): the loop
    evicts the least recently used idle factors, regenerates them on later use,
    and results are unchanged."""
    rows = mixed_rows(2, 30, 0.0)
    first = {}
    for i, r in enumerate(rows):
        f = first.setdefault(r["factor_id"], (300 + r["m"] % 300, r["precision"]))
        r["m"], r["precision"] = f  # one factor id keeps a single shape
        r["keep"] = 1 if i % 6 == 0 else 0
    reqs = serve.make_requests(rows)
    budget = 8 << 20
    srv = serve.Server(budget=budget, streams=2, max_inflight=8)
    try:
        say("serve small budget")
        lat, st = srv.run(reqs)
        assert st.completed == len(rows) and st.rejected == 0
        assert st.evictions >= 1
        assert st.peak_arena_bytes <= budget
        for i in range(0, len(rows), 6):
            check_result(reqs[i], srv.result(i))
    finally:
        srv.close()


def test_rejects_what_cannot_fit_and_bad_traces():
    srv = serve.Server(budget=1 << 20, streams=1)
    try:
        rows = [dict(factor_id=1, factor_seed=1, rhs_seed=1, arrival=0.0, m=600, n=4, precision=0,
                     keep=0),
                dict(factor_id=2, factor_seed=2, rhs_seed=2, arrival=0.0, m=32, n=4, precision=1,
                     keep=1)]
        lat, st = srv.run(serve.make_requests(rows))
        assert st.rejected == 1 and st.completed == 1
        assert lat[0] == -1.0 and lat[1] >= 0.0
        check_result(srv.requests[1], srv.result(1))
        bad = [dict(rows[1], arrival=1.0), dict(rows[1], arrival=0.5)]  # not sorted
        with pytest.raises(rb.RocblasError):
            srv.run(serve.make_requests(bad))
    finally:
        srv.close()


def test_poisson_trace_short():
    """A 0.25 s slice of the serving workload (2000 req/s, m log-uniform in
    [16, 4096], n in 1..128, FP64/FP32, 10% factor reuse): everything completes,
    latencies are reported, and sampled results meet the accuracy targets."""
    reqs = serve.poisson_trace(2000, 0.25, seed=7, keep_every=50)
    srv = serve.Server(budget=8 << 30, streams=4)
    try:
        say("serve poisson", len(reqs))
        lat, st = srv.run(reqs)
        assert st.completed == len(reqs)
        assert 0 < st.p50_latency <= st.p99_latency <= st.max_latency
        for i in range(0, len(reqs), 50):
            if reqs[i].m <= 1500:
                check_result(reqs[i], srv.result(i))
    finally:
        srv.close()
