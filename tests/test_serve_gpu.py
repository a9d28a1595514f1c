"""GPU tests of the continuous-batching serving loop (rocblas_amd_server_*).

Each request solves op(A) X = alpha B or X op(A) = alpha B in its own case
(side, uplo, trans, diag) and alpha, with a resident synthetic factor.
Results are regenerated on the host (bit-identical uniform streams) and
checked against the oracle and by backward error (<= 1e-12 FP64, <= 1e-5
FP32), alone and in a mix, under an admission budget small enough to force
evictions; the server's own per-request backward error (FP64, on the device)
is checked against the oracle's value of the same quantity.
"""
import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from rocblas_amd import serve
from tests.gpu_util import say

pytestmark = pytest.mark.gpu
S = rb.rocblas_status
CASES = [(s, u, t) for s in "LR" for u in "LU" for t in "NT"]
ENUM = {"L": 141, "R": 142}, {"U": 121, "L": 122}, {"N": 111, "T": 112}, {"N": 131, "U": 132}


@pytest.fixture(scope="module", autouse=True)
def gpu():
    assert torch.cuda.is_available()
    torch.cuda.init()
    oracle.build()


def mixed_rows(seed, count, arrival_step, all_cases=True):
    """Requests of mixed size, precision, case and alpha; 25% reuse a factor."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(count):
        side, uplo, trans = CASES[int(rng.integers(0, 8))] if all_cases else ("L", "L", "N")
        diag = "U" if all_cases and rng.uniform() < 0.3 else "N"
        alpha = float(rng.choice([1.0, -0.5, 2.0, 0.75])) if all_cases else 1.0
        if i >= 4 and rng.uniform() < 0.25:
            base = rows[int(rng.integers(0, i))]
            m, prec, fid, fseed = base["m"], base["precision"], base["factor_id"], base["factor_seed"]
            uplo_e = base["uplo"]
        else:
            m = int(np.exp(rng.uniform(np.log(16), np.log(900))))
            prec = int(rng.integers(0, 2))
            fid, fseed = 1000 + i, 77 + i
            uplo_e = ENUM[1][uplo]
        rows.append(dict(factor_id=fid, factor_seed=fseed, rhs_seed=5000 + i,
                         arrival=i * arrival_step, m=m, n=int(rng.integers(1, 48)),
                         precision=prec, keep=1, side=ENUM[0][side], uplo=uplo_e,
                         transA=ENUM[2][trans], diag=ENUM[3][diag], alpha=alpha))
    return rows


def oracle_case(req):
    side, uplo, trans, diag = serve.case_letters(req)
    rows, cols = serve.b_shape(req)
    return side, uplo, trans, diag, rows, cols


def check_result(req, X, eta_dev=None):
    """X against the oracle's solve and backward error; the server's own eta
    (if given) against the oracle's eta of the same X."""
    A, B = serve.host_problem(req)
    side, uplo, trans, diag, rows, cols = oracle_case(req)
    A64 = np.asfortranarray(A.astype(np.float64))
    B64 = np.asfortranarray(B.astype(np.float64))
    X64 = np.asfortranarray(X.astype(np.float64))
    e = oracle.backward_error(side, uplo, trans, diag, rows, cols, req.alpha, A64, B64, X64)
    tol = 1e-12 if req.precision == 0 else 1e-5
    assert e <= tol, (req.m, req.n, req.precision, e)
    ref = B64.copy(order="F")
    oracle.trsm(side, uplo, trans, diag, rows, cols, req.alpha, A64, ref)
    rel = np.abs(X64 - ref).max() / max(np.abs(ref).max(), 1e-300)
    assert rel <= (1e-10 if req.precision == 0 else 1e-3), (req.m, req.n, rel)
    if eta_dev is not None:
        # same quantity, FP64 on the device vs long double on the host: the
        # residual of an accurate X is at rounding level, so the two agree to
        # within rounding of the residual itself
        assert eta_dev >= 0
        assert abs(eta_dev - e) <= 0.25 * e + 4 * np.finfo(np.float64).eps, (eta_dev, e)


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
        assert st.checked == len(rows) and st.over_tolerance == 0
        for i in range(len(rows)):
            check_result(reqs[i], mix[i], srv.eta[i])
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


def test_poisson_trace_200gib_every_result_checked():
    """A 1 s slice of the serving workload at its full admission budget of
    200 GiB (2000 req/s, m log-uniform in [16, 4096], n in 1..128, FP64/FP32,
    all eight cases, unit and non-unit diagonals, alpha in +-[0.5, 2], 10%
    factor reuse): everything completes, every request's backward error is
    computed on the device and meets its target, and a sample of kept results
    (every tenth) is checked against the oracle, including the device eta."""
    reqs = serve.poisson_trace(2000, 1.0, seed=7, keep_every=10)
    srv = serve.Server(budget=200 << 30, streams=4)
    try:
        say("serve poisson 200 GiB", len(reqs))
        lat, st = srv.run(reqs)
        assert st.completed == len(reqs) and st.rejected == 0
        assert st.checked == st.completed
        assert st.over_tolerance == 0
        assert np.all(srv.eta >= 0)
        assert st.max_backward_error_fp64 <= 1e-12 and st.max_backward_error_fp32 <= 1e-5
        assert 0 < st.p50_latency <= st.p99_latency <= st.max_latency
        seen = set()
        for i in range(0, len(reqs), 10):
            seen.add(oracle_case(reqs[i])[:4])
            if reqs[i].m <= 2048:
                check_result(reqs[i], srv.result(i), srv.eta[i])
        assert len({c[:3] for c in seen}) == 8, seen
    finally:
        srv.close()


def test_verify_off_and_alpha_zero():
    """verify=0 computes no backward errors; alpha = 0 gives X = 0 (the factor is
    still generated and kept resident)."""
    rows = mixed_rows(3, 6, 0.0)
    rows[2]["alpha"] = 0.0
    srv = serve.Server(budget=1 << 30, streams=2, verify=False)
    try:
        lat, st = srv.run(serve.make_requests(rows))
        assert st.completed == 6 and st.checked == 0
        assert np.all(srv.eta == -1.0)
        assert np.all(srv.result(2) == 0)
        check_result(srv.requests[0], srv.result(0))
    finally:
        srv.close()
