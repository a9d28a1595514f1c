"""Python side of the serving loop (rocblas_amd_server_* in include/rocblas_amd_ext.h).

The loop itself runs in C++ inside the library: this module only builds
request traces and calls it.  ``poisson_trace`` draws the serving workload:
Poisson arrivals at ``rate`` requests/s for ``duration`` s, triangle order m
log-uniform in [m_lo, m_hi], n right-hand-side vectors uniform in
{n_lo..n_hi}, FP64 or FP32 with probability 1/2 each, one of the eight
side/uplo/trans cases (and a unit diagonal with probability ``unit``), alpha
uniform in [-2, -0.5] U [0.5, 2], and a fraction ``reuse`` of requests that
solve with a factor some earlier request already made resident (same uplo,
new right-hand side, case and alpha).  All randomness comes from the
counter-based SplitMix64 streams of rocblas_amd.synth, so a trace is a pure
function of its seed.
"""
import ctypes

import numpy as np

import rocblas_amd as rb
from rocblas_amd import synth


class Request(ctypes.Structure):
    """rocblas_amd_request"""
    _fields_ = [("factor_id", ctypes.c_int64), ("factor_seed", ctypes.c_uint64),
                ("rhs_seed", ctypes.c_uint64), ("arrival", ctypes.c_double),
                ("m", ctypes.c_int32), ("n", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("keep", ctypes.c_int32), ("side", ctypes.c_int32), ("uplo", ctypes.c_int32),
                ("transA", ctypes.c_int32), ("diag", ctypes.c_int32), ("alpha", ctypes.c_double)]


class Stats(ctypes.Structure):
    """rocblas_amd_serve_stats"""
    _fields_ = [(f, ctypes.c_int64) for f in
                ("completed", "rejected", "launches", "batched_launches", "factor_hits",
                 "factor_misses", "evictions", "checked", "over_tolerance")] + \
               [(f, ctypes.c_double) for f in
                ("wall_seconds", "p50_latency", "p99_latency", "max_latency", "mean_latency",
                 "peak_arena_bytes", "max_backward_error_fp64", "max_backward_error_fp32")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_BOUND = False


def _bind():
    global _BOUND
    if _BOUND:
        return rb.lib()
    L = rb.lib()
    vp = ctypes.c_void_p
    L.rocblas_amd_server_create.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.rocblas_amd_server_create.restype = ctypes.c_int32
    L.rocblas_amd_server_destroy.argtypes = [vp]
    L.rocblas_amd_server_destroy.restype = ctypes.c_int32
    L.rocblas_amd_server_run.argtypes = [vp, ctypes.POINTER(Request), ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(Stats)]
    L.rocblas_amd_server_run.restype = ctypes.c_int32
    L.rocblas_amd_server_result.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int32]
    L.rocblas_amd_server_result.restype = ctypes.c_int32
    _BOUND = True
    return L


class Server:
    """One serving loop: a device arena of ``budget`` bytes and ``streams`` streams;
    with ``verify`` every request's backward error is computed on the device."""

    def __init__(self, budget, streams=4, max_inflight=256, small_k=512, verify=True):
        L = _bind()
        self.ptr = ctypes.c_void_p()
        rb.check(L.rocblas_amd_server_create(ctypes.byref(self.ptr), int(budget), streams,
                                             max_inflight, small_k, int(bool(verify))),
                 "server_create")
        self.requests = None
        self.eta = None

    def close(self):
        if self.ptr:
            _bind().rocblas_amd_server_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, requests):
        """Replay ``requests`` (a ctypes Request array); returns (latency array, Stats).
        Per-request backward errors (-1 where not computed) are left in ``self.eta``."""
        n = len(requests)
        lat = (ctypes.c_double * max(n, 1))()
        eta = (ctypes.c_double * max(n, 1))()
        st = Stats()
        rb.check(_bind().rocblas_amd_server_run(self.ptr, requests, n, lat, eta,
                                                ctypes.byref(st)), "server_run")
        self.requests = requests
        self.eta = np.ctypeslib.as_array(eta)[:n].copy()
        return np.ctypeslib.as_array(lat)[:n].copy(), st

    def result(self, index):
        """Kept X of request ``index`` of the last run, as a Fortran array of B's
        shape ((m, n) for side left, (n, m) for side right)."""
        r = self.requests[index]
        dt = np.float64 if r.precision == 0 else np.float32
        rows, cols = b_shape(r)
        X = np.empty((rows, cols), dtype=dt, order="F")
        rb.check(_bind().rocblas_amd_server_result(self.ptr, index,
                                                   X.ctypes.data_as(ctypes.c_void_p), rows),
                 "server_result")
        return X


CASE_FIELDS = dict(side=int(rb.rocblas_side.left), uplo=int(rb.rocblas_fill.lower),
                   transA=int(rb.rocblas_operation.none), diag=int(rb.rocblas_diagonal.non_unit),
                   alpha=1.0)


def b_shape(req):
    """(rows, cols) of a request's B: m x n for side left, n x m for side right."""
    left = req.side in (0, int(rb.rocblas_side.left))
    return (req.m, req.n) if left else (req.n, req.m)


def case_letters(req):
    """The request's (side, uplo, trans, diag) as the oracle's letters."""
    side = "R" if req.side == int(rb.rocblas_side.right) else "L"
    uplo = "U" if req.uplo == int(rb.rocblas_fill.upper) else "L"
    trans = "N" if req.transA in (0, int(rb.rocblas_operation.none)) else "T"
    diag = "U" if req.diag == int(rb.rocblas_diagonal.unit) else "N"
    return side, uplo, trans, diag


def make_requests(rows):
    """rows: iterable of dicts with the Request fields -> ctypes array.  The case
    fields default to side left, lower, no transpose, non-unit, alpha 1."""
    rows = list(rows)
    arr = (Request * len(rows))()
    for i, r in enumerate(rows):
        for k, v in dict(CASE_FIELDS, **r).items():
            setattr(arr[i], k, v)
    return arr


def poisson_trace(rate, duration, seed=0, m_lo=16, m_hi=4096, n_lo=1, n_hi=128, reuse=0.1,
                  keep_every=0, all_cases=True, unit=0.25):
    """Seeded Poisson trace; returns a ctypes Request array sorted by arrival.
    ``keep_every`` > 0 keeps X of every keep_every-th request for checking.
    ``all_cases`` draws the side/uplo/trans case, the diagonal and alpha per
    request; otherwise every request is side left, lower, no transpose,
    non-unit, alpha 1."""
    key = synth.stream_key(seed, synth.TAG_SIZES)
    # exponential inter-arrival gaps until the duration is covered
    count_guess = int(rate * duration * 1.2 + 64)
    u = synth.uniform(key, 0, 9 * count_guess)
    gaps = -np.log1p(-u[:count_guess]) / rate
    t = np.cumsum(gaps)
    t = t[t < duration]
    cnt = len(t)
    um, un, up, ur, uf, uc, ud, ua = (u[count_guess * (i + 1):count_guess * (i + 1) + cnt]
                                      for i in range(8))
    ms = np.exp(np.log(m_lo) + um * (np.log(m_hi + 1) - np.log(m_lo))).astype(np.int64)
    ms = np.clip(ms, m_lo, m_hi)
    ns = np.clip((n_lo + un * (n_hi - n_lo + 1)).astype(np.int64), n_lo, n_hi)
    prec = (up < 0.5).astype(np.int64)  # 1 = FP32
    arr = (Request * cnt)()
    fid = np.empty(cnt, dtype=np.int64)
    next_id = 0
    for i in range(cnt):
        r = arr[i]
        reused = i > 0 and ur[i] < reuse
        if reused:
            # reuse a recent factor (same size and precision), new right-hand side
            j = max(0, i - 1 - int(uf[i] * min(i, 256)))
            fid[i] = fid[j]
            r.m, r.precision = arr[j].m, arr[j].precision
            r.factor_seed = arr[j].factor_seed
        else:
            fid[i] = next_id
            next_id += 1
            r.m, r.precision = int(ms[i]), int(prec[i])
            r.factor_seed = (seed << 32) + int(fid[i])
        r.factor_id = int(fid[i])
        if all_cases:
            c = min(int(uc[i] * 8), 7)
            r.side = int(rb.rocblas_side.right if c & 4 else rb.rocblas_side.left)
            r.transA = int(rb.rocblas_operation.transpose if c & 1
                           else rb.rocblas_operation.none)
            # a new factor takes its triangle from the draw; reuse keeps the factor's
            r.uplo = arr[j].uplo if reused else int(rb.rocblas_fill.upper if c & 2
                                                    else rb.rocblas_fill.lower)
            r.diag = int(rb.rocblas_diagonal.unit if ud[i] < unit
                         else rb.rocblas_diagonal.non_unit)
            a = 0.5 + 1.5 * ua[i]
            r.alpha = float(-a if (int(ua[i] * 2 ** 20) & 1) else a)
        else:
            r.side, r.uplo = int(rb.rocblas_side.left), int(rb.rocblas_fill.lower)
            r.transA, r.diag = int(rb.rocblas_operation.none), int(rb.rocblas_diagonal.non_unit)
            r.alpha = 1.0
        r.n = int(ns[i])
        r.rhs_seed = (seed << 32) + (1 << 31) + i
        r.arrival = float(t[i])
        r.keep = 1 if keep_every and i % keep_every == 0 else 0
    return arr


def host_problem(req):
    """The factor and right-hand side a request solves, regenerated on the host
    (the device generator's uniform streams are bit-identical)."""
    dt = np.float64 if req.precision == 0 else np.float32
    uplo = "U" if req.uplo == int(rb.rocblas_fill.upper) else "L"
    A = synth.tri_uniform(req.m, uplo, req.factor_seed, dtype=dt)
    rows, cols = b_shape(req)
    B = synth.rhs_uniform(rows, cols, req.rhs_seed, dtype=dt)
    return A, B
