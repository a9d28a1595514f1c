"""Additional measured workloads of bench.py (one JSON object each).

* single_512x256   -- rocblas_dtrsm L/L/N/N m=512 n=256 alpha=1, A lower with
                      diagonal U[1,2] and off-diagonal U[-1,1]/m, B U[-1,1],
                      seed 0: per-call latency p50/p99, TFLOP/s, backward error.
* large_16384      -- m=n=16384 L/U/N/N, off-diagonal N(0,1)/sqrt(m): the plain
                      FP64 blocked solve and the mixed form (FP32 MFMA trailing
                      updates, FP64 diagonal-block inverses, FP64 residual and
                      refinement to <= 1e-12), both credited with m*n*k flops.
* variable_batch   -- 4096 instances, side L lower, m log-uniform in [8, 512],
                      n uniform in {1..64}: one graph-captured vbatched launch vs
                      a per-instance loop of rocblas_dtrsm.
* serving          -- Poisson arrivals at 2000 req/s, m log-uniform in
                      [16, 4096], n in {1..128}, FP64/FP32, 10% factor reuse,
                      200 GB admission budget: p50/p99 latency and solves/s.

Where the vendor librocblas that ships with torch is present, its
rocblas_dtrsm is timed beside ours through its C API (a comparator only).
"""
import ctypes
import os
import time

import numpy as np

FP64_PEAK_TF = 78.6
FP32_PEAK_TF = 157.3


def _events(torch, n):
    return [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            for _ in range(n)]


def vendor_lib(torch):
    path = os.path.join(os.path.dirname(torch.__file__), "lib", "librocblas.so")
    if not os.path.exists(path):
        return None
    V = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    V.rocblas_dtrsm.argtypes = [vp] + [ctypes.c_int] * 6 + [vp, vp, ctypes.c_int, vp,
                                                             ctypes.c_int]
    V.rocblas_strsm.argtypes = V.rocblas_dtrsm.argtypes
    h = vp()
    if V.rocblas_create_handle(ctypes.byref(h)) != 0:
        return None
    V.rocblas_set_stream(h, vp(torch.cuda.current_stream().cuda_stream))
    return V, h


def eta_device(torch, side, uplo, A, B0, X, alpha=1.0):
    """Normwise backward error with a plain PyTorch float64 residual on the device."""
    T = torch.triu(A) if uplo == "U" else torch.tril(A)
    R = alpha * B0 - (T @ X if side == "L" else X @ T)
    n1 = lambda Z: Z.abs().sum(0).max()
    return float(n1(R) / (n1(T) * n1(X) + abs(alpha) * n1(B0)))


def leg_single(torch, rb, synth, oracle, vendor):
    m, n = 512, 256
    A = synth.tri_uniform(m, "L", 0)
    B = synth.rhs_uniform(m, n, 0)
    dA = torch.from_numpy(A.ravel(order="F").copy()).cuda()
    dB0 = torch.from_numpy(B.ravel(order="F").copy()).cuda()
    h = rb.Handle()
    rb.check(h.set_stream(torch.cuda.current_stream().cuda_stream))
    bufs = [dB0.clone() for _ in range(8)]
    out = {}
    for lib in ("ours", "vendor"):
        if lib == "vendor" and vendor is None:
            continue
        ts = []
        for r in range(105):
            dB = bufs[r % 8]
            dB.copy_(dB0)
            ev = _events(torch, 1)[0]
            ev[0].record()
            if lib == "ours":
                st = rb.trsm(h, "L", "L", "N", "N", m, n, 1.0, dA.data_ptr(), m, dB.data_ptr(), m)
            else:
                one = ctypes.c_double(1.0)
                V, vh = vendor
                st = V.rocblas_dtrsm(vh, 141, 122, 111, 131, m, n, ctypes.addressof(one),
                                     ctypes.c_void_p(dA.data_ptr()), m,
                                     ctypes.c_void_p(dB.data_ptr()), m)
            ev[1].record()
            torch.cuda.synchronize()
            assert st == 0, (lib, st)
            if r >= 5:
                ts.append(ev[0].elapsed_time(ev[1]) * 1e3)
        ts = np.array(ts)
        X = bufs[104 % 8].view(n, m).T
        e = eta_device(torch, "L", "L", dA.view(m, m).T, dB0.view(n, m).T, X)
        out[lib] = {"p50_us": float(np.percentile(ts, 50)), "p99_us": float(np.percentile(ts, 99)),
                    "tflops": m * n * m / (np.percentile(ts, 50) * 1e-6) / 1e12,
                    "backward_error": e}
    h.close()
    # oracle on one host core
    Bw = B.copy(order="F")
    t0 = time.perf_counter()
    oracle.trsm("L", "L", "N", "N", m, n, 1.0, A, Bw)
    t = time.perf_counter() - t0
    return {"workload": "dtrsm L/L/N/N m=512 n=256 alpha=1 (diag U[1,2], off-diag U[-1,1]/m)",
            "ours": out["ours"], "vendor_rocblas": out.get("vendor"),
            "cpu_oracle_us": t * 1e6, "cpu_cores": 1}


def leg_large(torch, rb, oracle, vendor, m=16384, reps=3):
    n = m
    h = rb.Handle()
    rb.check(h.set_stream(torch.cuda.current_stream().cuda_stream))
    dA = torch.empty(m * m, dtype=torch.float64, device="cuda")
    dB0 = torch.empty(m * n, dtype=torch.float64, device="cuda")
    rb.check(rb.synth_tri(h, "U", "normal", m, 0, dA.data_ptr(), m))
    rb.check(rb.synth_rhs(h, m, n, 0, dB0.data_ptr(), m))
    dB = torch.empty_like(dB0)
    flops = float(m) * n * m
    res = {}

    def timed(fn):
        ts = []
        for r in range(reps + 1):
            dB.copy_(dB0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            if r:
                ts.append(time.perf_counter() - t0)
        return min(ts), out

    t, _ = timed(lambda: rb.check(rb.trsm(h, "L", "U", "N", "N", m, n, 1.0, dA.data_ptr(), m,
                                          dB.data_ptr(), m)))
    Am, B0m = dA.view(m, m).T, dB0.view(n, m).T
    e = eta_device(torch, "L", "U", Am, B0m, dB.view(n, m).T)
    res["plain_fp64"] = {"ms": t * 1e3, "tflops": flops / t / 1e12,
                         "frac_fp64_peak": flops / t / 1e12 / FP64_PEAK_TF, "backward_error": e}
    info_box = {}
    # the precision form with each update form (tuned block size)
    for form in ("fp32", "bf16x3"):
        def mixed(form=form):
            st, info = rb.trsm_mixed(h, "L", "U", "N", "N", m, n, 1.0, dA.data_ptr(), m,
                                     dB.data_ptr(), m, max_refine=10, tol=5e-13, form=form)
            rb.check(st)
            info_box["info"] = info

        t, _ = timed(mixed)
        e = eta_device(torch, "L", "U", Am, B0m, dB.view(n, m).T)
        info = info_box["info"]
        res[f"mixed_{form}_refined"] = {"ms": t * 1e3, "tflops": flops / t / 1e12,
                                        "speedup_vs_plain": res["plain_fp64"]["ms"] / (t * 1e3),
                                        "refinement_steps": info.iterations,
                                        "reported_backward_error": info.backward_error,
                                        "backward_error": e}
    if vendor is not None:
        V, vh = vendor
        one = ctypes.c_double(1.0)
        t, _ = timed(lambda: V.rocblas_dtrsm(vh, 141, 121, 111, 131, m, n, ctypes.addressof(one),
                                             ctypes.c_void_p(dA.data_ptr()), m,
                                             ctypes.c_void_p(dB.data_ptr()), m))
        res["vendor_rocblas_fp64"] = {"ms": t * 1e3, "tflops": flops / t / 1e12}
    # the same problem in single precision (rocblas_strsm), FP32 copies of A and B
    dA32, dB032 = dA.float(), dB0.float()
    dB32 = torch.empty_like(dB032)

    def timed32(fn):
        ts = []
        for r in range(reps + 1):
            dB32.copy_(dB032)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            if r:
                ts.append(time.perf_counter() - t0)
        return min(ts)

    t = timed32(lambda: rb.check(rb.trsm(h, "L", "U", "N", "N", m, n, 1.0, dA32.data_ptr(), m,
                                         dB32.data_ptr(), m, precision="s")))
    e = eta_device(torch, "L", "U", dA32.double().view(m, m).T, dB032.double().view(n, m).T,
                   dB32.double().view(n, m).T)
    res["plain_fp32"] = {"ms": t * 1e3, "tflops": flops / t / 1e12,
                         "frac_fp32_peak": flops / t / 1e12 / FP32_PEAK_TF, "backward_error": e}
    if vendor is not None:
        V, vh = vendor
        one32 = ctypes.c_float(1.0)
        t = timed32(lambda: V.rocblas_strsm(vh, 141, 121, 111, 131, m, n, ctypes.addressof(one32),
                                            ctypes.c_void_p(dA32.data_ptr()), m,
                                            ctypes.c_void_p(dB32.data_ptr()), m))
        res["vendor_rocblas_fp32"] = {"ms": t * 1e3, "tflops": flops / t / 1e12}
    del dA32, dB032, dB32
    # CPU oracle on a bounded column sample, scaled to the full width
    A_host = dA.view(m, m).T.cpu().numpy()
    A_host = np.asfortranarray(A_host)
    cols = 8
    Bs = np.asfortranarray(dB0.view(n, m).T[:, :cols].cpu().numpy())
    t0 = time.perf_counter()
    oracle.trsm_omp("L", "U", "N", "N", m, cols, 1.0, A_host, Bs, oracle.num_threads())
    tc = time.perf_counter() - t0
    res["cpu_oracle"] = {"seconds_full_estimate": tc * n / cols, "cores": oracle.num_threads(),
                         "sample": f"{cols} of {n} right-hand sides, {tc:.1f} s, scaled"}
    h.close()
    del dA, dB0, dB, A_host
    torch.cuda.empty_cache()
    return dict({"workload": "dtrsm (and strsm) L/U/N/N m=n=16384, off-diagonal N(0,1)/sqrt(m)"},
                **res)


def leg_variable_batch(torch, rb, vendor, batch=4096, seed=0, eager=False):
    """The variable-size batch: one graph-captured plan + persistent launch
    (eager=True: the same call launched directly, for profilers whose counter
    collection does not follow graph launches), the backward error of every
    instance on the device, a per-instance loop of our rocblas_dtrsm and of
    the vendor's, and a per-instance loop of the mixed-precision form (FP32
    updates, refined to <= 1e-12)."""
    from rocblas_amd import workloads
    h = rb.Handle()
    s = torch.cuda.current_stream()
    rb.check(h.set_stream(s.cuda_stream))
    w = workloads.variable_batch(h, batch, 8, 512, 1, 64, seed)
    ms, ns, offA, offB = w["ms"], w["ns"], w["offA"], w["offB"]
    dA, dB0, dB = w["dA"], w["dB0"], w["dB"]
    dm, dn, pa, pb = w["m"], w["n"], w["pa"], w["pb"]
    flops = w["flops"]

    # graph-captured launch (plan + persistent solve), replayed
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    gs = torch.cuda.Stream()
    hg = rb.Handle()
    rb.check(hg.set_stream(gs.cuda_stream))

    def call(hh):
        rb.check(rb.trsm_vbatched(hh, "L", "L", "N", "N", dm.data_ptr(), dn.data_ptr(), 1.0,
                                  pa.data_ptr(), dm.data_ptr(), pb.data_ptr(), dm.data_ptr(),
                                  batch))

    if not eager:
        with torch.cuda.graph(g, stream=gs):
            call(hg)
    ts = []
    for r in range(11):
        dB.copy_(dB0)
        torch.cuda.synchronize()
        ev = _events(torch, 1)[0]
        ev[0].record()
        if eager:
            call(h)
        else:
            g.replay()
        ev[1].record()
        torch.cuda.synchronize()
        if r:
            ts.append(ev[0].elapsed_time(ev[1]) * 1e-3)
    t_graph = float(np.median(ts))
    # correctness: the backward error of every instance, on the device
    etas = []
    for i in range(batch):
        m, n = int(ms[i]), int(ns[i])
        A = torch.tril(dA[int(offA[i]):int(offA[i + 1])].view(m, m).T)
        B0 = dB0[int(offB[i]):int(offB[i + 1])].view(n, m).T
        X = dB[int(offB[i]):int(offB[i + 1])].view(n, m).T
        R = B0 - A @ X
        n1 = lambda Z: Z.abs().sum(0).max()
        etas.append(n1(R) / (n1(A) * n1(X) + n1(B0)))
    etas = torch.stack(etas)
    worst = float(etas.max())
    over = int((etas > 1e-12).sum())
    # per-instance loop of the plain entry point
    def loop(lib):
        dB.copy_(dB0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one = ctypes.c_double(1.0)
        for i in range(batch):
            m, n = int(ms[i]), int(ns[i])
            a = dA.data_ptr() + 8 * int(offA[i])
            b = dB.data_ptr() + 8 * int(offB[i])
            if lib == "ours":
                st = rb.trsm(h, "L", "L", "N", "N", m, n, 1.0, a, m, b, m)
            else:
                V, vh = vendor
                st = V.rocblas_dtrsm(vh, 141, 122, 111, 131, m, n, ctypes.addressof(one),
                                     ctypes.c_void_p(a), m, ctypes.c_void_p(b), m)
            assert st == 0
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    loop("ours")
    t_loop = min(loop("ours") for _ in range(2))
    key = "vbatched_eager" if eager else "vbatched_graph"
    res = {"workload": "4096 instances, side L lower, m log-uniform [8,512], n uniform {1..64}",
           "instances": batch, "gflop": flops / 1e9,
           key: {"ms": t_graph * 1e3, "solves_per_s": batch / t_graph,
                 "tflops": flops / t_graph / 1e12, "max_backward_error": worst,
                 "checked_instances": batch, "over_1e-12": over},
           "loop_of_dtrsm_ours": {"ms": t_loop * 1e3, "solves_per_s": batch / t_loop},
           "speedup_vs_loop": t_loop / t_graph}
    # the mixed-precision form, instance by instance (it has no batched form)
    dB.copy_(dB0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mixed_worst, mixed_steps = 0.0, 0
    for i in range(batch):
        m, n = int(ms[i]), int(ns[i])
        st, info = rb.trsm_mixed(h, "L", "L", "N", "N", m, n, 1.0, dA.data_ptr() + 8 * int(offA[i]),
                                 m, dB.data_ptr() + 8 * int(offB[i]), m, max_refine=10, tol=1e-12)
        rb.check(st, "dtrsm_mixed")
        mixed_worst = max(mixed_worst, info.backward_error)
        mixed_steps = max(mixed_steps, info.iterations)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    res["loop_of_mixed_fp32_refined"] = {"ms": t_mixed * 1e3, "solves_per_s": batch / t_mixed,
                                         "max_reported_backward_error": mixed_worst,
                                         "max_refinement_steps": mixed_steps}
    if vendor is not None:
        loop("vendor")
        t_v = min(loop("vendor") for _ in range(2))
        res["loop_of_dtrsm_vendor"] = {"ms": t_v * 1e3, "solves_per_s": batch / t_v}
        res["speedup_vs_vendor_loop"] = t_v / t_graph
    hg.close()
    h.close()
    del dA, dB0, dB, w
    torch.cuda.empty_cache()
    return res


def leg_trtri(torch, rb, vendor, sizes=(1024, 4096, 16384), reps=5):
    """Public rocblas_dtrtri (lower, non-unit) at several orders, ours and the
    vendor library's, credited n^3/3 flops (the reference's count,
    clients/include/flops.hpp:741-744); ||inv(A) A - I||_max checked on the
    device for each result.  Plus a strided batch of 20000 order-32 inverses."""
    import ctypes
    from rocblas_amd import synth
    out = {"workload": "dtrtri lower non-unit, A = diag U[1,2], off-diagonal U[-1,1]/n",
           "flops": "n^3/3"}
    h = rb.Handle()
    rb.check(h.set_stream(torch.cuda.current_stream().cuda_stream))
    if vendor is not None:
        V, vh = vendor
        V.rocblas_dtrtri.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return float(np.median(ts))

    for n in sizes:
        dA = torch.empty(n * n, dtype=torch.float64, device="cuda")
        rb.check(rb.synth_tri(h, "L", "uniform", n, 5, dA.data_ptr(), n, scale=1.0 / n),
                 "synth_tri")
        dX = torch.empty_like(dA)
        t = timed(lambda: rb.check(rb.trtri(h, "L", "N", n, dA.data_ptr(), n, dX.data_ptr(), n),
                                   "dtrtri"))
        A = torch.tril(dA.view(n, n).T)
        X = dX.view(n, n).T
        err = float((X @ A - torch.eye(n, dtype=torch.float64, device="cuda")).abs().max())
        row = {"ms": t * 1e3, "tflops": n ** 3 / 3 / t / 1e12, "max_abs_inv_times_A_minus_I": err}
        if vendor is not None:
            dY = torch.empty_like(dA)
            tv = timed(lambda: V.rocblas_dtrtri(vh, 122, 131, n, dA.data_ptr(), n, dY.data_ptr(),
                                                n))
            row["vendor_ms"] = tv * 1e3
            row["vendor_tflops"] = n ** 3 / 3 / tv / 1e12
            row["ours_over_vendor"] = tv / t
            del dY
        out[f"n{n}"] = row
        del dA, dX, A, X
        torch.cuda.empty_cache()
    n, batch = 32, 20000
    dA = torch.empty(batch * n * n, dtype=torch.float64, device="cuda")
    # one synthetic order-32 triangle repeated over the batch
    rb.check(rb.synth_tri(h, "L", "uniform", n, 9, dA.data_ptr(), n, scale=1.0 / n), "synth")
    dA.view(batch, n * n).copy_(dA[: n * n].view(1, n * n).expand(batch, n * n))
    dX = torch.empty_like(dA)
    t = timed(lambda: rb.check(rb.trtri_strided_batched(h, "L", "N", n, dA.data_ptr(), n, n * n,
                                                        dX.data_ptr(), n, n * n, batch),
                               "dtrtri_strided_batched"))
    out["strided_batched_20000_n32"] = {
        "us": t * 1e6, "inverses_per_s": batch / t,
        "hbm_gbs": batch * 8 * (n * (n + 1) / 2 + n * n) / t / 1e9}
    del dA, dX
    h.close()
    torch.cuda.empty_cache()
    return out


def _serve_once(serve, rate, duration, budget, seed=0):
    reqs = serve.poisson_trace(rate, duration, seed=seed)
    srv = serve.Server(budget=budget, streams=4, max_inflight=512)
    try:
        lat, st = srv.run(reqs)
    finally:
        srv.close()
    return reqs, st


def leg_serving(torch, rb, duration=60.0, rate=2000.0, budget=200 << 30,
                sweep=(4000.0, 8000.0, 16000.0, 32000.0), sweep_seconds=5.0, p99_bound_ms=10.0):
    """The serving workload (every request's backward error computed on the
    device, as part of the request), then shorter runs at higher offered rates
    to find the largest rate the loop sustains: completed rate >= 95% of the
    offered rate with p99 latency within p99_bound_ms."""
    from rocblas_amd import serve
    reqs, st = _serve_once(serve, rate, duration, budget)
    out = {"workload": f"Poisson {rate:.0f} req/s x {duration:.0f} s, m log-uniform [16,4096], "
                       f"n uniform {{1..128}}, FP64/FP32, all 8 side/uplo/trans cases, "
                       f"25% unit diagonal, alpha in +-[0.5,2], 10% factor reuse, "
                       f"{budget / 2**30:.0f} GiB budget",
           "requests": len(reqs), "completed": st.completed, "rejected": st.rejected,
           "solves_per_s": st.completed / st.wall_seconds, "wall_s": st.wall_seconds,
           "p50_ms": st.p50_latency * 1e3, "p99_ms": st.p99_latency * 1e3,
           "mean_ms": st.mean_latency * 1e3, "max_ms": st.max_latency * 1e3,
           "launches": st.launches, "batched_launches": st.batched_launches,
           "factor_hits": st.factor_hits, "factor_misses": st.factor_misses,
           "evictions": st.evictions, "peak_arena_gib": st.peak_arena_bytes / 2**30,
           "checked_results": st.checked, "over_tolerance": st.over_tolerance,
           "max_backward_error_fp64": st.max_backward_error_fp64,
           "max_backward_error_fp32": st.max_backward_error_fp32}
    curve = []
    best = rate if (st.completed >= 0.95 * len(reqs) * 1.0 and
                    st.p99_latency * 1e3 <= p99_bound_ms) else None
    for r in sweep:
        rq, s2 = _serve_once(serve, r, sweep_seconds, budget, seed=1)
        offered = len(rq) / sweep_seconds
        done_rate = s2.completed / s2.wall_seconds
        curve.append({"offered_rps": offered, "solves_per_s": done_rate,
                      "p50_ms": s2.p50_latency * 1e3, "p99_ms": s2.p99_latency * 1e3,
                      "checked": s2.checked, "over_tolerance": s2.over_tolerance})
        if done_rate >= 0.95 * offered and s2.p99_latency * 1e3 <= p99_bound_ms:
            best = r
        else:
            break
    out["rate_sweep"] = curve
    out["max_sustained_rps"] = best
    out["sustain_rule"] = (f"completed >= 95% of offered and p99 <= {p99_bound_ms:.0f} ms, "
                           f"{sweep_seconds:.0f} s per rate")
    return out
