"""Benchmark: plain FP64 rocblas_dtrsm_strided_batched on one MI355X.

Workload (one "step"): 20000 independent triangular solves of m=32, n=16 for
each of the eight side/uplo/trans cases {L,R} x {L,U} x {N,T}, diag=N, alpha=1,
lda = k, ldb = m, strides lda*k and ldb*n.  A = (diagonal U[1,2], off-diagonal
U[-1,1]/k) / k, B = U[-1,1], SplitMix64 seed 0 (rocblas_amd.synth).  Inputs are
resident in HBM before timing; each step solves in place (eight launches).

Printed JSON (one line): value = solves per second over the timed steps;
roofline = the side-left kernel's algorithmic HBM bytes (8*(k(k+1)/2 + 2mn) per
instance) over its average launch duration measured with HIP events on the
launch stream, beside the HBM rate a streaming kernel reaches for the same
read/write mix on this device (roofline.measured_ceiling); cpu_baseline = the
oracle's netlib loop nest, OpenMP over instances, on this host's cores (a
bounded sample of the same workload).

Further workloads (single 512x256 solve, 16384^2 plain and mixed-precision
solves, a 4096-instance variable-size batch, the Poisson serving loop) are
measured after the main line by bench_legs.py and reported under "workloads".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--legs all|none|single,large,vbatch,trtri,serve] [--serve-seconds S]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per side-left launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE
# (copy-calibrated); see the profiles/ summary named below.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "small_trsm_strided_batched_m32n16.traffic.json")
CASES = [(s, u, t) for s in "LR" for u in "LU" for t in "NT"]
BATCH, M, N = 20000, 32, 16


def algo_bytes(side):
    k = M if side == "L" else N
    return 8 * (k * (k + 1) // 2 + 2 * M * N)


def algo_flops(side):
    k = M if side == "L" else N
    return M * N * k


def make_inputs(torch, synth):
    data = {}
    Bsrc = {}
    for side in "LR":
        Bm = synth.rhs_uniform(M, N, 0, batch=BATCH)
        Bsrc[side] = np.ascontiguousarray(np.swapaxes(Bm, 1, 2)).ravel()
    for side, uplo, trans in CASES:
        k = M if side == "L" else N
        A = synth.tri_uniform(k, uplo, 0, batch=BATCH, scale=1.0 / k)
        Af = np.ascontiguousarray(np.swapaxes(A, 1, 2)).ravel()
        data[(side, uplo, trans)] = (Af, torch.from_numpy(Af).cuda(),
                                     torch.from_numpy(Bsrc[side].copy()).cuda())
    return data, Bsrc


def cpu_baseline(oracle, data, Bsrc, budget_s=12.0, cblas=False):
    """All eight cases of the workload on the host, OpenMP over instances,
    repeated until ~budget_s of wall time (at least one full pass): the
    oracle's netlib loop nest, or (cblas=True) the host CBLAS dtrsm loaded by
    the oracle library (the reference's own host comparator is cblas_dtrsm)."""
    threads = oracle.num_threads()
    solve = oracle.cblas_trsm_batched_omp if cblas else oracle.trsm_batched_omp
    reps, t_total, solves = 0, 0.0, 0
    while reps == 0 or t_total < budget_s:
        for (side, uplo, trans), (Af, _, _) in data.items():
            k = M if side == "L" else N
            Bw = Bsrc[side].copy()
            t0 = time.perf_counter()
            solve(side, uplo, trans, "N", M, N, 1.0, Af, k, k * k, Bw, M, M * N, BATCH, threads)
            t_total += time.perf_counter() - t0
            solves += BATCH
        reps += 1
        if reps >= 2000:
            break
    out = {"value": solves / t_total, "unit": "solves/s", "cores": threads, "kind": "port",
           "sample": f"{reps} pass(es) over all 8 cases x {BATCH} instances (m={M}, n={N}), "
                     f"{t_total:.1f} s"}
    if cblas:
        out["kind"] = "host CBLAS " + oracle.cblas_load()
    return out


def vendor_strided(torch, data, reps=20):
    """The vendor librocblas (shipped with torch) on the same eight launches,
    through its C API: solves/s over reps steps."""
    import ctypes
    import bench_legs as bl
    vendor = bl.vendor_lib(torch)
    if vendor is None:
        return None
    V, vh = vendor
    vp = ctypes.c_void_p
    V.rocblas_dtrsm_strided_batched.argtypes = [vp] + [ctypes.c_int] * 6 + [
        vp, vp, ctypes.c_int, ctypes.c_int64, vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    enums = {"L": 141, "R": 142}, {"U": 121, "L": 122}, {"N": 111, "T": 112}
    one = ctypes.c_double(1.0)

    def vstep():
        for (side, uplo, trans), (_, dA, dB) in data.items():
            k = M if side == "L" else N
            st = V.rocblas_dtrsm_strided_batched(
                vh, enums[0][side], enums[1][uplo], enums[2][trans], 131, M, N,
                ctypes.addressof(one), vp(dA.data_ptr()), k, k * k, vp(dB.data_ptr()), M, M * N,
                BATCH)
            assert st == 0, st
    for _ in range(3):
        vstep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        vstep()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    V.rocblas_destroy_handle(vh)
    return {"solves_per_s": len(CASES) * BATCH / t, "ms_per_step": t * 1e3,
            "library": "librocblas.so shipped with torch, C API"}


def check_accuracy(torch, rb, h, data, Bsrc):
    """Fresh solve of every case; per-instance normwise backward error in float64."""
    worst = 0.0
    for (side, uplo, trans), (Af, dA, _) in data.items():
        k = M if side == "L" else N
        dB = torch.from_numpy(Bsrc[side].copy()).cuda()
        rb.check(rb.trsm_strided_batched(h, side, uplo, trans, "N", M, N, 1.0, dA.data_ptr(), k,
                                         k * k, dB.data_ptr(), M, M * N, BATCH))
        X = dB.view(BATCH, N, M).transpose(1, 2)
        B0 = torch.from_numpy(Bsrc[side]).cuda().view(BATCH, N, M).transpose(1, 2)
        Am = dA.view(BATCH, k, k).transpose(1, 2)
        Am = torch.nan_to_num(Am)
        T = torch.triu(Am) if uplo == "U" else torch.tril(Am)
        op = T.transpose(1, 2).contiguous() if trans != "N" else T.contiguous()
        X = X.contiguous()
        R = B0 - (torch.bmm(op, X) if side == "L" else torch.bmm(X, op))
        n1 = lambda Z: Z.abs().sum(1).amax(1)
        eta = n1(R) / (n1(op) * n1(X) + n1(B0))
        worst = max(worst, float(eta.max()))
    return worst


def hbm_ceiling(torch, rb, h, units=12):
    """The HBM rate reachable for the side-left launch's read/write mix (about
    5:2 fetched to written): rocblas_amd_hbm_stream over 5 x units access
    units read and 2 x units written (about 1 GB and 0.4 GB), the best of six
    launch shapes, 5 passes each."""
    step = rb.hbm_stream_step_bytes(h)
    rd, wr = 5 * units * step, 2 * units * step
    src = torch.ones(rd // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty(wr // 4, dtype=torch.int32, device="cuda")
    sec = rb.hbm_stream(h, src.data_ptr(), rd, dst.data_ptr(), wr, iters=5)
    del src, dst
    return {"value": (rd + wr) / sec / 1e9, "unit": "GB/s", "mix": "read:write 5:2",
            "probe": "rocblas_amd_hbm_stream (16 B per lane per access; best of 4/8/16 "
                     "waves per SIMD, plain and non-temporal)"}


def headline_ratios(work):
    """The metric's comparisons with the plain FP64 solve, lifted from the
    workloads: the precision form over plain FP64 (TFLOP/s at 16384^2), the
    variable-size batch's graph launch over a per-instance loop of plain
    dtrsm (solves/s), and the latencies the metric names."""
    r = {}
    lg = work.get("large_16384", {})
    plain = lg.get("plain_fp64", {}).get("tflops")
    for form in ("mixed_fp32_refined", "mixed_bf16x3_refined"):
        f = lg.get(form, {})
        if plain and f.get("tflops"):
            r[f"large_16384_{form}_over_plain_tflops"] = f["tflops"] / plain
            r[f"large_16384_{form}_backward_error"] = f.get("backward_error")
    vb = work.get("variable_batch_4096", {})
    if "speedup_vs_loop" in vb:
        r["variable_batch_graph_over_dtrsm_loop_solves"] = vb["speedup_vs_loop"]
    if "speedup_vs_vendor_loop" in vb:
        r["variable_batch_graph_over_vendor_dtrsm_loop_solves"] = vb["speedup_vs_vendor_loop"]
    sv = work.get("serving_poisson_2000rps", {})
    if "p99_ms" in sv:
        r["serving_p50_ms"] = sv["p50_ms"]
        r["serving_p99_ms"] = sv["p99_ms"]
    sg = work.get("single_512x256", {}).get("ours", {})
    if "p99_us" in sg:
        r["single_512x256_p99_us"] = sg["p99_us"]
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="all",
                    help="further workloads: all, none, or a comma list of "
                         "single,large,vbatch,serve")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--vbatch-eager", action="store_true",
                    help="time the variable-size batch as a direct launch instead of a graph "
                         "replay (for profilers whose counter collection does not follow graph "
                         "launches)")
    ap.add_argument("--no-vendor", action="store_true",
                    help="do not time the vendor librocblas beside the further workloads")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain device copy of 248,320,000 bytes (the side-L "
                         "launch's algorithmic bytes), so a rocprofv3 --pmc run of this "
                         "command reads the counters of both")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the triangular-solve path is single-GPU here "
              f"(--gpus {args.gpus} not supported)", file=sys.stderr)
        sys.exit(2)

    import torch

    import rocblas_amd as rb
    from rocblas_amd import synth

    torch.cuda.set_device(0)
    h = rb.Handle()
    stream = torch.cuda.current_stream()
    rb.check(h.set_stream(stream.cuda_stream))
    data, Bsrc = make_inputs(torch, synth)

    def step(events=None):
        for idx, ((side, uplo, trans), (_, dA, dB)) in enumerate(data.items()):
            k = M if side == "L" else N
            if events is not None:
                events[idx][0].record(stream)
            st = rb.trsm_strided_batched(h, side, uplo, trans, "N", M, N, 1.0, dA.data_ptr(), k,
                                         k * k, dB.data_ptr(), M, M * N, BATCH)
            if events is not None:
                events[idx][1].record(stream)
            if st != rb.rocblas_status.success:
                raise rb.RocblasError(st, "trsm_strided_batched")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # Per-launch durations on the launch stream (each call is one kernel
    # launch): events bracket the four side-left launches and then the four
    # side-right launches of a step, so one event pair is amortised over four
    # back-to-back launches; each group's time / 4 is that kernel's average
    # launch duration.  Single-launch event pairs give the latency
    # percentiles.
    reps = max(10, min(args.steps, 50))
    left_us, right_us = [], []
    for _ in range(reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record(stream)
        for side_group, e in (("L", ev[1]), ("R", ev[2])):
            for (side, uplo, trans), (_, dA, dB) in data.items():
                if side != side_group:
                    continue
                k = M if side == "L" else N
                rb.check(rb.trsm_strided_batched(h, side, uplo, trans, "N", M, N, 1.0,
                                                 dA.data_ptr(), k, k * k, dB.data_ptr(), M,
                                                 M * N, BATCH))
            e.record(stream)
        torch.cuda.synchronize()
        left_us.append(ev[0].elapsed_time(ev[1]) * 1e3 / 4)
        right_us.append(ev[1].elapsed_time(ev[2]) * 1e3 / 4)
    left_us, right_us = np.array(left_us), np.array(right_us)
    per_case = {c: [] for c in data}
    for _ in range(reps):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in data]
        step(ev)
        torch.cuda.synchronize()
        for idx, c in enumerate(data):
            per_case[c].append(ev[idx][0].elapsed_time(ev[idx][1]) * 1e3)  # us
    all_us = np.concatenate([np.array(v) for v in per_case.values()])
    left_avg_s = float(left_us.mean()) * 1e-6
    achieved = BATCH * algo_bytes("L") / left_avg_s / 1e9

    eta = check_accuracy(torch, rb, h, data, Bsrc)
    ceiling = hbm_ceiling(torch, rb, h)
    if args.copy_calibration:
        # read 124,160,000 B + write 124,160,000 B
        src = torch.ones(BATCH * algo_bytes("L") // 16, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
    solves = len(CASES) * BATCH
    flops = sum(BATCH * algo_flops(c[0]) for c in CASES)
    with open(os.path.join(ROOT, "BASELINE.json")) as f:
        metric = json.load(f)["metric"]
    out = {
        "metric": metric,
        "value": solves / (ms_per_step * 1e-3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 seed 0)",
        "config": {"workload": "dtrsm_strided_batched batch=20000 m=32 n=16, "
                               "side/uplo/trans in {L,R}x{L,U}x{N,T}, diag=N, alpha=1",
                   "batch": BATCH, "m": M, "n": N, "cases": 8},
        "tflops": flops / (ms_per_step * 1e-3) / 1e12,
        "max_backward_error": eta,
        "latency_us": {"p50": float(np.percentile(all_us, 50)),
                       "p99": float(np.percentile(all_us, 99)),
                       "left_avg": float(left_us.mean()), "right_avg": float(right_us.mean())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "small_trsm_dpp_kernel<double,32,*,*,cols> (the 4 side-left cases)",
                     "algorithmic_bytes_per_launch": BATCH * algo_bytes("L"),
                     "avg_launch_us": float(left_us.mean()),
                     "timing": "HIP events on the launch stream around 4 back-to-back launches",
                     "measured_ceiling": ceiling},
    }
    if traffic and ceiling:
        rate = traffic / left_avg_s / 1e9
        out["roofline"]["traffic_rate"] = rate
        out["roofline"]["traffic_frac_of_measured_ceiling"] = rate / ceiling["value"]
    if not args.no_vendor:
        v = vendor_strided(torch, data)
        if v is not None:
            v["ours_over_vendor"] = out["value"] / v["solves_per_s"]
            out["vendor_rocblas"] = v
    import oracle
    oracle.build()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(oracle, data, Bsrc)
        if oracle.cblas_load() is not None:
            out["cpu_baseline_cblas"] = cpu_baseline(oracle, data, Bsrc, budget_s=8.0, cblas=True)
    h.close()
    legs = {"single", "large", "vbatch", "trtri", "serve"} if args.legs == "all" else \
        set() if args.legs == "none" else set(args.legs.split(","))
    if legs:
        import bench_legs as bl
        del data
        torch.cuda.empty_cache()
        vendor = None if args.no_vendor else bl.vendor_lib(torch)
        work = {}
        runs = [("single_512x256", "single", lambda: bl.leg_single(torch, rb, synth, oracle, vendor)),
                ("large_16384", "large", lambda: bl.leg_large(torch, rb, oracle, vendor)),
                ("variable_batch_4096", "vbatch",
                 lambda: bl.leg_variable_batch(torch, rb, vendor, eager=args.vbatch_eager)),
                ("trtri", "trtri", lambda: bl.leg_trtri(torch, rb, vendor)),
                ("serving_poisson_2000rps", "serve",
                 lambda: bl.leg_serving(torch, rb, duration=args.serve_seconds))]
        for name, key, fn in runs:
            if key not in legs:
                continue
            print(f"[bench] {name} ...", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            try:
                work[name] = fn()
            except Exception as e:  # report, keep the main line
                work[name] = {"error": f"{type(e).__name__}: {e}"}
            work[name]["leg_seconds"] = time.perf_counter() - t0
        out["workloads"] = work
        out["vs_plain_fp64"] = headline_ratios(work)
    if eta > 1e-12:
        print(f"backward error {eta:.3e} above 1e-12", file=sys.stderr)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
