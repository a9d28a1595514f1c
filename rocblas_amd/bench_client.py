"""rocblas-bench for the trsm path: the reference's flags
(clients/benchmarks/client.cpp:1430-1704), its timing loop
(testing_trsm.hpp:547-570) and its result line (argument_model.hpp:56-247),
over this library.  A bench line logged with ROCBLAS_LAYER=2 replays as is:

    ROCBLAS_LAYER=2 python app.py 2> calls.log      # ./rocblas-bench -f trsm ...
    python -m rocblas_amd.bench_client -f trsm -r f64_r --side L --uplo L \\
        --transposeA N --diag N -m 512 -n 256 --alpha 1 --lda 512 --ldb 512

``--yaml FILE`` times every case of a gtest-format YAML (or its JSON form)
instead.  Functions: trsm, trsm_batched, trsm_strided_batched and their _ex
forms, real single and double precision; and gemm, the update GEMM the
blocked solve runs on (rocblas_amd_{s,d}gemm).  ``--api 1`` calls the ILP64 _64
entry points; the Fortran apis call the C entry points (no Fortran module
is built).
"""
import argparse
import sys

from . import Handle
from . import clients as C


def parser():
    p = argparse.ArgumentParser(prog="rocblas-bench (trsm)", description=__doc__,
                                formatter_class=argparse.RawDescriptionHelpFormatter)
    a = p.add_argument
    a("-m", "--sizem", dest="M", type=int, default=128)
    a("-n", "--sizen", dest="N", type=int, default=128)
    a("-k", "--sizek", dest="K", type=int, default=128, help="not used by trsm")
    a("--lda", type=int, default=128)
    a("--ldb", type=int, default=128)
    a("--ldc", type=int, default=128, help="not used by trsm")
    a("--ldd", type=int, default=128, help="not used by trsm")
    a("--any_stride", type=int, default=0, help="keep strides smaller than a matrix")
    a("--stride_a", type=int, default=128 * 128)
    a("--stride_b", type=int, default=128 * 128)
    a("--alpha", type=float, default=1.0)
    a("--alphai", type=float, default=0.0, help="real types only: must be 0")
    a("--beta", type=float, default=0.0, help="not used by trsm")
    a("-f", "--function", default="trsm")
    a("-r", "--precision", default="f32_r")
    a("--a_type")
    a("--b_type")
    a("--c_type")
    a("--d_type")
    a("--compute_type")
    a("--initialization", "--init", default="hpl", choices=["rand_int", "trig_float", "hpl"])
    a("--transposeA", default="N")
    a("--transposeB", default="N", help="not used by trsm")
    a("--side", default="L")
    a("--uplo", default="U")
    a("--diag", default="N")
    a("--batch_count", type=int, default=1)
    a("-t", "--test", dest="unit_check", type=int, default=0)
    a("-v", "--verify", dest="norm_check", type=int, default=0)
    a("-i", "--iters", type=int, default=10)
    a("-j", "--cold_iters", type=int, default=2)
    a("--device", type=int, default=0)
    a("--api", type=int, default=0, help="0 C, 1 C_64, 2 FORTRAN, 3 FORTRAN_64")
    a("--fortran", action="store_true")
    a("--workspace", type=int, default=0, help="caller-owned workspace bytes")
    a("--flush_batch_count", type=int, default=1)
    a("--flush_memory_size", type=int, default=0)
    a("--log_function_name", action="store_true")
    a("--log_datatype", action="store_true")
    a("-z", type=int, default=0, help="set --log_function_name and --log_datatype")
    a("--atomics_allowed", action="store_true", help="accepted; trsm uses no atomics")
    a("--atomics_not_allowed", action="store_true", help="accepted; trsm uses no atomics")
    a("--yaml", help="time every case of a gtest-format YAML (or .json) file")
    a("-I", "--include", action="append", default=[], help="YAML include directory")
    a("--name_filter", default="")
    a("--function_filter", default="")
    a("--seed", type=int, default=0)
    return p


def bench_gemm(handle, ns, out=None):
    """-f gemm: the library's update GEMM (rocblas_amd_{s,d}gemm, the kernels
    of the blocked solve's trailing updates) with rocblas-bench's gemm flags,
    timing loop and result line (testing_gemm.hpp:430-500: transA, transB, M,
    N, K, alpha, lda, beta, ldb, ldc, then cold/hot iterations, rocblas-Gflops
    at 2 M N K flop, us; -v 1 adds the CPU figures and the relative 1-norm
    error against a host FP64 product)."""
    import time

    import numpy as np
    import torch

    import rocblas_amd as rb

    prec = C.PRECISIONS.get((ns.compute_type or ns.precision).lower())
    if prec is None:
        raise SystemExit("Invalid value for --precision (this library: s, d, f32_r, f64_r)")
    p = "d" if prec == "f64_r" else "s"
    dt = torch.float64 if p == "d" else torch.float32
    ta, tb = ns.transposeA.upper(), ns.transposeB.upper()
    M, N, K = ns.M, ns.N, ns.K
    ar, ac = (M, K) if ta == "N" else (K, M)
    br, bc = (K, N) if tb == "N" else (N, K)
    lda, ldb, ldc = max(ns.lda, ar), max(ns.ldb, br), max(ns.ldc, M)
    g = torch.Generator(device="cuda").manual_seed(ns.seed)
    A = torch.rand(lda * ac, dtype=dt, device="cuda", generator=g) - 0.5
    B = torch.rand(ldb * bc, dtype=dt, device="cuda", generator=g) - 0.5
    C0 = torch.rand(ldc * N, dtype=dt, device="cuda", generator=g) - 0.5
    Cd = C0.clone()
    call = lambda: rb.gemm(handle, ta, tb, M, N, K, ns.alpha, A.data_ptr(), lda,  # noqa: E731
                           B.data_ptr(), ldb, ns.beta, Cd.data_ptr(), ldc, precision=p)
    st = call()
    if st != rb.rocblas_status.success:
        raise SystemExit(f"rocblas_amd_{p}gemm: {st.name}")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for i in range(ns.cold_iters + ns.iters):
        if i == ns.cold_iters:
            e0.record()
        call()
    e1.record()
    torch.cuda.synchronize()
    gpu_us = e0.elapsed_time(e1) * 1e3 / max(ns.iters, 1)
    gflop = 2.0 * M * N * max(K, 1) / 1e9
    names = "transA,transB,M,N,K,alpha,lda,beta,ldb,ldc,cold_iters,hot_iters,rocblas-Gflops,us"
    vals = (f"{ta},{tb},{M},{N},{K},{ns.alpha:g},{lda},{ns.beta:g},{ldb},{ldc}, "
            f"{ns.cold_iters}, {ns.iters}, {gflop / gpu_us * 1e6:g}, {gpu_us:g}")
    if ns.norm_check or ns.unit_check:
        f = lambda t, r, c, ld: t.cpu().numpy().astype(np.float64).reshape(c, ld).T[:r]  # noqa
        opA = f(A, ar, ac, lda) if ta == "N" else f(A, ar, ac, lda).T
        opB = f(B, br, bc, ldb) if tb == "N" else f(B, br, bc, ldb).T
        t0 = time.perf_counter()
        want = ns.alpha * (opA @ opB) + ns.beta * f(C0, M, N, ldc)
        cpu_us = (time.perf_counter() - t0) * 1e6
        Cd.copy_(C0)
        call()
        torch.cuda.synchronize()
        err = np.abs(f(Cd, M, N, ldc) - want).sum() / max(np.abs(want).sum(), 1e-300)
        names += ",CPU-Gflops,CPU-us,norm_error_1"
        vals += f",{gflop / cpu_us * 1e6:g},{cpu_us:g},{err:g}"
    print(names + "\n" + vals, file=out or sys.stdout, flush=True)
    return {"gpu_us": gpu_us}


def args_from_cli(ns):
    prec = (ns.compute_type or ns.precision).lower()
    if prec not in C.PRECISIONS:
        raise SystemExit(f"Invalid value for --precision {prec} (this library: s, d, f32_r, f64_r)")
    if ns.function not in C.SOLVE_FUNCS:
        raise SystemExit(f"Invalid value for --function {ns.function} "
                         f"(this library: {', '.join(C.SOLVE_FUNCS)})")
    if ns.M < 0 or ns.N < 0:
        raise SystemExit("Invalid value for -m / -n")
    if ns.alphai:
        raise SystemExit("--alphai: real types only")
    api = ns.api or (2 if ns.fortran else 0)
    arg = C.Arguments(function=ns.function, precision=C.PRECISIONS[prec], side=ns.side.upper(),
                      uplo=ns.uplo.upper(), transA=ns.transposeA.upper(), diag=ns.diag.upper(),
                      M=ns.M, N=ns.N, lda=ns.lda, ldb=ns.ldb, stride_a=ns.stride_a,
                      stride_b=ns.stride_b, batch_count=ns.batch_count, alpha=ns.alpha, api=api,
                      user_allocated_workspace=ns.workspace, initialization=ns.initialization,
                      unit_check=ns.unit_check, norm_check=ns.norm_check, timing=1,
                      iters=ns.iters, cold_iters=ns.cold_iters,
                      flush_batch_count=ns.flush_batch_count,
                      flush_memory_size=ns.flush_memory_size, name="rocblas-bench", seed=ns.seed)
    if "strided" in arg.function and not ns.any_stride:
        arg.stride_a = max(arg.stride_a, arg.lda * arg.K)
        arg.stride_b = max(arg.stride_b, arg.ldb * arg.N)
    return arg


def bench_one(handle, arg, log_fn=False, log_dt=False, out=None):
    res = C.run_case(handle, arg)
    if res and "gpu_us" in res:
        print(C.log_args(arg, res, log_fn, log_dt), file=out or sys.stdout, flush=True)
    return res


def main(argv=None):
    ns = parser().parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("rocblas-bench: no GPU")
    torch.cuda.set_device(ns.device)
    log_fn = ns.log_function_name or bool(ns.z)
    log_dt = ns.log_datatype or bool(ns.z)
    with Handle() as h:
        if ns.yaml:
            for case in C.cases_from(ns.yaml, ns.include):
                if C.case_kind(case) != "run" or case["function"] not in C.SOLVE_FUNCS:
                    continue
                if ns.name_filter and ns.name_filter not in str(case.get("name", "")):
                    continue
                if ns.function_filter and ns.function_filter not in case["function"]:
                    continue
                arg = C.Arguments.from_case(case)
                arg.timing, arg.unit_check = 1, 0
                bench_one(h.ptr, arg, log_fn, log_dt)
        elif ns.function == "gemm":
            bench_gemm(h, ns)
        else:
            bench_one(h.ptr, args_from_cli(ns), log_fn, log_dt)
    return 0


if __name__ == "__main__":
    sys.exit(main())
