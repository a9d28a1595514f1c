"""GPU tests of the rocblas-test / rocblas-bench clients and of ROCBLAS_LAYER
logging: the reference's trsm gtest matrix (tests/golden/trsm_gtest_matrix.json,
from clients/gtest/trsm_gtest.yaml) runs against the library with the
reference's procedures and tolerances, a rocblas-bench command prints the
reference's result line, and the trace / bench / profile logs of a child
process read like the reference's and replay through the bench client."""
import io
import os
import subprocess
import sys
import textwrap

import pytest
import torch

import rocblas_amd as rb
from rocblas_amd import bench_client
from rocblas_amd import clients as C
from rocblas_amd import gtest_client
from tests.gpu_util import say

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MATRIX = os.path.join(HERE, "golden", "trsm_gtest_matrix.json")

GROUPS = {
    "trsm": ("trsm",),
    "trsm_batched": ("trsm_batched",),
    "trsm_strided_batched": ("trsm_strided_batched",),
    "ex": ("trsm_ex", "trsm_batched_ex", "trsm_strided_batched_ex"),
    "bad_arg_internal": C.BAD_ARG_FUNCS + C.INTERNAL_FUNCS,
}


@pytest.fixture(scope="module")
def handle():
    assert torch.cuda.is_available()
    h = rb.Handle()
    yield h
    h.close()


@pytest.fixture(scope="module")
def cases():
    return C.cases_from(MATRIX)


def _run(handle, args, what):
    say(what, len(args), "cases")
    out = io.StringIO()
    failures = gtest_client.run(args, handle.ptr, out=out)
    assert not failures, "\n".join(f"{l}: {m}" for l, m in failures[:20])


@pytest.mark.parametrize("group", list(GROUPS))
def test_gtest_matrix_quick(handle, cases, group):
    """Every runnable case of the reference's 'quick' category (real
    precisions; the FORTRAN apis fold onto the C entry points)."""
    args, _ = gtest_client.select(cases, ("quick",))
    args = [a for a in args if a.function in GROUPS[group]]
    assert args
    _run(handle, args, f"quick {group}")


@pytest.mark.parametrize("group", ["trsm", "trsm_batched", "trsm_strided_batched", "ex"])
def test_gtest_matrix_pre_checkin(handle, cases, group):
    """The 'pre_checkin' category (medium sizes, substitution sizes, the
    internal api with 2^32 element offsets, graph capture), every third case
    in a fixed order so the run stays short; the CLI runs all of them."""
    args, _ = gtest_client.select(cases, ("pre_checkin",))
    args = [a for a in args if a.function in GROUPS[group]][::3]
    assert args
    _run(handle, args, f"pre_checkin {group}")


def test_bench_client_line(capsys):
    """rocblas-bench -f trsm_strided_batched prints the reference's two lines."""
    rc = bench_client.main("-f trsm_strided_batched -r d --side R --uplo U --transposeA N "
                           "--diag N -m 32 -n 16 --lda 16 --ldb 32 --stride_a 256 "
                           "--stride_b 512 --batch_count 100 -i 3 -j 1 -v 1".split())
    assert rc == 0
    names, vals = capsys.readouterr().out.strip().split("\n")[-2:]
    assert names.startswith("side,uplo,transA,diag,M,N,alpha,lda,stride_a,ldb,stride_b,"
                            "batch_count,cold_iters,hot_iters,rocblas-Gflops,us,CPU-Gflops")
    v = vals.split(",")
    assert v[:12] == ["R", "U", "N", "N", "32", "16", "1", "16", "256", "32", "512", "100"]
    assert float(v[-2]) < 40 * 2.2e-16 * 32 and float(v[-1]) < 40 * 2.2e-16 * 32
    assert float(v[14]) > 0


def test_bench_client_gemm_line(capsys):
    """rocblas-bench -f gemm (testing_gemm.hpp's result line) on the update GEMM,
    with the host check."""
    rc = bench_client.main("-f gemm -r d --transposeA T --transposeB N -m 640 -n 520 -k 264 "
                           "--lda 300 --ldb 264 --ldc 640 --alpha -1 --beta 1 -i 2 -j 1 "
                           "-v 1".split())
    assert rc == 0
    names, vals = capsys.readouterr().out.strip().split("\n")[-2:]
    assert names == ("transA,transB,M,N,K,alpha,lda,beta,ldb,ldc,cold_iters,hot_iters,"
                     "rocblas-Gflops,us,CPU-Gflops,CPU-us,norm_error_1")
    v = vals.split(",")
    assert v[:10] == ["T", "N", "640", "520", "264", "-1", "300", "1", "264", "640"]
    assert float(v[-1]) < 1e-14


CHILD = textwrap.dedent("""
    import ctypes, sys, torch
    sys.path.insert(0, {root!r})
    import rocblas_amd as rb
    L = rb.lib()
    h = rb.Handle()
    A = torch.eye(64, dtype=torch.float64, device="cuda").flatten()
    B = torch.ones(64 * 8, dtype=torch.float64, device="cuda")
    a = ctypes.c_double(2.0)
    for _ in range(2):
        L.rocblas_dtrsm(h.ptr, 141, 122, 111, 131, 64, 8, ctypes.addressof(a), A.data_ptr(), 64,
                        B.data_ptr(), 64)
    L.rocblas_strsm_strided_batched(h.ptr, 142, 121, 112, 132, 8, 4, ctypes.addressof(
        ctypes.c_float(0.5)), A.data_ptr(), 4, 16, B.data_ptr(), 8, 32, 3)
    pa = torch.tensor([A.data_ptr()], dtype=torch.int64, device="cuda")
    pb = torch.tensor([B.data_ptr()], dtype=torch.int64, device="cuda")
    L.rocblas_dtrsm_batched_64(h.ptr, 141, 121, 113, 131, 16, 2, ctypes.addressof(a),
                               pa.data_ptr(), 64, pb.data_ptr(), 64, 1)
    d = torch.tensor([3.0], dtype=torch.float64, device="cuda")
    h.set_pointer_mode(1)
    L.rocblas_dtrsm(h.ptr, 141, 122, 111, 131, 8, 8, d.data_ptr(), A.data_ptr(), 64,
                    B.data_ptr(), 64)
    h.set_pointer_mode(0)
    L.rocblas_dtrsm(h.ptr, 141, 122, 111, 131, -1, 8, ctypes.addressof(a), A.data_ptr(), 64,
                    B.data_ptr(), 64)
    torch.cuda.synchronize()
    h.close()
    print("child done", flush=True)
""")


def test_rocblas_layer_logs(tmp_path):
    """ROCBLAS_LAYER=7: trace, bench and profile logs in the reference's
    formats (logging.hpp:150-275, rocblas_ostream.hpp:325-520, the trsm
    _imp.hpp files), one line per call including a call with bad arguments,
    alpha read from the device in device pointer mode, and the profile
    written at exit with call counts."""
    env = dict(os.environ, ROCBLAS_LAYER="7",
               ROCBLAS_LOG_TRACE_PATH=str(tmp_path / "trace.log"),
               ROCBLAS_LOG_BENCH_PATH=str(tmp_path / "bench.log"),
               ROCBLAS_LOG_PROFILE_PATH=str(tmp_path / "profile.yaml"))
    say("logging child")
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT)], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    trace = (tmp_path / "trace.log").read_text().splitlines()
    bench = (tmp_path / "bench.log").read_text().splitlines()
    prof = (tmp_path / "profile.yaml").read_text().splitlines()
    assert len(trace) == len(bench) == 6
    t = trace[0].split(",")
    assert t[:8] == ["rocblas_dtrsm", "L", "L", "N", "N", "64", "8", "2"]
    assert t[8].startswith("0x") and t[9] == "64" and t[11] == "64"
    assert t[-1] == "atomics_allowed"
    assert trace[2].split(",")[:8] == ["rocblas_strided_batched_strsm", "R", "U", "T", "U", "8",
                                       "4", "0.5"]
    assert trace[3].startswith("rocblas_batched_dtrsm_64,L,U,C,N,16,2,2,")
    assert trace[4].split(",")[7] == "3"  # device-mode alpha
    assert trace[5].split(",")[5] == "-1"  # logged before the argument check
    assert bench[0] == ("./rocblas-bench -f trsm -r f64_r --side L --uplo L --transposeA N "
                        "--diag N -m 64 -n 8 --alpha 2 --lda 64 --ldb 64")
    assert bench[2] == ("./rocblas-bench -f trsm_strided_batched -r f32_r --side R --uplo U "
                        "--transposeA T --diag U -m 8 -n 4 --alpha 0.5 --lda 4 --stride_a 16 "
                        "--ldb 8 --stride_b 32 --batch_count 3")
    assert bench[3].startswith("./rocblas-bench --api 1 -f trsm_batched -r f64_r")
    assert bench[3].endswith("--batch_count 1")
    first = [p for p in prof if '"rocblas_dtrsm"' in p and "m: 64," in p]
    assert first == ['- { rocblas_function: "rocblas_dtrsm", atomics_mode: atomics_allowed, '
                     "side: 'L', uplo: 'L', transA: 'N', diag: 'N', m: 64, n: 8, lda: 64, "
                     "ldb: 64, call_count: 2 }"]
    assert len(prof) == 5
    # every valid logged bench line replays through the bench client
    for line in bench[:5]:
        argv = line.split()[1:] + ["-i", "1", "-j", "0"]
        arg = bench_client.args_from_cli(bench_client.parser().parse_args(argv))
        with rb.Handle() as h:
            C.run_case(h.ptr, arg)
