"""CPU tests of the rocblas-bench / rocblas-test clients (rocblas_amd.clients):
the YAML case expansion against hand-counted documents and against the
reference's own trsm matrix, the reference's data generators, argument
mapping and the printed result line."""
import os

import numpy as np
import pytest

from rocblas_amd import bench_client
from rocblas_amd import clients as C
from rocblas_amd import gtest_client

HERE = os.path.dirname(os.path.abspath(__file__))
MATRIX = os.path.join(HERE, "golden", "trsm_gtest_matrix.json")
REF = "/root/reference"


def test_expansion_rules():
    """Dictionary lists merge item by item (in the declared order), the
    {function: precision} form pairs names with precisions, other lists form a
    cartesian product, ranges expand, Defaults fill the rest, and the
    strided trsm forms get strides from stride_scale (rocblas_gentest.py)."""
    doc = {
        "Defaults": {"alpha": 1.0, "lda": 7, "M": 3, "N": 2, "ldb": 5, "side": "*"},
        "Dictionary lists to expand": ["arguments", "matrix_size", "precision",
                                       {"function": "precision"}],
        "Lists to not expand": [],
        "Tests": [
            {"name": "a", "function": "trsm", "precision": [{"a_type": "f32_r"},
                                                            {"a_type": "f64_r"}],
             "side": ["L", "R"], "uplo": ["L", "U"],
             "matrix_size": [{"M": 4, "N": 4, "lda": 4, "ldb": 4}, {"M": 8, "N": 2}]},
            {"name": "b", "function": {"trsm_batched": {"a_type": "f64_r"},
                                       "trsm": [{"a_type": "f32_r"}, {"a_type": "f64_r"}]},
             "batch_count": "1..3"},
            {"name": "c", "function": "trsm_strided_batched", "side": "R", "N": 6, "lda": 9,
             "ldb": 10, "stride_scale": 2},
        ],
    }
    cases = C.expand_doc(doc)
    a = [c for c in cases if c["name"] == "a"]
    assert len(a) == 2 * 2 * 2 * 2
    assert {(c["M"], c["lda"]) for c in a} == {(4, 4), (8, 7)}  # default lda where not given
    b = [c for c in cases if c["name"] == "b"]
    assert len(b) == (1 + 2) * 3
    assert sorted((c["function"], c["a_type"]) for c in b if c["batch_count"] == 2) == \
        [("trsm", "f32_r"), ("trsm", "f64_r"), ("trsm_batched", "f64_r")]
    (c,) = [c for c in cases if c["name"] == "c"]
    assert c["stride_a"] == 6 * 9 * 2 and c["stride_b"] == 6 * 10 * 2


def test_committed_matrix_counts():
    """The committed matrix expands to the reference's case counts: hand
    counts of trsm_small and trsm_batched_small from the YAML's lists."""
    cases = C.cases_from(MATRIX)
    small = [c for c in cases if c["name"] == "trsm_small"]
    # 2 precisions x 2 sides x 2 uplo x 2 transA x 2 diag x 9 sizes x 2 alpha x 3 api x 2 ws
    assert len(small) == 2 * 16 * 9 * 2 * 3 * 2
    bsmall = [c for c in cases if c["name"] == "trsm_batched_small"]
    # 2 precisions x 4 argument sets x 9 sizes x 2 alpha x 4 batch counts x 3 api x 2 ws
    assert len(bsmall) == 2 * 4 * 9 * 2 * 4 * 3 * 2
    assert {c["transA"] for c in small} == {"N", "C"}
    kinds = {C.case_kind(c) for c in cases}
    assert "run" in kinds and any("f64_c" in k for k in kinds)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "clients", "gtest")),
                    reason="reference tree not present")
def test_committed_matrix_matches_the_reference_yaml():
    """The JSON fixture and the reference's YAML (with its includes) expand to
    the same cases."""
    from_yaml = C.cases_from(os.path.join(REF, "clients", "gtest", "trsm_gtest.yaml"),
                             [os.path.join(REF, "clients", "include")])
    from_json = C.cases_from(MATRIX)
    key = lambda c: repr(sorted((k, repr(v)) for k, v in c.items()))  # noqa: E731
    assert sorted(map(key, from_yaml)) == sorted(map(key, from_json))


def test_select_and_dedupe():
    cases = C.cases_from(MATRIX)
    args, skipped = gtest_client.select(cases, ("quick",))
    assert args and all(a.category == "quick" for a in args)
    assert all(a.precision in ("f32_r", "f64_r") for a in args)
    assert any("f32_c" in k for k in skipped)
    # FORTRAN apis collapse onto the C ones
    assert {a.api for a in args} <= {0, 1, 4, 5}
    labels = [a.label() for a in args]
    assert len(labels) == len(set(labels))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_generated_triangle(uplo, diag):
    """rocblas_init.hpp:263-327: strictly diagonally dominant by rows and
    columns (factor 1.01); client_utility.hpp:419-444 for a unit diagonal."""
    arg = C.Arguments(uplo=uplo, diag=diag, M=20, N=5, lda=23, side="L")
    A = C.init_triangle(arg, np.random.default_rng(1))
    assert A.shape == (23, 20)
    T = A[:20, :20]
    other = np.tril(T, -1) if uplo == "U" else np.triu(T, 1)
    assert not other.any() and not A[20:].any()
    off = np.abs(T - np.diag(np.diag(T)))
    if diag == "N":
        want = (off.sum(0) + off.sum(1)) * 1.01
        assert np.allclose(np.diag(T), np.where(want == 0, 1.0, want))
    else:
        # off-diagonal rows (L) / columns (U) scaled by the old diagonal:
        # dominance now reads |row or column sums| < 1
        sums = off.sum(1) if uplo == "L" else off.sum(0)
        assert (sums < 1).all()


def test_generated_rhs_alternates_sign():
    arg = C.Arguments(M=6, N=4, initialization="rand_int")
    X = C.init_general(arg, np.random.default_rng(2), 6, 4, 9)
    i, j = np.indices((6, 4))
    assert ((X[:6, :4] > 0) == ((i ^ j) & 1).astype(bool)).all()
    assert not X[6:].any()


def test_host_reference_closes_the_loop():
    """B = op(A) X / alpha then the host solve returns X (both sides, both
    transposes), so the checker's forward error is meaningful."""
    rng = np.random.default_rng(3)
    for side in "LR":
        for trans in "NT":
            arg = C.Arguments(side=side, transA=trans, uplo="U", M=9, N=7, lda=10, ldb=11,
                              alpha=-5.0, initialization="hpl")
            A = C.init_triangle(arg, rng)
            X = C.init_general(arg, rng, 9, 7, 11)
            B = C.host_trmm(arg, A, X, 1 / arg.alpha)
            Y = C.host_trsm(arg, A, B, arg.alpha)
            assert C.norm1_rel(X[:9, :7], Y) < 1e-13


def test_from_case_and_cli_mapping():
    case = {"function": "trsm_strided_batched", "a_type": "f32_r", "compute_type": "f32_r",
            "side": "R", "uplo": "U", "transA": "C", "diag": "U", "M": 5, "N": 6, "lda": 6,
            "ldb": 5, "stride_a": 72, "stride_b": 60, "batch_count": 3, "alpha": -5.0,
            "api": "FORTRAN_64", "pointer_mode_device": False, "name": "x"}
    a = C.Arguments.from_case(case)
    assert (a.precision, a.api, a.K, a.pointer_mode_device) == ("f32_r", 3, 6, False)
    ns = bench_client.parser().parse_args(
        "-f trsm_strided_batched -r d --side R --uplo L --transposeA T --diag N -m 32 -n 16 "
        "--alpha 1 --lda 16 --stride_a 10 --ldb 32 --stride_b 512 --batch_count 4".split())
    b = bench_client.args_from_cli(ns)
    assert (b.precision, b.M, b.N, b.K, b.batch_count) == ("f64_r", 32, 16, 16, 4)
    assert b.stride_a == 16 * 16 and b.stride_b == 512  # raised to a whole matrix
    with pytest.raises(SystemExit):
        bench_client.args_from_cli(bench_client.parser().parse_args(["-f", "gemm"]))


def test_result_line_format():
    """argument_model.hpp:56-247: names, then values with ', ' before the
    iteration counts and performance fields."""
    arg = C.Arguments(function="trsm_strided_batched", M=32, N=16, lda=32, ldb=32, stride_a=1024,
                      stride_b=512, batch_count=2, alpha=1.0, iters=10, cold_iters=2,
                      unit_check=0, norm_check=0)
    names, vals = C.log_args(arg, {"gpu_us": 100.0, "cpu_us": 1.0}).split("\n")
    assert names == ("side,uplo,transA,diag,M,N,alpha,lda,stride_a,ldb,stride_b,batch_count,"
                     "cold_iters,hot_iters,rocblas-Gflops,us")
    # 32*16*32 flop per instance, 2 instances, 10 us per call: 3.2768 GFLOP/s
    assert vals == "L,L,N,N,32,16,1,32,1024,32,512,2, 2, 10, 3.2768, 10"
    arg.norm_check = 1
    names, vals = C.log_args(arg, {"gpu_us": 100.0, "cpu_us": 4.0, "err_host": 1e-17,
                                   "err_device": 2e-17}).split("\n")
    assert names.endswith("us,CPU-Gflops,CPU-us,norm_error_1,norm_error_2")
    assert vals.endswith(",8.192,4,1e-17,2e-17")
