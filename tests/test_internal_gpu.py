"""GPU tests of the rocSOLVER-facing C++ templates (include/rocblas_amd_internal.hpp),
called through their mangled names: offsets and strides address the same data as
the public entry points, whose results they reproduce bit for bit; the workspace
queries report that nothing is needed (library/src/blas3/rocblas_trsm_kernels.hpp:
1650-1769 answers rocblas_status_continue with zero sizes for its substitution
paths); trtri with sub-batches and in place through w_C_tmp."""
import ctypes
import subprocess

import numpy as np
import pytest
import torch

import oracle
import rocblas_amd as rb
from tests.gpu_util import GpuHandle, dev, ptr_array, say

pytestmark = pytest.mark.gpu
S = rb.rocblas_status
_vp, _i32, _i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available()
    oracle.build()
    hh = GpuHandle()
    yield hh
    hh.close()


def mangled(prefix):
    out = subprocess.run(["nm", "-D", "--defined-only", rb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    dem = subprocess.run(["nm", "-DC", "--defined-only", rb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for raw, d in zip(out.splitlines(), dem.splitlines()):
        if d.split(" ", 2)[2].startswith(prefix):
            return raw.split()[-1]
    raise KeyError(prefix)


def fn(prefix, argtypes):
    f = getattr(rb.lib(), mangled(prefix))
    f.restype = ctypes.c_int32
    f.argtypes = argtypes
    return f


TRSM_ARGS = [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _i32, _i64, _vp, _i64, _i32,
             _i64, _i32, ctypes.c_bool, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i64]
TRTRI_ARGS = [_vp, _i32, _i32, _i32, _vp, _i64, _i32, _i64, _i64, _vp, _i64, _i32, _i64, _i64,
              _i32, _i32, _vp]
WS_ARGS = [_i32, _i32, _i32, _i32, _i32, _i32] + [ctypes.POINTER(ctypes.c_size_t)] * 5

SIDE = {"L": 141, "R": 142}
FILL = {"U": 121, "L": 122}
OPS = {"N": 111, "T": 112}
DIAG = {"N": 131, "U": 132}


@pytest.mark.parametrize("T", ["double", "float"])
def test_workspace_queries_need_nothing(h, T):
    for name in ("rocblas_internal_trsm_workspace_size", "rocblas_internal_trsm_batched_workspace_size"):
        f = fn(f"rocblas_status_ {name}<{T}>", WS_ARGS)
        sizes = [ctypes.c_size_t(7) for _ in range(5)]
        st = f(141, 111, 1000, 300, 4, 0, *[ctypes.byref(x) for x in sizes])
        assert st == S.continue_ and all(x.value == 0 for x in sizes)
        assert f(141, 111, 10, 10, 1, 0, None, *[ctypes.byref(x) for x in sizes[1:]]) \
            == S.invalid_pointer


@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("R", "U", "T"), ("L", "U", "T"),
                                             ("R", "L", "N")])
def test_trsm_template_offsets_and_strides(h, side, uplo, trans):
    """Offsets and strides of the template address the same instances as the
    strided-batched entry point on shifted pointers: bitwise equal results, for
    a small, a panel/chained and a recursive size."""
    rng = np.random.default_rng(11)
    f = fn("rocblas_status_ rocblas_internal_trsm_template<double>", TRSM_ARGS)
    fb = fn("rocblas_status_ rocblas_internal_trsm_batched_template<double>", TRSM_ARGS)
    for (m, n, batch) in [(40, 24, 3), (200, 70, 2), (700, 40, 1)]:
        say("internal trsm", side, uplo, trans, m, n, batch)
        k = m if side == "L" else n
        lda, ldb = k + 3, m + 1
        offA, offB = 5, 9
        sA, sB = lda * k + 11, ldb * n + 4
        A = np.full(offA + sA * batch, np.nan)
        for b in range(batch):
            Ab = oracle.diag_dominant_triangular(k, uplo, rng, lda=lda)
            A[offA + b * sA:offA + b * sA + lda * k] = Ab.ravel(order="F")
        B = rng.uniform(-1, 1, offB + sB * batch)
        dA = torch.from_numpy(A).cuda()
        want = torch.from_numpy(B.copy()).cuda()
        got = torch.from_numpy(B.copy()).cuda()
        assert rb.trsm_strided_batched(h, side, uplo, trans, "N", m, n, 0.5,
                                       dA.data_ptr() + 8 * offA, lda, sA,
                                       want.data_ptr() + 8 * offB, ldb, sB, batch) == S.success
        alpha = ctypes.c_double(0.5)
        assert f(h.ptr, SIDE[side], FILL[uplo], OPS[trans], DIAG["N"], m, n,
                 ctypes.addressof(alpha), dA.data_ptr(), offA, lda, sA, got.data_ptr(), offB, ldb,
                 sB, batch, True, None, None, None, None, None, 0, 0, 0) == S.success
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        # pointer arrays with the same offsets
        got2 = torch.from_numpy(B.copy()).cuda()
        pa = ptr_array([dA[b * sA:] for b in range(batch)])
        pb = ptr_array([got2[b * sB:] for b in range(batch)])
        assert fb(h.ptr, SIDE[side], FILL[uplo], OPS[trans], DIAG["N"], m, n,
                  ctypes.addressof(alpha), pa.data_ptr(), offA, lda, 0, pb.data_ptr(), offB, ldb, 0,
                  batch, True, None, None, None, None, None, 0, 0, 0) == S.success
        torch.cuda.synchronize()
        assert torch.equal(got2, want)


def test_trsm_template_device_alpha_and_quick_returns(h):
    rng = np.random.default_rng(12)
    f = fn("rocblas_status_ rocblas_internal_trsm_template<double>", TRSM_ARGS)
    m, n = 150, 30
    A = oracle.diag_dominant_triangular(m, "L", rng)
    B = rng.uniform(-1, 1, (m, n))
    dA, dB = dev(A), dev(B)
    ref = B.copy(order="F")
    oracle.trsm("L", "L", "N", "N", m, n, 2.0, np.asfortranarray(A), ref)
    dalpha = torch.tensor([2.0], dtype=torch.float64, device="cuda")
    rb.check(h.set_pointer_mode(1))
    try:
        assert f(h.ptr, 141, 122, 111, 131, m, n, dalpha.data_ptr(), dA.data_ptr(), 0, m, 0,
                 dB.data_ptr(), 0, m, 0, 1, True, None, None, None, None, None, 0, 0, 0) \
            == S.success
    finally:
        rb.check(h.set_pointer_mode(0))
    torch.cuda.synchronize()
    got = dB.cpu().numpy().reshape(n, m).T
    assert oracle.norm1_rel_diff(m, n, np.asfortranarray(got), ref) <= 40 * 2.2e-16 * m
    # zero batch / size: success, B untouched (no argument checks beyond the handle)
    before = dB.clone()
    one = ctypes.c_double(1.0)
    for (mm, nn, bb) in [(0, n, 1), (m, 0, 1), (m, n, 0)]:
        assert f(h.ptr, 141, 122, 111, 131, mm, nn, ctypes.addressof(one), dA.data_ptr(), 0, m, 0,
                 dB.data_ptr(), 0, m, 0, bb, True, None, None, None, None, None, 0, 0, 0) \
            == S.success
    torch.cuda.synchronize()
    assert torch.equal(dB, before)


def test_trtri_template_sub_batches_and_in_place(h):
    """batch_count instances of sub_batch_count triangles: each inverse equals
    the public rocblas_dtrtri's bit for bit; in place through w_C_tmp as well."""
    rng = np.random.default_rng(13)
    f = fn("rocblas_status_ rocblas_internal_trtri_template<double>", TRTRI_ARGS)
    fb = fn("rocblas_status_ rocblas_internal_trtri_batched_template<double>", TRTRI_ARGS)
    temp = getattr(rb.lib(), mangled("rocblas_internal_trtri_temp_elements"))
    temp.restype, temp.argtypes = ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]
    for uplo, diag in [("L", "N"), ("U", "U")]:
        n, lda, batch, sub = 90, 93, 2, 3
        say("internal trtri", uplo, diag, n, batch, sub)
        sub_stride = lda * n + 7
        stride = sub_stride * sub + 13
        off = 3
        A = np.full(off + stride * batch, 0.0)
        mats = []
        for b in range(batch):
            for s in range(sub):
                M = oracle.diag_dominant_triangular(n, uplo, rng, lda=lda, poison=0.0)
                o = off + b * stride + s * sub_stride
                A[o:o + lda * n] = M.ravel(order="F")
                mats.append((o, M))
        dA = torch.from_numpy(A).cuda()
        dI = torch.zeros_like(dA)
        assert f(h.ptr, FILL[uplo], DIAG[diag], n, dA.data_ptr(), off, lda, stride, sub_stride,
                 dI.data_ptr(), off, lda, stride, sub_stride, batch, sub, None) == S.success
        # in place: A itself, with the scratch the temp-elements query asks for
        dP = dA.clone()
        scratch = torch.empty(temp(n, batch), dtype=torch.float64, device="cuda")
        assert f(h.ptr, FILL[uplo], DIAG[diag], n, dP.data_ptr(), off, lda, stride, sub_stride,
                 dP.data_ptr(), off, lda, stride, sub_stride, batch, sub, scratch.data_ptr()) \
            == S.success
        # pointer-array form: one pointer per instance, sub-batches by sub_stride
        dI2 = torch.zeros_like(dA)
        pa = ptr_array([dA[b * stride:] for b in range(batch)])
        pi = ptr_array([dI2[b * stride:] for b in range(batch)])
        assert fb(h.ptr, FILL[uplo], DIAG[diag], n, pa.data_ptr(), off, lda, 0, sub_stride,
                  pi.data_ptr(), off, lda, 0, sub_stride, batch, sub, None) == S.success
        torch.cuda.synchronize()
        for (o, M) in mats:
            one = torch.zeros(lda * n, dtype=torch.float64, device="cuda")
            assert rb.trtri(h, uplo, diag, n, dev(M).data_ptr(), lda, one.data_ptr(), lda) \
                == S.success
            torch.cuda.synchronize()
            assert torch.equal(dI[o:o + lda * n], one), o
            assert torch.equal(dI2[o:o + lda * n], one), o
            # in place: the triangle of invA (the opposite one keeps A's values)
            got = dP[o:o + lda * n].cpu().numpy().reshape(n, lda).T[:n]
            inv = one.cpu().numpy().reshape(n, lda).T[:n]
            tri = np.tril if uplo == "L" else np.triu
            assert np.array_equal(tri(got), tri(inv)), o


TRSM64_ARGS = [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _vp, _vp, _i64, _i64, _i64, _vp, _i64,
               _i64, _i64, _i64, ctypes.c_bool, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64]


def test_trsm_template_64_matches_32_bit_form(h):
    """The ILP64 templates (library/src/src64/blas3/rocblas_trsm_kernels_64.cpp:31-225)
    give the 32-bit templates' results bit for bit on sizes both accept; a
    triangle beyond INT32_MAX is invalid_size and a leading dimension beyond it
    not_implemented, both before any launch."""
    rng = np.random.default_rng(14)
    f32 = fn("rocblas_status_ rocblas_internal_trsm_template<double>", TRSM_ARGS)
    f64 = fn("rocblas_status_ rocblas_internal_trsm_template_64<double>", TRSM64_ARGS)
    fb64 = fn("rocblas_status_ rocblas_internal_trsm_batched_template_64<double>", TRSM64_ARGS)
    alpha = ctypes.c_double(-1.5)
    for side, uplo, trans, m, n, batch in [("L", "U", "N", 300, 20, 2), ("R", "L", "T", 33, 90, 3)]:
        say("internal trsm_64", side, uplo, trans, m, n, batch)
        k = m if side == "L" else n
        lda, ldb = k, m
        sA, sB = lda * k, ldb * n
        A = np.concatenate([oracle.diag_dominant_triangular(k, uplo, rng).ravel(order="F")
                            for _ in range(batch)])
        B = rng.uniform(-1, 1, sB * batch)
        dA = torch.from_numpy(A).cuda()
        want = torch.from_numpy(B.copy()).cuda()
        got = torch.from_numpy(B.copy()).cuda()
        got2 = torch.from_numpy(B.copy()).cuda()
        args = (SIDE[side], FILL[uplo], OPS[trans], DIAG["N"], m, n, ctypes.addressof(alpha))
        assert f32(h.ptr, *args, dA.data_ptr(), 0, lda, sA, want.data_ptr(), 0, ldb, sB, batch,
                   True, None, None, None, None, None, 0, 0, 0) == S.success
        assert f64(h.ptr, *args, dA.data_ptr(), 0, lda, sA, got.data_ptr(), 0, ldb, sB, batch,
                   True, None, None, None, None, None, 0, 0, 0) == S.success
        pa = ptr_array([dA[b * sA:] for b in range(batch)])
        pb = ptr_array([got2[b * sB:] for b in range(batch)])
        assert fb64(h.ptr, *args, pa.data_ptr(), 0, lda, 0, pb.data_ptr(), 0, ldb, 0, batch,
                    True, None, None, None, None, None, 0, 0, 0) == S.success
        torch.cuda.synchronize()
        assert torch.equal(got, want) and torch.equal(got2, want)
    big = 1 << 31
    one = ctypes.c_double(1.0)
    common = (ctypes.addressof(one), dA.data_ptr(), 0)
    assert f64(h.ptr, 141, 122, 111, 131, big, 4, *common, big, 0, got.data_ptr(), 0, big, 0, 1,
               True, None, None, None, None, None, 0, 0, 0) == S.invalid_size
    assert f64(h.ptr, 142, 122, 111, 131, 4, big, *common, big, 0, got.data_ptr(), 0, 4, 0, 1,
               True, None, None, None, None, None, 0, 0, 0) == S.invalid_size
    assert f64(h.ptr, 141, 122, 111, 131, 8, 4, *common, big, 0, got.data_ptr(), 0, 8, 0, 1,
               True, None, None, None, None, None, 0, 0, 0) == S.not_implemented
    assert f64(h.ptr, 141, 122, 111, 131, 0, 4, *common, 8, 0, got.data_ptr(), 0, 8, 0, 1,
               True, None, None, None, None, None, 0, 0, 0) == S.success
