"""Device-side synthetic inputs (rocblas_amd_{s,d}synth_tri / _rhs) against the
host generator rocblas_amd/synth.py: bit-identical for the uniform
distributions, within a few ulps for the Box-Muller normals."""
import numpy as np
import pytest
import torch

import rocblas_amd as rb
from rocblas_amd import synth
from tests.gpu_util import GpuHandle, host, say

pytestmark = pytest.mark.gpu
S = rb.rocblas_status


@pytest.fixture(scope="module")
def h():
    assert torch.cuda.is_available()
    hd = GpuHandle()
    yield hd
    hd.close()


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_uniform_triangle_bitwise(h, uplo):
    for k, lda, seed, scale in [(1, 1, 0, 1.0), (37, 40, 5, 1.0), (300, 301, 9, 0.25)]:
        say("synth tri", uplo, k, lda)
        d = torch.empty(lda * k, dtype=torch.float64, device="cuda")
        assert rb.synth_tri(h, uplo, "uniform", k, seed, d.data_ptr(), lda, scale=scale,
                            poison=-7.0) == S.success
        torch.cuda.synchronize()
        want = synth.tri_uniform(k, uplo, seed, ld=lda, scale=scale, poison=-7.0)
        assert np.array_equal(host(d, (lda, k)), want)
        f = torch.empty(lda * k, dtype=torch.float32, device="cuda")
        assert rb.synth_tri(h, uplo, "uniform", k, seed, f.data_ptr(), lda, scale=scale,
                            poison=-7.0, precision="s") == S.success
        torch.cuda.synchronize()
        assert np.array_equal(host(f, (lda, k)), want.astype(np.float32))


def test_rhs_bitwise_and_normal_close(h):
    m, n, ldb = 129, 33, 131
    d = torch.full((ldb * n,), 5.0, dtype=torch.float64, device="cuda")
    assert rb.synth_rhs(h, m, n, 3, d.data_ptr(), ldb) == S.success
    torch.cuda.synchronize()
    assert np.array_equal(host(d, (ldb, n)), synth.rhs_uniform(m, n, 3, ld=ldb))
    k = 200
    a = torch.empty(k * k, dtype=torch.float64, device="cuda")
    assert rb.synth_tri(h, "U", "normal", k, 11, a.data_ptr(), k) == S.success
    torch.cuda.synchronize()
    want = synth.tri_normal(k, "U", 11)
    got = host(a, (k, k))
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15)
    assert np.array_equal(np.diag(got), np.diag(want))


def test_bad_args(h):
    L = rb.lib()
    d = torch.empty(16, dtype=torch.float64, device="cuda")
    assert L.rocblas_amd_dsynth_tri(None, 121, 0, 4, 0, 1.0, 0.0, d.data_ptr(), 4) == S.invalid_handle
    assert L.rocblas_amd_dsynth_tri(h.ptr, 120, 0, 4, 0, 1.0, 0.0, d.data_ptr(), 4) == S.invalid_value
    assert L.rocblas_amd_dsynth_tri(h.ptr, 121, 2, 4, 0, 1.0, 0.0, d.data_ptr(), 4) == S.invalid_value
    assert L.rocblas_amd_dsynth_tri(h.ptr, 121, 0, 4, 0, 1.0, 0.0, d.data_ptr(), 3) == S.invalid_size
    assert L.rocblas_amd_dsynth_tri(h.ptr, 121, 0, 4, 0, 1.0, 0.0, None, 4) == S.invalid_pointer
    assert L.rocblas_amd_dsynth_rhs(h.ptr, 4, 4, 0, 2, d.data_ptr(), 3) == S.invalid_size
    assert L.rocblas_amd_dsynth_rhs(h.ptr, 0, 4, 0, 2, None, 1) == S.success
