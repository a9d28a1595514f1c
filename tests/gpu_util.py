"""Helpers for GPU tests: numpy column-major <-> device buffers, one-call wrappers."""
import ctypes
import sys

import numpy as np
import torch

import rocblas_amd as rb


def say(*args):
    """Print a progress line that names the case before a GPU step (flushed)."""
    print(*args, file=sys.stderr, flush=True)


def dev(arr):
    """Fortran-ordered numpy array (ld, cols) -> 1-D device tensor with the same layout."""
    a = np.asarray(arr)
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda()


def host(t, shape):
    """1-D device tensor -> Fortran numpy array of ``shape`` (ld, cols)."""
    ld, cols = shape
    return np.asfortranarray(t.detach().cpu().numpy().reshape(cols, ld).T)


class GpuHandle(rb.Handle):
    """Handle bound to torch's current stream."""

    def __init__(self):
        super().__init__()
        rb.check(self.set_stream(torch.cuda.current_stream().cuda_stream), "set_stream")


def gpu_trsm(h, side, uplo, trans, diag, m, n, alpha, A, B, precision="d"):
    """Solve on the GPU through rocblas_{s,d}trsm; returns (status, X as Fortran array)."""
    dA = dev(A)
    dB = dev(B)
    st = rb.trsm(h, side, uplo, trans, diag, m, n, alpha, dA.data_ptr(), A.shape[0],
                 dB.data_ptr(), B.shape[0], precision)
    torch.cuda.synchronize()
    return st, host(dB, B.shape)


def eta_torch(side, uplo, trans, diag, alpha, A, B0, X):
    """Normwise backward error with a plain PyTorch float64 reference on the device:
    ||alpha B0 - op(A) X||_1 / (||op(A)||_1 ||X||_1 + |alpha| ||B0||_1).  Arguments
    are square-cropped float64 device tensors in row-major "transposed" form is
    avoided: A, B0, X are (rows, cols) logical matrices (torch layout)."""
    k = A.shape[0]
    T = torch.triu(A) if uplo == "U" else torch.tril(A)
    T = torch.nan_to_num(T, nan=0.0)
    if diag == "U":
        T = T - torch.diag(torch.diag(T)) + torch.eye(k, dtype=T.dtype, device=T.device)
    op = T.T.contiguous() if trans != "N" else T
    R = alpha * B0 - (op @ X if side == "L" else X @ op)
    n1 = lambda M: M.abs().sum(0).max()
    return float(n1(R) / (n1(op) * n1(X) + abs(alpha) * n1(B0)))


def logical(t, ld, rows, cols):
    """1-D column-major device buffer -> (rows, cols) logical torch matrix view copy."""
    return t.view(cols, ld)[:, :rows].T.contiguous()


def ptr_array(tensors):
    """Device array of device pointers (for the _batched forms)."""
    p = torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64)
    return p.cuda()


def c_double_p(x):
    return ctypes.addressof(x)
