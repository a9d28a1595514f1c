"""Seeded variable-size batch workloads built in device memory.

``variable_batch(batch, mlo, mhi, nlo, nhi, seed)``: instance i has m log-uniform
in [mlo, mhi] and n uniform in {nlo..nhi} (SplitMix64 stream TAG_SIZES of
``seed``), a lower-triangular A with diagonal U[1,2] and off-diagonal
U[-1,1]/m (seed + i) and B uniform in [-1,1] (seed + i), all generated on the
device by rocblas_amd_dsynth_tri / _rhs and packed back to back (lda = ldb = m).
Returns the device argument arrays of a side-left lower _vbatched call and the
pristine B to restore between timed runs.
"""
import numpy as np
import torch

import rocblas_amd as rb
from rocblas_amd import synth


def variable_sizes(batch, mlo, mhi, nlo, nhi, seed):
    key = synth.stream_key(seed, synth.TAG_SIZES)
    u = synth.uniform(key, 0, 2 * batch)
    ms = np.exp(np.log(mlo) + u[:batch] * (np.log(mhi + 1) - np.log(mlo))).astype(np.int64)
    ms = np.clip(ms, mlo, mhi)
    ns = np.clip((nlo + u[batch:] * (nhi - nlo + 1)).astype(np.int64), nlo, nhi)
    return ms, ns


def variable_batch(handle, batch=4096, mlo=8, mhi=512, nlo=1, nhi=64, seed=0):
    ms, ns = variable_sizes(batch, mlo, mhi, nlo, nhi, seed)
    offA = np.concatenate([[0], np.cumsum(ms * ms)])
    offB = np.concatenate([[0], np.cumsum(ms * ns)])
    dA = torch.empty(int(offA[-1]), dtype=torch.float64, device="cuda")
    dB0 = torch.empty(int(offB[-1]), dtype=torch.float64, device="cuda")
    for i in range(batch):
        m, n = int(ms[i]), int(ns[i])
        rb.check(rb.synth_tri(handle, "L", "uniform", m, seed + i,
                              dA.data_ptr() + 8 * int(offA[i]), m))
        rb.check(rb.synth_rhs(handle, m, n, seed + i, dB0.data_ptr() + 8 * int(offB[i]), m))
    dB = dB0.clone()
    i32 = lambda v: torch.from_numpy(np.asarray(v, dtype=np.int32)).cuda()
    dm, dn = i32(ms), i32(ns)
    pa = torch.tensor([dA.data_ptr() + 8 * int(o) for o in offA[:-1]], dtype=torch.int64).cuda()
    pb = torch.tensor([dB.data_ptr() + 8 * int(o) for o in offB[:-1]], dtype=torch.int64).cuda()
    return dict(ms=ms, ns=ns, offA=offA, offB=offB, dA=dA, dB0=dB0, dB=dB, m=dm, n=dn, lda=dm,
                ldb=dm, pa=pa, pb=pb, batch=batch, flops=float(np.sum(ms.astype(float) ** 2 * ns)))


def call_vbatched(handle, w):
    return rb.trsm_vbatched(handle, "L", "L", "N", "N", w["m"].data_ptr(), w["n"].data_ptr(), 1.0,
                            w["pa"].data_ptr(), w["lda"].data_ptr(), w["pb"].data_ptr(),
                            w["ldb"].data_ptr(), w["batch"])
