"""HBM rate this GPU reaches for several read/write mixes
(rocblas_amd_hbm_stream: best of 4/8/16 wavefronts per SIMD, plain and
non-temporal accesses).  The memory-bound solves quote their bandwidth against
these figures beside the published peak.

    python -m rocblas_amd.hbm_probe [--gib 4]
"""
import argparse
import sys

import torch

import rocblas_amd as rb

MIXES = [(1, 0), (0, 1), (1, 1), (2, 1), (5, 2), (3, 1)]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gib", type=int, default=4, help="bytes per buffer (GiB)")
    args = ap.parse_args(argv)
    assert torch.cuda.is_available(), "the probe runs on the GPU"
    h = rb.Handle()
    rb.check(h.set_stream(torch.cuda.current_stream().cuda_stream))
    step = rb.hbm_stream_step_bytes(h)
    cap = args.gib << 30
    src = torch.ones(cap // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty(cap // 4, dtype=torch.int32, device="cuda")
    for rq, wq in MIXES:
        units = cap // (step * max(rq, wq))
        if units < 1:
            continue
        rd, wr = rq * units * step, wq * units * step
        sec = rb.hbm_stream(h, src.data_ptr(), rd, dst.data_ptr(), wr, iters=5)
        print(f"read:write {rq}:{wq}  {(rd + wr) / 1e9:6.2f} GB per pass  {sec * 1e6:9.1f} us  "
              f"{(rd + wr) / sec / 1e12:.3f} TB/s", flush=True)
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
