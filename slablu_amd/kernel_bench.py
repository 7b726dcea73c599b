"""Single-kernel runs for hardware-counter collection.

    python -m slablu_amd.kernel_bench [--leaves 2500] [--p 22] [--reps 5]

Runs the leaf interior solve (slab_getrs_perm, one wavefront per leaf) on
leaf factors of the benchmark's shape (2500 leaves, (p-2)^2 = 400 interior
unknowns, [A_cc | A_cx] rows of 480 entries, FP32) and a plain device copy
of a known size (256 MiB read + 256 MiB written) whose FETCH_SIZE /
WRITE_SIZE readings calibrate the counters to bytes.  Under
`rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` this gives the HBM
traffic per leaf-solve launch that bench.py reports next to the kernel's
algorithmic bytes.
"""
import argparse
import json
import sys

import torch

from . import hip

COPY_BYTES = 256 * 2 ** 20


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--leaves", type=int, default=2500)
    ap.add_argument("--p", type=int, default=22)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    ni, nx = (args.p - 2) ** 2, 4 * (args.p - 2)
    g = torch.Generator(device="cuda").manual_seed(0)
    M = torch.randn(args.leaves, ni, ni + nx, device="cuda", generator=g)
    M[:, :, :ni] += 30 * torch.eye(ni, device="cuda")
    ipiv = torch.empty(args.leaves, ni, dtype=torch.int32, device="cuda")
    hip.getrf(M, ipiv, nelim=ni, npiv=ni)
    perm = hip.piv_to_perm(ipiv)
    x = torch.randn(args.leaves, ni, 1, device="cuda", generator=g)
    src = torch.randn(COPY_BYTES // 4, device="cuda", generator=g)
    dst = torch.empty_like(src)
    ms = []
    for _ in range(args.reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        hip.getrs(M[:, :, :ni], ipiv, x, perm=perm)
        e1.record()
        dst.copy_(src)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    print(json.dumps({"kernel": "k_getrs_wave<float,1>", "leaves": args.leaves, "ni": ni,
                      "avg_ms": sum(ms) / len(ms), "copy_bytes": COPY_BYTES}))
    sys.stdout.flush()


if __name__ == "__main__":
    main()
