"""Known-size device copy used to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE.

Copies ``--mib`` MiB between two device buffers with hipMemcpyDtoD-free
elementwise torch copies (one kernel per repetition) so that the per-dispatch
counter readings of a kernel whose HBM traffic is exactly known (read N bytes,
write N bytes) can scale the readings of the DMO kernels.
"""
from __future__ import annotations

import argparse


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    import torch
    n = a.mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda:0", dtype=torch.float32)
    dst = torch.empty_like(src)
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB x {a.reps}", flush=True)


if __name__ == "__main__":
    main()
