"""A plain device-to-device copy of known size, for calibrating HBM counters.

``python -m glasscamerapublic_amd.copy_probe [MiB] [repeats]`` copies a
buffer of MiB mebibytes to another `repeats` times (torch's copy kernel on
the current stream).  Under ``rocprofv3 --pmc FETCH_SIZE`` (or
``WRITE_SIZE``) each dispatch should read (write) MiB mebibytes; the ratio
of what the counter reports to that size is the scale applied to the same
counter on the denoise kernels.
"""
from __future__ import annotations

import sys

import torch


def main(argv) -> int:
    mib = int(argv[1]) if len(argv) > 1 else 256
    reps = int(argv[2]) if len(argv) > 2 else 5
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
