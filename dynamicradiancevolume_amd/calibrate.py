"""A plain device copy of known size, for calibrating hardware byte counters.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m dynamicradiancevolume_amd.calibrate

copies 256 MiB three times (torch's copy kernel); the per-dispatch FETCH_SIZE
and WRITE_SIZE that a profiler reports for it, divided into 256 MiB, give the
factor that converts those counters of the library's kernels into HBM bytes.
"""
import sys

import torch

N_BYTES = 256 * 1024 * 1024


def main() -> int:
    a = torch.ones(N_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {N_BYTES} bytes x3", file=sys.stderr)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
