"""Device-copy calibration for HBM traffic counters.

Runs a plain device-to-device copy of a known size (default 1 GiB, read
1 GiB + write 1 GiB) a few times.  Under ``rocprofv3 --pmc FETCH_SIZE`` /
``--pmc WRITE_SIZE`` the per-dispatch counter values of this copy, divided
by the known byte counts, give the factor that converts the counters of the
KMM kernels into bytes.

    python -m polymer_dns_amd.calibrate [--mib 1024] [--reps 3]
"""
from __future__ import annotations

import argparse
import json
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    import torch
    n = a.mib * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ts = []
    for _ in range(a.reps):
        t0 = time.perf_counter()
        dst.copy_(src)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    b = n * 8
    print(json.dumps({"bytes_read": b, "bytes_written": b, "best_s": min(ts),
                      "copy_GBs": 2 * b / min(ts) / 1e9}))


if __name__ == "__main__":
    main()
