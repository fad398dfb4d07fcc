"""Calibration program for HBM traffic counters.

`python -m changa_amd.roofline --copy-mib 1024` copies a buffer of known size
a few times. Run under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE), the
per-dispatch counter of this copy, divided by the bytes the copy must move,
is the factor that turns the same counters of the gravity kernels into bytes.
"""
import argparse
import json


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=4)
    args = ap.parse_args()
    import torch

    n = args.copy_mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "reps": args.reps}))


if __name__ == "__main__":
    main()
