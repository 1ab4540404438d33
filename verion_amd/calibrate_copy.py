"""A device-to-device copy of a known size, for calibrating rocprofv3's FETCH_SIZE / WRITE_SIZE
counters: the copy kernel reads and writes exactly `--mib` MiB, so the counter value it reports
divided by that size is the factor that turns the counters of other kernels into bytes.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace -- python -m verion_amd.calibrate_copy --mib 1024
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB three times", flush=True)


if __name__ == "__main__":
    main()
