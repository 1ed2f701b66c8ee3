"""Plain device-to-device copy of a known size, for scaling the FETCH_SIZE / WRITE_SIZE counters.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m sfsim_amd.calibrate [MiB]

Each copy reads and writes exactly MiB mebibytes of HBM (the buffers are far larger than the caches),
so the counter per copy dispatch divided by that size is the counter's scale factor.
"""
import sys


def main(mib=1024, repeats=3):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {repeats}", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1024)
