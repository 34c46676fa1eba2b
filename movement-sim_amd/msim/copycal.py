"""Copy of a known number of bytes on the device, for scaling rocprofv3's
FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) to bytes:

    rocprofv3 --pmc FETCH_SIZE -- python3 -m msim.copycal 268435456

reads and writes exactly N bytes once (torch's device-to-device copy)."""
import sys

import torch


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 28
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {n} bytes x3", flush=True)


if __name__ == "__main__":
    main()
