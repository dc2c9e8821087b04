"""Plain device copy of a known size, the calibration run for HBM byte counters.

python -m chemps2_amd.copy_probe [GiB]: copies a float64 buffer of the given size
(default 2 GiB) three times with torch and prints the bytes moved per copy.
"""
import sys

import torch


def main():
    gib = float(sys.argv[1]) if len(sys.argv) > 1 else 2.0
    n = int(gib * (1 << 30)) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy_probe: {n * 8} bytes read and {n * 8} bytes written per copy, 3 copies", flush=True)


if __name__ == "__main__":
    main()
