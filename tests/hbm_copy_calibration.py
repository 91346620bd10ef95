"""Plain device copy of a known size, for calibrating the HBM byte counters.

    rocprofv3 --pmc FETCH_SIZE -- python3 tests/hbm_copy_calibration.py

The copy kernel reads and writes exactly 512 MiB per launch; the ratio of
the counter's reading to that figure is the scale applied to the same
counter on the stage kernels (profiles/).
"""
import torch

N = 64 * 1024 * 1024  # doubles


def main():
    a = torch.ones(N, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy bytes read", N * 8, "written", N * 8, flush=True)


if __name__ == "__main__":
    main()
