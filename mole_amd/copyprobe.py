"""Runs the plain device copy used to calibrate HBM byte counters.

    python mole_amd/copyprobe.py [BYTES] [REPEATS]

copies BYTES (default 256 MiB) from one device buffer to another REPEATS
times (default 2) and prints the bytes moved per copy and the bandwidth.
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import mole_amd  # noqa: E402


def main():
    nbytes = int(sys.argv[1]) if len(sys.argv) > 1 else 256 << 20
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    n = nbytes // 8
    ms = mole_amd.debug_copy(n, reps)
    print(f"copy_probe_kernel: read {n * 8} B + write {n * 8} B per copy, {ms:.4f} ms, "
          f"{2 * n * 8 / ms / 1e6:.1f} GB/s")


if __name__ == "__main__":
    main()
