"""One device copy of a known size, for calibrating hardware traffic counters:
    python -m saruman_amd.calib [bytes]
prints the copy's size and kernel time."""
import ctypes as C
import sys

from . import _lib


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 256 << 20
    lib = _lib.load()
    ms = C.c_float()
    _lib.check(lib.srm_calibrate_copy(0, n, C.byref(ms)), "srm_calibrate_copy")
    print(f"copy_kernel bytes_read={n} bytes_written={n} ms={ms.value:.4f} "
          f"GB/s={2 * n / ms.value / 1e6:.1f}", flush=True)


if __name__ == "__main__":
    main()
