"""Calibration for HBM traffic counters: one device-to-device copy of a known
size (the HIP runtime's copy kernel), so that what rocprofv3's FETCH_SIZE
and WRITE_SIZE report for it can be compared with the bytes actually moved.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m frehdc_amd.copy_probe --mib 1024
"""
from __future__ import annotations

import argparse
import ctypes as C


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--repeat", type=int, default=3)
    a = ap.parse_args()
    hip = C.CDLL("libamdhip64.so")
    n = a.mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
    for _ in range(a.repeat):
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0    # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"copied {a.mib} MiB x {a.repeat} device to device", flush=True)


if __name__ == "__main__":
    main()
