"""Calibration runs for the kernel profiles.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m dynamicradiancevolume_amd.calibrate

copies 256 MiB three times (torch's copy kernel); the per-dispatch FETCH_SIZE
and WRITE_SIZE that a profiler reports for it, divided into 256 MiB, give the
factor that converts those counters of the library's kernels into HBM bytes.

    python3 -m dynamicradiancevolume_amd.calibrate --valu

times loops of independent FP32 vector instructions of the kinds the cone
trace and the SH gather use (``libdrv_calib.so``) and prints each kind's
time per wave64 instruction and SIMD, in ns, in cycles at the 2.4 GHz peak
clock, and relative to an FMA.
"""
import ctypes as C
import sys
from pathlib import Path

import torch

N_BYTES = 256 * 1024 * 1024
CALIB = Path(__file__).resolve().parent / "libdrv_calib.so"
PEAK_CLOCK_GHZ = 2.4
# (op, label, instructions of the measured kind per chain step, adds per step)
VALU_OPS = ((0, "v_fma_f32", 1, 0), (1, "v_add_f32", 1, 0), (2, "v_rsq_f32", 1, 0), (3, "v_rcp_f32", 1, 0),
            (4, "v_med3_f32", 1, 1), (5, "v_fract_f32", 1, 1), (6, "v_cvt_f32_u32 / v_cvt_u32_f32", 2, 1))


def copy() -> int:
    a = torch.ones(N_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {N_BYTES} bytes x3", file=sys.stderr)
    return 0


def valu(iters: int = 2000) -> dict:
    """ns per wave64 instruction per SIMD for each kind in VALU_OPS (a kind
    measured beside an add has the add's time taken off)."""
    if not CALIB.exists():
        raise RuntimeError(f"{CALIB} is missing: build first (python -m dynamicradiancevolume_amd.build)")
    lib = C.CDLL(str(CALIB))
    lib.drv_calib_valu.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_double)]
    lib.drv_calib_valu.restype = C.c_int
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    raw = {}
    for op, label, n, adds in VALU_OPS:
        ms, per_simd = C.c_float(), C.c_double()
        if lib.drv_calib_valu(op, iters, cus, C.byref(ms), C.byref(per_simd)) != 0:
            raise RuntimeError(f"drv_calib_valu({op}) failed")
        raw[op] = (ms.value * 1e6 / per_simd.value, label, n, adds)  # ns per chain step per SIMD
    add_ns = raw[1][0]
    out = {}
    for op, (ns, label, n, adds) in raw.items():
        out[label] = (ns - adds * add_ns) / n
    return out


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if "--valu" not in argv:
        return copy()
    rates = valu()
    fma = rates["v_fma_f32"]
    print(f"{'instruction':34s} {'ns/wave64 inst/SIMD':>20s} {'cycles at 2.4 GHz':>18s} {'x FMA':>6s}")
    for label, ns in rates.items():
        print(f"{label:34s} {ns:20.3f} {ns * PEAK_CLOCK_GHZ:18.2f} {ns / fma:6.2f}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
