"""Split arithmetic of the SH gather (csrc/split_plan.hpp), compiled for the
host: the partial-sum buffer always holds what the planner can pick.

The gather writes ceil(s / 4) x 27 x pitch floats of partials for s splits;
the device planner picks s as a power of two up to a cap.  The spot check of
the surrogate gathers a list of pitch 2 x ceil(capacity / k) into the same
buffer as the main list (27 x columns floats), so its cap must be the largest
power of two whose partials fit, whatever the main list's split count is."""
import ctypes as C
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "dynamicradiancevolume_amd" / "csrc"


@pytest.fixture(scope="module")
def sp(tmp_path_factory):
    d = tmp_path_factory.mktemp("split_plan")
    src = d / "sp.cpp"
    src.write_text('#include "split_plan.hpp"\nextern "C" {\n'
                   "uint32_t plan(uint32_t l, uint32_t n, uint32_t q, uint32_t m, uint32_t t)"
                   "{ return drvk::plan_split_count(l, n, q, m, t); }\n"
                   "uint64_t cols(uint32_t s, uint32_t p) { return drvk::gather_partial_columns(s, p); }\n"
                   "uint32_t fit(uint64_t c, uint32_t p, uint32_t m) { return drvk::max_fitting_splits(c, p, m); }\n"
                   "}\n")
    so = d / "libsp.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", str(CSRC), str(src), "-o", str(so)],
                   check=True)
    lib = C.CDLL(str(so))
    lib.plan.restype = C.c_uint32
    lib.plan.argtypes = [C.c_uint32] * 5
    lib.cols.restype = C.c_uint64
    lib.cols.argtypes = [C.c_uint32, C.c_uint32]
    lib.fit.restype = C.c_uint32
    lib.fit.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
    return lib


def test_planner_is_a_capped_power_of_two(sp):
    for live in (1, 63, 64, 6352, 262144):
        for n_vpl, quantum in ((4096, 64), (16384, 64), (65536, 256)):
            for cap in (1, 2, 4, 16, 256):
                for target in (1, 4096, 262144):
                    s = sp.plan(live, n_vpl, quantum, cap, target)
                    assert s >= 1 and s & (s - 1) == 0 and s <= cap
                    assert n_vpl // s >= quantum or s == 1
                    # it stops at the first count that reaches the target
                    if s > 1:
                        assert ((live + 63) // 64) * (s // 2) < target


def test_fitting_splits_fit_and_are_largest(sp):
    for columns in (0, 1, 100, 65536, 4 * 65536, 64 * 65536):
        for pitch in (1, 64, 1000, 6554, 131072):
            s = sp.fit(columns, pitch, 256)
            if s == 0:
                assert sp.cols(1, pitch) > columns
                continue
            assert s & (s - 1) == 0 and s <= 256
            assert sp.cols(s, pitch) <= columns
            if s < 256:
                assert sp.cols(2 * s, pitch) > columns


def test_spot_check_partials_stay_inside_the_buffer(sp):
    """The surrogate spot check at capacity 65536 (the 256x256 Cornell
    workload), 4096 VPLs, 1..4 main splits and a 5 % spot check: the old
    cap min(256, 4 groups x capacity / pitch2) overran the buffer; the
    cap from max_fitting_splits never does, for any live count."""
    cap = 65536
    for max_splits in (1, 2, 4, 256):
        for frac in (0.05, 0.01, 1.0, 1e-12):
            k = min(cap, max(1, round(1.0 / frac)))
            spot_cap = (cap + k - 1) // k
            pitch2 = 2 * spot_cap
            columns = max(sp.cols(max_splits, cap), pitch2)  # as drv_light_caches sizes the buffer
            old = min(256, 4 * ((max_splits + 3) // 4) * cap // pitch2)
            cap_splits = sp.fit(columns, pitch2, 256)
            assert cap_splits >= 1
            for live in (1, 64, spot_cap, 2 * spot_cap):
                s = sp.plan(live, 4096, 64, cap_splits, 4096 * 4)
                assert sp.cols(s, pitch2) <= columns
            if max_splits <= 4 and frac == 0.05:
                # the case the fix closes: the old cap (40 here), rounded up
                # to 64 by the planner, needed 1.6x the columns that exist
                s_old = sp.plan(2 * spot_cap, 4096, 64, 1 << (old - 1).bit_length(), 4096 * 4)
                assert sp.cols(s_old, pitch2) > columns
