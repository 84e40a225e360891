"""The C-ABI library: loads without a GPU, exports every entry point that
include/drv.h declares, struct layouts agree with the C compiler's, and the
host-only entry points match the oracle."""
import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dynamicradiancevolume_amd import _abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "drv.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(drv_\w+)\s*\(", text, flags=re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _abi.load_drv()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in _abi.DRV_FUNCTIONS, f"{n} not bound in _abi"
    assert set(_abi.DRV_FUNCTIONS) == set(names)


def test_struct_layouts_match_c(tmp_path):
    structs = ["drv_config", "drv_cascade", "drv_volume", "drv_frame", "drv_sh_constants", "drv_spot_params",
               "drv_spot_light", "drv_gbuffer", "drv_rsm", "drv_stage_times"]
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "drv.h"\nint main(void){\n' +
                   "".join(f'printf("%zu\\n", sizeof({s}));\n' for s in structs) + "return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    py = [_abi.DrvConfig, _abi.DrvCascade, _abi.DrvVolume, _abi.DrvFrame, _abi.DrvShConstants, _abi.DrvSpotParams,
          _abi.DrvSpotLight, _abi.DrvGBuffer, _abi.DrvRsm, _abi.DrvStageTimes]
    assert [C.sizeof(t) for t in py] == sizes


def test_invalid_arguments_are_rejected():
    lib = _abi.load_drv()
    ctx = C.c_void_p()
    assert lib.drv_create(None, C.byref(ctx)) == _abi.DRV_ERR_INVALID_ARG
    bad = _abi.DrvConfig(width=0, height=10, max_caches=1, cav_resolution=8, cav_cascades=1, voxel_resolution=8,
                         sh_bands=1, indirect_shadow=1, cav_transition_size=0, cav_cascade0_world_size=4)
    assert lib.drv_create(C.byref(bad), C.byref(ctx)) == _abi.DRV_ERR_INVALID_ARG
    bad.width = 10
    bad.voxel_resolution = 12  # not a power of two
    assert lib.drv_create(C.byref(bad), C.byref(ctx)) == _abi.DRV_ERR_INVALID_ARG
    bad.voxel_resolution = 8
    bad.cav_cascades = 5
    assert lib.drv_create(C.byref(bad), C.byref(ctx)) == _abi.DRV_ERR_INVALID_ARG
    bad.cav_cascades = 1
    bad.voxel_resolution = 1024  # the corner-expanded copy needs 32-bit byte offsets
    assert lib.drv_create(C.byref(bad), C.byref(ctx)) == _abi.DRV_ERR_INVALID_ARG
    bad.voxel_resolution = 8
    bad.indirect_specular = 1
    for size in (3, 64, -2):  # env-map side: a power of two up to 32 (0 = the default 16)
        bad.specular_envmap_size = size
        assert lib.drv_create(C.byref(bad), C.byref(ctx)) == _abi.DRV_ERR_INVALID_ARG
    bad.specular_envmap_size = 16
    bad.specular_hole_fill_level = -1
    assert lib.drv_create(C.byref(bad), C.byref(ctx)) == _abi.DRV_ERR_INVALID_ARG
    assert lib.drv_destroy(None) == _abi.DRV_OK
    assert lib.drv_set_voxel_overlap(None, 1) == _abi.DRV_ERR_INVALID_ARG
    # the atlas entry points refuse a null context
    assert lib.drv_get_specular_atlas(None, None, None, None) == _abi.DRV_ERR_INVALID_ARG
    assert lib.drv_read_specular_atlas(None, 0, None) == _abi.DRV_ERR_INVALID_ARG


def test_prepare_spot_light_matches_oracle_and_is_consistent():
    orc = pytest.importorskip("oracle.orc")
    lib = _abi.load_drv()
    for res, read_lod, lod, half in [(64, 0, 2, 0.45), (256, 1, 2, 0.35), (1000, 4, 1, 0.5)]:
        p = _abi.DrvSpotParams(intensity=_abi.f3(1, 2, 3), position=_abi.f3(0.5, 3.0, -1.0),
                               direction=_abi.f3(0.2, -1.0, 0.1), half_angle=half, rsm_resolution=res,
                               rsm_read_lod=read_lod, indirect_shadow_lod=lod, near_plane=0.1, far_plane=10000.0)
        got = _abi.DrvSpotLight()
        assert lib.drv_prepare_spot_light(C.byref(p), C.byref(got)) == _abi.DRV_OK
        ref = _abi.DrvSpotLight()
        orc.lib().orc_prepare_spot_scalars(C.byref(p), C.byref(ref))
        for f in ("rsm_render_resolution", "rsm_read_resolution", "val_area_factor", "indirect_shadow_lod",
                  "indirect_shadow_block_size", "indirect_shadow_sample_interval", "indirect_shadow_super_val_width",
                  "indirect_shadow_sampling_offset", "light_cos_half_angle"):
            assert getattr(got, f) == getattr(ref, f), f
        assert got.rsm_render_resolution == 1 << math.ceil(math.log2(res))
        # closed forms (renderer.cpp:694-719)
        rr = got.rsm_read_resolution
        assert got.val_area_factor == pytest.approx((2 * math.sin(half) / rr) ** 2, rel=1e-5)
        assert got.indirect_shadow_sampling_offset == pytest.approx(0.5 + math.sqrt(2) * (1 << lod) / 2)
        vp = np.array(got.light_view_projection[:]).reshape(4, 4)
        ivp = np.array(got.inverse_light_view_projection[:]).reshape(4, 4)
        np.testing.assert_allclose(vp @ ivp, np.eye(4), atol=2e-3)
        # the light direction projects to the RSM centre; a point one unit
        # ahead lies between the near (depth 1) and far (depth 0) planes
        ahead = np.array([0.5, 3.0, -1.0]) + np.array(got.light_direction[:])
        c = vp @ np.append(ahead, 1.0)
        assert abs(c[0] / c[3]) < 1e-4 and abs(c[1] / c[3]) < 1e-4 and 0 < c[2] / c[3] < 1
        # the cone edge maps to NDC +-tan(half)/tan(half) = 1 along the frustum
    # the 2:1 extension: half the rows, the texel solid angle over W x H texels
    for res, lod in ((512, 1), (64, 0)):
        sq = _abi.DrvSpotParams(intensity=_abi.f3(1, 1, 1), position=_abi.f3(0, 3, 0), direction=_abi.f3(0, -1, 0),
                                half_angle=0.4, rsm_resolution=res, rsm_read_lod=lod, indirect_shadow_lod=2,
                                near_plane=0.1, far_plane=100.0)
        wd = _abi.DrvSpotParams.from_buffer_copy(sq)
        wd.rsm_wide = 1
        a, b = _abi.DrvSpotLight(), _abi.DrvSpotLight()
        assert lib.drv_prepare_spot_light(C.byref(sq), C.byref(a)) == _abi.DRV_OK
        assert lib.drv_prepare_spot_light(C.byref(wd), C.byref(b)) == _abi.DRV_OK
        assert (a.rsm_render_height, a.rsm_read_height) == (a.rsm_render_resolution, a.rsm_read_resolution)
        assert (b.rsm_render_height, b.rsm_read_height) == (res // 2, (res >> lod) // 2)
        assert b.val_area_factor == pytest.approx(2.0 * a.val_area_factor, rel=1e-6)
        assert list(b.light_view_projection) == list(a.light_view_projection)   # the same square frustum
        ref = _abi.DrvSpotLight()
        orc.lib().orc_prepare_spot_scalars(C.byref(wd), C.byref(ref))
        assert (ref.rsm_read_height, ref.val_area_factor) == (b.rsm_read_height, b.val_area_factor)
    bad = _abi.DrvSpotParams(rsm_resolution=64, rsm_read_lod=0, indirect_shadow_lod=7, half_angle=0.5,
                             direction=_abi.f3(0, -1, 0), near_plane=0.1, far_plane=100.0)
    assert lib.drv_prepare_spot_light(C.byref(bad), C.byref(_abi.DrvSpotLight())) == _abi.DRV_ERR_INVALID_ARG


def test_calibration_library_exports_its_probe():
    """libdrv_calib.so (the VALU instruction-rate probe behind `calibrate
    --valu`) loads and exports drv_calib_valu; an unknown op is refused
    before any device call."""
    from dynamicradiancevolume_amd import calibrate

    lib = C.CDLL(str(calibrate.CALIB))
    assert hasattr(lib, "drv_calib_valu")
    ms, n = C.c_float(), C.c_double()
    assert lib.drv_calib_valu(7, 10, 1, C.byref(ms), C.byref(n)) == -1
    assert lib.drv_calib_valu(0, 0, 1, C.byref(ms), C.byref(n)) == -1
