"""Inputs rendered on the device (G-buffer fill, RSM fill and mip chain)
against the host producer in csrc/inputs (the same raster rules), and a
frame lit from device-rendered inputs against the same frame from host
inputs.

Depth, coverage, flux and depth/depth^2 are bit-identical (the same FP64
raster and FP32 shading in the same order).  Packed normals use atan2 (and,
in the mip chain, sin / cos of the unpacked angle), whose last-ulp results
differ between the host and device math libraries: a value may land on the
other side of an integer truncation, so a few normals may differ by one
unit; the same holds for sRGB albedo through pow.
"""
import numpy as np
import pytest
import torch

from canon import rel_l2
from conftest import announce, frame_inputs

pytestmark = pytest.mark.gpu


def _device(fi):
    from dynamicradiancevolume_amd.pipeline import DeviceFrame

    return DeviceFrame(fi)


def _renderer(fi):
    from dynamicradiancevolume_amd.pipeline import Renderer

    return Renderer(fi.workload.config())


def _off_by_one(a, b, max_frac, max_units=1):
    d = np.abs(a.astype(np.int64) - b.astype(np.int64))
    assert d.max() <= max_units, d.max()
    assert (d > 0).mean() <= max_frac, (d > 0).mean()


def _normals_close(a, b, max_frac, max_units):
    # packed (angle, z): the angle wraps at +-pi (32767 and -32768 are neighbours)
    d = a.astype(np.int64) - b.astype(np.int64)
    d[..., 0] = (d[..., 0] + 32768) % 65536 - 32768
    d = np.abs(d)
    assert d.max() <= max_units, d.max()
    assert (d > 0).sum() <= max(2, max_frac * d.size), (d > 0).sum()


@pytest.mark.parametrize("name", ["cornell_256", "boxes_1080p"])
def test_gbuffer_matches_host_producer(name):
    fi = frame_inputs(name)
    announce(f"device G-buffer {name}")
    r = _renderer(fi)
    df = _device(fi)
    r.render_gbuffer(df.scene(), fi.view_proj)
    depth, nrm, alb = r.read_gbuffer()
    np.testing.assert_array_equal(depth.view(np.uint32), fi.depth.view(np.uint32))
    _normals_close(nrm, fi.gnormal, 1e-4, 1)
    _off_by_one(alb, fi.galbedo, 1e-4)


@pytest.mark.parametrize("name", ["cornell_256", "boxes_1080p"])
def test_rsm_chain_matches_host_producer(name):
    fi = frame_inputs(name)
    announce(f"device RSM {name}")
    r = _renderer(fi)
    df = _device(fi)
    r.render_rsm(df.scene(), fi.light)
    res = int(fi.light.rsm_render_resolution)
    for level, (flux, nrm, dls) in enumerate(fi.rsm_levels):
        gf, gn, gd = r.read_rsm(level, res)
        # Flux and depth/depth^2 (and their 2x2 sums and means) are exact.
        np.testing.assert_array_equal(gf, flux)
        np.testing.assert_array_equal(gd, dls)
        # A mip normal is the repacked mean of four unpacked children, so
        # last-ulp sin / cos / atan2 differences reach it more often (0.3 %
        # of the 128^2 level of boxes_1080p) and compound over levels.
        if level == 0:
            _normals_close(gn, nrm, 1e-4, 1)
        else:
            _normals_close(gn, nrm, 2e-2, 4)


def test_frame_from_device_inputs_matches_host_inputs():
    from dynamicradiancevolume_amd.pipeline import entries_sh

    fi = frame_inputs("cornell_256")
    announce("frame from device-rendered inputs")
    df = _device(fi)
    r = _renderer(fi)
    r.render_frame(df)
    _, live = r.cache_count()
    e0 = r.read_caches(live)
    out0 = df.out.clone()
    # the same frame with both inputs rendered on the device
    r2 = _renderer(fi)
    w = fi.workload
    r2.set_frame(fi.frame())
    r2.update_volume(w.eye, fi.bbox[:3], fi.bbox[3:])
    r2.voxelize_scene(df.positions, len(fi.positions), adaption_factor=w.adaption_factor)
    gb = r2.render_gbuffer(df.scene(), fi.view_proj)
    rsm = r2.render_rsm(df.scene(), fi.light)
    r2.allocate_caches(gb)
    r2.light_caches_rsm(rsm, fi.light)
    out1 = torch.zeros_like(df.out)
    r2.apply_caches(gb, out1)
    _, live2 = r2.cache_count()
    e1 = r2.read_caches(live2)
    assert live2 == live
    o0 = np.lexsort(e0[:, :3].view(np.uint32).T)
    o1 = np.lexsort(e1[:, :3].view(np.uint32).T)
    np.testing.assert_array_equal(e0[o0, :3], e1[o1, :3])
    assert rel_l2(entries_sh(e1[o1], w.sh_bands), entries_sh(e0[o0], w.sh_bands)) <= 1e-4
    assert rel_l2(out1.cpu().numpy(), out0.cpu().numpy()) <= 1e-4
