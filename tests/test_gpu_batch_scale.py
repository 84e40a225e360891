"""The batched-views workload at its full size: 64 seeded views of the boxes
family (1280x720 to 3840x2160, 4096 or 16384 VPLs, own scene each; the
views the benchmark renders).

* the fused work-queue kernel (64 caches per wave, 1024-VPL tiles) and the
  split queue (with queued or per-view visibility) give every view the SH
  its own drv_light_caches gives, up to the order of the VPL sums;
* per-view passes give them bit for bit;
* a spread subset of the caches of the largest view matches the FP32 oracle.
"""
import os

import numpy as np
import pytest
import torch

from canon import rel_l2
from conftest import announce

pytestmark = pytest.mark.gpu

N_VIEWS = 64


def _sorted_sh(r):
    from dynamicradiancevolume_amd.pipeline import entries_sh

    _, live = r.cache_count()
    e = r.read_caches(live)
    order = np.lexsort(e[:, :3].view(np.uint32).T)
    return e[order, :3].copy(), entries_sh(e[order], r.cfg.sh_bands).astype(np.float64), e


@pytest.mark.timeout(600)
def test_batch_of_64_views_matches_single_views():
    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import Batch, DeviceFrame, Renderer

    views = W.batch_views(N_VIEWS, seed=0)
    announce(f"batch of {N_VIEWS} views: building inputs")
    frames = [W.make_inputs(w, n_threads=min(16, os.cpu_count() or 1)) for w in views]
    stream = torch.cuda.Stream()
    rs = [Renderer(w.config(), stream) for w in views]
    for r in rs:
        r.set_tuning(gather_max_splits=32)
    dfs = [DeviceFrame(fi) for fi in frames]
    b = Batch(rs)
    results = {}
    for cpw, tile in ((64, 1024), (0, 1024), (-1, 1024), (-2, 1024)):
        announce(f"batch of {N_VIEWS}: caches per wave {cpw}")
        b.set_tuning(cpw, tile)
        with torch.cuda.stream(stream):
            b.render_frame(dfs)
        stream.synchronize()
        results[cpw] = [_sorted_sh(r) for r in rs]
    lives = [len(x[0]) for x in results[0]]
    assert min(lives) > 1000
    big = int(np.argmax([w.width * w.height for w in views]))
    announce(f"batch of {N_VIEWS}: single views")
    for i, (w, fi) in enumerate(zip(views, frames)):
        one = Renderer(w.config())
        one.set_tuning(gather_max_splits=32)
        df = DeviceFrame(fi)
        one.render_frame(df)
        pos, sh, entries = _sorted_sh(one)
        for cpw in (64, 0, -1, -2):
            gp, gs, _ = results[cpw][i]
            np.testing.assert_array_equal(gp.view(np.uint32), pos.view(np.uint32))
        assert rel_l2(results[64][i][1], sh) <= 1e-5, (i, w.name)
        assert rel_l2(results[-1][i][1], sh) <= 1e-5, (i, w.name)
        assert rel_l2(results[-2][i][1], sh) <= 1e-5, (i, w.name)
        np.testing.assert_array_equal(results[0][i][1], sh)
        if i == big:
            from oracle import orc
            from dynamicradiancevolume_amd.pipeline import entries_sh

            announce(f"batch of {N_VIEWS}: view {i} ({w.width}x{w.height}) subset vs oracle")
            cfg = w.config()
            vol = orc.update_volume(cfg, w.eye, fi.bbox[:3], fi.bbox[3:])
            _, mips = orc.voxelize(w.voxel_resolution, vol, fi.positions, adaption_factor=w.adaption_factor)
            _, live = one.cache_count()
            raw = one.read_caches(live)
            ids = np.linspace(0, live - 1, 256).astype(np.uint32)
            ref = raw.copy()
            ref[:, 4:] = 0.0
            orc.gather_f32(cfg, vol, orc.sh_constants(), fi.light, fi, mips, ref, ids=ids)
            assert rel_l2(entries_sh(raw[ids], cfg.sh_bands), entries_sh(ref[ids], cfg.sh_bands)) < 1e-4
        one.close()
        del df
    b.close()
    for r in rs:
        r.close()


@pytest.mark.timeout(600)
def test_dense_batch_matches_single_views():
    """Batched views at the batched configuration's cache scale (4 x 240^3
    address volumes, tens of thousands of caches per view, the views of the
    benchmark's dense batch): per-view passes equal single views bit for
    bit; the split queue with per-view visibility and the fused queue within
    the VPL-summation-order tolerance."""
    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import Batch, DeviceFrame, Renderer

    views = [w.with_(max_caches=262144) for w in
             W.batch_views(3, seed=2, family=W.BOXES_1080P_DENSE, resolutions=W.DENSE_BATCH_RESOLUTIONS)]
    announce("dense batch of 3 views: building inputs")
    frames = [W.make_inputs(w, n_threads=min(16, os.cpu_count() or 1)) for w in views]
    stream = torch.cuda.Stream()
    rs = [Renderer(w.config(), stream) for w in views]
    dfs = [DeviceFrame(fi) for fi in frames]
    b = Batch(rs)
    results = {}
    for cpw in (0, -2, 64):
        announce(f"dense batch: caches per wave {cpw}")
        b.set_tuning(cpw, 1024)
        with torch.cuda.stream(stream):
            b.render_frame(dfs)
        stream.synchronize()
        results[cpw] = [_sorted_sh(r) for r in rs]
    lives = [len(x[0]) for x in results[0]]
    assert min(lives) > 20000, lives
    for i, (w, fi) in enumerate(zip(views, frames)):
        announce(f"dense batch: single view {i} ({w.width}x{w.height}, {lives[i]} caches)")
        one = Renderer(w.config())
        one.render_frame(DeviceFrame(fi))
        pos, sh, _ = _sorted_sh(one)
        for cpw in (0, -2, 64):
            np.testing.assert_array_equal(results[cpw][i][0].view(np.uint32), pos.view(np.uint32))
        np.testing.assert_array_equal(results[0][i][1], sh)
        assert rel_l2(results[-2][i][1], sh) <= 1e-5, (i, w.name)
        assert rel_l2(results[64][i][1], sh) <= 1e-5, (i, w.name)
        one.close()
    b.close()
    for r in rs:
        r.close()
