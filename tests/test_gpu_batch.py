"""Batched light pass and HIP graph capture.

* drv_batch_light_caches over uneven views equals each view's own
  drv_light_caches (the single-view path, itself pinned against the oracle)
  up to the order of the VPL sums, for every work decomposition, and bit for
  bit with per-view passes (caches per wave 0); one view is also checked
  against the FP32 oracle directly;
* the autotuner returns a valid decomposition and leaves the entries as it
  found them;
* a captured frame (allocate, batched light, apply for every view) replays
  bit-identically to the eager frame, several times, with eager work and
  synchronisations between the replays.
"""
import numpy as np
import pytest
import torch

from canon import rel_l2
from conftest import announce

pytestmark = pytest.mark.gpu

SH_TOL = 1e-5   # batch vs single view (same terms, different summation order)


def _views():
    from dynamicradiancevolume_amd import workloads as W

    c = W.CORNELL_256
    small_rsm = W.replace(c.light, rsm_resolution=128, rsm_read_lod=1)   # 64^2 VPLs from a 128^2 RSM
    b = W.BOXES_1080P
    return [
        c,
        c.with_(name="cornell_wide", seed=3, width=320, height=180, light=small_rsm),
        b.with_(name="boxes_720p", seed=5, width=1280, height=720, max_caches=65536,
                light=W.replace(b.light, rsm_resolution=128, rsm_read_lod=1)),
        b.with_(name="boxes_900p", seed=6, width=1600, height=900, max_caches=65536),
    ]


_cache = {}


def _frames(names=None):
    from dynamicradiancevolume_amd.workloads import make_inputs

    out = []
    for w in _views():
        if names and w.name not in names:
            continue
        if w.name not in _cache:
            _cache[w.name] = make_inputs(w)
        out.append(_cache[w.name])
    return out


def _by_position(r):
    from dynamicradiancevolume_amd.pipeline import entries_sh

    _, live = r.cache_count()
    e = r.read_caches(live)
    order = np.lexsort(e[:, :3].view(np.uint32).T)
    return e[order, :3].copy(), entries_sh(e[order], r.cfg.sh_bands).astype(np.float64)


def _group(frames, sh_bands, stream=None):
    from dynamicradiancevolume_amd.pipeline import DeviceFrame, Renderer

    rs = [Renderer(fi.workload.with_(sh_bands=sh_bands).config(), stream) for fi in frames]
    dfs = [DeviceFrame(fi) for fi in frames]
    return rs, dfs


def _single_view_sh(frames, sh_bands):
    rs, dfs = _group(frames, sh_bands)
    out = []
    for r, df in zip(rs, dfs):
        r.render_frame(df)
        out.append(_by_position(r))
        r.close()
    return out


@pytest.mark.parametrize("sh_bands", [1, 2])
def test_batch_matches_single_views(sh_bands):
    from dynamicradiancevolume_amd.pipeline import Batch

    frames = _frames()
    announce(f"batch of {len(frames)} views vs single-view path, SH{sh_bands}")
    ref = _single_view_sh(frames, sh_bands)
    rs, dfs = _group(frames, sh_bands)
    b = Batch(rs)
    b.render_frame(dfs)
    for (p0, s0), r, fi in zip(ref, rs, frames):
        p1, s1 = _by_position(r)
        np.testing.assert_array_equal(p0, p1)
        assert rel_l2(s1, s0) <= SH_TOL, fi.workload.name
    b.close()


def test_batch_per_view_passes_match_single_views():
    """Per-view passes (caches per wave 0) run the single-view kernels: the
    SH of every view are bit-identical to lighting the views one by one."""
    from dynamicradiancevolume_amd.pipeline import Batch

    frames = _frames()
    announce(f"batch of {len(frames)} views as per-view passes")
    ref = _single_view_sh(frames, 2)
    rs, dfs = _group(frames, 2)
    b = Batch(rs)
    b.set_tuning(0, 256)
    b.render_frame(dfs)
    for (p0, s0), r in zip(ref, rs):
        p1, s1 = _by_position(r)
        np.testing.assert_array_equal(p0, p1)
        np.testing.assert_array_equal(s1, s0)
    b.close()


def test_batch_view_matches_oracle():
    from dynamicradiancevolume_amd.pipeline import Batch
    from oracle import orc

    frames = _frames()
    announce("batch view 1 vs FP32 oracle")
    rs, dfs = _group(frames, 1)
    b = Batch(rs)
    b.render_frame(dfs)
    fi, r = frames[1], rs[1]
    res = orc.run_frame(fi)
    _, live = r.cache_count()
    assert live == res["live"]
    p1, s1 = _by_position(r)
    e = res["entries"][: res["live"]]
    order = np.lexsort(e[:, :3].view(np.uint32).T)
    from dynamicradiancevolume_amd.pipeline import entries_sh

    np.testing.assert_array_equal(e[order, :3], p1)
    assert rel_l2(s1, entries_sh(e[order], 1)) <= 1e-4


@pytest.mark.parametrize("cpw,tile", [(64, 256), (32, 1024), (16, 16384), (64, 65536), (0, 256), (-1, 1024),
                                      (-2, 1024)])
def test_batch_decompositions_agree(cpw, tile):
    from dynamicradiancevolume_amd.pipeline import Batch

    frames = _frames()
    announce(f"batch decomposition cpw={cpw} tile={tile}")
    res = []
    for t in (None, (cpw, tile)):
        rs, dfs = _group(frames, 2)
        b = Batch(rs)
        if t:
            b.set_tuning(*t)
        b.render_frame(dfs)
        res.append([_by_position(r) for r in rs])
        b.close()
    for (p0, s0), (p1, s1) in zip(*res):
        np.testing.assert_array_equal(p0, p1)
        assert rel_l2(s1, s0) <= SH_TOL


@pytest.mark.parametrize("mode", [-1, -2])
def test_split_queue_visibility_and_sh_equal_single_views(mode):
    """The split queue runs the single-view kernels' item bodies over all
    views (mode -2: the visibility of each view in a single-view launch):
    every view's per-pair visibility is bit-identical to its own
    drv_light_caches, and its SH differ only by the VPL split count (planned
    over all views together)."""
    from dynamicradiancevolume_amd.pipeline import Batch

    frames = _frames()
    announce(f"split queue (mode {mode}) of {len(frames)} views vs single views (visibility bit-exact)")
    ref = []
    rs1, dfs1 = _group(frames, 2)
    for r, df, fi in zip(rs1, dfs1, frames):
        r.render_frame(df)
        _, live = r.cache_count()
        nb = fi.n_vpl // int(fi.light.indirect_shadow_sample_interval)
        ref.append((_by_position(r), r.read_visibility(live, nb), r.read_caches(live)[:, :3].copy()))
        r.close()
    rs, dfs = _group(frames, 2)
    b = Batch(rs)
    b.set_tuning(mode, 1024)
    b.render_frame(dfs)
    for ((p0, s0), v0, pos0), r, fi in zip(ref, rs, frames):
        p1, s1 = _by_position(r)
        np.testing.assert_array_equal(p0, p1)
        assert rel_l2(s1, s0) <= SH_TOL, fi.workload.name
        _, live = r.cache_count()
        nb = v0.shape[1]
        # cache order by position (phantom caches append in atomic order;
        # caches at one position trace the same rays)
        pos1 = r.read_caches(live)[:, :3].copy()
        o0 = np.lexsort(pos0.view(np.uint32).T)
        o1 = np.lexsort(pos1.view(np.uint32).T)
        np.testing.assert_array_equal(pos1[o1], pos0[o0])
        np.testing.assert_array_equal(r.read_visibility(live, nb)[o1], v0[o0])
    b.close()


def test_split_queue_views_with_fewer_blocks_than_an_item():
    """The split queue's visibility items hold 4 shadow blocks; views with 1
    (an 8x8 read at shadow LOD 3) or 2 (a 16x8 read of the 2:1 extension)
    blocks end their only item early.  Their visibility and that of a
    4096-VPL view in the same batch equal each view's own light pass."""
    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import Batch

    c = W.CORNELL_256
    views = [
        c.with_(name="cornell_8x8_lod3", light=W.replace(c.light, rsm_resolution=64, rsm_read_lod=3,
                                                         indirect_shadow_lod=3)),
        c.with_(name="cornell_16x8_lod3", seed=4, light=W.replace(c.light, rsm_resolution=32, rsm_read_lod=1,
                                                                  rsm_wide=1, indirect_shadow_lod=3)),
        c,
    ]
    frames = []
    for w in views:
        if w.name not in _cache:
            _cache[w.name] = W.make_inputs(w)
        frames.append(_cache[w.name])
    blocks = [fi.n_vpl // int(fi.light.indirect_shadow_sample_interval) for fi in frames]
    assert blocks[:2] == [1, 2] and blocks[2] % 4 == 0, blocks
    announce(f"split queue over views of {blocks} shadow blocks vs single views")
    ref = []
    rs1, dfs1 = _group(frames, 2)
    for r, df, nb in zip(rs1, dfs1, blocks):
        r.render_frame(df)
        _, live = r.cache_count()
        ref.append((_by_position(r), r.read_visibility(live, nb), r.read_caches(live)[:, :3].copy()))
        r.close()
    rs, dfs = _group(frames, 2)
    b = Batch(rs)
    b.set_tuning(-1, 1024)
    b.render_frame(dfs)
    for ((p0, s0), v0, pos0), r, fi, nb in zip(ref, rs, frames, blocks):
        p1, s1 = _by_position(r)
        np.testing.assert_array_equal(p0, p1)
        assert rel_l2(s1, s0) <= SH_TOL, fi.workload.name
        _, live = r.cache_count()
        pos1 = r.read_caches(live)[:, :3].copy()
        o0 = np.lexsort(pos0.view(np.uint32).T)
        o1 = np.lexsort(pos1.view(np.uint32).T)
        np.testing.assert_array_equal(pos1[o1], pos0[o0])
        np.testing.assert_array_equal(r.read_visibility(live, nb)[o1], v0[o0], err_msg=fi.workload.name)
    b.close()
    for r in rs:
        r.close()


@pytest.mark.parametrize("mode", [0, -2, 64])
def test_batch_with_voxel_overlap_equals_without(mode):
    """With the voxel overlap on, each view's next voxelization starts after
    the batch's cone trace and runs beside the rest of the batch frame; over
    three batch frames the caches equal those of a batch without it, bit for
    bit, in every decomposition (compared by position: phantom caches are
    appended in atomic order)."""
    from dynamicradiancevolume_amd.pipeline import Batch

    frames = _frames()
    announce(f"batch mode {mode}: voxel overlap on vs off")
    out = []
    for overlap in (False, True):
        rs, dfs = _group(frames, 2)
        for r in rs:
            r.set_voxel_overlap(overlap)
        b = Batch(rs)
        b.set_tuning(mode, 1024)
        for _ in range(3):
            b.render_frame(dfs)
        torch.cuda.synchronize()
        out.append([_by_position(r) for r in rs])
        b.close()
        for r in rs:
            r.close()
    for (p0, s0), (p1, s1), fi in zip(out[0], out[1], frames):
        np.testing.assert_array_equal(p0, p1, err_msg=fi.workload.name)
        np.testing.assert_array_equal(s0, s1, err_msg=fi.workload.name)


def test_views_after_their_batch_is_destroyed():
    """A batch records one voxel-free event for all its views; once it is
    destroyed, the views voxelize and light on their own again (two frames
    each, voxel overlap on) with the caches of views that never joined a
    batch."""
    from dynamicradiancevolume_amd.pipeline import Batch

    frames = _frames()
    announce("views rendered on their own after their batch is destroyed")
    out = []
    for batched in (False, True):
        rs, dfs = _group(frames, 2)
        for r in rs:
            r.set_voxel_overlap(True)
        if batched:
            b = Batch(rs)
            b.set_tuning(0, 1024)
            b.render_frame(dfs)
            b.set_tuning(-2, 1024)
            b.render_frame(dfs)
            b.close()
        for _ in range(2):
            for r, df in zip(rs, dfs):
                r.render_frame(df)
        torch.cuda.synchronize()
        out.append([_by_position(r) for r in rs])
        for r in rs:
            r.close()
    for (p0, s0), (p1, s1), fi in zip(out[0], out[1], frames):
        np.testing.assert_array_equal(p0, p1, err_msg=fi.workload.name)
        np.testing.assert_array_equal(s0, s1, err_msg=fi.workload.name)


def test_batch_autotune_restores_entries():
    from dynamicradiancevolume_amd.pipeline import Batch

    frames = _frames()
    announce("batch autotune")
    rs, dfs = _group(frames, 2)
    b = Batch(rs)
    b.render_frame(dfs)
    before = [r.read_caches(r.cache_count()[1]) for r in rs]
    cpw, tile, ms = b.autotune(dfs, reps=1)
    assert cpw in (64, 32, 16, 0, -1, -2) and tile >= 256 and (tile & (tile - 1)) == 0 and ms > 0
    assert b.tuning()[:2] == (cpw, tile)
    for r, e in zip(rs, before):
        np.testing.assert_array_equal(r.read_caches(len(e)), e)


@pytest.mark.parametrize("cpw", [64, -1, -2])
def test_graph_replay_matches_eager(cpw):
    from dynamicradiancevolume_amd.pipeline import Batch, FrameGraph

    frames = _frames()
    announce(f"graph capture of a batched frame (caches per wave {cpw}), 3 replays")
    stream = torch.cuda.Stream()
    rs, dfs = _group(frames, 2, stream)
    b = Batch(rs)
    b.set_tuning(cpw, 1024)
    with torch.cuda.stream(stream):
        b.render_frame(dfs)           # eager frame (also grows every buffer)
        stream.synchronize()
        eager = [(_by_position(r), df.out.clone()) for r, df in zip(rs, dfs)]
        g = FrameGraph(rs[0])
        g.begin()
        b.render_frame(dfs)
        g.end()
        for rep in range(3):
            for df in dfs:
                df.out.fill_(-1.0)    # eager work between replays
            stream.synchronize()
            g.launch()
            stream.synchronize()
            for ((p0, s0), img0), r, df in zip(eager, rs, dfs):
                p1, s1 = _by_position(r)
                np.testing.assert_array_equal(p1, p0)
                np.testing.assert_array_equal(s1, s0)
                assert torch.equal(df.out, img0), f"replay {rep}"
    b.close()


def test_fused_queue_refuses_surrogate_views():
    """The fused queue cone-traces every pair: a batch with a view that has
    the surrogate enabled must use per-view passes (caches per wave 0), so
    the tuning and the light pass refuse the fused mode and the autotuner
    keeps per-view passes."""
    from dynamicradiancevolume_amd.pipeline import Batch, DrvError
    from dynamicradiancevolume_amd.surrogate import init_params

    announce("batch: fused queue refused with a surrogate view")
    frames = _frames(["cornell_256", "cornell_wide"])
    stream = torch.cuda.Stream()
    rs, dfs = _group(frames, 1, stream)
    b = Batch(rs)
    b.set_tuning(64, 1024)
    with torch.cuda.stream(stream):
        b.render_frame(dfs)          # fused, no surrogate yet
    rs[1].mlp_set_params(init_params(2))
    rs[1].set_surrogate(True, spot_fraction=0.05, fallback_threshold=1e9)
    with pytest.raises(DrvError):
        b.light_caches(dfs)           # still set to the fused mode
    with pytest.raises(DrvError):
        b.set_tuning(32, 1024)
    b.set_tuning(0, 1024)
    with torch.cuda.stream(stream):
        b.light_caches(dfs)
        cpw, _, _ = b.autotune(dfs, reps=1)
    assert cpw == 0 and b.tuning()[0] == 0
    stream.synchronize()
    b.close()
    for r in rs:
        r.close()
