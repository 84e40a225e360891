"""Surrogate visibility in a batch of views (the combined batched + surrogate
configuration), against the single-view surrogate path and the cone trace.

* a batch with the network lights each view exactly as that view's own
  drv_light_caches with the same network does (bit for bit, spot-check
  statistics included);
* with a fallback threshold of 0 every view's spot check fails and every view
  equals the cone-traced batch bit for bit;
* drv_batch_mlp_finetune trains one network on samples of every view and
  hands it to every view.
"""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu


def _views():
    from dynamicradiancevolume_amd import workloads as W

    c = W.CORNELL_256
    b = W.BOXES_1080P
    # a batch shares SH bands: every view with SH L2
    return [
        c.with_(sh_bands=2),
        c.with_(name="cornell_wide", seed=3, width=320, height=180, sh_bands=2,
                light=W.replace(c.light, rsm_resolution=128, rsm_read_lod=1)),
        b.with_(name="boxes_720p", seed=5, width=1280, height=720, max_caches=65536,
                light=W.replace(b.light, rsm_resolution=128, rsm_read_lod=1)),
    ]


_inputs = {}


def _frames():
    from dynamicradiancevolume_amd.workloads import make_inputs

    out = []
    for w in _views():
        if w.name not in _inputs:
            _inputs[w.name] = make_inputs(w)
        out.append(_inputs[w.name])
    return out


def _sorted_entries(r):
    _, live = r.cache_count()
    e = r.read_caches(live)
    return e[np.lexsort(e[:, :3].view(np.uint32).T)]


def _batch(frames, stream, params=None, threshold=1e9, spot=0.05):
    from dynamicradiancevolume_amd.pipeline import Batch, DeviceFrame, Renderer

    rs = [Renderer(fi.workload.config(), stream) for fi in frames]
    dfs = [DeviceFrame(fi) for fi in frames]
    b = Batch(rs)
    b.set_tuning(0, 1024)   # per-view passes (the surrogate switches to them too)
    if params is not None:
        b.set_surrogate(True, params, spot_fraction=spot, fallback_threshold=threshold)
    with torch.cuda.stream(stream):
        b.render_frame(dfs)
    stream.synchronize()
    return b, rs, dfs


def test_batch_surrogate_views_equal_single_views():
    from dynamicradiancevolume_amd import surrogate as S
    from dynamicradiancevolume_amd.pipeline import DeviceFrame, Renderer

    announce("batch surrogate == single-view surrogate (3 views)")
    params = S.load_weights()
    frames = _frames()
    stream = torch.cuda.Stream()
    b, rs, _ = _batch(frames, stream, params)
    assert b.tuning()[0] == 0   # the network runs in per-view passes
    for r, fi in zip(rs, frames):
        one = Renderer(fi.workload.config())
        one.mlp_set_params(params)
        one.set_surrogate(True, spot_fraction=0.05, fallback_threshold=1e9)
        one.render_frame(DeviceFrame(fi))
        np.testing.assert_array_equal(_sorted_entries(r), _sorted_entries(one))
        sb, so = r.surrogate_stats(), one.surrogate_stats()
        assert sb["spot_caches"] == so["spot_caches"] > 0 and not sb["fell_back"] and not so["fell_back"]
        # the spot set takes every k-th cache of the list, whose phantom
        # caches (corners outside the volume) arrive in atomic order
        assert sb["spot_error"] == pytest.approx(so["spot_error"], rel=1e-3)
        one.close()
    b.close()


def test_batch_surrogate_fallback_equals_traced_batch():
    from dynamicradiancevolume_amd import surrogate as S

    announce("batch surrogate with threshold 0 == traced batch")
    frames = _frames()
    stream = torch.cuda.Stream()
    bt, rt, _ = _batch(frames, stream)
    traced = [_sorted_entries(r) for r in rt]
    bs, rsur, _ = _batch(frames, stream, S.load_weights(), threshold=0.0)
    for r, e in zip(rsur, traced):
        assert r.surrogate_stats()["fell_back"]
        np.testing.assert_array_equal(_sorted_entries(r), e)
    bt.close()
    bs.close()


def test_batch_finetune_shares_one_network():
    from dynamicradiancevolume_amd import surrogate as S

    announce("batch fine-tune: one network for every view")
    params = S.load_weights()
    frames = _frames()
    stream = torch.cuda.Stream()
    b, rs, _ = _batch(frames, stream, params)
    with torch.cuda.stream(stream):
        loss = b.mlp_finetune(4096, 1024, 3, 1e-3, 5)
    stream.synchronize()
    assert np.isfinite(loss) and loss > 0
    p0 = rs[0].mlp_params()
    assert not np.array_equal(p0, params)
    for r in rs[1:]:
        np.testing.assert_array_equal(r.mlp_params(), p0)
    b.close()
