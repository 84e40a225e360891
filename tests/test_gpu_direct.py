"""Direct lighting (directdeferredlighting.frag), tonemap and PFM output on
the GPU against the oracle, with the shadow map the device rendered for the
light (its RSM depth), and a whole HDR frame composed on the device.

Tolerance: relative L2 1e-4 over the image, plus a bound on single pixels
whose PCF decision flips (a lookup sitting on a texel boundary or a depth
comparison within an ulp: sin/cos/pow differ in the last ulp between the
device and host math libraries).
"""
import numpy as np
import pytest
import torch

from canon import rel_l2
from conftest import announce, frame_inputs

pytestmark = pytest.mark.gpu


def _setup(fi):
    from dynamicradiancevolume_amd.pipeline import DeviceFrame, Renderer

    r = Renderer(fi.workload.config())
    df = DeviceFrame(fi)
    r.set_frame(fi.frame())
    r.render_rsm(df.scene(), fi.light)
    return r, df


@pytest.mark.parametrize("name", ["cornell_256", "boxes_1080p"])
def test_direct_lighting_matches_oracle(name):
    from oracle import orc

    fi = frame_inputs(name)
    announce(f"direct lighting {name}")
    r, df = _setup(fi)
    W, H = fi.workload.width, fi.workload.height
    rng = np.random.default_rng(9)
    rm = rng.integers(60, 256, (H, W, 2)).astype(np.uint8)
    rm[..., 1] = rng.integers(0, 256, (H, W))
    out = torch.zeros((H, W, 3), device="cuda")
    r.direct_lighting(df.gbuffer(), fi.light, out, torch.from_numpy(rm).cuda())
    got = out.cpu().numpy()
    ref = orc.direct_lighting(fi, r.read_shadow_map(), rm)
    assert np.abs(ref).sum() > 0
    diff = np.abs(got - ref).max(axis=2)
    flipped = diff > 1e-3 * max(ref.max(), 1e-30)
    assert flipped.mean() <= 1e-3, flipped.mean()
    keep = ~flipped
    assert rel_l2(got[keep], ref[keep]) <= 1e-4


def test_hdr_frame_tonemap_and_pfm(tmp_path):
    """Direct + indirect (cache path) into one HDR target, Drago tonemap, PFM."""
    fi = frame_inputs("cornell_256")
    announce("HDR frame: direct + indirect, tonemap, PFM")
    r, df = _setup(fi)
    r.render_frame(df)                      # indirect light into df.out
    indirect = df.out.clone()
    r.direct_lighting(df.gbuffer(), fi.light, df.out, accumulate=True)
    direct = torch.zeros_like(df.out)
    r.direct_lighting(df.gbuffer(), fi.light, direct)
    torch.cuda.synchronize()
    assert torch.allclose(df.out, indirect + direct, rtol=1e-6, atol=1e-7)
    ldr = torch.empty_like(df.out)
    r.tonemap(df.out, ldr, exposure=1.5, lmax=1.2)
    torch.cuda.synchronize()
    ref = torch.log2(df.out * 1.5 + 1.0) / np.log2(2.2)
    assert torch.allclose(ldr, ref, rtol=1e-5, atol=1e-6)
    path = tmp_path / "frame.pfm"
    r.save_pfm(df.out, path)
    raw = path.read_bytes()
    W, H = fi.workload.width, fi.workload.height
    header = f"PF\n{W} {H}\n-1.000000\n".encode()
    assert raw.startswith(header)
    img = np.frombuffer(raw[len(header):], dtype="<f4").reshape(H, W, 3)
    np.testing.assert_array_equal(img, df.out.cpu().numpy())
