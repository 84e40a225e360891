"""Pins of the CPU restatement of the surrogate network and its features.

The network oracle (oracle/mlp.py) is pinned by central finite differences of
its own loss in float64 (a dropped term, a wrong transpose or a wrong mask in
the backward pass fails them), its bfloat16 rounding by torch's RNE
conversion, and the feature restatement by closed forms on hand-built
volumes.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from dynamicradiancevolume_amd import _abi
from dynamicradiancevolume_amd.surrogate import OFFSETS, init_params
from oracle import mlp


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-20, 100), rng.normal(0, 1e20, 100),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8)]]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.bf16_round(a), ref)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert mlp.bf16_round(np.float32(1.0 + 2 ** -8)) == 1.0
    assert mlp.bf16_round(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def test_layout_matches_abi():
    assert mlp.N_PARAMS == _abi.DRV_MLP_PARAMS
    assert (mlp.W1, mlp.B1, mlp.W2, mlp.B2, mlp.W3, mlp.B3, mlp.W4, mlp.B4) == tuple(OFFSETS.values())


def _mean_loss(params, x, y):
    loss, _ = mlp.loss_and_grad(params, x, y, bf16=False)
    return loss / len(x)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = init_params(3).astype(np.float64)
    params[mlp.B1:mlp.W2] = rng.normal(0, 0.1, 64)  # non-zero biases exercise their gradients
    params[mlp.B2:mlp.W3] = rng.normal(0, 0.1, 64)
    params[mlp.B3:mlp.W4] = rng.normal(0, 0.1, 64)
    x = rng.normal(0, 1, (12, 32))
    y = rng.uniform(0, 1, 12)
    _, g = mlp.loss_and_grad(params, x, y, bf16=False)
    eps = 1e-6
    picks = []
    for lo, hi in [(mlp.W1, mlp.B1), (mlp.B1, mlp.W2), (mlp.W2, mlp.B2), (mlp.B2, mlp.W3), (mlp.W3, mlp.B3),
                   (mlp.B3, mlp.W4), (mlp.W4, mlp.B4), (mlp.B4, mlp.N_PARAMS)]:
        picks += list(rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False))
    for i in picks:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (_mean_loss(p1, x, y) - _mean_loss(p2, x, y)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_forward_is_sigmoid_of_layers():
    # A network whose hidden layers pass input 0 through (identity on the
    # positive part) reduces to sigmoid(w . relu(x0) + b).
    p = np.zeros(mlp.N_PARAMS)
    d = mlp.unpack(p)
    W1 = np.zeros((32, 64)); W1[0, 0] = 1.0
    W2 = np.eye(64); W3 = np.eye(64)
    w4 = np.zeros(64); w4[0] = 2.0
    p = mlp.pack({**d, "W1": W1, "W2": W2, "W3": W3, "w4": w4, "b4": np.float64(-1.0)})
    x = np.zeros((3, 32)); x[:, 0] = [-1.0, 0.5, 2.0]
    pr, _ = mlp.forward(p, x, bf16=False)
    np.testing.assert_allclose(pr, 1.0 / (1.0 + np.exp(-(2.0 * np.maximum(x[:, 0], 0) - 1.0))), rtol=1e-12)


def test_adam_first_step_is_signed_lr():
    g = np.array([1e-3, -2.0, 0.5])
    p, m, v = mlp.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-5)


# ---- feature restatement -----------------------------------------------------

def _volume(vmin=(0.0, 0.0, 0.0), voxel=0.5, res=16):
    v = _abi.DrvVolume()
    for i in range(3):
        v.volume_world_min[i] = vmin[i]
        v.volume_world_max[i] = vmin[i] + voxel * res
    v.voxel_size_in_world = voxel
    return v


def _features(vol, res, mips, cam, wp, blk):
    from oracle import orc

    out = (C.c_float * 32)()
    fp = C.POINTER(C.c_float)
    orc.lib().orc_mlp_features(C.byref(vol), res, mips.ctypes.data, (C.c_float * 3)(*cam), (C.c_float * 3)(*wp),
                               (C.c_float * 8)(*blk), out)
    return np.array(out[:], np.float64)


def _geometry(f):
    """slot order -> G[20], S[12]"""
    G = np.array([f[8 * (k // 5) + k % 5] for k in range(20)])
    S = np.array([f[8 * (k // 3) + 5 + k % 3] for k in range(12)])
    return G, S


@pytest.mark.parametrize("fill", [0, 255])
def test_features_closed_form(fill):
    from oracle import orc

    res = 16
    vol = _volume(res=res)
    mips = np.full(orc.mip_bytes(res), fill, np.uint8)
    wp = (2.0, 3.0, 4.0)
    blk = (6.0, 3.0, 4.0, 0.25, 0.0, 0.0, -1.0, 0.0)  # VAL 4 units (8 voxels) along +x
    cam = (2.0, 3.0, 7.0)
    G, S = _geometry(_features(vol, res, mips, cam, wp, blk))
    vs = 0.5 * res
    expect = [2.0 / vs, 3.0 / vs, 4.0 / vs, 0.0, 0.0, 1.0, 6.0 / vs, 3.0 / vs, 4.0 / vs, 0.0, 0.0, -1.0,
              1.0, 0.0, 0.0, 8.0 / 64, 0.25 * 16, 6.0 / 64, 0.0, 0.0]
    np.testing.assert_allclose(G, expect, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(S, np.full(12, fill / 255.0, np.float32))


def test_features_single_voxel_sample():
    from oracle import orc

    # One occupied texel on level 0 seen by the first fine sample: distance
    # 2 + 6 * 0.5 / 8 voxels along +x from the cache at texel centre (4.5, 8.5, 8.5).
    res = 16
    vol = _volume(res=res)
    mips = np.zeros(orc.mip_bytes(res), np.uint8)
    lvl0 = mips[: res ** 3].reshape(res, res, res)
    d0 = 2.0 + 6.0 * 0.5 / 8.0
    x0 = 4.5 + d0                       # sample position in texels
    ix = int(np.floor(x0 - 0.5))
    lvl0[8, 8, ix] = 255                 # z, y, x
    wp = (4.5 * 0.5, 8.5 * 0.5, 8.5 * 0.5)
    blk = (wp[0] + 4.0, wp[1], wp[2], 0.01, 0.0, 1.0, 0.0, 0.0)  # aperture small: level 0
    G, S = _geometry(_features(vol, res, mips, (0.0, 10.0, 0.0), wp, blk))
    frac = (x0 - 0.5) - ix
    assert S[0] == pytest.approx(1.0 - frac, abs=1e-6)
