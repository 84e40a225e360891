"""Surrogate visibility: parameters, weights file and the training driver.

The network (32 -> 64 -> 64 -> 64 -> 1, ReLU, sigmoid; BF16 operands, FP32
accumulation) predicts the shadowing that the cone trace of
shader/cacheLightingRSM.comp:167-231 computes for one (cache, shadow block)
pair.  Its inputs, forward and training kernels live in ``csrc/k_mlp.hip``;
this module only drives them through the C ABI:

* ``init_params``   He-normal initialisation (float32, the ABI layout);
* ``load_weights`` / ``save_weights``   the committed weights file;
* ``collect``       cone-traced labels from seeded scenes of a workload family,
                    rendered by the plain path and sampled on the device;
* ``train``         Adam on the device over a resident sample set.

Training runs on the GPU by construction: every step is a call of
``drv_mlp_train``.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from . import _abi
from .pipeline import DeviceFrame, Renderer
from .workloads import Workload, make_inputs

N_PARAMS = _abi.DRV_MLP_PARAMS
N_FEATURES = _abi.DRV_MLP_FEATURES
WEIGHTS = Path(__file__).resolve().parent / "weights" / "visibility_mlp.f32"

# Parameter layout (float offsets): W1[32][64], b1, W2[64][64], b2, W3, b3, w4, b4.
OFFSETS = {"W1": 0, "b1": 2048, "W2": 2112, "b2": 6208, "W3": 6272, "b3": 10368, "w4": 10432, "b4": 10496}


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal hidden layers, small output layer, zero hidden biases; the
    output bias starts at 2 because most pairs are unshadowed."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float64)
    p[0:2048] = rng.normal(0.0, np.sqrt(2.0 / 32), 2048)
    p[2112:6208] = rng.normal(0.0, np.sqrt(2.0 / 64), 4096)
    p[6272:10368] = rng.normal(0.0, np.sqrt(2.0 / 64), 4096)
    p[10432:10496] = rng.normal(0.0, np.sqrt(1.0 / 64), 64)
    p[10496] = 2.0
    return p.astype(np.float32)


def load_weights(path: Path | str = WEIGHTS) -> np.ndarray:
    w = np.fromfile(str(path), dtype="<f4")
    if w.size != N_PARAMS:
        raise ValueError(f"{path}: expected {N_PARAMS} float32 values, found {w.size}")
    return w


def save_weights(params: np.ndarray, path: Path | str = WEIGHTS) -> None:
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    np.ascontiguousarray(params, dtype="<f4").tofile(str(path))


class SampleSet:
    """Device-resident training samples x[n][32], y[n]."""

    def __init__(self, capacity: int, device: str = "cuda:0"):
        self.x = torch.empty((capacity, N_FEATURES), dtype=torch.float32, device=device)
        self.y = torch.empty((capacity,), dtype=torch.float32, device=device)
        self.n = 0

    @property
    def capacity(self) -> int:
        return self.x.shape[0]


def collect(family: Workload, seeds, per_view: int, samples: SampleSet, log=None) -> SampleSet:
    """Renders each seeded scene of `family` with the cone-traced path and
    draws `per_view` random (cache, block) samples from it."""
    r = Renderer.for_workload(family)
    try:
        for i, seed in enumerate(seeds):
            n = min(per_view, samples.capacity - samples.n)
            if n <= 0:
                break
            w = family.with_(seed=int(seed))
            df = DeviceFrame(make_inputs(w))
            r.render_frame(df)
            r.mlp_collect(n, int(seed) * 7919 + 1, samples.x[samples.n:], samples.y[samples.n:])
            samples.n += n
            if log:
                log(f"collect: scene seed {seed} ({i + 1}/{len(seeds)}), {samples.n} samples")
        torch.cuda.synchronize()
    finally:
        r.close()
    return samples


def train(r: Renderer, samples: SampleSet, steps: int, batch: int = 16384, lr: float = 2e-3,
          lr_final: float = 2e-4, seed: int = 0, chunk: int = 100, log=None) -> list[float]:
    """Adam on the device with a log-linear learning-rate decay, `chunk`
    steps per call; returns the mean minibatch loss after each chunk."""
    losses = []
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        frac = done / max(1, steps)
        rate = lr * (lr_final / lr) ** frac
        losses.append(r.mlp_train(samples.x, samples.y, samples.n, batch, k, rate, seed + done))
        done += k
        if log:
            log(f"train: step {done}/{steps} lr {rate:.2e} loss {losses[-1]:.4f}")
    return losses
