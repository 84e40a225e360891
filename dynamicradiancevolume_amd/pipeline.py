"""Python host-side mirror of the reference renderer's indirect-lighting passes.

``Renderer`` exposes the pass interface of ``Renderer`` /
``Voxelization`` (rendering/renderer.hpp:46-199, rendering/voxelization.hpp:37)
over the C ABI in ``include/drv.h``: ``voxelize_scene``, ``allocate_caches``,
``light_caches_rsm`` and ``apply_caches``, plus the host UBO updates
(``update_volume``, ``set_frame``) and read-back helpers.  Device memory is
held in torch tensors on ``cuda:0``; every computation runs in the HIP
library, which must be built (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from .workloads import FrameInputs, Workload

IDENTITY = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0)

# Canonical SH coefficient order and their std430 offsets
# (shader/lightcache.glsl:33-58): SH00, SH1-1, SH10, SH1+1, SH2-2, SH2-1, SH20, SH2+1, SH2+2.
SH_OFFSETS = np.array(
    [[7, 11, 15], [4, 5, 6], [8, 9, 10], [12, 13, 14], [16, 17, 18], [20, 21, 22], [19, 23, 27], [24, 25, 26],
     [28, 29, 30]], dtype=np.int64)


class DrvError(RuntimeError):
    pass


def entries_sh(entries: np.ndarray, sh_bands: int) -> np.ndarray:
    """[n, stride] std430 entries -> [n, 27] coefficients (zeros for absent bands)."""
    n = entries.shape[0]
    out = np.zeros((n, 27), entries.dtype)
    k = 9 if sh_bands >= 2 else 4
    out[:, : 3 * k] = entries[:, SH_OFFSETS[:k].reshape(-1)]
    return out


class DeviceFrame:
    """FrameInputs uploaded to device memory."""

    def __init__(self, fi: FrameInputs, device: str = "cuda:0"):
        self.inputs = fi
        dev = torch.device(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.positions = t(fi.positions)
        self.albedo = t(fi.albedo)
        self.normals = t(fi.normals)
        self.depth = t(fi.depth)
        self.gnormal = t(fi.gnormal)
        self.galbedo = t(fi.galbedo)
        self.groughmetal = t(fi.groughmetal) if fi.groughmetal is not None else None
        flux, rn, dls = fi.read_level
        self.flux = t(flux)
        self.rnormal = t(rn)
        self.dls = t(dls)
        self.dls_shadow = t(fi.shadow_dls)
        self.out = torch.zeros((fi.workload.height, fi.workload.width, 3), dtype=torch.float32, device=dev)

    def gbuffer(self) -> _abi.DrvGBuffer:
        w = self.inputs.workload
        rm = self.groughmetal.data_ptr() if self.groughmetal is not None else None
        return _abi.DrvGBuffer(self.depth.data_ptr(), self.gnormal.data_ptr(), self.galbedo.data_ptr(), w.width,
                               w.height, rm)

    def scene(self) -> _abi.DrvScene:
        return _abi.DrvScene(self.positions.data_ptr(), self.albedo.data_ptr(), self.normals.data_ptr(),
                             len(self.inputs.positions))

    def rsm(self) -> _abi.DrvRsm:
        return _abi.DrvRsm(self.flux.data_ptr(), self.rnormal.data_ptr(), self.dls.data_ptr(),
                           self.dls_shadow.data_ptr(), self.inputs.read_resolution, self.inputs.read_height)


class Renderer:
    def __init__(self, cfg: _abi.DrvConfig, stream: torch.cuda.Stream | None = None):
        self.lib = _abi.load_drv()
        self.cfg = cfg
        self.ctx = C.c_void_p()
        self._check(self.lib.drv_create(C.byref(cfg), C.byref(self.ctx)), "drv_create")
        self.stream = stream
        if stream is not None:
            self._check(self.lib.drv_set_stream(self.ctx, C.c_void_p(stream.cuda_stream)), "drv_set_stream")

    @classmethod
    def for_workload(cls, w: Workload, stream=None, **overrides) -> "Renderer":
        cfg = w.with_(**overrides).config() if overrides else w.config()
        return cls(cfg, stream)

    def close(self):
        if self.ctx:
            self.lib.drv_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != _abi.DRV_OK:
            msg = self.lib.drv_last_error(self.ctx).decode() if self.ctx else ""
            raise DrvError(f"{what} failed ({rc}): {msg}")

    @property
    def stride(self) -> int:
        return 32 if self.cfg.sh_bands >= 2 else 16

    # --- UBO updates -----------------------------------------------------
    def set_frame(self, frame: _abi.DrvFrame):
        self._check(self.lib.drv_set_frame(self.ctx, C.byref(frame)), "drv_set_frame")

    def update_volume(self, camera_position, scene_min, scene_max):
        self._check(self.lib.drv_update_volume(self.ctx, _abi.f3(*camera_position), _abi.f3(*scene_min),
                                               _abi.f3(*scene_max)), "drv_update_volume")

    def volume(self) -> _abi.DrvVolume:
        v = _abi.DrvVolume()
        self._check(self.lib.drv_get_volume(self.ctx, C.byref(v)), "drv_get_volume")
        return v

    def set_volume(self, v: _abi.DrvVolume):
        self._check(self.lib.drv_set_volume(self.ctx, C.byref(v)), "drv_set_volume")

    def sh_constants(self) -> _abi.DrvShConstants:
        s = _abi.DrvShConstants()
        self._check(self.lib.drv_get_sh_constants(self.ctx, C.byref(s)), "drv_get_sh_constants")
        return s

    # --- passes ------------------------------------------------------------
    def voxelize_scene(self, positions: torch.Tensor, n_tri: int, world=IDENTITY, adaption_factor: float = 1.0,
                       indices: torch.Tensor | None = None):
        self._check(self.lib.drv_voxelize_begin(self.ctx), "drv_voxelize_begin")
        self._check(self.lib.drv_voxelize_mesh(self.ctx, positions.data_ptr(),
                                               indices.data_ptr() if indices is not None else None, n_tri,
                                               _abi.fvec(world, 16)), "drv_voxelize_mesh")
        self._check(self.lib.drv_voxelize_end(self.ctx, adaption_factor), "drv_voxelize_end")

    def set_voxel_overlap(self, on: bool):
        """drv_set_voxel_overlap: the next voxelization may start right after
        the previous light pass (the scene's geometry must not be written on
        the stream after it)."""
        self._check(self.lib.drv_set_voxel_overlap(self.ctx, int(on)), "drv_set_voxel_overlap")

    def allocate_caches(self, gbuffer: _abi.DrvGBuffer):
        self._check(self.lib.drv_allocate_caches(self.ctx, C.byref(gbuffer)), "drv_allocate_caches")

    def light_caches_rsm(self, rsm: _abi.DrvRsm, light: _abi.DrvSpotLight):
        self._check(self.lib.drv_light_caches(self.ctx, C.byref(rsm), C.byref(light)), "drv_light_caches")

    def light_caches_noalloc(self, rsm: _abi.DrvRsm, light: _abi.DrvSpotLight):
        """drv_light_caches_noalloc: the light pass with no device allocation
        (static scratch of the library; at most 65536 VPLs)."""
        self._check(self.lib.drv_light_caches_noalloc(self.ctx, C.byref(rsm), C.byref(light)),
                    "drv_light_caches_noalloc")

    def apply_caches(self, gbuffer: _abi.DrvGBuffer, out: torch.Tensor, accumulate: bool = False):
        self._check(self.lib.drv_apply_caches(self.ctx, C.byref(gbuffer), out.data_ptr(), int(accumulate)),
                    "drv_apply_caches")

    def render_frame(self, df: DeviceFrame, voxelize: bool = True, accumulate: bool = False):
        """Renderer::Draw's DYN_RADIANCE_VOLUME branch (renderer.cpp:539-594) for one view."""
        fi = df.inputs
        w = fi.workload
        self.set_frame(fi.frame())
        self.update_volume(w.eye, fi.bbox[:3], fi.bbox[3:])
        if voxelize and self.cfg.indirect_shadow:
            self.voxelize_scene(df.positions, len(fi.positions), adaption_factor=w.adaption_factor)
        self.allocate_caches(df.gbuffer())
        self.light_caches_rsm(df.rsm(), fi.light)
        self.apply_caches(df.gbuffer(), df.out, accumulate)

    def bruteforce_rsm(self, gbuffer: _abi.DrvGBuffer, rsm: _abi.DrvRsm, light: _abi.DrvSpotLight,
                       out: torch.Tensor, diffuse_only: bool = True, roughness_metallic: torch.Tensor | None = None,
                       accumulate: bool = False):
        """Renderer::ApplyRSMsBruteForce (renderer.cpp:851-874) for one light."""
        rm = roughness_metallic.data_ptr() if roughness_metallic is not None else None
        self._check(self.lib.drv_bruteforce_rsm(self.ctx, C.byref(gbuffer), rm, C.byref(rsm), C.byref(light),
                                                int(diffuse_only), out.data_ptr(), int(accumulate)),
                    "drv_bruteforce_rsm")

    def shadow_map(self) -> tuple:
        """(device pointer, resolution) of the last rendered RSM's depth buffer."""
        ptr, res = C.c_void_p(), C.c_int32()
        self._check(self.lib.drv_get_shadow_map(self.ctx, C.byref(ptr), C.byref(res)), "drv_get_shadow_map")
        return ptr.value, res.value

    def read_shadow_map(self) -> np.ndarray:
        _, res = self.shadow_map()
        out = np.empty((res, res), np.float32)
        self._check(self.lib.drv_read_shadow_map(self.ctx, out.ctypes.data), "drv_read_shadow_map")
        return out

    def direct_lighting(self, gbuffer: _abi.DrvGBuffer, light: _abi.DrvSpotLight, out: torch.Tensor,
                        roughness_metallic: torch.Tensor | None = None, normal_offset_bias: float = 0.01,
                        shadow_bias: float = 0.0001, accumulate: bool = False):
        """Renderer::ApplyDirectLighting (renderer.cpp:827-849) with the last rendered RSM's depth."""
        ptr, res = self.shadow_map()
        rm = roughness_metallic.data_ptr() if roughness_metallic is not None else None
        self._check(self.lib.drv_direct_lighting(self.ctx, C.byref(gbuffer), rm, C.byref(light), ptr, res,
                                                 normal_offset_bias, shadow_bias, out.data_ptr(), int(accumulate)),
                    "drv_direct_lighting")

    def tonemap(self, hdr: torch.Tensor, out: torch.Tensor, exposure: float = 1.0, lmax: float = 1.2):
        self._check(self.lib.drv_tonemap(self.ctx, hdr.data_ptr(), exposure, lmax, out.data_ptr()), "drv_tonemap")

    def save_pfm(self, hdr: torch.Tensor, path: str):
        self._check(self.lib.drv_save_pfm(self.ctx, hdr.data_ptr(), str(path).encode()), "drv_save_pfm")

    # --- inputs rendered on the device ------------------------------------------
    def render_gbuffer(self, scene: _abi.DrvScene, view_proj) -> _abi.DrvGBuffer:
        out = _abi.DrvGBuffer()
        self._check(self.lib.drv_render_gbuffer(self.ctx, C.byref(scene), _abi.fvec(view_proj, 16), C.byref(out)),
                    "drv_render_gbuffer")
        return out

    def render_rsm(self, scene: _abi.DrvScene, light: _abi.DrvSpotLight) -> _abi.DrvRsm:
        out = _abi.DrvRsm()
        self._check(self.lib.drv_render_rsm(self.ctx, C.byref(scene), C.byref(light), C.byref(out)),
                    "drv_render_rsm")
        return out

    def read_rsm(self, level: int, render_resolution: int):
        r = render_resolution >> level
        flux = np.empty((r, r, 3), np.uint16)
        nrm = np.empty((r, r, 2), np.int16)
        dls = np.empty((r, r, 2), np.uint16)
        self._check(self.lib.drv_read_rsm(self.ctx, level, flux.ctypes.data, nrm.ctypes.data, dls.ctypes.data),
                    "drv_read_rsm")
        return flux, nrm, dls

    def read_gbuffer(self):
        """Host copies (depth, normal, albedo) of the last rendered G-buffer."""
        w, h = self.cfg.width, self.cfg.height
        depth = np.empty((h, w), np.float32)
        nrm = np.empty((h, w, 2), np.int16)
        alb = np.empty((h, w, 3), np.uint8)
        self._check(self.lib.drv_read_gbuffer(self.ctx, depth.ctypes.data, nrm.ctypes.data, alb.ctypes.data),
                    "drv_read_gbuffer")
        return depth, nrm, alb

    # --- read-back -----------------------------------------------------------
    def synchronize(self):
        self._check(self.lib.drv_synchronize(self.ctx), "drv_synchronize")

    def cache_count(self) -> tuple[int, int]:
        d, l_ = C.c_uint32(), C.c_uint32()
        self._check(self.lib.drv_get_cache_count(self.ctx, C.byref(d), C.byref(l_)), "drv_get_cache_count")
        return d.value, l_.value

    def read_caches(self, count: int | None = None) -> np.ndarray:
        if count is None:
            count = self.cache_count()[1]
        out = np.empty((count, self.stride), np.float32)
        self._check(self.lib.drv_read_caches(self.ctx, out.ctypes.data, count), "drv_read_caches")
        return out

    def read_cav(self) -> np.ndarray:
        R, Cn = self.cfg.cav_resolution, self.cfg.cav_cascades
        out = np.empty((R, R, Cn * R), np.uint32)
        self._check(self.lib.drv_read_cav(self.ctx, out.ctypes.data), "drv_read_cav")
        return out

    def read_voxels(self, level: int) -> np.ndarray:
        w = self.cfg.voxel_resolution >> max(level, 0)
        out = np.empty((w, w, w), np.uint8)
        self._check(self.lib.drv_read_voxels(self.ctx, level, out.ctypes.data), "drv_read_voxels")
        return out

    # --- specular env-map atlas ----------------------------------------------
    def specular_atlas_geometry(self) -> tuple[int, int, int]:
        """(width, caches per row, levels) of the specular atlas."""
        w, n, lv = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.drv_get_specular_atlas(self.ctx, C.byref(w), C.byref(n), C.byref(lv)),
                    "drv_get_specular_atlas")
        return w.value, n.value, lv.value

    def read_specular_atlas(self, level: int) -> np.ndarray:
        """Packed R11G11B10F texels of one atlas level, [w >> level, w >> level]."""
        w = self.specular_atlas_geometry()[0] >> level
        out = np.empty((w, w), np.uint32)
        self._check(self.lib.drv_read_specular_atlas(self.ctx, level, out.ctypes.data), "drv_read_specular_atlas")
        return out

    def read_visibility(self, n_caches: int, n_blocks: int) -> np.ndarray:
        out = np.empty((n_caches, n_blocks), np.float32)
        self._check(self.lib.drv_read_visibility(self.ctx, out.ctypes.data, out.size), "drv_read_visibility")
        return out

    # --- surrogate visibility ---------------------------------------------------
    def mlp_set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (_abi.DRV_MLP_PARAMS,)
        self._check(self.lib.drv_mlp_set_params(self.ctx, p.ctypes.data, p.size), "drv_mlp_set_params")

    def mlp_params(self) -> np.ndarray:
        out = np.empty(_abi.DRV_MLP_PARAMS, np.float32)
        self._check(self.lib.drv_mlp_get_params(self.ctx, out.ctypes.data, out.size), "drv_mlp_get_params")
        return out

    def set_surrogate(self, enabled: bool, spot_fraction: float = 0.01, fallback_threshold: float = 0.05):
        s = _abi.DrvSurrogate(int(enabled), spot_fraction, fallback_threshold)
        self._check(self.lib.drv_set_surrogate(self.ctx, C.byref(s)), "drv_set_surrogate")

    def surrogate_stats(self) -> dict:
        s = _abi.DrvSurrogateStats()
        self._check(self.lib.drv_get_surrogate_stats(self.ctx, C.byref(s)), "drv_get_surrogate_stats")
        return {"spot_error": s.spot_error, "fell_back": bool(s.fell_back), "spot_caches": s.spot_caches}

    def mlp_collect(self, n: int, seed: int, x: torch.Tensor, y: torch.Tensor):
        assert x.dtype == torch.float32 and x.is_contiguous() and x.numel() >= 32 * n and y.numel() >= n
        self._check(self.lib.drv_mlp_collect(self.ctx, n, seed, x.data_ptr(), y.data_ptr()), "drv_mlp_collect")

    def mlp_features(self, pairs: torch.Tensor, x: torch.Tensor):
        n = pairs.shape[0]
        assert pairs.dtype == torch.int32 and pairs.is_contiguous() and x.numel() >= 32 * n
        self._check(self.lib.drv_mlp_features(self.ctx, pairs.data_ptr(), n, x.data_ptr()), "drv_mlp_features")

    def mlp_forward(self, x: torch.Tensor, p: torch.Tensor):
        n = x.shape[0]
        assert x.dtype == torch.float32 and x.is_contiguous() and p.numel() >= n
        self._check(self.lib.drv_mlp_forward(self.ctx, x.data_ptr(), n, p.data_ptr()), "drv_mlp_forward")

    def mlp_gradient(self, x: torch.Tensor, y: torch.Tensor) -> np.ndarray:
        out = np.empty(_abi.DRV_MLP_PARAMS + 1, np.float32)
        self._check(self.lib.drv_mlp_gradient(self.ctx, x.data_ptr(), y.data_ptr(), x.shape[0], out.ctypes.data),
                    "drv_mlp_gradient")
        return out

    def mlp_train(self, x: torch.Tensor, y: torch.Tensor, n: int, batch: int, steps: int, lr: float,
                  seed: int = 0) -> float:
        loss = C.c_float()
        self._check(self.lib.drv_mlp_train(self.ctx, x.data_ptr(), y.data_ptr(), n, batch, steps, lr, seed,
                                           C.byref(loss)), "drv_mlp_train")
        return loss.value

    def mlp_finetune(self, n_samples: int, batch: int, steps: int, lr: float, seed: int) -> float:
        loss = C.c_float()
        self._check(self.lib.drv_mlp_finetune(self.ctx, n_samples, batch, steps, lr, seed, C.byref(loss)),
                    "drv_mlp_finetune")
        return loss.value

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._check(self.lib.drv_set_tuning(self.ctx, C.byref(t)), "drv_set_tuning")

    def tuning(self) -> _abi.DrvTuning:
        t = _abi.DrvTuning()
        self._check(self.lib.drv_get_tuning(self.ctx, C.byref(t)), "drv_get_tuning")
        return t

    def set_profiling(self, on: bool):
        self._check(self.lib.drv_set_profiling(self.ctx, int(on)), "drv_set_profiling")

    def stage_times(self) -> dict:
        t = _abi.DrvStageTimes()
        self._check(self.lib.drv_get_stage_times(self.ctx, C.byref(t)), "drv_get_stage_times")
        return {n: getattr(t, n) for n, _ in t._fields_}


class Batch:
    """Several views lit by one persistent work-queue kernel or by their own
    passes back to back (drv_batch_light_caches), each view a Renderer on the
    same stream."""

    def __init__(self, renderers: list):
        self.lib = _abi.load_drv()
        self.renderers = renderers
        arr = (C.c_void_p * len(renderers))(*[r.ctx.value for r in renderers])
        self.handle = C.c_void_p()
        rc = self.lib.drv_batch_create(arr, len(renderers), C.byref(self.handle))
        if rc != _abi.DRV_OK:
            raise DrvError(f"drv_batch_create failed ({rc}): views must share SH bands, shadow mode and stream")

    def close(self):
        if self.handle:
            self.lib.drv_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != _abi.DRV_OK:
            raise DrvError(f"{what} failed ({rc}): {self.lib.drv_batch_last_error(self.handle).decode()}")

    def set_tuning(self, caches_per_wave: int, vpl_tile: int, waves_per_cu: int = 64):
        self._check(self.lib.drv_batch_set_tuning(self.handle, caches_per_wave, vpl_tile, waves_per_cu),
                    "drv_batch_set_tuning")

    def tuning(self) -> tuple:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.drv_batch_get_tuning(self.handle, C.byref(a), C.byref(b), C.byref(c)),
                    "drv_batch_get_tuning")
        return a.value, b.value, c.value

    @staticmethod
    def _arrays(frames):
        rsms = (_abi.DrvRsm * len(frames))(*[df.rsm() for df in frames])
        lights = (_abi.DrvSpotLight * len(frames))(*[df.inputs.light for df in frames])
        return rsms, lights

    def light_caches(self, frames: list):
        rsms, lights = self._arrays(frames)
        self._check(self.lib.drv_batch_light_caches(self.handle, rsms, lights), "drv_batch_light_caches")

    def autotune(self, frames: list, reps: int = 3) -> tuple:
        rsms, lights = self._arrays(frames)
        cpw, tile, ms = C.c_int32(), C.c_int32(), C.c_float()
        self._check(self.lib.drv_batch_autotune(self.handle, rsms, lights, reps, C.byref(cpw), C.byref(tile),
                                                C.byref(ms)), "drv_batch_autotune")
        return cpw.value, tile.value, ms.value

    def set_surrogate(self, enabled: bool, params: np.ndarray | None = None, spot_fraction: float = 0.01,
                      fallback_threshold: float = 0.05):
        """drv_batch_set_surrogate: one network for every view (params, when
        given), each view with its own spot check and fallback."""
        s = _abi.DrvSurrogate(int(enabled), spot_fraction, fallback_threshold)
        ptr = None
        if params is not None:
            p = np.ascontiguousarray(params, np.float32)
            assert p.shape == (_abi.DRV_MLP_PARAMS,)
            ptr = p.ctypes.data_as(C.POINTER(C.c_float))
        self._check(self.lib.drv_batch_set_surrogate(self.handle, ptr, C.byref(s)), "drv_batch_set_surrogate")

    def mlp_finetune(self, n_samples: int, batch: int, steps: int, lr: float, seed: int) -> float:
        loss = C.c_float()
        self._check(self.lib.drv_batch_mlp_finetune(self.handle, n_samples, batch, steps, lr, seed, C.byref(loss)),
                    "drv_batch_mlp_finetune")
        return loss.value

    def render_frame(self, frames: list, voxelize: bool = True):
        """Every view's Renderer::Draw, with the light pass batched."""
        for r, df in zip(self.renderers, frames):
            fi = df.inputs
            w = fi.workload
            r.set_frame(fi.frame())
            r.update_volume(w.eye, fi.bbox[:3], fi.bbox[3:])
            if voxelize and r.cfg.indirect_shadow:
                r.voxelize_scene(df.positions, len(fi.positions), adaption_factor=w.adaption_factor)
            r.allocate_caches(df.gbuffer())
        self.light_caches(frames)
        for r, df in zip(self.renderers, frames):
            r.apply_caches(df.gbuffer(), df.out)


class FrameGraph:
    """A HIP graph of whatever a Renderer's stream receives between
    begin() and end() (drv_graph_begin / drv_graph_end), replayed by launch()."""

    def __init__(self, renderer: Renderer):
        self.r = renderer

    def begin(self):
        self.r._check(self.r.lib.drv_graph_begin(self.r.ctx), "drv_graph_begin")

    def end(self):
        self.r._check(self.r.lib.drv_graph_end(self.r.ctx), "drv_graph_end")

    def launch(self):
        self.r._check(self.r.lib.drv_graph_launch(self.r.ctx), "drv_graph_launch")
