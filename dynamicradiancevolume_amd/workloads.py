"""Named, seeded workloads and their pipeline inputs.

Single-view configurations of the indirect-lighting path:

``cornell_256``   Cornell box (5 walls, open front, 2 boxes; albedos U[0.2, 0.9]),
                  256x256 G-buffer, spot light with a 64x64 RSM read at LOD 0
                  (4096 VPLs), CAV 4 x 32^3, 64^3 voxels, SH L1, cone-traced
                  indirect shadows (shadow LOD 2).
``boxes_1080p``   2000 random boxes and ramps (~1M triangles) on a ground plane,
                  1920x1080 G-buffer, 256x256 RSM read at LOD 1 (16384 VPLs),
                  CAV 4 x 64^3, 128^3 voxels, SH L2, shadow LOD 2.
``boxes_1080p_dense``  the same frame with CAV 4 x 240^3 (62,920 live caches).

Inputs are produced by ``libdrv_inputs.so`` in the reference's texture
formats (see ``csrc/inputs/drv_inputs.h``).  Everything is seeded.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _abi


@dataclass(frozen=True)
class SpotSpec:
    intensity: tuple = (50.0, 50.0, 50.0)
    position: tuple = (0.0, 3.8, -0.5)
    direction: tuple = (0.15, -1.0, -0.3)
    half_angle: float = 0.45
    rsm_resolution: int = 64
    rsm_read_lod: int = 0
    indirect_shadow_lod: int = 2
    near_plane: float = 0.1   # Light::nearPlane (scene/scene.cpp:6)
    far_plane: float = 10000.0  # Light::farPlane (scene/scene.cpp:7)
    rsm_wide: int = 0         # 1: a 2:1 RSM (twice as wide as tall)


@dataclass(frozen=True)
class Workload:
    name: str
    scene_kind: int
    seed: int
    n_objects: int
    tess: int
    width: int
    height: int
    eye: tuple
    target: tuple
    hfov_deg: float
    near: float
    far: float
    light: SpotSpec
    cav_resolution: int
    cav_cascades: int
    cascade0_world_size: float
    voxel_resolution: int
    sh_bands: int
    indirect_shadow: int = 1
    transition_size: float = 2.0   # renderer.cpp:47
    max_caches: int = 262144
    adaption_factor: float = 1.0   # voxels snap to the target
    specular: int = 0              # INDIRECT_SPECULAR (renderer.cpp:49: off)
    specular_size: int = 16        # per-cache env-map side (renderer.cpp:43)
    hole_fill: int = 0             # hole-fill level (renderer.cpp:44)

    def config(self) -> _abi.DrvConfig:
        return _abi.DrvConfig(
            width=self.width, height=self.height, max_caches=self.max_caches,
            cav_resolution=self.cav_resolution, cav_cascades=self.cav_cascades,
            voxel_resolution=self.voxel_resolution, sh_bands=self.sh_bands,
            indirect_shadow=self.indirect_shadow, cav_transition_size=self.transition_size,
            cav_cascade0_world_size=self.cascade0_world_size, indirect_specular=self.specular,
            specular_envmap_size=self.specular_size, specular_hole_fill_level=self.hole_fill,
        )

    def with_(self, **kw) -> "Workload":
        return replace(self, **kw)


CORNELL_256 = Workload(
    name="cornell_256", scene_kind=0, seed=0, n_objects=2, tess=8, width=256, height=256,
    eye=(0.0, 0.3, 3.8), target=(0.0, -0.3, -4.0), hfov_deg=70.0, near=0.05, far=100.0,
    light=SpotSpec(), cav_resolution=32, cav_cascades=4, cascade0_world_size=4.0,
    voxel_resolution=64, sh_bands=1, max_caches=65536,
)

BOXES_1080P = Workload(
    name="boxes_1080p", scene_kind=1, seed=0, n_objects=2000, tess=7, width=1920, height=1080,
    eye=(0.0, 6.0, 6.0), target=(0.0, 0.0, -6.0), hfov_deg=75.0, near=0.1, far=500.0,
    light=SpotSpec(intensity=(2000.0, 2000.0, 2000.0), position=(15.0, 30.0, 12.0),
                   direction=(-15.0, -30.0, -18.0), half_angle=0.35, rsm_resolution=256,
                   rsm_read_lod=1, indirect_shadow_lod=2),
    cav_resolution=64, cav_cascades=4, cascade0_world_size=4.0, voxel_resolution=128, sh_bands=2,
    max_caches=262144,
)

# The boxes_1080p frame at the cache scale its configuration quotes (~60k
# live caches).  With 64^3 cells per cascade a cell covers at least ~39
# pixels across at 1080p for every surface inside its cascade (the cascade
# reaches size/2 from the camera, a cell is size/64), so the 4 x 64^3 clipmap
# holds only ~4-8k distinct surface caches in any 1080p view; 240^3 cells per
# cascade give 62,920 live caches on this frame.
BOXES_1080P_DENSE = BOXES_1080P.with_(name="boxes_1080p_dense", cav_resolution=240)

WORKLOADS = {w.name: w for w in (CORNELL_256, BOXES_1080P, BOXES_1080P_DENSE)}

# Batched views: resolutions of 16:9 frames from 720p to 4K, and RSM reads of
# 4096, 16384 or 32768 VPLs: square power-of-two reads (the Morton order of
# cacheLightingRSM.comp:142) or the 2:1 extension (256 x 128).
BATCH_RESOLUTIONS = ((1280, 720), (1600, 900), (1920, 1080), (2560, 1440), (3200, 1800), (3840, 2160))
# (render width, read LOD, 2:1) -> 64^2, 128^2 or 256 x 128 VPLs
BATCH_RSM = ((128, 1, 0), (256, 1, 0), (512, 1, 1))


# Many small views (thumbnails, probes): 320x180 with 4096 VPLs each, where
# one launch per view and stage leaves most of the GPU idle.
SMALL_RESOLUTIONS = ((320, 180),)
# Views of the batch at the tens-of-thousands-of-caches scale (4 x 240^3 CAV).
DENSE_BATCH_RESOLUTIONS = ((1920, 1080), (2560, 1440))
SMALL_RSM = ((128, 1, 0),)


def batch_views(n: int, seed: int = 0, family: Workload = BOXES_1080P, resolutions=BATCH_RESOLUTIONS,
                rsm=BATCH_RSM) -> list:
    """`n` views of seeded scenes of `family`, each with its own scene seed,
    camera, resolution and VPL count (drawn with numpy's PCG64 from `seed`).
    Capacities are sized for the resolution."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w_, h_ = resolutions[int(rng.integers(len(resolutions)))]
        rsm_res, lod, wide = rsm[int(rng.integers(len(rsm)))]
        ang = float(rng.uniform(-0.6, 0.6))
        eye = (6.0 * math.sin(ang), float(rng.uniform(4.0, 8.0)), 6.0 * math.cos(ang))
        target = (float(rng.uniform(-3, 3)), 0.0, float(rng.uniform(-8, -4)))
        light = replace(family.light, rsm_resolution=rsm_res, rsm_read_lod=lod, rsm_wide=wide)
        cap = 65536 if w_ * h_ <= 1920 * 1080 else 131072
        out.append(family.with_(name=f"{family.name}_view{i}", seed=int(rng.integers(1 << 30)), width=w_, height=h_,
                                eye=eye, target=target, light=light, max_caches=cap))
    return out


@dataclass
class FrameInputs:
    """Host-side inputs of one view: scene, G-buffer, light and RSM pyramid."""

    workload: Workload
    positions: np.ndarray      # float32 [n_tri, 9]
    albedo: np.ndarray         # float32 [n_tri, 3]
    normals: np.ndarray        # float32 [n_tri, 3]
    bbox: np.ndarray           # float32 [6]
    view_proj: np.ndarray      # float32 [16]
    inv_view_proj: np.ndarray  # float32 [16]
    depth: np.ndarray          # float32 [H, W]
    gnormal: np.ndarray        # int16 [H, W, 2]
    galbedo: np.ndarray        # uint8 [H, W, 3]
    light: _abi.DrvSpotLight
    rsm_levels: list = field(default_factory=list)  # [(flux u16 [r,r,3], normal i16 [r,r,2], dls u16 [r,r,2])]
    groughmetal: np.ndarray | None = None  # uint8 [H, W, 2]: roughness, metallic (RG8)

    @property
    def read_level(self):
        return self.rsm_levels[self.workload.light.rsm_read_lod]

    @property
    def shadow_dls(self):
        return self.rsm_levels[self.workload.light.rsm_read_lod + self.workload.light.indirect_shadow_lod][2]

    @property
    def read_resolution(self) -> int:
        """Width of the RSM read level."""
        return int(self.light.rsm_read_resolution)

    @property
    def read_height(self) -> int:
        """Rows of the RSM read level (the width, or half of it for a 2:1 RSM)."""
        return int(self.light.rsm_read_height)

    @property
    def n_vpl(self) -> int:
        return self.read_resolution * self.read_height

    def frame(self) -> _abi.DrvFrame:
        f = _abi.DrvFrame()
        for i in range(16):
            f.inverse_view_projection[i] = float(self.inv_view_proj[i])
        for i in range(3):
            f.camera_position[i] = float(self.workload.eye[i])
        return f


def spot_light(spec: SpotSpec) -> _abi.DrvSpotLight:
    lib = _abi.load_drv()
    p = _abi.DrvSpotParams(
        intensity=_abi.f3(*spec.intensity), position=_abi.f3(*spec.position),
        direction=_abi.f3(*spec.direction), half_angle=spec.half_angle,
        rsm_resolution=spec.rsm_resolution, rsm_read_lod=spec.rsm_read_lod,
        indirect_shadow_lod=spec.indirect_shadow_lod, near_plane=spec.near_plane, far_plane=spec.far_plane,
        rsm_wide=spec.rsm_wide,
    )
    out = _abi.DrvSpotLight()
    rc = lib.drv_prepare_spot_light(C.byref(p), C.byref(out))
    if rc != _abi.DRV_OK:
        raise ValueError(f"drv_prepare_spot_light failed ({rc})")
    return out


def build_scene(w: Workload):
    lib = _abi.load_inputs()
    n = lib.drvin_scene_build(w.scene_kind, w.seed, w.n_objects, w.tess, None, None, None, 0, None)
    if n < 0:
        raise ValueError("bad scene parameters")
    pos = np.empty((n, 9), np.float32)
    alb = np.empty((n, 3), np.float32)
    nrm = np.empty((n, 3), np.float32)
    bbox = np.empty(6, np.float32)
    lib.drvin_scene_build(w.scene_kind, w.seed, w.n_objects, w.tess, pos.ctypes.data, alb.ctypes.data,
                          nrm.ctypes.data, n, bbox.ctypes.data)
    return pos, alb, nrm, bbox


def make_inputs(w: Workload, n_threads: int = 0) -> FrameInputs:
    lib = _abi.load_inputs()
    pos, alb, nrm, bbox = build_scene(w)
    vp = np.empty(16, np.float32)
    ivp = np.empty(16, np.float32)
    rc = lib.drvin_camera(_abi.f3(*w.eye), _abi.f3(*w.target), math.radians(w.hfov_deg), w.width / w.height,
                          w.near, w.far, vp.ctypes.data, ivp.ctypes.data)
    if rc != 0:
        raise ValueError("singular camera")
    H, W = w.height, w.width
    depth = np.empty((H, W), np.float32)
    gn = np.empty((H, W, 2), np.int16)
    ga = np.empty((H, W, 3), np.uint8)
    lib.drvin_raster_gbuffer(pos.ctypes.data, alb.ctypes.data, nrm.ctypes.data, len(pos), vp.ctypes.data, W, H,
                             depth.ctypes.data, gn.ctypes.data, ga.ctypes.data, n_threads)
    light = spot_light(w.light)
    r = int(light.rsm_render_resolution)
    rh = int(light.rsm_render_height)
    flux = np.empty((rh, r, 3), np.uint16)
    rn = np.empty((rh, r, 2), np.int16)
    dls = np.empty((rh, r, 2), np.uint16)
    lib.drvin_raster_rsm(pos.ctypes.data, alb.ctypes.data, nrm.ctypes.data, len(pos), C.byref(light),
                         flux.ctypes.data, rn.ctypes.data, dls.ctypes.data, n_threads)
    levels = [(flux, rn, dls)]
    while rh > 1:
        r2, rh2 = r // 2, rh // 2
        f2 = np.empty((rh2, r2, 3), np.uint16)
        n2 = np.empty((rh2, r2, 2), np.int16)
        d2 = np.empty((rh2, r2, 2), np.uint16)
        pf, pn, pd = levels[-1]
        lib.drvin_rsm_downsample(r, rh, pf.ctypes.data, pn.ctypes.data, pd.ctypes.data, f2.ctypes.data,
                                 n2.ctypes.data, d2.ctypes.data)
        levels.append((f2, n2, d2))
        r, rh = r2, rh2
    return FrameInputs(workload=w, positions=pos, albedo=alb, normals=nrm, bbox=bbox, view_proj=vp,
                       inv_view_proj=ivp, depth=depth, gnormal=gn, galbedo=ga, light=light, rsm_levels=levels,
                       groughmetal=roughness_metallic(w))


def roughness_metallic(w: Workload) -> np.ndarray:
    """RG8 material texture of the view: roughness uniform in [0.1, 1];
    metallic 0 (60 %), 1 (20 %) or uniform (20 %); drawn per pixel with
    numpy's PCG64 seeded by the scene seed."""
    rng = np.random.default_rng(1000 + w.seed)
    rough = rng.uniform(0.1, 1.0, (w.height, w.width))
    kind = rng.uniform(0.0, 1.0, (w.height, w.width))
    metal = np.where(kind < 0.6, 0.0, np.where(kind < 0.8, 1.0, rng.uniform(0.0, 1.0, (w.height, w.width))))
    return np.stack([np.rint(rough * 255), np.rint(metal * 255)], axis=-1).astype(np.uint8)
