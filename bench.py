"""Benchmark of the indirect-lighting frame on one MI355X.

One step = one frame of the boxes_1080p workload through the HIP library:
voxelize (1M triangles, 128^3 + mips) -> allocate caches (1920x1080 G-buffer,
CAV 4 x 64^3) -> gather 16384 VPLs per cache into SH L2 with cone-traced
indirect shadows -> apply to every pixel.  Inputs (scene, G-buffer, RSM) are
produced once on the host, uploaded, and stay resident in HBM; the timed
region covers the whole frame.

value = cache.VPL interactions per second of whole frames
        (live caches x VPLs / frame time); ms_per_step = ms per frame.

Clock settle: the GPU raises its clocks over the first ~20 ms of
back-to-back work (boxes_1080p frames go from 0.52 to 0.47 ms over ~35
frames, and again after an idle second).  Each timed section therefore runs
its frame untimed for --settle-s seconds of wall time before its warm-up
steps, so the timed steps measure a renderer that draws continuously.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# MI355X published peaks.
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# Algorithmic work per unit (FMA = 2 FLOP; SFU not counted).
FLOP_SH = {1: 49, 2: 93}          # per cache.VPL interaction
FLOP_CONE_SETUP = 90              # per (cache, shadow block) trace
FLOP_CONE_STEP = 60               # per march step

# Online fine-tuning of the surrogate (every 16 frames).
FINETUNE_SAMPLES, FINETUNE_BATCH, FINETUNE_STEPS, FINETUNE_LR = 8192, 2048, 4, 1e-4


def _renderer(cfg, stream):
    """A renderer of the bench's static scenes: their geometry is never
    rewritten on the stream, so the next frame's voxelization may start right
    after the previous frame's cone trace (drv_set_voxel_overlap)."""
    from dynamicradiancevolume_amd.pipeline import Renderer
    r = Renderer(cfg, stream)
    r.set_voxel_overlap(True)
    return r


def _host_cpu() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "os_cpu_count": os.cpu_count(),
            "sched_affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def _cpu_baseline(fi, r, n_threads: int, budget_s: float):
    """The oracle's FP32 gather (VPL sum + cone tracing) on a bounded sample of
    the same frame's caches, timed on the host cores.  The oracle's other
    stages of the frame (voxelization, allocation, apply) are timed once
    beside it, so a whole-frame CPU time can be stated too."""
    import numpy as np

    from oracle import orc

    w = fi.workload
    cfg = w.config()
    t0 = time.perf_counter()
    vol = orc.update_volume(cfg, w.eye, fi.bbox[:3], fi.bbox[3:])
    _, mips = orc.voxelize(w.voxel_resolution, vol, fi.positions, adaption_factor=w.adaption_factor)
    t_vox = time.perf_counter() - t0
    t0 = time.perf_counter()
    count, cav, o_entries = orc.allocate(cfg, vol, fi.frame(), fi.depth)
    t_alloc = time.perf_counter() - t0
    _, live = r.cache_count()
    entries = r.read_caches(live)
    shc = orc.sh_constants()
    t0 = time.perf_counter()
    orc.apply(cfg, vol, fi.frame(), shc, fi, cav, o_entries, min(count, cfg.max_caches), n_threads=n_threads)
    t_apply = time.perf_counter() - t0
    n_vpl = fi.n_vpl
    # Calibrate on a few caches, then size the sample to the time budget.
    probe = np.linspace(0, live - 1, min(live, n_threads * 2)).astype(np.uint32)
    t0 = time.perf_counter()
    orc.gather_f32(cfg, vol, shc, fi.light, fi, mips, entries.copy(), ids=probe, n_threads=n_threads)
    per_cache = (time.perf_counter() - t0) / len(probe)
    n = int(max(len(probe), min(live, budget_s / max(per_cache, 1e-9))))
    ids = np.linspace(0, live - 1, n).astype(np.uint32)
    t0 = time.perf_counter()
    orc.gather_f32(cfg, vol, shc, fi.light, fi, mips, entries.copy(), ids=ids, n_threads=n_threads)
    dt = time.perf_counter() - t0
    gather_full_s = dt * live / n
    frame_s = t_vox + t_alloc + gather_full_s + t_apply
    return {
        "value": n * n_vpl / dt,
        "unit": "cache.VPL interactions/s",
        "cores": n_threads,
        "kind": "port",
        "stage": "gather only (VPL sum + cone-traced shadows); the GPU value is whole frames",
        "sample": f"{n} of {live} caches x {n_vpl} VPLs of one {w.name} frame (SH L{w.sh_bands}, cone-traced "
                  f"shadows), oracle FP32 gather, {dt:.1f} s on {n_threads} threads",
        "host": dict(_host_cpu(), threads_used=n_threads,
                     threads_note="16 threads: the host CPU share of one GPU on this machine"),
        "whole_frame": {"voxelize_s": round(t_vox, 3), "allocate_s": round(t_alloc, 3),
                        "gather_s_extrapolated": round(gather_full_s, 3), "apply_s": round(t_apply, 3),
                        "frame_s": round(frame_s, 3),
                        "interactions_per_s": live * n_vpl / frame_s},
    }


# SH-error statements against the FP64 oracle: the GPU results are captured
# during the run (cache positions and SH of a seeded subset, the device's
# voxel mips), and the oracle evaluates them on the host in the CPU leg at
# the end, as a checker.
_FP64_PENDING = []


def _fp64_capture(r, fi, n_sub: int, seed: int) -> dict:
    """Captures a seeded subset of the renderer's current caches (one light's
    gather) for the FP64 statement; returns the placeholder the CPU leg fills
    with the SH relative L2 against the FP64 oracle at the same positions,
    voxels (read back from the device, bit-exact with the oracle's, checked by
    the tests) and light."""
    import numpy as np

    from dynamicradiancevolume_amd.pipeline import entries_sh

    w = fi.workload
    _, live = r.cache_count()
    e = r.read_caches(live)
    rng = np.random.default_rng(seed)
    ids = np.sort(rng.choice(live, size=min(n_sub, live), replace=False))
    res = w.voxel_resolution
    mips = np.concatenate([r.read_voxels(lv).reshape(-1) for lv in range(int(np.log2(res)) + 1)])
    cfg = type(r.cfg).from_buffer_copy(r.cfg)
    placeholder = {"caches": int(len(ids)), "seed": seed, "sh_rel_l2_vs_fp64": None}
    _FP64_PENDING.append((placeholder, dict(cfg=cfg, vol=r.volume(), fi=fi, mips=mips, pos=e[ids, :3].copy(),
                                            sh=entries_sh(e[ids], cfg.sh_bands).astype(np.float64))))
    return placeholder


def _fp64_leg(n_threads: int) -> float:
    """Evaluates every captured subset with the oracle's FP64 gather."""
    import numpy as np

    from oracle import orc

    t0 = time.perf_counter()
    shc = orc.sh_constants()
    for placeholder, c in _FP64_PENDING:
        sh64, _ = orc.gather_f64(c["cfg"], c["vol"], shc, c["fi"].light, c["fi"], c["mips"], c["pos"],
                                 n_threads=n_threads)
        placeholder["sh_rel_l2_vs_fp64"] = float(np.linalg.norm(c["sh"] - sh64) / np.linalg.norm(sh64))
    _FP64_PENDING.clear()
    return time.perf_counter() - t0


def _traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed byte-counter
    summary (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload,
    scaled by a known-size copy); None when absent."""
    p = ROOT / "profiles" / "pmc_boxes_1080p.json"
    if not p.exists():
        return None
    k = json.loads(p.read_text()).get("kernels", {}).get(kernel)
    return k["hbm_bytes"] if k else None


SETTLE_S = 0.15


def _settle(fn, stream, seconds: float) -> int:
    """Runs `fn` back to back, untimed, for `seconds` of wall time (the GPU
    clock ramp); returns the number of calls."""
    t0 = time.perf_counter()
    n = 0
    while seconds > 0:
        fn()
        n += 1
        if n % 8 == 0:
            stream.synchronize()
            if time.perf_counter() - t0 >= seconds:
                break
    stream.synchronize()
    return n


def _time_frames(r, df, stream, steps: int, finetune_every: int = 0):
    """ms per frame over `steps` whole frames (HIP events on the library's
    stream), optionally with an online fine-tuning step every N frames."""
    import torch

    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    start.record(stream)
    for i in range(steps):
        r.render_frame(df)
        if finetune_every and (i + 1) % finetune_every == 0:
            r.mlp_finetune(FINETUNE_SAMPLES, FINETUNE_BATCH, FINETUNE_STEPS, FINETUNE_LR, 1000 + i)
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / steps, time.perf_counter() - t_wall


def _sh_by_position(r, bands: int):
    """{cache position bits: 27 SH coefficients} of the current cache list
    (the append order of the allocation differs between frames)."""
    import numpy as np

    from dynamicradiancevolume_amd.pipeline import entries_sh

    _, live = r.cache_count()
    e = r.read_caches(live)
    keys = [tuple(k) for k in e[:, :3].view(np.uint32).tolist()]
    return dict(zip(keys, entries_sh(e, bands).astype(np.float64)))


def _surrogate(r, df, fi, stream, w, steps: int, warmup: int, traced_ms: float, threshold: float, settle_s: float,
               n_threads: int):
    """The same frames with the surrogate visibility network (committed
    weights, 1 % spot check, fine-tuning every 16 frames)."""
    import numpy as np
    import torch

    from dynamicradiancevolume_amd import surrogate as S

    if not S.WEIGHTS.exists():
        return None
    ref = _sh_by_position(r, w.sh_bands)
    out = {"weights": S.WEIGHTS.name, "spot_fraction": 0.01, "finetune_every": 16}
    for label, thr in (("no_fallback", 1e30), ("fallback", threshold)):
        r.mlp_set_params(S.load_weights())
        r.set_surrogate(True, spot_fraction=0.01, fallback_threshold=thr)
        _settle(lambda: r.render_frame(df), stream, settle_s)
        for _ in range(warmup):
            r.render_frame(df)
        torch.cuda.synchronize()
        ms, _ = _time_frames(r, df, stream, steps, finetune_every=16)
        st = r.surrogate_stats()
        out[label] = {"ms_per_frame": ms, "speedup_vs_traced": traced_ms / ms, "threshold": thr,
                      "spot_error_last_frame": st["spot_error"], "fell_back_last_frame": st["fell_back"]}
    # SH error of one surrogate frame against the cone-traced frame
    r.set_surrogate(True, spot_fraction=0.01, fallback_threshold=1e30)
    r.mlp_set_params(S.load_weights())
    r.render_frame(df)
    sur = _sh_by_position(r, w.sh_bands)
    keys = [k for k in ref if k in sur]
    a = np.stack([sur[k] for k in keys])
    b = np.stack([ref[k] for k in keys])
    out["sh_rel_l2_vs_traced"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["fp64"] = _fp64_capture(r, fi, 1 << 30, 7)
    r.set_surrogate(False)
    return out


def _batch(n_views: int, steps: int, stream, n_threads: int, threshold: float, fp64_caches: int = 256,
           with_surrogate: bool = True):
    """The batched-views workload: `n_views` seeded views of the boxes family
    (720p to 4K, 4096 or 16384 VPLs, own scene each) rendered (a) view after
    view through the single-view path, (b) as one batch (autotuned work
    decomposition), (c) as a replayed HIP graph of the batched frame; then
    (d) the same batch with surrogate visibility (one network for every view,
    1 % spot check and fallback per view, fine-tuning every 16 frames).  The
    SH error of every view against the FP64 oracle is stated for (b) and (d)
    on a seeded subset of its caches."""
    import numpy as np
    import torch

    from dynamicradiancevolume_amd import surrogate as S
    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import Batch, DeviceFrame, FrameGraph

    views = W.batch_views(n_views, seed=0)
    frames = []
    for i, w in enumerate(views):
        frames.append(W.make_inputs(w, n_threads=n_threads))
        if (i + 1) % 8 == 0:
            print(f"[bench] batch inputs {i + 1}/{n_views}", file=sys.stderr, flush=True)
    rs = [_renderer(w.config(), stream) for w in views]
    for r in rs:
        r.set_tuning(gather_max_splits=32)
    dfs = [DeviceFrame(fi) for fi in frames]
    torch.cuda.synchronize()

    def timed(fn, k):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / k

    def per_view_fp64():
        return [_fp64_capture(r, fi, fp64_caches, 100 + i) for i, (r, fi) in enumerate(zip(rs, frames))]

    k = max(2, steps // 4)
    seq_ms = timed(lambda: [r.render_frame(df) for r, df in zip(rs, dfs)], k)
    lives = [r.cache_count()[1] for r in rs]
    interactions = sum(lv * df.inputs.n_vpl for lv, df in zip(lives, dfs))
    b = Batch(rs)
    b.render_frame(dfs)
    # Light pass of each decomposition: per-view passes, split queue with the
    # visibility queued over all views, split queue with per-view visibility
    # launches, fused queue (64 caches per wave, 1024-VPL tiles).
    by_mode = {}
    for label, mode in (("per_view_passes", 0), ("split_queue", -1), ("split_gather", -2), ("fused_queue", 64)):
        b.set_tuning(mode, 1024)
        by_mode[label] = timed(lambda: b.light_caches(dfs), k)
    cpw, tile, tune_ms = b.autotune(dfs, reps=2)
    batch_ms = timed(lambda: b.render_frame(dfs), k)
    traced_err = per_view_fp64()
    g = FrameGraph(rs[0])
    with torch.cuda.stream(stream):
        b.render_frame(dfs)
        stream.synchronize()
        g.begin()
        b.render_frame(dfs)
        g.end()
    graph_ms = timed(g.launch, k)
    print(f"[bench] batch of {n_views}: sequential {seq_ms:.2f} ms, batched {batch_ms:.2f} ms, graph {graph_ms:.2f} ms",
          file=sys.stderr, flush=True)
    out = {
        "views": n_views,
        "resolutions": sorted({f"{w.width}x{w.height}" for w in views}),
        "vpls_per_view": sorted({df.inputs.n_vpl for df in dfs}),
        "live_caches": [min(lives), max(lives), sum(lives)],
        "interactions_per_batch": interactions,
        "sequential_ms": seq_ms,
        "batched_ms": batch_ms,
        "graph_ms": graph_ms,
        "speedup_graph_vs_sequential": seq_ms / graph_ms,
        "interactions_per_s_graph": interactions / (graph_ms * 1e-3),
        "autotuned": {"caches_per_wave": cpw, "vpl_tile": tile, "light_pass_ms": tune_ms},
        "light_pass_ms_by_decomposition": by_mode,
        "frames_timed": k,
        "sh_rel_l2_vs_fp64_per_view": traced_err,
        "fp64_caches_per_view": fp64_caches,
    }
    sur = None
    if with_surrogate and S.WEIGHTS.exists():
        print(f"[bench] batch of {n_views} with surrogate visibility", file=sys.stderr, flush=True)
        sur = {"weights": S.WEIGHTS.name, "spot_fraction": 0.01, "finetune_every": 16, "views": n_views}
        n_frames = 16
        for label, thr in (("no_fallback", 1e30), ("fallback", threshold)):
            b.set_surrogate(True, S.load_weights(), spot_fraction=0.01, fallback_threshold=thr)
            b.render_frame(dfs)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for i in range(n_frames):
                b.render_frame(dfs)
                if (i + 1) % 16 == 0:
                    with torch.cuda.stream(stream):
                        b.mlp_finetune(8192, 2048, 4, 1e-4, 2000 + i)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1) / n_frames
            st = [r.surrogate_stats() for r in rs]
            sur[label] = {"ms_per_batch_frame": ms, "speedup_vs_batched_traced": batch_ms / ms, "threshold": thr,
                          "frames_timed": n_frames,
                          "views_fell_back_last_frame": int(sum(x["fell_back"] for x in st)),
                          "spot_error_last_frame": [round(x["spot_error"], 5) for x in st]}
        # one more frame without fallback: its per-view SH error against FP64
        b.set_surrogate(True, S.load_weights(), spot_fraction=0.01, fallback_threshold=1e30)
        b.render_frame(dfs)
        stream.synchronize()
        sur["sh_rel_l2_vs_fp64_per_view"] = per_view_fp64()
        b.set_surrogate(False)
    b.close()
    for r in rs:
        r.close()
    return out, sur


def _batch_dense(n_views: int, steps: int, stream, n_threads: int):
    """Batched views at the cache scale of the batched configuration: seeded
    boxes views at 1080p and 1440p with the 4 x 240^3 address volume (tens of
    thousands of caches per view), rendered view after view and as one batch
    (the autotuned decomposition), with each decomposition's light pass."""
    import torch

    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import Batch, DeviceFrame

    views = [w.with_(max_caches=262144) for w in
             W.batch_views(n_views, seed=2, family=W.BOXES_1080P_DENSE, resolutions=W.DENSE_BATCH_RESOLUTIONS)]
    frames = [W.make_inputs(w, n_threads=n_threads) for w in views]
    rs = [_renderer(w.config(), stream) for w in views]
    dfs = [DeviceFrame(fi) for fi in frames]
    torch.cuda.synchronize()

    def timed(fn, k):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / k

    k = max(2, steps // 4)
    seq_ms = timed(lambda: [r.render_frame(df) for r, df in zip(rs, dfs)], k)
    lives = [r.cache_count()[1] for r in rs]
    interactions = sum(lv * df.inputs.n_vpl for lv, df in zip(lives, dfs))
    b = Batch(rs)
    b.render_frame(dfs)
    by_mode = {}
    for label, mode in (("per_view_passes", 0), ("split_queue", -1), ("split_gather", -2), ("fused_queue", 64)):
        b.set_tuning(mode, 1024)
        by_mode[label] = timed(lambda: b.light_caches(dfs), k)
    cpw, tile, tune_ms = b.autotune(dfs, reps=2)
    batch_ms = timed(lambda: b.render_frame(dfs), k)
    print(f"[bench] dense batch of {n_views}: sequential {seq_ms:.2f} ms, batched {batch_ms:.2f} ms "
          f"({sum(lives)} caches)", file=sys.stderr, flush=True)
    out = {
        "views": n_views,
        "cav": "4x240^3",
        "resolutions": sorted({f"{w.width}x{w.height}" for w in views}),
        "vpls_per_view": sorted({df.inputs.n_vpl for df in dfs}),
        "live_caches": [min(lives), max(lives), sum(lives)],
        "interactions_per_batch": interactions,
        "sequential_ms": seq_ms,
        "batched_ms": batch_ms,
        "speedup_batched_vs_sequential": seq_ms / batch_ms,
        "interactions_per_s_batched": interactions / (batch_ms * 1e-3),
        "autotuned": {"caches_per_wave": cpw, "vpl_tile": tile, "light_pass_ms": tune_ms},
        "light_pass_ms_by_decomposition": by_mode,
        "frames_timed": k,
    }
    b.close()
    for r in rs:
        r.close()
    return out


def _batch_small(n_views: int, steps: int, stream, n_threads: int):
    """Many small views (320x180, 4096 VPLs each, own scene each): the light
    pass of the whole batch view by view (per-view passes) against the split
    queue (visibility queued over all views, or launched per view) and the
    fused queue, with the rest of each view's frame (voxelize, allocate,
    apply) identical in all of them."""
    import torch

    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import Batch, DeviceFrame, FrameGraph

    views = W.batch_views(n_views, seed=1, resolutions=W.SMALL_RESOLUTIONS, rsm=W.SMALL_RSM)
    frames = [W.make_inputs(w.with_(max_caches=16384), n_threads=n_threads) for w in views]
    rs = [_renderer(fi.workload.config(), stream) for fi in frames]
    for r in rs:
        r.set_tuning(gather_max_splits=32)
    dfs = [DeviceFrame(fi) for fi in frames]
    b = Batch(rs)
    b.render_frame(dfs)
    torch.cuda.synchronize()

    def timed(fn, k):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / k

    k = max(4, steps)
    out = {"views": n_views, "resolution": "320x180", "vpls_per_view": 4096}
    lives = [r.cache_count()[1] for r in rs]
    out["live_caches"] = [min(lives), max(lives), sum(lives)]
    for label, cpw in (("per_view_passes", 0), ("split_queue", -1), ("split_gather", -2), ("fused_queue", 64)):
        b.set_tuning(cpw, 1024)
        out[label + "_light_ms"] = timed(lambda: b.light_caches(dfs), k)
        out[label + "_frame_ms"] = timed(lambda: b.render_frame(dfs), k)
    out["split_vs_per_view_light"] = out["per_view_passes_light_ms"] / min(out["split_queue_light_ms"],
                                                                           out["split_gather_light_ms"])
    cpw, tile, ms = b.autotune(dfs, reps=2)
    out["autotuned"] = {"caches_per_wave": cpw, "vpl_tile": tile, "light_pass_ms": ms}
    # The eager frames above are issued view by view from Python (about ten
    # library calls per view); a replayed graph of the autotuned batch frame
    # issues the same work without that host time.
    g = FrameGraph(rs[0])
    with torch.cuda.stream(stream):
        b.render_frame(dfs)
        stream.synchronize()
        g.begin()
        b.render_frame(dfs)
        g.end()
    out["graph_frame_ms"] = timed(g.launch, k)
    out["interactions_per_s_graph"] = sum(lives) * 4096 / (out["graph_frame_ms"] * 1e-3)
    print(f"[bench] {n_views} small views: light pass per-view {out['per_view_passes_light_ms']:.2f} ms, "
          f"split queue {out['split_queue_light_ms']:.2f} ms, split gather {out['split_gather_light_ms']:.2f} ms, "
          f"fused {out['fused_queue_light_ms']:.2f} ms; graph frame {out['graph_frame_ms']:.2f} ms",
          file=sys.stderr, flush=True)
    b.close()
    for r in rs:
        r.close()
    return out


def _inputs_and_reference(fi, df, stream, w):
    """The gather's inputs rendered on the device (G-buffer, RSM + mips) and
    the brute-force RSM reference (unshadowed, diffuse) of the same frame,
    with the end-to-end error of the unshadowed cache path against it."""
    import numpy as np
    import torch


    r = _renderer(w.with_(indirect_shadow=0).config(), stream)
    scene = df.scene()

    def timed(fn, k=5):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / k

    gb_ms = timed(lambda: r.render_gbuffer(scene, fi.view_proj))
    rsm_ms = timed(lambda: r.render_rsm(scene, fi.light))
    r.render_frame(df)
    stream.synchronize()  # torch's own work below runs on its default stream
    cache_rgb = df.out.clone()
    bf = torch.zeros_like(cache_rgb)
    torch.cuda.synchronize()
    bf_ms = timed(lambda: r.bruteforce_rsm(df.gbuffer(), df.rsm(), fi.light, bf, True), k=2)
    stream.synchronize()
    lit = torch.from_numpy(fi.depth >= 1e-5).to(bf.device)
    a = cache_rgb[lit].double()
    b = bf[lit].double()
    err = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    n_px = int(np.count_nonzero(fi.depth >= 1e-5))
    r.close()
    return {
        "gbuffer_ms": gb_ms,
        "rsm_ms": rsm_ms,
        "triangles": int(len(fi.positions)),
        "bruteforce_ms": bf_ms,
        "bruteforce_pixel_val_interactions_per_s": w.width * w.height * fi.n_vpl / (bf_ms * 1e-3),
        "unshadowed_cache_path_vs_bruteforce_rel_l2": err,
        "lit_pixels": n_px,
    }


def _specular(fi, df, stream, w, steps: int, warmup: int, n_vpl: int, settle_s: float):
    """The same frame with INDIRECT_SPECULAR (16 x 16 env-map per cache, the
    reference defaults): frame time, the env-map stage, and the end-to-end
    error of the unshadowed cache path (diffuse SH + specular env-maps)
    against the brute-force RSM with the full BRDF."""
    import numpy as np
    import torch


    ws = w.with_(specular=1)
    r = _renderer(ws.config(), stream)
    _settle(lambda: r.render_frame(df), stream, settle_s)
    for _ in range(warmup):
        r.render_frame(df)
    ms, _ = _time_frames(r, df, stream, steps)
    r.set_profiling(1)
    keys = ("light_caches_ms", "gather_ms", "specular_ms", "apply_ms")
    acc = {k: 0.0 for k in keys}
    n_prof = max(3, min(10, steps))
    for _ in range(n_prof):
        for _ in range(8):  # back to back, as in the timed region; the last frame's stage times
            r.render_frame(df)
        t = r.stage_times()
        for k in keys:
            acc[k] += t[k]
    r.set_profiling(0)
    stages = {k: round(v / n_prof, 4) for k, v in acc.items()}
    _, live = r.cache_count()
    width, per_row, levels = r.specular_atlas_geometry()
    r.close()

    ru = _renderer(ws.with_(indirect_shadow=0).config(), stream)
    ru.render_frame(df)
    stream.synchronize()
    cache_rgb = df.out.clone()
    bf = torch.zeros_like(cache_rgb)
    torch.cuda.synchronize()
    ru.bruteforce_rsm(df.gbuffer(), df.rsm(), fi.light, bf, False, roughness_metallic=df.groughmetal)
    stream.synchronize()
    lit = torch.from_numpy(fi.depth >= 1e-5).to(bf.device)
    a, b = cache_rgb[lit].double(), bf[lit].double()
    err = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    ru.close()
    # the same unshadowed frame without the specular term, against the same reference
    rd = _renderer(w.with_(indirect_shadow=0).config(), stream)
    rd.render_frame(df)
    stream.synchronize()
    d = df.out[lit].double()
    err_diffuse = float(torch.linalg.norm(d - b) / torch.linalg.norm(b))
    rd.close()
    return {
        "envmap_size": ws.specular_size,
        "hole_fill_level": ws.hole_fill,
        "atlas": {"width": width, "caches_per_row": per_row, "levels": levels},
        "ms_per_frame": ms,
        "stage_ms": stages,
        "envmap_interactions_per_s": live * n_vpl / (stages["specular_ms"] * 1e-3) if stages["specular_ms"] > 0
        else None,
        "roughness_metallic": "per pixel, roughness U[0.1, 1], metallic 0 / 1 / U[0, 1] (60/20/20 %)",
        "unshadowed_cache_path_vs_bruteforce_full_brdf_rel_l2": err,
        "unshadowed_diffuse_only_cache_path_vs_bruteforce_full_brdf_rel_l2": err_diffuse,
        "lit_pixels": int(np.count_nonzero(fi.depth >= 1e-5)),
    }


def _frame_section(w, stream, steps: int, warmup: int, settle_s: float, n_threads: int):
    """Whole frames of another single-view workload: ms/frame, interactions/s,
    stage times, the k_visibility roofline and the FP64 SH error."""
    import torch

    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import DeviceFrame

    fi = W.make_inputs(w, n_threads=n_threads)
    df = DeviceFrame(fi)
    r = _renderer(w.config(), stream)
    _settle(lambda: r.render_frame(df), stream, settle_s)
    for _ in range(warmup):
        r.render_frame(df)
    torch.cuda.synchronize()
    ms, _ = _time_frames(r, df, stream, steps)
    _, live = r.cache_count()
    n_vpl = fi.n_vpl
    stages, traces, steps_, steps_exec = _stage_profile(r, df, steps)
    vis_flop = FLOP_CONE_SETUP * traces + FLOP_CONE_STEP * steps_
    vis_exec = FLOP_CONE_SETUP * traces + FLOP_CONE_STEP * steps_exec
    vms = stages["visibility_ms"]
    gms = stages["gather_ms"]
    out = {
        "workload": w.name,
        "cav": f"{w.cav_cascades}x{w.cav_resolution}^3",
        "live_caches": live,
        "interactions_per_frame": live * n_vpl,
        "ms_per_frame": ms,
        "interactions_per_s": live * n_vpl / (ms * 1e-3),
        "stage_ms": stages,
        "cone_traces": traces,
        "cone_steps": steps_,
        "cone_steps_executed": steps_exec,
        "k_visibility_tflops": vis_flop / (vms * 1e-3) / 1e12 if vms > 0 else None,
        "k_visibility_frac": vis_flop / (vms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS if vms > 0 else None,
        "k_visibility_frac_executed": vis_exec / (vms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS if vms > 0 else None,
        "k_gather_frac": FLOP_SH[w.sh_bands] * live * n_vpl / (gms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS
        if gms > 0 else None,
        "fp64": _fp64_capture(r, fi, 4096, 11),
    }
    r.close()
    return out


def _stage_profile(r, df, steps: int, frames_per_round: int = 8):
    """Per-stage ms (HIP events inside the library) and the march-step
    counts of one more frame with counting on.  Each round renders frames
    back to back, as the timed region does, and reads the stage times of the
    last one (the library keeps the latest frame's events): a frame read
    right after a host synchronization would follow an idle GPU."""
    r.set_profiling(1)
    keys = ("voxelize_ms", "voxel_mips_ms", "allocate_ms", "light_caches_ms", "prepare_ms", "visibility_ms",
            "gather_ms", "reduce_ms", "apply_ms")
    acc = {k: 0.0 for k in keys}
    n_prof = max(3, min(10, steps))
    for _ in range(n_prof):
        for _ in range(frames_per_round):
            r.render_frame(df)
        t = r.stage_times()
        for k in keys:
            acc[k] += t[k]
    stages = {k: round(v / n_prof, 4) for k, v in acc.items()}
    r.set_profiling(2)
    r.render_frame(df)
    t = r.stage_times()
    r.set_profiling(0)
    return stages, t["cone_traces"], t["cone_steps"], t["cone_steps_executed"]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--settle-s", type=float, default=SETTLE_S,
                    help="seconds of untimed frames before each section's warm-up (GPU clock ramp)")
    ap.add_argument("--workload", default="boxes_1080p")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-specular", action="store_true")
    ap.add_argument("--batch-views", type=int, default=64, help="views of the batched workload (0 = skip)")
    ap.add_argument("--small-views", type=int, default=64, help="views of the small-view batch (0 = skip)")
    ap.add_argument("--dense-views", type=int, default=8,
                    help="views of the batch with 4 x 240^3 address volumes (0 = skip)")
    ap.add_argument("--dense", default="boxes_1080p_dense",
                    help="second single-view workload at the ~60k-cache scale ('' = skip)")
    ap.add_argument("--sh-threshold", type=float, default=2e-2, help="surrogate spot-check fallback threshold")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the path is not sharded across "
              "GPUs here", file=sys.stderr)
        return 2

    import torch

    from dynamicradiancevolume_amd import workloads as W
    from dynamicradiancevolume_amd.pipeline import DeviceFrame

    w = W.WORKLOADS[args.workload]
    print(f"[bench] building inputs for {w.name}", file=sys.stderr, flush=True)
    t_in = time.perf_counter()
    fi = W.make_inputs(w, n_threads=min(16, os.cpu_count() or 1))
    host_inputs_s = time.perf_counter() - t_in
    df = DeviceFrame(fi)
    stream = torch.cuda.Stream()
    r = _renderer(w.config(), stream)
    torch.cuda.synchronize()
    n_vpl = fi.n_vpl

    settled = _settle(lambda: r.render_frame(df), stream, args.settle_s)
    print(f"[bench] clock settle {settled} frames, warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        r.render_frame(df)
    torch.cuda.synchronize()

    ms, wall = _time_frames(r, df, stream, args.steps)
    demanded, live = r.cache_count()
    # The same frame captured once into a HIP graph and replayed (no host
    # launch overhead; a static camera, or a graph updated per frame).
    from dynamicradiancevolume_amd.pipeline import FrameGraph

    g = FrameGraph(r)
    with torch.cuda.stream(stream):
        g.begin()
        r.render_frame(df)
        g.end()
        g.launch()
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            g.launch()
        e1.record(stream)
        e1.synchronize()
    graph_ms = e0.elapsed_time(e1) / args.steps
    interactions = live * n_vpl
    print(f"[bench] timed {args.steps} frames: {ms:.3f} ms/frame (wall {wall:.2f} s), caches {live}",
          file=sys.stderr, flush=True)

    # Per-kernel durations: HIP events on the library's stream, averaged over
    # a profiled pass of the same frames; the march-step counts come from one
    # more frame with step counting on (which slows the trace, so it is not
    # timed).
    stages, traces, steps_, steps_exec = _stage_profile(r, df, args.steps)
    n_threads = min(16, os.cpu_count() or 1)
    plain_fp64 = _fp64_capture(r, fi, 1 << 30, 7)

    vis_flop = FLOP_CONE_SETUP * traces + FLOP_CONE_STEP * steps_
    vis_flop_exec = FLOP_CONE_SETUP * traces + FLOP_CONE_STEP * steps_exec
    gather_flop = FLOP_SH[w.sh_bands] * interactions
    gather_name = f"k_gather<{'true' if w.sh_bands >= 2 else 'false'}, {'true' if w.indirect_shadow else 'false'}>"
    kernels = {
        "k_visibility": (stages["visibility_ms"], vis_flop, vis_flop_exec),
        gather_name: (stages["gather_ms"], gather_flop, gather_flop),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    dms, dflop, dflop_exec = kernels[dom]
    achieved = dflop / (dms * 1e-3) / 1e12 if dms > 0 else 0.0
    achieved_exec = dflop_exec / (dms * 1e-3) / 1e12 if dms > 0 else 0.0
    roofline = {
        "bound": "valu",
        "pipe": "fp32 vector ALU (VALU work and per-step load latency; the FP32 vector peak is the roofline)",
        "frac_executed_steps": round(achieved_exec / PEAK_FP32_TFLOPS, 4),
        "frac_note": "frac counts 60 FLOP for every march step the reference takes; frac_executed_steps counts "
                     "only the steps sampled before a ray turns opaque",
        "kernel": dom,
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "traffic": _traffic(dom) if w.name == "boxes_1080p" else None,
        "traffic_source": "profiles/hbm_bytes_boxes_1080p.md (rocprofv3 --pmc, copy-calibrated)",
        "kernel_ms": dms,
        "algorithmic_flop": dflop,
    }

    sur = None
    if not args.no_surrogate:
        print("[bench] surrogate", file=sys.stderr, flush=True)
        sur = _surrogate(r, df, fi, stream, w, args.steps, args.warmup, ms, args.sh_threshold, args.settle_s,
                         n_threads)

    dense = None
    if args.dense:
        print("[bench] dense address volume (~60k caches)", file=sys.stderr, flush=True)
        dense = _frame_section(W.WORKLOADS[args.dense], stream, args.steps, args.warmup, args.settle_s, n_threads)

    spec = None
    if not args.no_specular:
        print("[bench] indirect specular", file=sys.stderr, flush=True)
        spec = _specular(fi, df, stream, w, args.steps, args.warmup, n_vpl, args.settle_s)

    print("[bench] device inputs and brute-force reference", file=sys.stderr, flush=True)
    ref = _inputs_and_reference(fi, df, stream, w)
    ref["host_producer_s"] = host_inputs_s

    batch = batch_sur = None
    if args.batch_views > 0:
        print(f"[bench] batch of {args.batch_views} views", file=sys.stderr, flush=True)
        batch, batch_sur = _batch(args.batch_views, args.steps, stream, n_threads, args.sh_threshold,
                                  with_surrogate=not args.no_surrogate)

    batch_small = None
    if args.small_views > 0:
        print(f"[bench] batch of {args.small_views} small views", file=sys.stderr, flush=True)
        batch_small = _batch_small(args.small_views, args.steps, stream, n_threads)

    batch_dense = None
    if args.dense_views > 0:
        print(f"[bench] batch of {args.dense_views} views with 4 x 240^3 address volumes", file=sys.stderr,
              flush=True)
        batch_dense = _batch_dense(args.dense_views, args.steps, stream, n_threads)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline and FP64 error statements (oracle on the host)", file=sys.stderr, flush=True)
        r.render_frame(df)
        stream.synchronize()
        cpu = _cpu_baseline(fi, r, n_threads=n_threads, budget_s=args.cpu_budget)
        cpu["fp64_statements_s"] = round(_fp64_leg(n_threads), 2)
        for sec in (batch, batch_sur):
            if sec and "sh_rel_l2_vs_fp64_per_view" in sec:
                errs = [p["sh_rel_l2_vs_fp64"] for p in sec["sh_rel_l2_vs_fp64_per_view"]]
                sec["sh_rel_l2_vs_fp64_per_view"] = [round(e, 7) for e in errs]
                sec["sh_rel_l2_vs_fp64_max"] = max(errs)
                sec["sh_rel_l2_vs_fp64_median"] = sorted(errs)[len(errs) // 2]
    _FP64_PENDING.clear()

    line = {
        "metric": "cache.VPL interactions/s (whole frame, 1080p, 16k VPLs, cone-traced)",
        "value": interactions / (ms * 1e-3),
        "unit": "cache.VPL interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, CPU-rasterized G-buffer/RSM)",
        "config": {
            "workload": w.name,
            "resolution": [w.width, w.height],
            "vpls": n_vpl,
            "live_caches": live,
            "interactions_per_frame": interactions,
            "cav": f"{w.cav_cascades}x{w.cav_resolution}^3",
            "voxels": w.voxel_resolution,
            "sh_bands": w.sh_bands,
            "shadow_lod": w.light.indirect_shadow_lod,
            "triangles": int(len(fi.positions)),
            "cone_traces": traces,
            "cone_steps": steps_,
            "cone_steps_executed": steps_exec,
            "stage_ms": stages,
            "path": "plain cone-traced",
            "fp64": plain_fp64,
            "graph_replay_ms_per_frame": graph_ms,
            "clock_settle_s": args.settle_s,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "surrogate": sur,
        "dense": dense,
        "specular": spec,
        "batch": batch,
        "batch_surrogate": batch_sur,
        "batch_small": batch_small,
        "batch_dense": batch_dense,
        "inputs_and_reference": ref,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
