#!/usr/bin/env python3
"""Benchmark of the plain wavefront path tracer on one MI355X.

Workload (the real-shape configuration): the instanced scene - 300 instances of five seeded
meshes (icospheres L2/L4/L6, a 512x512 noise terrain, a 200k-triangle soup; ~2.95M
triangles, per-instance materials, 20 emissive spheres) - at 1920x1080, max depth 8, fp32.

A step is one 64-spp frame at 1920x1080 (pt_render(64): 64 sample passes, each traced
exactly as the reference's 1-spp submits, src/pathtrace/ptdispatch.cpp:101, as the
reference's headless `-s 64` run renders it). Inputs are resident in HBM before the timed
region (scene, BVH, camera); the timed region is K back-to-back pt_render(64) calls
bracketed by stream synchronisation.

Other workloads (--workload views | scenes | cache): 64 views of the instanced scene (or, with
--scene box, the box scene) in one persistent wavefront run; 16 instanced scenes of 10k-20M
triangles with 4 views each in one run (pt_batch_*; a step is one render of every view); the
learned radiance cache on a 64-frame camera path at 1-4 spp per frame (a step is the path).

The JSON line of the default workload also carries:
  roofline      the dominant stage (k_extend, closest-hit traversal) against the bound that
                holds for it, VALU instruction issue: its VALU wave-instructions per launch
                (the per-ray count of the committed rocprofv3 summary x the launch's rays) over
                the launch's HIP-event mean time, against 1024 SIMDs x 2.4 GHz / 4 cycles.
                Beside it: the measured HBM traffic per launch (rocprofv3 FETCH_SIZE +
                WRITE_SIZE, calibrated on a copy of known size) against 8 TB/s, and the demand
                bytes of the per-ray model (32 ray + 20 hit + 128 per node line + 48 per
                triangle record, fetch counts from an instrumented replay of the same sample
                indices), which the caches serve and which therefore exceed the HBM peak.
  convergence   per scene (instanced and box, 1920x1080, depth 8): seconds of rendering until
                the relative MSE against a 4096-spp plain reference of disjoint sample indices
                (samples 4096..8191) is <= 1e-3, for the plain estimator (16-spp render calls)
                and the learned-cache estimator (16-spp render calls, the network trained in
                flight between wavefront iterations, paths ended at their second vertex with its
                prediction; a 16-spp warm-up counts towards its time). Error evaluation is
                outside the clock.
  cpu_baseline  the CPU oracle (this project's C++ restatement of the reference shaders,
                OpenMP) timed on whole 1-spp frames of the same workload for ~12 s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import pt_amd  # noqa: E402
from pt_amd.scenes import generate  # noqa: E402
from pt_amd.scenes import workloads as wl  # noqa: E402

W, H, DEPTH = 1920, 1080, 8
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, cam, threads, target_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob  # the checker, timed here as the CPU baseline only

    t = time.time()
    orc = ob.Oracle(scene)
    log("cpu baseline: oracle BVH built in %.1f s" % (time.time() - t))
    acc = np.zeros((H, W, 4), np.float32)
    # probe with every 8th row, then trace whole 1-spp frames for about target_s seconds
    t = time.perf_counter()
    orc.render(cam, W, H, 1, DEPTH, accum=acc, rows_step=8, threads=threads)
    frame_s = (time.perf_counter() - t) * 8
    frames = int(max(1, min(64, round(target_s / max(frame_s, 1e-3)))))
    rays = paths = 0
    t = time.perf_counter()
    for k in range(frames):
        _, c = orc.render(cam, W, H, 1, DEPTH, sample_index=k, accum=acc, threads=threads)
        rays += c["extension_rays"] + c["shadow_rays"]
        paths += c["paths"]
    dt = time.perf_counter() - t
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "%d full 1920x1080 frames at 1 spp, depth 8 (%d paths, %d rays) in %.1f s"
                      % (frames, paths, rays, dt)}


def reference_image(tr, cam, spp=4096, chunk=512):
    """The plain estimator's mean over sample indices [spp, 2 spp) of every pixel: 2 A(2 spp) -
    A(spp) of the running means after spp and 2 spp samples. The timed runs use indices from 0,
    so the reference shares no sample with them."""
    tr.set_camera(cam)
    for _ in range(spp // chunk):
        tr.render(chunk)
    a1 = tr.read()[..., :3].astype(np.float64)
    for _ in range(spp // chunk):
        tr.render(chunk)
    a2 = tr.read()[..., :3].astype(np.float64)
    out = np.zeros(a1.shape[:2] + (4,), np.float32)
    out[..., :3] = 2.0 * a2 - a1
    out[..., 3] = 1.0
    return out


def time_to_threshold(tr, cam, ref, chunk, limit_spp, limit_s, threshold, before=None):
    """Renders chunk-spp batches until the relative MSE against ref is <= threshold; returns
    (seconds of rendering incl. before(), spp, final rel-MSE, trajectory). Readback and the
    error evaluation are outside the clock."""
    tr.set_camera(cam)
    tr.synchronize()
    el = 0.0
    if before is not None:
        t = time.perf_counter()
        before()
        tr.synchronize()
        el += time.perf_counter() - t
        tr.set_camera(cam)  # restart the accumulation
    done, e, traj = 0, float("inf"), []
    while done < limit_spp and el < limit_s:
        t = time.perf_counter()
        tr.render(chunk)
        tr.synchronize()
        el += time.perf_counter() - t
        done += chunk
        e = pt_amd.rel_mse(tr.read(), ref)
        traj.append((done, round(el, 4), float("%.4e" % e)))
        if e <= threshold:
            break
    return el, done, e, traj


# The learned-cache estimator of the convergence runs: in-flight training (three Adam steps of
# 8192 records after every wavefront iteration), one path in 16 trains; the others end at their
# second vertex with the prediction of the weights' moving average (decay 0.9), unless they
# arrived there by a direction sampled with a pdf above 2 (a sharp glossy lobe, which goes on).
# Learning rate 3e-3: the best of a sweep over 1e-3..1e-2, 2..4 steps, batches of 8K..16K
# records, strides 4..16, continuation probabilities 0..0.75, decays 0..0.999 and pdf limits
# 0..5 on the instanced scene. The same estimator queried at the fourth vertex is timed beside
# it: a later vertex leaves less of each path to the cache's error, which pays where the plain
# tracer converges fast (the box).
CACHE_CONVERGENCE = dict(query=True, train_stride=16, train_steps=3, batch=8192, learning_rate=3e-3, frame_spp=0,
                         continue_prob=0.0, seed=1, query_vertex=2, ema_decay=0.9, query_max_pdf=2.0)
CACHE_LATER_VERTEX = 4


def convergence(tr, cam, threshold=1e-3, ref_spp=4096):
    """Seconds to rel-MSE <= threshold against a ref_spp plain reference of disjoint sample
    indices, plain vs cached."""
    t = time.perf_counter()
    ref = reference_image(tr, cam, ref_spp)
    ref_s = time.perf_counter() - t
    log("convergence: %d-spp reference (samples %d..%d) in %.1f s" % (ref_spp, ref_spp, 2 * ref_spp - 1, ref_s))
    plain_s, plain_spp, plain_e, plain_traj = time_to_threshold(tr, cam, ref, 16, 4096, 60.0, threshold)
    log("convergence: plain %.3f s, %d spp, rel-MSE %.3e" % (plain_s, plain_spp, plain_e))
    warm = 16
    runs = []
    for vertex in (CACHE_CONVERGENCE["query_vertex"], CACHE_LATER_VERTEX):
        cfg = dict(CACHE_CONVERGENCE, query_vertex=vertex)
        tr.cache_configure(**cfg)
        cached_s, cached_spp, cached_e, cached_traj = time_to_threshold(
            tr, cam, ref, 16, 4096, max(30.0, 3 * plain_s), threshold, before=lambda: tr.render(warm))
        cs = tr.cache_stats()
        tr.cache_configure(enable=False)
        log("convergence: cached (vertex %d) %.3f s, %d spp, rel-MSE %.3e" % (vertex, cached_s, cached_spp, cached_e))
        runs.append({"cached_s": round(cached_s, 4), "cached_spp": cached_spp, "cached_relmse": cached_e,
                     "cached_over_plain_time": round(cached_s / plain_s, 3) if plain_s > 0 else None,
                     "cache_config": dict(cfg, warmup_spp=warm, steps_done=cs["steps"]),
                     "cached_trajectory": cached_traj[-4:]})
    out = {"threshold": threshold, "reference_spp": ref_spp, "reference_samples": [ref_spp, 2 * ref_spp - 1],
           "reference_s": round(ref_s, 3),
           "plain_s": round(plain_s, 4), "plain_spp": plain_spp, "plain_relmse": plain_e,
           "plain_trajectory": plain_traj[-4:]}
    out.update(runs[0])
    out["later_vertex"] = runs[1]
    return out


def profile_figure(name, kernel):
    """A figure of a kernel from the committed rocprofv3 summaries (machine-readable lines
    `name kernel value` in profiles/*.txt and *.md), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        if not fn.endswith((".txt", ".md")):
            continue
        for line in open(os.path.join(d, fn)):
            f = line.split()
            if len(f) == 3 and f[0] == name and f[1] == kernel:
                return float(f[2])
    return None


PEAK_VALU_GINST = 1024 * 2.4 / 4.0  # 1024 SIMDs x 2.4 GHz, a wave64 VALU instruction per 4 cycles


def stage_roofline(kernel, d):
    """The roofline of a traversal stage by the bound that holds. Its rays visit BVH node lines
    that the caches serve (L1/L2/MALL), so its demand bytes can exceed the HBM peak while the
    measured HBM traffic is a few percent of it; what limits it is VALU instruction issue along
    each wave's dependent node chain. achieved = VALU wave-instructions per launch (per-ray count
    of the committed rocprofv3 summary x the launch's rays) / the launch's HIP-event mean time,
    against 1024 SIMDs x 2.4 GHz / 4 cycles; the HBM figures and the demand bytes are beside it."""
    launches = max(1, d["launches"])
    avg_s = d["ms"] / 1e3 / launches
    rays = d["rays"] / launches
    inst_per_ray = profile_figure("valu_inst_per_ray", kernel)
    hbm_per_ray = profile_figure("hbm_bytes_per_ray", kernel)
    achieved = inst_per_ray * rays / avg_s / 1e9 if inst_per_ray and avg_s > 0 else None
    traffic = hbm_per_ray * rays if hbm_per_ray else None
    demand = d["bytes"] / launches
    return {"bound": "valu", "achieved": round(achieved, 2) if achieved else None, "peak": PEAK_VALU_GINST,
            "unit": "G wave-instructions/s", "frac": round(achieved / PEAK_VALU_GINST, 4) if achieved else None,
            "traffic": round(traffic) if traffic else None,
            "kernel": kernel, "launches": d["launches"], "avg_launch_ms": round(avg_s * 1e3, 4),
            "rays_per_launch": round(rays), "valu_inst_per_ray": inst_per_ray,
            "lane_utilisation_pct": profile_figure("valu_lane_util_pct", kernel),
            "hbm": {"achieved": round(traffic / avg_s / 1e9, 1) if traffic else None, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None},
            "demand_bytes_per_launch": round(demand),
            "demand_gbs_cache_served": round(demand / avg_s / 1e9, 1) if avg_s > 0 else None,
            "l2_hit_pct": profile_figure("l2_hit_pct", kernel)}


def run_instanced(a):
    t0 = time.time()
    scn = generate.cached_scene("instanced", root=a.scene_dir)
    scene = pt_amd.Scene(scn)
    info = scene.info()
    cam = scene.camera(W, H)
    log("scene: %d instances, %d triangles (%.1f s)" % (info["instances"], info["total_triangles"], time.time() - t0))
    tr = pt_amd.Tracer(scene, W, H, max_bounces=DEPTH)
    tr.set_camera(cam)
    st0 = tr.stats()
    log("BVH: %.1f ms build, %d nodes, %.1f MB" % (st0["bvh_build_ms"], st0["node_count"], st0["bvh_bytes"] / 1e6))

    # launch configuration picked by the autotuner (untimed set-up, like the BVH build),
    # starting from any configuration given on the command line
    if a.leaf_size:
        tr.set_option(pt_amd.PT_OPT_LEAF_SIZE, a.leaf_size)
    if a.blocks_per_cu:
        tr.set_option(pt_amd.PT_OPT_BLOCKS_PER_CU, a.blocks_per_cu)
    if a.pool_slots:
        tr.set_option(pt_amd.PT_OPT_POOL_SLOTS, a.pool_slots)
    if a.tlas_open >= 0:
        tr.set_option(pt_amd.PT_OPT_TLAS_OPEN, a.tlas_open)
    if a.node_width:
        tr.set_option(pt_amd.PT_OPT_NODE_WIDTH, a.node_width)
    if a.lds_stack:
        tr.set_option(pt_amd.PT_OPT_LDS_STACK, a.lds_stack)
    st0 = tr.stats()
    tune = None
    if not a.no_autotune:
        tr.render(1)
        # trials of 16 spp (four pools of paths at 1080p): long enough to time the streaming
        # steady state the timed 64-spp steps run in, not one pool's fill and drain
        tune = tr.autotune(16)
        log("autotune: %d trials, blocks/CU %d, LDS stack %d, leaf %d, pool %d, TLAS opening %d: %.3f -> %.3f ns/ray"
            % (tune["trials"], tune["blocks_per_cu"], tune["lds_stack"], tune["leaf_size"], tune["pool_slots"],
               tune["tlas_open"], tune["ns_per_ray_start"], tune["ns_per_ray_best"]))
        st0 = tr.stats()

    tr.set_option(pt_amd.PT_OPT_SORT_HITS, a.sort_hits)

    # warm-up
    for _ in range(max(1, a.warmup)):
        tr.render(a.spp_per_step)
    tr.synchronize()

    # instrumented replay of the first steps' sample indices: fetch counts per ray
    tr.set_option(pt_amd.PT_OPT_COUNT_FETCHES, 1)
    tr.set_camera(cam)
    tr.render(a.spp_per_step)
    cs = tr.stats()
    tr.set_option(pt_amd.PT_OPT_COUNT_FETCHES, 0)
    n_node_ext = cs["ext_node_fetches"] / max(1, cs["extension_rays"])
    n_tri_ext = cs["ext_tri_fetches"] / max(1, cs["extension_rays"])
    n_node_sh = cs["shadow_node_fetches"] / max(1, cs["shadow_rays"])
    n_tri_sh = cs["shadow_tri_fetches"] / max(1, cs["shadow_rays"])
    per_ray = lambda k, n: round(cs[k] / max(1, cs[n]), 3)  # noqa: E731
    topo = {"ext_tlas_nodes": per_ray("ext_tlas_node_fetches", "extension_rays"),
            "ext_instances_entered": per_ray("ext_instance_entries", "extension_rays"),
            "shadow_tlas_nodes": per_ray("shadow_tlas_node_fetches", "shadow_rays"),
            "shadow_instances_entered": per_ray("shadow_instance_entries", "shadow_rays")}
    log("per ray: %s" % topo)
    bytes_ext = 32 + 20 + 128 * n_node_ext + 48 * n_tri_ext
    bytes_sh = 32 + 4 + 128 * n_node_sh + 48 * n_tri_sh
    log("fetches/ray: extension %.1f nodes %.1f tris (%.0f B), shadow %.1f nodes %.1f tris (%.0f B)"
        % (n_node_ext, n_tri_ext, bytes_ext, n_node_sh, n_tri_sh, bytes_sh))

    # timed region: exactly K steps, production settings (iterations replayed from HIP graphs)
    tr.set_camera(cam)
    tr.synchronize()
    t_start = time.perf_counter()
    for _ in range(a.steps):
        tr.render(a.spp_per_step)
    tr.synchronize()
    elapsed = time.perf_counter() - t_start
    st = tr.stats()
    rays = st["extension_rays"] + st["shadow_rays"]
    value = rays / elapsed / 1e6
    log("timed: %d steps in %.3f s, %d ext + %d shadow rays" % (a.steps, elapsed, st["extension_rays"],
                                                               st["shadow_rays"]))

    # the same steps again with every stage launch timed by HIP events (direct launches)
    tr.set_option(pt_amd.PT_OPT_PROFILE_KERNELS, 1)
    tr.set_camera(cam)
    for _ in range(a.steps):
        tr.render(a.spp_per_step)
    sp = tr.stats()
    tr.set_option(pt_amd.PT_OPT_PROFILE_KERNELS, 0)
    log("profiled pass: stages extend %.1f ms, shade %.1f ms, shadow %.1f ms, refill %.1f ms (render %.1f ms)"
        % (sp["extend_ms"], sp["shade_ms"], sp["shadow_ms"], sp["refill_ms"], sp["render_ms"]))

    # Demand bytes (the cache-independent per-ray model): extension ray 32 in + 20 out + 128 per
    # node line + 48 per triangle record; shadow ray 32 + 4 + the same. The 8-bounce shade stage
    # reads per hit: hit 20 + path state 32 + instance 128 + material 96 + indices 12 + 3 vertices
    # 96 + light bins 64 + emitter vertices 36, writes next ray 36 + path state 32 + shadow
    # record 64.
    stages = {
        "extend": {"ms": sp["extend_ms"], "launches": sp["extend_launches"], "rays": sp["extension_rays"],
                   "bytes": sp["extension_rays"] * bytes_ext},
        "shadow": {"ms": sp["shadow_ms"], "launches": sp["shadow_launches"], "rays": sp["shadow_rays"],
                   "bytes": sp["shadow_rays"] * bytes_sh},
        "shade": {"ms": sp["shade_ms"], "launches": sp["shade_launches"], "rays": sp["extension_rays"],
                  "bytes": sp["extension_rays"] * 616},
    }
    dom = max(stages, key=lambda k: stages[k]["ms"])
    roofline = stage_roofline("k_" + dom, stages[dom])

    conv = None
    if not a.no_convergence:
        conv = {"instanced": convergence(tr, cam)}
        tr.close()
        box = pt_amd.Scene(generate.cached_scene("box", root=a.scene_dir, seed=1))
        tb = pt_amd.Tracer(box, W, H, max_bounces=DEPTH)
        conv["box"] = convergence(tb, box.camera(W, H))
        tb.close()

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(scene, cam, a.cpu_threads)

    return {
        "metric": "Mrays/s", "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "instanced scene, 300 instances / %d triangles, 1920x1080, %d spp per step, depth %d"
                               % (info["total_triangles"], a.spp_per_step, DEPTH),
                   "spp_total": a.steps * a.spp_per_step, "width": W, "height": H, "max_depth": DEPTH},
        "samples_per_s": round(a.steps * a.spp_per_step / elapsed, 3),
        "pixel_samples_per_s": round(a.steps * a.spp_per_step * W * H / elapsed, 1),
        "rays": {"extension": st["extension_rays"], "shadow": st["shadow_rays"]},
        "fetches_per_ray": dict({"ext_nodes": round(n_node_ext, 3), "ext_tris": round(n_tri_ext, 3),
                                 "shadow_nodes": round(n_node_sh, 3), "shadow_tris": round(n_tri_sh, 3)}, **topo),
        "stages_ms": dict({k: round(v["ms"], 3) for k, v in stages.items()}, refill=round(sp["refill_ms"], 3)),
        "stage_launches": dict({k: v["launches"] for k, v in stages.items()}, refill=sp["refill_launches"]),
        # rays of the profiled pass (shade and refill work per extension ray)
        "stage_rays": {"extension": sp["extension_rays"], "shadow": sp["shadow_rays"]},
        "bvh_build_ms": round(st0["bvh_build_ms"], 2),
        "autotune": tune,
        "sort_hits": bool(a.sort_hits),
        "roofline": roofline,
        "convergence": conv,
        "cpu_baseline": cpu,
    }


def _timed_batch(batch, steps, warmup):
    for _ in range(warmup):
        batch.render()
    batch.reset_stats()
    t = time.perf_counter()
    for _ in range(steps):
        batch.render()
    st = batch.stats()  # synchronises
    el = time.perf_counter() - t
    return el, st


def run_views(a):
    """64 views of one scene in one persistent wavefront run (pt_amd/scenes/workloads.py
    views_batch): resolution log-uniform in 128^2..1024^2 pixels, aspect U(0.5, 2), spp
    log-uniform in 4..256, eye uniform in the scene box, direction uniform on the sphere,
    depth 8. The instanced scene by default (its box U(-50, 50)^3), the box scene with
    --scene box. A step is one pt_batch_render of all views."""
    if a.scene == "box":
        scene = pt_amd.Scene(generate.cached_scene("box", root=a.scene_dir, seed=1))
        views = wl.views_batch(box=((-1.0, -2.0, -1.0), (1.0, 0.0, 1.0)), focus_dist=1.0, lens_radius=0.01)
    else:
        scene = pt_amd.Scene(generate.cached_scene("instanced", root=a.scene_dir))
        views = wl.views_batch()
    tr = pt_amd.Tracer(scene, 64, 64)
    paths = sum(w * h * spp for w, h, spp, _, _ in views)
    log("views: 64 views of the %s scene, %d pixel-samples per step" % (a.scene, paths))
    batch = pt_amd.Batch([tr], [(0,) + v for v in views])
    el, st = _timed_batch(batch, a.steps, a.warmup)
    rays = st["extension_rays"] + st["shadow_rays"]
    batch.close()
    # the same views one after another, each its own one-view batch
    seq = 0.0
    for v in views:
        b = pt_amd.Batch([tr], [(0,) + v])
        b.render()
        b.stats()
        t = time.perf_counter()
        for _ in range(a.steps):
            b.render()
        b.stats()
        seq += time.perf_counter() - t
        b.close()
    return {
        "metric": "Mrays/s", "value": round(rays / el / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(el / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "64 views of the %s scene in one batch (res 128^2-1024^2, spp 4-256, depth 8)"
                               % a.scene, "pixel_samples_per_step": paths},
        "pixel_samples_per_s": round(paths * a.steps / el, 1),
        "sequential_views_s": round(seq, 4), "batched_s": round(el, 4),
        "batched_speedup": round(seq / el, 3),
    }


def run_scenes(a):
    """16 seeded instanced scenes with triangle totals log-uniformly spaced over 10k..20M
    (workloads.scene_sizes), each with its own device-built BVH and 4 views (log-uniform
    256^2..1920x1080 pixels, 16-128 spp, depth 8), all traced in one persistent wavefront run;
    then each scene's views alone. Per scene: BVH build time, Mrays/s, and the L2 hit rate of
    its traversal from the committed rocprofv3 summary (l2_hit_pct k_extend/scene<k>)."""
    sizes = wl.scene_sizes()
    only = a.only_scene
    tracers, scenes_info, views = [], [], []
    for k, target in enumerate(sizes):
        if only >= 0 and k != only:
            tracers.append(None)
            scenes_info.append(None)
            continue
        t = time.time()
        scn = generate.cached_scene("instanced", root=a.scene_dir, seed=100 + k,
                                    counts=generate.counts_for_triangles(target))
        scene = pt_amd.Scene(scn)
        tr = pt_amd.Tracer(scene, 64, 64)
        s0 = tr.stats()
        info = scene.info()
        scenes_info.append({"triangles": info["total_triangles"], "instances": info["instances"],
                            "bvh_build_ms": round(s0["bvh_build_ms"], 2), "bvh_mb": round(s0["bvh_bytes"] / 1e6, 1)})
        log("scene %d: %d triangles, %d instances, BVH %.1f ms (%.1f s incl. load)"
            % (k, info["total_triangles"], info["instances"], s0["bvh_build_ms"], time.time() - t))
        tracers.append(tr)
        views += [(k,) + v for v in wl.scene_views(100 + k)]
    if only >= 0:
        # one scene's four views alone (per-scene profiling)
        vk = [(0,) + v[1:] for v in views]
        b = pt_amd.Batch([tracers[only]], vk)
        ek, sk = _timed_batch(b, a.steps, a.warmup)
        rays = sk["extension_rays"] + sk["shadow_rays"]
        b.close()
        print("scene %d alone: %.1f Mrays/s" % (only, rays / ek / 1e6), file=sys.stderr, flush=True)
        return {"metric": "Mrays/s", "value": round(rays / ek / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
                "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ek / a.steps * 1e3, 3),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic (seeded procedural scenes)",
                "config": {"workload": "scene %d of the 16-scene batch alone (4 views, depth 8)" % only},
                "scene": scenes_info[only]}
    paths = sum(w * h * spp for _, w, h, spp, _, _ in views)
    log("scenes: 16 scenes, 64 views, %d pixel-samples per step" % paths)
    batch = pt_amd.Batch(tracers, views)
    el, st = _timed_batch(batch, a.steps, a.warmup)
    rays = st["extension_rays"] + st["shadow_rays"]
    batch.close()
    # each scene's views on their own
    seq_s = 0.0
    for k in range(16):
        vk = [(0,) + v[1:] for v in views if v[0] == k]
        b = pt_amd.Batch([tracers[k]], vk)
        ek, sk = _timed_batch(b, 1, 1)
        seq_s += ek
        scenes_info[k]["alone_s"] = round(ek, 4)
        scenes_info[k]["mrays_s"] = round((sk["extension_rays"] + sk["shadow_rays"]) / ek / 1e6, 2)
        scenes_info[k]["l2_hit_pct"] = profile_figure("l2_hit_pct", "k_extend/scene%d" % k)
        b.close()
    return {
        "metric": "Mrays/s", "value": round(rays / el / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(el / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": "16 instanced scenes (10k-20M triangles) x 4 views in one batch, depth 8",
                   "pixel_samples_per_step": paths},
        "pixel_samples_per_s": round(paths * a.steps / el, 1),
        # one step of the batch against the same 16 scenes' views rendered scene by scene
        "batched_s_per_step": round(el / a.steps, 4), "sequential_s_per_step": round(seq_s, 4),
        "batched_speedup": round(seq_s / (el / a.steps), 3) if el > 0 else None,
        "scenes": scenes_info,
    }


# The learned cache of the camera-path workload: Adam at learning rate 1e-3 on 4 minibatches of
# 16384 records after every frame (65,536 record samples per frame), one path in 4 recording
# its suffix from the second vertex, the others ending there with the prediction.
CACHE_PATH = dict(query=True, train_stride=4, train_steps=4, batch=16384, learning_rate=1e-3, frame_spp=4,
                  continue_prob=0.0, seed=1)


def run_cache(a):
    """The learned radiance cache on the instanced scene: 64 frames of 1920x1080 at 1-4 spp
    along a camera path through 8 random keyframes (workloads.camera_path), depth 8, the cache
    trained online after every frame on that frame's path suffixes and queried at the second
    vertex. The same frames with the plain estimator for comparison, and at 4 frames the
    relative MSE of both against a 4096-spp plain reference of disjoint sample indices. A step
    is the 64-frame path."""
    scene = pt_amd.Scene(generate.cached_scene("instanced", root=a.scene_dir))
    tr = pt_amd.Tracer(scene, W, H, max_bounces=DEPTH)
    path = wl.camera_path()
    check = [] if a.no_reference else [15, 31, 47, 63]
    refs = {}
    t = time.perf_counter()
    for f in check:
        refs[f] = reference_image(tr, path[f][1], 4096)
    log("cache: %d references in %.1f s" % (len(check), time.perf_counter() - t))

    def run(cached):
        frames, errs, losses, records = [], {}, [], []
        if cached:
            tr.cache_configure(**dict(CACHE_PATH, ema_decay=a.cache_ema))
        rays = 0
        for f, (spp, cam) in enumerate(path):
            tr.set_camera(cam)
            tr.synchronize()
            t0 = time.perf_counter()
            tr.render(spp)
            tr.synchronize()
            frames.append(time.perf_counter() - t0)
            st = tr.stats()
            rays += st["extension_rays"] + st["shadow_rays"]
            if cached:
                cs = tr.cache_stats()
                losses.append(round(cs["last_loss"], 5))
                records.append(cs["last_records"])
            if f in refs:
                errs[f] = float("%.4e" % pt_amd.rel_mse(tr.read(), refs[f]))
        if cached:
            tr.cache_configure(enable=False)
        return frames, errs, losses, records, rays

    for _ in range(max(0, a.warmup)):
        run(False)
    pf, perr, _, _, prays = run(False)
    cf, cerr, loss, recs, crays = run(True)
    log("cache: plain %.1f ms/frame, cached %.1f ms/frame; rel-MSE plain %s cached %s"
        % (1e3 * np.mean(pf), 1e3 * np.mean(cf), perr, cerr))
    total = sum(cf)
    return {
        "metric": "Mrays/s", "value": round(crays / total / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": 1,
        "warmup": a.warmup, "ms_per_step": round(total * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (cache network bf16 weights, fp32 accumulation)",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "instanced scene, 64 frames of 1920x1080 at 1-4 spp on a keyframed camera path, "
                               "depth 8, learned radiance cache trained online", "cache": dict(CACHE_PATH, ema_decay=a.cache_ema)},
        "frame_ms": {"plain_mean": round(1e3 * np.mean(pf), 3), "cached_mean": round(1e3 * np.mean(cf), 3)},
        "frame_spp": [spp for spp, _ in path],
        "relmse_vs_4096spp": {"frames": check, "plain": [perr[f] for f in check], "cached": [cerr[f] for f in check]},
        "cache_loss_first_last": [loss[:4], loss[-4:]],
        "records_per_frame": int(np.median(recs)) if recs else 0,
        "plain_mrays_s": round(prays / sum(pf) / 1e6, 2),
        # the network's kernels against the dense BF16 MFMA peak (committed rocprofv3 summary)
        "mfma": {k: {"tflops": profile_figure("mfma_tflops", k), "busy_pct": profile_figure("mfma_busy_pct", k),
                     "frac_of_2500_tflops": (round(profile_figure("mfma_tflops", k) / 2500.0, 5)
                                             if profile_figure("mfma_tflops", k) else None)}
                 for k in ("k_mlp_train", "k_mlp_infer")},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--spp-per-step", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["instanced", "views", "scenes", "cache"], default="instanced")
    ap.add_argument("--scene", choices=["instanced", "box"], default="instanced", help="views workload: the scene")
    ap.add_argument("--scene-dir", default=os.environ.get("PT_SCENE_DIR", "/tmp/pt_scenes"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-convergence", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-reference", action="store_true", help="cache workload: no 4096-spp references")
    ap.add_argument("--leaf-size", type=int, default=0, help="BLAS leaf size to start from (0: the library's)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="traversal blocks per CU to start from")
    ap.add_argument("--pool-slots", type=int, default=0, help="path pool slots to start from")
    ap.add_argument("--tlas-open", type=int, default=-1, help="TLAS opening depth to start from (-1: the library's)")
    ap.add_argument("--node-width", type=int, default=0, help="BVH node width to start from (0: the library's)")
    ap.add_argument("--lds-stack", type=int, default=0, help="LDS stack depth to start from (0: the library's)")
    ap.add_argument("--cache-ema", type=float, default=0.0,
                    help="camera-path workload: decay of the weights' moving average the queries read (0: none)")
    ap.add_argument("--only-scene", type=int, default=-1, help="scenes workload: render scene k alone")
    ap.add_argument("--sort-hits", type=int, default=0, help="1: shade hits in material order")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; every box here has one, so --gpus %d is not run"
              % a.gpus, file=sys.stderr)
        return 2
    out = {"instanced": run_instanced, "views": run_views, "scenes": run_scenes, "cache": run_cache}[a.workload](a)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
