#!/usr/bin/env python3
"""Benchmark of the plain wavefront path tracer on one MI355X.

Workload (the real-shape configuration): the instanced scene - 300 instances of five seeded
meshes (icospheres L2/L4/L6, a 512x512 noise terrain, a 200k-triangle soup; ~2.95M
triangles, per-instance materials, 20 emissive spheres) - at 1920x1080, max depth 8, fp32.

A step is one batch of 8 samples per pixel over the whole 1920x1080 frame (pt_render(8):
8 sample passes, each traced exactly as the reference's 1-spp submits,
src/pathtrace/ptdispatch.cpp:101); the default 8 steps are the 64-spp frame. Inputs are
resident in HBM before the timed region (scene, BVH, camera); the timed region is K
back-to-back pt_render(8) calls bracketed by stream synchronisation.

The JSON line also carries:
  roofline      the dominant stage's algorithmic bytes per launch / its average HIP-event
                duration, against the 8 TB/s HBM peak. Bytes per ray: 32 (ray) + 20 (hit)
                + 128 x node fetches + 48 x triangle-record fetches (extension) and
                32 + 4 + 128 x nodes + 48 x tris (shadow); the fetch counts come from an
                instrumented replay of the same sample indices.
  convergence   seconds of rendering until the relative MSE against a 4096-spp plain
                reference is <= 1e-3, for the plain estimator (16-spp render calls) and for the
                learned-cache estimator (16-spp render calls, the network trained in flight
                between wavefront iterations, paths ended at their second vertex with its
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


def convergence(tr, cam, threshold=1e-3, ref_spp=4096):
    """Seconds to rel-MSE <= threshold against a ref_spp plain reference, plain vs cached."""
    t = time.perf_counter()
    tr.set_camera(cam)
    for _ in range(ref_spp // 512):
        tr.render(512)
    tr.synchronize()
    ref_s = time.perf_counter() - t
    ref = tr.read()
    log("convergence: %d-spp reference in %.1f s" % (ref_spp, ref_s))
    plain_s, plain_spp, plain_e, plain_traj = time_to_threshold(tr, cam, ref, 16, 4096, 60.0, threshold)
    log("convergence: plain %.3f s, %d spp, rel-MSE %.3e" % (plain_s, plain_spp, plain_e))
    # in-flight training (two Adam steps of 8192 records after every wavefront iteration), one
    # path in 16 trains; the others end at their second vertex with the prediction
    cfg = dict(query=True, train_stride=16, train_steps=2, batch=8192, learning_rate=1e-3, frame_spp=0,
               continue_prob=0.0, seed=1)
    tr.cache_configure(**cfg)
    warm = 16
    cached_s, cached_spp, cached_e, cached_traj = time_to_threshold(
        tr, cam, ref, 16, 4096, max(60.0, 3 * plain_s), threshold, before=lambda: tr.render(warm))
    cs = tr.cache_stats()
    tr.cache_configure(enable=False)
    log("convergence: cached %.3f s, %d spp, rel-MSE %.3e" % (cached_s, cached_spp, cached_e))
    return {"threshold": threshold, "reference_spp": ref_spp, "reference_s": round(ref_s, 3),
            "plain_s": round(plain_s, 4), "plain_spp": plain_spp, "plain_relmse": plain_e,
            "cached_s": round(cached_s, 4), "cached_spp": cached_spp, "cached_relmse": cached_e,
            "cached_over_plain_time": round(cached_s / plain_s, 3) if plain_s > 0 else None,
            "cache_config": dict(cfg, warmup_spp=warm, steps_done=cs["steps"]),
            "plain_trajectory": plain_traj[-4:], "cached_trajectory": cached_traj[-4:]}


def measured_traffic(kernel):
    """HBM bytes per launch of a stage from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "k_extend_instanced_1080p_pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--spp-per-step", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scene-dir", default=os.environ.get("PT_SCENE_DIR", "/tmp/pt_scenes"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-convergence", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; every box here has one, so --gpus %d is not run"
              % a.gpus, file=sys.stderr)
        return 2

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
    bytes_ext = 32 + 20 + 128 * n_node_ext + 48 * n_tri_ext
    bytes_sh = 32 + 4 + 128 * n_node_sh + 48 * n_tri_sh
    log("fetches/ray: extension %.1f nodes %.1f tris (%.0f B), shadow %.1f nodes %.1f tris (%.0f B)"
        % (n_node_ext, n_tri_ext, bytes_ext, n_node_sh, n_tri_sh, bytes_sh))

    # timed region: exactly K steps
    tr.set_option(pt_amd.PT_OPT_PROFILE_KERNELS, 1)
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
    log("timed: %d steps in %.3f s, %d ext + %d shadow rays; stages: extend %.1f ms, shade %.1f ms, shadow %.1f ms"
        % (a.steps, elapsed, st["extension_rays"], st["shadow_rays"], st["extend_ms"], st["shade_ms"],
           st["shadow_ms"]))

    # the 8-bounce shade stage reads per hit: hit 20 + path state 32 + instance 128 + material
    # 96 + indices 12 + 3 vertices 96 + light bins 64 + emitter vertices 36, writes next ray
    # 36 + path state 32 + shadow record 64.
    stages = {
        "extend": {"ms": st["extend_ms"], "launches": st["extend_launches"],
                   "bytes": st["extension_rays"] * bytes_ext},
        "shadow": {"ms": st["shadow_ms"], "launches": st["shadow_launches"],
                   "bytes": st["shadow_rays"] * bytes_sh},
        "shade": {"ms": st["shade_ms"], "launches": st["shade_launches"],
                  "bytes": st["extension_rays"] * 616},
    }
    dom = max(stages, key=lambda k: stages[k]["ms"])
    d = stages[dom]
    achieved = d["bytes"] / (d["ms"] / 1e3) / 1e9 if d["ms"] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic("k_" + dom),
                "kernel": "k_" + dom,
                "bytes_per_launch": d["bytes"] / max(1, d["launches"]),
                "avg_launch_ms": d["ms"] / max(1, d["launches"])}

    conv = None
    if not a.no_convergence:
        tr.set_option(pt_amd.PT_OPT_PROFILE_KERNELS, 0)
        conv = convergence(tr, cam)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(scene, cam, a.cpu_threads)

    out = {
        "metric": "Mrays/s", "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "instanced scene, 300 instances / %d triangles, 1920x1080, %d spp per step, depth %d"
                               % (info["total_triangles"], a.spp_per_step, DEPTH),
                   "spp_total": a.steps * a.spp_per_step, "width": W, "height": H, "max_depth": DEPTH},
        "samples_per_s": round(a.steps * a.spp_per_step / elapsed, 3),
        "pixel_samples_per_s": round(a.steps * a.spp_per_step * W * H / elapsed, 1),
        "rays": {"extension": st["extension_rays"], "shadow": st["shadow_rays"]},
        "fetches_per_ray": {"ext_nodes": round(n_node_ext, 3), "ext_tris": round(n_tri_ext, 3),
                            "shadow_nodes": round(n_node_sh, 3), "shadow_tris": round(n_tri_sh, 3)},
        "stages_ms": {k: round(v["ms"], 3) for k, v in stages.items()},
        "bvh_build_ms": round(st0["bvh_build_ms"], 2),
        "roofline": roofline,
        "convergence": conv,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
