"""The batched and learned-cache workloads, as seeded lists of views, scenes and camera paths.

  views_batch   64 views of one scene: resolution log-uniform in 128^2..1024^2 pixels (aspect
                U(0.5, 2)), spp log-uniform in 4..256, eye uniform in the scene box, view
                direction uniform on the sphere, depth 8.
  scene_sizes   16 triangle totals log-uniformly spaced from 10k to 20M (both ends included).
  scene_views   4 views of one of those scenes: log-uniform 256^2..1920x1080 pixels, 16-128
                spp, depth 8.
  camera_path   64 frames of 1920x1080 at 1-4 spp along a path through 8 random keyframes
                (Catmull-Rom through the eye positions, look-at points likewise).

Cameras are pt_camera blocks from camera_block, which takes the .scn position and lookdir:
the reference's view matrix negates both (src/scene/camera.h:16), so an eye e looking along f
is written as position -e, lookdir -f.
"""
import numpy as np

import pt_amd
from pt_amd.scenes.rng import Xshiro128

DEPTH = 8
INSTANCED_BOX = ((-50.0, -50.0, -50.0), (50.0, 50.0, 50.0))  # instance positions U(-50, 50)^3


def random_dir(rng):
    z = rng.uniform(-1.0, 1.0)
    phi = rng.uniform(0.0, 2 * np.pi)
    r = (1 - z * z) ** 0.5
    return (r * np.cos(phi), r * np.sin(phi), z)


def camera(eye, fwd, w, h, focus_dist=30.0, lens_radius=0.03, fov_deg=60.0):
    return pt_amd.camera_block([-float(e) for e in eye], [-float(f) for f in fwd], fov_deg=fov_deg, aspect=w / h,
                               focus_dist=focus_dist, lens_radius=lens_radius)


def _size(rng, lo_area, hi_area):
    area = rng.log_uniform(lo_area, hi_area)
    aspect = rng.uniform(0.5, 2.0)
    w = max(1, int(round((area * aspect) ** 0.5)))
    h = max(1, int(round(area / w)))
    return w, h


def views_batch(seed=3, n=64, box=INSTANCED_BOX, focus_dist=30.0, lens_radius=0.03):
    """[(w, h, spp, depth, camera)] of the 64-view batch over a scene whose box is `box`."""
    rng = Xshiro128.seeded(seed)
    lo, hi = box
    out = []
    for _ in range(n):
        w, h = _size(rng, 128.0 ** 2, 1024.0 ** 2)
        spp = int(round(rng.log_uniform(4.0, 256.0)))
        eye = [rng.uniform(lo[k], hi[k]) for k in range(3)]
        fwd = random_dir(rng)
        out.append((w, h, spp, DEPTH, camera(eye, fwd, w, h, focus_dist, lens_radius)))
    return out


def scene_sizes(n=16, lo=1e4, hi=2e7):
    """n triangle totals log-uniformly spaced over [lo, hi], both ends included."""
    return [float(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * k / (n - 1))) for k in range(n)]


def scene_views(seed, n=4, spp_lo=16, spp_hi=128):
    """[(w, h, spp, depth, camera)] of one scene of the batch of scenes."""
    rng = Xshiro128.seeded(seed)
    out = []
    for _ in range(n):
        w, h = _size(rng, 256.0 ** 2, 1920.0 * 1080.0)
        spp = int(round(rng.log_uniform(spp_lo, spp_hi)))
        eye = [rng.uniform(-20, 20) for _ in range(3)]
        fwd = random_dir(rng)
        out.append((w, h, spp, DEPTH, camera(eye, fwd, w, h)))
    return out


def _catmull_rom(p0, p1, p2, p3, t):
    t2, t3 = t * t, t * t * t
    return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 + (-p0 + 3 * p1 - 3 * p2 + p3) * t3)


def camera_path(seed=5, frames=64, keys=8, w=1920, h=1080, spp_lo=1, spp_hi=4, radius=35.0):
    """[(spp, camera)] of a camera path through `keys` random keyframes: eyes uniform in a ball of
    `radius` around the scene's centre looking at points within a quarter of it (so that frames
    see geometry), interpolated by a Catmull-Rom spline over `frames` frames; spp uniform in
    spp_lo..spp_hi per frame."""
    rng = Xshiro128.seeded(seed)

    def in_ball(r):
        while True:
            p = np.array([rng.uniform(-1, 1) for _ in range(3)])
            if np.dot(p, p) <= 1.0:
                return p * r
    eyes = np.array([in_ball(radius) for _ in range(keys)])
    looks = np.array([in_ball(radius / 4) for _ in range(keys)])
    out = []
    for f in range(frames):
        x = f * (keys - 1) / (frames - 1)
        k = min(int(x), keys - 2)
        t = x - k
        idx = [max(0, k - 1), k, k + 1, min(keys - 1, k + 2)]
        eye = _catmull_rom(*(eyes[i] for i in idx), t)
        at = _catmull_rom(*(looks[i] for i in idx), t)
        fwd = at - eye
        fwd = fwd / np.linalg.norm(fwd)
        spp = spp_lo + int(rng.next_float() * (spp_hi - spp_lo + 1) * 0.999999)
        out.append((spp, camera(eye, fwd, w, h)))
    return out
