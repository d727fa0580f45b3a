"""FP64 adjudication of hit mismatches between the device traversal and the oracle.

The device intersects in fp32 (watertight test of Woop, Benthin and Wald 2013, rays moved to
object space with an fp32 inverse transform); the oracle in fp64. They may disagree only where
fp32 rounding can decide: a ray passing within a few fp32 ulps of a triangle's edge (the
triangle is hit on one side and missed on the other), two triangles at the same depth along
the ray (intersecting or coincident surfaces), or a hit at the ends of the ray's interval.
Every mismatch is recomputed here in fp64 from the scene's world-space triangles and accepted
only if it is one of those cases; anything else - a hit the fp64 test rejects by more than the
rounding bound, or a nearer hit missed - is a traversal error.

Tolerance: distances in world units of EDGE_ULPS fp32 ulps of the largest coordinate involved
(ray origin, hit point, triangle corners), i.e. EDGE_ULPS * 2^-24 * scale.
"""
import numpy as np

import pt_amd

EDGE_ULPS = 64.0
U32 = 2.0 ** -24


class WorldTriangles:
    """World-space corners (fp64) of any (instance, primitive) of a packed scene."""

    def __init__(self, scene):
        a = scene.arrays()
        self.v = a["vertices"][:, :3].astype(np.float64)
        self.idx = a["indices"]
        inst = a["instances"].reshape(-1, 128)
        self.m = inst[:, 16:80].copy().view(np.float32).reshape(-1, 4, 4).astype(np.float64)
        self.meshes = a["meshes"]
        self.inst_mesh = a["instance_mesh"]

    def corners(self, inst, prim):
        vs, is_, nt, nv = self.meshes[int(self.inst_mesh[inst])]
        assert prim < nt, (inst, prim, nt)
        ids = self.idx[is_ + 3 * prim:is_ + 3 * prim + 3].astype(np.int64)
        p = self.v[vs + ids]
        m = self.m[inst]
        return p @ m[:3, :3].T + m[:3, 3]


def ray_triangle(o, d, tri):
    """(t, edge distance): ray parameter of the hit with the triangle's plane, and the signed
    world distance of that point from the triangle's boundary (positive inside)."""
    a, b, c = tri
    n = np.cross(b - a, c - a)
    area2 = np.linalg.norm(n)
    denom = np.dot(n, d)
    if area2 == 0.0 or denom == 0.0:
        return np.inf, -np.inf
    t = np.dot(n, a - o) / denom
    p = o + t * d
    # barycentric weights, each times the altitude over its opposite edge = distance to it
    dist = []
    for v0, v1, v2 in ((a, b, c), (b, c, a), (c, a, b)):
        e = v2 - v1
        w = np.dot(np.cross(v2 - v1, p - v1), n) / (area2 * area2)  # weight of v0
        dist.append(w * area2 / max(np.linalg.norm(e), 1e-300))
    return float(t), float(min(dist))


class Adjudicator:
    def __init__(self, scene, edge_ulps=EDGE_ULPS):
        self.tris = WorldTriangles(scene)
        self.edge_ulps = edge_ulps
        self.reasons = {}

    def _info(self, o, d, hit):
        if int(hit["instance"]) == pt_amd.PT_NO_HIT:
            return None
        tri = self.tris.corners(int(hit["instance"]), int(hit["primitive"]))
        t, dist = ray_triangle(o, d, tri)
        p = o + t * d if np.isfinite(t) else o
        scale = max(np.abs(o).max(), np.abs(p).max(), np.abs(tri).max(), 1.0)
        return {"t": t, "dist": dist, "eps": self.edge_ulps * U32 * scale}

    def ok(self, ray, g, o_hit, any_hit=False):
        """True when the device hit g and the oracle hit o_hit of this ray differ only where fp32
        rounding decides. Records the reason in self.reasons."""
        o = np.array(ray["origin"], np.float64)
        d = np.array(ray["dir"], np.float64)
        tmin, tmax = float(ray["tmin"]), float(ray["tmax"])
        dn = np.linalg.norm(d)
        G, O = self._info(o, d, g), self._info(o, d, o_hit)

        def graze(x):  # the ray passes within the rounding bound of the triangle's edge
            return abs(x["dist"]) <= x["eps"]

        def valid(x):  # an fp64 hit, up to rounding, inside [tmin, tmax] up to rounding
            tol = x["eps"] / dn
            return x["dist"] >= -x["eps"] and tmin - tol <= x["t"] <= tmax + tol

        def at_end(x):
            tol = x["eps"] / dn
            return abs(x["t"] - tmax) <= tol or abs(x["t"] - tmin) <= tol

        def note(r):
            self.reasons[r] = self.reasons.get(r, 0) + 1
            return True

        for side in (G, O):
            if side is not None and not valid(side):
                return False
        if any_hit:  # only occlusion is compared: the side that found a hit must have been marginal
            side = G if G is not None else O
            return side is not None and (graze(side) and note("edge") or at_end(side) and note("interval end"))
        if G is not None and O is not None:
            tol = max(G["eps"], O["eps"]) / dn
            if abs(G["t"] - O["t"]) <= 8 * tol:
                return note("depth tie")
            nearer, farther = (G, O) if G["t"] < O["t"] else (O, G)
            # the nearer triangle was hit on one side only: it must be grazed (or at the interval end)
            return graze(nearer) and note("edge") or at_end(nearer) and note("interval end")
        side = G if G is not None else O
        return graze(side) and note("edge") or at_end(side) and note("interval end")
