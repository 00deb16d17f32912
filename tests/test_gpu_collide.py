# SPDX-License-Identifier: Apache-2.0
"""The reference's pair queries (collide.h:96-121) on the device: the
manifold of every convex pair (spheres, capsules, boxes, 12-vertex hulls) at
seeded random poses near contact, with and without the cache, and the
manifold reduction of seeded candidate sets, bit for bit against the
reference's own functions (oracle/_ref); and the batched query against the
single one."""
import ctypes as C
import math

import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

pytestmark = pytest.mark.gpu


def _geometry(r: S.Pcg32, kind: int, hull_bank):
    if kind == 0:
        return P.geometry_sphere(float(r.next_range(0.2, 0.6))), 0.6
    if kind == 1:
        rad = float(r.next_range(0.2, 0.4))
        hh = float(r.next_range(0.3, 0.8))
        return P.geometry_capsule(hh, rad), hh + rad
    if kind == 2:
        h = [float(r.next_range(0.2, 0.7)) for _ in range(3)]
        return P.geometry_box(*h), math.sqrt(sum(x * x for x in h))
    return P.geometry_hull(hull_bank.ptrs[r.next_bounded(len(hull_bank))]), 0.6


def _pairs(seed, n, hull_bank):
    r = S.Pcg32(seed)
    out = []
    for i in range(n):
        ka, kb = i % 4, (i // 4) % 4
        ga, ra = _geometry(r, ka, hull_bank)
        gb, rb = _geometry(r, kb, hull_bank)
        pa = P.Pose(P.Vec3(*(float(r.next_range(-1.0, 1.0)) for _ in range(3))),
                    P.Quat(*S.rotation(r)))
        d = [float(r.next_range(-1.0, 1.0)) for _ in range(3)]
        norm = math.sqrt(sum(x * x for x in d)) or 1.0
        dist = float(r.next_range(0.3, 1.15)) * (ra + rb)
        pos = [pa.position.x + d[0] / norm * dist, pa.position.y + d[1] / norm * dist,
               pa.position.z + d[2] / norm * dist]
        pb = P.Pose(P.Vec3(*pos), P.Quat(*S.rotation(r)))
        spec = float(r.next_range(0.0, 0.08))
        out.append((ga, pa, gb, pb, spec))
    return out


def _manifold_bytes(m: P.Manifold):
    n = m.point_count
    return bytes(m)[:16 + 20 * n]


def test_pairs_match_the_reference(gpu_api, ref_api, hull_bank):
    print("collide: 2 000 random convex pairs, twice each with the cache", flush=True)
    g = P.bind_collide(gpu_api)
    f = P.bind_collide(ref_api)
    touching = 0
    for ga, pa, gb, pb, spec in _pairs(31, 2000, hull_bank):
        cg, cr = P.CollideCache(), P.CollideCache()
        for _ in range(2):  # the second call starts from the first's cache
            mg, mr = P.Manifold(), P.Manifold()
            g.collide(C.byref(ga), pa, C.byref(gb), pb, spec, C.byref(cg), C.byref(mg))
            f.collide(C.byref(ga), pa, C.byref(gb), pb, spec, C.byref(cr), C.byref(mr))
            assert _manifold_bytes(mg) == _manifold_bytes(mr)
            assert bytes(cg) == bytes(cr)
        touching += mr.point_count > 0
    assert touching > 500
    for ta in range(4):
        for tb in range(4):
            assert g.collide_supported(ta, tb) == f.collide_supported(ta, tb)
    assert not g.collide_supported(P.SHAPE_MESH, 0)


def test_many_pairs_in_one_launch_equal_one_at_a_time(gpu_api, hull_bank):
    g = P.bind_collide(gpu_api)
    pairs = _pairs(32, 1500, hull_bank)
    q = (P.CollideQuery * len(pairs))()
    for i, (ga, pa, gb, pb, spec) in enumerate(pairs):
        q[i].a, q[i].pose_a, q[i].b, q[i].pose_b = ga, pa, gb, pb
        q[i].speculative_distance = spec
    out = (P.Manifold * len(pairs))()
    caches = (P.CollideCache * len(pairs))()
    assert g.collide_many(q, len(pairs), out, caches) == P.OK
    for i, (ga, pa, gb, pb, spec) in enumerate(pairs):
        m, c = P.Manifold(), P.CollideCache()
        g.collide(C.byref(ga), pa, C.byref(gb), pb, spec, C.byref(c), C.byref(m))
        assert _manifold_bytes(m) == _manifold_bytes(out[i]) and bytes(c) == bytes(caches[i]), i


def test_manifold_reduction_matches_the_reference(gpu_api, ref_api):
    """Seeded candidate sets of 5 to 64 points (some coplanar, some lines,
    ties), reduced to at most four by both libraries."""
    g = P.bind_collide(gpu_api)
    f = P.bind_collide(ref_api)
    r = S.Pcg32(33)
    for case in range(300):
        n = 5 + r.next_bounded(60)
        pts = (P.ManifoldPoint * n)()
        flat = case % 3 == 0
        line = case % 7 == 0
        for i in range(n):
            x = float(r.next_range(-1.0, 1.0))
            z = 0.0 if line else float(r.next_range(-1.0, 1.0))
            y = 0.0 if flat else float(r.next_range(-0.05, 0.05))
            pts[i].point = P.Vec3(x, y, z)
            pts[i].separation = -0.01 if flat else float(r.next_range(-0.05, 0.01))
            pts[i].id = i
        normal = P.Vec3(0.0, 1.0, 0.0)
        a = (P.ManifoldPoint * n).from_buffer_copy(pts)
        b = (P.ManifoldPoint * n).from_buffer_copy(pts)
        kg = g.manifold_reduce(a, n, normal)
        kr = f.manifold_reduce(b, n, normal)
        assert kg == kr and bytes(a)[:20 * kg] == bytes(b)[:20 * kr], case
