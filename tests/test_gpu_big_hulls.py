# SPDX-License-Identifier: Apache-2.0
"""Hulls of more than 32 vertices (up to the reference's 256): their pairs
run in the large-hull narrow phase (collide_big.hip), the same collide code
built for 256-vertex hulls and the reference's clip capacity. Pair queries
against every shape type, a world of mixed bodies and large hulls on a
ground box and on a mesh terrain, and a batch with such a world, each bit
for bit against the reference (oracle/_ref). The hulls are built by this
library's builder, whose blocks equal the reference's (tests/test_hull.py)."""
import ctypes as C
import math
import os

import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big_bank(gpu_api):
    return S.built_hull_bank(gpu_api, (33, 40, 64, 100, 160, 256))


@pytest.fixture(scope="module")
def small_big_bank(gpu_api):
    return S.built_hull_bank(gpu_api, (33, 48, 64), seed=9)


def _geometry(r, kind, bank):
    if kind == 0:
        return P.geometry_sphere(float(r.next_range(0.2, 0.6))), 0.6
    if kind == 1:
        rad = float(r.next_range(0.2, 0.4))
        hh = float(r.next_range(0.3, 0.8))
        return P.geometry_capsule(hh, rad), hh + rad
    if kind == 2:
        h = [float(r.next_range(0.2, 0.7)) for _ in range(3)]
        return P.geometry_box(*h), math.sqrt(sum(x * x for x in h))
    return P.geometry_hull(bank.ptrs[r.next_bounded(len(bank.ptrs))]), 0.62


def _pairs(seed, n, bank):
    r = S.Pcg32(seed)
    out = []
    for i in range(n):
        ka, kb = 3, i % 4  # a large hull against every type
        if i % 2:
            ka, kb = kb, ka
        ga, ra = _geometry(r, ka, bank)
        gb, rb = _geometry(r, kb, bank)
        pa = P.Pose(P.Vec3(*(float(r.next_range(-1.0, 1.0)) for _ in range(3))),
                    P.Quat(*S.rotation(r)))
        d = [float(r.next_range(-1.0, 1.0)) for _ in range(3)]
        norm = math.sqrt(sum(x * x for x in d)) or 1.0
        dist = float(r.next_range(0.3, 1.1)) * (ra + rb)
        pb = P.Pose(P.Vec3(*(pa.position.x + d[0] / norm * dist, pa.position.y + d[1] / norm * dist,
                             pa.position.z + d[2] / norm * dist)), P.Quat(*S.rotation(r)))
        out.append((ga, pa, gb, pb, float(r.next_range(0.0, 0.08))))
    return out


def _manifold_bytes(m):
    return bytes(m)[:16 + 20 * m.point_count]


def test_large_hull_pairs_match_the_reference(gpu_api, ref_api, big_bank):
    print("big hulls: 400 pairs against every shape type, twice each with the cache", flush=True)
    g = P.bind_collide(gpu_api)
    f = P.bind_collide(ref_api)
    touching = 0
    for ga, pa, gb, pb, spec in _pairs(41, 400, big_bank):
        cg, cr = P.CollideCache(), P.CollideCache()
        for _ in range(2):
            mg, mr = P.Manifold(), P.Manifold()
            g.collide(C.byref(ga), pa, C.byref(gb), pb, spec, C.byref(cg), C.byref(mg))
            f.collide(C.byref(ga), pa, C.byref(gb), pb, spec, C.byref(cr), C.byref(mr))
            assert _manifold_bytes(mg) == _manifold_bytes(mr)
            assert bytes(cg) == bytes(cr)
        touching += mr.point_count > 0
    assert touching > 100


def test_mixed_launch_equals_single_queries(gpu_api, big_bank, hull_bank):
    """Large and 12-vertex hull pairs in one nph_amd_collide_many launch:
    each as its own query gives."""
    g = P.bind_collide(gpu_api)
    pairs = _pairs(42, 200, big_bank) + _pairs(43, 200, hull_bank)
    qs = (P.CollideQuery * len(pairs))()
    for q, (ga, pa, gb, pb, spec) in zip(qs, pairs):
        q.a, q.pose_a, q.b, q.pose_b, q.speculative_distance = ga, pa, gb, pb, spec
    out = (P.Manifold * len(pairs))()
    caches = (P.CollideCache * len(pairs))()
    assert g.collide_many(qs, len(pairs), out, caches) == 0
    for k, (ga, pa, gb, pb, spec) in enumerate(pairs):
        m, c = P.Manifold(), P.CollideCache()
        g.collide(C.byref(ga), pa, C.byref(gb), pb, spec, C.byref(c), C.byref(m))
        assert _manifold_bytes(out[k]) == _manifold_bytes(m), k
        assert bytes(caches[k]) == bytes(c), k


def _drop(api, bank, world=None, meshes=None):
    if meshes is not None:
        return S.mesh_terrain_drop(api, meshes, 40, 61, bank, world=world)
    return S.bin_drop(api, 60, 62, bank, half=3.5, height=9.0, y_lo=1.0, world=world)


def test_worlds_with_large_hulls_match_the_reference(gpu_api, ref_api, small_big_bank):
    meshes = S.load_mesh_bank(os.path.join(GOLDEN, "meshes.bin"))
    for label, mesh in (("bin", None), ("mesh terrain", meshes)):
        print(f"big hulls: {label} drop, 90 steps against the reference", flush=True)
        w = _drop(gpu_api, small_big_bank, meshes=mesh)
        r = _drop(ref_api, small_big_bank, meshes=mesh)
        for step in range(90):
            w.step(S.DT)
            r.step(S.DT)
            assert w.hash() == r.hash(), f"{label}: hash differs at step {step + 1}"
        assert w.stats().touching_contacts > 20
        w.close()
        r.close()


def test_a_batch_with_large_hulls_matches_the_reference(gpu_api, ref_api, small_big_bank,
                                                         hull_bank):
    print("big hulls: a batch of 3 worlds, 60 steps against the reference", flush=True)
    batch = P.Batch(gpu_api, 3, S.DESC)
    fills = [lambda w: S.bin_drop(None, 60, 62, small_big_bank, half=3.5, height=9.0, y_lo=1.0,
                                  world=w),
             lambda w: S.bin_drop(None, 80, 63, hull_bank, half=3.5, height=9.0, y_lo=1.0,
                                  world=w),
             lambda w: S.pyramid(None, 6, world=w)]
    refs = []
    for fill, w in zip(fills, batch.worlds):
        fill(w)
        r = P.World(ref_api, S.DESC)
        fill(r)
        refs.append(r)
    for step in range(60):
        batch.step(S.DT)
        for k, (w, r) in enumerate(zip(batch.worlds, refs)):
            r.step(S.DT)
            assert w.hash() == r.hash(), f"world {k}: hash differs at step {step + 1}"
    batch.close()


def test_hull_faces_beyond_64_vertices_are_refused(gpu_api):
    """A 100-segment cylinder's caps are 100-gons: refused
    (NPH_RESULT_UNSUPPORTED); its 60-segment sibling (120 vertices, 60-gon
    caps) is taken."""
    import numpy as np
    shapes = {}
    for seg in (100, 60):
        a = 2.0 * np.pi * np.arange(seg) / seg
        pts = np.zeros((2 * seg, 3), np.float32)
        pts[0::2] = np.stack([np.cos(a), np.full(seg, -0.5), np.sin(a)], 1)
        pts[1::2] = np.stack([np.cos(a), np.full(seg, 0.5), np.sin(a)], 1)
        rc, block = P.build_hull(gpu_api, pts, 256)
        assert rc == 0
        shapes[seg] = S.HullBank(block, [0])
    w = P.World(gpu_api, S.DESC)
    b = w.add_body(P.BODY_DYNAMIC, (0.0, 2.0, 0.0))
    with pytest.raises(RuntimeError, match="failed: 8"):
        w.add_shape(b, P.geometry_hull(shapes[100].ptrs[0]))
    w.add_shape(b, P.geometry_hull(shapes[60].ptrs[0]))
    w.step(S.DT)
    w.close()
    # The pair queries refuse the 100-gon caps too (an empty manifold, and
    # NPH_RESULT_UNSUPPORTED from the batched queries), and take the 60-gon's.
    g = P.bind_collide(gpu_api)
    low = P.Pose(P.Vec3(0.0, 0.0, 0.0), P.Quat(0.0, 0.0, 0.0, 1.0))
    high = P.Pose(P.Vec3(0.0, 0.95, 0.0), P.Quat(0.0, 0.0, 0.0, 1.0))
    for seg, touching in ((100, False), (60, True)):
        geom = P.geometry_hull(shapes[seg].ptrs[0])
        m = P.Manifold()
        g.collide(C.byref(geom), low, C.byref(geom), high, 0.02, None, C.byref(m))
        assert (m.point_count > 0) == touching, seg
        q = (P.CollideQuery * 1)()
        q[0].a, q[0].pose_a, q[0].b, q[0].pose_b = geom, low, geom, high
        q[0].speculative_distance = 0.02
        out = (P.Manifold * 1)()
        rc = g.collide_many(q, 1, out, None)
        assert rc == (P.OK if touching else P.UNSUPPORTED), seg
