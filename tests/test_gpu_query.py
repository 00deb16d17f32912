# SPDX-License-Identifier: Apache-2.0
"""Casts, overlaps and world queries against the reference's (engine/physics/
src/cast.c, query.c; collide.h:240-318, query.h:51-172) on seeded scenes:
both worlds stepped alike (equal hashes), then the same queries on each;
every reported hit, shape and manifold, in order, bit for bit.

Scope: meshes in the world are met by rays, box overlaps and shape
overlaps; shape casts and collide_shape pass over them in this library, so
those two are pinned on convex-only worlds."""
import ctypes as C
import os

import numpy as np
import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DT = 1.0 / 60.0


@pytest.fixture(scope="module")
def mesh_bank():
    return S.load_mesh_bank(os.path.join(GOLDEN, "meshes.bin"))


def _pair(gpu_api, ref_api, build, steps):
    wg, wr = build(gpu_api), build(ref_api)
    for i in range(steps):
        wg.step(DT)
        wr.step(DT)
    assert wg.hash() == wr.hash()
    return wg, wr


def _rays(rng, n, half, top):
    out = []
    for _ in range(n):
        o = (float(rng.uniform(-half, half)), float(rng.uniform(top * 0.5, top)),
             float(rng.uniform(-half, half)))
        d = (float(rng.uniform(-half, half)) - o[0], -float(rng.uniform(top * 0.6, top * 1.5)),
             float(rng.uniform(-half, half)) - o[2])
        out.append((o, tuple(np.float32(x) for x in d)))
    return out


def _movers(hull_bank):
    return [P.geometry_sphere(0.3), P.geometry_box(0.4, 0.25, 0.3), P.geometry_capsule(0.35, 0.2),
            P.geometry_hull(hull_bank.ptrs[7])]


def _filters(api, w):
    f_all = api.query_filter_default()
    f_dyn = api.query_filter_default()
    f_dyn.body_types = 1 << P.BODY_DYNAMIC
    f_static = api.query_filter_default()
    f_static.body_types = 1 << P.BODY_STATIC
    f_ignore = api.query_filter_default()
    f_ignore.ignore = w.bodies[7]
    return [None, f_all, f_dyn, f_static, f_ignore]


def test_world_raycasts_match_the_reference(gpu_api, ref_api, hull_bank):
    build = lambda a: S.bin_drop(a, 200, 17, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    wg, wr = _pair(gpu_api, ref_api, build, 60)
    P.bind_query(gpu_api)
    P.bind_query(ref_api)
    rng = np.random.default_rng(5)
    fg, fr = _filters(gpu_api, wg), _filters(ref_api, wr)
    hits = 0
    for k, (o, d) in enumerate(_rays(rng, 160, 4.5, 12.0)):
        for fi in range(len(fg)):
            for answer in (1.0, "fraction"):
                got = P.world_raycast_all(gpu_api, wg.ptr, o, d, fg[fi], answer)
                want = P.world_raycast_all(ref_api, wr.ptr, o, d, fr[fi], answer)
                assert got == want, f"ray {k} filter {fi} answer {answer}"
                hits += len(got)
        hg, hr = P.QueryHit(), P.QueryHit()
        bg = gpu_api.world_raycast_closest(wg.ptr, P.DVec3(*o), P.Vec3(*d), None, C.byref(hg))
        br = ref_api.world_raycast_closest(wr.ptr, P.DVec3(*o), P.Vec3(*d), None, C.byref(hr))
        assert bg == br and (not bg or P._hit_tuple(hg) == P._hit_tuple(hr)), f"closest {k}"
    assert hits > 500
    # A callback that ends the walk after two hits.
    o, d = (0.0, 11.0, 0.0), (0.3, -12.0, 0.2)
    assert P.world_raycast_all(gpu_api, wg.ptr, o, d, stop_after=2) == \
        P.world_raycast_all(ref_api, wr.ptr, o, d, stop_after=2)


def test_world_shape_casts_match_the_reference(gpu_api, ref_api, hull_bank):
    build = lambda a: S.bin_drop(a, 150, 23, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    wg, wr = _pair(gpu_api, ref_api, build, 60)
    rng = np.random.default_rng(9)
    hits = 0
    for k, (o, d) in enumerate(_rays(rng, 40, 4.0, 11.0)):
        q = S.rotation(S.Pcg32(100 + k))
        for mi, g in enumerate(_movers(hull_bank)):
            for margin in (0.0, 0.05):
                for answer in (1.0, "fraction"):
                    got = P.world_cast_all(gpu_api, wg.ptr, g, o, q, d, margin, None, answer)
                    want = P.world_cast_all(ref_api, wr.ptr, g, o, q, d, margin, None, answer)
                    assert got == want, f"cast {k} mover {mi} margin {margin} answer {answer}"
                    hits += len(got)
    assert hits > 200


def test_world_overlaps_and_contacts_match_the_reference(gpu_api, ref_api, hull_bank):
    build = lambda a: S.bin_drop(a, 200, 29, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    wg, wr = _pair(gpu_api, ref_api, build, 90)
    rng = np.random.default_rng(13)
    found = 0
    manifolds = 0
    for k in range(60):
        p = (float(rng.uniform(-3.5, 3.5)), float(rng.uniform(0.2, 3.0)),
             float(rng.uniform(-3.5, 3.5)))
        q = S.rotation(S.Pcg32(300 + k))
        for mi, g in enumerate(_movers(hull_bank)):
            got = P.world_overlaps(gpu_api, wg.ptr, g, p, q)
            want = P.world_overlaps(ref_api, wr.ptr, g, p, q)
            assert got == want, f"overlap {k} shape {mi}"
            found += len(got)
            for spec in (0.0, 0.02):
                cg = P.world_contacts_with(gpu_api, wg.ptr, g, p, q, spec)
                cr = P.world_contacts_with(ref_api, wr.ptr, g, p, q, spec)
                assert cg == cr, f"collide_shape {k} shape {mi} speculative {spec}"
                manifolds += len(cg)
        e = float(rng.uniform(0.2, 1.5))
        box = ((p[0] - e, p[1] - e, p[2] - e), (p[0] + e, p[1] + e, p[2] + e))
        assert P.world_overlaps(gpu_api, wg.ptr, box=box) == P.world_overlaps(ref_api, wr.ptr,
                                                                             box=box)
    assert found > 100 and manifolds > 50


def test_queries_over_meshes_match_the_reference(gpu_api, ref_api, hull_bank, mesh_bank):
    build = lambda a: S.mesh_terrain_drop(a, mesh_bank, 120, 41, hull_bank)
    wg, wr = _pair(gpu_api, ref_api, build, 45)
    rng = np.random.default_rng(21)
    tri_hits = 0
    for k, (o, d) in enumerate(_rays(rng, 120, 14.0, 14.0)):
        got = P.world_raycast_all(gpu_api, wg.ptr, o, d)
        want = P.world_raycast_all(ref_api, wr.ptr, o, d)
        assert got == want, f"ray {k}"
        tri_hits += sum(1 for h in got if h[9] != P.NULL32)
    assert tri_hits > 50
    for k in range(40):
        p = (float(rng.uniform(-12.0, 12.0)), float(rng.uniform(0.0, 2.5)),
             float(rng.uniform(-12.0, 12.0)))
        q = S.rotation(S.Pcg32(500 + k))
        for mi, g in enumerate(_movers(hull_bank)):
            assert P.world_overlaps(gpu_api, wg.ptr, g, p, q) == \
                P.world_overlaps(ref_api, wr.ptr, g, p, q), f"overlap {k} shape {mi}"
        box = ((p[0] - 1.0, p[1] - 1.0, p[2] - 1.0), (p[0] + 1.0, p[1] + 1.0, p[2] + 1.0))
        assert P.world_overlaps(gpu_api, wg.ptr, box=box) == P.world_overlaps(ref_api, wr.ptr,
                                                                             box=box)


def test_single_geometry_casts_match_the_reference(gpu_api, ref_api, hull_bank):
    """nph_physics_raycast, _shape_cast_margin and _overlap on seeded pairs."""
    P.bind_query(gpu_api)
    P.bind_query(ref_api)
    rng = np.random.default_rng(31)
    geoms = _movers(hull_bank) + [P.geometry_hull(hull_bank.ptrs[2]), P.geometry_sphere(0.8)]
    casts = 0
    for k in range(300):
        ga = geoms[k % len(geoms)]
        gb = geoms[(k * 7 + 3) % len(geoms)]
        qa, qb = S.rotation(S.Pcg32(k)), S.rotation(S.Pcg32(k + 1000))
        pa = P.Pose(P.Vec3(0.0, 0.0, 0.0), P.Quat(*qa))
        pb = P.Pose(P.Vec3(*[float(x) for x in rng.uniform(-1.2, 1.2, 3)]), P.Quat(*qb))
        d = P.Vec3(*[float(x) for x in rng.uniform(-3.0, 3.0, 3)])
        hg, hr = P.CastHit(), P.CastHit()
        bg = gpu_api.raycast(C.byref(gb), pb, P.Vec3(-1.5, 0.3, 0.1), d, C.byref(hg))
        br = ref_api.raycast(C.byref(gb), pb, P.Vec3(-1.5, 0.3, 0.1), d, C.byref(hr))
        assert bg == br and (not bg or bytes(hg) == bytes(hr)), f"ray {k}"
        for margin in (0.0, 0.03):
            hg, hr = P.CastHit(), P.CastHit()
            bg = gpu_api.shape_cast_margin(C.byref(ga), pa, d, margin, C.byref(gb), pb, C.byref(hg))
            br = ref_api.shape_cast_margin(C.byref(ga), pa, d, margin, C.byref(gb), pb, C.byref(hr))
            assert bg == br and (not bg or bytes(hg) == bytes(hr)), f"cast {k} margin {margin}"
            casts += int(bg)
        assert gpu_api.overlap(C.byref(ga), pa, C.byref(gb), pb) == \
            ref_api.overlap(C.byref(ga), pa, C.byref(gb), pb), f"overlap {k}"
    assert casts > 50
