# SPDX-License-Identifier: Apache-2.0
"""Debug drawing of single shapes (nph_physics_geometry_draw, world.h:320-325,
world.c:700-838) against the reference's own (oracle/_ref): every line's
ends and colour, in call order, for spheres, capsules, boxes, hulls and
meshes at seeded poses, bases far from the draw's origin included. The
line count of each type is also fixed by construction: three 16-segment
circles for a sphere, four circles and four sides for a capsule, twelve
box edges, one line per hull edge pair and three per mesh triangle."""
import ctypes as C
import os

import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def libs(ref_api):
    meshes = S.load_mesh_bank(os.path.join(GOLDEN, "meshes.bin"))
    return P.bind_draw(P.product_api()), P.bind_draw(ref_api), meshes


def _lines(api, geom, base, pose, rgba, origin):
    rec = P.LineRecorder(origin)
    api.geometry_draw(C.byref(rec.draw), C.byref(geom), base, pose, rgba)
    return rec.lines


def test_geometry_outlines_match_the_reference(libs, hull_bank):
    g, f, meshes = libs
    r = S.Pcg32(411)
    counts = {}
    for i in range(150):
        k = i % 5
        if k == 0:
            geom = P.geometry_sphere(float(r.next_range(0.1, 1.0)))
            want = 3 * 16
        elif k == 1:
            geom = P.geometry_capsule(float(r.next_range(0.0, 0.8)), float(r.next_range(0.1, 0.5)))
            want = 4 * 16 + 4
        elif k == 2:
            geom = P.geometry_box(*(float(r.next_range(0.1, 0.8)) for _ in range(3)))
            want = 12
        elif k == 3:
            hull = hull_bank.ptrs[r.next_bounded(len(hull_bank.ptrs))]
            geom = P.geometry_hull(hull)
            want = C.cast(hull, C.POINTER(C.c_uint32))[2] // 2  # the block's edge_count
        else:
            mesh = meshes.ptrs[r.next_bounded(len(meshes.ptrs))]
            geom = P.geometry_mesh(mesh)
            want = 3 * C.cast(mesh, C.POINTER(P.Mesh)).contents.triangle_count
        far = 1e6 if i % 3 == 0 else 10.0
        base = P.DVec3(*(float(r.next_range(-far, far)) for _ in range(3)))
        origin = tuple(float(r.next_range(-far, far)) for _ in range(3))
        pose = P.Pose(P.Vec3(*(float(r.next_range(-2.0, 2.0)) for _ in range(3))),
                      P.Quat(*S.rotation(r)))
        rgba = 0xFF000000 | r.next_bounded(1 << 24)
        mine = _lines(g, geom, base, pose, rgba, origin)
        theirs = _lines(f, geom, base, pose, rgba, origin)
        assert mine == theirs, (i, k)
        assert len(mine) == want, (i, k, len(mine), want)
        counts[k] = counts.get(k, 0) + len(mine)
    assert len(counts) == 5
