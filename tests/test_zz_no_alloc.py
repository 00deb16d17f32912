# SPDX-License-Identifier: Apache-2.0
"""nph_amd_collide_many_static, the pair query that allocates nothing (its
pairs, results and hull blocks live in one fixed device array), against
nph_amd_collide_many, bit for bit: two calls in a row, the second smaller,
each with caches carried in; hulls of up to 256 vertices (the large-hull
kernel) among them; and its refusals, made before the device is touched."""
import ctypes as C
import math

import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

pytestmark = pytest.mark.gpu


def _queries(seed, n, banks):
    r = S.Pcg32(seed)
    q = (P.CollideQuery * n)()
    for i in range(n):
        geoms = []
        for kind in (i % 5, (i // 5) % 5):
            if kind == 0:
                geoms.append((P.geometry_sphere(float(r.next_range(0.2, 0.6))), 0.6))
            elif kind == 1:
                rad, hh = float(r.next_range(0.2, 0.4)), float(r.next_range(0.3, 0.8))
                geoms.append((P.geometry_capsule(hh, rad), hh + rad))
            elif kind == 2:
                h = [float(r.next_range(0.2, 0.7)) for _ in range(3)]
                geoms.append((P.geometry_box(*h), math.sqrt(sum(x * x for x in h))))
            else:
                # At most 100 + 3 distinct blocks: a call takes 256.
                bank = banks[kind - 3]
                pick = r.next_bounded(min(len(bank.ptrs), 100))
                geoms.append((P.geometry_hull(bank.ptrs[pick]), 0.62))
        (ga, ra), (gb, rb) = geoms
        pa = P.Pose(P.Vec3(*(float(r.next_range(-1.0, 1.0)) for _ in range(3))),
                    P.Quat(*S.rotation(r)))
        d = [float(r.next_range(-1.0, 1.0)) for _ in range(3)]
        norm = math.sqrt(sum(x * x for x in d)) or 1.0
        dist = float(r.next_range(0.3, 1.15)) * (ra + rb)
        pb = P.Pose(P.Vec3(pa.position.x + d[0] / norm * dist, pa.position.y + d[1] / norm * dist,
                           pa.position.z + d[2] / norm * dist), P.Quat(*S.rotation(r)))
        q[i].a, q[i].pose_a, q[i].b, q[i].pose_b = ga, pa, gb, pb
        q[i].speculative_distance = float(r.next_range(0.0, 0.08))
    return q


def _bytes(ms, caches):
    out = []
    for m, c in zip(ms, caches):
        out.append(bytes(m)[:16 + 20 * m.point_count] + bytes(c))
    return out


@pytest.fixture(scope="module")
def banks(gpu_api, hull_bank):
    return (hull_bank, S.built_hull_bank(gpu_api, (33, 64, 256), seed=21))


def test_static_queries_equal_the_allocating_ones(gpu_api, banks):
    g = P.bind_collide(gpu_api)
    for call, (seed, n) in enumerate(((71, 3000), (72, 700))):
        print(f"no-alloc: call {call + 1}, seed {seed}, {n} pairs", flush=True)
        q = _queries(seed, n, banks)
        want_m, got_m = (P.Manifold * n)(), (P.Manifold * n)()
        want_c, got_c = (P.CollideCache * n)(), (P.CollideCache * n)()
        for rnd in range(2):  # the second round starts from the first's caches
            assert g.collide_many(q, n, want_m, want_c) == P.OK
            assert g.collide_many_static(q, n, got_m, got_c) == P.OK
            want, got = _bytes(want_m, want_c), _bytes(got_m, got_c)
            bad = [i for i in range(n) if want[i] != got[i]]
            assert not bad, f"call {call + 1} round {rnd + 1}: pairs {bad[:10]} differ"
        touching = sum(1 for m in got_m if m.point_count > 0)
        assert touching > n // 5
    # Without caches: the same manifolds as with fresh zero caches.
    q = _queries(73, 200, banks)
    a, b = (P.Manifold * 200)(), (P.Manifold * 200)()
    assert g.collide_many(q, 200, a, None) == P.OK
    assert g.collide_many_static(q, 200, b, None) == P.OK
    assert [bytes(x)[:16 + 20 * x.point_count] for x in a] == \
        [bytes(x)[:16 + 20 * x.point_count] for x in b]


def test_static_queries_refuse_what_does_not_fit(gpu_api, hull_bank):
    g = P.bind_collide(gpu_api)
    # Over the 16 MiB array: 100 000 sphere pairs need 22 MB.
    n = 100_000
    q = (P.CollideQuery * n)()
    for i in range(n):
        q[i].a = P.geometry_sphere(0.5)
        q[i].b = P.geometry_sphere(0.5)
        q[i].pose_a = P.Pose(P.Vec3(0.0, 0.0, 0.0), P.Quat(0.0, 0.0, 0.0, 1.0))
        q[i].pose_b = P.Pose(P.Vec3(0.9, 0.0, 0.0), P.Quat(0.0, 0.0, 0.0, 1.0))
    out = (P.Manifold * n)()
    assert g.collide_many_static(q, n, out, None) == P.INVALID_ARGUMENT
    # 16 000 pairs fit, and after the refusal the array still serves.
    assert g.collide_many_static(q, 16_000, out, None) == P.OK
    assert all(out[i].point_count == 1 for i in range(16_000))
    # A mesh is not a convex type.
    q[0].b = P.geometry_mesh(None)
    assert g.collide_many_static(q, 1, out, None) == P.INVALID_ARGUMENT
    # A hull with 100-gon faces is taken (the wide-clip build), as by the
    # allocating query.
    import numpy as np
    seg = 100
    ang = 2.0 * np.pi * np.arange(seg) / seg
    pts = np.zeros((2 * seg, 3), np.float32)
    pts[0::2] = np.stack([np.cos(ang), np.full(seg, -0.5), np.sin(ang)], 1)
    pts[1::2] = np.stack([np.cos(ang), np.full(seg, 0.5), np.sin(ang)], 1)
    rc, block = P.build_hull(gpu_api, pts, 256)
    assert rc == 0
    wide = S.HullBank(block, [0])
    q[0].a = P.geometry_hull(wide.ptrs[0])
    q[0].b = P.geometry_hull(wide.ptrs[0])
    q[0].pose_b = P.Pose(P.Vec3(0.05, 0.98, 0.0), P.Quat(0.0, 0.0871557, 0.0, 0.9961947))
    q[0].speculative_distance = 0.02
    assert g.collide_many_static(q, 1, out, None) == P.OK
    alloc = (P.Manifold * 1)()
    assert g.collide_many(q, 1, alloc, None) == P.OK
    assert out[0].point_count == 4
    assert bytes(out[0]) == bytes(alloc[0])
    # More than 256 distinct hull blocks in one call.
    many = S.built_hull_bank(gpu_api, [8] * 260, seed=23)
    for i in range(260):
        q[i].a = P.geometry_hull(many.ptrs[i])
        q[i].b = P.geometry_sphere(0.5)
    assert g.collide_many_static(q, 260, out, None) == P.INVALID_ARGUMENT
    assert g.collide_many_static(q, 250, out, None) == P.OK
