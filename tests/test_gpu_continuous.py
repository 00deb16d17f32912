# SPDX-License-Identifier: Apache-2.0
"""Continuous collision on the GPU path (continuous.c:1-409): fast bodies
marked in the device's finalize, swept on the host against static shapes
and the triangles of static meshes (bullets also against bodies at their
final poses), set back to the first impact, their boxes refreshed before
the grown proxies; bit for bit against the reference's world hashes."""
import json
import os

import pytest

from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _runs():
    with open(os.path.join(GOLDEN, "continuous_hashes.json")) as f:
        return json.load(f)["runs"]


def _check(w, want, name):
    swept = impacts = 0
    for i, h in enumerate(want):
        w.step(1.0 / 60.0)
        st = w.stats()
        swept += st.swept_bodies
        impacts += st.impact_bodies
        assert "%016x" % w.hash() == h, f"{name}: hash differs at step {i + 1}"
    return swept, impacts


def test_projectiles_through_a_thin_wall_bit_exact(gpu_api):
    print("projectiles, 40 shots, 120 steps", flush=True)
    swept, impacts = _check(S.projectiles(gpu_api, 40, 5), _runs()["projectiles_40"],
                            "projectiles_40")
    assert swept > 100 and impacts > 50  # the sweeps and the set-backs both ran


def test_bin_drop_from_30_m_with_continuous_bit_exact(gpu_api, hull_bank):
    print("bin drop 300 bodies from 30 m, continuous on, 200 steps", flush=True)
    w = S.bin_drop(gpu_api, 300, 21, hull_bank, half=4.0, height=30.0, y_lo=1.0,
                   desc={"enable_continuous": 1})
    swept, _ = _check(w, _runs()["bin_300_21_h30"], "bin_300_21_h30")
    assert swept > 1000


def test_projectiles_match_reference_stats(gpu_api, ref_api):
    wg = S.projectiles(gpu_api, 40, 9)
    wr = S.projectiles(ref_api, 40, 9)
    for i in range(60):
        wg.step(1.0 / 60.0)
        wr.step(1.0 / 60.0)
        sg, sr = wg.stats(), wr.stats()
        assert (sg.swept_bodies, sg.impact_bodies) == (sr.swept_bodies, sr.impact_bodies)
        assert wg.hash() == wr.hash(), f"differs at step {i + 1}"


def _mesh_shots(api, meshes, hulls):
    """Fast mixed bodies (including hulls) thrown down at 40-90 m/s onto a
    one-sided terrain and a two-sided patch, spinning, every fourth a bullet;
    a few slow ones resting there as bullet targets."""
    from nepheline_amd import physics as P
    w = P.World(api, dict(S.DESC, enable_sleep=0, enable_continuous=1))
    ground = w.add_body(P.BODY_STATIC, (0.0, 0.0, 0.0))
    w.add_shape(ground, P.geometry_mesh(meshes.ptrs[0]), friction=0.5)
    patch = w.add_body(P.BODY_STATIC, (20.0, 4.0, 0.0), S.quat_axis_angle((1.0, 0.0, 0.0), 3.0))
    w.add_shape(patch, P.geometry_mesh(meshes.ptrs[4]), friction=0.5)
    r = S.Pcg32(17)
    S.drop_mixed(w, r, 12, hulls, -6.0, 6.0, 2.0, 4.0)
    for i in range(48):
        x = float(r.next_range(-12.0, 26.0))
        z = float(r.next_range(-12.0, 12.0))
        y = float(r.next_range(8.0, 20.0))
        kind = i % 4
        if kind == 0:
            g = P.geometry_sphere(float(r.next_range(0.05, 0.3)))
        elif kind == 1:
            g = P.geometry_capsule(float(r.next_range(0.1, 0.4)), float(r.next_range(0.05, 0.2)))
        elif kind == 2:
            g = P.geometry_box(*(float(r.next_range(0.05, 0.4)) for _ in range(3)))
        else:
            g = P.geometry_hull(hulls.ptrs[r.next_bounded(len(hulls.ptrs))])
        v = (float(r.next_range(-10.0, 10.0)), -float(r.next_range(40.0, 90.0)),
             float(r.next_range(-10.0, 10.0)))
        spin = tuple(float(r.next_range(-20.0, 20.0)) for _ in range(3))
        b = w.add_body(P.BODY_DYNAMIC, (x, y, z), S.rotation(r), velocity=v,
                       angular_velocity=spin, bullet=1 if i % 8 == 3 else 0)
        w.add_shape(b, g, friction=0.5)
    w._keep.extend([meshes, hulls])
    return w


def test_fast_bodies_onto_meshes_match_the_reference(gpu_api, ref_api, hull_bank):
    """The sweeps against mesh triangles (continuous.c:211-300): one-sided
    terrain and an upside-down two-sided patch, 120 steps, the sweep and
    impact counts and the world hash equal to the reference's every step."""
    meshes = S.load_mesh_bank(os.path.join(GOLDEN, "meshes.bin"))
    print("continuous onto meshes: 60 bodies, 120 steps against the reference", flush=True)
    wg = _mesh_shots(gpu_api, meshes, hull_bank)
    wr = _mesh_shots(ref_api, meshes, hull_bank)
    swept = impacts = 0
    for i in range(120):
        wg.step(1.0 / 60.0)
        wr.step(1.0 / 60.0)
        sg, sr = wg.stats(), wr.stats()
        assert (sg.swept_bodies, sg.impact_bodies) == (sr.swept_bodies, sr.impact_bodies), i
        assert wg.hash() == wr.hash(), f"differs at step {i + 1}"
        swept += sg.swept_bodies
        impacts += sg.impact_bodies
    assert swept > 40 and impacts > 20


def _compound_shots(api):
    """Fast compound bodies of 70 and 100 small shapes (more than 64, the
    sweep's stack list) fired at the thin wall, one a bullet."""
    from nepheline_amd import physics as P
    w = S.projectiles(api, 4, 13)
    for k, (count, bullet) in enumerate(((70, 0), (100, 1))):
        b = w.add_body(P.BODY_DYNAMIC, (-1.0 + 2.0 * k, 2.0, -1.0), velocity=(0.0, 0.0, 120.0),
                       bullet=bullet)
        for i in range(count):
            offset = (0.02 * (i % 5) - 0.04, 0.02 * ((i // 5) % 5) - 0.04, 0.02 * (i // 25))
            w.add_shape(b, P.geometry_sphere(0.02), pose=(offset, (0.0, 0.0, 0.0, 1.0)),
                        density=100.0)
    return w


def test_compound_fast_bodies_match_the_reference(gpu_api, ref_api):
    print("compound fast bodies of 70 and 100 shapes", flush=True)
    wg = _compound_shots(gpu_api)
    wr = _compound_shots(ref_api)
    swept = 0
    for i in range(40):
        wg.step(1.0 / 60.0)
        wr.step(1.0 / 60.0)
        swept += wg.stats().swept_bodies
        assert wg.hash() == wr.hash(), f"differs at step {i + 1}"
    assert swept > 10
