# SPDX-License-Identifier: Apache-2.0
"""The bodies' and shapes' API between steps, against the reference: forces,
torques and impulses, teleports, mass, damping, gravity scale, locks, sleep
settings, surfaces, density, filters, shape and body destruction (with the
contacts and joints that go with them) and slot reuse, applied to the same
world in both libraries, then the world hash compared after every step."""
import ctypes as C

import numpy as np
import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

pytestmark = pytest.mark.gpu


def _mutate(w: P.World, r: S.Pcg32, hull_bank):
    a = w.api
    bodies = w.bodies
    shapes = w.shapes
    dyn = bodies[5:]  # the bin's floor and walls come first
    for k in range(0, 40, 4):
        b = dyn[k]
        a.body_apply_force(w.ptr, b, P.Vec3(30.0, 400.0, -20.0),
                           P.DVec3(float(k) * 0.01, 3.0, 0.0))
        a.body_apply_torque(w.ptr, dyn[k + 1], P.Vec3(0.0, 5.0, 1.0))
        a.body_apply_force_at_center(w.ptr, dyn[k + 2], P.Vec3(0.0, 0.0, 90.0))
        a.body_apply_linear_impulse(w.ptr, dyn[k + 3], P.Vec3(1.5, 2.0, 0.0),
                                    P.DVec3(0.25, 2.0, -0.5))
    a.body_apply_angular_impulse(w.ptr, dyn[41], P.Vec3(0.2, 0.0, -0.3))
    a.body_set_transform(w.ptr, dyn[42], P.DVec3(0.5, 9.0, -0.5), P.Quat(0.0, 0.38268343, 0.0,
                                                                            0.92387953))
    a.body_set_gravity_scale(w.ptr, dyn[43], 0.25)
    a.body_set_damping(w.ptr, dyn[44], 0.5, 2.0)
    a.body_set_locks(w.ptr, dyn[45], 0x38)  # angular x, y, z
    m = P.MassData()
    m.mass = 3.0
    m.center = P.Vec3(0.0, 0.05, 0.0)
    m.inertia.c[0] = P.Vec3(0.2, 0.0, 0.0)
    m.inertia.c[1] = P.Vec3(0.0, 0.3, 0.0)
    m.inertia.c[2] = P.Vec3(0.0, 0.0, 0.25)
    a.body_set_mass_data(w.ptr, dyn[46], C.byref(m))
    a.body_set_sleep(w.ptr, dyn[47], False, 0.1)
    a.body_set_awake(w.ptr, dyn[48], True)
    a.body_set_bullet(w.ptr, dyn[49], True)
    a.shape_set_friction(w.ptr, shapes[5 + 50], 0.05)
    a.shape_set_restitution(w.ptr, shapes[5 + 51], 0.8)
    a.shape_set_friction_combine(w.ptr, shapes[5 + 52], 4)
    a.shape_set_density(w.ptr, shapes[5 + 53], 250.0)
    f = a.shape_filter(w.ptr, shapes[5 + 54])
    f.group = -3
    a.shape_set_filter(w.ptr, shapes[5 + 54], f)
    a.shape_set_filter(w.ptr, shapes[5 + 55], f)  # the two no longer collide
    f2 = a.shape_filter(w.ptr, shapes[5 + 56])
    f2.mask = 0
    a.shape_set_filter(w.ptr, shapes[5 + 56], f2)  # collides with nothing
    a.shape_destroy(w.ptr, shapes[5 + 57])
    gone = [dyn[k] for k in (60, 61, 62, 70, 90)]
    for b in gone:
        a.body_destroy(w.ptr, b)
        w.bodies.remove(b)  # the reference asserts on stale handles
    # New bodies take the freed slots (LIFO), as in the reference.
    S.drop_mixed(w, r, 8, hull_bank, -2.0, 2.0, 6.0, 9.0)


def _run(api, hull_bank, steps_before=40, steps_after=100):
    w = S.bin_drop(api, 200, 13, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    hashes = []
    for _ in range(steps_before):
        w.step(1.0 / 60.0)
    _mutate(w, S.Pcg32(77), hull_bank)
    for _ in range(steps_after):
        w.step(1.0 / 60.0)
        hashes.append("%016x" % w.hash())
    return w, hashes


def test_mutations_between_steps_match_the_reference(gpu_api, ref_api, hull_bank):
    print("api: mutations after 40 steps of a 200-body drop, 100 steps compared", flush=True)
    wg, hg = _run(gpu_api, hull_bank)
    wr, hr = _run(ref_api, hull_bank)
    for i, (a, b) in enumerate(zip(hg, hr)):
        assert a == b, f"hash differs at step {i + 1} after the mutations"
    np.testing.assert_array_equal(wg.body_state(), wr.body_state())


def test_getters_report_the_settings(gpu_api, hull_bank):
    w = S.bin_drop(gpu_api, 20, 14, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    a = gpu_api
    s = w.shapes[6]
    a.shape_set_friction(w.ptr, s, 0.25)
    a.shape_set_friction_combine(w.ptr, s, 2)
    a.shape_set_density(w.ptr, s, 123.0)
    assert a.shape_friction(w.ptr, s) == pytest.approx(0.25)
    assert a.shape_friction_combine(w.ptr, s) == 2
    assert a.shape_density(w.ptr, s) == pytest.approx(123.0)
    b = w.bodies[6]
    a.body_set_user_data(w.ptr, b, 0xDEADBEEF)
    assert a.body_user_data(w.ptr, b) == 0xDEADBEEF
    out = (C.c_uint32 * 4)()
    assert a.body_shapes(w.ptr, b, out, 4) == 1 and out[0] == w.shapes[6]
    a.body_set_bullet(w.ptr, b, True)
    assert a.body_is_bullet(w.ptr, b)
    a.body_destroy(w.ptr, b)
    a.lib.nph_physics_body_is_valid.restype = C.c_bool
    a.lib.nph_physics_body_is_valid.argtypes = [C.c_void_p, C.c_uint32]
    assert not a.lib.nph_physics_body_is_valid(w.ptr, b)
    w.step(1.0 / 60.0)
    w.close()


def _eventful(api, hull_bank):
    """A small bin drop whose shapes all ask for contact and hit events."""
    w = P.World(api, S.DESC)
    add = w.add_shape
    w.add_shape = lambda b, g, **kw: add(b, g, hit_events=1, **kw)
    floor = w.add_body(P.BODY_STATIC, (0.0, -1.0, 0.0))
    w.add_shape(floor, P.geometry_box(5.0, 1.0, 5.0), friction=0.5)
    S.drop_mixed(w, S.Pcg32(15), 120, hull_bank, -3.0, 3.0, 1.0, 8.0)
    api.world_set_hit_threshold(w.ptr, 0.5)
    return w


def test_events_match_the_reference(gpu_api, ref_api, hull_bank):
    """Begin, end, hit and move events of every step, in order, bit for bit."""
    print("api: events of a 120-body drop with hit events, 150 steps", flush=True)
    wg = _eventful(gpu_api, hull_bank)
    wr = _eventful(ref_api, hull_bank)
    hits = begins = 0
    for i in range(150):
        wg.step(1.0 / 60.0)
        wr.step(1.0 / 60.0)
        eg, er = wg.events(), wr.events()
        for kind, a, b in zip(("begin", "end", "hit", "move"), eg, er):
            assert a == b, f"{kind} events differ at step {i + 1}"
        begins += len(eg[0])
        hits += len(eg[2])
    assert begins > 100 and hits > 20
