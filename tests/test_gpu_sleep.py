# SPDX-License-Identifier: Apache-2.0
"""Islands and sleep on the GPU against the reference's step with sleep on
(engine/physics/src/island.c:139-735, body.c:132-180 and :567-593,
solver.c:845-910, world.c:351-359).

Islands only show through sleep: which bodies leave the awake list and in
what order (the solver's body order), which contacts leave the graph, and
the move events' fell_asleep marks.  Every test steps the HIP world and the
reference's world side by side and checks, after every step, the world hash
(engine/physics/src/world.c:436-462), every event list, the awake flag of
every body and the island counts; the bookkeeping's invariants hold
throughout (nph_physics_world_validate)."""
import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

pytestmark = pytest.mark.gpu

SLEEP = {"enable_sleep": 1, "enable_continuous": 0}
DT = 1.0 / 60.0


def _state(w):
    s = w.stats()
    awake = tuple(bool(w.api.body_is_awake(w.ptr, h)) for h in w.bodies)
    return (s.awake_bodies, s.islands, s.awake_islands, s.touching_contacts), awake


def _lockstep(wg, wr, steps, action=None, label=""):
    """Steps both worlds `steps` times; `action(i, w)` runs on each before
    step i. Returns the number of fell_asleep marks seen."""
    asleep = 0
    for i in range(steps):
        if action is not None:
            action(i, wg)
            action(i, wr)
        wg.step(DT)
        wr.step(DT)
        where = f"{label} step {i + 1}"
        assert "%016x" % wg.hash() == "%016x" % wr.hash(), f"{where}: hash"
        eg, er = wg.events(), wr.events()
        for k, name in enumerate(("begin", "end", "hit", "move")):
            assert eg[k] == er[k], f"{where}: {name} events"
        asleep += sum(1 for m in eg[3] if m[1])
        assert _state(wg) == _state(wr), f"{where}: awake flags or island counts"
        assert wg.api.world_validate(wg.ptr), f"{where}: invariants"
    return asleep


def _pair(api_g, api_r, build):
    return build(api_g), build(api_r)


def test_default_world_desc_is_accepted(gpu_api):
    """The reference's defaults (sleep and continuous collision on) make a
    world, as they do in the reference."""
    w = P.World(gpu_api, {})
    assert w.desc.enable_sleep == 1 and w.desc.enable_continuous == 1
    b = w.add_body(P.BODY_DYNAMIC, (0.0, 1.0, 0.0))
    w.add_shape(b, P.geometry_sphere(0.5))
    w.step(DT)
    w.close()


def test_pyramid_falls_asleep_and_a_falling_ball_wakes_it(gpu_api, ref_api):
    wg, wr = _pair(gpu_api, ref_api, lambda a: S.pyramid(a, rows=8, desc=SLEEP))
    asleep = _lockstep(wg, wr, 150, label="settle")
    assert asleep == 36, "every box falls asleep once"
    assert wg.stats().awake_bodies == 0

    def drop(_, w):
        b = w.add_body(P.BODY_DYNAMIC, (0.2, 10.0, 0.0))
        w.add_shape(b, P.geometry_sphere(0.4), density=3000.0)
    drop(0, wg)
    drop(0, wr)
    _lockstep(wg, wr, 150, label="impact")


def test_bin_drop_merges_splits_and_sleeps(gpu_api, ref_api, hull_bank):
    build = lambda a: S.bin_drop(a, 150, 11, hull_bank, half=4.0, height=10.0, y_lo=1.0,
                                 desc=SLEEP)
    wg, wr = _pair(gpu_api, ref_api, build)
    asleep = _lockstep(wg, wr, 420, label="bin")
    assert asleep > 0


def test_ragdoll_pile_sleeps_through_its_joints(gpu_api, ref_api):
    wg, wr = _pair(gpu_api, ref_api, lambda a: S.ragdoll_pile(a, count=3, desc=SLEEP))
    asleep = _lockstep(wg, wr, 360, label="ragdolls")
    assert asleep > 0


def test_wake_and_sleep_calls(gpu_api, ref_api):
    """set_awake both ways, a velocity set on a sleeper, a body destroyed out
    of a sleeping pile, a teleport, sleep turned off for one body, and a body
    made asleep, between steps."""
    def build(a):
        w = S.pyramid(a, rows=5, desc=SLEEP)
        z = w.add_body(P.BODY_DYNAMIC, (6.0, 0.5, 0.0), awake=0)
        w.add_shape(z, P.geometry_box(0.5, 0.5, 0.5))
        return w
    wg, wr = _pair(gpu_api, ref_api, build)

    def act(i, w):
        a, p, bodies = w.api, w.ptr, w.bodies
        if i == 20:
            a.body_set_awake(p, bodies[15], False)  # the top box: its island sleeps now
        elif i == 45:
            a.body_set_linear_velocity(p, bodies[3], P.Vec3(0.0, 2.0, 0.0))
        elif i == 120:
            a.body_destroy(p, bodies[4])
        elif i == 150:
            a.body_set_transform(p, bodies[15], P.DVec3(0.0, 7.0, 0.0), P.Quat(0.0, 0.0, 0.0, 1.0))
        elif i == 160:
            a.body_set_sleep(p, bodies[1], False, 0.05)
        elif i == 200:
            a.body_set_awake(p, bodies[16], True)  # the one made asleep
    _lockstep(wg, wr, 260, action=act, label="calls")


def test_kinematic_platform_wakes_its_sleepers_and_motors_keep_awake(gpu_api, ref_api):
    """A kinematic platform that falls asleep with a box on it, then moves
    (it wakes the box: solver.c:893-909); two boxes on a driven hinge never
    sleep (solver.c:850-857), two on an idle one do."""
    def build(a):
        w = S.pyramid(a, rows=1, desc=SLEEP)
        k = w.add_body(P.BODY_KINEMATIC, (4.0, 0.25, 0.0))
        w.add_shape(k, P.geometry_box(1.5, 0.25, 1.5))
        b = w.add_body(P.BODY_DYNAMIC, (4.0, 1.0, 0.0))
        w.add_shape(b, P.geometry_box(0.5, 0.5, 0.5))
        for x, motor in ((-4.0, 1), (-8.0, 0)):
            p = w.add_body(P.BODY_DYNAMIC, (x, 0.5, 0.0))
            w.add_shape(p, P.geometry_box(0.5, 0.5, 0.5))
            q = w.add_body(P.BODY_DYNAMIC, (x, 1.5, 0.0))
            w.add_shape(q, P.geometry_box(0.3, 0.5, 0.3))
            w.add_joint(P.JOINT_HINGE, p, q, (x, 1.0, 0.0), (0.0, 1.0, 0.0),
                        hinge={"enable_motor": motor, "motor_speed": 1.5,
                               "max_motor_torque": 20.0})
        return w
    wg, wr = _pair(gpu_api, ref_api, build)

    def act(i, w):
        if i == 90:
            assert not w.api.body_is_awake(w.ptr, w.bodies[3]), "the rider sleeps first"
            w.api.body_set_linear_velocity(w.ptr, w.bodies[2], P.Vec3(0.5, 0.0, 0.0))
        elif i == 150:
            w.api.body_set_linear_velocity(w.ptr, w.bodies[2], P.Vec3(0.0, 0.0, 0.0))
    _lockstep(wg, wr, 260, action=act, label="kinematic")
    driven, idle = wg.bodies[4:6], wg.bodies[6:8]
    assert all(wg.api.body_is_awake(wg.ptr, h) for h in driven)
    assert not any(wg.api.body_is_awake(wg.ptr, h) for h in idle)
