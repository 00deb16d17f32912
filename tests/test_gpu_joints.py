# SPDX-License-Identifier: Apache-2.0
"""Joints on the GPU (ball, hinge and fixed; joint_solver.c:32-907): the
reference's ragdolls (test_world_support.h:281-379) stepped bit for bit
against the reference's world hashes, and the joint queries (angle, length,
reaction) equal to the reference's after the same steps."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _runs():
    with open(os.path.join(GOLDEN, "ragdoll_hashes.json")) as f:
        return json.load(f)["runs"]


def _check(w, want, name):
    for i, h in enumerate(want):
        w.step(1.0 / 60.0)
        assert "%016x" % w.hash() == h, f"{name}: hash differs at step {i + 1}"


def test_ragdoll_pile_bit_exact(gpu_api):
    print("ragdoll pile of 4, seed 2026, 240 steps", flush=True)
    _check(S.ragdoll_pile(gpu_api, 4, 2026), _runs()["pile_4"], "pile_4")


def test_standing_ragdoll_bit_exact(gpu_api):
    print("one standing ragdoll, 180 steps", flush=True)
    w = P.World(gpu_api, S.DESC)
    S.ground(w, 0.0)
    S.ragdoll(w, (0.0, 0.95, 0.0))
    _check(w, _runs()["standing_1"], "standing_1")


def test_joint_queries_match_reference(gpu_api, ref_api):
    print("joint queries after 90 steps of a pile of 3", flush=True)
    wg = S.ragdoll_pile(gpu_api, 3, 7)
    wr = S.ragdoll_pile(ref_api, 3, 7)
    for _ in range(90):
        wg.step(1.0 / 60.0)
        wr.step(1.0 / 60.0)
    assert wg.hash() == wr.hash()
    for (_, jg), (_, jr) in zip(wg.ragdolls, wr.ragdolls):
        for a, b in zip(jg, jr):
            assert gpu_api.joint_length(wg.ptr, a) == ref_api.joint_length(wr.ptr, b)
            assert gpu_api.joint_hinge_angle(wg.ptr, a) == ref_api.joint_hinge_angle(wr.ptr, b)
            fg, tg, fr, tr = P.Vec3(), P.Vec3(), P.Vec3(), P.Vec3()
            gpu_api.joint_reaction(wg.ptr, a, C.byref(fg), C.byref(tg))
            ref_api.joint_reaction(wr.ptr, b, C.byref(fr), C.byref(tr))
            assert (fg.x, fg.y, fg.z, tg.x, tg.y, tg.z) == (fr.x, fr.y, fr.z, tr.x, tr.y, tr.z)
    assert wg.stats().joints == wr.stats().joints == 30


def test_joint_create_validates_like_the_reference(gpu_api):
    w = P.World(gpu_api, S.DESC)
    a = w.add_body(P.BODY_DYNAMIC, (0.0, 1.0, 0.0))
    w.add_shape(a, P.geometry_sphere(0.2))
    b = w.add_body(P.BODY_DYNAMIC, (0.0, 2.0, 0.0))
    w.add_shape(b, P.geometry_sphere(0.2))
    s = w.add_body(P.BODY_STATIC, (0.0, 3.0, 0.0))
    api = gpu_api
    h = C.c_uint32()
    d = api.joint_desc_default(P.JOINT_BALL, a, a)
    assert api.joint_create(w.ptr, C.byref(d), C.byref(h)) == P.INVALID_ARGUMENT
    d = api.joint_desc_default(P.JOINT_SLIDER, a, b)
    assert api.joint_create(w.ptr, C.byref(d), C.byref(h)) == P.UNSUPPORTED
    d = api.joint_desc_default(P.JOINT_HINGE, a, b)
    d.hinge.enable_limit = 1
    d.hinge.lower, d.hinge.upper = 1.0, -1.0
    assert api.joint_create(w.ptr, C.byref(d), C.byref(h)) == P.INVALID_ARGUMENT
    d = api.joint_desc_default(P.JOINT_FIXED, b, s)
    assert api.joint_create(w.ptr, C.byref(d), C.byref(h)) == P.OK
    is_valid = api.lib.nph_physics_joint_is_valid
    is_valid.restype = C.c_bool
    is_valid.argtypes = [C.c_void_p, C.c_uint32]
    assert is_valid(w.ptr, h.value)
    api.joint_destroy(w.ptr, h.value)
    assert not is_valid(w.ptr, h.value)


def test_fixed_and_hinge_chain_matches_reference(gpu_api, ref_api):
    """A chain hung from a static body: a fixed joint, then hinges with a
    spring and a motor, then balls, 120 steps, bit for bit."""
    print("joint chain, 120 steps", flush=True)
    ws = []
    for api in (gpu_api, ref_api):
        w = P.World(api, S.DESC)
        top = w.add_body(P.BODY_STATIC, (0.0, 5.0, 0.0))
        w.add_shape(top, P.geometry_box(0.2, 0.2, 0.2))
        prev = top
        for i in range(8):
            b = w.add_body(P.BODY_DYNAMIC, (0.3 * (i + 1), 5.0, 0.1 * i))
            w.add_shape(b, P.geometry_capsule(0.1, 0.05))
            at = (0.3 * i + 0.15, 5.0, 0.1 * i)
            if i == 0:
                w.add_joint(P.JOINT_FIXED, prev, b, at, (1.0, 0.0, 0.0))
            elif i < 4:
                w.add_joint(P.JOINT_HINGE, prev, b, at, (0.0, 0.0, 1.0), hinge={
                    "enable_spring": 1, "spring_hertz": 2.0, "spring_damping_ratio": 0.3,
                    "spring_target": 0.2, "enable_motor": 1, "motor_speed": 0.5,
                    "max_motor_torque": 1.0, "enable_limit": 1, "lower": -0.8, "upper": 0.8})
            else:
                w.add_joint(P.JOINT_BALL, prev, b, at, (1.0, 0.0, 0.0), ball={
                    "enable_swing_limit": 1, "swing_y": 0.5, "swing_z": 0.7,
                    "enable_twist_limit": 1, "twist_lower": -0.3, "twist_upper": 0.4})
            prev = b
        ws.append(w)
    for i in range(120):
        for w in ws:
            w.step(1.0 / 60.0)
        assert ws[0].hash() == ws[1].hash(), f"chain differs at step {i + 1}"
    np.testing.assert_array_equal(ws[0].body_state(), ws[1].body_state())
