# SPDX-License-Identifier: Apache-2.0
"""Seeded scenes of the world-step workloads, built through the C API so the
same calls drive this package's library and, in tests, the reference's.

* pyramid: a 20-row box pyramid (210 unit boxes, 1000 kg/m^3, friction 0.6)
  on a static ground box, as engine/physics/tests/test_world.c:365-379.
* bin: mixed spheres, capsules, boxes and 12-vertex hulls dropped into an
  open static bin (the 10k-body workload and smaller versions of it).
* ragdolls: the reference's eleven-body ragdoll (ball joints with cones and
  twists, hinges at elbows and knees, test_world_support.h:281-379) and its
  criss-cross pile (test_ragdoll.c:269-303).

Random draws come from PCG32 (engine/math/include/nepheline/math/random.h:
38-74, O'Neill's pcg32 XSH-RR with the reference's seeding), one draw per
statement, x before y before z.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import physics as P

_MASK64 = (1 << 64) - 1


class Pcg32:
    """pcg32 (XSH-RR 64/32), seeded as nph_pcg32_seed(state, stream)."""

    MULT = 6364136223846793005

    def __init__(self, seed: int, stream: int = 0x9E3779B97F4A7C15):
        self.state = 0
        self.inc = ((stream << 1) | 1) & _MASK64
        self.next_u32()
        self.state = (self.state + seed) & _MASK64
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * self.MULT + self.inc) & _MASK64
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def next_f32(self) -> float:
        """[0, 1) with 24 bits."""
        return np.float32((self.next_u32() >> 8) * (1.0 / 16777216.0))

    def next_range(self, lo: float, hi: float) -> float:
        t = self.next_f32()
        return np.float32(np.float32(lo) + np.float32(np.float32(hi) - np.float32(lo)) * t)

    def next_bounded(self, bound: int) -> int:
        """Unbiased [0, bound): Lemire's multiply-shift with rejection, as
        nph_pcg32_next_bounded (engine/math/src/random.c:20-38)."""
        m = self.next_u32() * bound
        low = m & 0xFFFFFFFF
        if low < bound:
            t = (-bound & 0xFFFFFFFF) % bound
            while low < t:
                m = self.next_u32() * bound
                low = m & 0xFFFFFFFF
        return m >> 32


def rotation(r: Pcg32):
    """A uniform rotation (Shoemake, Graphics Gems III), normalized in f32."""
    u1 = float(r.next_f32())
    u2 = float(r.next_f32()) * 6.2831853
    u3 = float(r.next_f32()) * 6.2831853
    a = math.sqrt(1.0 - u1)
    b = math.sqrt(u1)
    q = np.array([a * math.sin(u2), a * math.cos(u2), b * math.sin(u3), b * math.cos(u3)],
                 dtype=np.float32)
    q = q / np.float32(np.sqrt(np.sum(q.astype(np.float64) ** 2)))
    return tuple(float(x) for x in q.astype(np.float32))


DESC = {"enable_sleep": 0, "enable_continuous": 0}
DT = 1.0 / 60.0


def _world(api, desc: dict | None, world: P.World | None) -> P.World:
    """The world to fill: `world` (one a batch made) or a new one."""
    if world is not None:
        return world
    d = dict(DESC)
    d.update(desc or {})
    return P.World(api, d)


def pyramid(api, rows: int = 20, desc: dict | None = None, world: P.World | None = None,
            ground_half: float = 50.0) -> P.World:
    w = _world(api, desc, world)
    g = w.add_body(P.BODY_STATIC, (0.0, -1.0, 0.0))
    w.add_shape(g, P.geometry_box(ground_half, 1.0, ground_half), friction=0.6)
    for row in range(rows):
        count = rows - row
        for i in range(count):
            x = (float(i) - float(count - 1) * 0.5) * 1.05
            b = w.add_body(P.BODY_DYNAMIC, (x, 0.5 + float(row), 0.0))
            w.add_shape(b, P.geometry_box(0.5, 0.5, 0.5), friction=0.6)
    return w


def hull_points(r: Pcg32, count: int):
    """`count` points in a ball of radius U[0.3, 0.6]: a uniform direction
    (two draws, z = 1 - 2u, phi = 2 pi v) times radius * sqrt(0.5 + 0.5 w)."""
    radius = float(r.next_range(0.3, 0.6))
    pts = []
    for _ in range(count):
        u = float(r.next_f32())
        v = float(r.next_f32())
        z = 1.0 - 2.0 * u
        s = math.sqrt(max(0.0, 1.0 - z * z))
        phi = 2.0 * math.pi * v
        d = radius * math.sqrt(float(r.next_f32()) * 0.5 + 0.5)
        pts.append((s * math.cos(phi) * d, s * math.sin(phi) * d, z * d))
    return np.array(pts, dtype=np.float32)


class HullBank:
    """Hull blocks (the reference's byte format) kept alive for the worlds
    that point at them."""

    def __init__(self, blob: bytes, offsets):
        self._buf = C.create_string_buffer(blob, len(blob))
        base = C.addressof(self._buf)
        self.ptrs = [base + int(o) for o in offsets]

    def __len__(self):
        return len(self.ptrs)


def built_hull_bank(api, vertex_counts, seed: int = 5) -> HullBank:
    """Hulls built by the library's own nph_physics_hull_build: for each
    budget, the hull of 4 * budget points on a sphere of radius U[0.3, 0.6]
    (slightly jittered so faces stay triangles), the budget filled."""
    r = Pcg32(seed)
    blocks = []
    for n in vertex_counts:
        radius = float(r.next_range(0.3, 0.6))
        pts = []
        for _ in range(4 * n):
            u, v = float(r.next_f32()), float(r.next_f32())
            z = 1.0 - 2.0 * u
            sxy = math.sqrt(max(0.0, 1.0 - z * z))
            phi = 2.0 * math.pi * v
            d = radius * (1.0 + 0.01 * float(r.next_f32()))
            pts.append((sxy * math.cos(phi) * d, sxy * math.sin(phi) * d, z * d))
        rc, block = P.build_hull(api, np.array(pts, dtype=np.float32), n)
        if rc != 0:
            raise RuntimeError(f"hull build failed: {rc}")
        blocks.append(block)
    offsets, blob = [], b""
    for b in blocks:
        offsets.append(len(blob))
        blob += b + b"\0" * ((16 - len(b) % 16) % 16)
    return HullBank(blob, offsets)


def load_hull_bank(path: str) -> HullBank:
    """A bank file: u32 count, u32 offsets[count], then 16-byte aligned blocks."""
    data = open(path, "rb").read()
    n = int(np.frombuffer(data[:4], dtype=np.uint32)[0])
    offsets = np.frombuffer(data[4:4 + 4 * n], dtype=np.uint32)
    return HullBank(data, offsets)


def bin_drop(api, bodies: int, seed: int, hulls: HullBank | None, half: float = 20.0,
             height: float = 20.0, desc: dict | None = None, y_lo: float = 5.0,
             world: P.World | None = None) -> P.World:
    """`bodies` mixed shapes dropped into an open bin of static boxes
    (floor and four walls), x, z in [-half, half], y in [y_lo, height]."""
    w = _world(api, desc, world)
    floor = w.add_body(P.BODY_STATIC, (0.0, -1.0, 0.0))
    w.add_shape(floor, P.geometry_box(half + 1.0, 1.0, half + 1.0), friction=0.5)
    wall_h = height * 0.5
    for sx, sz, hx, hz in ((1, 0, 0.5, half + 1.0), (-1, 0, 0.5, half + 1.0),
                           (0, 1, half + 1.0, 0.5), (0, -1, half + 1.0, 0.5)):
        wb = w.add_body(P.BODY_STATIC, (sx * (half + 0.5), wall_h, sz * (half + 0.5)))
        w.add_shape(wb, P.geometry_box(hx, wall_h, hz), friction=0.5)
    drop_mixed(w, Pcg32(seed), bodies, hulls, -half + 1.0, half - 1.0, y_lo, height)
    return w


def drop_mixed(w: P.World, r: Pcg32, bodies: int, hulls: HullBank | None, lo: float, hi: float,
               y_lo: float, y_hi: float):
    """`bodies` dynamic bodies of mixed shapes (sphere, capsule, box, hull)
    at x, z in [lo, hi], y in [y_lo, y_hi], random rotations, friction 0.5."""
    kinds = 4 if hulls is not None and len(hulls) > 0 else 3
    for _ in range(bodies):
        kind = r.next_bounded(kinds)
        x = float(r.next_range(lo, hi))
        y = float(r.next_range(y_lo, y_hi))
        z = float(r.next_range(lo, hi))
        q = rotation(r)
        if kind == 0:
            g = P.geometry_sphere(float(r.next_range(0.2, 0.6)))
        elif kind == 1:
            rad = float(r.next_range(0.2, 0.4))
            hh = float(r.next_range(0.3, 0.8))
            g = P.geometry_capsule(hh, rad)
        elif kind == 2:
            hx = float(r.next_range(0.2, 0.7))
            hy = float(r.next_range(0.2, 0.7))
            hz = float(r.next_range(0.2, 0.7))
            g = P.geometry_box(hx, hy, hz)
        else:
            g = P.geometry_hull(hulls.ptrs[r.next_bounded(len(hulls))])
        b = w.add_body(P.BODY_DYNAMIC, (x, y, z), q)
        w.add_shape(b, g, friction=0.5)
    w._keep.append(hulls)


# ---------------------------------------------------------------------------
# Meshes: blocks built by the reference's mesh builder (tests/golden), and
# scenes of bodies dropped on them.
# ---------------------------------------------------------------------------
class MeshBank(HullBank):
    """Mesh blocks kept alive, each 64-byte aligned (the blocks' nodes are
    read with aligned four-wide loads)."""

    def __init__(self, blob: bytes, offsets):
        self._buf = C.create_string_buffer(len(blob) + 64)
        base = (C.addressof(self._buf) + 63) & ~63
        C.memmove(base, blob, len(blob))
        self.ptrs = [base + int(o) for o in offsets]


def load_mesh_bank(path: str) -> MeshBank:
    """The bank format of load_hull_bank, blocks 64-byte aligned in the file."""
    data = open(path, "rb").read()
    n = int(np.frombuffer(data[:4], dtype=np.uint32)[0])
    offsets = np.frombuffer(data[4:4 + 4 * n], dtype=np.uint32)
    return MeshBank(data, offsets)


def terrain_grid(r: Pcg32, nx: int, nz: int, cell: float, amp: float):
    """A height grid centred on the origin: (nx + 1)(nz + 1) vertices, heights
    amp * U[0, 1) row by row, two triangles a cell, counter-clockwise seen
    from +y.  Returns (vertices float32 (n, 3), indices uint32 (m, 3))."""
    verts = []
    for j in range(nz + 1):
        for i in range(nx + 1):
            h = np.float32(np.float32(amp) * r.next_f32())
            verts.append((np.float32(i * cell - nx * cell * 0.5), h,
                          np.float32(j * cell - nz * cell * 0.5)))
    idx = []
    for j in range(nz):
        for i in range(nx):
            v00 = j * (nx + 1) + i
            v10 = v00 + 1
            v01 = v00 + nx + 1
            v11 = v01 + 1
            idx.append((v00, v01, v10))
            idx.append((v10, v01, v11))
    return np.array(verts, dtype=np.float32), np.array(idx, dtype=np.uint32)


def mesh_terrain_drop(api, meshes: MeshBank, bodies: int, seed: int, hulls: HullBank | None,
                      desc: dict | None = None, world: P.World | None = None) -> P.World:
    """Mixed bodies dropped from 3-12 m onto the bank's first mesh (a
    32 x 32 m height grid) on a static body at the origin."""
    w = _world(api, desc, world)
    ground = w.add_body(P.BODY_STATIC, (0.0, 0.0, 0.0))
    w.add_shape(ground, P.geometry_mesh(meshes.ptrs[0]), friction=0.5)
    drop_mixed(w, Pcg32(seed), bodies, hulls, -14.0, 14.0, 3.0, 12.0)
    w._keep.append(meshes)
    return w


def mesh_field(api, meshes: MeshBank, bodies: int, seed: int, hulls: HullBank | None,
               desc: dict | None = None) -> P.World:
    """Four 64-triangle patches (the bank's meshes 1-4) on static bodies, one
    tilted, one turned, the two-sided one upside down, over a ground slab;
    mixed bodies dropped onto them."""
    d = dict(DESC)
    d.update(desc or {})
    w = P.World(api, d)
    ground(w, -2.0)
    places = (((-7.0, 0.0, -7.0), (0.0, 0.0, 0.0, 1.0)),
              ((7.0, 0.5, -7.0), quat_axis_angle((1.0, 0.0, 0.0), 0.35)),
              ((-7.0, 0.0, 7.0), quat_axis_angle((0.0, 1.0, 0.0), 0.785398)),
              ((7.0, 1.0, 7.0), quat_axis_angle((1.0, 0.0, 0.0), 3.14159265)))
    for k, (pos, q) in enumerate(places):
        b = w.add_body(P.BODY_STATIC, pos, q)
        w.add_shape(b, P.geometry_mesh(meshes.ptrs[1 + k]), friction=0.5)
    drop_mixed(w, Pcg32(seed), bodies, hulls, -12.0, 12.0, 4.0, 14.0)
    w._keep.append(meshes)
    return w


# ---------------------------------------------------------------------------
# Ragdolls.  Every float operation below is the reference's, in single
# precision where it uses f32 (numpy float32 rounds each operation), f64
# where it uses f64, and sinf/cosf from the C library.
# ---------------------------------------------------------------------------
_F = np.float32
_libm = C.CDLL("libm.so.6")
_libm.sinf.restype = C.c_float
_libm.sinf.argtypes = [C.c_float]
_libm.cosf.restype = C.c_float
_libm.cosf.argtypes = [C.c_float]

K_PI = _F(3.14159265)
DEG = _F(K_PI / _F(180.0))


def quat_axis_angle(axis, angle):
    """nph_quat_from_axis_angle (quat.h:121-127)."""
    half = _F(_F(0.5) * _F(angle))
    sn = _F(_libm.sinf(float(half)))
    cs = _F(_libm.cosf(float(half)))
    return (_F(_F(axis[0]) * sn), _F(_F(axis[1]) * sn), _F(_F(axis[2]) * sn), cs)


def quat_mul(a, b):
    """Hamilton product (quat.h:70-76), left to right in f32."""
    ax, ay, az, aw = (_F(v) for v in a)
    bx, by, bz, bw = (_F(v) for v in b)
    return (_F(_F(_F(aw * bx) + _F(ax * bw)) + _F(ay * bz)) - _F(az * by),
            _F(_F(_F(aw * by) - _F(ax * bz)) + _F(ay * bw)) + _F(az * bx),
            _F(_F(_F(aw * bz) + _F(ax * by)) - _F(ay * bx)) + _F(az * bw),
            _F(_F(_F(aw * bw) - _F(ax * bx)) - _F(ay * by)) - _F(az * bz))


def _cross(a, b, T):
    return (T(T(a[1] * b[2]) - T(a[2] * b[1])), T(T(a[2] * b[0]) - T(a[0] * b[2])),
            T(T(a[0] * b[1]) - T(a[1] * b[0])))


def quat_rotate(q, v):
    """v + 2w (u x v) + 2 u x (u x v) in f32 (quat.h:157-164)."""
    u = (_F(q[0]), _F(q[1]), _F(q[2]))
    v = tuple(_F(x) for x in v)
    c = _cross(u, v, _F)
    t = tuple(_F(x * _F(2.0)) for x in c)
    c2 = _cross(u, t, _F)
    return tuple(_F(_F(v[i] + _F(t[i] * _F(q[3]))) + c2[i]) for i in range(3))


def ragdoll_turn(q, v):
    """`v` turned by `q` in f64 (test_world_support.h:293-299)."""
    u = (float(q[0]), float(q[1]), float(q[2]))
    c = _cross(u, v, float)
    t = tuple(x * 2.0 for x in c)
    c2 = _cross(u, t, float)
    return tuple((v[i] + t[i] * float(q[3])) + c2[i] for i in range(3))


_RAGDOLL_PARTS = (  # x, y, geometry, sideways
    (0.0, 1.0, ("box", (0.15, 0.1, 0.1)), False),
    (0.0, 1.32, ("box", (0.16, 0.17, 0.1)), False),
    (0.0, 1.62, ("sphere", 0.1), False),
    (-0.33, 1.4, ("capsule", (0.11, 0.05)), True),
    (0.33, 1.4, ("capsule", (0.11, 0.05)), True),
    (-0.62, 1.4, ("capsule", (0.11, 0.045)), True),
    (0.62, 1.4, ("capsule", (0.11, 0.045)), True),
    (-0.1, 0.68, ("capsule", (0.14, 0.07)), False),
    (0.1, 0.68, ("capsule", (0.14, 0.07)), False),
    (-0.1, 0.25, ("capsule", (0.14, 0.06)), False),
    (0.1, 0.25, ("capsule", (0.14, 0.06)), False),
)

_RAGDOLL_JOINTS = (  # a, b, x, y, axis, hinge, p0, p1, p2
    (0, 1, 0.0, 1.12, (0.0, 1.0, 0.0), False, 30.0, 20.0, 20.0),
    (1, 2, 0.0, 1.5, (0.0, 1.0, 0.0), False, 40.0, 40.0, 45.0),
    (1, 3, -0.17, 1.42, (-1.0, 0.0, 0.0), False, 80.0, 60.0, 45.0),
    (1, 4, 0.17, 1.42, (1.0, 0.0, 0.0), False, 80.0, 60.0, 45.0),
    (3, 5, -0.48, 1.4, (0.0, 1.0, 0.0), True, 0.0, 140.0, 0.0),
    (4, 6, 0.48, 1.4, (0.0, -1.0, 0.0), True, 0.0, 140.0, 0.0),
    (0, 7, -0.1, 0.9, (0.0, -1.0, 0.0), False, 60.0, 30.0, 30.0),
    (0, 8, 0.1, 0.9, (0.0, -1.0, 0.0), False, 60.0, 30.0, 30.0),
    (7, 9, -0.1, 0.46, (1.0, 0.0, 0.0), True, 0.0, 140.0, 0.0),
    (8, 10, 0.1, 0.46, (1.0, 0.0, 0.0), True, 0.0, 140.0, 0.0),
)


def _geometry(g):
    kind, v = g
    if kind == "box":
        return P.geometry_box(*v)
    if kind == "sphere":
        return P.geometry_sphere(v)
    return P.geometry_capsule(*v)


def ragdoll(w: P.World, base, turn=(0.0, 0.0, 0.0, 1.0)):
    """An eleven-body ragdoll standing in a T-pose at `base`, turned by `turn`
    (build_ragdoll, test_world_support.h:301-379). Returns (bodies, joints)."""
    side = quat_axis_angle((0.0, 0.0, 1.0), _F(_F(90.0) * DEG))
    turn = tuple(_F(x) for x in turn)
    bodies = []
    for x, y, g, sideways in _RAGDOLL_PARTS:
        t = ragdoll_turn(turn, (x, y, 0.0))
        pos = (base[0] + t[0], base[1] + t[1], base[2] + t[2])
        rot = quat_mul(turn, side if sideways else (0.0, 0.0, 0.0, 1.0))
        b = w.add_body(P.BODY_DYNAMIC, pos, tuple(float(c) for c in rot))
        w.add_shape(b, _geometry(g), friction=0.6, restitution=0.0)
        bodies.append(b)
    joints = []
    for a, b, x, y, axis, hinge, p0, p1, p2 in _RAGDOLL_JOINTS:
        t = ragdoll_turn(turn, (x, y, 0.0))
        at = (base[0] + t[0], base[1] + t[1], base[2] + t[2])
        ax = tuple(float(c) for c in quat_rotate(turn, axis))
        if hinge:
            j = w.add_joint(P.JOINT_HINGE, bodies[a], bodies[b], at, ax, hinge={
                "enable_limit": 1, "lower": float(_F(_F(p0) * DEG)),
                "upper": float(_F(_F(p1) * DEG)), "enable_motor": 1, "max_motor_torque": 2.0})
        else:
            j = w.add_joint(P.JOINT_BALL, bodies[a], bodies[b], at, ax, ball={
                "enable_swing_limit": 1, "swing_y": float(_F(_F(p0) * DEG)),
                "swing_z": float(_F(_F(p1) * DEG)), "enable_twist_limit": 1,
                "twist_lower": float(_F(_F(-p2) * DEG)), "twist_upper": float(_F(_F(p2) * DEG)),
                "friction_torque": 2.0})
        joints.append(j)
    return bodies, joints


def throw_ragdoll(w: P.World, bodies, v, omega):
    """Moving as one rigid body: `v` at the pelvis, turning at `omega`
    (test_ragdoll.c:269-281)."""
    a = w.api
    about = a.body_center(w.ptr, bodies[0])
    om = tuple(_F(c) for c in omega)
    for b in bodies:
        c = a.body_center(w.ptr, b)
        arm = (_F(c.x - about.x), _F(c.y - about.y), _F(c.z - about.z))
        cr = _cross(om, arm, _F)
        lin = tuple(float(_F(_F(v[i]) + cr[i])) for i in range(3))
        a.body_set_linear_velocity(w.ptr, b, P.Vec3(*lin))
        a.body_set_angular_velocity(w.ptr, b, P.Vec3(*(float(c) for c in om)))


def ground(w: P.World, top=0.0):
    """A static slab whose top face is y = top (test_world_support.h:63-68)."""
    g = w.add_body(P.BODY_STATIC, (0.0, top - 1.0, 0.0))
    w.add_shape(g, P.geometry_box(50.0, 1.0, 50.0), friction=0.6)
    return g


def _pile(w: P.World, r: Pcg32, count: int, x: float = 0.0, z: float = 0.0):
    """`count` ragdolls lying criss-cross one above another at (x, z), each a
    quarter turn from the last and thrown at random (build_pile,
    test_ragdoll.c:289-303)."""
    lying = quat_axis_angle((1.0, 0.0, 0.0), _F(_F(-90.0) * DEG))
    for k in range(count):
        turn = quat_mul(quat_axis_angle((0.0, 1.0, 0.0), _F(_F(_F(k) * _F(90.0)) * DEG)), lying)
        pelvis = (x, 0.15 + 0.3 * float(k), z)
        off = ragdoll_turn(turn, (0.0, 1.0, 0.0))
        base = (pelvis[0] - off[0], pelvis[1] - off[1], pelvis[2] - off[2])
        bodies, joints = ragdoll(w, base, turn)
        v = tuple(float(r.next_range(-0.5, 0.5)) for _ in range(3))
        om = tuple(float(r.next_range(-1.0, 1.0)) for _ in range(3))
        throw_ragdoll(w, bodies, v, om)
        w.ragdolls.append((bodies, joints))


def ragdoll_pile(api, count: int = 4, seed: int = 2026, desc: dict | None = None,
                 world: P.World | None = None) -> P.World:
    """`count` ragdolls in one criss-cross pile on a ground slab."""
    w = _world(api, desc, world)
    ground(w, 0.0)
    w.ragdolls = []
    _pile(w, Pcg32(seed), count)
    return w


def ragdoll_piles(api, ragdolls: int, seed: int, desc: dict | None = None,
                  world: P.World | None = None, per_pile: int = 4,
                  spacing: float = 4.0) -> P.World:
    """`ragdolls` ragdolls in piles of `per_pile` (the pile above) on a
    square grid `spacing` m apart, on one ground slab."""
    w = _world(api, desc, world)
    piles = (ragdolls + per_pile - 1) // per_pile
    side = max(1, int(math.ceil(math.sqrt(piles))))
    half = 0.5 * spacing * float(side)
    g = w.add_body(P.BODY_STATIC, (0.0, -1.0, 0.0))
    w.add_shape(g, P.geometry_box(half + 5.0, 1.0, half + 5.0), friction=0.6)
    w.ragdolls = []
    r = Pcg32(seed)
    left = ragdolls
    for p in range(piles):
        x = -half + spacing * (float(p % side) + 0.5)
        z = -half + spacing * (float(p // side) + 0.5)
        _pile(w, r, min(per_pile, left), x, z)
        left -= per_pile
    return w


def batch_plan(seed: int, worlds: int):
    """Each world of the batched workload: (kind, bodies, seed). Bodies =
    floor(16 * 1250^u), u uniform in [0, 1) (log-uniform in [16, 20 000]);
    kind 0 pyramid, 1 bin drop, 2 ragdoll piles."""
    r = Pcg32(seed)
    out = []
    for _ in range(worlds):
        u = r.next_f32()
        n = int(math.floor(16.0 * 1250.0 ** u))
        kind = r.next_bounded(3)
        out.append((kind, n, r.next_u32()))
    return out


def fill_batch_world(w: P.World, kind: int, bodies: int, seed: int, hulls: HullBank | None):
    """A world of the batched workload with about `bodies` dynamic bodies: a
    pyramid (the most rows whose boxes fit), a bin drop (bin area scaled
    with the count, 40 x 40 m at 10 000), or ragdolls (bodies // 11) in
    piles of four."""
    if kind == 0:
        rows = max(1, int((math.isqrt(8 * bodies + 1) - 1) // 2))
        pyramid(None, rows, world=w, ground_half=max(50.0, 0.55 * rows + 5.0))
    elif kind == 1:
        half = max(3.0, 20.0 * math.sqrt(bodies / 10000.0))
        bin_drop(None, bodies, seed, hulls, half=half, world=w)
    else:
        ragdoll_piles(None, max(1, bodies // 11), seed, world=w)
    return w


def field(api, meshes: MeshBank, hulls: HullBank | None, seed: int = 3, mixed: int = 78000,
          ragdolls: int = 2000, patches: int = 200, half: float = 75.0,
          desc: dict | None = None) -> P.World:
    """A 150 x 150 m field: a ground slab; `patches` static mesh patches (the
    bank's 64-triangle meshes 1-4 in turn) at random places and headings;
    `ragdolls` eleven-body ragdolls dropped lying at random headings from
    3-12 m; `mixed` bodies as the bin drop's from 2-20 m."""
    d = dict(DESC)
    d.update(desc or {})
    w = P.World(api, d)
    g = w.add_body(P.BODY_STATIC, (0.0, -1.0, 0.0))
    w.add_shape(g, P.geometry_box(half + 1.0, 1.0, half + 1.0), friction=0.6)
    r = Pcg32(seed)
    for k in range(patches):
        x = float(r.next_range(-half + 6.0, half - 6.0))
        z = float(r.next_range(-half + 6.0, half - 6.0))
        yaw = quat_axis_angle((0.0, 1.0, 0.0), _F(r.next_range(0.0, 6.2831853)))
        b = w.add_body(P.BODY_STATIC, (x, 0.05, z), tuple(float(c) for c in yaw))
        w.add_shape(b, P.geometry_mesh(meshes.ptrs[1 + k % 4]), friction=0.5)
    lying = quat_axis_angle((1.0, 0.0, 0.0), _F(_F(-90.0) * DEG))
    for _ in range(ragdolls):
        x = float(r.next_range(-half + 3.0, half - 3.0))
        y = float(r.next_range(3.0, 12.0))
        z = float(r.next_range(-half + 3.0, half - 3.0))
        turn = quat_mul(quat_axis_angle((0.0, 1.0, 0.0), _F(r.next_range(0.0, 6.2831853))), lying)
        off = ragdoll_turn(turn, (0.0, 1.0, 0.0))
        ragdoll(w, (x - off[0], y - off[1], z - off[2]), turn)
    drop_mixed(w, r, mixed, hulls, -half + 3.0, half - 3.0, 2.0, 20.0)
    w._keep.append(meshes)
    return w


def projectiles(api, count: int = 40, seed: int = 5, bullets: bool = True,
                desc: dict | None = None) -> P.World:
    """Continuous collision: spheres, capsules and boxes shot at 60-150 m/s
    along +Z, spinning, at a 4 cm static wall 5 m away; every fifth a bullet
    (swept against bodies too); a stack of six dynamic boxes between them."""
    d = {"enable_sleep": 0, "enable_continuous": 1}
    d.update(desc or {})
    w = P.World(api, d)
    ground(w, 0.0)
    wall = w.add_body(P.BODY_STATIC, (0.0, 2.0, 5.0))
    w.add_shape(wall, P.geometry_box(3.0, 2.0, 0.02))
    r = Pcg32(seed)
    for i in range(count):
        x = float(r.next_range(-2.0, 2.0))
        y = float(r.next_range(0.5, 3.5))
        q = rotation(r)
        vz = float(r.next_range(60.0, 150.0))
        wx = float(r.next_range(-20.0, 20.0))
        b = w.add_body(P.BODY_DYNAMIC, (x, y, 0.0), q, velocity=(0.0, 0.0, vz),
                       angular_velocity=(wx, 0.0, 0.0),
                       bullet=1 if (bullets and i % 5 == 0) else 0)
        k = i % 3
        g = (P.geometry_sphere(0.1) if k == 0 else
             P.geometry_capsule(0.15, 0.05) if k == 1 else P.geometry_box(0.1, 0.05, 0.2))
        w.add_shape(b, g)
    for j in range(6):
        b = w.add_body(P.BODY_DYNAMIC, (0.0, 0.25 + 0.5 * j, 2.5))
        w.add_shape(b, P.geometry_box(0.25, 0.25, 0.25))
    return w
