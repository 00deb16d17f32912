# SPDX-License-Identifier: Apache-2.0
"""Seeded scenes of the world-step workloads, built through the C API so the
same calls drive this package's library and, in tests, the reference's.

* pyramid: a 20-row box pyramid (210 unit boxes, 1000 kg/m^3, friction 0.6)
  on a static ground box, as engine/physics/tests/test_world.c:365-379.
* bin: mixed spheres, capsules, boxes and 12-vertex hulls dropped into an
  open static bin (the 10k-body workload and smaller versions of it).

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


def pyramid(api, rows: int = 20, desc: dict | None = None) -> P.World:
    d = dict(DESC)
    d.update(desc or {})
    w = P.World(api, d)
    g = w.add_body(P.BODY_STATIC, (0.0, -1.0, 0.0))
    w.add_shape(g, P.geometry_box(50.0, 1.0, 50.0), friction=0.6)
    for row in range(rows):
        count = rows - row
        for i in range(count):
            x = (float(i) - float(count - 1) * 0.5) * 1.05
            b = w.add_body(P.BODY_DYNAMIC, (x, 0.5 + float(row), 0.0))
            w.add_shape(b, P.geometry_box(0.5, 0.5, 0.5), friction=0.6)
    return w


class HullBank:
    """Hull blocks (the reference's byte format) kept alive for the worlds
    that point at them."""

    def __init__(self, blob: bytes, offsets):
        self._buf = C.create_string_buffer(blob, len(blob))
        base = C.addressof(self._buf)
        self.ptrs = [base + int(o) for o in offsets]

    def __len__(self):
        return len(self.ptrs)


def load_hull_bank(path: str) -> HullBank:
    """A bank file: u32 count, u32 offsets[count], then 16-byte aligned blocks."""
    data = open(path, "rb").read()
    n = int(np.frombuffer(data[:4], dtype=np.uint32)[0])
    offsets = np.frombuffer(data[4:4 + 4 * n], dtype=np.uint32)
    return HullBank(data, offsets)


def bin_drop(api, bodies: int, seed: int, hulls: HullBank | None, half: float = 20.0,
             height: float = 20.0, desc: dict | None = None, y_lo: float = 5.0) -> P.World:
    """`bodies` mixed shapes dropped into an open bin of static boxes
    (floor and four walls), x, z in [-half, half], y in [y_lo, height]."""
    d = dict(DESC)
    d.update(desc or {})
    w = P.World(api, d)
    floor = w.add_body(P.BODY_STATIC, (0.0, -1.0, 0.0))
    w.add_shape(floor, P.geometry_box(half + 1.0, 1.0, half + 1.0), friction=0.5)
    wall_h = height * 0.5
    for sx, sz, hx, hz in ((1, 0, 0.5, half + 1.0), (-1, 0, 0.5, half + 1.0),
                           (0, 1, half + 1.0, 0.5), (0, -1, half + 1.0, 0.5)):
        wb = w.add_body(P.BODY_STATIC, (sx * (half + 0.5), wall_h, sz * (half + 0.5)))
        w.add_shape(wb, P.geometry_box(hx, wall_h, hz), friction=0.5)
    r = Pcg32(seed)
    kinds = 4 if hulls is not None and len(hulls) > 0 else 3
    for _ in range(bodies):
        kind = r.next_bounded(kinds)
        x = float(r.next_range(-half + 1.0, half - 1.0))
        y = float(r.next_range(y_lo, height))
        z = float(r.next_range(-half + 1.0, half - 1.0))
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
    return w
