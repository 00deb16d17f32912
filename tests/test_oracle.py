# SPDX-License-Identifier: Apache-2.0
"""The oracle (oracle/oracle.c, the C restatement of the solver's arithmetic)
pinned against the reference: its eight-lane contact kernels bit for bit
against the reference's own kernels (compiled from its sources, oracle/_ref)
on seeded random groups; the soft constraint coefficients against the
reference and their closed form; the state hash against the reference's
regression vectors (engine/core/tests/test_hash.c:21-33)."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")

L8 = 8
GROUP = np.dtype([
    ("body_a", "<u4", (8,)), ("body_b", "<u4", (8,)), ("contact", "<u4", (8,)),
    ("write_mask_a", "<u4"), ("write_mask_b", "<u4"), ("lane_count", "<u4"),
    ("point_count", "<u4"), ("first_point", "<u4"), ("scale_hits", "<u4"), ("pad", "<u4", (2,)),
    ("inv_mass_a", "<f4", (8,)), ("inv_mass_b", "<f4", (8,)), ("normal", "<f4", (3, 8)),
    ("tangent1", "<f4", (3, 8)), ("tangent2", "<f4", (3, 8)), ("friction", "<f4", (8,)),
    ("restitution", "<f4", (8,)), ("bias_rate", "<f4", (8,)), ("mass_scale", "<f4", (8,)),
    ("impulse_scale", "<f4", (8,)), ("impact_lanes", "<f4", (8,)), ("impacted", "<f4", (8,))])
POINTS = np.dtype([
    ("anchor_a", "<f4", (3, 8)), ("anchor_b", "<f4", (3, 8)), ("adjusted_separation", "<f4", (8,)),
    ("normal_mass", "<f4", (8,)), ("normal_impulse", "<f4", (8,)),
    ("max_normal_impulse", "<f4", (8,)), ("relative_velocity", "<f4", (8,)),
    ("substep_velocity", "<f4", (8,)), ("tangent_mass", "<f4", (3, 8)),
    ("tangent_impulse", "<f4", (2, 8)), ("ang_a_n", "<f4", (3, 8)), ("ang_b_n", "<f4", (3, 8)),
    ("ang_a_t1", "<f4", (3, 8)), ("ang_b_t1", "<f4", (3, 8)), ("ang_a_t2", "<f4", (3, 8)),
    ("ang_b_t2", "<f4", (3, 8))])
PREPARE = np.dtype([
    ("inv_inertia_a", "<f4", (9, 8)), ("inv_inertia_b", "<f4", (9, 8)),
    ("tangent_impulse", "<f4", (4, 3, 8)), ("active", "<f4", (4, 8))])


def _lib(path, what):
    if not os.path.exists(path):
        pytest.skip(f"{what} not built ({path})")
    return C.CDLL(path, mode=os.RTLD_NOW | os.RTLD_LOCAL)


@pytest.fixture(scope="module")
def oracle():
    return _lib(ORACLE, "the oracle")


@pytest.fixture(scope="module")
def ref():
    return _lib(REF, "the reference (oracle/_ref)")


def test_layouts_are_the_references(ref):
    ref.ref_shim_sizes.restype = C.c_uint32
    assert ref.ref_shim_sizes(0) == 64
    assert ref.ref_shim_sizes(1) == GROUP.itemsize
    assert ref.ref_shim_sizes(2) == POINTS.itemsize
    assert ref.ref_shim_sizes(3) == PREPARE.itemsize


def _unit(rng, n):
    v = rng.normal(size=(n, 3)).astype(np.float64)
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _random_case(rng, n_bodies=12):
    """One group of up to eight lanes over a few bodies (the last is the
    static dummy), with its point rows and prepare inputs."""
    bodies = np.zeros((n_bodies + 1, 16), dtype=np.float32)
    bodies[:, 0:3] = rng.uniform(-2, 2, (n_bodies + 1, 3))
    bodies[:, 4:7] = rng.uniform(-3, 3, (n_bodies + 1, 3))
    bodies[:, 8:11] = rng.uniform(-0.02, 0.02, (n_bodies + 1, 3))
    q = rng.normal(size=(n_bodies + 1, 4))
    q[:, 3] += 30.0  # small rotations
    bodies[:, 12:16] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    bodies[n_bodies] = 0.0
    bodies[n_bodies, 15] = 1.0
    g = np.zeros(1, dtype=GROUP)[0]
    lanes = int(rng.integers(1, 9))
    pc = rng.integers(1, 5, size=8)
    g["lane_count"] = lanes
    g["point_count"] = int(pc[:lanes].max())
    perm = rng.permutation(n_bodies)
    for l in range(8):
        if l < lanes:
            g["body_a"][l] = perm[l % n_bodies] if rng.random() > 0.2 else n_bodies
            g["body_b"][l] = perm[(l + 5) % n_bodies]
            if g["body_a"][l] != n_bodies:
                g["write_mask_a"] |= 1 << l
            g["write_mask_b"] |= 1 << l
            g["inv_mass_a"][l] = 0.0 if g["body_a"][l] == n_bodies else rng.uniform(0.01, 2)
            g["inv_mass_b"][l] = rng.uniform(0.01, 2)
        else:
            g["body_a"][l] = g["body_b"][l] = n_bodies
    n = _unit(rng, 8)
    t1 = _unit(rng, 8)
    t1 = t1 - (t1 * n).sum(1, keepdims=True) * n
    t1 = (t1 / np.linalg.norm(t1, axis=1, keepdims=True)).astype(np.float32)
    t2 = np.cross(n, t1).astype(np.float32)
    g["normal"] = n.T
    g["tangent1"] = t1.T
    g["tangent2"] = t2.T
    g["friction"] = rng.uniform(0, 1, 8)
    g["restitution"] = rng.uniform(0, 1, 8)
    g["bias_rate"] = rng.uniform(10, 60, 8)
    g["mass_scale"] = rng.uniform(0.3, 1, 8)
    g["impulse_scale"] = rng.uniform(0, 0.7, 8)
    for l in range(lanes, 8):
        for f in ("normal", "tangent1", "tangent2"):
            g[f][:, l] = 0.0
        for f in ("friction", "restitution", "bias_rate", "mass_scale", "impulse_scale"):
            g[f][l] = 0.0
    pts = np.zeros(4, dtype=POINTS)
    pin = np.zeros(1, dtype=PREPARE)[0]
    for k in range(4):
        for l in range(lanes):
            if k < pc[l]:
                pts[k]["anchor_a"][:, l] = rng.uniform(-0.7, 0.7, 3)
                pts[k]["anchor_b"][:, l] = rng.uniform(-0.7, 0.7, 3)
                pts[k]["adjusted_separation"][l] = rng.uniform(-0.05, 0.03)
                pts[k]["normal_impulse"][l] = rng.uniform(0, 5) if rng.random() > 0.3 else 0.0
                pin["tangent_impulse"][k, :, l] = rng.uniform(-1, 1, 3)
                pin["active"][k, l] = 1.0
    for l in range(lanes):
        for key in ("inv_inertia_a", "inv_inertia_b"):
            m = rng.uniform(-0.2, 0.2, (3, 3))
            m = m @ m.T + np.eye(3) * rng.uniform(0.2, 2)
            pin[key][:, l] = m.T.reshape(9).astype(np.float32)  # column-major
        if g["body_a"][l] == n_bodies:
            pin["inv_inertia_a"][:, l] = 0.0
    return bodies, g, pts, pin


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def _run(lib, prefix, case, passes, inv_h=240.0, push_max=3.0):
    bodies, g, pts, pin = (x.copy() for x in case)
    g = np.array([g], dtype=GROUP)
    pin = np.array([pin], dtype=PREPARE)
    f = lambda name: getattr(lib, prefix + name)  # noqa: E731
    f("prepare")(_ptr(bodies), _ptr(g), _ptr(pts), _ptr(pin))
    for kind in passes:
        if kind == "warm":
            f("warm_start")(_ptr(bodies), _ptr(g), _ptr(pts))
        elif kind in ("solve", "relax"):
            f("solve")(_ptr(bodies), _ptr(g), _ptr(pts), C.c_float(inv_h), C.c_float(push_max),
                       C.c_float(0.0), C.c_uint32(1 if kind == "solve" else 0))
        else:
            f("restitution")(_ptr(bodies), _ptr(g), _ptr(pts), C.c_float(1.0))
    return bodies.tobytes(), g.tobytes(), pts.tobytes()


PASSES = ["warm", "solve", "relax", "warm", "solve", "relax", "restitution"]


@pytest.mark.parametrize("seed", range(40))
def test_kernels_match_reference_bit_for_bit(oracle, ref, seed):
    rng = np.random.default_rng(1000 + seed)
    case = _random_case(rng)
    # Some cases with impacts (fast approach) and bounces.
    if seed % 3 == 0:
        case[0][:, 0:3] *= 8.0
    got = _run(oracle, "oracle_", case, PASSES)
    want = _run(ref, "ref_shim_", case, PASSES)
    for name, a, b in zip(("bodies", "group", "points"), got, want):
        assert a == b, f"{name} differ (seed {seed})"


def test_make_soft_matches_reference_and_closed_form(oracle, ref):
    out_o = (C.c_float * 3)()
    out_r = (C.c_float * 3)()
    for hertz, zeta, h in ((30.0, 10.0, 1 / 240), (60.0, 10.0, 1 / 240), (15.0, 2.0, 1 / 60),
                           (0.0, 10.0, 1 / 240)):
        oracle.oracle_make_soft(C.c_float(hertz), C.c_float(zeta), C.c_float(h), out_o)
        ref.ref_shim_make_soft(C.c_float(hertz), C.c_float(zeta), C.c_float(h), out_r)
        assert list(out_o) == list(out_r)
        if hertz > 0:
            # Catto, "Soft Constraints": bias = w / (2z + hw), mass = a2/(1+a2), a2 = hw(2z+hw)
            w = 2 * np.pi * hertz
            a2 = h * w * (2 * zeta + h * w)
            np.testing.assert_allclose(list(out_o), [w / (2 * zeta + h * w), a2 / (1 + a2),
                                                     1 / (1 + a2)], rtol=1e-5)
        else:
            assert list(out_o) == [0.0, 1.0, 0.0]


HASH_VECTORS = [  # engine/core/tests/test_hash.c:21-33
    (b"", 0, 0x7B4CCDCEFE90894C),
    (b"", 0x4E65706865, 0x411C93B9BCA3ED3A),
    (b"a", 0, 0x943DB8306D803398),
    (b"abc", 0, 0x834B087AACAB5094),
    (b"message digest", 0, 0x491AC5EEDACE5058),
    (b"abcdefghijklmnop", 0, 0x637305800D88373D),
    (b"abcdefghijklmnopq", 0, 0x6C4803A35122813B),
    (b"The quick brown fox jumps over the lazy dog", 0, 0x57C55C4DB52C6985),
    (b"The quick brown fox jumps over the lazy dog", 0x4E65706865, 0xF01C3FDC8A981C4F),
    (b"0123456789012345678901234567890123456789012345678901234567890123456789", 7,
     0xEA0E9338CABB8FB5),
]


@pytest.mark.parametrize("data,seed,want", HASH_VECTORS)
def test_hash64_regression_vectors(oracle, data, seed, want):
    oracle.oracle_hash64.restype = C.c_uint64
    oracle.oracle_hash64.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64]
    assert oracle.oracle_hash64(data, len(data), seed) == want


def test_hash_default_seed_is_the_references():
    # NPH_HASH_DEFAULT_SEED ("Nephe"), engine/core/include/nepheline/core/hash.h
    assert 0x4E65706865 == int.from_bytes(b"Nephe", "big")
