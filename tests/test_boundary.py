# SPDX-License-Identifier: Apache-2.0
"""The C ABI without a GPU: the library exports every function the header
declares, rejects what it does not support with the reference's error
codes, refuses to run without a HIP device (no CPU fallback), and the
fixtures and scene generators agree with the reference."""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nepheline_physics.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b(nph_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = P.PRODUCT_LIB
    if not os.path.exists(lib):
        pytest.skip("libnepheline_amd.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in _declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    assert len(_declared_functions()) >= 30


LAYOUT_TYPES = {
    "NphPhysicsPose": P.Pose, "NphPhysicsMassData": P.MassData, "NphPhysicsGeometry": P.Geometry,
    "NphPhysicsWorldDesc": P.WorldDesc, "NphPhysicsBodyDesc": P.BodyDesc,
    "NphPhysicsFilter": P.Filter, "NphPhysicsShapeDesc": P.ShapeDesc, "NphPhysicsStats": P.Stats,
    "NphPhysicsHull": None, "NphPhysicsEvents": P.Events,
    "NphPhysicsContactHitEvent": P.HitEvent, "NphPhysicsBodyMoveEvent": P.MoveEvent, "NphPhysicsJointDesc": P.JointDesc,
    "NphPhysicsBallJoint": P.BallJoint, "NphPhysicsHingeJoint": P.HingeJoint,
    "NphPhysicsD6Joint": P.D6Joint, "NphPhysicsMesh": P.Mesh, "NphPhysicsMeshNode": P.MeshNode,
    "NphPhysicsContactBeginEvent": P.PairEvent, "NphPhysicsSensorEvent": P.SensorEvent,
    "NphPhysicsManifold": P.Manifold, "NphPhysicsManifoldPoint": P.ManifoldPoint,
    "NphPhysicsCollideCache": P.CollideCache,
}


def _layouts(tmp_path, includes, header):
    """sizeof and the offset of every ctypes-declared field, as the C
    compiler lays the header's structs out."""
    lines = [f"#include {header}", "#include <stdio.h>", "#include <stddef.h>", "int main(void){"]
    for name, ct in LAYOUT_TYPES.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for fname, *_ in getattr(ct, "_fields_", []):
            if fname == "u":
                continue
            lines.append(f'printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c17", *[f"-I{i}" for i in includes], "-o", str(exe), str(src)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict(line.split() for line in out.splitlines())


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirror lays out every struct as the C header does."""
    got = _layouts(tmp_path, [os.path.join(ROOT, "include")], "<nepheline_physics.h>")
    for name, ct in LAYOUT_TYPES.items():
        if ct is None:
            continue
        assert int(got[name]) == C.sizeof(ct), name
        for fname, *_ in ct._fields_:
            if fname != "u":
                assert int(got[f"{name}.{fname}"]) == getattr(ct, fname).offset, (name, fname)


REF_INCLUDE = ["/root/reference/engine/physics/include", "/root/reference/engine/core/include",
               "/root/reference/engine/math/include", "/root/reference/engine/jobs/include",
               "/root/reference/engine/platform/include"]


def test_header_layouts_are_the_references(tmp_path):
    """Every public struct of include/nepheline_physics.h has the size and
    field offsets of the reference's (engine/physics/include/nepheline/
    physics/{types,hull,world,body,shape,joint,mesh}.h)."""
    if not os.path.isdir(REF_INCLUDE[0]):
        pytest.skip("reference sources not present")
    mine = _layouts(tmp_path, [os.path.join(ROOT, "include")], "<nepheline_physics.h>")
    hdr = "<nepheline/physics/world.h>\n#include <nepheline/physics/body.h>\n" \
          "#include <nepheline/physics/shape.h>\n#include <nepheline/physics/hull.h>\n" \
          "#include <nepheline/physics/joint.h>\n#include <nepheline/physics/mesh.h>\n" \
          "#include <nepheline/physics/collide.h>"
    ref = _layouts(tmp_path, REF_INCLUDE, hdr)
    assert mine == ref


def test_world_refuses_to_run_without_a_gpu():
    """No HIP device here: creation fails loudly with NPH_RESULT_UNSUPPORTED
    instead of stepping on the CPU."""
    if not os.path.exists(P.PRODUCT_LIB):
        pytest.skip("libnepheline_amd.so not built")
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    a = P.product_api()
    d = a.world_desc_default()
    d.alloc = a.allocator()
    d.enable_sleep = 0
    d.enable_continuous = 0
    w = C.c_void_p()
    assert a.world_create(C.byref(d), C.byref(w)) == P.UNSUPPORTED


def test_unsupported_settings_are_rejected_like_the_reference_rejects_bad_ones():
    if not os.path.exists(P.PRODUCT_LIB):
        pytest.skip("libnepheline_amd.so not built")
    a = P.product_api()
    d = a.world_desc_default()
    d.alloc = a.allocator()
    w = C.c_void_p()
    d.substeps = 0
    assert a.world_create(C.byref(d), C.byref(w)) == P.INVALID_ARGUMENT
    d.substeps = 4
    d.solver = 1  # TGS
    d.enable_sleep = 0
    d.enable_continuous = 0
    assert a.world_create(C.byref(d), C.byref(w)) == P.UNSUPPORTED
    d.solver = 2  # XPBD
    assert a.world_create(C.byref(d), C.byref(w)) == P.UNSUPPORTED


def test_desc_defaults_match_the_reference(ref_api):
    if not os.path.exists(P.PRODUCT_LIB):
        pytest.skip("libnepheline_amd.so not built")
    a = P.product_api()
    for f in ("world_desc_default", "body_desc_default"):
        assert bytes(getattr(a, f)()) == bytes(getattr(ref_api, f)())
    g = P.geometry_box(0.5, 0.5, 0.5)
    assert bytes(a.shape_desc_default(g)) == bytes(ref_api.shape_desc_default(g))


def test_geometry_mass_matches_the_reference(ref_api, hull_bank):
    a = P.product_api()
    geoms = [P.geometry_sphere(0.37), P.geometry_capsule(0.6, 0.25),
             P.geometry_box(0.3, 0.55, 0.2), P.geometry_hull(hull_bank.ptrs[3])]
    for g in geoms:
        for rho in (1000.0, 7.5):
            assert bytes(a.geometry_mass(C.byref(g), rho)) == bytes(
                ref_api.geometry_mass(C.byref(g), rho))


def test_pcg32_reference_vectors():
    # engine/math/tests/test_random.c:13-31 (O'Neill's pcg32_srandom_r(42, 54))
    r = S.Pcg32(42, 54)
    assert (r.state, r.inc) == (0x185706B82C2E03F8, 0x6D)
    assert [r.next_u32() for _ in range(8)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293,
                                                0xBFA4784B, 0xCBED606E, 0xBFC6A3AD, 0x812FFF6D]
    r = S.Pcg32(0, 0)
    assert [r.next_u32(), r.next_u32()] == [0xE4C14788, 0x379C6516]


def test_bounded_draws_match_the_reference(ref_api):
    lib = ref_api.lib

    class Pcg(C.Structure):
        _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]

    lib.nph_pcg32_seed.argtypes = [C.POINTER(Pcg), C.c_uint64, C.c_uint64]
    lib.nph_pcg32_next_bounded.argtypes = [C.POINTER(Pcg), C.c_uint32]
    lib.nph_pcg32_next_bounded.restype = C.c_uint32
    st = Pcg()
    lib.nph_pcg32_seed(C.byref(st), 2, 0x9E3779B97F4A7C15)
    mine = S.Pcg32(2)
    for bound in (4, 3, 7, 1000, 0x80000001):
        for _ in range(50):
            assert lib.nph_pcg32_next_bounded(C.byref(st), bound) == mine.next_bounded(bound)


def test_hull_bank_blocks_are_valid_reference_hulls(ref_api, hull_bank):
    lib = ref_api.lib
    lib.nph_physics_hull_validate.argtypes = [C.c_void_p, C.c_uint32, C.c_float]
    lib.nph_physics_hull_validate.restype = C.c_bool
    for p in hull_bank.ptrs[:64]:
        size = C.cast(p, C.POINTER(C.c_uint32))[0]
        vcount = C.cast(p, C.POINTER(C.c_uint32))[1]
        assert 4 <= vcount <= 32
        assert lib.nph_physics_hull_validate(p, size, 0.0)


def test_golden_pyramid_hashes_are_the_references(ref_api):
    """The committed world hashes are what the reference's step gives (its
    step is deterministic: the fixture is regenerated, first 40 steps)."""
    want = json.load(open(os.path.join(GOLDEN, "pyramid_hashes.json")))["hashes"]
    w = S.pyramid(ref_api)
    for i in range(40):
        w.step(1.0 / 60.0)
        assert "%016x" % w.hash() == want[i]


def test_combine_rules_match_the_reference(ref_api):
    """shape.c:29-42 (and world_internal.h:119-155): every pair of modes on
    a grid of values, bit for bit."""
    if not os.path.exists(P.PRODUCT_LIB):
        pytest.skip("libnepheline_amd.so not built")
    a = P.product_api()
    vals = [0.0, 0.1, 0.3, 0.5, 0.77, 1.0]
    for ma in range(5):
        for mb in range(5):
            for x in vals:
                for y in vals:
                    assert a.combine_friction(x * 2, ma, y * 2, mb) == \
                        ref_api.combine_friction(x * 2, ma, y * 2, mb)
                    assert a.combine_restitution(x, ma, y, mb) == \
                        ref_api.combine_restitution(x, ma, y, mb)


def test_batch_creation_rules():
    """A batch needs worlds and refuses continuous collision; without a HIP
    device it refuses to exist at all (no CPU fallback)."""
    if not os.path.exists(P.PRODUCT_LIB):
        pytest.skip("libnepheline_amd.so not built")
    a = P.product_api()
    P._bind_batch(a)
    d = P.world_desc(a, S.DESC)
    b = C.c_void_p()
    assert a.batch_create(C.byref(d), 0, C.byref(b)) == P.INVALID_ARGUMENT
    d.enable_continuous = 1
    assert a.batch_create(C.byref(d), 4, C.byref(b)) == P.UNSUPPORTED
    try:
        import torch
        if torch.cuda.is_available():
            return
    except ImportError:
        pass
    d.enable_continuous = 0
    assert a.batch_create(C.byref(d), 4, C.byref(b)) == P.UNSUPPORTED


def test_batch_plan_is_log_uniform_over_the_three_kinds():
    plan = S.batch_plan(4, 1024)
    sizes = [n for _, n, _ in plan]
    assert min(sizes) >= 16 and max(sizes) <= 20000
    assert 2.4e6 < sum(sizes) < 3.3e6  # mean 16 * 1249 / ln 1250 = 2 803
    assert {k for k, _, _ in plan} == {0, 1, 2}
    small = sum(1 for n in sizes if n < 566)  # 16 * sqrt(1250): the median
    assert 440 < small < 584


def test_collide_query_layout():
    """The extension's query struct as the header lays it out."""
    import tempfile, pathlib
    with tempfile.TemporaryDirectory() as d:
        src = pathlib.Path(d) / "q.c"
        src.write_text("#include <nepheline_physics.h>\n#include <stdio.h>\n#include <stddef.h>\n"
                       "int main(void){printf(\"%zu %zu %zu\\n\", sizeof(NphAmdCollideQuery), "
                       "offsetof(NphAmdCollideQuery, b), offsetof(NphAmdCollideQuery, "
                       "speculative_distance));return 0;}")
        exe = pathlib.Path(d) / "q"
        subprocess.run(["gcc", "-std=c17", f"-I{os.path.join(ROOT, 'include')}", "-o", str(exe),
                        str(src)], check=True, capture_output=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(x) for x in out] == [C.sizeof(P.CollideQuery), P.CollideQuery.b.offset,
                                     P.CollideQuery.speculative_distance.offset]
