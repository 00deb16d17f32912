# SPDX-License-Identifier: Apache-2.0
"""Debug drawing of whole worlds (nph_physics_world_draw, world.h:317-318,
world.c:840-900) against the reference's, line for line (ends and colour in
call order), after stepping both worlds bit-exactly: a mixed drop onto mesh
patches (shapes, fat boxes, centres of mass, touching contacts), and a
ragdoll pile with sleep on (joints, sleeping bodies' colour), each with and
without NPH_PHYSICS_DRAW_SLEEP, from an origin far from the scene."""
import ctypes as C
import os

import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

EVERYTHING = (P.DRAW_SHAPES | P.DRAW_AABBS | P.DRAW_CONTACTS | P.DRAW_MASS | P.DRAW_JOINTS)


def _draw(w, flags, origin):
    P.bind_draw(w.api)
    rec = P.LineRecorder(origin, flags)
    w.api.world_draw(w.ptr, C.byref(rec.draw))
    return rec.lines


def _compare(wg, wr, origin=(1000.5, -20.25, 3.0)):
    for flags in (EVERYTHING, EVERYTHING | P.DRAW_SLEEP, P.DRAW_CONTACTS):
        mine = _draw(wg, flags, origin)
        theirs = _draw(wr, flags, origin)
        assert len(mine) == len(theirs), flags
        assert mine == theirs, flags
    return len(mine)


def test_world_draw_matches_the_reference_on_meshes(gpu_api, ref_api, hull_bank):
    print("world draw: mesh field", flush=True)
    meshes = S.load_mesh_bank(os.path.join(GOLDEN, "meshes.bin"))
    wg = S.mesh_field(gpu_api, meshes, 150, 31, hull_bank)
    wr = S.mesh_field(ref_api, meshes, 150, 31, hull_bank)
    for _ in range(150):
        wg.step(S.DT)
        wr.step(S.DT)
    assert wg.hash() == wr.hash()
    contact_lines = _compare(wg, wr)
    assert contact_lines > 50


def test_world_draw_matches_the_reference_on_sleeping_ragdolls(gpu_api, ref_api):
    print("world draw: ragdoll pile with sleep", flush=True)
    desc = {"enable_sleep": 1}
    wg = S.ragdoll_pile(gpu_api, 4, desc=desc)
    wr = S.ragdoll_pile(ref_api, 4, desc=desc)
    for _ in range(240):
        wg.step(S.DT)
        wr.step(S.DT)
    assert wg.hash() == wr.hash()
    _compare(wg, wr)
    asleep = _draw(wg, P.DRAW_SHAPES | P.DRAW_SLEEP, (0.0, 0.0, 0.0))
    assert any(line[-4:] == (0xFF806040).to_bytes(4, "little") for line in asleep)
