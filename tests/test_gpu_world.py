# SPDX-License-Identifier: Apache-2.0
"""World-step parity on the GPU: the HIP step against the reference's step
(world hash after every step, engine/physics/src/world.c:436-462; pair and
contact-id lists after canonical sort; poses and impulses).  Bit-exact is
the bar: the hash covers every body's f64 origin, rotation and velocities."""
import json
import os

import numpy as np
import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_backend_is_hip(gpu_api):
    w = S.pyramid(gpu_api, rows=2)
    gpu_api.lib.nph_amd_backend_name.restype = __import__("ctypes").c_char_p
    assert gpu_api.lib.nph_amd_backend_name(w.ptr) == b"hip:gfx950"


def test_pyramid_300_steps_bit_exact(gpu_api):
    want = _golden("pyramid_hashes.json")["hashes"]
    w = S.pyramid(gpu_api)
    for i, h in enumerate(want):
        w.step(1.0 / 60.0)
        got = "%016x" % w.hash()
        assert got == h, f"pyramid hash differs at step {i + 1}"


@pytest.mark.parametrize("run", ["bin_64_11", "bin_400_12"])
def test_bin_drop_bit_exact(gpu_api, hull_bank, run):
    _, bodies, seed = run.split("_")
    want = _golden("bin_hashes.json")["runs"][run]
    w = S.bin_drop(gpu_api, int(bodies), int(seed), hull_bank, half=4.0, height=10.0, y_lo=1.0)
    for i, h in enumerate(want):
        w.step(1.0 / 60.0)
        got = "%016x" % w.hash()
        assert got == h, f"{run}: hash differs at step {i + 1}"


def test_pairs_contacts_and_impulses_match_reference(gpu_api, ref_api, hull_bank):
    """Pair list and contact-id list bit-exact after canonical sort, every
    point's impulses and anchors equal, after 60 steps of a 300-body drop."""
    wg = S.bin_drop(gpu_api, 300, 21, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    wr = S.bin_drop(ref_api, 300, 21, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    for _ in range(60):
        wg.step(1.0 / 60.0)
        wr.step(1.0 / 60.0)
    pg, ig, bg = P.contacts_as_arrays(wg.contacts())
    pr, ir, br = P.contacts_as_arrays(wr.contacts())
    assert pg == pr
    assert ig == ir
    for key, c in br.items():
        g = bg[key]
        assert g.color == c.color and g.point_count == c.point_count
        for k in range(c.point_count):
            a, b = c.points[k], g.points[k]
            assert (a.normal_impulse, a.id, a.anchor_a.x, a.anchor_b.z) == \
                   (b.normal_impulse, b.id, b.anchor_a.x, b.anchor_b.z)
    np.testing.assert_array_equal(wg.body_state(), wr.body_state())
