# SPDX-License-Identifier: Apache-2.0
"""Continuous collision on the GPU path (continuous.c:1-409): fast bodies
marked in the device's finalize, swept on the host against static shapes
(bullets also against bodies at their final poses), set back to the first
impact, their boxes refreshed before the grown proxies; bit for bit against
the reference's world hashes."""
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
