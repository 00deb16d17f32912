# SPDX-License-Identifier: Apache-2.0
"""Parity at the benchmark workloads' full sizes, against the reference's own
step (oracle/_ref) run in the build container (tests/golden/
make_full_fixtures.py wrote tests/golden/full_hashes.json):

  * the 10 000-body mixed bin drop, seed 2, all 600 steps: the world hash
    (world.c:436-462) after every step, and the canonical pair and
    contact-id lists after the last (physics.canonical_digests);
  * the 100 201-body field (200 mesh patches, 2 000 ragdolls, 78 000 mixed
    bodies), seed 3, 300 steps: every step's hash and the final lists;
  * the 1 024-world batch plan (seed 4), every world stepped in one batch:
    each world's hash after steps 50, 100 and 200 against that world stepped
    alone by the reference.
Bit-exact is the bar: the hash covers every body's f64 origin, rotation and
velocities."""
import json
import os

import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    with open(os.path.join(GOLDEN, "full_hashes.json")) as f:
        return json.load(f)


def _check_run(w, want, label):
    for i, h in enumerate(want["hashes"]):
        w.step(S.DT)
        got = "%016x" % w.hash()
        assert got == h, f"{label}: hash differs at step {i + 1}"
        if (i + 1) % 100 == 0:
            print(f"{label}: step {i + 1}/{len(want['hashes'])} equal", flush=True)
    d = P.canonical_digests(w)
    for k in ("pairs", "ids", "pairs_sha256", "ids_sha256"):
        assert d[k] == want[k], f"{label}: {k} differs after the last step"


def test_bin_drop_10000_bodies_600_steps(gpu_api, hull_bank, full):
    print("full size: bin drop, 10 000 bodies, seed 2, 600 steps", flush=True)
    w = S.bin_drop(gpu_api, 10000, 2, hull_bank, half=20.0, height=20.0, y_lo=5.0)
    _check_run(w, full["bin_10000_2"], "bin_10000_2")
    w.close()


def test_field_100k_bodies_300_steps(gpu_api, hull_bank, full):
    print("full size: field, 100 201 bodies, seed 3, 300 steps", flush=True)
    meshes = S.load_mesh_bank(os.path.join(GOLDEN, "meshes.bin"))
    w = S.field(gpu_api, meshes, hull_bank, seed=3)
    assert len(w.bodies) == full["field_3"]["bodies"]
    _check_run(w, full["field_3"], "field_3")
    w.close()


def test_batch_1024_worlds_200_steps(gpu_api, hull_bank, full):
    want = full["batch_1024_4"]
    plan = S.batch_plan(4, 1024)
    print("full size: batch of 1 024 worlds, seed 4, 200 steps", flush=True)
    batch = P.Batch(gpu_api, len(plan), S.DESC)
    for (kind, n, seed), w, ref in zip(plan, batch.worlds, want["worlds"]):
        assert (kind, n, seed) == (ref["kind"], ref["plan_bodies"], ref["seed"])
        S.fill_batch_world(w, kind, n, seed, hull_bank)
    print("batch built", flush=True)
    checks = {s: k for k, s in enumerate(want["steps"])}
    for step in range(1, want["steps"][-1] + 1):
        batch.step(S.DT)
        if step in checks:
            k = checks[step]
            bad = [i for i, (w, ref) in enumerate(zip(batch.worlds, want["worlds"]))
                   if "%016x" % w.hash() != ref["hashes"][k]]
            assert not bad, f"step {step}: worlds {bad[:10]} ({len(bad)} in all) differ"
            print(f"batch: step {step}, 1 024 world hashes equal", flush=True)
    batch.close()
