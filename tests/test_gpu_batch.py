# SPDX-License-Identifier: Apache-2.0
"""Batched worlds (nph_amd_batch_*): worlds of every kind and size stepped
together, each world's hash after every step equal to what the reference
gives that world stepped alone (oracle/_ref), whatever launch choices the
batch runs under; and the batch's own rules."""
import ctypes as C
import os

import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mesh_bank():
    return S.load_mesh_bank(os.path.join(GOLDEN, "meshes.bin"))


def _hit_events(w: P.World):
    add = w.add_shape
    w.add_shape = lambda b, g, **kw: add(b, g, hit_events=1, **kw)


# (label, filler): each filler fills a given world; the same fillers build
# the reference's worlds one by one.
def _worlds(hull_bank, mesh_bank):
    return [
        ("empty", lambda w: w),
        ("ground only", lambda w: S.ground(w)),
        ("pyramid 3 rows", lambda w: S.pyramid(None, 3, world=w)),
        ("pyramid 12 rows", lambda w: S.pyramid(None, 12, world=w)),
        ("bin 16", lambda w: S.bin_drop(None, 16, 21, hull_bank, half=3.0, world=w)),
        ("bin 500", lambda w: S.bin_drop(None, 500, 22, hull_bank, half=5.0, height=12.0,
                                         y_lo=1.0, world=w)),
        ("ragdolls 9", lambda w: S.ragdoll_piles(None, 9, 23, world=w)),
        ("mesh terrain 150", lambda w: S.mesh_terrain_drop(None, mesh_bank, 150, 24, hull_bank,
                                                           world=w)),
        ("hit events 80", lambda w: (_hit_events(w), S.ground(w),
                                     S.drop_mixed(w, S.Pcg32(25), 80, hull_bank, -2.0, 2.0, 1.0,
                                                  6.0))),
        ("plan 6", lambda w: S.fill_batch_world(w, *S.batch_plan(4, 8)[6], hull_bank)),
        ("plan 7", lambda w: S.fill_batch_world(w, *S.batch_plan(4, 8)[7], hull_bank)),
    ]


# (pass block, body block, graph, groups): the worlds change groups and
# streams as the tuning changes.
TUNINGS = [(256, 256, 0, 2), (64, 64, 1, 1), (128, 256, 1, 4), (256, 64, 0, 3)]


def test_batch_matches_each_world_alone(gpu_api, ref_api, hull_bank, mesh_bank):
    """Every world of a mixed batch (empty, static only, pyramids, bin drops
    of 16 and 500 bodies, ragdoll piles, a mesh terrain, hit events, two
    worlds of the batched workload's plan), 150 steps, the launch choices changing
    every step: hashes and events equal the reference's, world by world."""
    specs = _worlds(hull_bank, mesh_bank)
    print(f"batch: {len(specs)} worlds, 150 steps against the reference", flush=True)
    batch = P.Batch(gpu_api, len(specs), S.DESC)
    refs = []
    for (label, fill), w in zip(specs, batch.worlds):
        fill(w)
        r = P.World(ref_api, S.DESC)
        fill(r)
        refs.append(r)
    for k, w in enumerate(batch.worlds):
        if specs[k][0] == "hit events 80":
            gpu_api.world_set_hit_threshold(w.ptr, 0.5)
            ref_api.world_set_hit_threshold(refs[k].ptr, 0.5)
    for step in range(150):
        batch.set_tuning(*TUNINGS[step % len(TUNINGS)])
        batch.step(S.DT)
        for (label, _), w, r in zip(specs, batch.worlds, refs):
            r.step(S.DT)
            assert w.hash() == r.hash(), f"{label}: hash differs at step {step + 1}"
            if label == "hit events 80":
                for kind, a, b in zip(("begin", "end", "hit", "move"), w.events(), r.events()):
                    assert a == b, f"{label}: {kind} events differ at step {step + 1}"
    st = batch.stats()
    assert st.solved_contacts > 1000 and st.awake_bodies > 1000
    batch.close()


def test_autotune_picks_a_candidate_and_results_stay_exact(gpu_api, ref_api, hull_bank):
    print("batch: autotune on 6 bin drops, then 20 steps against the reference", flush=True)
    batch = P.Batch(gpu_api, 6, S.DESC)
    refs = []
    for k, w in enumerate(batch.worlds):
        S.bin_drop(None, 40 + 60 * k, 30 + k, hull_bank, half=4.0, height=10.0, y_lo=1.0, world=w)
        refs.append(S.bin_drop(ref_api, 40 + 60 * k, 30 + k, hull_bank, half=4.0, height=10.0,
                               y_lo=1.0))
    t = batch.autotune(S.DT, rounds=1)  # 32 candidates, one step each
    assert t.pass_block in (128, 256) and t.body_block in (64, 256) and t.use_graph in (0, 1)
    assert 1 <= t.groups <= 4
    got = batch.tuning()
    assert (got.pass_block, got.body_block, got.use_graph, got.groups) == (
        t.pass_block, t.body_block, t.use_graph, t.groups)
    for _ in range(32):
        for r in refs:
            r.step(S.DT)
    for step in range(20):
        batch.step(S.DT)
        for k, (w, r) in enumerate(zip(batch.worlds, refs)):
            r.step(S.DT)
            assert w.hash() == r.hash(), f"world {k}: hash differs {step + 1} steps after tuning"
    batch.close()


def test_a_batch_world_steps_alone_too(gpu_api, ref_api, hull_bank):
    """A batch's world also steps by itself (nph_physics_world_step), then
    with the batch again, exactly as the reference's world."""
    batch = P.Batch(gpu_api, 2, S.DESC)
    for w in batch.worlds:
        S.bin_drop(None, 120, 41, hull_bank, half=4.0, height=10.0, y_lo=1.0, world=w)
    r = S.bin_drop(ref_api, 120, 41, hull_bank, half=4.0, height=10.0, y_lo=1.0)
    for step in range(60):
        if step % 3 == 0:
            batch.worlds[0].step(S.DT)
            batch.worlds[1].step(S.DT)
        else:
            batch.step(S.DT)
        r.step(S.DT)
        assert batch.worlds[0].hash() == r.hash() == batch.worlds[1].hash(), step
    batch.close()


def _network(seed):
    """A small fixed network (random weights, unit normalisation)."""
    import numpy as np
    from nepheline_amd import learned as L
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(L.PARAMS) * 0.05).astype(np.float32)
    ws = L.WarmStart()
    for i in range(16):
        ws.feature_shift[i] = 0.0
        ws.feature_scale[i] = 1.0
    for i in range(4):
        ws.output_scale[i] = 0.05
    ws.mode = L.NEW_POINTS
    ws.gate = True
    return ws, params


def test_learned_batch_matches_each_world_alone(gpu_api, hull_bank):
    """One network for the batch: every world's hash after every step equals
    the same world stepped alone with that network (the network's rows are
    independent of one another), and the batch's residual statistics add up
    the worlds' own."""
    from nepheline_amd import learned as L
    print("batch: learned warm start, 5 worlds, 80 steps against the same worlds alone",
          flush=True)
    ws, params = _network(5)
    sizes = [(60, 51), (200, 52), (20, 53), (400, 54), (5, 55)]
    batch = P.Batch(gpu_api, len(sizes), S.DESC)
    alone = []
    for (n, seed), w in zip(sizes, batch.worlds):
        S.bin_drop(None, n, seed, hull_bank, half=3.0 + n / 100.0, height=8.0, y_lo=1.0, world=w)
        a = S.bin_drop(gpu_api, n, seed, hull_bank, half=3.0 + n / 100.0, height=8.0, y_lo=1.0)
        L.set_warm_start(a, ws, params)
        L.measure_residual(a, True)
        alone.append(a)
    L.batch_set_warm_start(batch, ws, params)
    L.batch_measure_residual(batch, True)
    L.batch_warm_stats(batch, reset=True)
    learned = 0
    for step in range(80):
        batch.step(S.DT)
        tot_p = tot_l = 0
        for k, (w, a) in enumerate(zip(batch.worlds, alone)):
            a.step(S.DT)
            assert w.hash() == a.hash(), f"world {k}: hash differs at step {step + 1}"
            st = L.warm_stats(a)
            tot_p += st.contacts_plain
            tot_l += st.contacts_learned
        learned += tot_l
        bs = L.batch_warm_stats(batch, reset=True)
        assert (bs.contacts_plain, bs.contacts_learned) == (tot_p, tot_l), step
    assert learned > 100
    # A second network on one world: the batch refuses to mix them.
    ws2, params2 = _network(6)
    L.set_warm_start(batch.worlds[1], ws2, params2)
    assert gpu_api.batch_step(batch.ptr, S.DT) == P.UNSUPPORTED
    for a in alone:
        a.close()
    batch.close()


def test_the_contact_log_is_not_batched(gpu_api, hull_bank):
    from nepheline_amd import learned as L
    batch = P.Batch(gpu_api, 2, S.DESC)
    for w in batch.worlds:
        S.bin_drop(None, 30, 3, hull_bank, half=3.0, world=w)
    L.log_contacts(batch.worlds[1], 1000)
    assert gpu_api.batch_step(batch.ptr, S.DT) == P.UNSUPPORTED
    L.log_contacts(batch.worlds[1], 0)
    batch.step(S.DT)
    batch.close()
