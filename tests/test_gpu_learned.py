# SPDX-License-Identifier: Apache-2.0
"""The learned warm start in the world step, on the GPU.

The plain step is the reference's; the learned start is this library's
addition, so its tests pin what must not change (logging, measuring and a
network that predicts nothing leave the step bit-exact), what the log holds
(the impulses the step stored, the contacts' normals and friction), that the
residual gate hands a contact back to the plain start for good, and what the
trained network gains on a held-out scene over the plain start at the same
substep count (the ablation)."""
import numpy as np
import pytest

from nepheline_amd import learned as L
from nepheline_amd import scenes as S

pytestmark = pytest.mark.gpu


def _null_net(mode=L.NEW_POINTS, gate=True):
    ws = L.WarmStart()
    for j in range(16):
        ws.feature_scale[j] = 1.0
    for j in range(4):
        ws.output_scale[j] = 1.0
    ws.mode = mode
    ws.gate = gate
    return ws, np.zeros(L.PARAMS, np.float32)


def test_logging_and_measuring_leave_the_step_unchanged(gpu_api, hull_bank):
    print("learned: log + measure vs plain, bin 300 seed 3", flush=True)
    hashes = []
    for instrument in (False, True):
        w = S.bin_drop(gpu_api, 300, 3, hull_bank)
        if instrument:
            L.measure_residual(w, True)
            L.log_contacts(w, 200_000)
        for _ in range(80):
            w.step(S.DT)
        hashes.append(w.hash())
        if instrument:
            rows, _, _ = L.contact_log(w)
            assert rows > 0
            s = L.warm_stats(w)
            assert s.contacts_plain > 0 and s.contacts_learned == 0
        w.close()
    assert hashes[0] == hashes[1]


def test_a_network_that_predicts_nothing_is_the_plain_step(gpu_api, hull_bank):
    """New points start at zero either way; the network's lanes are flagged."""
    print("learned: zero network vs plain, bin 300 seed 4", flush=True)
    w = S.bin_drop(gpu_api, 300, 4, hull_bank)
    for _ in range(80):
        w.step(S.DT)
    plain = w.hash()
    w.close()
    ws, params = _null_net()
    w = S.bin_drop(gpu_api, 300, 4, hull_bank)
    L.set_warm_start(w, ws, params)
    learned = 0
    for _ in range(80):
        w.step(S.DT)
        learned += L.warm_stats(w).contacts_learned
    assert w.hash() == plain
    assert learned > 0
    w.close()


def test_the_log_holds_the_stored_impulses(gpu_api, hull_bank):
    print("learned: one logged step vs contact readback, bin 400 seed 5", flush=True)
    w = S.bin_drop(gpu_api, 400, 5, hull_bank)
    for _ in range(90):
        w.step(S.DT)
    L.log_contacts(w, 100_000)
    w.step(S.DT)
    feats, targs = L.download_log(w)
    contacts = [c for c in w.contacts() if c.point_count > 0]
    pts = [(c, c.points[k]) for c in contacts for k in range(c.point_count)]
    assert len(pts) == len(feats) > 100
    # Targets: the normal impulses the step stored, bit for bit.
    assert np.array_equal(np.sort(targs[:, 0]),
                          np.sort(np.array([p.normal_impulse for _, p in pts], np.float32)))
    # Friction impulses: the tangent pair's length is the stored vector's.
    want = np.sort(np.array([np.hypot(np.hypot(p.tangent_impulse.x, p.tangent_impulse.y),
                                      p.tangent_impulse.z) for _, p in pts]))
    got = np.sort(np.hypot(targs[:, 1].astype(np.float64), targs[:, 2]))
    assert np.allclose(got, want, rtol=1e-5, atol=1e-9)
    # Features: each point's manifold normal and friction.
    nrm = sorted((c.normal.x, c.normal.y, c.normal.z, c.friction) for c, _ in pts)
    got = sorted(tuple(float(v) for v in r) for r in feats[:, [4, 5, 6, 9]])
    assert np.array_equal(np.array(got, np.float32), np.array(nrm, np.float32))
    # The carried-nothing flag is 0 or 1, and carried points show their impulse.
    assert set(np.unique(feats[:, 15])) <= {0.0, 1.0}
    assert np.all((feats[:, 15] == 1.0) <= (feats[:, 14] == 0.0))
    w.close()


def test_the_gate_hands_contacts_back_for_good(gpu_api):
    """A network that pushes every point on a settled pyramid: the gate takes
    contacts back from it (for the rest of their lives), so fewer and fewer
    contacts start from the network, and the stack stays whole."""
    print("learned: gate on a settled pyramid with a bad network", flush=True)
    w = S.pyramid(gpu_api, rows=10)
    for _ in range(120):
        w.step(S.DT)
    ws, params = _null_net(mode=L.ALL_POINTS, gate=True)
    params[6272] = 0.05  # b3[0]: every normal impulse 0.05 m/s times the normal mass
    L.set_warm_start(w, ws, params)
    w.step(S.DT)
    s1 = L.warm_stats(w)
    assert s1.contacts_learned > 0 and s1.gated > 0
    assert s1.contacts_plain > 0  # the control sample
    gated = s1.gated
    for _ in range(60):
        w.step(S.DT)
        s = L.warm_stats(w)
        gated += s.gated
    assert s.contacts_learned <= s1.contacts_learned - s1.gated
    assert gated >= s1.gated
    state = w.body_state()
    assert np.all(np.isfinite(state))
    assert state[:, 1].max() < 12.0  # nothing thrown out of the stack
    w.close()


def test_the_trained_start_lowers_the_residual(gpu_api, hull_bank):
    """Ablation on a held-out scene: a network trained on the plain solver's
    converged impulses from other seeds, applied to points that carry no
    impulse, against the plain start, same scene, same 3 substeps."""
    print("learned: train on bin drops 100-103, evaluate bin 2000 seed 7", flush=True)
    ws, params, hist = L.train(gpu_api, hull_bank, seeds=(100, 101, 102, 103), bodies=2000,
                               steps=200, epochs=8, min_flag=1.0)
    assert hist[-1] < 0.7 * hist[0], hist
    means = {}
    for name in ("plain", "learned"):
        w = S.bin_drop(gpu_api, 2000, 7, hull_bank, desc={"substeps": 3})
        if name == "plain":
            L.measure_residual(w, True)
        else:
            ws.mode = L.NEW_POINTS
            ws.gate = True
            L.set_warm_start(w, ws, params)
        total, count = 0.0, 0
        for _ in range(200):
            w.step(S.DT)
            s = L.warm_stats(w)
            total += s.residual_sum_plain + s.residual_sum_learned
            count += s.contacts_plain + s.contacts_learned
        means[name] = total / count
        w.close()
    print("mean residual", means, flush=True)
    assert means["learned"] < means["plain"], means
