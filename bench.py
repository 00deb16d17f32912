# SPDX-License-Identifier: Apache-2.0
"""Contact solves per second of the world step on one MI355X.

Workload: 10,000 mixed bodies (spheres r 0.2-0.6, capsules r 0.2-0.4 / half
height 0.3-0.8, boxes half 0.2-0.7, 12-vertex hulls) dropped from 5-20 m
into an open 40 x 40 x 20 m bin of static boxes; friction 0.5, dt 1/60,
4 substeps, FP32 with f64 body centres.  A step is one world step; its
contact solves are the constraint-graph contacts the solver took times the
substeps (solver.c:810-843).  The timed region is whole steps: pairs,
narrow phase, serial bookkeeping, solver, finalize.

Beside it: the roofline of the dominant kernel, the colour pass (one launch
per colour per constraint pass; HIP events around each launch, in a window of
steps right after the timed region so the headline stays uninstrumented),
the reference's own CPU step (oracle/_ref, its job system on up to 16 host
threads) on a bounded sample of the same scene, and the learned warm start:
a network trained on the device from plain bin drops of other seeds, the
same scene and steps timed with it, the post-solve residual and penetration
of both starts, and the network's forward kernel against the MFMA peak; the
100 000-body field (meshes, ragdolls, mixed bodies) on the GPU beside the
reference's CPU step on a bounded sample of it; and 1 024 worlds of 16 to
20 000 bodies stepped as one batch, beside the reference stepping a sample
of them.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nepheline_amd import physics as P  # noqa: E402
from nepheline_amd import scenes as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA
# The network's forward flops per contact point: 2 (14*64 + 64*64 + 64*3).
MLP_FLOPS_PER_ROW = 2 * (14 * 64 + 64 * 64 + 64 * 3)
DT = 1.0 / 60.0
SUBSTEPS = 4
# Algorithmic bytes per contact-substep of the three constraint passes
# (warm start, solve, relax), FP32 structure-of-arrays:
# B = 3 (C + P (K + 2 M)) + E (3 * 2 V + 2 X), C = 68 contact constants,
# K = 116 point constants, M = 16 point accumulators (read and written),
# V = 24 (v, w) read and written per pass, X = 28 (dp, dq) read by solve and
# relax; P points, E dynamic endpoints of the contact.
C_BYTES, K_BYTES, M_BYTES, V_BYTES, X_BYTES = 68, 116, 16, 24, 28
# HBM bytes per contact-substep of the colour passes, from rocprofv3 counters
# (FETCH_SIZE and WRITE_SIZE, scaled by a device copy of known size).
TRAFFIC_FILE = "k_color_pass_bin10k_traffic.json"
BATCH_TRAFFIC_FILE = "k_b_color_pass_batch1024_traffic.json"


def contact_bytes(points, endpoints):
    return 3 * (C_BYTES + points * (K_BYTES + 2 * M_BYTES)) + endpoints * (
        3 * 2 * V_BYTES + 2 * X_BYTES)


def graph_mix(world):
    """Mean algorithmic bytes per graph contact-substep, from the contacts the
    graph holds now (touching ones with a colour)."""
    tot = 0.0
    n = 0
    for c in world.contacts():
        if c.color == P.NULL32 or c.point_count == 0:
            continue
        # Dynamic endpoints: the ground and the walls are bodies 0..4.
        e = (1 if c.body_a >= 5 else 0) + (1 if c.body_b >= 5 else 0)
        tot += contact_bytes(c.point_count, e)
        n += 1
    return (tot / n if n else 0.0), n


def graph_bytes(world, static_slots):
    """Algorithmic bytes of one substep over the graph contacts the world
    holds now, and their count (numpy over the contact readback). Bodies in
    slots below `static_slots` are the scene's static ones."""
    n = world.api.world_contacts(world.ptr, None, 0)
    if n == 0:
        return 0.0, 0
    buf = (P.ContactInfo * n)()
    world.api.world_contacts(world.ptr, buf, n)
    arr = np.frombuffer(buf, dtype=np.dtype(P.ContactInfo), count=n)
    sel = (arr["color"] != P.NULL32) & (arr["point_count"] > 0)
    p = arr["point_count"][sel].astype(np.int64)
    e = (arr["body_a"][sel] >= static_slots).astype(np.int64) + (
        arr["body_b"][sel] >= static_slots).astype(np.int64)
    b = 3 * (C_BYTES + p * (K_BYTES + 2 * M_BYTES)) + e * (3 * 2 * V_BYTES + 2 * X_BYTES)
    return float(b.sum()), int(sel.sum())


# Static bodies each kind of batched world makes first (ground; bin floor
# and walls; ground).
BATCH_STATIC = {0: 1, 1: 5, 2: 1}


def build(api, bodies, seed, bank):
    return S.bin_drop(api, bodies, seed, bank, half=20.0, height=20.0, y_lo=5.0)


def run_gpu(args, bank):
    a = P.product_api()
    a.lib.nph_amd_solved_contacts.restype = C.c_uint32
    a.lib.nph_amd_solved_contacts.argtypes = [C.c_void_p]
    a.lib.nph_amd_device_times.restype = None
    a.lib.nph_amd_device_times.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    w = build(a, args.bodies, args.seed, bank)
    for _ in range(args.warmup):
        w.step(DT)
    times = (C.c_double * 3)()
    solved = 0
    passes_ms = 0.0
    phases = np.zeros(6)
    t0 = time.perf_counter()
    for i in range(args.steps):
        w.step(DT)
        solved += a.lib.nph_amd_solved_contacts(w.ptr)
        a.lib.nph_amd_device_times(w.ptr, times)
        passes_ms += times[2]
        st = a.world_stats(w.ptr) if args.phases else None
        if st is not None:
            phases += (st.pairs_ms, st.collide_ms, st.solve_ms, times[0], times[1], times[2])
        if args.progress and (i + 1) % 100 == 0:
            print(f"# gpu step {i + 1}/{args.steps}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    mean_bytes, graph_n = graph_mix(w)
    # The roofline window: the next steps of the same run with every
    # colour-pass launch bracketed by HIP events (kept out of the timed
    # region above, whose throughput is the headline).
    a.lib.nph_amd_time_kernels.argtypes = [C.c_void_p, C.c_bool]
    a.lib.nph_amd_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    a.lib.nph_amd_time_kernels(w.ptr, True)
    k_solved = 0
    for _ in range(args.roofline_steps):
        w.step(DT)
        k_solved += a.lib.nph_amd_solved_contacts(w.ptr)
    ktimes = (C.c_double * 2)()
    a.lib.nph_amd_kernel_times(w.ptr, ktimes)
    a.lib.nph_amd_time_kernels(w.ptr, False)
    k_bytes, _ = graph_mix(w)
    return {
        "wall_s": wall,
        "solved": solved,
        "passes_ms": passes_ms,
        "mean_bytes": mean_bytes,
        "graph_contacts_end": graph_n,
        "kernel_ms": ktimes[0],
        "kernel_launches": int(ktimes[1]),
        "kernel_solved": k_solved,
        "kernel_bytes": k_bytes,
        "hash": "%016x" % w.hash(),
        "phases_ms": {k: round(v / max(args.steps, 1), 4) for k, v in zip(
            ("host_pairs", "host_collide", "host_solve", "dev_collide", "dev_solve", "dev_passes"),
            phases)},
    }


def residual_run(a, bank, args, ws=None, params=None):
    """Mean and largest post-solve residual (m/s) and the largest penetration
    over the headline's steps, reading the world's residual statistics after
    every step (outside any timed region)."""
    from nepheline_amd import learned as L
    w = build(a, args.bodies, args.seed, bank)
    if ws is None:
        L.measure_residual(w, True)
    else:
        L.set_warm_start(w, ws, params)
    total, count, worst, pen, learned, gated = 0.0, 0, 0.0, 0.0, 0, 0
    for i in range(args.warmup + args.steps):
        w.step(DT)
        if i < args.warmup:
            continue
        s = L.warm_stats(w)
        total += s.residual_sum_plain + s.residual_sum_learned
        count += s.contacts_plain + s.contacts_learned
        worst = max(worst, s.residual_max_plain, s.residual_max_learned)
        pen = max(pen, s.penetration_max)
        learned += s.contacts_learned
        gated += s.gated
    w.close()
    return {"residual_mean": total / max(count, 1), "residual_max": worst,
            "penetration_max": pen, "contacts_from_network": learned, "gated": gated}


def run_learned(args, bank):
    """The learned warm start: trained on the device from plain bin drops of
    other seeds, then the headline's scene and steps timed with it, and the
    residuals of both starts."""
    from nepheline_amd import learned as L
    a = P.product_api()
    t0 = time.perf_counter()
    tr = L.Trainer(a, capacity=args.train_capacity, seed=1)
    logged = L.log_bin_drops(a, tr, bank, args.train_seeds, args.bodies, args.train_steps,
                             min_flag=1.0, log_capacity=args.train_log)
    tr.prepare()
    loss0 = tr.loss()
    for e in range(args.train_epochs):
        tr.epoch(8192, 2e-3 if e < args.train_epochs - 2 else 6e-4)
    loss1 = tr.loss()
    rows = tr.rows
    ws, params = tr.warm_start(mode=L.NEW_POINTS, gate=True)
    args.network = (ws, params)  # the batched worlds take the same network
    tr.close()
    train_s = time.perf_counter() - t0
    # Throughput with the network, same scene and steps as the headline.
    a.lib.nph_amd_solved_contacts.restype = C.c_uint32
    a.lib.nph_amd_solved_contacts.argtypes = [C.c_void_p]
    w = build(a, args.bodies, args.seed, bank)
    L.set_warm_start(w, ws, params)
    for _ in range(args.warmup):
        w.step(DT)
    solved = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w.step(DT)
        solved += a.lib.nph_amd_solved_contacts(w.ptr)
    wall = time.perf_counter() - t0
    w.close()
    plain = residual_run(a, bank, args)
    learned = residual_run(a, bank, args, ws, params)
    # The network's forward kernel alone on 4M rows (MFMA-bound).
    n = 1 << 22
    x = L.DeviceArray(a, shape=(n, L.INPUTS), dtype=np.uint16)
    x.upload(np.zeros((n, L.INPUTS), np.uint16))
    net = L.Network(a, params)
    y = L.DeviceArray(a, shape=(n, L.OUTPUTS), dtype=np.float32)
    f = net.f
    reps = 20
    for _ in range(3):
        f.mlp_forward(x.ptr, n, net.params.ptr, y.ptr, None)
    y.numpy()
    t0 = time.perf_counter()
    for _ in range(reps):
        f.mlp_forward(x.ptr, n, net.params.ptr, y.ptr, None)
    y.numpy()
    fwd_s = (time.perf_counter() - t0) / reps
    flops = MLP_FLOPS_PER_ROW * n
    return {
        "value": solved * SUBSTEPS / wall,
        "unit": "contact-substeps/s",
        "ms_per_step": wall * 1e3 / args.steps,
        "mode": "points that carry no impulse start from the network; residual gate on",
        "residual_plain": plain,
        "residual_learned": learned,
        "training": {"seeds": list(args.train_seeds), "steps_per_seed": args.train_steps,
                     "rows_logged": logged, "rows_trained": rows,
                     "epochs": args.train_epochs, "loss_before": loss0, "loss_after": loss1,
                     "seconds": train_s},
        "forward_roofline": {"bound": "mfma", "achieved": flops / fwd_s / 1e12,
                             "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": flops / fwd_s / 1e12 / BF16_PEAK_TFLOPS,
                             "rows": n, "flop_per_row": MLP_FLOPS_PER_ROW,
                             "us_per_launch": fwd_s * 1e6},
    }


def run_field(args, bank):
    """The 100 000-body field (150 x 150 m: 200 mesh patches, 2 000 ragdolls
    of 11 capsules and 10 joints, 78 000 mixed bodies): the GPU's contact
    solves/s over the timed steps, and the reference's own CPU step on the
    same scene over a bounded sample of the same steps."""
    meshes = S.load_mesh_bank(os.path.join(ROOT, "tests", "golden", "meshes.bin"))
    a = P.product_api()
    a.lib.nph_amd_solved_contacts.restype = C.c_uint32
    a.lib.nph_amd_solved_contacts.argtypes = [C.c_void_p]
    t0 = time.perf_counter()
    w = S.field(a, meshes, bank, seed=3)
    setup = time.perf_counter() - t0
    for _ in range(args.field_warmup):
        w.step(DT)
    solved = 0
    t0 = time.perf_counter()
    for _ in range(args.field_steps):
        w.step(DT)
        solved += a.lib.nph_amd_solved_contacts(w.ptr)
    wall = time.perf_counter() - t0
    bodies = len(w.bodies)
    w.close()
    out = {"workload": "field: %d bodies (200 mesh patches of 64 triangles, 2000 ragdolls, "
                       "78000 mixed), 150 x 150 m, dt 1/60, 4 substeps" % bodies,
           "value": solved * SUBSTEPS / wall, "unit": "contact-substeps/s",
           "steps": args.field_steps, "warmup": args.field_warmup,
           "ms_per_step": wall * 1e3 / args.field_steps, "setup_s": setup, "cpu_baseline": None}
    path = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")
    if args.no_cpu or not os.path.exists(path):
        return out
    ra = P.api(path, "ref_shim_allocator")
    ra.lib.ref_shim_graph_contacts.restype = C.c_uint32
    ra.lib.ref_shim_graph_contacts.argtypes = [C.c_void_p]
    ra.lib.ref_shim_jobs_create.restype = C.c_void_p
    ra.lib.ref_shim_jobs_create.argtypes = [C.c_uint32]
    cores = min(len(os.sched_getaffinity(0)), 16)
    jobs = ra.lib.ref_shim_jobs_create(max(cores - 1, 1)) if cores > 1 else None
    wr = S.field(ra, meshes, bank, seed=3)
    for _ in range(args.field_warmup):
        wr.step(DT, jobs)
    solved = 0
    steps = 0
    t0 = time.perf_counter()
    while steps < args.field_steps:
        wr.step(DT, jobs)
        solved += ra.lib.ref_shim_graph_contacts(wr.ptr)
        steps += 1
        if time.perf_counter() - t0 > args.cpu_seconds:
            break
    cwall = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": solved * SUBSTEPS / cwall, "unit": "contact-substeps/s",
                           "cores": cores, "kind": "reference",
                           "sample": "the reference's step (job system, %d threads) on the same "
                                     "field, steps %d-%d (%.1f s)" % (
                                         cores, args.field_warmup + 1,
                                         args.field_warmup + steps, cwall)}
    out["vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def _batch_residuals(st):
    """Mean and largest post-solve residual and largest penetration over the
    measured worlds and steps."""
    n = st.contacts_plain + st.contacts_learned
    return {"mean": (st.residual_sum_plain + st.residual_sum_learned) / n if n else 0.0,
            "mean_plain_started": st.residual_sum_plain / st.contacts_plain
            if st.contacts_plain else None,
            "mean_learned_started": st.residual_sum_learned / st.contacts_learned
            if st.contacts_learned else None,
            "max": max(st.residual_max_plain, st.residual_max_learned),
            "penetration_max": st.penetration_max, "contact_steps": n,
            "learned_fraction": st.contacts_learned / n if n else 0.0,
            "gated": st.gated, "steps": st.steps}


def _batch_run(args, bank, plan, network=None, roofline=False):
    """One batch of the plan: built, autotuned, warmed up, then the timed
    steps with residual measurement on (the same for both starts)."""
    from nepheline_amd import learned as L
    a = P.product_api()
    t0 = time.perf_counter()
    batch = P.Batch(a, len(plan), S.DESC)
    for (kind, n, seed), w in zip(plan, batch.worlds):
        S.fill_batch_world(w, kind, n, seed, bank)
    setup = time.perf_counter() - t0
    bodies = sum(len(w.bodies) for w in batch.worlds)
    print(f"# batch: {len(plan)} worlds, {bodies} bodies built in {setup:.1f} s"
          + (" (learned start)" if network else ""), file=sys.stderr, flush=True)
    if network is not None:
        L.batch_set_warm_start(batch, network[0], network[1])
    # Warm-up steps first, so the autotuner sees the pile the timed steps
    # will (the choice that wins while bodies fall is not the one that wins
    # once they rest).
    for _ in range(args.batch_warmup):
        batch.step(DT)
    if args.batch_fixed:
        batch.set_tuning(256, 64, 0, 1)
        tune = batch.tuning()
        warm = args.batch_warmup
    else:
        rounds = 2
        tune = batch.autotune(DT, rounds=rounds)  # 32 candidates a round, one step each
        warm = args.batch_warmup + 32 * rounds
    L.batch_measure_residual(batch, True)
    L.batch_warm_stats(batch, reset=True)
    solved = 0
    phases = np.zeros(11)
    t0 = time.perf_counter()
    for i in range(args.batch_steps):
        batch.step(DT)
        st = batch.stats()
        solved += st.solved_contacts
        phases += (st.step_ms, st.narrow_ms, st.pairs_ms, st.changes_ms, st.solve_ms, st.post_ms,
                   st.device_solve_ms, st.device_passes_ms, st.device_collide_ms,
                   st.wait_narrow_ms, st.wait_solve_ms)
        if (i + 1) % 25 == 0:
            print(f"# batch step {i + 1}/{args.batch_steps}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    res = _batch_residuals(L.batch_warm_stats(batch, reset=True))
    L.batch_measure_residual(batch, False)
    names = ("step", "narrow", "host_pairs", "host_changes", "solve", "host_post", "dev_solve",
             "dev_passes", "dev_collide", "wait_narrow", "wait_solve")
    out = {"value": solved * SUBSTEPS / wall, "unit": "contact-substeps/s",
           "steps": args.batch_steps, "warmup": warm, "ms_per_step": wall * 1e3 / args.batch_steps,
           "setup_s": setup, "bodies": bodies, "worlds": len(plan),
           "tuning": {"pass_block": tune.pass_block, "body_block": tune.body_block,
                      "use_graph": bool(tune.use_graph), "groups": tune.groups},
           "phases_ms_per_step": {k: round(v / max(args.batch_steps, 1), 3)
                                  for k, v in zip(names, phases)},
           "residual": res}
    if roofline:
        # The tuned blocks, launched from one stream (one group, so no two
        # colour passes overlap) with HIP events around each.
        batch.set_tuning(tune.pass_block, tune.body_block, 0, 1)
        a.batch_time_kernels(batch.ptr, True)
        k_solved = 0
        for _ in range(args.batch_roofline_steps):
            batch.step(DT)
            k_solved += batch.stats().solved_contacts
        kt = (C.c_double * 2)()
        a.batch_kernel_times(batch.ptr, kt)
        a.batch_time_kernels(batch.ptr, False)
        tot_b = 0.0
        tot_n = 0
        for (kind, _, _), w in zip(plan, batch.worlds):
            b, n = graph_bytes(w, BATCH_STATIC[kind])
            tot_b += b
            tot_n += n
        mean_bytes = tot_b / tot_n if tot_n else 0.0
        launches = max(int(kt[1]), 1)
        bytes_per_launch = k_solved * SUBSTEPS * mean_bytes / launches
        avg_launch_s = kt[0] * 1e-3 / launches
        achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
        traffic = None
        tpath = os.path.join(ROOT, "profiles", BATCH_TRAFFIC_FILE)
        if os.path.exists(tpath):
            t = json.load(open(tpath))
            traffic = t["hbm_bytes_per_contact_substep"] * k_solved * SUBSTEPS / launches
        out["roofline"] = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_b_color_pass (one colour of one constraint pass over every world), "
                      "HIP events around each launch over %d steps after the timed region, "
                      "one group" % args.batch_roofline_steps,
            "launches": int(kt[1]), "avg_launch_us": avg_launch_s * 1e6,
            "bytes_per_launch": bytes_per_launch, "bytes_per_contact_substep": mean_bytes}
    batch.close()
    return out, warm


def run_batch(args, bank):
    """The batched workload: 1 024 worlds (seed 4), each a pyramid, a bin
    drop or ragdoll piles of floor(16 * 1250^u) bodies, u uniform, stepped
    together (nph_amd_batch_step) after the autotuner picks the launch
    choices: contact solves/s over the timed steps with the plain start and
    with the learned start (the network the learned leg trained), the
    residuals of both, the batched colour pass's roofline, and the
    reference's CPU step on a sample of the worlds."""
    plan = S.batch_plan(args.batch_seed, args.batch_worlds)
    out, warm = _batch_run(args, bank, plan, roofline=True)
    out = dict({"workload": "batch: %d worlds (seed %d; pyramids, bin drops, ragdoll piles of "
                            "floor(16*1250^u) bodies), %d bodies, dt 1/60, 4 substeps, plain "
                            "start; residual measured in the timed steps"
                            % (len(plan), args.batch_seed, out["bodies"])}, **out)
    net = getattr(args, "network", None)
    if net is not None:
        lo, _ = _batch_run(args, bank, plan, network=net)
        lo["vs_plain_batch"] = lo["value"] / out["value"]
        lo["mode"] = "points that carry no impulse start from the network; residual gate on"
        out["learned"] = lo
    out["cpu_baseline"] = None
    if not args.no_cpu:
        cpu = run_batch_cpu(args, bank, plan, warm)
        if cpu is not None:
            out["cpu_baseline"] = cpu
            out["vs_cpu"] = out["value"] / cpu["value"]
    return out


def run_batch_cpu(args, bank, plan, warm):
    """The reference's step on every 16th world of the batch (worlds in
    parallel on up to 16 host threads, each world's step on one), the same
    steps as the GPU's timed region, stopped after `cpu_seconds`."""
    from concurrent.futures import ThreadPoolExecutor
    path = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")
    if not os.path.exists(path):
        return None
    ra = P.api(path, "ref_shim_allocator")
    ra.lib.ref_shim_graph_contacts.restype = C.c_uint32
    ra.lib.ref_shim_graph_contacts.argtypes = [C.c_void_p]
    cores = min(len(os.sched_getaffinity(0)), 16)
    sample = plan[::16]
    worlds = []
    for kind, n, seed in sample:
        w = P.World(ra, S.DESC)
        S.fill_batch_world(w, kind, n, seed, bank)
        worlds.append(w)
    bodies = sum(len(w.bodies) for w in worlds)
    order = sorted(range(len(worlds)), key=lambda k: -len(worlds[k].bodies))

    def one(k):
        worlds[k].step(DT)
        return ra.lib.ref_shim_graph_contacts(worlds[k].ptr)

    with ThreadPoolExecutor(max_workers=cores) as pool:
        for _ in range(warm):
            list(pool.map(one, order))
        solved = 0
        steps = 0
        t0 = time.perf_counter()
        while steps < args.batch_steps:
            solved += sum(pool.map(one, order))
            steps += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        cwall = time.perf_counter() - t0
    for w in worlds:
        w.close()
    return {"value": solved * SUBSTEPS / cwall, "unit": "contact-substeps/s", "cores": cores,
            "kind": "reference",
            "sample": "the reference's step on every 16th world of the batch (%d worlds, %d "
                      "bodies), worlds in parallel on %d threads, steps %d-%d (%.1f s)" % (
                          len(worlds), bodies, cores, warm + 1, warm + steps, cwall)}


def run_cpu(args, bank):
    path = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")
    if not os.path.exists(path):
        return None
    a = P.api(path, "ref_shim_allocator")
    a.lib.ref_shim_graph_contacts.restype = C.c_uint32
    a.lib.ref_shim_graph_contacts.argtypes = [C.c_void_p]
    a.lib.ref_shim_jobs_create.restype = C.c_void_p
    a.lib.ref_shim_jobs_create.argtypes = [C.c_uint32]
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    cores = min(cores, 16)
    jobs = a.lib.ref_shim_jobs_create(max(cores - 1, 1)) if cores > 1 else None
    w = build(a, args.bodies, args.seed, bank)
    for _ in range(args.warmup):
        w.step(DT, jobs)
    solved = 0
    steps = 0
    t0 = time.perf_counter()
    # A bounded sample: the same steps as the GPU's timed region, stopped
    # once the CPU has worked `cpu_seconds`.
    while steps < args.steps:
        w.step(DT, jobs)
        solved += a.lib.ref_shim_graph_contacts(w.ptr)
        steps += 1
        if time.perf_counter() - t0 > args.cpu_seconds:
            break
    wall = time.perf_counter() - t0
    return {"wall_s": wall, "solved": solved, "steps": steps, "cores": cores,
            "hash": "%016x" % w.hash()}


def copy_calibration():
    import torch
    n = 256 << 20
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-steps", type=int, default=30,
                    help="steps after the timed region with per-launch kernel timing")
    ap.add_argument("--progress", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned warm start's training and runs")
    ap.add_argument("--no-field", action="store_true",
                    help="skip the 100 000-body field")
    ap.add_argument("--field-steps", type=int, default=150)
    ap.add_argument("--field-warmup", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 1 024-world batch")
    ap.add_argument("--batch-worlds", type=int, default=1024)
    ap.add_argument("--batch-seed", type=int, default=4)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-warmup", type=int, default=40)
    ap.add_argument("--batch-roofline-steps", type=int, default=5)
    ap.add_argument("--batch-fixed", action="store_true",
                    help="no autotuner: colour-pass blocks of 256, body blocks of 64, stream "
                         "launches, one group (one stream), as the profiles are taken")
    ap.add_argument("--train-seeds", type=int, nargs="+", default=[100, 101])
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--train-log", type=int, default=8_000_000,
                    help="contact-log rows a training world holds between flushes")
    ap.add_argument("--train-capacity", type=int, default=40_000_000,
                    help="rows the trainer holds")
    ap.add_argument("--phases", action="store_true",
                    help="also report host and device time per step phase (reads stats "
                         "each step, which adds a contact readback to the timed region)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only; the world step has no multi-GPU path "
              "(batched worlds would shard by cost across GPUs, one large world by slabs)",
              file=sys.stderr)
        sys.exit(2)
    bank = S.load_hull_bank(os.path.join(ROOT, "tests", "golden", "hulls_12v.bin"))
    g = run_gpu(args, bank)
    value = g["solved"] * SUBSTEPS / g["wall_s"]
    # Roofline of the dominant kernel, the colour pass: algorithmic bytes of
    # the launches in the roofline window over their event-timed durations.
    launches = max(g["kernel_launches"], 1)
    k_cs = g["kernel_solved"] * SUBSTEPS
    bytes_per_launch = k_cs * g["kernel_bytes"] / launches
    avg_launch_s = g["kernel_ms"] * 1e-3 / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", TRAFFIC_FILE)
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        traffic = t["hbm_bytes_per_contact_substep"] * k_cs / launches
    learned = None if args.no_learned else run_learned(args, bank)
    field = None if args.no_field else run_field(args, bank)
    batched = None if args.no_batch else run_batch(args, bank)
    if batched is not None and "learned" in batched:
        # The batched learned start against the plain one-world solver above.
        batched["learned"]["vs_plain_single_world"] = batched["learned"]["value"] / value
    cpu = None if args.no_cpu else run_cpu(args, bank)
    line = {
        "metric": "contact solves/sec (contacts x substeps/s) plain vs learned warm start, "
                  "equal residual",
        "value": value,
        "unit": "contact-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "bin drop: %d mixed bodies (spheres, capsules, boxes, 12-vertex "
                               "hulls) into a 40x40x20 m bin, dt 1/60, 4 substeps, plain "
                               "solver" % args.bodies,
                   "bodies": args.bodies, "substeps": SUBSTEPS, "seed": args.seed,
                   "sleep": False, "continuous": False},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_color_pass (one colour of one constraint pass), HIP events "
                               "around each launch over %d steps after the timed region"
                               % args.roofline_steps,
                     "launches": g["kernel_launches"],
                     "avg_launch_us": avg_launch_s * 1e6,
                     "bytes_per_launch": bytes_per_launch,
                     "bytes_per_contact_substep": g["kernel_bytes"]},
        "cpu_baseline": None if cpu is None else {
            "value": cpu["solved"] * SUBSTEPS / cpu["wall_s"],
            "unit": "contact-substeps/s",
            "cores": cpu["cores"],
            "kind": "reference",
            "sample": "the reference's step (oracle/_ref, job system with %d threads) on the "
                      "same scene, steps %d-%d of the timed window (%.1f s)" % (
                          cpu["cores"], args.warmup + 1, args.warmup + cpu["steps"], cpu["wall_s"]),
        },
        "learned_warm_start": learned,
        "field_100k": field,
        "batch_1024": batched,
        "gpu_hash": g["hash"],
        "phases_ms_per_step": g["phases_ms"] if args.phases else None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
