# SPDX-License-Identifier: Apache-2.0
"""Contact solves per second of the world step on one MI355X.

Headline workload: 1 024 independent worlds stepped as one batch
(nph_amd_batch_step), body counts floor(16 * 1250^u) with u uniform (16 to
20 000 bodies, log-uniform; seed 4), each world a box pyramid, a mixed bin
drop (spheres, capsules, boxes, 12-vertex hulls) or ragdoll piles; dt 1/60,
4 substeps, FP32 with f64 body centres.  A step is one batch step; its
contact solves are the constraint-graph contacts the solver took in every
world times the substeps (solver.c:810-843).  The timed region is whole
batch steps (pairs, narrow phase, bookkeeping, solver, finalize), the
post-solve residual measured in it, as it is for the learned start.

The same line carries: the learned warm start on the same batch and steps
(a network trained on the device from plain bin drops of other seeds), with
the residual, largest penetration and mechanical energy of both starts; the
roofline of the two heaviest kernels of the batch, the convex narrow phase
(k_b_narrow_lanes then k_b_narrow_teams) and the colour pass
(k_b_color_pass), from dispatch-stamped HIP
events on the stream they run on, in a window of steps right after the
timed region; the reference's own CPU step (oracle/_ref) on a sample of the
batch's worlds, on every host thread the run may use and on one thread; and
two one-world workloads beside it: the 10 000-body bin drop and the
100 000-body field (mesh patches, ragdolls, mixed bodies), each with both
starts and the reference's CPU step.
"""
import argparse
import ctypes as C
import json
import os
import platform
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
# Host threads one GPU's share of the box allows (the CPU baseline's pool).
HOST_THREADS_PER_GPU = 16
# Algorithmic bytes per contact-substep of the three constraint passes
# (warm start, solve, relax), FP32 structure-of-arrays:
# B = 3 (C + P (K + 2 M)) + E (3 * 2 V + 2 X), C = 68 contact constants,
# K = 116 point constants, M = 16 point accumulators (read and written),
# V = 24 (v, w) read and written per pass, X = 28 (dp, dq) read by solve and
# relax; P points, E dynamic endpoints of the contact.
C_BYTES, K_BYTES, M_BYTES, V_BYTES, X_BYTES = 68, 116, 16, 24, 28
# A joint in a colour, per pass: its configuration (112) and step state
# (228) read, its eight accumulated impulses (56) written; its dynamic
# bodies as a contact's (V, X above).
J_BYTES = 112 + 228 + 56
# Algorithmic bytes per pair of the narrow phase: the contact record read
# and written (280 + 280), both shapes' geometry and relative pose (2 x 48),
# both bodies' f64 origin and rotation (2 x 40), the sorted entry (16) and
# the change code written (4).
COLLIDE_PAIR_BYTES = 280 + 280 + 2 * 48 + 2 * 40 + 16 + 4
# Static bodies each kind of batched world makes first (ground; bin floor
# and walls; ground).
BATCH_STATIC = {0: 1, 1: 5, 2: 1}


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0))
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "cpus_allowed": allowed,
            "threads": min(allowed, HOST_THREADS_PER_GPU)}


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
    # Colours only: the overflow (colour 24) runs in its own kernel.
    sel = (arr["color"] < 24) & (arr["point_count"] > 0)
    p = arr["point_count"][sel].astype(np.int64)
    e = (arr["body_a"][sel] >= static_slots).astype(np.int64) + (
        arr["body_b"][sel] >= static_slots).astype(np.int64)
    b = 3 * (C_BYTES + p * (K_BYTES + 2 * M_BYTES)) + e * (3 * 2 * V_BYTES + 2 * X_BYTES)
    return float(b.sum()), int(sel.sum())


def bind_extras(a):
    L = a.lib
    L.nph_amd_solved_contacts.restype = C.c_uint32
    L.nph_amd_solved_contacts.argtypes = [C.c_void_p]
    L.nph_amd_world_energy.restype = None
    L.nph_amd_world_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.nph_amd_time_kernels.argtypes = [C.c_void_p, C.c_bool]
    L.nph_amd_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.nph_amd_batch_collide_times.restype = None
    L.nph_amd_batch_collide_times.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    return a


def energy(worlds):
    """Mechanical energy (J) of the dynamic bodies of the worlds, summed:
    linear and angular kinetic, potential (nph_amd_world_energy)."""
    out = np.zeros(3)
    e = (C.c_double * 3)()
    for w in worlds:
        w.api.lib.nph_amd_world_energy(w.ptr, e)
        out += (e[0], e[1], e[2])
    return {"kinetic": float(out[0] + out[1]), "potential": float(out[2]),
            "total": float(out.sum())}


def drift(e0, e1):
    return {"start": e0["total"], "end": e1["total"], "change": e1["total"] - e0["total"],
            "relative_change": (e1["total"] - e0["total"]) / abs(e0["total"])
            if e0["total"] else None}


def roofline(bound, achieved, peak, unit, **extra):
    out = {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
           "frac": achieved / peak if peak else 0.0, "traffic": None}
    out.update(extra)
    return out


# ---------------------------------------------------------------------------
# The learned warm start's network
# ---------------------------------------------------------------------------
def train_network(args, bank):
    """Logs plain bin drops of other seeds on the device, trains the network
    there (the library's forward, gradient and Adam kernels) and times its
    forward kernel alone against the MFMA peak."""
    from nepheline_amd import learned as L
    a = P.product_api()
    t0 = time.perf_counter()
    tr = L.Trainer(a, capacity=args.train_capacity, seed=1)
    logged = L.log_bin_drops(a, tr, bank, args.train_seeds, args.bodies, args.train_steps,
                             min_flag=1.0, log_capacity=args.train_log, progress=log)
    tr.prepare()
    loss0 = tr.loss()
    for e in range(args.train_epochs):
        tr.epoch(8192, 2e-3 if e < args.train_epochs - 2 else 6e-4)
    loss1 = tr.loss()
    rows = tr.rows
    ws, params = tr.warm_start(mode=L.NEW_POINTS, gate=True)
    tr.close()
    train_s = time.perf_counter() - t0
    log(f"network trained on {rows} rows in {train_s:.1f} s (loss {loss0:.3f} -> {loss1:.3f})")
    # The forward kernel alone on 4M rows.
    n = 1 << 22
    x = L.DeviceArray(a, shape=(n, L.INPUTS), dtype=np.uint16)
    x.upload(np.zeros((n, L.INPUTS), np.uint16))
    net = L.Network(a, params)
    y = L.DeviceArray(a, shape=(n, L.OUTPUTS), dtype=np.float32)
    f = net.f
    reps = 200
    head = np.empty(4, np.float32)

    def wait():  # a 16-byte copy after the null stream's launches
        f.device_copy(head.ctypes.data, y.ptr, head.nbytes)

    for _ in range(3):
        f.mlp_forward(x.ptr, n, net.params.ptr, y.ptr, None)
    wait()
    t0 = time.perf_counter()
    for _ in range(reps):
        f.mlp_forward(x.ptr, n, net.params.ptr, y.ptr, None)
    wait()
    fwd_s = (time.perf_counter() - t0) / reps
    flops = MLP_FLOPS_PER_ROW * n
    info = {"seeds": list(args.train_seeds), "bodies": args.bodies,
            "steps_per_seed": args.train_steps, "rows_logged": logged, "rows_trained": rows,
            "epochs": args.train_epochs, "loss_before": loss0, "loss_after": loss1,
            "seconds": train_s,
            "forward_roofline": roofline("mfma", flops / fwd_s / 1e12, BF16_PEAK_TFLOPS,
                                         "TFLOP/s", rows=n, flop_per_row=MLP_FLOPS_PER_ROW,
                                         us_per_launch=fwd_s * 1e6)}
    return (ws, params), info


# ---------------------------------------------------------------------------
# The batch (headline)
# ---------------------------------------------------------------------------
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "k_b_color_pass_batch1024_traffic.json")
# HBM read requests by size (the L2's misses to the memory fabric) and the
# written kilobytes, per dispatch: two rocprofv3 --pmc passes.
READ_COUNTERS = {"TCC_EA0_RDREQ_32B_sum": 32, "TCC_EA0_RDREQ_64B_sum": 64,
                 "TCC_EA0_RDREQ_128B_sum": 128}


def measured_traffic(key):
    """HBM bytes per launch of the batched colour pass from the committed
    counter summary, when it was taken on this workload; else None."""
    try:
        with open(TRAFFIC_JSON) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("kernel") != "k_b_color_pass" or d.get("workload_key") != key:
        return None
    return d


def _profile_rows(directory, suffix):
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True)):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    if not rows:
        raise SystemExit("no %s under %s" % (suffix, directory))
    return rows


def _timed_window(rows, groups=2, warmup=5, steps=20, roofline_steps=10):
    """Dispatch ids of the plain batch's colour passes in its roofline
    window: the steps right after its warm-up and timed steps (each step
    ends with one k_b_finalize a world group), up to the end of the window
    and before the learned batch's first k_b_learned_features."""
    ids = sorted({int(r["Dispatch_Id"]): r["Kernel_Name"] for r in rows}.items())
    first = groups * (warmup + steps)
    last = first + groups * roofline_steps
    finals, window = 0, []
    for did, name in ids:
        if name.startswith("k_b_learned_features"):
            break
        if name.startswith("k_b_finalize"):
            finals += 1
        elif name.startswith("k_b_color_pass") and first <= finals < last:
            window.append(did)
    return set(window)


def summarize_profiles(trace_dir, read_dir, write_dir, copy_read_dir, copy_write_dir, key):
    """The batched colour pass's HBM traffic and duration over the plain
    batch's timed steps, from rocprofv3 runs of the bench's own command: a
    kernel trace, and two counter passes (read requests by size; WRITE_SIZE)
    restricted to k_b_color_pass, k_b_finalize and k_b_learned_features.
    The same counters over a 256 MiB device copy (--copy-calibration) show
    what they read for a known byte count."""
    trace = [r for r in _profile_rows(trace_dir, "kernel_trace.csv")
             if r["Kernel_Name"].startswith(("k_b_color_pass", "k_b_finalize",
                                             "k_b_learned_features"))]
    win = _timed_window(trace, warmup=key["warmup"], steps=key["steps"],
                        roofline_steps=key["roofline_steps"])
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in trace
            if int(r["Dispatch_Id"]) in win]

    def per_dispatch(directory, names):
        rows = _profile_rows(directory, "counter_collection.csv")
        w = _timed_window(rows, warmup=key["warmup"], steps=key["steps"],
                          roofline_steps=key["roofline_steps"])
        acc = {}
        for r in rows:
            if int(r["Dispatch_Id"]) in w and r["Counter_Name"] in names:
                acc.setdefault(int(r["Dispatch_Id"]), {})[r["Counter_Name"]] = float(
                    r["Counter_Value"])
        return acc

    reads = per_dispatch(read_dir, READ_COUNTERS)
    writes = per_dispatch(write_dir, {"WRITE_SIZE"})
    read_b = [sum(v[c] * READ_COUNTERS[c] for c in READ_COUNTERS) for v in reads.values()]
    write_b = [v["WRITE_SIZE"] * 1024.0 for v in writes.values()]

    def copy_bytes(directory, fn):
        rows = _profile_rows(directory, "counter_collection.csv")
        acc = {}
        for r in rows:
            if "copy" in r["Kernel_Name"].lower():
                acc.setdefault(int(r["Dispatch_Id"]), {})[r["Counter_Name"]] = float(
                    r["Counter_Value"])
        return [fn(v) for v in acc.values()]

    copy_read = copy_bytes(copy_read_dir,
                           lambda v: sum(v.get(c, 0.0) * READ_COUNTERS[c] for c in READ_COUNTERS))
    copy_write = copy_bytes(copy_write_dir, lambda v: v.get("WRITE_SIZE", 0.0) * 1024.0)
    out = {
        "kernel": "k_b_color_pass",
        "workload": "the bench's 1 024-world batch (seed 4), plain start: its roofline window",
        "workload_key": key,
        "window": "k_b_color_pass dispatches of the plain batch's roofline window: the "
                  "steps right after its warm-up and timed steps (one k_b_finalize a world "
                  "group a step), before the learned batch's first k_b_learned_features",
        "dispatches": {"trace": len(durs), "read_counters": len(read_b),
                       "write_counters": len(write_b)},
        "avg_duration_us": float(np.mean(durs)) / 1e3 if durs else None,
        "read_bytes_per_dispatch": float(np.mean(read_b)) if read_b else None,
        "write_bytes_per_dispatch": float(np.mean(write_b)) if write_b else None,
        "read_counters": "32 TCC_EA0_RDREQ_32B + 64 TCC_EA0_RDREQ_64B + 128 TCC_EA0_RDREQ_128B",
        "write_counter": "WRITE_SIZE (KiB)",
        "copy_check": {"copy_bytes": 256 << 20, "read_bytes": copy_read,
                       "write_bytes": copy_write},
    }
    out["hbm_bytes_per_launch"] = (out["read_bytes_per_dispatch"] or 0.0) + (
        out["write_bytes_per_dispatch"] or 0.0)
    return out


def batch_rooflines(a, batch, plan, solved, window, key=None, joints=(0, 0)):
    """The colour pass's and the convex narrow phase's rooflines from the
    dispatch-stamped HIP events around every launch of `window` (kernel
    timing was on for its steps, which solved `solved` graph contacts and
    `joints` = (joints in colours, their dynamic ends), summed over steps)."""
    kt = (C.c_double * 2)()
    a.batch_kernel_times(batch.ptr, kt)
    ct = (C.c_double * 3)()
    a.lib.nph_amd_batch_collide_times(batch.ptr, ct)
    a.batch_time_kernels(batch.ptr, False)
    tot_b, tot_n = 0.0, 0
    for (kind, _, _), w in zip(plan, batch.worlds):
        b, n = graph_bytes(w, BATCH_STATIC[kind])
        tot_b += b
        tot_n += n
    mean_bytes = tot_b / tot_n if tot_n else 0.0
    launches = max(int(kt[1]), 1)
    joint_bytes = SUBSTEPS * (joints[0] * 3 * J_BYTES + joints[1] * (3 * 2 * V_BYTES + 2 * X_BYTES))
    bytes_per_launch = (solved * SUBSTEPS * mean_bytes + joint_bytes) / launches
    avg_s = kt[0] * 1e-3 / launches
    color = roofline("hbm", bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0, HBM_PEAK_GBS,
                     "GB/s", kernel="k_b_color_pass",
                     window=window, launches=int(kt[1]), avg_launch_us=avg_s * 1e6, device_ms=kt[0],
                     bytes_per_launch=bytes_per_launch, bytes_per_contact_substep=mean_bytes,
                     contact_substeps=solved * SUBSTEPS, joint_substeps=joints[0] * SUBSTEPS,
                     joint_bytes=joint_bytes)
    counters = measured_traffic(key)
    if counters is not None:
        color["traffic"] = counters["hbm_bytes_per_launch"]
        color["traffic_source"] = ("profiles/k_b_color_pass_batch1024_traffic.json: "
                                   "rocprofv3 counters of this command, %s" % counters["window"])
        color["traffic_over_algorithmic"] = counters["hbm_bytes_per_launch"] / bytes_per_launch
    claunches = max(int(ct[1]), 1)
    cavg_s = ct[0] * 1e-3 / claunches
    cbytes = ct[2] * COLLIDE_PAIR_BYTES / claunches
    collide = roofline("hbm", cbytes / cavg_s / 1e9 if cavg_s > 0 else 0.0, HBM_PEAK_GBS,
                       "GB/s", kernel="k_b_narrow_lanes + k_b_narrow_teams (the lane "
                       "pass's start to the team pass's end)",
                       window=window, launches=int(ct[1]), avg_launch_us=cavg_s * 1e6, device_ms=ct[0],
                       pairs_per_launch=ct[2] / claunches, bytes_per_pair=COLLIDE_PAIR_BYTES,
                       bytes_per_launch=cbytes)
    return color, collide


def run_batch(args, bank, plan, network=None, rooflines=False):
    """One batch of the plan: built, warmed up (the autotuner inside the
    warm-up when it is long enough for two rounds), then the timed steps
    with residual measurement on, the same for both starts."""
    from nepheline_amd import learned as L
    a = bind_extras(P.product_api())
    t0 = time.perf_counter()
    batch = P.Batch(a, len(plan), S.DESC)
    for (kind, n, seed), w in zip(plan, batch.worlds):
        S.fill_batch_world(w, kind, n, seed, bank)
    setup = time.perf_counter() - t0
    bodies = sum(len(w.bodies) for w in batch.worlds)
    log(f"batch: {len(plan)} worlds, {bodies} bodies built in {setup:.1f} s"
        + (" (learned start)" if network else " (plain start)"))
    if network is not None:
        L.batch_set_warm_start(batch, network[0], network[1])
    tune_rounds = 2
    autotune = args.warmup >= 32 * tune_rounds + 8
    plain_warm = args.warmup - 32 * tune_rounds if autotune else args.warmup
    for _ in range(plain_warm):
        batch.step(DT)
    if autotune:
        tune = batch.autotune(DT, rounds=tune_rounds)  # 32 candidates a round, a step each
        how = "autotuned (%d rounds of 32 candidates inside the warm-up)" % tune_rounds
    else:
        batch.set_tuning(256, 64, 0, args.batch_groups)
        tune = batch.tuning()
        how = "fixed (the warm-up is too short for the autotuner)"
    L.batch_measure_residual(batch, True)
    L.batch_warm_stats(batch, reset=True)
    # The timed steps run uninstrumented; the rooflines' kernel timing
    # (dispatch-stamped events on each launch's own stream) runs over a
    # window of steps right after them.
    e0 = energy(batch.worlds)
    solved = 0
    joints = [0, 0]
    phases = np.zeros(19)
    t0 = time.perf_counter()
    for i in range(args.steps):
        batch.step(DT)
        st = batch.stats()
        solved += st.solved_contacts
        joints[0] += st.colored_joints
        joints[1] += st.colored_joint_ends
        phases += (st.step_ms, st.narrow_ms, st.pairs_ms, st.changes_ms, st.solve_ms, st.post_ms,
                   st.device_solve_ms, st.device_passes_ms, st.device_collide_ms,
                   st.wait_narrow_ms, st.wait_solve_ms, st.device_pairs_ms, st.pairs_trees_ms,
                   st.pairs_search_ms, st.pairs_create_ms, st.rebuild_cpu_ms,
                   st.pairs_gather_ms, st.pairs_upload_ms, st.pairs_upload_bytes / 1e6)
        if (i + 1) % 10 == 0:
            log(f"batch step {i + 1}/{args.steps}")
    wall = time.perf_counter() - t0
    e1 = energy(batch.worlds)
    st = L.batch_warm_stats(batch, reset=True)
    L.batch_measure_residual(batch, False)
    n = st.contacts_plain + st.contacts_learned
    names = ("step", "narrow", "pairs", "host_changes", "solve", "host_post", "dev_solve",
             "dev_passes", "dev_collide", "wait_narrow", "wait_solve", "dev_pairs",
             "pairs_trees", "pairs_search", "pairs_create", "rebuild_cpu",
             "pairs_gather", "pairs_upload", "pairs_upload_mb")
    out = {"value": solved * SUBSTEPS / wall, "ms_per_step": wall * 1e3 / args.steps,
           "contact_substeps": solved * SUBSTEPS, "setup_s": setup, "bodies": bodies,
           "tuning": {"how": how, "pass_block": tune.pass_block, "body_block": tune.body_block,
                      "use_graph": bool(tune.use_graph), "groups": tune.groups},
           "phases_ms_per_step": {k: round(v / max(args.steps, 1), 3)
                                  for k, v in zip(names, phases)},
           "residual_mean": (st.residual_sum_plain + st.residual_sum_learned) / n if n else 0.0,
           "residual_max": max(st.residual_max_plain, st.residual_max_learned),
           "penetration_max": st.penetration_max,
           "learned_fraction": st.contacts_learned / n if n else 0.0, "gated": st.gated,
           "energy": drift(e0, e1)}
    if rooflines:
        batch.set_tuning(tune.pass_block, tune.body_block, 0, tune.groups)
        a.batch_time_kernels(batch.ptr, True)
        solved_window = 0
        joints_window = [0, 0]
        for _ in range(args.roofline_steps):
            batch.step(DT)
            st = batch.stats()
            solved_window += st.solved_contacts
            joints_window[0] += st.colored_joints
            joints_window[1] += st.colored_joint_ends
        window = "the %d steps right after the timed ones (stream launches)" % (
            args.roofline_steps)
        key = {"worlds": len(plan), "seed": args.batch_seed, "steps": args.steps,
               "warmup": args.warmup, "roofline_steps": args.roofline_steps}
        out["rooflines"] = batch_rooflines(a, batch, plan, solved_window, window, key,
                                           joints_window)
        a.batch_time_kernels(batch.ptr, False)
    batch.close()
    return out


def batch_cpu(args, bank, plan, threads, stride):
    """The reference's step on every `stride`-th world of the batch, worlds
    side by side on `threads` host threads (each world's step on one), after
    the same warm-up as the GPU's, over at least the GPU's timed steps and
    at least `cpu_seconds` of stepping (the sample's steps are then the
    GPU's and the ones after them)."""
    from concurrent.futures import ThreadPoolExecutor
    path = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")
    ra = P.api(path, "ref_shim_allocator")
    ra.lib.ref_shim_graph_contacts.restype = C.c_uint32
    ra.lib.ref_shim_graph_contacts.argtypes = [C.c_void_p]
    sample = plan[::stride]
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

    with ThreadPoolExecutor(max_workers=threads) as pool:
        for _ in range(args.warmup):
            list(pool.map(one, order))
        solved, steps = 0, 0
        t0 = time.perf_counter()
        while (steps < args.steps or time.perf_counter() - t0 < args.cpu_seconds) and \
                steps < 100 * max(args.steps, 1):
            solved += sum(pool.map(one, order))
            steps += 1
        cwall = time.perf_counter() - t0
    for w in worlds:
        w.close()
    return {"value": solved * SUBSTEPS / cwall, "unit": "contact-substeps/s", "threads": threads,
            "sample": "the reference's step on every %dth world of the batch (%d worlds, %d "
                      "bodies), worlds side by side on %d thread(s), steps %d-%d (%.1f s)" % (
                          stride, len(worlds), bodies, threads, args.warmup + 1,
                          args.warmup + steps, cwall)}


# ---------------------------------------------------------------------------
# One-world workloads beside the batch
# ---------------------------------------------------------------------------
def world_run(a, make, steps, warmup, network=None, measure=True, kernel_steps=0):
    """One world: warm-up, then timed steps (residual measured), energy
    before and after, and optionally a window of colour-pass event timing."""
    from nepheline_amd import learned as L
    w = make(a)
    if network is not None:
        L.set_warm_start(w, network[0], network[1])
    elif measure:
        L.measure_residual(w, True)
    for _ in range(warmup):
        w.step(DT)
    e0 = energy([w])
    solved = 0
    res_sum, res_n, res_max, pen, learned, gated = 0.0, 0, 0.0, 0.0, 0, 0
    t_wall = 0.0
    phases = np.zeros(4)
    for i in range(steps):
        t0 = time.perf_counter()
        w.step(DT)
        t_wall += time.perf_counter() - t0
        st = w.stats()  # outside the timed region: it reads the contacts back
        phases += (st.pairs_ms, st.collide_ms, st.solve_ms, st.sleep_ms)
        solved += a.lib.nph_amd_solved_contacts(w.ptr)
        if measure or network is not None:
            s = L.warm_stats(w)
            res_sum += s.residual_sum_plain + s.residual_sum_learned
            res_n += s.contacts_plain + s.contacts_learned
            res_max = max(res_max, s.residual_max_plain, s.residual_max_learned)
            pen = max(pen, s.penetration_max)
            learned += s.contacts_learned
            gated += s.gated
    e1 = energy([w])
    out = {"value": solved * SUBSTEPS / t_wall, "ms_per_step": t_wall * 1e3 / steps,
           "steps": steps, "warmup": warmup, "residual_mean": res_sum / max(res_n, 1),
           "residual_max": res_max, "penetration_max": pen, "contacts_from_network": learned,
           "gated": gated, "energy": drift(e0, e1),
           "phases_ms_per_step": {k: round(v / max(steps, 1), 3) for k, v in
                                  zip(("pairs", "collide", "solve", "islands"), phases)}}
    if kernel_steps:
        a.lib.nph_amd_time_kernels(w.ptr, True)
        k_solved = 0
        for _ in range(kernel_steps):
            w.step(DT)
            k_solved += a.lib.nph_amd_solved_contacts(w.ptr)
        kt = (C.c_double * 2)()
        a.lib.nph_amd_kernel_times(w.ptr, kt)
        a.lib.nph_amd_time_kernels(w.ptr, False)
        b, n = graph_bytes(w, 5)
        mean_bytes = b / n if n else 0.0
        launches = max(int(kt[1]), 1)
        per_launch = k_solved * SUBSTEPS * mean_bytes / launches
        avg_s = kt[0] * 1e-3 / launches
        out["roofline"] = roofline("hbm", per_launch / avg_s / 1e9 if avg_s > 0 else 0.0,
                                   HBM_PEAK_GBS, "GB/s", kernel="k_color_pass",
                                   window="%d steps after the timed ones" % kernel_steps,
                                   launches=int(kt[1]), avg_launch_us=avg_s * 1e6,
                                   bytes_per_launch=per_launch,
                                   bytes_per_contact_substep=mean_bytes)
    out["bodies"] = len(w.bodies)
    w.close()
    return out


def world_cpu(make, steps, warmup, threads, seconds):
    """The reference's step of the same world with its job system on
    `threads` threads (deterministic by design: the same bits on any count)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")
    ra = P.api(path, "ref_shim_allocator")
    ra.lib.ref_shim_graph_contacts.restype = C.c_uint32
    ra.lib.ref_shim_graph_contacts.argtypes = [C.c_void_p]
    ra.lib.ref_shim_jobs_create.restype = C.c_void_p
    ra.lib.ref_shim_jobs_create.argtypes = [C.c_uint32]
    jobs = ra.lib.ref_shim_jobs_create(threads - 1) if threads > 1 else None
    w = make(ra)
    for _ in range(warmup):
        w.step(DT, jobs)
    solved, done = 0, 0
    t0 = time.perf_counter()
    while done < steps and time.perf_counter() - t0 < seconds:
        w.step(DT, jobs)
        solved += ra.lib.ref_shim_graph_contacts(w.ptr)
        done += 1
    wall = time.perf_counter() - t0
    w.close()
    return {"value": solved * SUBSTEPS / wall, "unit": "contact-substeps/s", "threads": threads,
            "kind": "reference",
            "sample": "the reference's step (job system, %d thread(s)), steps %d-%d (%.1f s)" % (
                threads, warmup + 1, warmup + done, wall)}


def run_bin(args, bank, network, cpu):
    make = lambda api: S.bin_drop(api, args.bodies, args.seed, bank, half=20.0, height=20.0,
                                  y_lo=5.0)
    out = {"workload": "bin drop: %d mixed bodies (spheres, capsules, boxes, 12-vertex hulls) "
                       "from 5-20 m into a 40 x 40 x 20 m bin, seed %d, dt 1/60, 4 substeps"
                       % (args.bodies, args.seed)}
    out["plain"] = world_run(bind_extras(P.product_api()), make, args.bin_steps, args.bin_warmup,
                             kernel_steps=args.roofline_steps)
    if network is not None:
        lo = world_run(bind_extras(P.product_api()), make, args.bin_steps, args.bin_warmup,
                       network)
        lo["vs_plain"] = lo["value"] / out["plain"]["value"]
        out["learned"] = lo
    if cpu is not None:
        out["cpu_baseline"] = world_cpu(make, args.bin_steps, args.bin_warmup, cpu["threads"],
                                        args.cpu_seconds)
        out["vs_cpu"] = out["plain"]["value"] / out["cpu_baseline"]["value"]
    return out


def run_field(args, bank, network, cpu):
    meshes = S.load_mesh_bank(os.path.join(ROOT, "tests", "golden", "meshes.bin"))
    make = lambda api: S.field(api, meshes, bank, seed=3)
    out = {"workload": "field: 100 201 bodies (200 mesh patches of 64 triangles, 2 000 ragdolls "
                       "of 11 capsules and 10 joints, 78 000 mixed), 150 x 150 m, seed 3, "
                       "dt 1/60, 4 substeps"}
    out["plain"] = world_run(bind_extras(P.product_api()), make, args.field_steps,
                             args.field_warmup)
    if network is not None:
        lo = world_run(bind_extras(P.product_api()), make, args.field_steps, args.field_warmup,
                       network)
        lo["vs_plain"] = lo["value"] / out["plain"]["value"]
        out["learned"] = lo
    if cpu is not None:
        out["cpu_baseline"] = world_cpu(make, args.field_steps, args.field_warmup,
                                        cpu["threads"], args.cpu_seconds)
        out["vs_cpu"] = out["plain"]["value"] / out["cpu_baseline"]["value"]
    return out


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
    ap.add_argument("--steps", type=int, default=30, help="timed batch steps")
    ap.add_argument("--warmup", type=int, default=10, help="batch warm-up steps (the autotuner "
                    "runs inside it when it has at least 72)")
    ap.add_argument("--batch-worlds", type=int, default=1024)
    ap.add_argument("--batch-seed", type=int, default=4)
    ap.add_argument("--batch-groups", type=int, default=2,
                    help="world groups the batch step is pipelined over when not autotuned")
    ap.add_argument("--roofline-steps", type=int, default=10,
                    help="steps right after the timed ones with per-launch kernel timing")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-stride", type=int, default=16,
                    help="the CPU baseline steps every n-th world of the batch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned warm start's training and runs")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the one-world workloads (bin drop, field)")
    ap.add_argument("--bodies", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--bin-steps", type=int, default=600)
    ap.add_argument("--bin-warmup", type=int, default=20)
    ap.add_argument("--field-steps", type=int, default=150)
    ap.add_argument("--field-warmup", type=int, default=20)
    ap.add_argument("--train-seeds", type=int, nargs="+", default=[100, 101, 102, 103])
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--train-log", type=int, default=8_000_000,
                    help="contact-log rows a training world holds between flushes")
    ap.add_argument("--train-capacity", type=int, default=40_000_000,
                    help="rows the trainer holds")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (a known byte count "
                         "for the memory counters) and exit")
    ap.add_argument("--summarize-profiles", nargs=5, metavar=("TRACE", "READS", "WRITES",
                                                              "COPY_READS", "COPY_WRITES"),
                    help="write the colour pass's traffic summary from rocprofv3 output "
                         "directories of this command and of --copy-calibration, and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.summarize_profiles:
        key = {"worlds": args.batch_worlds, "seed": args.batch_seed, "steps": args.steps,
               "warmup": args.warmup, "roofline_steps": args.roofline_steps}
        out = summarize_profiles(*args.summarize_profiles, key)
        with open(TRAFFIC_JSON, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out, indent=1))
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only; the world step has no multi-GPU path "
              "(batched worlds would shard by cost across GPUs, one large world by slabs)",
              file=sys.stderr)
        sys.exit(2)
    bank = S.load_hull_bank(os.path.join(ROOT, "tests", "golden", "hulls_12v.bin"))
    plan = S.batch_plan(args.batch_seed, args.batch_worlds)
    ref_lib = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")
    cpu = None if args.no_cpu or not os.path.exists(ref_lib) else cpu_info()

    network, training = (None, None) if args.no_learned else train_network(args, bank)
    plain = run_batch(args, bank, plan, rooflines=True)
    learned = None
    if network is not None:
        learned = run_batch(args, bank, plan, network)
        learned["vs_plain"] = learned["value"] / plain["value"]
        learned["residual_vs_plain"] = learned["residual_mean"] / plain["residual_mean"] \
            if plain["residual_mean"] else None
        learned["energy_vs_plain"] = (learned["energy"]["end"] - plain["energy"]["end"]) / abs(
            plain["energy"]["end"]) if plain["energy"]["end"] else None
        learned["mode"] = "points that carry no impulse start from the network; residual gate on"
        learned["training"] = training
    color, collide = plain.pop("rooflines")
    dominant = color  # the batch's largest kernel by device time
    cpu_line = None
    if cpu is not None:
        log(f"CPU baseline: {cpu['threads']} threads ({cpu['cpu_model']})")
        many = batch_cpu(args, bank, plan, cpu["threads"], args.cpu_stride)
        one = batch_cpu(args, bank, plan, 1, args.cpu_stride * 4)
        cpu_line = {"value": many["value"], "unit": "contact-substeps/s",
                    "cores": cpu["threads"], "kind": "reference", "sample": many["sample"],
                    "cpu_model": cpu["cpu_model"], "logical_cpus": cpu["logical_cpus"],
                    "cpus_allowed": cpu["cpus_allowed"],
                    "threads_note": "all host threads one GPU's share of the box allows "
                                    "(min(cpus allowed, %d))" % HOST_THREADS_PER_GPU,
                    "one_thread": one}
    extra = {}
    if not args.no_extra:
        extra["bin_10k"] = run_bin(args, bank, network, cpu)
        extra["field_100k"] = run_field(args, bank, network, cpu)
    bodies = plain["bodies"]
    line = {
        "metric": "contact solves/sec (contacts x substeps/s) plain vs learned warm start, "
                  "equal residual",
        "value": plain["value"],
        "unit": "contact-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch: %d worlds of floor(16*1250^u) bodies (seed %d; box "
                               "pyramids, mixed bin drops, ragdoll piles), %d bodies, dt 1/60, "
                               "4 substeps, plain start" % (len(plan), args.batch_seed, bodies),
                   "worlds": len(plan), "bodies": bodies, "substeps": SUBSTEPS,
                   "seed": args.batch_seed, "sleep": False, "continuous": False},
        "roofline": dominant,
        "cpu_baseline": cpu_line,
        "vs_cpu": plain["value"] / cpu_line["value"] if cpu_line else None,
        "plain": plain,
        "learned_warm_start": learned,
        "rooflines": {"narrow_phase": collide, "k_b_color_pass": color},
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
