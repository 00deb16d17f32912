// SPDX-License-Identifier: Apache-2.0
// Many independent worlds stepped together.
//
// Each world's bookkeeping is its own, as when it steps alone: the same
// functions, in the same order, on the same records.  What changes is how
// the phases meet the device.  The one-world step alternates host phases
// and device launches; the batch runs each host phase for every world (the
// worlds in parallel on OpenMP threads: they share nothing), then one
// launch sequence for all of them (device_batch.hip):
//
//   begin        events cleared (every world)
//   narrow 0     the contacts alive at step start, queued   [device]
//   pairs        new pairs of moved proxies (worlds in parallel, overlapping
//                narrow 0 as the one-world step overlaps them)
//   narrow 1     the new contacts; change codes of both halves back
//   changes      begin/end, colouring, awake lists; the solver's input
//   solve        prepare, substeps, restitution, store, finalize, grown
//                proxies                                    [device]
//   post         grown proxies into the trees, hit events
//
// A world's results depend only on its own records and the order of its own
// operations, so each world's step is bit for bit the one it takes alone.

#include "world_impl.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

struct NphAmdBatch {
    NphAllocator *alloc;
    NphPhysicsWorldDesc desc;
    u32 count;
    NphPhysicsWorld **worlds;
    NphDevice **devs;
    NphDeviceBatch *dev;
    NphCollidePart *parts;
    NphSolveInput *inputs;
    u32 *first;  // per world: awake contacts when the step started
    u32 *off_a;  // per world: its first change code of each narrow-phase half
    u32 *off_b;
    u32 *active; // the worlds with awake bodies, in world order
    u64 *order;  // worlds by descending awake count (host phases' schedule):
                 // (~count << 32) | world
    NphResult *results;
    NphAmdBatchStats stats;
};

static void *batch_alloc(NphAllocator *a, u64 bytes)
{
    NphCallsite cs = {__FILE__, (u32)__LINE__};
    void *p = a->vtable->alloc(a, bytes > 0U ? bytes : 1U, 16U, cs);
    if (p != NULL) {
        memset(p, 0, (size_t)(bytes > 0U ? bytes : 1U));
    }
    return p;
}

static void batch_free(NphAllocator *a, void *p, u64 bytes)
{
    if (p != NULL) {
        NphCallsite cs = {__FILE__, (u32)__LINE__};
        a->vtable->free(a, p, bytes > 0U ? bytes : 1U, 16U, cs);
    }
}

void nph_amd_batch_destroy(NphAmdBatch *b)
{
    if (b == NULL) {
        return;
    }
    NphAllocator *a = b->alloc;
    const u64 n = b->count;
    if (b->worlds != NULL) {
        for (u32 k = 0U; k < b->count; ++k) {
            nph_physics_world_destroy(b->worlds[k]);
        }
    }
    nph_devb_destroy(b->dev);
    batch_free(a, b->worlds, n * sizeof *b->worlds);
    batch_free(a, b->devs, n * sizeof *b->devs);
    batch_free(a, b->parts, n * sizeof *b->parts);
    batch_free(a, b->inputs, n * sizeof *b->inputs);
    batch_free(a, b->first, n * sizeof *b->first);
    batch_free(a, b->off_a, n * sizeof *b->off_a);
    batch_free(a, b->off_b, n * sizeof *b->off_b);
    batch_free(a, b->active, n * sizeof *b->active);
    batch_free(a, b->order, n * sizeof *b->order);
    batch_free(a, b->results, n * sizeof *b->results);
    batch_free(a, b, sizeof *b);
}

NphResult nph_amd_batch_create(const NphPhysicsWorldDesc *desc, u32 world_count,
                               NphAmdBatch **out)
{
    if (desc == NULL || out == NULL || desc->alloc == NULL || world_count == 0U ||
        world_count > (1U << 20U)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    if (desc->enable_continuous != 0) {
        return NPH_RESULT_UNSUPPORTED;
    }
    NphAllocator *a = desc->alloc;
    NphAmdBatch *b = (NphAmdBatch *)batch_alloc(a, sizeof *b);
    if (b == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    b->alloc = a;
    b->desc = *desc;
    b->count = world_count;
    const u64 n = world_count;
    b->worlds = (NphPhysicsWorld **)batch_alloc(a, n * sizeof *b->worlds);
    b->devs = (NphDevice **)batch_alloc(a, n * sizeof *b->devs);
    b->parts = (NphCollidePart *)batch_alloc(a, n * sizeof *b->parts);
    b->inputs = (NphSolveInput *)batch_alloc(a, n * sizeof *b->inputs);
    b->first = (u32 *)batch_alloc(a, n * sizeof *b->first);
    b->off_a = (u32 *)batch_alloc(a, n * sizeof *b->off_a);
    b->off_b = (u32 *)batch_alloc(a, n * sizeof *b->off_b);
    b->active = (u32 *)batch_alloc(a, n * sizeof *b->active);
    b->order = (u64 *)batch_alloc(a, n * sizeof *b->order);
    b->results = (NphResult *)batch_alloc(a, n * sizeof *b->results);
    NphResult r = NPH_RESULT_OK;
    if (b->worlds == NULL || b->devs == NULL || b->parts == NULL || b->inputs == NULL ||
        b->first == NULL || b->off_a == NULL || b->off_b == NULL || b->active == NULL ||
        b->order == NULL || b->results == NULL) {
        r = NPH_RESULT_OUT_OF_MEMORY;
    }
    if (r == NPH_RESULT_OK) {
        r = nph_devb_create(&b->dev);
    }
    for (u32 k = 0U; k < world_count && r == NPH_RESULT_OK; ++k) {
        r = world_create(desc, b->dev, &b->worlds[k]);
        if (r == NPH_RESULT_OK) {
            b->devs[k] = b->worlds[k]->dev;
        }
    }
    if (r != NPH_RESULT_OK) {
        nph_amd_batch_destroy(b);
        return r;
    }
    *out = b;
    return NPH_RESULT_OK;
}

u32 nph_amd_batch_world_count(const NphAmdBatch *b) { return b->count; }

NphPhysicsWorld *nph_amd_batch_world(NphAmdBatch *b, u32 index)
{
    return index < b->count ? b->worlds[index] : NULL;
}

NphAmdBatchStats nph_amd_batch_stats(const NphAmdBatch *b) { return b->stats; }

void nph_amd_batch_set_tuning(NphAmdBatch *b, const NphAmdBatchTuning *t)
{
    NphBatchTuning d;
    d.pass_block = t->pass_block;
    d.body_block = t->body_block;
    d.use_graph = t->use_graph;
    d.pad = 0U;
    nph_devb_set_tuning(b->dev, &d);
}

NphAmdBatchTuning nph_amd_batch_tuning(const NphAmdBatch *b)
{
    const NphBatchTuning d = nph_devb_tuning(b->dev);
    NphAmdBatchTuning t;
    t.pass_block = d.pass_block;
    t.body_block = d.body_block;
    t.use_graph = d.use_graph;
    t.pad = 0U;
    return t;
}

void nph_amd_batch_time_kernels(NphAmdBatch *b, bool enable)
{
    nph_devb_time_kernels(b->dev, enable ? 1 : 0);
}

void nph_amd_batch_kernel_times(const NphAmdBatch *b, f64 out[2])
{
    nph_devb_kernel_times(b->dev, out);
}

static int u64_ascending(const void *x, const void *y)
{
    const u64 a = *(const u64 *)x, c = *(const u64 *)y;
    return a < c ? -1 : (a > c ? 1 : 0);
}

// The worlds' first error in world order, or OK.
static NphResult first_error(const NphAmdBatch *b)
{
    for (u32 k = 0U; k < b->count; ++k) {
        if (b->results[k] != NPH_RESULT_OK) {
            return b->results[k];
        }
    }
    return NPH_RESULT_OK;
}

NphResult nph_amd_batch_step(NphAmdBatch *b, f32 dt)
{
    if (!(dt > 0.0F) || !nph_isfinitef(dt)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const u32 n = b->count;
    for (u32 k = 0U; k < n; ++k) {
        if (nph_dev_logging(b->devs[k])) {
            return NPH_RESULT_UNSUPPORTED;
        }
    }
    const f64 t0 = world_now_ms();
    // The host phases take the largest worlds first.
    for (u32 k = 0U; k < n; ++k) {
        b->order[k] = ((u64)(~b->worlds[k]->awake_bodies.len) << 32U) | k;
    }
    qsort(b->order, n, sizeof *b->order, u64_ascending);
    for (u32 k = 0U; k < n; ++k) {
        NphPhysicsWorld *w = b->worlds[k];
        world_step_begin(w, dt);
        NphCollidePart *p = &b->parts[k];
        p->dev = w->dev;
        p->inits = VB(w->inits, NphContactInit);
        p->init_count = w->inits.len;
        p->list = VB(w->awake_contacts, u32);
        p->first = 0U;
        p->count = w->awake_contacts.len;
        b->first[k] = w->awake_contacts.len;
    }
    const u32 *codes[2] = {NULL, NULL};
    NphResult r = nph_devb_collide(b->dev, b->parts, n, dt, 0U, codes);
    for (u32 k = 0U; k < n; ++k) {
        b->worlds[k]->inits.len = 0U; // copied for the device
    }
    f64 t = world_now_ms();
    if (r == NPH_RESULT_OK) {
#pragma omp parallel for schedule(dynamic, 1)
        for (u32 i = 0U; i < n; ++i) {
            const u32 k = (u32)b->order[i];
            b->results[k] = contacts_update_pairs(b->worlds[k]);
        }
        r = first_error(b);
    }
    b->stats.pairs_ms = world_now_ms() - t;
    if (r == NPH_RESULT_OK) {
        for (u32 k = 0U; k < n; ++k) {
            NphPhysicsWorld *w = b->worlds[k];
            NphCollidePart *p = &b->parts[k];
            p->inits = VB(w->inits, NphContactInit);
            p->init_count = w->inits.len;
            p->list = VB(w->awake_contacts, u32);
            p->first = b->first[k];
            p->count = w->awake_contacts.len;
        }
        r = nph_devb_collide(b->dev, b->parts, n, dt, 1U, codes);
        for (u32 k = 0U; k < n; ++k) {
            b->worlds[k]->inits.len = 0U;
        }
    }
    t = world_now_ms();
    if (r == NPH_RESULT_OK) {
        // Each world's codes: its entries of both halves, in world order.
        u64 oa = 0U, ob = 0U;
        for (u32 k = 0U; k < n; ++k) {
            b->off_a[k] = (u32)oa;
            b->off_b[k] = (u32)ob;
            oa += b->first[k];
            ob += b->parts[k].count - b->first[k];
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (u32 i = 0U; i < n; ++i) {
            const u32 k = (u32)b->order[i];
            NphPhysicsWorld *w = b->worlds[k];
            const NphCollidePart *p = &b->parts[k];
            NphResult rk = contacts_apply_codes(w, codes[0] + b->off_a[k], b->first[k],
                                                codes[1] + b->off_b[k], p->count - b->first[k]);
            if (rk == NPH_RESULT_OK) {
                rk = world_solve_input(w, dt, &b->inputs[k]);
            }
            b->results[k] = rk;
        }
        r = first_error(b);
    }
    b->stats.changes_ms = world_now_ms() - t;
    const NphGrown *grown = NULL;
    const u32 *starts = NULL;
    t = world_now_ms();
    u32 active = 0U;
    if (r == NPH_RESULT_OK) {
        // Worlds with nothing awake skip the solve, as the one-world step does.
        u64 solved = 0U, awake = 0U;
        for (u32 k = 0U; k < n; ++k) {
            if (b->inputs[k].awake_count > 0U) {
                b->devs[active] = b->worlds[k]->dev;
                if (active != k) {
                    b->inputs[active] = b->inputs[k];
                }
                b->active[active] = k;
                active += 1U;
                solved += b->worlds[k]->solved_contacts;
                awake += b->worlds[k]->awake_bodies.len;
            } else {
                b->worlds[k]->solved_contacts = 0U;
            }
        }
        b->stats.solved_contacts = solved;
        b->stats.awake_bodies = awake;
        r = nph_devb_solve(b->dev, b->devs, b->inputs, active, &grown, &starts);
    }
    b->stats.solve_ms = world_now_ms() - t;
    t = world_now_ms();
    if (r == NPH_RESULT_OK) {
#pragma omp parallel for schedule(dynamic, 1)
        for (u32 e = 0U; e < active; ++e) {
            const u32 k = b->active[e];
            b->results[k] = world_solve_post(b->worlds[k], grown + starts[e],
                                             starts[e + 1U] - starts[e]);
        }
        r = first_error(b);
    }
    b->stats.post_ms = world_now_ms() - t;
    for (u32 k = 0U; k < n; ++k) {
        b->devs[k] = b->worlds[k]->dev;
        b->results[k] = NPH_RESULT_OK;
        NphPhysicsWorld *w = b->worlds[k];
        w->stepping = 0;
        w->stats.steps += 1U;
    }
    f64 dev[3];
    nph_devb_times(b->dev, dev);
    b->stats.device_solve_ms = dev[0];
    b->stats.device_passes_ms = dev[1];
    b->stats.device_collide_ms = dev[2];
    b->stats.step_ms = world_now_ms() - t0;
    return r;
}

NphResult nph_amd_batch_autotune(NphAmdBatch *b, f32 dt, u32 rounds, NphAmdBatchTuning *chosen)
{
    static const u32 pass_blocks[3] = {64U, 128U, 256U};
    static const u32 body_blocks[2] = {64U, 256U};
    enum { CANDIDATES = 12 };
    NphAmdBatchTuning cand[CANDIDATES];
    f64 total[CANDIDATES];
    u32 m = 0U;
    for (u32 g = 0U; g < 2U; ++g) {
        for (u32 p = 0U; p < 3U; ++p) {
            for (u32 q = 0U; q < 2U; ++q) {
                cand[m].pass_block = pass_blocks[p];
                cand[m].body_block = body_blocks[q];
                cand[m].use_graph = g;
                cand[m].pad = 0U;
                total[m] = 0.0;
                m += 1U;
            }
        }
    }
    const NphAmdBatchTuning before = nph_amd_batch_tuning(b);
    for (u32 round = 0U; round < (rounds > 0U ? rounds : 1U); ++round) {
        for (u32 c = 0U; c < m; ++c) {
            nph_amd_batch_set_tuning(b, &cand[c]);
            const NphResult r = nph_amd_batch_step(b, dt);
            if (r != NPH_RESULT_OK) {
                nph_amd_batch_set_tuning(b, &before);
                return r;
            }
            total[c] += b->stats.solve_ms;
        }
    }
    u32 best = 0U;
    for (u32 c = 1U; c < m; ++c) {
        if (total[c] < total[best]) {
            best = c;
        }
    }
    nph_amd_batch_set_tuning(b, &cand[best]);
    if (chosen != NULL) {
        *chosen = cand[best];
    }
    return NPH_RESULT_OK;
}

NphResult nph_amd_batch_set_warm_start(NphAmdBatch *b, const NphAmdWarmStart *ws)
{
    for (u32 k = 0U; k < b->count; ++k) {
        const NphResult r = nph_amd_world_set_warm_start(b->worlds[k], ws);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    return NPH_RESULT_OK;
}

void nph_amd_batch_measure_residual(NphAmdBatch *b, bool enable)
{
    for (u32 k = 0U; k < b->count; ++k) {
        nph_amd_world_measure_residual(b->worlds[k], enable);
    }
}

NphResult nph_amd_batch_warm_stats(NphAmdBatch *b, NphAmdBatchWarmStats *out, bool reset)
{
    f64 v[8];
    u32 steps = 0U;
    const NphResult r = nph_devb_warm_stats(b->dev, v, &steps, reset ? 1 : 0);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    out->residual_sum_plain = v[0];
    out->residual_sum_learned = v[1];
    out->contacts_plain = v[2];
    out->contacts_learned = v[3];
    out->residual_max_plain = v[4];
    out->residual_max_learned = v[5];
    out->penetration_max = v[6];
    out->gated = v[7];
    out->steps = steps;
    out->pad = 0U;
    return NPH_RESULT_OK;
}
