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
// The worlds are split into groups (up to four, balanced by awake bodies
// every step), each on its own device stream, and the step is pipelined over
// them: while the device solves group 0, the host finds group 1's new
// pairs, and so on; while it solves the last group, the host finishes the
// others (grown proxies, events).
//
// A world's results depend only on its own records and the order of its own
// operations, so each world's step is bit for bit the one it takes alone.

#include "world_impl.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define NPH_BATCH_GROUPS 4U

// Per-world arrays are indexed by position: group 0's worlds first, then
// group 1's and so on, each group's by descending awake bodies.
struct NphAmdBatch {
    NphAllocator *alloc;
    NphPhysicsWorldDesc desc;
    u32 count;
    u32 streams;                      // device batches made: min(4, count)
    u32 groups;                       // groups this step
    u32 want_groups;                  // the tuning's
    u32 start[NPH_BATCH_GROUPS + 1U]; // each group's first position
    NphPhysicsWorld **worlds;          // by world index
    u32 *world_at;                     // position -> world index
    NphDevice **devs;                  // by position
    NphDeviceBatch *dev[NPH_BATCH_GROUPS];
    NphCollidePart *parts;
    NphSolveInput *inputs;
    u32 *first;  // awake contacts when the step started
    u32 *off_a;  // the position's first change code of each narrow-phase half
    u32 *off_b;
    u32 *active; // per group, from its start: positions with awake bodies
    u32 active_count[NPH_BATCH_GROUPS];
    u32 begun[NPH_BATCH_GROUPS]; // the group's solve was queued this step
    u64 *order;  // worlds by descending awake count: (~count << 32) | world
    NphResult *results;
    NphPairSearch *pairs; // one search a group, over the group's worlds that moved
    NphPairInput *pin;    // by position
    u32 *pin_pos;         // the search's worlds: their positions
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
    nph_pairs_destroy(b->pairs);
    for (u32 g = 0U; g < NPH_BATCH_GROUPS; ++g) {
        nph_devb_destroy(b->dev[g]);
    }
    batch_free(a, b->pin, n * sizeof *b->pin);
    batch_free(a, b->pin_pos, n * sizeof *b->pin_pos);
    batch_free(a, b->worlds, n * sizeof *b->worlds);
    batch_free(a, b->world_at, n * sizeof *b->world_at);
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
    b->world_at = (u32 *)batch_alloc(a, n * sizeof *b->world_at);
    b->devs = (NphDevice **)batch_alloc(a, n * sizeof *b->devs);
    b->parts = (NphCollidePart *)batch_alloc(a, n * sizeof *b->parts);
    b->inputs = (NphSolveInput *)batch_alloc(a, n * sizeof *b->inputs);
    b->first = (u32 *)batch_alloc(a, n * sizeof *b->first);
    b->off_a = (u32 *)batch_alloc(a, n * sizeof *b->off_a);
    b->off_b = (u32 *)batch_alloc(a, n * sizeof *b->off_b);
    b->active = (u32 *)batch_alloc(a, n * sizeof *b->active);
    b->order = (u64 *)batch_alloc(a, n * sizeof *b->order);
    b->results = (NphResult *)batch_alloc(a, n * sizeof *b->results);
    b->pin = (NphPairInput *)batch_alloc(a, n * sizeof *b->pin);
    b->pin_pos = (u32 *)batch_alloc(a, n * sizeof *b->pin_pos);
    NphResult r = NPH_RESULT_OK;
    if (b->worlds == NULL || b->world_at == NULL || b->devs == NULL || b->parts == NULL ||
        b->inputs == NULL ||
        b->first == NULL || b->off_a == NULL || b->off_b == NULL || b->active == NULL ||
        b->order == NULL || b->results == NULL || b->pin == NULL || b->pin_pos == NULL) {
        r = NPH_RESULT_OUT_OF_MEMORY;
    }
    b->streams = world_count < NPH_BATCH_GROUPS ? world_count : NPH_BATCH_GROUPS;
    b->want_groups = 2U;
    for (u32 g = 0U; g < b->streams && r == NPH_RESULT_OK; ++g) {
        r = nph_devb_create(&b->dev[g]);
    }
    for (u32 k = 0U; k < world_count && r == NPH_RESULT_OK; ++k) {
        r = world_create(desc, b->dev[0], &b->worlds[k]);
        if (r == NPH_RESULT_OK) {
            b->world_at[k] = k;
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
    for (u32 g = 0U; g < b->streams; ++g) {
        nph_devb_set_tuning(b->dev[g], &d);
    }
    b->want_groups = t->groups < 1U ? 1U : (t->groups > NPH_BATCH_GROUPS ? NPH_BATCH_GROUPS
                                                                          : t->groups);
}

NphAmdBatchTuning nph_amd_batch_tuning(const NphAmdBatch *b)
{
    const NphBatchTuning d = nph_devb_tuning(b->dev[0]);
    NphAmdBatchTuning t;
    t.pass_block = d.pass_block;
    t.body_block = d.body_block;
    t.use_graph = d.use_graph;
    t.groups = b->want_groups;
    return t;
}

void nph_amd_batch_time_kernels(NphAmdBatch *b, bool enable)
{
    for (u32 g = 0U; g < b->streams; ++g) {
        nph_devb_time_kernels(b->dev[g], enable ? 1 : 0);
    }
}

void nph_amd_batch_kernel_times(const NphAmdBatch *b, f64 out[2])
{
    out[0] = out[1] = 0.0;
    for (u32 g = 0U; g < b->streams; ++g) {
        f64 t[5];
        nph_devb_kernel_times(b->dev[g], t);
        out[0] += t[0];
        out[1] += t[1];
    }
}

void nph_amd_batch_collide_times(const NphAmdBatch *b, f64 out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    for (u32 g = 0U; g < b->streams; ++g) {
        f64 t[5];
        nph_devb_kernel_times(b->dev[g], t);
        out[0] += t[2];
        out[1] += t[3];
        out[2] += t[4];
    }
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

// The narrow phase's first half for group g: the contacts alive when the
// step starts, queued.
static NphResult narrow_begin(NphAmdBatch *b, u32 g, f32 dt)
{
    const u32 s0 = b->start[g], m = b->start[g + 1U] - s0;
    for (u32 i = s0; i < s0 + m; ++i) {
        NphPhysicsWorld *w = b->worlds[b->world_at[i]];
        NphCollidePart *p = &b->parts[i];
        p->dev = w->dev;
        p->inits = VB(w->inits, NphContactInit);
        p->init_count = w->inits.len;
        p->list = VB(w->awake_contacts, u32);
        p->first = 0U;
        p->count = w->awake_contacts.len;
        b->first[i] = w->awake_contacts.len;
    }
    const u32 *codes[2] = {NULL, NULL};
    const NphResult r = nph_devb_collide(b->dev[g], b->parts + s0, m, dt, 0U, codes);
    for (u32 i = s0; i < s0 + m; ++i) {
        b->worlds[b->world_at[i]]->inits.len = 0U; // copied for the device
    }
    return r;
}

// Group g from its new pairs to its queued solve.
static NphResult group_to_solve(NphAmdBatch *b, u32 g, f32 dt)
{
    const u32 s0 = b->start[g], m = b->start[g + 1U] - s0;
    f64 t = world_now_ms();
    // New pairs: each world's trees and change lists (worlds in parallel),
    // one device search over the worlds whose move buffers hold proxies,
    // then each world's new contacts.
#pragma omp parallel for schedule(dynamic, 1)
    for (u32 i = 0U; i < m; ++i) {
        const u32 pos = s0 + i;
        NphPhysicsWorld *w = b->worlds[b->world_at[pos]];
        b->results[pos] = w->move_buffer.len == 0U ? contacts_update_pairs(w)
                                                   : pairs_prepare(w, &b->pin[pos]);
    }
    NphResult r = first_error(b);
    f64 t1 = world_now_ms();
    b->stats.pairs_trees_ms += t1 - t;
    for (u32 i = s0; i < s0 + m; ++i) {
        b->stats.rebuild_cpu_ms += b->worlds[b->world_at[i]]->rebuild_ms;
    }
    u32 searched = 0U;
    for (u32 i = s0; i < s0 + m && r == NPH_RESULT_OK; ++i) {
        if (b->worlds[b->world_at[i]]->move_buffer.len > 0U) {
            b->pin[searched] = b->pin[i]; // packed to the front (searched <= i)
            b->pin_pos[searched++] = i;
        }
    }
    const u32 *pairs = NULL;
    const u32 *starts = NULL;
    if (r == NPH_RESULT_OK && searched > 0U) {
        if (b->pairs == NULL) {
            r = nph_pairs_create(&b->pairs);
        }
        if (r == NPH_RESULT_OK) {
            r = nph_pairs_find(b->pairs, b->pin, searched, &pairs, &starts);
        }
        if (r == NPH_RESULT_OK) {
            b->stats.device_pairs_ms += nph_pairs_last_ms(b->pairs);
            f64 detail[3];
            nph_pairs_last_detail(b->pairs, detail);
            b->stats.pairs_gather_ms += detail[0];
            b->stats.pairs_upload_ms += detail[1];
            b->stats.pairs_upload_bytes += (u64)detail[2];
        }
    }
    const f64 t2 = world_now_ms();
    b->stats.pairs_search_ms += t2 - t1;
    if (r == NPH_RESULT_OK && searched > 0U) {
#pragma omp parallel for schedule(dynamic, 1)
        for (u32 k = 0U; k < searched; ++k) {
            NphPhysicsWorld *w = b->worlds[b->world_at[b->pin_pos[k]]];
            pairs_sent(w);
            b->results[b->pin_pos[k]] =
                pairs_finish(w, pairs + 2ULL * starts[k], starts[k + 1U] - starts[k]);
        }
        r = first_error(b);
    }
    b->stats.pairs_create_ms += world_now_ms() - t2;
    b->stats.pairs_ms += world_now_ms() - t;
    if (r != NPH_RESULT_OK) {
        return r;
    }
    t = world_now_ms();
    for (u32 i = s0; i < s0 + m; ++i) {
        NphPhysicsWorld *w = b->worlds[b->world_at[i]];
        NphCollidePart *p = &b->parts[i];
        p->inits = VB(w->inits, NphContactInit);
        p->init_count = w->inits.len;
        p->list = VB(w->awake_contacts, u32);
        p->first = b->first[i];
        p->count = w->awake_contacts.len;
    }
    const u32 *codes[2] = {NULL, NULL};
    r = nph_devb_collide(b->dev[g], b->parts + s0, m, dt, 1U, codes);
    for (u32 i = s0; i < s0 + m; ++i) {
        b->worlds[b->world_at[i]]->inits.len = 0U;
    }
    b->stats.narrow_ms += world_now_ms() - t;
    if (r != NPH_RESULT_OK) {
        return r;
    }
    t = world_now_ms();
    // Each position's codes: its entries of both halves, in group order.
    u64 oa = 0U, ob = 0U;
    for (u32 i = s0; i < s0 + m; ++i) {
        b->off_a[i] = (u32)oa;
        b->off_b[i] = (u32)ob;
        oa += b->first[i];
        ob += b->parts[i].count - b->first[i];
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (u32 i = 0U; i < m; ++i) {
        const u32 pos = s0 + i;
        NphPhysicsWorld *w = b->worlds[b->world_at[pos]];
        const NphCollidePart *p = &b->parts[pos];
        NphResult rk = contacts_apply_codes(w, codes[0] + b->off_a[pos], b->first[pos],
                                            codes[1] + b->off_b[pos], p->count - b->first[pos]);
        if (rk == NPH_RESULT_OK) {
            rk = world_solve_input(w, dt, &b->inputs[pos]);
        }
        b->results[pos] = rk;
    }
    r = first_error(b);
    b->stats.changes_ms += world_now_ms() - t;
    if (r != NPH_RESULT_OK) {
        return r;
    }
    // Worlds with nothing awake skip the solve, as the one-world step does;
    // the rest are packed to the group's front.
    t = world_now_ms();
    u32 active = 0U;
    for (u32 i = s0; i < s0 + m; ++i) {
        NphPhysicsWorld *w = b->worlds[b->world_at[i]];
        if (b->inputs[i].awake_count > 0U) {
            b->devs[s0 + active] = w->dev;
            if (s0 + active != i) {
                b->inputs[s0 + active] = b->inputs[i];
            }
            b->active[s0 + active] = i;
            active += 1U;
            b->stats.solved_contacts += w->solved_contacts;
            b->stats.colored_joints += w->solved_joints;
            b->stats.colored_joint_ends += w->solved_joint_ends;
            b->stats.awake_bodies += w->awake_bodies.len;
        } else {
            w->solved_contacts = 0U;
        }
    }
    b->active_count[g] = active;
    r = nph_devb_solve_begin(b->dev[g], b->devs + s0, b->inputs + s0, active);
    b->begun[g] = r == NPH_RESULT_OK ? 1U : 0U;
    b->stats.solve_ms += world_now_ms() - t;
    return r;
}

// Group g's solve awaited, then its grown proxies and events.
static NphResult group_finish(NphAmdBatch *b, u32 g)
{
    const u32 s0 = b->start[g];
    const u32 active = b->active_count[g];
    const NphGrown *grown = NULL;
    const u32 *starts = NULL;
    f64 t = world_now_ms();
    NphResult r = nph_devb_solve_end(b->dev[g], &grown, &starts);
    b->stats.solve_ms += world_now_ms() - t;
    if (r != NPH_RESULT_OK || !b->begun[g] || active == 0U) {
        return r; // nothing solved for this group this step
    }
    t = world_now_ms();
#pragma omp parallel for schedule(dynamic, 1)
    for (u32 e = 0U; e < active; ++e) {
        const u32 pos = b->active[s0 + e];
        b->results[pos] = world_solve_post(b->worlds[b->world_at[pos]], grown + starts[e],
                                           starts[e + 1U] - starts[e]);
    }
    r = first_error(b);
    b->stats.post_ms += world_now_ms() - t;
    return r;
}

NphResult nph_amd_batch_step(NphAmdBatch *b, f32 dt)
{
    if (!(dt > 0.0F) || !nph_isfinitef(dt)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const u32 n = b->count;
    for (u32 k = 0U; k < n; ++k) {
        if (nph_dev_logging(b->worlds[k]->dev)) {
            return NPH_RESULT_UNSUPPORTED;
        }
    }
    const f64 t0 = world_now_ms();
    NphAmdBatchStats *st = &b->stats;
    st->solved_contacts = 0U;
    st->awake_bodies = 0U;
    st->colored_joints = 0U;
    st->colored_joint_ends = 0U;
    st->pairs_gather_ms = st->pairs_upload_ms = 0.0;
    st->pairs_upload_bytes = 0U;
    st->narrow_ms = st->pairs_ms = st->changes_ms = st->solve_ms = st->post_ms = 0.0;
    st->device_pairs_ms = st->pairs_trees_ms = st->pairs_search_ms = st->pairs_create_ms = 0.0;
    st->rebuild_cpu_ms = 0.0;
    // The groups: worlds by descending awake bodies, each to the group with
    // the fewest so far; inside a group, positions keep that order (the host
    // phases take the largest worlds first).
    b->groups = b->want_groups < b->streams ? b->want_groups : b->streams;
    for (u32 k = 0U; k < n; ++k) {
        b->order[k] = ((u64)(~b->worlds[k]->awake_bodies.len) << 32U) | k;
    }
    qsort(b->order, n, sizeof *b->order, u64_ascending);
    u64 load[NPH_BATCH_GROUPS] = {0U, 0U, 0U, 0U};
    u32 members[NPH_BATCH_GROUPS] = {0U, 0U, 0U, 0U};
    for (u32 i = 0U; i < n; ++i) {
        const u32 k = (u32)b->order[i];
        u32 g = 0U;
        for (u32 h = 1U; h < b->groups; ++h) {
            g = load[h] < load[g] ? h : g;
        }
        load[g] += (u64)b->worlds[k]->awake_bodies.len + 1U;
        members[g] += 1U;
        b->active[i] = g; // the group of the i-th largest world, for a moment
    }
    b->start[0] = 0U;
    for (u32 g = 0U; g < NPH_BATCH_GROUPS; ++g) {
        b->start[g + 1U] = b->start[g] + (g < b->groups ? members[g] : 0U);
        members[g] = 0U;
    }
    for (u32 i = 0U; i < n; ++i) {
        const u32 k = (u32)b->order[i];
        const u32 g = b->active[i];
        const u32 pos = b->start[g] + members[g]++;
        nph_devb_adopt(b->dev[g], b->worlds[k]->dev);
        b->world_at[pos] = k;
        b->devs[pos] = b->worlds[k]->dev;
    }
    for (u32 k = 0U; k < n; ++k) {
        world_step_begin(b->worlds[k], dt);
    }
    NphResult r = NPH_RESULT_OK;
    for (u32 g = 0U; g < NPH_BATCH_GROUPS; ++g) {
        b->begun[g] = 0U;
        b->active_count[g] = 0U;
    }
    f64 t = world_now_ms();
    // Group 1's first narrow-phase half queues behind group 0's: group 0's
    // codes are needed first.
    for (u32 g = 0U; g < b->groups && r == NPH_RESULT_OK; ++g) {
        if (g > 0U) {
            r = nph_devb_follow(b->dev[g], b->dev[g - 1U]);
        }
        if (r == NPH_RESULT_OK) {
            r = narrow_begin(b, g, dt);
        }
    }
    st->narrow_ms += world_now_ms() - t;
    // Group 0 to its solve; group 1 to its solve while the device solves
    // group 0; group 0 finished while the device solves group 1.
    for (u32 g = 0U; g < b->groups && r == NPH_RESULT_OK; ++g) {
        r = group_to_solve(b, g, dt);
    }
    for (u32 g = 0U; g < b->groups; ++g) {
        const NphResult rg = group_finish(b, g); // every begun solve is ended
        r = r == NPH_RESULT_OK ? rg : r;
    }
    if (r == NPH_RESULT_OK) {
#pragma omp parallel for schedule(dynamic, 8)
        for (u32 k = 0U; k < n; ++k) {
            b->results[k] = world_step_end(b->worlds[k]); // host lists only
        }
        r = first_error(b);
    }
    st->device_solve_ms = st->device_passes_ms = st->device_collide_ms = 0.0;
    st->wait_narrow_ms = st->wait_solve_ms = 0.0;
    for (u32 g = 0U; g < b->groups; ++g) {
        f64 dev[5];
        nph_devb_times(b->dev[g], dev);
        st->device_solve_ms += dev[0];
        st->device_passes_ms += dev[1];
        st->device_collide_ms += dev[2];
        st->wait_narrow_ms += dev[3];
        st->wait_solve_ms += dev[4];
    }
    for (u32 i = 0U; i < n; ++i) {
        b->devs[i] = b->worlds[b->world_at[i]]->dev;
        b->results[i] = NPH_RESULT_OK;
    }
    for (u32 k = 0U; k < n; ++k) {
        NphPhysicsWorld *w = b->worlds[k];
        w->stepping = 0;
        w->stats.steps += 1U;
    }
    st->step_ms = world_now_ms() - t0;
    return r;
}

NphResult nph_amd_batch_autotune(NphAmdBatch *b, f32 dt, u32 rounds, NphAmdBatchTuning *chosen)
{
    static const u32 pass_blocks[2] = {128U, 256U};
    static const u32 body_blocks[2] = {64U, 256U};
    enum { CANDIDATES = 32 };
    NphAmdBatchTuning cand[CANDIDATES];
    f64 total[CANDIDATES];
    u32 m = 0U;
    for (u32 groups = 1U; groups <= NPH_BATCH_GROUPS; ++groups) {
        for (u32 g = 0U; g < 2U; ++g) {
            for (u32 p = 0U; p < 2U; ++p) {
                for (u32 q = 0U; q < 2U; ++q) {
                    cand[m].pass_block = pass_blocks[p];
                    cand[m].body_block = body_blocks[q];
                    cand[m].use_graph = g;
                    cand[m].groups = groups;
                    total[m] = 0.0;
                    m += 1U;
                }
            }
        }
    }
    const NphAmdBatchTuning before = nph_amd_batch_tuning(b);
    // Round-robin, every other round backwards, so a trend in the step's
    // cost over the rounds weighs on every candidate alike.
    for (u32 round = 0U; round < (rounds > 0U ? rounds : 1U); ++round) {
        for (u32 i = 0U; i < m; ++i) {
            const u32 c = (round & 1U) != 0U ? m - 1U - i : i;
            nph_amd_batch_set_tuning(b, &cand[c]);
            const NphResult r = nph_amd_batch_step(b, dt);
            if (r != NPH_RESULT_OK) {
                nph_amd_batch_set_tuning(b, &before);
                return r;
            }
            total[c] += b->stats.step_ms;
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
    memset(v, 0, sizeof v);
    u32 steps = 0U;
    for (u32 g = 0U; g < b->streams; ++g) {
        f64 x[8];
        u32 sg = 0U;
        const NphResult r = nph_devb_warm_stats(b->dev[g], x, &sg, reset ? 1 : 0);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        for (u32 j = 0U; j < 8U; ++j) {
            v[j] = (j >= 4U && j <= 6U) ? (x[j] > v[j] ? x[j] : v[j]) : v[j] + x[j];
        }
        steps = sg > steps ? sg : steps;
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
