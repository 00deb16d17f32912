// SPDX-License-Identifier: Apache-2.0
// The world, its bodies and shapes, and the step's order, on the host.
//
//   world create / step / stats / hash   world.c:85-462
//   slots and handles                    world.c:18-49, world_internal.h:20-32
//   bodies and their mass                body.c:12-130, :392-448
//   shapes, world boxes, proxies         shape.c:12-263
//   the solver's host part               solver.c:961-1017 (softness, groups' lists)
//
// The step runs the reference's phases in its order: new pairs (host tree),
// the narrow phase (device), the changes it found (host), the solver and
// finalize (device), continuous sweeps of the fast bodies (host), the
// proxies that grew (host), then islands and sleep (host).  The TGS and
// XPBD solvers, and sleep in a batch's worlds, are NPH_RESULT_UNSUPPORTED.
#define _POSIX_C_SOURCE 200809L
#include "world_impl.h"

#include <stdio.h>
#include <time.h>

// ---------------------------------------------------------------------------
// Allocation
// ---------------------------------------------------------------------------
static void *sys_alloc(NphAllocator *self, u64 size, u32 align, NphCallsite site)
{
    (void)self;
    (void)site;
    void *p = NULL;
    if (posix_memalign(&p, align < 16U ? 16U : align, size == 0U ? 16U : size) != 0) {
        return NULL;
    }
    return p;
}

static void *sys_realloc(NphAllocator *self, void *ptr, u64 old_size, u64 new_size, u32 align,
                         NphCallsite site)
{
    void *p = sys_alloc(self, new_size, align, site);
    if (p != NULL && ptr != NULL) {
        memcpy(p, ptr, old_size < new_size ? old_size : new_size);
        free(ptr);
    }
    return p;
}

static void sys_free(NphAllocator *self, void *ptr, u64 size, u32 align, NphCallsite site)
{
    (void)self;
    (void)size;
    (void)align;
    (void)site;
    free(ptr);
}

static const NphAllocatorVtable g_sys_vtable = {sys_alloc, sys_realloc, sys_free};
static NphAllocator g_sys_alloc = {&g_sys_vtable, "nph_amd.system"};

NphAllocator *nph_amd_allocator_system(void) { return &g_sys_alloc; }

NphResult vbuf_reserve_at(NphAllocator *a, NphVecBuf *v, u32 elem, u32 cap, NphCallsite site)
{
    if (cap <= v->cap) {
        return NPH_RESULT_OK;
    }
    u32 c = v->cap < 16U ? 16U : v->cap;
    while (c < cap) {
        c *= 2U;
    }
    void *p = v->data == NULL ? a->vtable->alloc(a, (u64)c * elem, 16U, site)
                              : a->vtable->realloc(a, v->data, (u64)v->cap * elem, (u64)c * elem,
                                                   16U, site);
    if (p == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    v->data = p;
    v->cap = c;
    return NPH_RESULT_OK;
}

NphResult vbuf_push_at(NphAllocator *a, NphVecBuf *v, u32 elem, const void *item, NphCallsite site)
{
    NphResult r = vbuf_reserve_at(a, v, elem, v->len + 1U, site);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    memcpy((u8 *)v->data + (u64)v->len * elem, item, elem);
    v->len += 1U;
    return NPH_RESULT_OK;
}

void vbuf_free(NphAllocator *a, NphVecBuf *v, u32 elem)
{
    if (v->data != NULL) {
        NphCallsite cs = {__FILE__, (u32)__LINE__};
        a->vtable->free(a, v->data, (u64)v->cap * elem, 16U, cs);
    }
    memset(v, 0, sizeof *v);
}

// ---------------------------------------------------------------------------
// Slots and handles (world.c:18-80)
// ---------------------------------------------------------------------------
#define SLOT_BITS 20U
#define SLOT_MASK ((1U << SLOT_BITS) - 1U)
#define GEN_MAX 4095U

u32 phys_handle(u32 slot, u32 generation) { return (generation << SLOT_BITS) | slot; }
u32 phys_slot(u32 handle) { return handle & SLOT_MASK; }

NphResult slots_alloc(NphAllocator *a, NphSlots *s, u32 *slot, u32 *generation)
{
    if (s->free.len > 0U) {
        const u32 k = VB(s->free, u32)[s->free.len - 1U];
        s->free.len -= 1U;
        *slot = k;
        *generation = VB(s->generation, u32)[k];
        return NPH_RESULT_OK;
    }
    if (s->generation.len >= NPH_PHYSICS_MAX_OBJECTS) {
        return NPH_RESULT_FULL;
    }
    const u32 one = 1U;
    NphResult r = vbuf_push(a, &s->generation, 4U, &one);
    if (r == NPH_RESULT_OK) {
        r = vbuf_reserve(a, &s->free, 4U, s->generation.len);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    *slot = s->generation.len - 1U;
    *generation = 1U;
    return NPH_RESULT_OK;
}

void slots_release(NphAllocator *a, NphSlots *s, u32 slot)
{
    u32 *g = &VB(s->generation, u32)[slot];
    *g = *g >= GEN_MAX ? 1U : *g + 1U;
    (void)vbuf_push(a, &s->free, 4U, &slot);
}

HBody *world_body(const NphPhysicsWorld *w, NphPhysicsBody h)
{
    const u32 slot = phys_slot(h.value);
    if (h.value == 0U || slot >= w->bodies.len || VB(w->bodies, HBody)[slot].handle != h.value) {
        return NULL;
    }
    return &VB(w->bodies, HBody)[slot];
}

HShape *world_shape(const NphPhysicsWorld *w, NphPhysicsShape h)
{
    const u32 slot = phys_slot(h.value);
    if (h.value == 0U || slot >= w->shapes.len || VB(w->shapes, HShape)[slot].handle != h.value) {
        return NULL;
    }
    return &VB(w->shapes, HShape)[slot];
}

static NphBodyRec *rec_at(NphPhysicsWorld *w, u32 slot) { return &VB(w->body_recs, NphBodyRec)[slot]; }

// solver.c:886-896: the step's move events, one per body of the awake list
// as the step left it (its first `move_pending` entries), in its order, with
// the poses and velocities the step wrote. Built from the host copy the
// first time it is brought up to date after the step, so before any change
// to a body or to the awake list.
static NphResult build_move_events(NphPhysicsWorld *w)
{
    const u32 n = w->move_pending;
    NphResult r = vbuf_reserve(w->alloc, &w->move_events, sizeof(NphPhysicsBodyMoveEvent),
                               n > 0U ? n : 1U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    for (u32 i = 0U; i < n; ++i) {
        const u32 slot = VB(w->awake_bodies, u32)[i];
        HBody *hb = &VB(w->bodies, HBody)[slot];
        const NphBodyRec *b = rec_at(w, slot);
        NphPhysicsBodyMoveEvent *m = &VB(w->move_events, NphPhysicsBodyMoveEvent)[i];
        m->body.value = hb->handle;
        m->fell_asleep = 0U;
        m->user_data = hb->user_data;
        m->position = b->origin;
        m->rotation = b->rotation;
        m->linear_velocity = b->v;
        m->angular_velocity = b->w;
        hb->move_event = i;
    }
    w->move_events.len = n;
    w->move_events_ready = 1;
    return NPH_RESULT_OK;
}

NphResult world_sync_bodies(NphPhysicsWorld *w)
{
    if (w->bodies_stale) {
        NphResult r = nph_dev_read_bodies(w->dev, VB(w->body_recs, NphBodyRec), w->bodies.len);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        w->bodies_stale = 0;
    }
    return w->move_events_ready ? NPH_RESULT_OK : build_move_events(w);
}

NphResult world_upload_body(NphPhysicsWorld *w, u32 slot)
{
    return nph_dev_write_body(w->dev, slot, rec_at(w, slot));
}

// ---------------------------------------------------------------------------
// World lifecycle (world.c:85-215)
// ---------------------------------------------------------------------------
NphPhysicsWorldDesc nph_physics_world_desc_default(void)
{
    NphPhysicsWorldDesc d;
    memset(&d, 0, sizeof d);
    d.gravity = v3(0.0F, -9.81F, 0.0F);
    d.substeps = 4U;
    d.contact_hertz = 0.0F;
    d.contact_damping_ratio = 10.0F;
    d.joint_hertz = 0.0F;
    d.joint_damping_ratio = 2.0F;
    d.contact_push_max_velocity = 3.0F;
    d.restitution_threshold = 1.0F;
    d.max_linear_speed = 400.0F;
    d.time_to_sleep = 0.5F;
    d.enable_sleep = 1;
    d.enable_continuous = 1;
    d.body_capacity = 64U;
    d.shape_capacity = 64U;
    d.contact_capacity = 256U;
    return d;
}

static int finite_nonneg(f32 v) { return nph_isfinitef(v) && v >= 0.0F; }

void nph_physics_world_destroy(NphPhysicsWorld *w);

NphResult world_create(const NphPhysicsWorldDesc *desc, NphDeviceBatch *batch,
                       NphPhysicsWorld **out)
{
    if (desc == NULL || out == NULL || desc->alloc == NULL || desc->substeps == 0U ||
        (u32)desc->solver >= (u32)NPH_PHYSICS_SOLVER_COUNT || desc->substeps > 64U ||
        !nph_isfinitef(desc->gravity.x) || !nph_isfinitef(desc->gravity.y) ||
        !nph_isfinitef(desc->gravity.z) || !finite_nonneg(desc->contact_hertz) ||
        !finite_nonneg(desc->contact_damping_ratio) || !finite_nonneg(desc->joint_hertz) ||
        !finite_nonneg(desc->joint_damping_ratio) ||
        !finite_nonneg(desc->contact_push_max_velocity) ||
        !finite_nonneg(desc->restitution_threshold) || !(desc->max_linear_speed > 0.0F) ||
        !nph_isfinitef(desc->max_linear_speed) || !finite_nonneg(desc->time_to_sleep)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    // A batch's worlds step without sleep (their records are read back only
    // when a step's events need them).
    if (desc->solver != NPH_PHYSICS_SOLVER_SOFT_STEP || (batch != NULL && desc->enable_sleep != 0)) {
        return NPH_RESULT_UNSUPPORTED;
    }
    NphAllocator *a = desc->alloc;
    NphCallsite cs = {__FILE__, (u32)__LINE__};
    NphPhysicsWorld *w = (NphPhysicsWorld *)a->vtable->alloc(a, sizeof *w, 16U, cs);
    if (w == NULL) {
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    memset(w, 0, sizeof *w);
    w->alloc = a;
    w->desc = *desc;
    w->hit_threshold = 1.0F;
    w->move_events_ready = 1;
    NphResult r = batch != NULL ? nph_devb_world(batch, &w->dev) : nph_dev_create(&w->dev);
    for (u32 t = 0U; t < 3U && r == NPH_RESULT_OK; ++t) {
        r = tree_init(&w->trees[t], a, 16U);
    }
    if (r == NPH_RESULT_OK) {
        r = pairmap_init(a, &w->pairs);
    }
    if (r == NPH_RESULT_OK) {
        r = nph_dev_reserve(w->dev, desc->body_capacity, desc->shape_capacity,
                            desc->contact_capacity);
    }
    if (r != NPH_RESULT_OK) {
        nph_physics_world_destroy(w);
        return r;
    }
    *out = w;
    return NPH_RESULT_OK;
}

NphResult nph_physics_world_create(const NphPhysicsWorldDesc *desc, NphPhysicsWorld **out)
{
    return world_create(desc, NULL, out);
}

void nph_physics_world_destroy(NphPhysicsWorld *w)
{
    if (w == NULL) {
        return;
    }
    NphAllocator *a = w->alloc;
    nph_dev_destroy(w->dev);
    vbuf_free(a, &w->body_slots.generation, 4U);
    vbuf_free(a, &w->body_slots.free, 4U);
    vbuf_free(a, &w->bodies, sizeof(HBody));
    vbuf_free(a, &w->body_recs, sizeof(NphBodyRec));
    vbuf_free(a, &w->shape_slots.generation, 4U);
    vbuf_free(a, &w->shape_slots.free, 4U);
    vbuf_free(a, &w->shapes, sizeof(HShape));
    vbuf_free(a, &w->contacts, sizeof(HContact));
    vbuf_free(a, &w->free_contacts, 4U);
    vbuf_free(a, &w->awake_bodies, 4U);
    vbuf_free(a, &w->awake_contacts, 4U);
    vbuf_free(a, &w->inits, sizeof(NphContactInit));
    vbuf_free(a, &w->sweep_shapes, sizeof(NphShapeRec));
    vbuf_free(a, &w->move_buffer, 4U);
    nph_pairs_destroy(w->pair_search);
    vbuf_free(a, &w->pf_dirty_shapes, 4U);
    vbuf_free(a, &w->pf_touched, 8U);
    vbuf_free(a, &w->pf_keys, 8U);
    vbuf_free(a, &w->pf_key_in, 1U);
    vbuf_free(a, &w->pf_shape_info, sizeof(NphPairShape));
    vbuf_free(a, &w->pf_moved_shapes, 4U);
    for (u32 c = 0U; c < NPH_GRAPH_COLORS; ++c) {
        vbuf_free(a, &w->color_bits[c], 8U);
    }
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        vbuf_free(a, &w->color_contacts[c], 4U);
    }
    vbuf_free(a, &w->graph, 4U);
    joints_free(w);
    vbuf_free(a, &w->changed, 4U);
    vbuf_free(a, &w->codes, 4U);
    vbuf_free(a, &w->begin_events, sizeof(NphPhysicsContactBeginEvent));
    vbuf_free(a, &w->hit_events, sizeof(NphPhysicsContactHitEvent));
    vbuf_free(a, &w->end_events, sizeof(NphPhysicsContactEndEvent));
    vbuf_free(a, &w->move_events, sizeof(NphPhysicsBodyMoveEvent));
    vbuf_free(a, &w->contact_view, sizeof(NphContactRec));
    vbuf_free(a, &w->shape_view, sizeof(NphShapeRec));
    islands_free(w);
    for (u32 t = 0U; t < 3U; ++t) {
        tree_free(&w->trees[t]);
    }
    pairmap_free(a, &w->pairs);
    NphCallsite cs = {__FILE__, (u32)__LINE__};
    a->vtable->free(a, w, sizeof *w, 16U, cs);
}

void nph_physics_world_set_gravity(NphPhysicsWorld *w, NphVec3 gravity) { w->desc.gravity = gravity; }
NphVec3 nph_physics_world_gravity(const NphPhysicsWorld *w) { return w->desc.gravity; }
void nph_physics_world_set_hit_threshold(NphPhysicsWorld *w, f32 speed) { w->hit_threshold = speed; }
const char *nph_amd_backend_name(const NphPhysicsWorld *w)
{
    (void)w;
    return nph_dev_name();
}

u32 nph_amd_solved_contacts(const NphPhysicsWorld *w) { return w->solved_contacts; }

void nph_amd_device_times(const NphPhysicsWorld *w, f64 out[3])
{
    out[0] = nph_dev_last_collide_ms(w->dev);
    out[1] = nph_dev_last_solve_ms(w->dev);
    out[2] = nph_dev_last_passes_ms(w->dev);
}

void nph_amd_time_kernels(NphPhysicsWorld *w, bool enable) { nph_dev_time_kernels(w->dev, enable); }

void nph_amd_kernel_times(const NphPhysicsWorld *w, f64 out[2]) { nph_dev_kernel_times(w->dev, out); }

NphResult nph_amd_world_set_warm_start(NphPhysicsWorld *w, const NphAmdWarmStart *ws)
{
    if (w == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    return nph_dev_set_warm_start(w->dev, ws);
}

void nph_amd_world_measure_residual(NphPhysicsWorld *w, bool enable)
{
    nph_dev_measure_residual(w->dev, enable ? 1 : 0);
}

NphResult nph_amd_world_warm_stats(const NphPhysicsWorld *w, NphAmdWarmStats *out)
{
    if (w == NULL || out == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    return nph_dev_warm_stats(w->dev, out);
}

NphResult nph_amd_world_log_contacts(NphPhysicsWorld *w, u32 capacity)
{
    if (w == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    return nph_dev_log_contacts(w->dev, capacity);
}

NphResult nph_amd_world_contact_log(const NphPhysicsWorld *w, u32 *rows, const f32 **features,
                                    const f32 **targets)
{
    if (w == NULL || rows == NULL || features == NULL || targets == NULL) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    return nph_dev_contact_log(w->dev, rows, features, targets);
}

// ---------------------------------------------------------------------------
// Bodies (body.c:12-130, :392-448)
// ---------------------------------------------------------------------------
NphPhysicsBodyDesc nph_physics_body_desc_default(void)
{
    NphPhysicsBodyDesc d;
    memset(&d, 0, sizeof d);
    d.type = NPH_PHYSICS_BODY_DYNAMIC;
    d.rotation = q_identity();
    d.gravity_scale = 1.0F;
    d.sleep_threshold = 0.05F;
    d.enable_sleep = 1;
    d.awake = 1;
    d.gyroscopic = 1;
    return d;
}

static int quat_is_normalized(NphQuat q, f32 tol)
{
    return nph_absf(q_len_sq(q) - 1.0F) <= 2.0F * tol;
}

static int body_desc_valid(const NphPhysicsBodyDesc *d)
{
    const NphDVec3 p = d->position;
    const NphQuat q = d->rotation;
    return (u32)d->type <= (u32)NPH_PHYSICS_BODY_DYNAMIC && __builtin_isfinite(p.x) &&
           __builtin_isfinite(p.y) && __builtin_isfinite(p.z) && nph_isfinitef(q.x) &&
           nph_isfinitef(q.y) && nph_isfinitef(q.z) && nph_isfinitef(q.w) &&
           quat_is_normalized(q, 1e-3F) && nph_isfinitef(d->linear_velocity.x) &&
           nph_isfinitef(d->linear_velocity.y) && nph_isfinitef(d->linear_velocity.z) &&
           nph_isfinitef(d->angular_velocity.x) && nph_isfinitef(d->angular_velocity.y) &&
           nph_isfinitef(d->angular_velocity.z) && finite_nonneg(d->linear_damping) &&
           finite_nonneg(d->angular_damping) && nph_isfinitef(d->gravity_scale) &&
           finite_nonneg(d->sleep_threshold) && (d->locks & ~0x3FU) == 0U;
}

static NphResult grow_color_bits(NphPhysicsWorld *w, u32 slots)
{
    const u32 words = (slots + 63U) / 64U;
    for (u32 c = 0U; c < NPH_GRAPH_COLORS; ++c) {
        NphVecBuf *b = &w->color_bits[c];
        if (b->len < words) {
            NphResult r = vbuf_reserve(w->alloc, b, 8U, words);
            if (r != NPH_RESULT_OK) {
                return r;
            }
            memset(VB(*b, u64) + b->len, 0, (size_t)(words - b->len) * 8U);
            b->len = words;
        }
    }
    return NPH_RESULT_OK;
}

static HShape *shape_slot(NphPhysicsWorld *w, u32 s) { return &VB(w->shapes, HShape)[s]; }

// The hull or mesh block a geometry points at (host memory), or NULL.
static const void *geometry_block(const NphPhysicsGeometry *g)
{
    if (g->type == NPH_PHYSICS_SHAPE_HULL) {
        return g->hull;
    }
    if (g->type == NPH_PHYSICS_SHAPE_MESH) {
        return g->mesh;
    }
    return NULL;
}

static NphGeom host_geom(const HShape *s) { return geom_from(&s->geometry, geometry_block(&s->geometry)); }

NphBodyRec *world_rec(NphPhysicsWorld *w, u32 slot) { return rec_at(w, slot); }

// body.c:392-448 (the caller's mass properties kept when set)
static void update_mass(NphPhysicsWorld *w, u32 slot)
{
    HBody *hb = &VB(w->bodies, HBody)[slot];
    NphBodyRec *b = rec_at(w, slot);
    f32 extent = 0.0F;
    if ((hb->flags & HB_MASS_SET) == 0U) {
        NphPhysicsMassData total;
        memset(&total, 0, sizeof total);
        for (u32 s = hb->shape_head; s != NPH_NULL32; s = shape_slot(w, s)->next) {
            const HShape *shape = shape_slot(w, s);
            if (shape->density <= 0.0F || (shape->flags & (HS_SENSOR | HS_GHOST)) != 0U) {
                continue;
            }
            const NphGeom g = host_geom(shape);
            total = mass_add(total, mass_transform(geom_mass(&g, shape->density), shape->pose));
        }
        b->mass = total.mass;
        b->local_center = total.center;
        b->inertia = total.inertia;
    }
    f32 inner = NPH_F32_MAX;
    for (u32 s = hb->shape_head; s != NPH_NULL32; s = shape_slot(w, s)->next) {
        const HShape *shape = shape_slot(w, s);
        const NphGeom g = host_geom(shape);
        const f32 d = v3_dist(shape->pose.position, b->local_center) + geom_outer_radius(&g);
        extent = nph_maxf(extent, d);
        inner = nph_minf(inner, geom_inner_radius(&g));
    }
    b->max_extent = extent;
    b->min_extent = inner;
    if (hb->type == NPH_PHYSICS_BODY_DYNAMIC) {
        if (b->mass > 0.0F) {
            b->inv_mass = 1.0F / b->mass;
        } else {
            b->mass = 1.0F;
            b->inv_mass = 1.0F;
            b->inertia = m3_diag(v3_splat(0.4F));
            b->local_center = v3_zero();
        }
        if (!m3_inverse(b->inertia, &b->inv_inertia)) {
            const f32 k = 0.4F * b->mass * nph_maxf(extent * extent, 1e-4F);
            b->inertia = m3_diag(v3_splat(k));
            b->inv_inertia = m3_diag(v3_splat(1.0F / k));
        }
    } else {
        b->inv_mass = 0.0F;
        b->inv_inertia = m3_zero();
    }
    b->center = dv3_add_v3(b->origin, q_rotate(b->rotation, b->local_center));
}

NphResult nph_physics_body_create(NphPhysicsWorld *w, const NphPhysicsBodyDesc *desc,
                                  NphPhysicsBody *out)
{
    if (w == NULL || out == NULL || desc == NULL || !body_desc_valid(desc)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    u32 slot = 0U;
    u32 generation = 0U;
    NphResult r = slots_alloc(w->alloc, &w->body_slots, &slot, &generation);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (slot >= w->bodies.len) {
        r = vbuf_reserve(w->alloc, &w->bodies, sizeof(HBody), slot + 1U);
        if (r == NPH_RESULT_OK) {
            r = vbuf_reserve(w->alloc, &w->body_recs, sizeof(NphBodyRec), slot + 1U);
        }
        if (r == NPH_RESULT_OK) {
            r = grow_color_bits(w, slot + 1U);
        }
        if (r == NPH_RESULT_OK) {
            r = nph_dev_reserve(w->dev, slot + 1U, 0U, 0U);
        }
        if (r != NPH_RESULT_OK) {
            slots_release(w->alloc, &w->body_slots, slot);
            return r;
        }
        w->bodies.len = slot + 1U;
        w->body_recs.len = slot + 1U;
    }
    HBody *hb = &VB(w->bodies, HBody)[slot];
    NphBodyRec *b = rec_at(w, slot);
    memset(hb, 0, sizeof *hb);
    memset(b, 0, sizeof *b);
    hb->type = (u32)desc->type;
    b->type = (u32)desc->type;
    w->kinematic_bodies += desc->type == NPH_PHYSICS_BODY_KINEMATIC ? 1U : 0U;
    b->locks = desc->locks;
    b->origin = desc->position;
    b->rotation = q_normalize(desc->rotation);
    b->v = desc->type == NPH_PHYSICS_BODY_STATIC ? v3_zero() : desc->linear_velocity;
    b->w = desc->type == NPH_PHYSICS_BODY_STATIC ? v3_zero() : desc->angular_velocity;
    b->linear_damping = desc->linear_damping;
    b->angular_damping = desc->angular_damping;
    b->gravity_scale = desc->gravity_scale;
    b->sleep_threshold = desc->sleep_threshold;
    hb->flags = (desc->enable_sleep != 0 ? HB_ENABLE_SLEEP : 0U) |
                (desc->gyroscopic != 0 ? HB_GYROSCOPIC : 0U) |
                (desc->bullet != 0 && desc->type == NPH_PHYSICS_BODY_DYNAMIC ? HB_BULLET : 0U);
    b->flags = hb->flags & (HB_ENABLE_SLEEP | HB_GYROSCOPIC | HB_BULLET);
    hb->shape_head = NPH_NULL32;
    b->shape_head = NPH_NULL32;
    hb->contact_head = NPH_NULL32;
    hb->joint_head = NPH_NULL32;
    hb->awake_slot = NPH_NULL32;
    hb->island = NPH_NULL32;
    hb->island_prev = NPH_NULL32;
    hb->island_next = NPH_NULL32;
    hb->move_event = NPH_NULL32;
    hb->user_data = desc->user_data;
    hb->handle = phys_handle(slot, generation);
    update_mass(w, slot);
    r = world_upload_body(w, slot);
    if (r == NPH_RESULT_OK && desc->type == NPH_PHYSICS_BODY_DYNAMIC) {
        r = islands_add_body(w, slot);
    }
    if (r == NPH_RESULT_OK && desc->type != NPH_PHYSICS_BODY_STATIC) {
        // A body made asleep sleeps alone: its own island, already still.
        r = world_set_awake_list(w, slot, desc->awake != 0);
    }
    if (r != NPH_RESULT_OK) {
        hb = &VB(w->bodies, HBody)[slot];
        if (hb->island != NPH_NULL32) {
            islands_remove_body(w, slot);
        }
        hb->handle = 0U;
        w->kinematic_bodies -= hb->type == NPH_PHYSICS_BODY_KINEMATIC ? 1U : 0U;
        slots_release(w->alloc, &w->body_slots, slot);
        return r;
    }
    out->value = hb->handle;
    return NPH_RESULT_OK;
}

bool nph_physics_body_is_valid(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    return world_body(w, body) != NULL;
}

NphPhysicsBodyType nph_physics_body_type(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const HBody *b = world_body(w, body);
    return b != NULL ? (NphPhysicsBodyType)b->type : NPH_PHYSICS_BODY_STATIC;
}

static const NphBodyRec *body_rec(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    if (world_body(w, body) == NULL) {
        return NULL;
    }
    NphPhysicsWorld *mw = (NphPhysicsWorld *)w;
    if (world_sync_bodies(mw) != NPH_RESULT_OK) {
        return NULL;
    }
    return rec_at(mw, phys_slot(body.value));
}

NphDVec3 nph_physics_body_position(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const NphBodyRec *b = body_rec(w, body);
    return b != NULL ? b->origin : dv3(0.0, 0.0, 0.0);
}

NphQuat nph_physics_body_rotation(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const NphBodyRec *b = body_rec(w, body);
    return b != NULL ? b->rotation : q_identity();
}

NphDVec3 nph_physics_body_center_of_mass(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const NphBodyRec *b = body_rec(w, body);
    return b != NULL ? b->center : dv3(0.0, 0.0, 0.0);
}

NphVec3 nph_physics_body_linear_velocity(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const NphBodyRec *b = body_rec(w, body);
    return b != NULL ? b->v : v3_zero();
}

NphVec3 nph_physics_body_angular_velocity(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const NphBodyRec *b = body_rec(w, body);
    return b != NULL ? b->w : v3_zero();
}

static void set_velocity(NphPhysicsWorld *w, NphPhysicsBody body, NphVec3 v, int angular)
{
    HBody *hb = world_body(w, body);
    if (hb == NULL || hb->type == NPH_PHYSICS_BODY_STATIC || world_sync_bodies(w) != NPH_RESULT_OK) {
        return;
    }
    const u32 slot = phys_slot(body.value);
    if (angular) {
        rec_at(w, slot)->w = v;
    } else {
        rec_at(w, slot)->v = v;
    }
    if (v3_dot(v, v) > 0.0F) {
        (void)world_wake_body(w, slot);
    }
    (void)world_upload_body(w, slot);
}

void nph_physics_body_set_linear_velocity(NphPhysicsWorld *w, NphPhysicsBody body, NphVec3 v)
{
    set_velocity(w, body, v, 0);
}

void nph_physics_body_set_angular_velocity(NphPhysicsWorld *w, NphPhysicsBody body, NphVec3 v)
{
    set_velocity(w, body, v, 1);
}

NphPhysicsMassData nph_physics_body_mass_data(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    NphPhysicsMassData m;
    memset(&m, 0, sizeof m);
    const NphBodyRec *b = body_rec(w, body);
    if (b != NULL) {
        m.mass = b->mass;
        m.center = b->local_center;
        m.inertia = b->inertia;
    }
    return m;
}

u64 nph_physics_body_user_data(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const HBody *b = world_body(w, body);
    return b != NULL ? b->user_data : 0U;
}

bool nph_physics_body_is_awake(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const HBody *b = world_body(w, body);
    return b != NULL && (b->flags & HB_AWAKE) != 0U;
}

// ---------------------------------------------------------------------------
// Geometry (geometry.c) through the public calls
// ---------------------------------------------------------------------------
NphPhysicsMassData nph_physics_geometry_mass(const NphPhysicsGeometry *g, f32 density)
{
    const NphGeom gg = geom_from(g, geometry_block(g));
    return geom_mass(&gg, density);
}

NphAABB nph_physics_geometry_aabb(const NphPhysicsGeometry *g, NphPhysicsPose pose)
{
    const NphGeom gg = geom_from(g, geometry_block(g));
    return geom_aabb(&gg, pose);
}

static int size_ok(f32 v) { return v > 0.0F && v <= NPH_PHYSICS_MAX_COORDINATE; }

bool nph_physics_geometry_is_valid(const NphPhysicsGeometry *g)
{
    if (g == NULL) {
        return false;
    }
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        return size_ok(g->sphere.radius);
    case NPH_PHYSICS_SHAPE_CAPSULE:
        return size_ok(g->capsule.radius) && g->capsule.half_height >= 0.0F &&
               g->capsule.half_height <= NPH_PHYSICS_MAX_COORDINATE;
    case NPH_PHYSICS_SHAPE_BOX:
        return size_ok(g->box.half_extents.x) && size_ok(g->box.half_extents.y) &&
               size_ok(g->box.half_extents.z);
    case NPH_PHYSICS_SHAPE_HULL:
        return g->hull != NULL && g->hull->vertex_count >= 4U;
    case NPH_PHYSICS_SHAPE_MESH:
        return g->mesh != NULL && g->mesh->triangle_count > 0U;
    default:
        return false;
    }
}

// ---------------------------------------------------------------------------
// Shapes (shape.c:12-263)
// ---------------------------------------------------------------------------
NphPhysicsFilter nph_physics_filter_default(void)
{
    NphPhysicsFilter f;
    f.category = 1U;
    f.mask = ~0ULL;
    f.group = 0;
    return f;
}

bool nph_physics_filter_should_collide(NphPhysicsFilter a, NphPhysicsFilter b)
{
    if (a.group == b.group && a.group != 0) {
        return a.group > 0;
    }
    return (a.category & b.mask) != 0U && (b.category & a.mask) != 0U;
}

NphPhysicsShapeDesc nph_physics_shape_desc_default(NphPhysicsGeometry geometry)
{
    NphPhysicsShapeDesc d;
    memset(&d, 0, sizeof d);
    d.geometry = geometry;
    d.pose = pose_identity();
    d.density = 1000.0F;
    d.friction = 0.6F;
    d.filter = nph_physics_filter_default();
    d.contact_events = 1;
    d.sensor_events = 1;
    return d;
}

static NphShapeRec shape_record(NphPhysicsWorld *w, u32 s, const void *device_block)
{
    const HShape *h = shape_slot(w, s);
    NphShapeRec r;
    memset(&r, 0, sizeof r);
    r.geom = geom_from(&h->geometry, device_block);
    r.pose = h->pose;
    r.aabb = h->aabb;
    r.fat_aabb = h->fat_aabb;
    r.friction = h->friction;
    r.restitution = h->restitution;
    r.body = h->body;
    r.next = h->next;
    r.combine = h->combine;
    if (h->geometry.type == NPH_PHYSICS_SHAPE_HULL &&
        h->geometry.hull->vertex_count > NPH_HULL_MAX_VERTICES) {
        r.flags = NPH_SHAPE_BIG_HULL;
    }
    return r;
}

// The device copy of a geometry's hull or mesh block (made once a block).
static NphResult device_block_of(NphPhysicsWorld *w, const NphPhysicsGeometry *g,
                                 const void **out)
{
    *out = NULL;
    if (g->type == NPH_PHYSICS_SHAPE_HULL) {
        return nph_dev_block(w->dev, g->hull, g->hull->byte_size, out);
    }
    if (g->type == NPH_PHYSICS_SHAPE_MESH) {
        return nph_dev_block(w->dev, g->mesh, g->mesh->byte_size, out);
    }
    return NPH_RESULT_OK;
}

NphResult world_shape_record(NphPhysicsWorld *w, u32 s, NphShapeRec *out)
{
    const HShape *h = shape_slot(w, s);
    const void *device_block = NULL;
    NphResult r = device_block_of(w, &h->geometry, &device_block);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    *out = shape_record(w, s, device_block);
    return NPH_RESULT_OK;
}

NphGeom world_host_geom(const HShape *s) { return host_geom(s); }

void world_update_mass(NphPhysicsWorld *w, u32 slot) { update_mass(w, slot); }

NphResult world_refresh_shape(NphPhysicsWorld *w, u32 s)
{
    NphShapeRec rec;
    NphResult r = world_shape_record(w, s, &rec);
    w->shape_view_ok = 0;
    return r == NPH_RESULT_OK ? nph_dev_write_shape(w->dev, s, &rec) : r;
}

NphResult nph_physics_shape_create(NphPhysicsWorld *w, NphPhysicsBody body,
                                   const NphPhysicsShapeDesc *desc, NphPhysicsShape *out)
{
    if (w == NULL || out == NULL || desc == NULL || !nph_physics_geometry_is_valid(&desc->geometry) ||
        !nph_isfinitef(desc->pose.position.x) || !nph_isfinitef(desc->pose.position.y) ||
        !nph_isfinitef(desc->pose.position.z) || !quat_is_normalized(desc->pose.rotation, 1e-3F) ||
        !(desc->density >= 0.0F) || !nph_isfinitef(desc->density) || !(desc->friction >= 0.0F) ||
        !nph_isfinitef(desc->friction) || !(desc->restitution >= 0.0F) ||
        !(desc->restitution <= 1.0F) || (u32)desc->friction_combine >= NPH_PHYSICS_COMBINE_COUNT ||
        (u32)desc->restitution_combine >= NPH_PHYSICS_COMBINE_COUNT) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    HBody *hb = world_body(w, body);
    if (hb == NULL) {
        return NPH_RESULT_NOT_FOUND;
    }
    // A mesh bounds no volume, so it has no mass to move a body with
    // (shape.c:196-201).
    if (desc->geometry.type == NPH_PHYSICS_SHAPE_MESH &&
        (hb->type == NPH_PHYSICS_BODY_DYNAMIC || desc->sensor != 0)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    if (desc->sensor != 0 || desc->ghost != 0) {
        return NPH_RESULT_UNSUPPORTED; // sensors and ghosts: not on this path
    }
    // The device walks a mesh's tree with a fixed stack: only a block the
    // validator accepts (every one the builder writes) fits it.
    if (desc->geometry.type == NPH_PHYSICS_SHAPE_MESH &&
        (desc->geometry.mesh == NULL ||
         !nph_physics_mesh_validate(desc->geometry.mesh, desc->geometry.mesh->byte_size))) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    NphResult r = world_sync_bodies(w);
    const void *device_block = NULL;
    if (r == NPH_RESULT_OK) {
        r = device_block_of(w, &desc->geometry, &device_block);
    }
    const u32 body_slot = phys_slot(body.value);
    u32 slot = 0U;
    u32 generation = 0U;
    if (r == NPH_RESULT_OK) {
        r = slots_alloc(w->alloc, &w->shape_slots, &slot, &generation);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (slot >= w->shapes.len) {
        r = vbuf_reserve(w->alloc, &w->shapes, sizeof(HShape), slot + 1U);
        if (r == NPH_RESULT_OK) {
            r = nph_dev_reserve(w->dev, 0U, slot + 1U, 0U);
        }
        if (r != NPH_RESULT_OK) {
            slots_release(w->alloc, &w->shape_slots, slot);
            return r;
        }
        w->shapes.len = slot + 1U;
    }
    hb = &VB(w->bodies, HBody)[body_slot];
    HShape *s = shape_slot(w, slot);
    memset(s, 0, sizeof *s);
    s->body = body_slot;
    s->geometry = desc->geometry;
    s->pose = pose_make(desc->pose.position, q_normalize(desc->pose.rotation));
    s->density = desc->density;
    s->friction = desc->friction;
    s->restitution = desc->restitution;
    s->combine = (u32)desc->friction_combine | ((u32)desc->restitution_combine << 8U);
    s->filter = desc->filter;
    s->flags = (desc->contact_events != 0 ? HS_CONTACT_EVENTS : 0U) |
               (desc->hit_events != 0 ? HS_HIT_EVENTS : 0U);
    w->hit_event_shapes += desc->hit_events != 0 ? 1U : 0U;
    s->user_data = desc->user_data;
    s->next = hb->shape_head;
    hb->shape_head = slot;
    hb->shape_count += 1U;
    NphBodyRec *b = rec_at(w, body_slot);
    b->shape_head = slot;
    update_mass(w, body_slot);
    // The proxy (shape.c:128-146): static shapes only a speculative margin.
    s->tree = hb->type == NPH_PHYSICS_BODY_STATIC      ? NPH_TREE_STATIC
              : hb->type == NPH_PHYSICS_BODY_KINEMATIC ? NPH_TREE_KINEMATIC
                                                        : NPH_TREE_DYNAMIC;
    {
        NphShapeRec tmp = shape_record(w, slot, device_block);
        tmp.geom = host_geom(s); // the host's own hull address for the box
        s->aabb = shape_world_aabb(b, &tmp);
    }
    const f32 margin =
        s->tree == NPH_TREE_STATIC ? NPH_PHYSICS_SPECULATIVE_DISTANCE : NPH_PHYSICS_AABB_MARGIN;
    s->fat_aabb = aabb_grow(s->aabb, margin);
    r = tree_create_proxy(&w->trees[s->tree], s->fat_aabb, s->filter.category, slot, &s->proxy);
    if (r == NPH_RESULT_OK) {
        pairs_mark_shape(w, slot);
        r = move_buffer_push(w, slot);
    }
    if (r == NPH_RESULT_OK) {
        const NphShapeRec rec = shape_record(w, slot, device_block);
        w->shape_view_ok = 0;
        r = nph_dev_write_shape(w->dev, slot, &rec);
        if (desc->geometry.type == NPH_PHYSICS_SHAPE_MESH) {
            nph_dev_note_mesh_shape(w->dev);
        }
        if ((rec.flags & NPH_SHAPE_BIG_HULL) != 0U) {
            nph_dev_note_big_hull_shape(w->dev);
        }
    }
    if (r == NPH_RESULT_OK && hb->type != NPH_PHYSICS_BODY_STATIC) {
        r = world_wake_body(w, body_slot);
    }
    if (r == NPH_RESULT_OK) {
        r = world_upload_body(w, body_slot);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    s->handle = phys_handle(slot, generation);
    out->value = s->handle;
    return NPH_RESULT_OK;
}

bool nph_physics_shape_is_valid(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    return world_shape(w, shape) != NULL;
}

NphPhysicsBody nph_physics_shape_body(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    NphPhysicsBody b = {0U};
    const HShape *s = world_shape(w, shape);
    if (s != NULL) {
        b.value = VB(w->bodies, HBody)[s->body].handle;
    }
    return b;
}

NphAABB nph_physics_shape_aabb(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    if (s == NULL) {
        return aabb_make(v3_zero(), v3_zero());
    }
    NphShapeRec r;
    if (nph_dev_read_shapes(w->dev, NULL, 0U) == NPH_RESULT_OK) {
        // Tight boxes of moving bodies live on the device; read this one.
        NphShapeRec *all = (NphShapeRec *)malloc((size_t)w->shapes.len * sizeof(NphShapeRec));
        if (all != NULL && nph_dev_read_shapes(w->dev, all, w->shapes.len) == NPH_RESULT_OK) {
            r = all[phys_slot(shape.value)];
            free(all);
            return r.aabb;
        }
        free(all);
    }
    return s->aabb;
}

// ---------------------------------------------------------------------------
// The step (world.c:307-383, solver.c:961-1105)
// ---------------------------------------------------------------------------
static f64 now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (f64)ts.tv_sec * 1e3 + (f64)ts.tv_nsec * 1e-6;
}

static NphResult read_contacts(NphPhysicsWorld *w);

NphResult world_read_contacts(NphPhysicsWorld *w) { return read_contacts(w); }

// solver.c:919-958: after the step, one hit a pair for the graph's
// contacts that asked for them, in colour order: the pair's fastest
// approaching point (over a mesh pair's chain) that carried load, when it
// beats the world's threshold.  The records are read back only when some
// contact in the graph wants hits.
static NphResult hit_events(NphPhysicsWorld *w)
{
    // A contact's hit flag comes from its shapes, which outlive it.
    if (w->hit_event_shapes == 0U) {
        return NPH_RESULT_OK;
    }
    int any = 0;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS && !any; ++c) {
        const u32 *list = VB(w->color_contacts[c], u32);
        for (u32 i = 0U; i < w->color_contacts[c].len; ++i) {
            if ((VB(w->contacts, HContact)[list[i]].flags & HC_HIT_EVENTS) != 0U) {
                any = 1;
                break;
            }
        }
    }
    if (!any) {
        return NPH_RESULT_OK;
    }
    NphResult r = read_contacts(w);
    if (r == NPH_RESULT_OK) {
        r = world_sync_bodies(w);
    }
    if (r != NPH_RESULT_OK) {
        return r;
    }
    const NphContactRec *recs = VB(w->contact_view, NphContactRec);
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        const u32 *list = VB(w->color_contacts[c], u32);
        for (u32 i = 0U; i < w->color_contacts[c].len; ++i) {
            const HContact *contact = &VB(w->contacts, HContact)[list[i]];
            if ((contact->flags & HC_HIT_EVENTS) == 0U) {
                continue;
            }
            f32 best = w->hit_threshold;
            u32 hit = NPH_NULL32;
            u32 best_point = NPH_NULL32;
            for (u32 k = list[i]; k != NPH_NULL32; k = VB(w->contacts, HContact)[k].child) {
                const NphContactRec *record = &recs[k];
                for (u32 j = 0U; j < record->point_count; ++j) {
                    const NphPointRec *p = &record->points[j];
                    const f32 approach = -p->relative_velocity;
                    if (approach > best && p->max_normal_impulse > 0.0F) {
                        best = approach;
                        hit = k;
                        best_point = j;
                    }
                }
            }
            if (hit == NPH_NULL32) {
                continue;
            }
            NphPhysicsContactHitEvent e;
            memset(&e, 0, sizeof e);
            e.shape_a.value = VB(w->shapes, HShape)[contact->shape_a].handle;
            e.shape_b.value = VB(w->shapes, HShape)[contact->shape_b].handle;
            e.point = dv3_add_v3(rec_at(w, contact->body_a)->center,
                                 recs[hit].points[best_point].anchor_a);
            e.normal = recs[hit].normal;
            e.approach_speed = best;
            r = vbuf_push(w->alloc, &w->hit_events, sizeof e, &e);
            if (r != NPH_RESULT_OK) {
                return r;
            }
        }
    }
    return NPH_RESULT_OK;
}

NphResult world_solve_input(NphPhysicsWorld *w, f32 dt, NphSolveInput *out)
{
    NphSolveInput in;
    memset(&in, 0, sizeof in);
    const u32 n = w->awake_bodies.len;
    w->stats.swept_bodies = 0U;
    w->stats.impact_bodies = 0U;
    if (n == 0U) {
        *out = in;
        return NPH_RESULT_OK;
    }
    in.awake_bodies = VB(w->awake_bodies, u32);
    in.awake_count = n;
    in.body_slots = w->bodies.len;
    u32 graph = 0U;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        graph += w->color_contacts[c].len;
    }
    NphResult r = vbuf_reserve(w->alloc, &w->graph, 4U, graph > 0U ? graph : 1U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    u32 at = 0U;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        in.color_count[c] = w->color_contacts[c].len;
        memcpy(VB(w->graph, u32) + at, w->color_contacts[c].data, (size_t)in.color_count[c] * 4U);
        at += in.color_count[c];
    }
    in.graph_contacts = VB(w->graph, u32);
    w->solved_contacts = graph;
    // The joints, colour by colour, the overflow last (solver.c:791-806).
    u32 joints = 0U;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        joints += w->color_joints[c].len;
    }
    r = vbuf_reserve(w->alloc, &w->step_joints, 4U, joints > 0U ? joints : 1U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    at = 0U;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        in.joint_color_count[c] = w->color_joints[c].len;
        memcpy(VB(w->step_joints, u32) + at, w->color_joints[c].data,
               (size_t)in.joint_color_count[c] * 4U);
        at += in.joint_color_count[c];
    }
    in.graph_joints = VB(w->step_joints, u32);
    // The colour passes' joints and their dynamic ends, for measurement.
    w->solved_joints = 0U;
    w->solved_joint_ends = 0U;
    for (u32 c = 0U; c < NPH_GRAPH_COLORS; ++c) {
        w->solved_joints += w->color_joints[c].len;
        for (u32 i = 0U; i < w->color_joints[c].len; ++i) {
            const HJoint *j = &VB(w->joints, HJoint)[VB(w->color_joints[c], u32)[i]];
            w->solved_joint_ends +=
                (VB(w->bodies, HBody)[j->body_a].type == (u32)NPH_PHYSICS_BODY_DYNAMIC ? 1U : 0U) +
                (VB(w->bodies, HBody)[j->body_b].type == (u32)NPH_PHYSICS_BODY_DYNAMIC ? 1U : 0U);
        }
    }
    NphStepParams *sp = &in.params;
    sp->gravity = w->desc.gravity;
    sp->dt = dt;
    sp->h = dt / (f32)w->desc.substeps;
    sp->inv_h = 1.0F / sp->h;
    sp->push_max = w->desc.contact_push_max_velocity;
    sp->restitution_threshold = w->desc.restitution_threshold;
    sp->max_linear_speed = w->desc.max_linear_speed;
    sp->substeps = w->desc.substeps;
    const f32 contact_hertz =
        w->desc.contact_hertz > 0.0F ? w->desc.contact_hertz : 0.125F * sp->inv_h;
    const f32 hertz = nph_minf(contact_hertz, 0.25F * sp->inv_h);
    make_soft(hertz, w->desc.contact_damping_ratio, sp->h, &sp->contact_bias, &sp->contact_mass,
              &sp->contact_impulse);
    make_soft(2.0F * hertz, w->desc.contact_damping_ratio, sp->h, &sp->static_bias,
              &sp->static_mass, &sp->static_impulse);
    // Joints at most a quarter of the substep rate stiff (solver.c:998-1005).
    const f32 joint_hertz = w->desc.joint_hertz > 0.0F ? w->desc.joint_hertz : 0.25F * sp->inv_h;
    make_soft(nph_minf(joint_hertz, 0.25F * sp->inv_h), w->desc.joint_damping_ratio, sp->h,
              &sp->joint_bias, &sp->joint_mass, &sp->joint_impulse);
    w->last_substep = sp->h;
    sp->continuous = w->desc.enable_continuous != 0 ? 1U : 0U;
    *out = in;
    return NPH_RESULT_OK;
}

static int joint_is_driven(const HJoint *j)
{
    // joint_solver.c:1410-1433, for the joint types built here.
    return j->type == (u32)NPH_PHYSICS_JOINT_HINGE && j->s.hinge.enable_motor != 0 &&
           j->s.hinge.motor_speed != 0.0F && j->s.hinge.max_motor_torque > 0.0F;
}

static NphResult wake_unless_awake(NphPhysicsWorld *w, u32 body)
{
    const HBody *b = &VB(w->bodies, HBody)[body];
    return (b->flags & HB_AWAKE) == 0U ? world_wake_body(w, body) : NPH_RESULT_OK;
}

// solver.c:845-910, after the device's finalize, in the awake list's order:
// a motor driving towards a speed keeps its bodies' sleep clocks at zero;
// the move events cover the awake list as it stands; a kinematic body that
// moved wakes the sleepers it touches or is jointed to. The host copy of
// the records is fetched only when some joint is driven or some kinematic
// body is awake.
static NphResult post_finalize(NphPhysicsWorld *w)
{
    w->move_pending = w->awake_bodies.len;
    w->move_events_ready = 0;
    w->move_events.len = 0U;
    int driven = 0;
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS && !driven; ++c) {
        for (u32 i = 0U; i < w->color_joints[c].len && !driven; ++i) {
            driven = joint_is_driven(&VB(w->joints, HJoint)[VB(w->color_joints[c], u32)[i]]);
        }
    }
    int kinematic = 0;
    for (u32 i = 0U; w->kinematic_bodies != 0U && i < w->awake_bodies.len && !kinematic; ++i) {
        kinematic = VB(w->bodies, HBody)[VB(w->awake_bodies, u32)[i]].type ==
                    NPH_PHYSICS_BODY_KINEMATIC;
    }
    if (!driven && !kinematic) {
        return NPH_RESULT_OK;
    }
    NphResult r = world_sync_bodies(w); // the move events too
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS && driven && r == NPH_RESULT_OK; ++c) {
        for (u32 i = 0U; i < w->color_joints[c].len && r == NPH_RESULT_OK; ++i) {
            const HJoint *j = &VB(w->joints, HJoint)[VB(w->color_joints[c], u32)[i]];
            if (joint_is_driven(j)) {
                rec_at(w, j->body_a)->sleep_time = 0.0F;
                rec_at(w, j->body_b)->sleep_time = 0.0F;
                r = world_upload_body(w, j->body_a);
                if (r == NPH_RESULT_OK) {
                    r = world_upload_body(w, j->body_b);
                }
            }
        }
    }
    for (u32 i = 0U; i < w->awake_bodies.len && kinematic && r == NPH_RESULT_OK; ++i) {
        const u32 slot = VB(w->awake_bodies, u32)[i];
        const HBody *b = &VB(w->bodies, HBody)[slot];
        if (b->type != NPH_PHYSICS_BODY_KINEMATIC || rec_at(w, slot)->sleep_time > 0.0F) {
            continue;
        }
        for (u32 key = b->contact_head; key != NPH_NULL32 && r == NPH_RESULT_OK;) {
            const HContact *c = &VB(w->contacts, HContact)[key >> 1U];
            const u32 other = (key & 1U) == 0U ? c->body_b : c->body_a;
            key = c->next[key & 1U];
            if ((c->flags & HC_TOUCHING) != 0U) {
                r = wake_unless_awake(w, other);
            }
        }
        for (u32 key = b->joint_head; key != NPH_NULL32 && r == NPH_RESULT_OK;) {
            const HJoint *j = &VB(w->joints, HJoint)[key >> 1U];
            const u32 other = (key & 1U) == 0U ? j->body_b : j->body_a;
            key = j->next[key & 1U];
            r = wake_unless_awake(w, other);
        }
    }
    return r;
}

NphResult world_solve_post(NphPhysicsWorld *w, const NphGrown *grown, u32 grown_count)
{
    w->bodies_stale = 1;
    NphResult r = contacts_post_solve(w, grown, grown_count);
    if (r == NPH_RESULT_OK) {
        r = post_finalize(w);
    }
    return r == NPH_RESULT_OK ? hit_events(w) : r;
}

NphResult world_step_end(NphPhysicsWorld *w)
{
    // Islands merge whether or not they may sleep (world.c:351-359).
    NphResult r = islands_merge_pending(w);
    if (r == NPH_RESULT_OK && w->desc.enable_sleep != 0) {
        r = islands_sleep_and_split(w);
    }
    return r;
}

static NphResult solve(NphPhysicsWorld *w, f32 dt)
{
    NphSolveInput in;
    NphResult r = world_solve_input(w, dt, &in);
    if (r != NPH_RESULT_OK || in.awake_count == 0U) {
        return r;
    }
    const NphFastBody *fast = NULL;
    u32 fast_count = 0U;
    r = nph_dev_solve(w->dev, &in, &fast, &fast_count);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    w->bodies_stale = 1;
    if (fast_count > 0U) {
        r = continuous_sweep(w, fast, fast_count);
        if (r != NPH_RESULT_OK) {
            return r;
        }
    }
    const NphGrown *grown = NULL;
    u32 grown_count = 0U;
    r = nph_dev_solve_finish(w->dev, &grown, &grown_count);
    w->bodies_stale = 1; // the sweeps may have read the records mid-step
    if (r != NPH_RESULT_OK) {
        return r;
    }
    return world_solve_post(w, grown, grown_count);
}

void world_step_begin(NphPhysicsWorld *w, f32 dt)
{
    w->shape_view_ok = 0; // the step moves the tight boxes
    w->stepping = 1;
    w->step_dt = dt;
    w->begin_events.len = 0U;
    w->hit_events.len = 0U;
    w->end_events.len = 0U;
    w->move_events.len = 0U;
    w->move_events_ready = 1; // none until the solve's end
    w->move_pending = 0U;
}

f64 world_now_ms(void) { return now_ms(); }

NphResult nph_physics_world_step(NphPhysicsWorld *w, f32 dt, NphJobSystem *jobs)
{
    (void)jobs; // the device does the parallel work; results never depend on it
    if (!(dt > 0.0F) || !nph_isfinitef(dt)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const f64 t0 = now_ms();
    NphResult r = NPH_RESULT_OK;
    world_step_begin(w, dt);
    // The existing contacts' narrow phase runs on the device while the host
    // finds new pairs; the new contacts' follows (world.c:330-345 order of
    // results).
    f64 t = now_ms();
    r = contacts_collide_begin(w);
    f64 collide_ms = now_ms() - t;
    const int queued = r == NPH_RESULT_OK;
    if (r == NPH_RESULT_OK) {
        t = now_ms();
        r = contacts_update_pairs(w);
        w->stats.pairs_ms = now_ms() - t;
    }
    if (r == NPH_RESULT_OK) {
        r = contacts_reserve_changes(w);
    }
    if (r != NPH_RESULT_OK && queued) {
        // Abandoned before the queued narrow phase's changes reached the
        // host (world.h:161-166: consistent, not advanced): its records go
        // back to how the step found them, so the next step finds the same
        // changes. Contacts the pair search made keep their staged inits.
        (void)nph_dev_contacts_restore(w->dev);
    }
    if (r == NPH_RESULT_OK) {
        t = now_ms();
        r = contacts_collide_end(w);
        w->stats.collide_ms = collide_ms + now_ms() - t;
    }
    if (r == NPH_RESULT_OK) {
        t = now_ms();
        r = solve(w, dt);
        w->stats.solve_ms = now_ms() - t;
    }
    if (r == NPH_RESULT_OK) {
        t = now_ms();
        r = world_step_end(w);
        w->stats.sleep_ms = now_ms() - t;
    }
    w->stepping = 0;
    w->stats.steps += 1U;
    w->stats.step_ms = now_ms() - t0;
    return r;
}

// ---------------------------------------------------------------------------
// Readback: events, stats, hash, contacts
// ---------------------------------------------------------------------------
NphPhysicsEvents nph_physics_world_events(const NphPhysicsWorld *cw)
{
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    NphPhysicsEvents e;
    memset(&e, 0, sizeof e);
    if (!w->move_events_ready) {
        (void)world_sync_bodies(w); // builds them
    }
    e.begin = VB(w->begin_events, NphPhysicsContactBeginEvent);
    e.begin_count = w->begin_events.len;
    e.end = VB(w->end_events, NphPhysicsContactEndEvent);
    e.end_count = w->end_events.len;
    e.hit = VB(w->hit_events, NphPhysicsContactHitEvent);
    e.hit_count = w->hit_events.len;
    e.move = VB(w->move_events, NphPhysicsBodyMoveEvent);
    e.move_count = w->move_events.len;
    return e;
}

static NphResult read_contacts(NphPhysicsWorld *w)
{
    NphResult r = vbuf_reserve(w->alloc, &w->contact_view, sizeof(NphContactRec),
                               w->contacts.len > 0U ? w->contacts.len : 1U);
    if (r == NPH_RESULT_OK) {
        r = nph_dev_read_contacts(w->dev, VB(w->contact_view, NphContactRec), w->contacts.len);
    }
    return r;
}

NphPhysicsStats nph_physics_world_stats(const NphPhysicsWorld *cw)
{
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    NphPhysicsStats s = w->stats;
    s.bodies = w->body_slots.generation.len - w->body_slots.free.len;
    s.awake_bodies = w->awake_bodies.len;
    s.shapes = w->shape_slots.generation.len - w->shape_slots.free.len;
    s.contacts = w->contacts.len - w->free_contacts.len;
    u32 touching = 0U;
    u32 points = 0U;
    if (read_contacts(w) == NPH_RESULT_OK) {
        for (u32 i = 0U; i < w->awake_contacts.len; ++i) {
            const u32 k = VB(w->awake_contacts, u32)[i];
            if ((VB(w->contacts, HContact)[k].flags & HC_TOUCHING) != 0U) {
                touching += 1U;
                points += VB(w->contact_view, NphContactRec)[k].point_count;
            }
        }
    }
    s.touching_contacts = touching;
    s.manifold_points = points;
    s.joints = w->joint_count;
    s.islands = w->islands.len - w->free_islands.len;
    s.awake_islands = w->awake_islands.len;
    u32 colors = 0U;
    for (u32 c = 0U; c < NPH_GRAPH_COLORS; ++c) {
        colors += w->color_contacts[c].len > 0U ? 1U : 0U;
    }
    s.colors_used = colors;
    s.overflow_constraints = w->color_contacts[NPH_GRAPH_OVERFLOW].len;
    s.workers = 1U;
    return s;
}

// nph_hash64 (engine/core/src/hash.c:73-120, the wyhash construction): the
// state fingerprint's hash, restated for the records below.
static void mul128(u64 *a, u64 *b)
{
    const unsigned __int128 r = (unsigned __int128)*a * (unsigned __int128)*b;
    *a = (u64)r;
    *b = (u64)(r >> 64U);
}

static u64 hmix(u64 a, u64 b)
{
    mul128(&a, &b);
    return a ^ b;
}

static u64 rd8(const u8 *p)
{
    u64 v;
    memcpy(&v, p, 8U);
    return v;
}

static u64 rd4(const u8 *p)
{
    u32 v;
    memcpy(&v, p, 4U);
    return v;
}

static u64 nph_hash64(const void *data, u64 len, u64 seed)
{
    const u64 s0 = 0x9E3779B97F4A7C15ULL;
    const u64 s1 = 0x6A09E667F3BCC909ULL;
    const u64 s2 = 0xBB67AE8584CAA73BULL;
    const u64 s3 = 0x3C6EF372FE94F82BULL;
    const u8 *p = (const u8 *)data;
    seed ^= hmix(seed ^ s0, s1);
    u64 a;
    u64 b;
    if (len <= 16U) {
        if (len >= 4U) {
            a = (rd4(p) << 32U) | rd4(p + ((len >> 3U) << 2U));
            b = (rd4(p + len - 4U) << 32U) | rd4(p + len - 4U - ((len >> 3U) << 2U));
        } else if (len > 0U) {
            a = ((u64)p[0] << 16U) | ((u64)p[len >> 1U] << 8U) | (u64)p[len - 1U];
            b = 0U;
        } else {
            a = 0U;
            b = 0U;
        }
    } else {
        u64 i = len;
        if (i > 48U) {
            u64 see1 = seed;
            u64 see2 = seed;
            do {
                seed = hmix(rd8(p) ^ s1, rd8(p + 8) ^ seed);
                see1 = hmix(rd8(p + 16) ^ s2, rd8(p + 24) ^ see1);
                see2 = hmix(rd8(p + 32) ^ s3, rd8(p + 40) ^ see2);
                p += 48;
                i -= 48U;
            } while (i > 48U);
            seed ^= see1 ^ see2;
        }
        while (i > 16U) {
            seed = hmix(rd8(p) ^ s1, rd8(p + 8) ^ seed);
            i -= 16U;
            p += 16;
        }
        a = rd8(p + i - 16U);
        b = rd8(p + i - 8U);
    }
    a ^= s1;
    b ^= seed;
    mul128(&a, &b);
    return hmix(a ^ s0 ^ len, b ^ s1);
}

// world.c:436-462
u64 nph_physics_world_hash(const NphPhysicsWorld *cw)
{
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    if (world_sync_bodies(w) != NPH_RESULT_OK) {
        return 0U;
    }
    u64 h = 0x6E70687973696373ULL;
    for (u32 i = 0U; i < w->bodies.len; ++i) {
        const HBody *hb = &VB(w->bodies, HBody)[i];
        if (hb->handle == 0U) {
            continue;
        }
        const NphBodyRec *b = rec_at(w, i);
        struct {
            u32 handle;
            u32 flags;
            NphDVec3 origin;
            NphQuat rotation;
            NphVec3 v;
            NphVec3 w;
        } rec;
        memset(&rec, 0, sizeof rec);
        rec.handle = hb->handle;
        rec.flags = hb->flags & HB_AWAKE;
        rec.origin = b->origin;
        rec.rotation = b->rotation;
        rec.v = b->v;
        rec.w = b->w;
        h = nph_hash64(&rec, sizeof rec, h);
    }
    return h;
}

void nph_amd_world_energy(const NphPhysicsWorld *cw, f64 out[3])
{
    // Mechanical energy of the dynamic bodies, in f64: linear kinetic
    // 1/2 m |v|^2, angular kinetic 1/2 w_b . I w_b with w_b the angular
    // velocity in body axes, and potential -m s (g . c) with s the body's
    // gravity scale and c its centre of mass.
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    out[0] = out[1] = out[2] = 0.0;
    if (world_sync_bodies(w) != NPH_RESULT_OK) {
        return;
    }
    const NphVec3 g = w->desc.gravity;
    for (u32 i = 0U; i < w->bodies.len; ++i) {
        const HBody *hb = &VB(w->bodies, HBody)[i];
        const NphBodyRec *b = rec_at(w, i);
        if (hb->handle == 0U || b->type != (u32)NPH_PHYSICS_BODY_DYNAMIC || !(b->mass > 0.0F)) {
            continue;
        }
        const f64 m = b->mass;
        const f64 v[3] = {b->v.x, b->v.y, b->v.z};
        out[0] += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        // Body-axis angular velocity: the rotation's inverse applied to w.
        const f64 qx = -b->rotation.x, qy = -b->rotation.y, qz = -b->rotation.z,
                  qw = b->rotation.w;
        const f64 wx = b->w.x, wy = b->w.y, wz = b->w.z;
        const f64 tx = 2.0 * (qy * wz - qz * wy), ty = 2.0 * (qz * wx - qx * wz),
                  tz = 2.0 * (qx * wy - qy * wx);
        const f64 wb[3] = {wx + qw * tx + (qy * tz - qz * ty), wy + qw * ty + (qz * tx - qx * tz),
                           wz + qw * tz + (qx * ty - qy * tx)};
        const NphMat3 *I = &b->inertia;
        for (u32 r = 0U; r < 3U; ++r) {
            const f64 row = (f64)(&I->c[0].x)[r] * wb[0] + (f64)(&I->c[1].x)[r] * wb[1] +
                            (f64)(&I->c[2].x)[r] * wb[2];
            out[1] += 0.5 * wb[r] * row;
        }
        out[2] -= m * (f64)b->gravity_scale *
                  ((f64)g.x * b->center.x + (f64)g.y * b->center.y + (f64)g.z * b->center.z);
    }
}

bool nph_physics_world_validate(const NphPhysicsWorld *w)
{
    // The bookkeeping's invariants (world.c:482-695, the parts kept here):
    // awake slots point back, colour slots point back, pairs are in the map.
    for (u32 i = 0U; i < w->awake_contacts.len; ++i) {
        const u32 k = VB(w->awake_contacts, u32)[i];
        if (k >= w->contacts.len || VB(w->contacts, HContact)[k].awake_slot != i) {
            return false;
        }
    }
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        for (u32 i = 0U; i < w->color_contacts[c].len; ++i) {
            const u32 k = VB(w->color_contacts[c], u32)[i];
            const HContact *hc = &VB(w->contacts, HContact)[k];
            if (hc->color != c || hc->color_slot != i || (hc->flags & HC_TOUCHING) == 0U) {
                return false;
            }
        }
    }
    for (u32 i = 0U; i < w->awake_bodies.len; ++i) {
        const u32 s = VB(w->awake_bodies, u32)[i];
        if (VB(w->bodies, HBody)[s].awake_slot != i) {
            return false;
        }
    }
    // Islands: dynamic bodies only; linked contacts touch between two
    // dynamic bodies, linked joints join two; every list agrees with its
    // members and its count; merges are done by the step's end.
    const HBody *bodies = VB(w->bodies, HBody);
    for (u32 i = 0U; i < w->bodies.len; ++i) {
        const HBody *b = &bodies[i];
        if (b->handle != 0U && ((b->type == NPH_PHYSICS_BODY_DYNAMIC) != (b->island != NPH_NULL32) ||
                                ((b->flags & HB_AWAKE) != 0U) != (b->awake_slot != NPH_NULL32))) {
            return false;
        }
    }
    for (u32 i = 0U; i < w->contacts.len; ++i) {
        const HContact *c = &VB(w->contacts, HContact)[i];
        if (c->shape_a == NPH_NULL32 || c->parent != NPH_NULL32) {
            continue;
        }
        const int linked = (c->flags & HC_TOUCHING) != 0U &&
                           bodies[c->body_a].type == NPH_PHYSICS_BODY_DYNAMIC &&
                           bodies[c->body_b].type == NPH_PHYSICS_BODY_DYNAMIC;
        if (linked != ((c->flags & HC_ISLAND) != 0U) || linked != (c->island != NPH_NULL32)) {
            return false;
        }
    }
    for (u32 i = 0U; i < w->joints.len; ++i) {
        const HJoint *j = &VB(w->joints, HJoint)[i];
        if (j->handle == 0U) {
            continue;
        }
        const int linked = bodies[j->body_a].type == NPH_PHYSICS_BODY_DYNAMIC &&
                           bodies[j->body_b].type == NPH_PHYSICS_BODY_DYNAMIC;
        if (linked != ((j->flags & HJ_ISLAND) != 0U) || linked != (j->island != NPH_NULL32)) {
            return false;
        }
    }
    for (u32 i = 0U; i < w->islands.len; ++i) {
        const HIsland *is = &VB(w->islands, HIsland)[i];
        if (is->live == 0U) {
            continue;
        }
        u32 n = 0U;
        for (u32 b = is->body_head; b != NPH_NULL32; b = bodies[b].island_next, ++n) {
            if (bodies[b].island != i) {
                return false;
            }
        }
        u32 m = 0U;
        for (u32 c = is->contact_head; c != NPH_NULL32;
             c = VB(w->contacts, HContact)[c].island_next, ++m) {
            if (VB(w->contacts, HContact)[c].island != i) {
                return false;
            }
        }
        u32 nj = 0U;
        for (u32 j = is->joint_head; j != NPH_NULL32; j = VB(w->joints, HJoint)[j].island_next, ++nj) {
            if (VB(w->joints, HJoint)[j].island != i) {
                return false;
            }
        }
        if (n != is->body_count || n == 0U || m != is->contact_count || nj != is->joint_count ||
            is->parent != NPH_NULL32 || (is->awake != 0U) != (is->awake_slot != NPH_NULL32)) {
            return false;
        }
    }
    return true;
}

u32 nph_physics_world_contacts(const NphPhysicsWorld *cw, NphPhysicsContactInfo *out, u32 capacity)
{
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    if (read_contacts(w) != NPH_RESULT_OK) {
        return 0U;
    }
    u32 n = 0U;
    for (u32 i = 0U; i < w->contacts.len; ++i) {
        const HContact *hc = &VB(w->contacts, HContact)[i];
        if (hc->shape_a == NPH_NULL32) {
            continue;
        }
        if (n < capacity && out != NULL) {
            const NphContactRec *c = &VB(w->contact_view, NphContactRec)[i];
            NphPhysicsContactInfo *o = &out[n];
            memset(o, 0, sizeof *o);
            o->record = i;
            o->shape_a = hc->shape_a;
            o->shape_b = hc->shape_b;
            o->body_a = hc->body_a;
            o->body_b = hc->body_b;
            o->flags = hc->flags;
            o->color = hc->color;
            o->color_slot = hc->color_slot;
            o->parent = hc->parent;
            o->child = hc->child;
            o->friction = c->friction;
            o->restitution = c->restitution;
            o->normal = c->normal;
            o->point_count = c->point_count;
            for (u32 k = 0U; k < c->point_count; ++k) {
                memcpy(&o->points[k], &c->points[k], sizeof(NphPhysicsContactPointInfo));
            }
        }
        n += 1U;
    }
    return n;
}
