// SPDX-License-Identifier: Apache-2.0
// The rest of the bodies' and shapes' API on the host: destruction,
// teleports, forces and impulses, mass and motion settings (body.c:132-620),
// and the shapes' settings and destruction (shape.c:29-440).  Each call edits
// the host's records and stages the changed device records; contacts and
// joints go through the same bookkeeping as the step's.
#define _POSIX_C_SOURCE 200809L
#include "world_impl.h"

static HBody *hbody(NphPhysicsWorld *w, u32 slot) { return &VB(w->bodies, HBody)[slot]; }
static HShape *hshape(NphPhysicsWorld *w, u32 slot) { return &VB(w->shapes, HShape)[slot]; }
static int finite_nonneg(f32 v) { return nph_isfinitef(v) && v >= 0.0F; }

bool nph_physics_body_is_null(NphPhysicsBody b) { return b.value == 0U; }
bool nph_physics_shape_is_null(NphPhysicsShape s) { return s.value == 0U; }

// The body's record, current on the host, for a change; NULL for a stale
// handle.
static NphBodyRec *editable(NphPhysicsWorld *w, NphPhysicsBody body, HBody **out)
{
    HBody *hb = world_body(w, body);
    if (hb == NULL || world_sync_bodies(w) != NPH_RESULT_OK) {
        return NULL;
    }
    if (out != NULL) {
        *out = hb;
    }
    return world_rec(w, phys_slot(body.value));
}

static NphBodyRec *dynamic_rec(NphPhysicsWorld *w, NphPhysicsBody body)
{
    HBody *hb = NULL;
    NphBodyRec *b = editable(w, body, &hb);
    return b != NULL && hb->type == NPH_PHYSICS_BODY_DYNAMIC ? b : NULL;
}

static void wake_and_upload(NphPhysicsWorld *w, NphPhysicsBody body)
{
    const u32 slot = phys_slot(body.value);
    (void)world_wake_body(w, slot);
    (void)world_upload_body(w, slot);
}

// shape.c:148-166: the proxy out of its tree, and out of the move buffer
// with the buffer's order kept.
static void destroy_proxy(NphPhysicsWorld *w, u32 s)
{
    HShape *shape = hshape(w, s);
    if (shape->proxy == NPH_NULL32) {
        return;
    }
    if ((shape->flags & HS_IN_MOVE_BUFFER) != 0U) {
        const u32 key = (shape->proxy << 2U) | shape->tree;
        u32 *buf = VB(w->move_buffer, u32);
        for (u32 i = 0U; i < w->move_buffer.len; ++i) {
            if (buf[i] == key) {
                memmove(&buf[i], &buf[i + 1U], (size_t)(w->move_buffer.len - i - 1U) * 4U);
                w->move_buffer.len -= 1U;
                break;
            }
        }
        shape->flags &= ~(u32)HS_IN_MOVE_BUFFER;
    }
    tree_destroy_proxy(&w->trees[shape->tree], shape->proxy);
    shape->proxy = NPH_NULL32;
}

// shape.c:265-283: the shape's contacts end without events; the bodies
// that touched it wake.
static void destroy_shape_contacts(NphPhysicsWorld *w, u32 s)
{
    const u32 body = hshape(w, s)->body;
    u32 key = hbody(w, body)->contact_head;
    while (key != NPH_NULL32) {
        const u32 c = key >> 1U;
        const HContact *contact = contacts_at(w, c);
        key = contact->next[key & 1U];
        if (contact->shape_a == s || contact->shape_b == s) {
            const u32 other = contact->body_a == body ? contact->body_b : contact->body_a;
            const int touching = (contact->flags & HC_TOUCHING) != 0U;
            contacts_destroy(w, c, 0);
            if (touching) {
                (void)world_wake_body(w, other);
            }
        }
    }
}

// body.c:132-180
void nph_physics_body_destroy(NphPhysicsWorld *w, NphPhysicsBody body)
{
    HBody *b = NULL;
    if (w == NULL || w->stepping || editable(w, body, &b) == NULL) {
        return;
    }
    const u32 slot = phys_slot(body.value);
    while (b->joint_head != NPH_NULL32) {
        const u32 j = b->joint_head >> 1U;
        const HJoint *joint = &VB(w->joints, HJoint)[j];
        const u32 other = joint->body_a == slot ? joint->body_b : joint->body_a;
        joints_destroy(w, j, 0);
        (void)world_wake_body(w, other);
    }
    while (b->contact_head != NPH_NULL32) {
        const u32 c = b->contact_head >> 1U;
        const HContact *contact = contacts_at(w, c);
        const u32 other = contact->body_a == slot ? contact->body_b : contact->body_a;
        const int touching = (contact->flags & HC_TOUCHING) != 0U;
        contacts_destroy(w, c, 0);
        if (touching) {
            (void)world_wake_body(w, other);
        }
    }
    while (b->shape_head != NPH_NULL32) {
        const u32 s = b->shape_head;
        HShape *shape = hshape(w, s);
        b->shape_head = shape->next;
        destroy_proxy(w, s);
        shape->handle = 0U;
        w->hit_event_shapes -= (shape->flags & HS_HIT_EVENTS) != 0U ? 1U : 0U;
        slots_release(w->alloc, &w->shape_slots, s);
    }
    b->shape_count = 0U;
    world_rec(w, slot)->shape_head = NPH_NULL32;
    if (b->awake_slot != NPH_NULL32) {
        (void)world_set_awake_list(w, slot, 0);
    }
    islands_remove_body(w, slot);
    b = hbody(w, slot);
    b->handle = 0U;
    w->kinematic_bodies -= b->type == NPH_PHYSICS_BODY_KINEMATIC ? 1U : 0U;
    slots_release(w->alloc, &w->body_slots, slot);
    (void)world_upload_body(w, slot);
}

// body.c:219-262
void nph_physics_body_set_transform(NphPhysicsWorld *w, NphPhysicsBody body, NphDVec3 position,
                                    NphQuat rotation)
{
    HBody *hb = NULL;
    NphBodyRec *b = w != NULL && !w->stepping ? editable(w, body, &hb) : NULL;
    if (b == NULL) {
        return;
    }
    const u32 slot = phys_slot(body.value);
    b->origin = position;
    b->rotation = q_normalize(rotation);
    b->center = dv3_add_v3(b->origin, q_rotate(b->rotation, b->local_center));
    for (u32 s = hb->shape_head; s != NPH_NULL32; s = hshape(w, s)->next) {
        HShape *shape = hshape(w, s);
        NphShapeRec tmp;
        if (world_shape_record(w, s, &tmp) != NPH_RESULT_OK) {
            return;
        }
        tmp.geom = world_host_geom(shape); // the host's block address for the box
        shape->aabb = shape_world_aabb(b, &tmp);
        const f32 margin = hb->type == NPH_PHYSICS_BODY_STATIC ? NPH_PHYSICS_SPECULATIVE_DISTANCE
                                                               : NPH_PHYSICS_AABB_MARGIN;
        shape->fat_aabb = aabb_grow(shape->aabb, margin);
        tree_move_proxy(&w->trees[shape->tree], shape->proxy, shape->fat_aabb);
        (void)move_buffer_push(w, s);
        (void)world_refresh_shape(w, s);
    }
    if (hb->type == NPH_PHYSICS_BODY_STATIC) {
        for (u32 key = hb->contact_head; key != NPH_NULL32;) {
            const HContact *c = contacts_at(w, key >> 1U);
            const u32 other = (key & 1U) == 0U ? c->body_b : c->body_a;
            key = c->next[key & 1U];
            (void)world_wake_body(w, other);
        }
    } else {
        (void)world_wake_body(w, slot);
    }
    (void)world_upload_body(w, slot);
}

// body.c:316-386
void nph_physics_body_apply_force(NphPhysicsWorld *w, NphPhysicsBody body, NphVec3 force,
                                  NphDVec3 point)
{
    NphBodyRec *b = dynamic_rec(w, body);
    if (b == NULL) {
        return;
    }
    b->force = v3_add(b->force, force);
    b->torque = v3_add(b->torque, v3_cross(dv3_relative(point, b->center), force));
    wake_and_upload(w, body);
}

void nph_physics_body_apply_force_at_center(NphPhysicsWorld *w, NphPhysicsBody body, NphVec3 force)
{
    NphBodyRec *b = dynamic_rec(w, body);
    if (b == NULL) {
        return;
    }
    b->force = v3_add(b->force, force);
    wake_and_upload(w, body);
}

void nph_physics_body_apply_torque(NphPhysicsWorld *w, NphPhysicsBody body, NphVec3 torque)
{
    NphBodyRec *b = dynamic_rec(w, body);
    if (b == NULL) {
        return;
    }
    b->torque = v3_add(b->torque, torque);
    wake_and_upload(w, body);
}

// R I^-1 R^T
static NphMat3 world_inv_inertia(const NphBodyRec *b)
{
    const NphMat3 r = m3_from_quat(b->rotation);
    return m3_mul(m3_mul(r, b->inv_inertia), m3_transpose(r));
}

void nph_physics_body_apply_linear_impulse(NphPhysicsWorld *w, NphPhysicsBody body,
                                           NphVec3 impulse, NphDVec3 point)
{
    NphBodyRec *b = dynamic_rec(w, body);
    if (b == NULL) {
        return;
    }
    b->v = v3_mul_add(b->v, impulse, b->inv_mass);
    const NphVec3 r = dv3_relative(point, b->center);
    b->w = v3_add(b->w, m3_mul_v3(world_inv_inertia(b), v3_cross(r, impulse)));
    wake_and_upload(w, body);
}

void nph_physics_body_apply_angular_impulse(NphPhysicsWorld *w, NphPhysicsBody body,
                                            NphVec3 impulse)
{
    NphBodyRec *b = dynamic_rec(w, body);
    if (b == NULL) {
        return;
    }
    b->w = v3_add(b->w, m3_mul_v3(world_inv_inertia(b), impulse));
    wake_and_upload(w, body);
}

// body.c:463-481
void nph_physics_body_set_mass_data(NphPhysicsWorld *w, NphPhysicsBody body,
                                    const NphPhysicsMassData *mass)
{
    HBody *hb = NULL;
    NphBodyRec *b = w != NULL && mass != NULL ? editable(w, body, &hb) : NULL;
    if (b == NULL) {
        return;
    }
    if (!(mass->mass > 0.0F)) {
        hb->flags &= ~(u32)HB_MASS_SET;
    } else {
        hb->flags |= HB_MASS_SET;
        b->mass = mass->mass;
        b->local_center = mass->center;
        b->inertia = mass->inertia;
    }
    const u32 slot = phys_slot(body.value);
    world_update_mass(w, slot);
    (void)world_upload_body(w, slot);
}

bool nph_physics_body_is_bullet(const NphPhysicsWorld *w, NphPhysicsBody body)
{
    const HBody *b = world_body(w, body);
    return b != NULL && (b->flags & HB_BULLET) != 0U;
}

void nph_physics_body_set_bullet(NphPhysicsWorld *w, NphPhysicsBody body, bool bullet)
{
    HBody *hb = NULL;
    NphBodyRec *b = editable(w, body, &hb);
    if (b == NULL || hb->type != NPH_PHYSICS_BODY_DYNAMIC ||
        ((hb->flags & HB_BULLET) != 0U) == bullet) {
        return;
    }
    if (bullet) {
        hb->flags |= HB_BULLET;
        b->flags |= HB_BULLET;
    } else {
        hb->flags &= ~(u32)HB_BULLET;
        b->flags &= ~(u32)HB_BULLET;
    }
    (void)world_upload_body(w, phys_slot(body.value));
}

void nph_physics_body_set_damping(NphPhysicsWorld *w, NphPhysicsBody body, f32 linear, f32 angular)
{
    NphBodyRec *b = editable(w, body, NULL);
    if (b == NULL || !finite_nonneg(linear) || !finite_nonneg(angular)) {
        return;
    }
    b->linear_damping = linear;
    b->angular_damping = angular;
    (void)world_upload_body(w, phys_slot(body.value));
}

void nph_physics_body_set_gravity_scale(NphPhysicsWorld *w, NphPhysicsBody body, f32 scale)
{
    HBody *hb = NULL;
    NphBodyRec *b = editable(w, body, &hb);
    if (b == NULL || !nph_isfinitef(scale) || b->gravity_scale == scale) {
        return;
    }
    b->gravity_scale = scale;
    if (hb->type == NPH_PHYSICS_BODY_DYNAMIC) {
        (void)world_wake_body(w, phys_slot(body.value));
    }
    (void)world_upload_body(w, phys_slot(body.value));
}

void nph_physics_body_set_locks(NphPhysicsWorld *w, NphPhysicsBody body, u32 locks)
{
    NphBodyRec *b = editable(w, body, NULL);
    if (b == NULL || (locks & ~0x3FU) != 0U) {
        return;
    }
    b->locks = locks;
    (void)world_upload_body(w, phys_slot(body.value));
}

void nph_physics_body_set_sleep(NphPhysicsWorld *w, NphPhysicsBody body, bool enable,
                                f32 threshold)
{
    HBody *hb = NULL;
    NphBodyRec *b = editable(w, body, &hb);
    if (b == NULL || !finite_nonneg(threshold)) {
        return;
    }
    b->sleep_threshold = threshold;
    if (enable) {
        hb->flags |= HB_ENABLE_SLEEP;
        b->flags |= HB_ENABLE_SLEEP;
    } else {
        hb->flags &= ~(u32)HB_ENABLE_SLEEP;
        b->flags &= ~(u32)HB_ENABLE_SLEEP;
        b->sleep_time = 0.0F;
        if (hb->type != NPH_PHYSICS_BODY_STATIC) {
            (void)world_wake_body(w, phys_slot(body.value));
        }
    }
    (void)world_upload_body(w, phys_slot(body.value));
}

void nph_physics_body_set_awake(NphPhysicsWorld *w, NphPhysicsBody body, bool awake)
{
    HBody *hb = world_body(w, body);
    if (w == NULL || w->stepping || hb == NULL || hb->type == NPH_PHYSICS_BODY_STATIC ||
        world_sync_bodies(w) != NPH_RESULT_OK) {
        return;
    }
    const u32 slot = phys_slot(body.value);
    if (awake) {
        wake_and_upload(w, body);
        return;
    }
    if ((hb->flags & HB_AWAKE) == 0U) {
        return;
    }
    // Asleep on request: the whole island (what touches it would keep it
    // awake otherwise); a kinematic body alone.
    if (hb->island != NPH_NULL32) {
        (void)islands_sleep_now(w, islands_root(w, hb->island));
    } else {
        (void)world_set_awake_list(w, slot, 0);
    }
}

void nph_physics_body_set_user_data(NphPhysicsWorld *w, NphPhysicsBody body, u64 user_data)
{
    HBody *b = world_body(w, body);
    if (b != NULL) {
        b->user_data = user_data;
    }
}

u32 nph_physics_body_shapes(const NphPhysicsWorld *cw, NphPhysicsBody body, NphPhysicsShape *out,
                            u32 capacity)
{
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    const HBody *b = world_body(w, body);
    if (b == NULL) {
        return 0U;
    }
    u32 n = 0U;
    for (u32 s = b->shape_head; s != NPH_NULL32; s = hshape(w, s)->next) {
        if (n < capacity && out != NULL) {
            out[n].value = hshape(w, s)->handle;
        }
        n += 1U;
    }
    return n;
}

// ---------------------------------------------------------------------------
// Shapes
// ---------------------------------------------------------------------------
f32 nph_physics_combine_friction(f32 a, NphPhysicsCombine mode_a, f32 b, NphPhysicsCombine mode_b)
{
    return combine_values(a, (u32)mode_a, b, (u32)mode_b, 0);
}

f32 nph_physics_combine_restitution(f32 a, NphPhysicsCombine mode_a, f32 b,
                                    NphPhysicsCombine mode_b)
{
    return combine_values(a, (u32)mode_a, b, (u32)mode_b, 1);
}

// shape.c:285-311
void nph_physics_shape_destroy(NphPhysicsWorld *w, NphPhysicsShape shape)
{
    HShape *s = w != NULL && !w->stepping ? world_shape(w, shape) : NULL;
    if (s == NULL || world_sync_bodies(w) != NPH_RESULT_OK) {
        return;
    }
    const u32 slot = phys_slot(shape.value);
    const u32 body = s->body;
    destroy_shape_contacts(w, slot);
    destroy_proxy(w, slot);
    HBody *b = hbody(w, body);
    u32 prev = NPH_NULL32;
    for (u32 k = b->shape_head; k != slot; k = hshape(w, k)->next) {
        prev = k;
    }
    if (prev == NPH_NULL32) {
        b->shape_head = s->next;
        world_rec(w, body)->shape_head = s->next;
    } else {
        hshape(w, prev)->next = s->next;
        (void)world_refresh_shape(w, prev);
    }
    b->shape_count -= 1U;
    s->handle = 0U;
    w->hit_event_shapes -= (s->flags & HS_HIT_EVENTS) != 0U ? 1U : 0U;
    slots_release(w->alloc, &w->shape_slots, slot);
    world_update_mass(w, body);
    if (b->type != NPH_PHYSICS_BODY_STATIC) {
        (void)world_wake_body(w, body);
    }
    (void)world_upload_body(w, body);
}

NphPhysicsGeometry nph_physics_shape_geometry(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    if (s != NULL) {
        return s->geometry;
    }
    NphPhysicsGeometry g;
    memset(&g, 0, sizeof g);
    g.type = NPH_PHYSICS_SHAPE_SPHERE;
    return g;
}

NphPhysicsPose nph_physics_shape_pose(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? s->pose : pose_identity();
}

f32 nph_physics_shape_friction(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? s->friction : 0.0F;
}

f32 nph_physics_shape_restitution(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? s->restitution : 0.0F;
}

NphPhysicsCombine nph_physics_shape_friction_combine(const NphPhysicsWorld *w,
                                                     NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? (NphPhysicsCombine)(s->combine & 0xFFU) : NPH_PHYSICS_COMBINE_DEFAULT;
}

NphPhysicsCombine nph_physics_shape_restitution_combine(const NphPhysicsWorld *w,
                                                        NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? (NphPhysicsCombine)(s->combine >> 8U) : NPH_PHYSICS_COMBINE_DEFAULT;
}

// A surface setting reaches the device's shape record (the narrow phase
// reads it every step, shape.h).
static void surface_changed(NphPhysicsWorld *w, NphPhysicsShape shape)
{
    (void)world_refresh_shape(w, phys_slot(shape.value));
}

void nph_physics_shape_set_friction(NphPhysicsWorld *w, NphPhysicsShape shape, f32 friction)
{
    HShape *s = world_shape(w, shape);
    if (s != NULL) {
        s->friction = friction;
        surface_changed(w, shape);
    }
}

void nph_physics_shape_set_restitution(NphPhysicsWorld *w, NphPhysicsShape shape, f32 restitution)
{
    HShape *s = world_shape(w, shape);
    if (s != NULL) {
        s->restitution = restitution;
        surface_changed(w, shape);
    }
}

void nph_physics_shape_set_friction_combine(NphPhysicsWorld *w, NphPhysicsShape shape,
                                            NphPhysicsCombine mode)
{
    HShape *s = world_shape(w, shape);
    if (s != NULL && (u32)mode < NPH_PHYSICS_COMBINE_COUNT) {
        s->combine = (s->combine & ~0xFFU) | (u32)mode;
        surface_changed(w, shape);
    }
}

void nph_physics_shape_set_restitution_combine(NphPhysicsWorld *w, NphPhysicsShape shape,
                                               NphPhysicsCombine mode)
{
    HShape *s = world_shape(w, shape);
    if (s != NULL && (u32)mode < NPH_PHYSICS_COMBINE_COUNT) {
        s->combine = (s->combine & 0xFFU) | ((u32)mode << 8U);
        surface_changed(w, shape);
    }
}

f32 nph_physics_shape_density(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? s->density : 0.0F;
}

void nph_physics_shape_set_density(NphPhysicsWorld *w, NphPhysicsShape shape, f32 density)
{
    HShape *s = w != NULL && !w->stepping ? world_shape(w, shape) : NULL;
    if (s == NULL || !(nph_isfinitef(density) && density >= 0.0F) ||
        world_sync_bodies(w) != NPH_RESULT_OK) {
        return;
    }
    s->density = density;
    world_update_mass(w, s->body);
    (void)world_upload_body(w, s->body);
}

NphPhysicsFilter nph_physics_shape_filter(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? s->filter : nph_physics_filter_default();
}

// shape.c:404-436
void nph_physics_shape_set_filter(NphPhysicsWorld *w, NphPhysicsShape shape,
                                  NphPhysicsFilter filter)
{
    HShape *s = w != NULL && !w->stepping ? world_shape(w, shape) : NULL;
    if (s == NULL) {
        return;
    }
    const u32 slot = phys_slot(shape.value);
    s->filter = filter;
    tree_set_category_bits(&w->trees[s->tree], s->proxy, filter.category);
    pairs_mark_shape(w, slot);
    const u32 body = s->body;
    for (u32 key = hbody(w, body)->contact_head; key != NPH_NULL32;) {
        const u32 c = key >> 1U;
        const HContact *contact = contacts_at(w, c);
        key = contact->next[key & 1U];
        if (contact->shape_a != slot && contact->shape_b != slot) {
            continue;
        }
        const u32 other = contact->shape_a == slot ? contact->shape_b : contact->shape_a;
        if (!nph_physics_filter_should_collide(filter, hshape(w, other)->filter)) {
            const u32 other_body = hshape(w, other)->body;
            const int touching = (contact->flags & HC_TOUCHING) != 0U;
            contacts_destroy(w, c, 1);
            if (touching) {
                (void)world_wake_body(w, other_body);
            }
        }
    }
    (void)move_buffer_push(w, slot);
    if (hbody(w, body)->type != NPH_PHYSICS_BODY_STATIC) {
        (void)world_wake_body(w, body);
    }
}

u64 nph_physics_shape_user_data(const NphPhysicsWorld *w, NphPhysicsShape shape)
{
    const HShape *s = world_shape(w, shape);
    return s != NULL ? s->user_data : 0U;
}

void nph_physics_shape_set_user_data(NphPhysicsWorld *w, NphPhysicsShape shape, u64 user_data)
{
    HShape *s = world_shape(w, shape);
    if (s != NULL) {
        s->user_data = user_data;
    }
}
