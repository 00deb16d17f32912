// SPDX-License-Identifier: Apache-2.0
// Joints on the host (engine/physics/src/joint.c): creation, destruction,
// the bodies' joint lists, graph membership and the public queries.  The
// configuration goes to the device as a record by joint slot; the impulses
// live there and are read back only for the reaction query.
#define _POSIX_C_SOURCE 200809L
#include "world_impl.h"

#include <math.h>

static HJoint *joint_at(const NphPhysicsWorld *w, u32 slot) { return &VB(w->joints, HJoint)[slot]; }
static HBody *jbody(const NphPhysicsWorld *w, u32 slot) { return &VB(w->bodies, HBody)[slot]; }

static HJoint *world_joint(const NphPhysicsWorld *w, NphPhysicsJoint h)
{
    const u32 slot = phys_slot(h.value);
    if (h.value == 0U || slot >= w->joints.len || joint_at(w, slot)->handle != h.value) {
        return NULL;
    }
    return joint_at(w, slot);
}

void joints_free(NphPhysicsWorld *w)
{
    NphAllocator *a = w->alloc;
    vbuf_free(a, &w->joint_slots.generation, 4U);
    vbuf_free(a, &w->joint_slots.free, 4U);
    vbuf_free(a, &w->joints, sizeof(HJoint));
    for (u32 c = 0U; c <= NPH_GRAPH_COLORS; ++c) {
        vbuf_free(a, &w->color_joints[c], 4U);
    }
    vbuf_free(a, &w->step_joints, 4U);
}

// ---------------------------------------------------------------------------
// Descriptions (joint.c:37-107)
// ---------------------------------------------------------------------------
NphPhysicsJointDesc nph_physics_joint_desc_default(NphPhysicsJointType type, NphPhysicsBody a,
                                                   NphPhysicsBody b)
{
    NphPhysicsJointDesc d;
    memset(&d, 0, sizeof d);
    d.type = type;
    d.body_a = a;
    d.body_b = b;
    d.frame_a = pose_identity();
    d.frame_b = pose_identity();
    const f32 pi = 3.14159265F;
    if (type == NPH_PHYSICS_JOINT_BALL) {
        d.ball.swing_y = 0.25F * pi;
        d.ball.swing_z = 0.25F * pi;
        d.ball.twist_lower = -0.25F * pi;
        d.ball.twist_upper = 0.25F * pi;
    } else if (type == NPH_PHYSICS_JOINT_HINGE) {
        d.hinge.lower = -0.5F * pi;
        d.hinge.upper = 0.5F * pi;
        d.hinge.spring_hertz = 1.0F;
        d.hinge.spring_damping_ratio = 1.0F;
    } else if (type == NPH_PHYSICS_JOINT_SLIDER) {
        d.slider.lower = -1.0F;
        d.slider.upper = 1.0F;
        d.slider.spring_hertz = 1.0F;
        d.slider.spring_damping_ratio = 1.0F;
    } else if (type == NPH_PHYSICS_JOINT_DISTANCE) {
        d.distance.length = 1.0F;
        d.distance.max_length = 1.0F;
        d.distance.spring_hertz = 1.0F;
        d.distance.spring_damping_ratio = 1.0F;
    } else if (type == NPH_PHYSICS_JOINT_D6) {
        for (u32 i = 0U; i < NPH_PHYSICS_D6_AXES; ++i) {
            const f32 range = i < 3U ? 1.0F : 0.25F * pi;
            d.d6.axes[i].lower = -range;
            d.d6.axes[i].upper = range;
            d.d6.axes[i].spring_hertz = 1.0F;
            d.d6.axes[i].spring_damping_ratio = 1.0F;
        }
    }
    return d;
}

// Shepperd's method, the largest of the four candidates as the divisor
// (mat.c:31-73).
static NphQuat quat_from_basis(NphVec3 x, NphVec3 y, NphVec3 z)
{
    const f32 m00 = x.x, m10 = x.y, m20 = x.z;
    const f32 m01 = y.x, m11 = y.y, m21 = y.z;
    const f32 m02 = z.x, m12 = z.y, m22 = z.z;
    const f32 trace = m00 + m11 + m22;
    NphQuat q;
    if (trace > 0.0F) {
        const f32 s = nph_sqrtf(trace + 1.0F) * 2.0F;
        q = q4((m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s, 0.25F * s);
    } else if (m00 > m11 && m00 > m22) {
        const f32 s = nph_sqrtf(1.0F + m00 - m11 - m22) * 2.0F;
        q = q4(0.25F * s, (m01 + m10) / s, (m02 + m20) / s, (m21 - m12) / s);
    } else if (m11 > m22) {
        const f32 s = nph_sqrtf(1.0F + m11 - m00 - m22) * 2.0F;
        q = q4((m01 + m10) / s, 0.25F * s, (m12 + m21) / s, (m02 - m20) / s);
    } else {
        const f32 s = nph_sqrtf(1.0F + m22 - m00 - m11) * 2.0F;
        q = q4((m02 + m20) / s, (m12 + m21) / s, 0.25F * s, (m10 - m01) / s);
    }
    return q_normalize(q);
}

void nph_physics_joint_frames_at(const NphPhysicsWorld *w, NphPhysicsBody a, NphPhysicsBody b,
                                 NphDVec3 anchor, NphVec3 axis, NphPhysicsPose *frame_a,
                                 NphPhysicsPose *frame_b)
{
    const HBody *ha = world_body(w, a);
    const HBody *hb = world_body(w, b);
    if (ha == NULL || hb == NULL || world_sync_bodies((NphPhysicsWorld *)w) != NPH_RESULT_OK) {
        *frame_a = pose_identity();
        *frame_b = pose_identity();
        return;
    }
    const NphBodyRec *ba = &VB(w->body_recs, NphBodyRec)[phys_slot(a.value)];
    const NphBodyRec *bb = &VB(w->body_recs, NphBodyRec)[phys_slot(b.value)];
    const NphVec3 x = v3_normalize_or(axis, v3(1.0F, 0.0F, 0.0F));
    const NphVec3 y = v3_perpendicular(x);
    const NphVec3 z = v3_cross(x, y);
    const NphQuat q = q_normalize(quat_from_basis(x, y, z));
    frame_a->rotation = q_normalize(q_mul(q_conj(ba->rotation), q));
    frame_b->rotation = q_normalize(q_mul(q_conj(bb->rotation), q));
    frame_a->position = q_rotate_inv(ba->rotation, dv3_relative(anchor, ba->origin));
    frame_b->position = q_rotate_inv(bb->rotation, dv3_relative(anchor, bb->origin));
}

// ---------------------------------------------------------------------------
// Validation (joint.c:109-217)
// ---------------------------------------------------------------------------
static int pose_valid(NphPhysicsPose p)
{
    const NphQuat q = p.rotation;
    return nph_isfinitef(p.position.x) && nph_isfinitef(p.position.y) &&
           nph_isfinitef(p.position.z) && nph_isfinitef(q.x) && nph_isfinitef(q.y) &&
           nph_isfinitef(q.z) && nph_isfinitef(q.w) &&
           nph_absf(q_len_sq(q) - 1.0F) <= 2.0F * 1e-3F;
}

static int finite_nonneg(f32 v) { return nph_isfinitef(v) && v >= 0.0F; }

static int angle_range_valid(f32 lower, f32 upper)
{
    const f32 pi = 3.14159265F;
    return nph_isfinitef(lower) && nph_isfinitef(upper) && lower <= upper && lower >= -pi &&
           upper <= pi;
}

static int ball_valid(const NphPhysicsBallJoint *b)
{
    const f32 pi = 3.14159265F;
    if (b->enable_swing_limit != 0 &&
        !(b->swing_y > 0.0F && b->swing_y < pi && b->swing_z > 0.0F && b->swing_z < pi)) {
        return 0;
    }
    if (b->enable_twist_limit != 0 && !angle_range_valid(b->twist_lower, b->twist_upper)) {
        return 0;
    }
    return finite_nonneg(b->friction_torque);
}

static int hinge_valid(const NphPhysicsHingeJoint *h)
{
    if (h->enable_limit != 0 && !angle_range_valid(h->lower, h->upper)) {
        return 0;
    }
    if (h->enable_motor != 0 &&
        (!nph_isfinitef(h->motor_speed) || !finite_nonneg(h->max_motor_torque))) {
        return 0;
    }
    if (h->enable_spring != 0 &&
        (!(h->spring_hertz > 0.0F) || !nph_isfinitef(h->spring_hertz) ||
         !finite_nonneg(h->spring_damping_ratio) || !nph_isfinitef(h->spring_target))) {
        return 0;
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Lists and membership (joint.c:219-293)
// ---------------------------------------------------------------------------
static void list_insert(NphPhysicsWorld *w, u32 joint, u32 side, u32 body)
{
    HBody *b = jbody(w, body);
    HJoint *j = joint_at(w, joint);
    const u32 key = (joint << 1U) | side;
    j->prev[side] = NPH_NULL32;
    j->next[side] = b->joint_head;
    if (b->joint_head != NPH_NULL32) {
        joint_at(w, b->joint_head >> 1U)->prev[b->joint_head & 1U] = key;
    }
    b->joint_head = key;
    b->joint_count += 1U;
    pairs_mark_body(w, body);
}

static void list_remove(NphPhysicsWorld *w, u32 joint, u32 side, u32 body)
{
    HBody *b = jbody(w, body);
    HJoint *j = joint_at(w, joint);
    const u32 prev = j->prev[side];
    const u32 next = j->next[side];
    if (prev != NPH_NULL32) {
        joint_at(w, prev >> 1U)->next[prev & 1U] = next;
    } else {
        b->joint_head = next;
    }
    if (next != NPH_NULL32) {
        joint_at(w, next >> 1U)->prev[next & 1U] = prev;
    }
    b->joint_count -= 1U;
    pairs_mark_body(w, body);
}

static int body_awake(const HBody *b) { return (b->flags & HB_AWAKE) != 0U; }
static int body_solvable(const HBody *b)
{
    return body_awake(b) || b->type != (u32)NPH_PHYSICS_BODY_DYNAMIC;
}

// island.c:106-118
static NphResult graph_add_joint(NphPhysicsWorld *w, u32 joint)
{
    HJoint *j = joint_at(w, joint);
    const u32 color = graph_pick_color(w, j->body_a, j->body_b);
    NphVecBuf *list = &w->color_joints[color];
    NphResult r = vbuf_push(w->alloc, list, 4U, &joint);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    graph_set_bodies(w, color, j->body_a, j->body_b, 1);
    j->color = color;
    j->color_slot = list->len - 1U;
    return NPH_RESULT_OK;
}

// island.c:120-135
static void graph_remove_joint(NphPhysicsWorld *w, u32 joint)
{
    HJoint *j = joint_at(w, joint);
    const u32 color = j->color;
    NphVecBuf *list = &w->color_joints[color];
    u32 *d = VB(*list, u32);
    const u32 slot = j->color_slot;
    const u32 last = d[list->len - 1U];
    d[slot] = last;
    joint_at(w, last)->color_slot = slot;
    list->len -= 1U;
    graph_set_bodies(w, color, j->body_a, j->body_b, 0);
    j->color = NPH_NULL32;
    j->color_slot = NPH_NULL32;
}

// joint.c:257-271
NphResult joints_sync(NphPhysicsWorld *w, u32 joint)
{
    HJoint *j = joint_at(w, joint);
    const HBody *a = jbody(w, j->body_a);
    const HBody *b = jbody(w, j->body_b);
    const int in_graph = (body_awake(a) || body_awake(b)) && body_solvable(a) && body_solvable(b);
    if (in_graph && j->color == NPH_NULL32) {
        return graph_add_joint(w, joint);
    }
    if (!in_graph && j->color != NPH_NULL32) {
        graph_remove_joint(w, joint);
    }
    return NPH_RESULT_OK;
}

// joint.c:273-291: no joint between the two bodies keeps their shapes apart.
int joints_allow_contact(const NphPhysicsWorld *w, u32 body_a, u32 body_b)
{
    const HBody *a = jbody(w, body_a);
    const HBody *b = jbody(w, body_b);
    const u32 body = a->joint_count <= b->joint_count ? body_a : body_b;
    const u32 other = body == body_a ? body_b : body_a;
    for (u32 key = jbody(w, body)->joint_head; key != NPH_NULL32;) {
        const HJoint *j = joint_at(w, key >> 1U);
        const u32 far = (key & 1U) == 0U ? j->body_b : j->body_a;
        if (far == other && (j->flags & NPH_JOINT_COLLIDE_CONNECTED) == 0U) {
            return 0;
        }
        key = j->next[key & 1U];
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Lifetime (joint.c:357-473)
// ---------------------------------------------------------------------------
static NphResult upload_conf(NphPhysicsWorld *w, u32 slot)
{
    const HJoint *j = joint_at(w, slot);
    NphJointConf c;
    memset(&c, 0, sizeof c);
    c.type = j->type;
    c.flags = j->flags;
    c.body_a = j->body_a;
    c.body_b = j->body_b;
    c.frame_a = j->frame_a;
    c.frame_b = j->frame_b;
    c.s = j->s;
    return nph_dev_write_joint(w->dev, slot, &c, 0);
}

static void joint_destroy(NphPhysicsWorld *w, u32 joint, int wake)
{
    HJoint *j = joint_at(w, joint);
    const u32 a = j->body_a;
    const u32 b = j->body_b;
    if (j->color != NPH_NULL32) {
        graph_remove_joint(w, joint);
    }
    if ((j->flags & HJ_ISLAND) != 0U) {
        islands_unlink_joint(w, joint);
    }
    list_remove(w, joint, 0U, a);
    list_remove(w, joint, 1U, b);
    pairs_touch_key(w, nph_pair_key(a, b) | NPH_PAIR_BODY_KEY);
    if ((j->flags & NPH_JOINT_COLLIDE_CONNECTED) == 0U) {
        // The shapes may touch now: A's look for pairs next step.
        for (u32 s = jbody(w, a)->shape_head; s != NPH_NULL32;
             s = VB(w->shapes, HShape)[s].next) {
            (void)move_buffer_push(w, s);
        }
    }
    j->handle = 0U;
    slots_release(w->alloc, &w->joint_slots, joint);
    w->joint_count -= 1U;
    if (wake) {
        (void)world_wake_body(w, a);
        (void)world_wake_body(w, b);
    }
}

NphResult nph_physics_joint_create(NphPhysicsWorld *w, const NphPhysicsJointDesc *desc,
                                   NphPhysicsJoint *out)
{
    if (w == NULL || out == NULL || w->stepping) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    if (desc == NULL || (u32)desc->type >= (u32)NPH_PHYSICS_JOINT_TYPE_COUNT ||
        !pose_valid(desc->frame_a) || !pose_valid(desc->frame_b) ||
        desc->body_a.value == desc->body_b.value) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    if (desc->type == NPH_PHYSICS_JOINT_SLIDER || desc->type == NPH_PHYSICS_JOINT_DISTANCE ||
        desc->type == NPH_PHYSICS_JOINT_D6) {
        return NPH_RESULT_UNSUPPORTED;
    }
    if ((desc->type == NPH_PHYSICS_JOINT_BALL && !ball_valid(&desc->ball)) ||
        (desc->type == NPH_PHYSICS_JOINT_HINGE && !hinge_valid(&desc->hinge))) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const HBody *ba = world_body(w, desc->body_a);
    const HBody *bb = world_body(w, desc->body_b);
    if (ba == NULL || bb == NULL) {
        return NPH_RESULT_NOT_FOUND;
    }
    if (ba->type != (u32)NPH_PHYSICS_BODY_DYNAMIC && bb->type != (u32)NPH_PHYSICS_BODY_DYNAMIC) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const u32 slot_a = phys_slot(desc->body_a.value);
    const u32 slot_b = phys_slot(desc->body_b.value);
    u32 slot = 0U;
    u32 generation = 0U;
    NphResult r = slots_alloc(w->alloc, &w->joint_slots, &slot, &generation);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    if (slot >= w->joints.len) {
        r = vbuf_reserve(w->alloc, &w->joints, sizeof(HJoint), slot + 1U);
        if (r == NPH_RESULT_OK) {
            r = nph_dev_reserve_joints(w->dev, slot + 1U);
        }
        if (r != NPH_RESULT_OK) {
            slots_release(w->alloc, &w->joint_slots, slot);
            return r;
        }
        w->joints.len = slot + 1U;
    }
    HJoint *j = joint_at(w, slot);
    memset(j, 0, sizeof *j);
    j->type = (u32)desc->type;
    j->body_a = slot_a;
    j->body_b = slot_b;
    j->flags = desc->collide_connected != 0 ? NPH_JOINT_COLLIDE_CONNECTED : 0U;
    j->color = NPH_NULL32;
    j->color_slot = NPH_NULL32;
    j->island = NPH_NULL32;
    j->island_prev = NPH_NULL32;
    j->island_next = NPH_NULL32;
    j->frame_a = desc->frame_a;
    j->frame_b = desc->frame_b;
    j->frame_a.rotation = q_normalize(desc->frame_a.rotation);
    j->frame_b.rotation = q_normalize(desc->frame_b.rotation);
    if (desc->type == NPH_PHYSICS_JOINT_BALL) {
        j->s.ball = desc->ball;
    } else if (desc->type == NPH_PHYSICS_JOINT_HINGE) {
        j->s.hinge = desc->hinge;
    }
    j->user_data = desc->user_data;
    j->handle = phys_handle(slot, generation);
    w->joint_count += 1U;
    list_insert(w, slot, 0U, slot_a);
    list_insert(w, slot, 1U, slot_b);
    pairs_touch_key(w, nph_pair_key(slot_a, slot_b) | NPH_PAIR_BODY_KEY);
    // Joined bodies keep apart unless asked: their contacts go now.
    if (desc->collide_connected == 0) {
        for (u32 key = jbody(w, slot_a)->contact_head; key != NPH_NULL32;) {
            const u32 c = key >> 1U;
            const HContact *contact = contacts_at(w, c);
            const u32 other = (key & 1U) == 0U ? contact->body_b : contact->body_a;
            key = contact->next[key & 1U];
            if (other == slot_b) {
                contacts_destroy(w, c, 1);
            }
        }
    }
    // Both ends awake, their islands linked (merged at the step's end).
    r = world_wake_body(w, slot_a);
    if (r == NPH_RESULT_OK) {
        r = world_wake_body(w, slot_b);
    }
    if (r == NPH_RESULT_OK && jbody(w, slot_a)->type == NPH_PHYSICS_BODY_DYNAMIC &&
        jbody(w, slot_b)->type == NPH_PHYSICS_BODY_DYNAMIC) {
        islands_link_joint(w, slot);
    }
    if (r == NPH_RESULT_OK) {
        r = joints_sync(w, slot);
    }
    if (r == NPH_RESULT_OK) {
        r = upload_conf(w, slot);
    }
    if (r == NPH_RESULT_OK) {
        r = nph_dev_write_joint(w->dev, slot, NULL, 1); // impulses from nothing
    }
    if (r != NPH_RESULT_OK) {
        joint_destroy(w, slot, 0);
        return r;
    }
    out->value = j->handle;
    return NPH_RESULT_OK;
}

void joints_destroy(NphPhysicsWorld *w, u32 joint, int wake) { joint_destroy(w, joint, wake); }

bool nph_physics_joint_is_null(NphPhysicsJoint j) { return j.value == 0U; }

void nph_physics_joint_destroy(NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    if (w == NULL || w->stepping || world_joint(w, joint) == NULL) {
        return;
    }
    joint_destroy(w, phys_slot(joint.value), 1);
}

bool nph_physics_joint_is_valid(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    return world_joint(w, joint) != NULL;
}

NphPhysicsJointDesc nph_physics_joint_desc(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    NphPhysicsJointDesc d;
    memset(&d, 0, sizeof d);
    const HJoint *j = world_joint(w, joint);
    if (j == NULL) {
        return d;
    }
    d.type = (NphPhysicsJointType)j->type;
    d.body_a.value = jbody(w, j->body_a)->handle;
    d.body_b.value = jbody(w, j->body_b)->handle;
    d.frame_a = j->frame_a;
    d.frame_b = j->frame_b;
    d.collide_connected = (j->flags & NPH_JOINT_COLLIDE_CONNECTED) != 0U ? 1 : 0;
    d.user_data = j->user_data;
    if (j->type == (u32)NPH_PHYSICS_JOINT_BALL) {
        d.ball = j->s.ball;
    } else if (j->type == (u32)NPH_PHYSICS_JOINT_HINGE) {
        d.hinge = j->s.hinge;
    }
    return d;
}

static void wake_ends(NphPhysicsWorld *w, const HJoint *j)
{
    (void)world_wake_body(w, j->body_a);
    (void)world_wake_body(w, j->body_b);
}

void nph_physics_joint_set_ball(NphPhysicsWorld *w, NphPhysicsJoint joint,
                                const NphPhysicsBallJoint *ball)
{
    HJoint *j = w->stepping ? NULL : world_joint(w, joint);
    if (j == NULL || j->type != (u32)NPH_PHYSICS_JOINT_BALL || !ball_valid(ball)) {
        return;
    }
    j->s.ball = *ball;
    (void)upload_conf(w, phys_slot(joint.value));
    wake_ends(w, j);
}

void nph_physics_joint_set_hinge(NphPhysicsWorld *w, NphPhysicsJoint joint,
                                 const NphPhysicsHingeJoint *hinge)
{
    HJoint *j = w->stepping ? NULL : world_joint(w, joint);
    if (j == NULL || j->type != (u32)NPH_PHYSICS_JOINT_HINGE || !hinge_valid(hinge)) {
        return;
    }
    j->s.hinge = *hinge;
    (void)upload_conf(w, phys_slot(joint.value));
    wake_ends(w, j);
}

// Joints of these types cannot be made here (NPH_RESULT_UNSUPPORTED), so no
// handle names one: the setters have nothing to change.
void nph_physics_joint_set_slider(NphPhysicsWorld *w, NphPhysicsJoint joint,
                                  const NphPhysicsSliderJoint *slider)
{
    (void)w;
    (void)joint;
    (void)slider;
}

void nph_physics_joint_set_distance(NphPhysicsWorld *w, NphPhysicsJoint joint,
                                    const NphPhysicsDistanceJoint *distance)
{
    (void)w;
    (void)joint;
    (void)distance;
}

void nph_physics_joint_set_d6(NphPhysicsWorld *w, NphPhysicsJoint joint,
                              const NphPhysicsD6Joint *d6)
{
    (void)w;
    (void)joint;
    (void)d6;
}

// ---------------------------------------------------------------------------
// Queries between steps (joint.c:597-693, joint_solver.c:1316-1445)
// ---------------------------------------------------------------------------
static const NphBodyRec *jrec(const NphPhysicsWorld *w, u32 slot)
{
    return &VB(w->body_recs, NphBodyRec)[slot];
}

// The frames, anchors and B's anchor's offset from A's now.
static void frames_now(const NphPhysicsWorld *w, const HJoint *j, NphQuat *qa, NphQuat *qb,
                       NphVec3 *ra, NphVec3 *rb, NphVec3 *d)
{
    const NphBodyRec *a = jrec(w, j->body_a);
    const NphBodyRec *b = jrec(w, j->body_b);
    *qa = q_mul(a->rotation, j->frame_a.rotation);
    *qb = q_mul(b->rotation, j->frame_b.rotation);
    *ra = q_rotate(a->rotation, v3_sub(j->frame_a.position, a->local_center));
    *rb = q_rotate(b->rotation, v3_sub(j->frame_b.position, b->local_center));
    *d = v3_add(dv3_relative(b->center, a->center), v3_sub(*rb, *ra));
}

static const HJoint *query_joint(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    const HJoint *j = world_joint(w, joint);
    if (j == NULL || world_sync_bodies((NphPhysicsWorld *)w) != NPH_RESULT_OK) {
        return NULL;
    }
    return j;
}

f32 nph_physics_joint_hinge_angle(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    const HJoint *j = query_joint(w, joint);
    if (j == NULL) {
        return 0.0F;
    }
    return joint_twist_angle(q_mul(jrec(w, j->body_a)->rotation, j->frame_a.rotation),
                           q_mul(jrec(w, j->body_b)->rotation, j->frame_b.rotation));
}

f32 nph_physics_joint_hinge_speed(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    const HJoint *j = query_joint(w, joint);
    if (j == NULL) {
        return 0.0F;
    }
    const NphBodyRec *a = jrec(w, j->body_a);
    const NphBodyRec *b = jrec(w, j->body_b);
    const NphVec3 axis = q_axis_x(q_mul(a->rotation, j->frame_a.rotation));
    return v3_dot(v3_sub(b->w, a->w), axis);
}

f32 nph_physics_joint_slider_position(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    (void)w;
    (void)joint;
    return 0.0F;
}

f32 nph_physics_joint_slider_speed(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    (void)w;
    (void)joint;
    return 0.0F;
}

f32 nph_physics_joint_length(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    const HJoint *j = query_joint(w, joint);
    if (j == NULL) {
        return 0.0F;
    }
    NphQuat qa, qb;
    NphVec3 ra, rb, d;
    frames_now(w, j, &qa, &qb, &ra, &rb, &d);
    return v3_len(d);
}

void nph_physics_joint_reaction(const NphPhysicsWorld *w, NphPhysicsJoint joint, NphVec3 *force,
                                NphVec3 *torque)
{
    *force = v3_zero();
    *torque = v3_zero();
    const HJoint *j = query_joint(w, joint);
    const f32 h = w->last_substep;
    if (j == NULL || !(h > 0.0F)) {
        return;
    }
    const u32 slot = phys_slot(joint.value);
    NphJointState st;
    if (nph_dev_read_joint(w->dev, slot, &st) != NPH_RESULT_OK) {
        return;
    }
    NphJointConf cf;
    memset(&cf, 0, sizeof cf);
    cf.type = j->type;
    cf.s = j->s;
    NphQuat qa, qb;
    NphVec3 ra, rb, d;
    frames_now(w, j, &qa, &qb, &ra, &rb, &d);
    // Per-substep impulses over the substep: the last substep's force and
    // torque (joint.c:646-665).
    const f32 inv_h = 1.0F / h;
    const NphBodyRec *b = jrec(w, j->body_b);
    *force = v3_scale(st.linear_impulse, inv_h);
    const NphVec3 rbn = q_rotate(b->rotation, v3_sub(j->frame_b.position, b->local_center));
    *torque = v3_add(v3_scale(joint_angular_impulse(&cf, &st, qa, qb), inv_h), v3_cross(rbn, *force));
}

u64 nph_physics_joint_user_data(const NphPhysicsWorld *w, NphPhysicsJoint joint)
{
    const HJoint *j = world_joint(w, joint);
    return j != NULL ? j->user_data : 0U;
}

void nph_physics_joint_set_user_data(NphPhysicsWorld *w, NphPhysicsJoint joint, u64 user_data)
{
    HJoint *j = world_joint(w, joint);
    if (j != NULL) {
        j->user_data = user_data;
    }
}

u32 nph_physics_body_joints(const NphPhysicsWorld *w, NphPhysicsBody body, NphPhysicsJoint *out,
                            u32 capacity)
{
    const HBody *b = world_body(w, body);
    if (b == NULL) {
        return 0U;
    }
    u32 n = 0U;
    for (u32 key = b->joint_head; key != NPH_NULL32;) {
        const HJoint *j = joint_at(w, key >> 1U);
        if (n < capacity && out != NULL) {
            out[n].value = j->handle;
        }
        n += 1U;
        key = j->next[key & 1U];
    }
    return n;
}
