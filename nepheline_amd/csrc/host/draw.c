// SPDX-License-Identifier: Apache-2.0
// Debug drawing (world.h:291-325; world.c:698-900): a world's shapes, fat
// boxes, centres of mass, joints and touching contact points as lines
// through the caller's callback, from the host's copies of the device's
// body poses and contact records.
//
// Every line's ends are 64-bit world points narrowed to 32 bits relative to
// the draw's origin. The order of the calls and their colours are the
// reference's: bodies by slot, each body's shapes in list order (outline,
// then fat box), then its centre of mass; joints by slot; contacts by
// record, points in order.
#include "world_impl.h"

#include <math.h>

#define RING_SEGMENTS 16U
#define MASS_ARM 0.1

#define RGBA_AWAKE 0xFF80C080U
#define RGBA_STATIC 0xFF808080U
#define RGBA_KINEMATIC 0xFFC08080U
#define RGBA_ASLEEP 0xFF806040U
#define RGBA_FAT_BOX 0xFF40A0FFU
#define RGBA_MASS 0xFFFFFFFFU
#define RGBA_ANCHOR 0xFF00C0FFU
#define RGBA_JOINT_ERROR 0xFF0000FFU
#define RGBA_FRAME_A 0xFFFFFF00U
#define RGBA_FRAME_B 0xFFFF00FFU
#define RGBA_SPECULATIVE 0xFF808000U
#define RGBA_PENETRATING 0xFF0000FFU
#define INACTIVE_EDGE_MASK 0xFF7F7F7FU // an inactive mesh edge: colour halved

static NphVec3 from_origin(const NphPhysicsDebugDraw *d, NphDVec3 p)
{
    return v3((f32)(p.x - d->origin.x), (f32)(p.y - d->origin.y), (f32)(p.z - d->origin.z));
}

static void segment(const NphPhysicsDebugDraw *d, NphDVec3 a, NphDVec3 b, u32 rgba)
{
    d->line(d->user, from_origin(d, a), from_origin(d, b), rgba);
}

// A line between two points given relative to `base`.
static void segment_at(const NphPhysicsDebugDraw *d, NphDVec3 base, NphVec3 a, NphVec3 b, u32 rgba)
{
    segment(d, dv3_add_v3(base, a), dv3_add_v3(base, b), rgba);
}

// The twelve edges of a box given by its corners in bit order (bit 0: +x,
// bit 1: +y, bit 2: +z): the x edges, then y, then z, each in the order of
// their lower corner.
static void box_edges(const NphPhysicsDebugDraw *d, const NphDVec3 corner[8], u32 rgba)
{
    for (u32 axis = 1U; axis <= 4U; axis <<= 1U) {
        for (u32 c = 0U; c < 8U; ++c) {
            if ((c & axis) == 0U) {
                segment(d, corner[c], corner[c | axis], rgba);
            }
        }
    }
}

// A circle of radius r about `centre` in the plane of unit vectors u, v,
// from centre + r u, as RING_SEGMENTS chords.
static void ring(const NphPhysicsDebugDraw *d, NphDVec3 base, NphVec3 centre, NphVec3 u, NphVec3 v,
                 f32 r, u32 rgba)
{
    NphVec3 from = v3_mul_add(centre, u, r);
    for (u32 i = 1U; i <= RING_SEGMENTS; ++i) {
        const f32 angle = 6.2831853F * (f32)i / (f32)RING_SEGMENTS;
        const NphVec3 to =
            v3_add(centre, v3_add(v3_scale(u, r * cosf(angle)), v3_scale(v, r * sinf(angle))));
        segment_at(d, base, from, to, rgba);
        from = to;
    }
}

// A triangle's edges in corner order, inactive ones (bit k of `active`
// clear for edge k -> k + 1) dimmed.
static void triangle_edges(const NphPhysicsDebugDraw *d, NphDVec3 base, NphPhysicsPose pose,
                           const NphVec3 t[3], u32 active, u32 rgba)
{
    for (u32 k = 0U; k < 3U; ++k) {
        const NphVec3 p = pose_point(pose, t[k]);
        const NphVec3 q = pose_point(pose, t[k == 2U ? 0U : k + 1U]);
        segment_at(d, base, p, q, ((active >> k) & 1U) != 0U ? rgba : (rgba & INACTIVE_EDGE_MASK));
    }
}

void nph_physics_geometry_draw(const NphPhysicsDebugDraw *d, const NphPhysicsGeometry *g,
                               NphDVec3 base, NphPhysicsPose pose, u32 rgba)
{
    if (d == NULL || d->line == NULL || g == NULL) {
        return;
    }
    const NphVec3 ax = q_axis_x(pose.rotation);
    const NphVec3 ay = q_axis_y(pose.rotation);
    const NphVec3 az = q_axis_z(pose.rotation);
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        ring(d, base, pose.position, ax, ay, g->sphere.radius, rgba);
        ring(d, base, pose.position, ay, az, g->sphere.radius, rgba);
        ring(d, base, pose.position, az, ax, g->sphere.radius, rgba);
        break;
    case NPH_PHYSICS_SHAPE_CAPSULE: {
        // Both end circles across the axis, both along it, then four
        // straight sides.
        const f32 r = g->capsule.radius;
        const NphVec3 half = v3_scale(ay, g->capsule.half_height);
        const NphVec3 end[2] = {v3_sub(pose.position, half), v3_add(pose.position, half)};
        ring(d, base, end[0], ax, az, r, rgba);
        ring(d, base, end[1], ax, az, r, rgba);
        ring(d, base, end[0], ax, ay, r, rgba);
        ring(d, base, end[1], ax, ay, r, rgba);
        const NphVec3 side[4] = {ax, v3_neg(ax), az, v3_neg(az)};
        for (u32 k = 0U; k < 4U; ++k) {
            segment_at(d, base, v3_mul_add(end[0], side[k], r), v3_mul_add(end[1], side[k], r),
                       rgba);
        }
        break;
    }
    case NPH_PHYSICS_SHAPE_BOX: {
        const NphVec3 h = g->box.half_extents;
        NphDVec3 corner[8];
        for (u32 c = 0U; c < 8U; ++c) {
            const NphVec3 local = v3((c & 1U) != 0U ? h.x : -h.x, (c & 2U) != 0U ? h.y : -h.y,
                                     (c & 4U) != 0U ? h.z : -h.z);
            corner[c] = dv3_add_v3(base, pose_point(pose, local));
        }
        box_edges(d, corner, rgba);
        break;
    }
    case NPH_PHYSICS_SHAPE_HULL: {
        // Each edge once: the half-edges come in twin pairs.
        const NphVec3 *v = hull_vertices(g->hull);
        const NphPhysicsHullEdge *e = hull_edges(g->hull);
        for (u32 i = 0U; i + 1U < g->hull->edge_count; i += 2U) {
            segment_at(d, base, pose_point(pose, v[e[i].origin]),
                       pose_point(pose, v[e[i + 1U].origin]), rgba);
        }
        break;
    }
    case NPH_PHYSICS_SHAPE_MESH: {
        const u8 *active = mesh_edges(g->mesh);
        for (u32 t = 0U; t < g->mesh->triangle_count; ++t) {
            NphVec3 corners[3];
            mesh_corners(g->mesh, t, corners);
            triangle_edges(d, base, pose, corners, active[t], rgba);
        }
        break;
    }
    default:
        break; // heightfields are not built here
    }
}

static u32 body_rgba(const HBody *b, u32 flags)
{
    if (b->type == (u32)NPH_PHYSICS_BODY_STATIC) {
        return RGBA_STATIC;
    }
    if (b->type == (u32)NPH_PHYSICS_BODY_KINEMATIC) {
        return RGBA_KINEMATIC;
    }
    return (b->flags & HB_AWAKE) == 0U && (flags & NPH_PHYSICS_DRAW_SLEEP) != 0U ? RGBA_ASLEEP
                                                                                 : RGBA_AWAKE;
}

// A shape's outline at its pose relative to its body's origin
// (shape.c:62-67 with the body's own origin: the body's offset is +0).
static void shape_outline(const NphPhysicsDebugDraw *d, const NphBodyRec *b, const HShape *s,
                          u32 rgba)
{
    NphPhysicsPose pose;
    pose.position = v3_add(v3_zero(), q_rotate(b->rotation, s->pose.position));
    pose.rotation = q_mul(b->rotation, s->pose.rotation);
    nph_physics_geometry_draw(d, &s->geometry, b->origin, pose, rgba);
}

static void fat_box(const NphPhysicsDebugDraw *d, NphAABB box)
{
    NphDVec3 corner[8];
    for (u32 c = 0U; c < 8U; ++c) {
        corner[c] = dv3((c & 1U) != 0U ? (f64)box.max.x : (f64)box.min.x,
                        (c & 2U) != 0U ? (f64)box.max.y : (f64)box.min.y,
                        (c & 4U) != 0U ? (f64)box.max.z : (f64)box.min.z);
    }
    box_edges(d, corner, RGBA_FAT_BOX);
}

static void mass_cross(const NphPhysicsDebugDraw *d, NphDVec3 c)
{
    segment(d, dv3(c.x + -MASS_ARM, c.y + 0.0, c.z + 0.0), dv3(c.x + MASS_ARM, c.y + 0.0, c.z + 0.0),
            RGBA_MASS);
    segment(d, dv3(c.x + 0.0, c.y + -MASS_ARM, c.z + 0.0), dv3(c.x + 0.0, c.y + MASS_ARM, c.z + 0.0),
            RGBA_MASS);
    segment(d, dv3(c.x + 0.0, c.y + 0.0, c.z + -MASS_ARM), dv3(c.x + 0.0, c.y + 0.0, c.z + MASS_ARM),
            RGBA_MASS);
}

// A joint: each body's centre to its anchor, the anchors' gap, and each
// frame's x axis from its anchor.
static void joint_lines(const NphPhysicsDebugDraw *d, const NphBodyRec *a, const NphBodyRec *b,
                        const HJoint *j)
{
    const NphDVec3 pa = dv3_add_v3(a->origin, q_rotate(a->rotation, j->frame_a.position));
    const NphDVec3 pb = dv3_add_v3(b->origin, q_rotate(b->rotation, j->frame_b.position));
    segment(d, a->center, pa, RGBA_ANCHOR);
    segment(d, b->center, pb, RGBA_ANCHOR);
    segment(d, pa, pb, RGBA_JOINT_ERROR);
    const NphVec3 xa = q_axis_x(q_mul(a->rotation, j->frame_a.rotation));
    const NphVec3 xb = q_axis_x(q_mul(b->rotation, j->frame_b.rotation));
    segment(d, pa, dv3_add_v3(pa, v3_scale(xa, 0.2F)), RGBA_FRAME_A);
    segment(d, pb, dv3_add_v3(pb, v3_scale(xb, 0.15F)), RGBA_FRAME_B);
}

void nph_physics_world_draw(const NphPhysicsWorld *cw, const NphPhysicsDebugDraw *d)
{
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    if (w == NULL || d == NULL || d->line == NULL || w->stepping ||
        world_sync_bodies(w) != NPH_RESULT_OK) {
        return;
    }
    const HBody *hb = VB(w->bodies, HBody);
    const NphBodyRec *rec = VB(w->body_recs, NphBodyRec);
    const HShape *shapes = VB(w->shapes, HShape);
    for (u32 i = 0U; i < w->bodies.len; ++i) {
        if (hb[i].handle == 0U) {
            continue;
        }
        const u32 rgba = body_rgba(&hb[i], d->flags);
        for (u32 s = hb[i].shape_head; s != NPH_NULL32; s = shapes[s].next) {
            if ((d->flags & NPH_PHYSICS_DRAW_SHAPES) != 0U) {
                const u32 f = shapes[s].flags;
                shape_outline(d, &rec[i], &shapes[s],
                              (f & HS_SENSOR) != 0U  ? NPH_PHYSICS_DRAW_SENSOR_RGBA
                              : (f & HS_GHOST) != 0U ? NPH_PHYSICS_DRAW_GHOST_RGBA
                                                     : rgba);
            }
            if ((d->flags & NPH_PHYSICS_DRAW_AABBS) != 0U) {
                fat_box(d, shapes[s].fat_aabb);
            }
        }
        if ((d->flags & NPH_PHYSICS_DRAW_MASS) != 0U &&
            hb[i].type == (u32)NPH_PHYSICS_BODY_DYNAMIC) {
            mass_cross(d, rec[i].center);
        }
    }
    if ((d->flags & NPH_PHYSICS_DRAW_JOINTS) != 0U) {
        const HJoint *joints = VB(w->joints, HJoint);
        for (u32 i = 0U; i < w->joints.len; ++i) {
            if (joints[i].handle != 0U) {
                joint_lines(d, &rec[joints[i].body_a], &rec[joints[i].body_b], &joints[i]);
            }
        }
    }
    if ((d->flags & NPH_PHYSICS_DRAW_CONTACTS) != 0U && world_read_contacts(w) == NPH_RESULT_OK) {
        const HContact *hc = VB(w->contacts, HContact);
        const NphContactRec *cr = VB(w->contact_view, NphContactRec);
        for (u32 i = 0U; i < w->contacts.len; ++i) {
            if (hc[i].shape_a == NPH_NULL32 || (hc[i].flags & HC_TOUCHING) == 0U) {
                continue;
            }
            const NphDVec3 centre = rec[hc[i].body_a].center;
            for (u32 p = 0U; p < cr[i].point_count; ++p) {
                const NphDVec3 at = dv3_add_v3(centre, cr[i].points[p].anchor_a);
                segment(d, at, dv3_add_v3(at, v3_scale(cr[i].normal, 0.1F)),
                        cr[i].points[p].separation > 0.0F ? RGBA_SPECULATIVE : RGBA_PENETRATING);
            }
        }
    }
}
