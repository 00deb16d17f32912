// SPDX-License-Identifier: Apache-2.0
// Continuous collision (engine/physics/src/continuous.c:1-409): a body that
// moved far in a step is swept over it, its centre along a line and its
// rotation by slerp, against the static shapes (bullets also against
// kinematic and dynamic bodies at their final poses), and set back to the
// first time any comes within a slop of it; conservative advancement on GJK
// distances with a bound on rotation.
//
// The device marks the fast bodies in finalize and leaves their boxes
// alone; the sweeps run here, on the host, between finalize and the grown
// proxies, because the sweep's slerp takes the C library's acosf and sinf,
// whose bits the device cannot reproduce, and because the trees they query
// live on the host.  Fast bodies are few (the bodies that moved farther in a
// step than half their thinnest shape); each sweep reads only static shapes
// (or, for bullets, every body at its final pose) and writes only its own
// body, so the sweeps run in parallel and the results do not depend on
// their order.
#define _POSIX_C_SOURCE 200809L
#include "world_impl.h"

#include <math.h>

#define TOI_TARGET NPH_PHYSICS_LINEAR_SLOP
#define TOI_TOLERANCE (0.25F * NPH_PHYSICS_LINEAR_SLOP)
#define TOI_STEPS 32U
#define INSET_FRACTION 0.5F
#define CORE_FRACTION 0.25F

// ---------------------------------------------------------------------------
// Quaternion interpolation (quat.h:214-218, quat.c:74-93) with the C
// library's acosf and sinf, as the reference calls them.
// ---------------------------------------------------------------------------
static NphQuat qneg(NphQuat q) { return q4(-q.x, -q.y, -q.z, -q.w); }
static f32 qdot(NphQuat a, NphQuat b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

static NphQuat nlerp(NphQuat a, NphQuat b, f32 t)
{
    const NphQuat bb = qdot(a, b) < 0.0F ? qneg(b) : b;
    return q_normalize(q_add(q_scale(a, 1.0F - t), q_scale(bb, t)));
}

static NphQuat slerp(NphQuat a, NphQuat b, f32 t)
{
    f32 cos_theta = qdot(a, b);
    NphQuat bb = b;
    if (cos_theta < 0.0F) {
        bb = qneg(b);
        cos_theta = -cos_theta;
    }
    if (cos_theta > 0.9995F) {
        return nlerp(a, bb, t);
    }
    const f32 theta = acosf(cos_theta < -1.0F ? -1.0F : (cos_theta > 1.0F ? 1.0F : cos_theta));
    const f32 inv_sin = 1.0F / sinf(theta);
    const f32 wa = sinf((1.0F - t) * theta) * inv_sin;
    const f32 wb = sinf(t * theta) * inv_sin;
    return q_add(q_scale(a, wa), q_scale(bb, wb));
}

static f32 quat_angle(NphQuat q) { return 2.0F * atan2f(v3_len(q_xyz(q)), nph_absf(q.w)); }

// ---------------------------------------------------------------------------
// The sweep (continuous.c:64-219)
// ---------------------------------------------------------------------------
typedef struct Sweep {
    NphVec3 dc;
    NphQuat q0;
    NphQuat q1;
    f32 spin;
    f32 reach;
} Sweep;

static void sweep_at(const Sweep *s, f32 t, NphVec3 *centre, NphQuat *rotation)
{
    *centre = v3_scale(s->dc, t);
    *rotation = t <= 0.0F ? s->q0 : (t >= 1.0F ? s->q1 : slerp(s->q0, s->q1, t));
}

typedef struct Mover {
    NphProxy proxy;
    NphVec3 storage[8];
    NphProxy inset;
    NphProxy core;
    NphVec3 core_point;
    NphVec3 local;
    NphQuat local_q;
} Mover;

static void mover_init(Mover *mv, const NphBodyRec *b, const HShape *s)
{
    mv->proxy = proxy_of(&s->geometry, mv->storage);
    const NphGeom g = world_host_geom(s);
    const f32 inner = geom_inner_radius(&g);
    mv->inset = mv->proxy;
    mv->inset.radius = mv->proxy.radius - INSET_FRACTION * inner;
    mv->core_point =
        s->geometry.type == NPH_PHYSICS_SHAPE_HULL ? s->geometry.hull->centroid : v3_zero();
    mv->core.points = &mv->core_point;
    mv->core.count = 1U;
    mv->core.radius = CORE_FRACTION * inner;
    mv->local = v3_sub(s->pose.position, b->local_center);
    mv->local_q = s->pose.rotation;
}

static NphPhysicsPose mover_at(const Mover *mv, const Sweep *s, f32 t)
{
    NphVec3 c;
    NphQuat q;
    sweep_at(s, t, &c, &q);
    return pose_make(v3_add(c, q_rotate(q, mv->local)), q_mul(q, mv->local_q));
}

static f32 advance(const Mover *mv, const NphProxy *m, const Sweep *s, const NphProxy *t,
                   NphPhysicsPose pose_t, f32 limit, NphVec3 *axis)
{
    const f32 radius = m->radius + t->radius;
    NphSimplexCache cache;
    memset(&cache, 0, sizeof cache);
    f32 alpha = 0.0F;
    for (u32 step = 0U; step < TOI_STEPS; ++step) {
        const NphPhysicsPose pm = mover_at(mv, s, alpha);
        const NphPhysicsPose at_m = pose_make(v3_zero(), pm.rotation);
        const NphPhysicsPose at_t = pose_make(v3_sub(pose_t.position, pm.position), pose_t.rotation);
        NphProxyDistance pd;
        proxy_distance(t, at_t, m, at_m, 0, &cache, &pd);
        const NphDistance d = pd.d;
        const f32 separation = d.distance - radius;
        if (!(d.distance > 0.0F)) {
            return alpha;
        }
        *axis = d.normal;
        if (separation <= TOI_TARGET + TOI_TOLERANCE) {
            *axis = proxy_contact_axis(m, at_m, t, at_t, &cache, d.normal);
            return alpha;
        }
        const f32 closing = -v3_dot(s->dc, d.normal) + s->spin;
        if (!(closing > 0.0F)) {
            return -1.0F;
        }
        alpha += (separation - TOI_TARGET) / closing;
        if (!(alpha < limit)) {
            return -1.0F;
        }
    }
    return alpha;
}

static f32 gap_at(const Mover *mv, const NphProxy *m, const Sweep *s, const NphProxy *t,
                  NphPhysicsPose pose_t, NphVec3 axis, f32 alpha)
{
    const NphPhysicsPose pm = mover_at(mv, s, alpha);
    return proxy_gap(m, pose_make(v3_zero(), pm.rotation), t,
                     pose_make(v3_sub(pose_t.position, pm.position), pose_t.rotation), axis);
}

static f32 shape_toi(const Mover *mv, const Sweep *s, const NphProxy *t, NphPhysicsPose pose_t,
                     f32 limit)
{
    NphVec3 axis = v3(0.0F, 1.0F, 0.0F);
    const f32 f = advance(mv, &mv->proxy, s, t, pose_t, limit, &axis);
    if (f > 0.0F) {
        const f32 angle = s->reach > 0.0F ? s->spin / s->reach : 0.0F;
        const f32 bow = 0.125F * s->spin * angle;
        const f32 low = nph_minf(gap_at(mv, &mv->proxy, s, t, pose_t, axis, 0.0F),
                                 gap_at(mv, &mv->proxy, s, t, pose_t, axis, 1.0F));
        return low - bow >= -NPH_PHYSICS_LINEAR_SLOP ? -1.0F : f;
    }
    if (f < 0.0F) {
        return f;
    }
    const f32 g = advance(mv, &mv->inset, s, t, pose_t, limit, &axis);
    if (g != 0.0F) {
        return g > 0.0F ? g : -1.0F;
    }
    const f32 h = advance(mv, &mv->core, s, t, pose_t, limit, &axis);
    return h > 0.0F ? h : -1.0F;
}

// ---------------------------------------------------------------------------
// A fast body (continuous.c:299-409)
// ---------------------------------------------------------------------------
typedef struct BodySweep {
    NphPhysicsWorld *w;
    const NphBodyRec *b; // at the step's end
    u32 slot;
    const NphBodyRec *records; // every body's record (targets)
    const u8 *bullet;          // per body slot: a bullet, as the step began
    NphDVec3 start;
    Sweep sweep;
    f32 best;
} BodySweep;

static void sweep_target(BodySweep *bs, u32 target_body, const HShape *ts)
{
    NphPhysicsWorld *w = bs->w;
    const HBody *tb = &VB(w->bodies, HBody)[target_body];
    if (target_body == bs->slot ||
        (tb->type != (u32)NPH_PHYSICS_BODY_STATIC && bs->bullet[target_body] != 0U) ||
        !joints_allow_contact(w, bs->slot, target_body)) {
        return;
    }
    NphVec3 storage[8];
    const NphProxy tp = proxy_of(&ts->geometry, storage);
    NphShapeRec trec;
    memset(&trec, 0, sizeof trec);
    trec.pose = ts->pose;
    const NphPhysicsPose pose_t = shape_pose_relative(&bs->records[target_body], &trec, bs->start);
    const HBody *hb = &VB(w->bodies, HBody)[bs->slot];
    for (u32 s = hb->shape_head; s != NPH_NULL32; s = VB(w->shapes, HShape)[s].next) {
        const HShape *ms = &VB(w->shapes, HShape)[s];
        if (!nph_physics_filter_should_collide(ms->filter, ts->filter)) {
            continue;
        }
        Mover mv;
        mover_init(&mv, bs->b, ms);
        const f32 f = shape_toi(&mv, &bs->sweep, &tp, pose_t, bs->best);
        if (f >= 0.0F && f < bs->best) {
            bs->best = f;
        }
    }
}

static int sweep_leaf(u32 proxy, u64 leaf, void *context)
{
    (void)proxy;
    BodySweep *bs = (BodySweep *)context;
    const HShape *ts = &VB(bs->w->shapes, HShape)[(u32)leaf];
    sweep_target(bs, ts->body, ts);
    return 1;
}

// The body's sweep; on an impact its record is set back (the boxes are the
// caller's).  Returns 1 on an impact.
static int sweep_body(NphPhysicsWorld *w, const NphFastBody *f, NphBodyRec *b,
                      const NphBodyRec *records, const u8 *bullet, int bullets)
{
    BodySweep bs;
    bs.w = w;
    bs.b = b;
    bs.slot = f->slot;
    bs.records = records;
    bs.bullet = bullet;
    const NphVec3 dp = sb_dp(&f->sb);
    const NphQuat dq = sb_dq(&f->sb);
    bs.start = dv3_add_v3(b->center, v3_neg(dp));
    bs.sweep.dc = dp;
    bs.sweep.q0 = f->params.q0;
    bs.sweep.q1 = b->rotation;
    bs.sweep.reach = b->max_extent;
    bs.sweep.spin = quat_angle(dq) * b->max_extent;
    bs.best = 1.0F;
    u64 mask = 0U;
    const HBody *hb = &VB(w->bodies, HBody)[f->slot];
    for (u32 s = hb->shape_head; s != NPH_NULL32; s = VB(w->shapes, HShape)[s].next) {
        mask |= VB(w->shapes, HShape)[s].filter.mask;
    }
    const f32 grow = b->max_extent + TOI_TARGET + TOI_TOLERANCE;
    NphAABB rel;
    rel.min = v3_sub(v3_min(v3_zero(), dp), v3_splat(grow));
    rel.max = v3_add(v3_max(v3_zero(), dp), v3_splat(grow));
    const NphAABB box = nph_world_box(bs.start, rel);
    tree_query(&w->trees[NPH_TREE_STATIC], box, mask, sweep_leaf, &bs);
    if (bullets) {
        tree_query(&w->trees[NPH_TREE_KINEMATIC], box, mask, sweep_leaf, &bs);
        tree_query(&w->trees[NPH_TREE_DYNAMIC], box, mask, sweep_leaf, &bs);
    }
    if (!(bs.best < 1.0F)) {
        return 0;
    }
    NphVec3 c;
    NphQuat q;
    sweep_at(&bs.sweep, bs.best, &c, &q);
    b->center = dv3_add_v3(bs.start, c);
    b->rotation = q_normalize(q);
    b->origin = dv3_add_v3(b->center, v3_neg(q_rotate(b->rotation, b->local_center)));
    return 1;
}

// The body's shapes' boxes at pose `b` over the fat boxes in `recs`
// (solver.c:351-369), into `recs`.
static void refresh(NphBodyRec *b, NphShapeRec *recs, u32 n)
{
    for (u32 i = 0U; i < n; ++i) {
        NphShapeRec *shape = &recs[i];
        shape->aabb = shape_world_aabb(b, shape);
        const NphAABB fat = shape->fat_aabb;
        const NphAABB box = shape->aabb;
        const int inside = fat.min.x <= box.min.x && fat.min.y <= box.min.y &&
                           fat.min.z <= box.min.z && box.max.x <= fat.max.x &&
                           box.max.y <= fat.max.y && box.max.z <= fat.max.z;
        if (!inside) {
            shape->fat_aabb = aabb_grow(box, NPH_PHYSICS_AABB_MARGIN);
            shape->flags |= NPH_SHAPE_ENLARGED;
            b->enlarged = 1U;
        }
    }
}

// One fast body: its shapes' records as the device holds them after
// finalize (a bullet's boxes refreshed there at its end pose), the sweep,
// the refresh, and the writes.
static NphResult finish_body(NphPhysicsWorld *w, const NphFastBody *f, NphBodyRec *b,
                             const NphBodyRec *records, const u8 *bullet, int is_bullet,
                             u32 *impacts)
{
    NphShapeRec recs[64];
    u32 n = 0U;
    const HBody *hb = &VB(w->bodies, HBody)[f->slot];
    for (u32 s = hb->shape_head; s != NPH_NULL32 && n < 64U; s = VB(w->shapes, HShape)[s].next) {
        NphResult r = world_shape_record(w, s, &recs[n]);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        n += 1U;
    }
    if (n < hb->shape_count) {
        return NPH_RESULT_UNSUPPORTED; // more shapes than a fast body's sweep holds
    }
    if (is_bullet) {
        refresh(b, recs, n); // as finalize did on the device
    }
    if (sweep_body(w, f, b, records, bullet, is_bullet)) {
        *impacts += 1U;
        refresh(b, recs, n);
    } else if (!is_bullet) {
        refresh(b, recs, n);
    }
    NphResult r = nph_dev_write_body(w->dev, f->slot, b);
    u32 i = 0U;
    for (u32 s = hb->shape_head; s != NPH_NULL32 && r == NPH_RESULT_OK && i < n;
         s = VB(w->shapes, HShape)[s].next, ++i) {
        w->shape_view_ok = 0;
        r = nph_dev_write_shape(w->dev, s, &recs[i]);
    }
    return r;
}

NphResult continuous_sweep(NphPhysicsWorld *w, const NphFastBody *fast, u32 count)
{
    // Bullets as the step began, by body slot.
    NphVecBuf flags = {NULL, 0U, 0U};
    NphResult r = vbuf_reserve(w->alloc, &flags, 1U, w->bodies.len > 0U ? w->bodies.len : 1U);
    if (r != NPH_RESULT_OK) {
        return r;
    }
    u8 *bullet = VB(flags, u8);
    memset(bullet, 0, w->bodies.len);
    for (u32 i = 0U; i < w->awake_bodies.len; ++i) {
        const u32 slot = VB(w->awake_bodies, u32)[i];
        bullet[slot] = (VB(w->body_recs, NphBodyRec)[slot].flags & NPH_BODY_BULLET) != 0U ? 1U : 0U;
    }
    // Static targets: their records never change on the device.
    const NphBodyRec *records = VB(w->body_recs, NphBodyRec);
    NphBodyRec *bodies = (NphBodyRec *)malloc((size_t)count * sizeof(NphBodyRec));
    if (bodies == NULL) {
        vbuf_free(w->alloc, &flags, 1U);
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    u32 impacts = 0U;
    u32 bullets = 0U;
    for (u32 k = 0U; k < count && r == NPH_RESULT_OK; ++k) {
        bodies[k] = fast[k].body;
        if ((fast[k].params.flags & NPH_PARAM_BULLET) != 0U) {
            bullets += 1U;
            continue;
        }
        r = finish_body(w, &fast[k], &bodies[k], records, bullet, 0, &impacts);
    }
    if (r == NPH_RESULT_OK && bullets > 0U) {
        // Against every body at its final pose: the device's records with the
        // sweeps above written.
        w->bodies_stale = 1;
        r = world_sync_bodies(w);
        records = VB(w->body_recs, NphBodyRec);
        for (u32 k = 0U; k < count && r == NPH_RESULT_OK; ++k) {
            if ((fast[k].params.flags & NPH_PARAM_BULLET) != 0U) {
                r = finish_body(w, &fast[k], &bodies[k], records, bullet, 1, &impacts);
            }
        }
    }
    w->stats.swept_bodies = count;
    w->stats.impact_bodies = impacts;
    free(bodies);
    vbuf_free(w->alloc, &flags, 1U);
    return r;
}
