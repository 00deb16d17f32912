// SPDX-License-Identifier: Apache-2.0
// Continuous collision (engine/physics/src/continuous.c:1-409): a body that
// moved far in a step is swept back over the step against what it could
// have passed through, and set back to the first moment it comes within a
// slop of any of it.
//
// Where it runs. The device marks the fast bodies in finalize and leaves
// their boxes alone; the sweeps run here, between finalize and the grown
// proxies, for two reasons: the motion's rotation is a slerp through the C
// library's acosf and sinf, whose bits the device does not reproduce, and
// the trees the sweeps query live on the host. Fast bodies are few, each
// sweep reads only static shapes (bullets: every body at its final pose)
// and writes only its own body, so the results do not depend on the order.
//
// How a sweep is organised here:
//  * a Motion is the body's step as a function of the fraction t: its
//    centre moves along a line, its rotation turns by slerp, and `arc`
//    bounds how fast a point of its shapes moves because of the turn;
//  * a Probe is one of the body's shapes in three erosion stages: the shape
//    itself, its points with a radius smaller by half the inner radius, and
//    a ball of a quarter of the inner radius at its centroid. A stage that
//    starts the step already within reach of the target hands over to the
//    next, thinner one (the contacts push out what touches at the start);
//  * each stage advances conservatively: the GJK gap, less the target gap,
//    over the fastest the gap can close, until the gap is within tolerance
//    (Mirtich's conservative advancement);
//  * an impact of the full shape that the step's contacts already caught
//    (the gap along the impact axis stays within a slop at both ends of the
//    step, less the bow of the arc) is no impact;
//  * static targets are convex shapes or the triangles of a mesh, taken in
//    the mesh's frame (a triangle the body does not approach from its front
//    face is skipped).
#define _POSIX_C_SOURCE 200809L
#include "world_impl.h"

#include <math.h>

#define TOI_TARGET NPH_PHYSICS_LINEAR_SLOP
#define TOI_TOLERANCE (0.25F * NPH_PHYSICS_LINEAR_SLOP)
#define TOI_STEPS 32U
#define INSET_FRACTION 0.5F
#define CORE_FRACTION 0.25F
#define TRIANGLE_APPROACH NPH_PHYSICS_LINEAR_SLOP

// ---------------------------------------------------------------------------
// The motion (quat.h:214-218 and quat.c:74-93 for the slerp, with the C
// library's acosf and sinf, as the reference calls them)
// ---------------------------------------------------------------------------
static f32 q4_dot(NphQuat a, NphQuat b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

static NphQuat q4_blend(NphQuat a, f32 wa, NphQuat b, f32 wb)
{
    return q_add(q_scale(a, wa), q_scale(b, wb));
}

static NphQuat turn_between(NphQuat a, NphQuat b, f32 t)
{
    f32 c = q4_dot(a, b);
    if (c < 0.0F) { // q and -q: the shorter arc
        b = q4(-b.x, -b.y, -b.z, -b.w);
        c = -c;
    }
    if (c > 0.9995F) { // nearly parallel: the normalized blend
        const NphQuat bb = q4_dot(a, b) < 0.0F ? q4(-b.x, -b.y, -b.z, -b.w) : b;
        return q_normalize(q4_blend(a, 1.0F - t, bb, t));
    }
    const f32 theta = acosf(c < -1.0F ? -1.0F : (c > 1.0F ? 1.0F : c));
    const f32 over_sin = 1.0F / sinf(theta);
    return q4_blend(a, sinf((1.0F - t) * theta) * over_sin, b, sinf(t * theta) * over_sin);
}

typedef struct Motion {
    NphVec3 shift;      // the centre's displacement over the step
    NphQuat from, to;   // the rotation at the step's start and end
    f32 arc;            // the turn's angle times the reach
    f32 reach;          // how far the body's shapes reach from its centre
} Motion;

static NphQuat motion_turn(const Motion *m, f32 t)
{
    if (t <= 0.0F) {
        return m->from;
    }
    return t >= 1.0F ? m->to : turn_between(m->from, m->to, t);
}

// The frame at fraction t of something fixed on the body at `local` (about
// the centre of mass) turned by `local_q`.
static NphPhysicsPose motion_frame(const Motion *m, f32 t, NphVec3 local, NphQuat local_q)
{
    const NphQuat q = motion_turn(m, t);
    return pose_make(v3_add(v3_scale(m->shift, t), q_rotate(q, local)), q_mul(q, local_q));
}

// The motion seen in a frame turned by `inv` (a surface's).
static Motion motion_in(const Motion *m, NphQuat inv)
{
    Motion r = *m;
    r.shift = q_rotate(inv, m->shift);
    r.from = q_mul(inv, m->from);
    r.to = q_mul(inv, m->to);
    return r;
}

// ---------------------------------------------------------------------------
// A shape of the swept body, in its three erosion stages
// ---------------------------------------------------------------------------
enum { STAGE_SHAPE = 0, STAGE_INSET, STAGE_CORE, STAGES };

typedef struct Probe {
    NphProxy stage[STAGES];
    NphVec3 storage[8];
    NphVec3 centroid; // the core ball's centre (the hull's centroid, or the origin)
    NphVec3 local;    // the shape's place about the centre of mass
    NphQuat local_q;
} Probe;

static void probe_make(Probe *p, const NphBodyRec *b, const HShape *s)
{
    const NphGeom g = world_host_geom(s);
    const f32 inner = geom_inner_radius(&g);
    p->stage[STAGE_SHAPE] = proxy_of(&s->geometry, p->storage);
    p->stage[STAGE_INSET] = p->stage[STAGE_SHAPE];
    p->stage[STAGE_INSET].radius = p->stage[STAGE_SHAPE].radius - INSET_FRACTION * inner;
    p->centroid =
        s->geometry.type == NPH_PHYSICS_SHAPE_HULL ? s->geometry.hull->centroid : v3_zero();
    p->stage[STAGE_CORE].points = &p->centroid;
    p->stage[STAGE_CORE].count = 1U;
    p->stage[STAGE_CORE].radius = CORE_FRACTION * inner;
    p->local = v3_sub(s->pose.position, b->local_center);
    p->local_q = s->pose.rotation;
}

// A still target: its proxy and its pose in the motion's frame.
typedef struct Target {
    const NphProxy *proxy;
    NphPhysicsPose pose;
} Target;

// The GJK gap between stage `k` of the probe at fraction t and the target,
// measured in a frame at the probe's place (as the casts measure it).
static NphDistance gap_probe(const Probe *p, u32 k, const Motion *m, const Target *tg, f32 t,
                             NphSimplexCache *cache, NphPhysicsPose *at_p, NphPhysicsPose *at_t)
{
    const NphPhysicsPose f = motion_frame(m, t, p->local, p->local_q);
    *at_p = pose_make(v3_zero(), f.rotation);
    *at_t = pose_make(v3_sub(tg->pose.position, f.position), tg->pose.rotation);
    NphProxyDistance pd;
    proxy_distance(tg->proxy, *at_t, &p->stage[k], *at_p, 0, cache, &pd);
    return pd.d;
}

// Conservative advancement of stage `k` towards the target, short of
// `limit`: the fraction where the gap first comes within the target
// tolerance (0 when it starts there), or -1 when it does not before
// `limit` (or the gap never closes); `axis` gets the axis between them
// there, from the target towards the probe.
static f32 advance_stage(const Probe *p, u32 k, const Motion *m, const Target *tg, f32 limit,
                         NphVec3 *axis)
{
    const f32 reach = p->stage[k].radius + tg->proxy->radius;
    NphSimplexCache cache;
    memset(&cache, 0, sizeof cache);
    f32 t = 0.0F;
    for (u32 i = 0U; i < TOI_STEPS; ++i) {
        NphPhysicsPose at_p, at_t;
        const NphDistance d = gap_probe(p, k, m, tg, t, &cache, &at_p, &at_t);
        if (!(d.distance > 0.0F)) {
            return t; // the cores touch: no better place to stop
        }
        const f32 gap = d.distance - reach;
        *axis = d.normal;
        if (gap <= TOI_TARGET + TOI_TOLERANCE) {
            *axis = proxy_contact_axis(&p->stage[k], at_p, tg->proxy, at_t, &cache, d.normal);
            return t;
        }
        // The gap closes no faster than the centre comes along the normal
        // plus the fastest a point turns.
        const f32 closing = -v3_dot(m->shift, d.normal) + m->arc;
        if (!(closing > 0.0F)) {
            return -1.0F;
        }
        t += (gap - TOI_TARGET) / closing;
        if (!(t < limit)) {
            return -1.0F;
        }
    }
    return t; // still approaching, grazing: the last safe fraction
}

// The shape's gap to the target along `axis` at fraction t.
static f32 gap_along(const Probe *p, const Motion *m, const Target *tg, NphVec3 axis, f32 t)
{
    const NphPhysicsPose f = motion_frame(m, t, p->local, p->local_q);
    return proxy_gap(&p->stage[STAGE_SHAPE], pose_make(v3_zero(), f.rotation), tg->proxy,
                     pose_make(v3_sub(tg->pose.position, f.position), tg->pose.rotation), axis);
}

// The step's contacts already hold an impact whose gap along its axis is
// within a slop at both ends of the step once the arc's bow (reach angle^2
// / 8) is taken off: the shape never passed through the target.
static bool impact_caught(const Probe *p, const Motion *m, const Target *tg, NphVec3 axis)
{
    const f32 angle = m->reach > 0.0F ? m->arc / m->reach : 0.0F;
    const f32 bow = 0.125F * m->arc * angle;
    const f32 low = nph_minf(gap_along(p, m, tg, axis, 0.0F), gap_along(p, m, tg, axis, 1.0F));
    return low - bow >= -NPH_PHYSICS_LINEAR_SLOP;
}

// The probe's time of impact with the target before `limit`, or -1.
static f32 probe_toi(const Probe *p, const Motion *m, const Target *tg, f32 limit)
{
    NphVec3 axis = v3(0.0F, 1.0F, 0.0F);
    for (u32 k = STAGE_SHAPE; k < STAGES; ++k) {
        const f32 t = advance_stage(p, k, m, tg, limit, &axis);
        if (t == 0.0F && k + 1U < STAGES) {
            continue; // within reach at the start: the thinner stage decides
        }
        if (k == STAGE_SHAPE && t > 0.0F && impact_caught(p, m, tg, axis)) {
            return -1.0F;
        }
        return t > 0.0F ? t : -1.0F;
    }
    return -1.0F;
}

// ---------------------------------------------------------------------------
// Mesh targets (continuous.c:211-300): the motion in the mesh's frame, the
// triangles under the box the body's reach sweeps
// ---------------------------------------------------------------------------
static f32 mesh_toi(const Probe *p, const Motion *world, NphDVec3 start, const NphBodyRec *mb,
                    const HShape *ms, f32 limit)
{
    const NphPhysicsMesh *mesh = ms->geometry.mesh;
    const NphQuat qs = q_mul(mb->rotation, ms->pose.rotation);
    const NphDVec3 ps = dv3_add_v3(mb->origin, q_rotate(mb->rotation, ms->pose.position));
    const NphDVec3 origin = drotate_inverse(qs, dv3(start.x - ps.x, start.y - ps.y, start.z - ps.z));
    const Motion m = motion_in(world, q_conj(qs));
    const NphVec3 at = v3((f32)origin.x, (f32)origin.y, (f32)origin.z);
    const NphVec3 aa = v3_abs(at);
    const f32 grow = world->reach + TOI_TARGET + TOI_TOLERANCE +
                     2.4e-7F * nph_maxf(aa.x, nph_maxf(aa.y, aa.z));
    const NphVec3 end = v3_add(at, m.shift);
    const NphAABB box = aabb_make(v3_sub(v3_min(at, end), v3_splat(grow)),
                                  v3_add(v3_max(at, end), v3_splat(grow)));
    const int two_sided = (mesh->flags & NPH_PHYSICS_MESH_TWO_SIDED) != 0U;
    // The centroid's place at both ends of the step (constant per probe).
    const NphPhysicsPose f0 = motion_frame(&m, 0.0F, p->local, p->local_q);
    const NphPhysicsPose f1 = motion_frame(&m, 1.0F, p->local, p->local_q);
    const NphVec3 c0 = v3_add(q_rotate(f0.rotation, p->centroid), f0.position);
    const NphVec3 c1 = v3_add(q_rotate(f1.rotation, p->centroid), f1.position);
    f32 best = limit;
    NphBoxWalk walk;
    box_walk_begin(&walk, mesh, box);
    for (u32 t = box_walk_next(&walk); t != NPH_NULL32; t = box_walk_next(&walk)) {
        NphVec3 v[3];
        mesh_corners(mesh, t, v);
        for (u32 k = 0U; k < 3U; ++k) {
            v[k] = relative_vertex(v[k], origin);
        }
        NphVec3 n = v3_cross(v3_sub(v[1], v[0]), v3_sub(v[2], v[0]));
        const f32 len = v3_len(n);
        if (!(len > 0.0F)) {
            continue;
        }
        n = v3_scale(n, 1.0F / len);
        // Heights of the centroid over the plane at the ends: a triangle
        // the body is behind (one-sided) or does not come towards is not
        // one it could tunnel through.
        f32 h0 = v3_dot(n, v3_sub(c0, v[0]));
        f32 h1 = v3_dot(n, v3_sub(c1, v[0]));
        if (two_sided && h0 < 0.0F) {
            h0 = -h0;
            h1 = -h1;
        }
        if (h0 < -NPH_PHYSICS_LINEAR_SLOP || !(h0 - h1 > TRIANGLE_APPROACH)) {
            continue;
        }
        const NphProxy tri = {v, 3U, 0.0F};
        const Target tg = {&tri, pose_identity()};
        const f32 toi = probe_toi(p, &m, &tg, best);
        if (toi >= 0.0F && toi < best) {
            best = toi;
        }
    }
    return best < limit ? best : -1.0F;
}

// ---------------------------------------------------------------------------
// One fast body (continuous.c:302-409)
// ---------------------------------------------------------------------------
typedef struct Sweeper {
    NphPhysicsWorld *w;
    const NphBodyRec *b;       // the body at the step's end
    u32 slot;
    const NphBodyRec *records; // every body's record (the targets)
    const u8 *bullet;          // per body slot: a bullet as the step began
    NphDVec3 start;            // the centre of mass at the step's start
    Motion motion;             // world axes, about `start`
    f32 best;                  // the first impact so far (1: none)
} Sweeper;

static void sweep_against(Sweeper *sw, const HShape *ts)
{
    NphPhysicsWorld *w = sw->w;
    const u32 target = ts->body;
    const HBody *tb = &VB(w->bodies, HBody)[target];
    // Not its own shapes, nor a bullet's (bullets do not sweep against each
    // other), nor a body a joint keeps it from touching.
    if (target == sw->slot ||
        (tb->type != (u32)NPH_PHYSICS_BODY_STATIC && sw->bullet[target] != 0U) ||
        !joints_allow_contact(w, sw->slot, target)) {
        return;
    }
    const int mesh = ts->geometry.type == NPH_PHYSICS_SHAPE_MESH;
    NphVec3 storage[8];
    NphProxy tp = {storage, 1U, 0.0F};
    Target tg = {&tp, pose_identity()};
    if (!mesh) {
        tp = proxy_of(&ts->geometry, storage);
        NphShapeRec rec;
        memset(&rec, 0, sizeof rec);
        rec.pose = ts->pose;
        tg.pose = shape_pose_relative(&sw->records[target], &rec, sw->start);
    }
    const HBody *hb = &VB(w->bodies, HBody)[sw->slot];
    for (u32 s = hb->shape_head; s != NPH_NULL32; s = VB(w->shapes, HShape)[s].next) {
        const HShape *ms = &VB(w->shapes, HShape)[s];
        if (((ms->flags | ts->flags) & (HS_SENSOR | HS_GHOST)) != 0U ||
            !nph_physics_filter_should_collide(ms->filter, ts->filter)) {
            continue;
        }
        Probe p;
        probe_make(&p, sw->b, ms);
        const f32 toi = mesh ? mesh_toi(&p, &sw->motion, sw->start, &sw->records[target], ts,
                                        sw->best)
                             : probe_toi(&p, &sw->motion, &tg, sw->best);
        if (toi >= 0.0F && toi < sw->best) {
            sw->best = toi;
        }
    }
}

static int sweep_leaf(u32 proxy, u64 leaf, void *context)
{
    (void)proxy;
    Sweeper *sw = (Sweeper *)context;
    sweep_against(sw, &VB(sw->w->shapes, HShape)[(u32)leaf]);
    return 1;
}

// The body's sweep; on an impact its record is set back to it (the boxes
// are the caller's). Returns 1 on an impact.
static int sweep_body(NphPhysicsWorld *w, const NphFastBody *f, NphBodyRec *b,
                      const NphBodyRec *records, const u8 *bullet, int bullets)
{
    const NphVec3 dp = sb_dp(&f->sb);
    Sweeper sw;
    sw.w = w;
    sw.b = b;
    sw.slot = f->slot;
    sw.records = records;
    sw.bullet = bullet;
    sw.start = dv3_add_v3(b->center, v3_neg(dp));
    sw.motion.shift = dp;
    sw.motion.from = f->params.q0;
    sw.motion.to = b->rotation;
    sw.motion.reach = b->max_extent;
    const NphQuat dq = sb_dq(&f->sb);
    sw.motion.arc = 2.0F * atan2f(v3_len(q_xyz(dq)), nph_absf(dq.w)) * b->max_extent;
    sw.best = 1.0F;
    // Targets in a category some shape of the body accepts (each pair then
    // checks both filters), under the box the body's reach sweeps.
    u64 mask = 0U;
    const HBody *hb = &VB(w->bodies, HBody)[f->slot];
    for (u32 s = hb->shape_head; s != NPH_NULL32; s = VB(w->shapes, HShape)[s].next) {
        mask |= VB(w->shapes, HShape)[s].filter.mask;
    }
    const f32 grow = b->max_extent + TOI_TARGET + TOI_TOLERANCE;
    NphAABB rel;
    rel.min = v3_sub(v3_min(v3_zero(), dp), v3_splat(grow));
    rel.max = v3_add(v3_max(v3_zero(), dp), v3_splat(grow));
    const NphAABB box = nph_world_box(sw.start, rel);
    tree_query(&w->trees[NPH_TREE_STATIC], box, mask, sweep_leaf, &sw);
    if (bullets) {
        tree_query(&w->trees[NPH_TREE_KINEMATIC], box, mask, sweep_leaf, &sw);
        tree_query(&w->trees[NPH_TREE_DYNAMIC], box, mask, sweep_leaf, &sw);
    }
    if (!(sw.best < 1.0F)) {
        return 0;
    }
    // Back to the first impact: the next step's contacts take it from there.
    const NphQuat q = motion_turn(&sw.motion, sw.best);
    b->center = dv3_add_v3(sw.start, v3_scale(sw.motion.shift, sw.best));
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
    // The records on the stack for up to 64 shapes, else in the world's
    // scratch list.
    NphShapeRec few[64];
    const HBody *hb = &VB(w->bodies, HBody)[f->slot];
    NphShapeRec *recs = few;
    if (hb->shape_count > 64U) {
        const NphResult g = vbuf_reserve(w->alloc, &w->sweep_shapes, sizeof(NphShapeRec),
                                         hb->shape_count);
        if (g != NPH_RESULT_OK) {
            return g;
        }
        recs = VB(w->sweep_shapes, NphShapeRec);
    }
    u32 n = 0U;
    for (u32 s = hb->shape_head; s != NPH_NULL32; s = VB(w->shapes, HShape)[s].next) {
        NphResult r = world_shape_record(w, s, &recs[n]);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        n += 1U;
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
