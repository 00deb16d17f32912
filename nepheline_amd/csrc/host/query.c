// SPDX-License-Identifier: Apache-2.0
// World queries (engine/physics/src/query.c; query.h): what a segment, a
// moving convex shape, a placed one or a box meets among a world's shapes.
//
// A query walks the broadphase's trees on the host (static, kinematic,
// then dynamic shapes, as the reference orders them) and tests each shape
// it finds exactly, in a frame centred on the query's origin, so that a
// query keeps its precision anywhere in a world of 64-bit positions.  A
// cast walks a tree along its segment (the nearer child first, every box
// grown by the moving shape's extent and the rounding) and clips the search
// by each answer; an overlap walks it by box.  The bodies' poses and the
// shapes' tight boxes are the device's, read once per step at most.
// collide_shape gathers its candidate pairs in walk order, runs all of
// them in one narrow-phase launch, and reports them in that order.
#include "world_impl.h"

#include <float.h>

#define CAST_REACH NPH_PHYSICS_LINEAR_SLOP // physics_internal.h:175
#define QUERY_STACK 1024U                  // aabb_tree.h:50

NphPhysicsQueryFilter nph_physics_query_filter_default(void)
{
    NphPhysicsQueryFilter f;
    memset(&f, 0, sizeof f);
    f.category = 1U;
    f.mask = ~0ULL;
    f.body_types = NPH_PHYSICS_QUERY_ALL_BODIES;
    return f;
}

// The trees in walk order, with the kind of body each holds.
static const u32 WALK[3][2] = {
    {NPH_TREE_STATIC, NPH_PHYSICS_QUERY_STATIC},
    {NPH_TREE_KINEMATIC, NPH_PHYSICS_QUERY_KINEMATIC},
    {NPH_TREE_DYNAMIC, NPH_PHYSICS_QUERY_DYNAMIC},
};

typedef struct Query {
    NphPhysicsWorld *w;
    NphPhysicsQueryFilter filter;
    NphDVec3 origin; // the frame shapes are tested in
    const NphBodyRec *bodies;
    const NphShapeRec *shapes; // tight boxes
    int ended;
} Query;

static NphResult shape_view(NphPhysicsWorld *w)
{
    if (w->shape_view_ok) {
        return NPH_RESULT_OK;
    }
    NphResult r = vbuf_reserve(w->alloc, &w->shape_view, sizeof(NphShapeRec),
                               w->shapes.len > 0U ? w->shapes.len : 1U);
    if (r == NPH_RESULT_OK) {
        r = nph_dev_read_shapes(w->dev, VB(w->shape_view, NphShapeRec), w->shapes.len);
    }
    w->shape_view_ok = r == NPH_RESULT_OK;
    return r;
}

static int query_begin(Query *q, const NphPhysicsWorld *cw, const NphPhysicsQueryFilter *filter,
                       NphDVec3 origin)
{
    NphPhysicsWorld *w = (NphPhysicsWorld *)cw;
    if (w == NULL || w->stepping || world_sync_bodies(w) != NPH_RESULT_OK ||
        shape_view(w) != NPH_RESULT_OK) {
        return 0;
    }
    q->w = w;
    q->filter = filter != NULL ? *filter : nph_physics_query_filter_default();
    q->origin = origin;
    q->bodies = VB(w->body_recs, NphBodyRec);
    q->shapes = VB(w->shape_view, NphShapeRec);
    q->ended = 0;
    return 1;
}

static int walks(const Query *q, u32 k)
{
    return !q->ended && (q->filter.body_types & WALK[k][1]) != 0U;
}

// query.c:56-73: the leaf's shape when the filter passes it.
static const HShape *admit(const Query *q, u64 leaf, u32 *slot)
{
    const HShape *s = &VB(q->w->shapes, HShape)[(u32)leaf];
    const HBody *b = &VB(q->w->bodies, HBody)[s->body];
    if ((s->filter.category & q->filter.mask) == 0U ||
        (s->filter.mask & q->filter.category) == 0U || b->handle == q->filter.ignore.value ||
        ((s->flags & HS_SENSOR) != 0U && q->filter.sensors == 0)) {
        return NULL;
    }
    *slot = (u32)leaf;
    return s;
}

static u32 body_handle(const Query *q, const HShape *s)
{
    return VB(q->w->bodies, HBody)[s->body].handle;
}

// The shape's rotation, and its origin less the query's in 64 bits.
static void shape_frame(const Query *q, const HShape *s, NphQuat *rotation, NphDVec3 *offset)
{
    const NphBodyRec *b = &q->bodies[s->body];
    *rotation = q_mul(b->rotation, s->pose.rotation);
    const NphDVec3 at = dv3_add_v3(b->origin, q_rotate(b->rotation, s->pose.position));
    *offset = dv3(at.x - q->origin.x, at.y - q->origin.y, at.z - q->origin.z);
}

static NphDVec3 origin_in_shape(NphQuat rotation, NphDVec3 offset)
{
    return drotate_inverse(rotation, dv3(-offset.x, -offset.y, -offset.z));
}

static NphVec3 narrow(NphDVec3 v) { return v3((f32)v.x, (f32)v.y, (f32)v.z); }

static f32 max_abs(NphVec3 v) { return nph_maxf(nph_absf(v.x), nph_maxf(nph_absf(v.y), nph_absf(v.z))); }

// What a 32-bit stand-in's box test adds to stay conservative.
static f32 rounding(NphVec3 at, NphVec3 translation)
{
    return 1e-6F * (max_abs(at) + max_abs(translation)) + 1e-6F;
}

static int dfinite(NphDVec3 v)
{
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

static int vfinite(NphVec3 v) { return nph_isfinitef(v.x) && nph_isfinitef(v.y) && nph_isfinitef(v.z); }

static int shape_input_ok(const NphPhysicsGeometry *g, NphDVec3 position, NphQuat rotation)
{
    return g != NULL && (u32)g->type <= (u32)NPH_PHYSICS_SHAPE_HULL &&
           nph_physics_geometry_is_valid(g) && dfinite(position) &&
           nph_absf(q_len_sq(rotation) - 1.0F) <= 2.0F * 1e-3F;
}

// ---------------------------------------------------------------------------
// The segment walk (aabb_tree.c:562-632): every box grown by `extent`,
// entered by the segment from `at` along `d` before `max`; the nearer
// child first.  A leaf's answer below zero ends the walk; below `max`
// clips it.
// ---------------------------------------------------------------------------
typedef f32 (*SegmentLeafFn)(void *ctx, u64 leaf);

// geometry.c:128-166 (math): the slab test of o + t d, t in [0, t_max].
static int segment_enters(NphVec3 o, NphVec3 d, NphAABB b, f32 t_max, f32 *t_enter)
{
    f32 lo_t = 0.0F;
    f32 hi_t = t_max;
    for (u32 axis = 0U; axis < 3U; ++axis) {
        const f32 oa = v3_at(o, axis);
        const f32 da = v3_at(d, axis);
        const f32 lo = v3_at(b.min, axis);
        const f32 hi = v3_at(b.max, axis);
        if (da == 0.0F) {
            if (oa < lo || oa > hi) {
                return 0;
            }
            continue;
        }
        const f32 inv = 1.0F / da;
        f32 t1 = (lo - oa) * inv;
        f32 t2 = (hi - oa) * inv;
        if (t1 > t2) {
            const f32 x = t1;
            t1 = t2;
            t2 = x;
        }
        lo_t = nph_maxf(lo_t, t1);
        hi_t = nph_minf(hi_t, t2);
        if (lo_t > hi_t) {
            return 0;
        }
    }
    if (t_enter != NULL) {
        *t_enter = lo_t;
    }
    return 1;
}

static NphAABB grown(NphAABB b, NphVec3 e) { return aabb_make(v3_sub(b.min, e), v3_add(b.max, e)); }

static void segment_walk(const NphTree *t, NphVec3 at, NphVec3 d, f32 max, NphVec3 extent, u64 mask,
                         SegmentLeafFn leaf, void *ctx)
{
    if (t->root == NPH_TREE_NULL) {
        return;
    }
    u32 stack[QUERY_STACK];
    u32 top = 0U;
    stack[top++] = t->root;
    while (top > 0U) {
        const u32 i = stack[--top];
        const NphTreeNode *n = &t->nodes[i];
        if ((n->category_bits & mask) == 0U || !segment_enters(at, d, grown(n->aabb, extent), max, NULL)) {
            continue;
        }
        if (n->height == 0U) {
            const f32 v = leaf(ctx, tree_user_data(t, i));
            if (v < 0.0F) {
                return;
            }
            if (v < max) {
                max = v;
            }
            continue;
        }
        f32 t1 = FLT_MAX;
        f32 t2 = FLT_MAX;
        const int h1 = segment_enters(at, d, grown(t->nodes[n->child1].aabb, extent), max, &t1);
        const int h2 = segment_enters(at, d, grown(t->nodes[n->child2].aabb, extent), max, &t2);
        if (top + 2U > QUERY_STACK) {
            continue;
        }
        if (h1 && h2) {
            const int first = t1 <= t2;
            stack[top++] = first ? n->child2 : n->child1;
            stack[top++] = first ? n->child1 : n->child2;
        } else if (h1) {
            stack[top++] = n->child1;
        } else if (h2) {
            stack[top++] = n->child2;
        }
    }
}

// ---------------------------------------------------------------------------
// Casts (query.c:120-345)
// ---------------------------------------------------------------------------
typedef struct Cast {
    Query q;
    const NphPhysicsGeometry *g; // NULL: a segment
    NphQuat rotation;
    NphVec3 translation;
    f32 margin;
    f32 max;
    NphPhysicsQueryCastFn fn;
    void *user;
} Cast;

static f32 report(Cast *c, const HShape *s, const NphPhysicsCastHit *local)
{
    NphPhysicsQueryHit hit;
    hit.shape.value = s->handle;
    hit.body.value = body_handle(&c->q, s);
    hit.fraction = local->fraction;
    hit.point = dv3_add_v3(c->q.origin, local->point);
    hit.normal = local->normal;
    hit.triangle = local->triangle;
    const f32 r = c->fn(c->user, &hit);
    if (r < 0.0F) {
        c->q.ended = 1;
        return -1.0F;
    }
    c->max = nph_minf(c->max, r);
    return c->max;
}

static f32 ray_leaf(void *ctx, u64 leaf)
{
    Cast *c = (Cast *)ctx;
    u32 slot;
    const HShape *s = admit(&c->q, leaf, &slot);
    if (s == NULL) {
        return c->max;
    }
    NphQuat rotation;
    NphDVec3 offset;
    shape_frame(&c->q, s, &rotation, &offset);
    NphPhysicsCastHit local;
    if (!raycast_local(&s->geometry, narrow(origin_in_shape(rotation, offset)),
                       q_rotate_inv(rotation, c->translation), &local) ||
        local.fraction > c->max) {
        return c->max;
    }
    local.point = v3_scale(c->translation, local.fraction);
    local.normal = q_rotate(rotation, local.normal);
    return report(c, s, &local);
}

static f32 cast_leaf(void *ctx, u64 leaf)
{
    Cast *c = (Cast *)ctx;
    u32 slot;
    const HShape *s = admit(&c->q, leaf, &slot);
    if (s == NULL || s->geometry.type == NPH_PHYSICS_SHAPE_MESH) {
        return c->max; // casts against meshes are not built here
    }
    NphQuat rotation;
    NphDVec3 offset;
    shape_frame(&c->q, s, &rotation, &offset);
    NphPhysicsCastHit local;
    if (!cast_convex(c->g, pose_make(v3_zero(), c->rotation), c->translation, c->margin,
                     &s->geometry, pose_make(narrow(offset), rotation), c->max, &local)) {
        return c->max;
    }
    return report(c, s, &local);
}

static void cast_walk(Cast *c, NphVec3 at, NphVec3 extent, SegmentLeafFn leaf)
{
    for (u32 k = 0U; k < 3U; ++k) {
        if (walks(&c->q, k)) {
            segment_walk(&c->q.w->trees[WALK[k][0]], at, c->translation, c->max, extent,
                         c->q.filter.mask, leaf, c);
        }
    }
}

typedef struct Closest {
    NphPhysicsQueryHit hit;
    int found;
} Closest;

static f32 keep_closest(void *user, const NphPhysicsQueryHit *hit)
{
    Closest *c = (Closest *)user;
    if (!c->found || hit->fraction < c->hit.fraction) {
        c->hit = *hit;
        c->found = 1;
    }
    return c->hit.fraction > 0.0F ? c->hit.fraction : -1.0F; // nothing comes before the start
}

void nph_physics_world_raycast(const NphPhysicsWorld *world, NphDVec3 origin, NphVec3 translation,
                               const NphPhysicsQueryFilter *filter, NphPhysicsQueryCastFn fn,
                               void *user)
{
    Cast c;
    if (fn == NULL || !dfinite(origin) || !vfinite(translation) ||
        !(v3_len_sq(translation) > 0.0F) || !query_begin(&c.q, world, filter, origin)) {
        return;
    }
    c.g = NULL;
    c.rotation = q_identity();
    c.translation = translation;
    c.margin = 0.0F;
    c.max = 1.0F;
    c.fn = fn;
    c.user = user;
    const NphVec3 at = narrow(origin);
    cast_walk(&c, at, v3_splat(rounding(at, translation)), ray_leaf);
}

bool nph_physics_world_raycast_closest(const NphPhysicsWorld *world, NphDVec3 origin,
                                       NphVec3 translation, const NphPhysicsQueryFilter *filter,
                                       NphPhysicsQueryHit *out)
{
    Closest c;
    memset(&c, 0, sizeof c);
    nph_physics_world_raycast(world, origin, translation, filter, keep_closest, &c);
    if (c.found && out != NULL) {
        *out = c.hit;
    }
    return c.found != 0;
}

void nph_physics_world_cast_shape_margin(const NphPhysicsWorld *world,
                                         const NphPhysicsGeometry *geometry, NphDVec3 position,
                                         NphQuat rotation, NphVec3 translation, f32 margin,
                                         const NphPhysicsQueryFilter *filter,
                                         NphPhysicsQueryCastFn fn, void *user)
{
    Cast c;
    if (fn == NULL || !shape_input_ok(geometry, position, rotation) || !vfinite(translation) ||
        !(v3_len_sq(translation) > 0.0F) || !(margin >= 0.0F) || !nph_isfinitef(margin) ||
        !query_begin(&c.q, world, filter, position)) {
        return;
    }
    c.g = geometry;
    c.rotation = q_normalize(rotation);
    c.translation = translation;
    c.margin = margin;
    c.max = 1.0F;
    c.fn = fn;
    c.user = user;
    // The shape's box swept along the translation, grown by the narrow
    // phase's reach, the margin and the rounding.
    const NphAABB rel = nph_physics_geometry_aabb(geometry, pose_make(v3_zero(), c.rotation));
    const NphVec3 mid = v3_scale(v3_add(rel.min, rel.max), 0.5F);
    const NphVec3 half = v3_scale(v3_sub(rel.max, rel.min), 0.5F);
    const NphVec3 at = narrow(dv3_add_v3(position, mid));
    const NphVec3 extent =
        v3_add(half, v3_splat(CAST_REACH + margin + rounding(at, translation)));
    cast_walk(&c, at, extent, cast_leaf);
}

void nph_physics_world_cast_shape(const NphPhysicsWorld *world, const NphPhysicsGeometry *geometry,
                                  NphDVec3 position, NphQuat rotation, NphVec3 translation,
                                  const NphPhysicsQueryFilter *filter, NphPhysicsQueryCastFn fn,
                                  void *user)
{
    nph_physics_world_cast_shape_margin(world, geometry, position, rotation, translation, 0.0F,
                                        filter, fn, user);
}

bool nph_physics_world_cast_shape_margin_closest(const NphPhysicsWorld *world,
                                                 const NphPhysicsGeometry *geometry,
                                                 NphDVec3 position, NphQuat rotation,
                                                 NphVec3 translation, f32 margin,
                                                 const NphPhysicsQueryFilter *filter,
                                                 NphPhysicsQueryHit *out)
{
    Closest c;
    memset(&c, 0, sizeof c);
    nph_physics_world_cast_shape_margin(world, geometry, position, rotation, translation, margin,
                                        filter, keep_closest, &c);
    if (c.found && out != NULL) {
        *out = c.hit;
    }
    return c.found != 0;
}

bool nph_physics_world_cast_shape_closest(const NphPhysicsWorld *world,
                                          const NphPhysicsGeometry *geometry, NphDVec3 position,
                                          NphQuat rotation, NphVec3 translation,
                                          const NphPhysicsQueryFilter *filter,
                                          NphPhysicsQueryHit *out)
{
    return nph_physics_world_cast_shape_margin_closest(world, geometry, position, rotation,
                                                       translation, 0.0F, filter, out);
}

// ---------------------------------------------------------------------------
// Overlaps and contacts (query.c:347-566)
// ---------------------------------------------------------------------------
typedef struct Shapes {
    Query q;
    const NphPhysicsGeometry *g;
    NphQuat rotation;
    NphAABB box; // the query's, in the trees' frame
    f32 speculative;
    NphPhysicsQueryShapeFn fn;
    NphVecBuf pairs;   // collide_shape's candidate pairs (NphAmdCollideQuery)
    NphVecBuf targets; // and their shape slots
    NphResult err;
    void *user;
} Shapes;

static int found(Shapes *sq, const HShape *s)
{
    const NphPhysicsShape shape = {s->handle};
    const NphPhysicsBody body = {body_handle(&sq->q, s)};
    if (!sq->fn(sq->user, shape, body)) {
        sq->q.ended = 1;
        return 0;
    }
    return 1;
}

static int tight_overlaps(const Shapes *sq, u32 slot)
{
    return aabb_overlaps(sq->q.shapes[slot].aabb, sq->box);
}

static int overlap_leaf(u32 proxy, u64 leaf, void *ctx)
{
    (void)proxy;
    Shapes *sq = (Shapes *)ctx;
    u32 slot;
    const HShape *s = admit(&sq->q, leaf, &slot);
    if (s == NULL || !tight_overlaps(sq, slot)) {
        return 1;
    }
    NphQuat rotation;
    NphDVec3 offset;
    shape_frame(&sq->q, s, &rotation, &offset);
    int touch;
    if (s->geometry.type == NPH_PHYSICS_SHAPE_MESH) {
        touch = overlap_mesh_local(sq->g, q_mul(q_conj(rotation), sq->rotation),
                                   origin_in_shape(rotation, offset), s->geometry.mesh);
    } else {
        touch = overlap_convex(sq->g, pose_make(v3_zero(), sq->rotation), &s->geometry,
                               pose_make(narrow(offset), rotation));
    }
    return !touch || found(sq, s);
}

static int aabb_leaf(u32 proxy, u64 leaf, void *ctx)
{
    (void)proxy;
    Shapes *sq = (Shapes *)ctx;
    u32 slot;
    const HShape *s = admit(&sq->q, leaf, &slot);
    return s == NULL || !tight_overlaps(sq, slot) || found(sq, s);
}

static int collide_leaf(u32 proxy, u64 leaf, void *ctx)
{
    (void)proxy;
    Shapes *sq = (Shapes *)ctx;
    u32 slot;
    const HShape *s = admit(&sq->q, leaf, &slot);
    if (s == NULL || s->geometry.type == NPH_PHYSICS_SHAPE_MESH) {
        return 1; // a mesh's manifolds are not built here
    }
    NphQuat rotation;
    NphDVec3 offset;
    shape_frame(&sq->q, s, &rotation, &offset);
    NphAmdCollideQuery p;
    memset(&p, 0, sizeof p);
    p.a = *sq->g;
    p.pose_a = pose_make(v3_zero(), sq->rotation);
    p.b = s->geometry;
    p.pose_b = pose_make(narrow(offset), rotation);
    p.speculative_distance = sq->speculative;
    sq->err = vbuf_push(sq->q.w->alloc, &sq->pairs, sizeof p, &p);
    if (sq->err == NPH_RESULT_OK) {
        sq->err = vbuf_push(sq->q.w->alloc, &sq->targets, 4U, &slot);
    }
    return sq->err == NPH_RESULT_OK;
}

static void shape_walk(Shapes *sq, NphAABB box, NphTreeQueryFn leaf)
{
    for (u32 k = 0U; k < 3U; ++k) {
        if (walks(&sq->q, k)) {
            tree_query(&sq->q.w->trees[WALK[k][0]], box, sq->q.filter.mask, leaf, sq);
        }
    }
}

static int shapes_begin(Shapes *sq, const NphPhysicsWorld *world, const NphPhysicsGeometry *g,
                        NphDVec3 position, NphQuat rotation, const NphPhysicsQueryFilter *filter,
                        void *user)
{
    memset(sq, 0, sizeof *sq);
    if (!query_begin(&sq->q, world, filter, position)) {
        return 0;
    }
    sq->g = g;
    sq->rotation = q_normalize(rotation);
    sq->user = user;
    return 1;
}

void nph_physics_world_overlap_shape(const NphPhysicsWorld *world,
                                     const NphPhysicsGeometry *geometry, NphDVec3 position,
                                     NphQuat rotation, const NphPhysicsQueryFilter *filter,
                                     NphPhysicsQueryShapeFn fn, void *user)
{
    Shapes sq;
    if (fn == NULL || !shape_input_ok(geometry, position, rotation) ||
        !shapes_begin(&sq, world, geometry, position, rotation, filter, user)) {
        return;
    }
    sq.fn = fn;
    sq.box = nph_world_box(position,
                           nph_physics_geometry_aabb(geometry, pose_make(v3_zero(), sq.rotation)));
    shape_walk(&sq, sq.box, overlap_leaf);
}

void nph_physics_world_overlap_aabb(const NphPhysicsWorld *world, NphAABB box,
                                    const NphPhysicsQueryFilter *filter, NphPhysicsQueryShapeFn fn,
                                    void *user)
{
    Shapes sq;
    if (fn == NULL || !vfinite(box.min) || !vfinite(box.max) ||
        !shapes_begin(&sq, world, NULL, dv3(0.0, 0.0, 0.0), q_identity(), filter, user)) {
        return;
    }
    sq.fn = fn;
    sq.box = box;
    shape_walk(&sq, box, aabb_leaf);
}

void nph_physics_world_collide_shape(const NphPhysicsWorld *world,
                                     const NphPhysicsGeometry *geometry, NphDVec3 position,
                                     NphQuat rotation, f32 speculative_distance,
                                     const NphPhysicsQueryFilter *filter,
                                     NphPhysicsQueryContactFn fn, void *user)
{
    Shapes sq;
    if (fn == NULL || !shape_input_ok(geometry, position, rotation) ||
        !(speculative_distance >= 0.0F) || !nph_isfinitef(speculative_distance) ||
        !shapes_begin(&sq, world, geometry, position, rotation, filter, user)) {
        return;
    }
    sq.speculative = speculative_distance;
    const NphAABB rel = nph_physics_geometry_aabb(geometry, pose_make(v3_zero(), sq.rotation));
    shape_walk(&sq, nph_world_box(position, aabb_grow(rel, speculative_distance)), collide_leaf);
    const u32 n = sq.pairs.len;
    NphPhysicsManifold *m = NULL;
    if (sq.err == NPH_RESULT_OK && n > 0U) {
        m = (NphPhysicsManifold *)malloc((size_t)n * sizeof *m);
        if (m == NULL || nph_amd_collide_many(VB(sq.pairs, NphAmdCollideQuery), n, m, NULL) !=
                             NPH_RESULT_OK) {
            free(m);
            m = NULL;
        }
    }
    for (u32 i = 0U; m != NULL && i < n; ++i) {
        if (m[i].point_count == 0U) {
            continue;
        }
        const HShape *s = &VB(sq.q.w->shapes, HShape)[VB(sq.targets, u32)[i]];
        NphPhysicsQueryContact c;
        c.shape.value = s->handle;
        c.body.value = body_handle(&sq.q, s);
        c.manifold = m[i];
        if (!fn(user, &c)) {
            break;
        }
    }
    free(m);
    vbuf_free(sq.q.w->alloc, &sq.pairs, sizeof(NphAmdCollideQuery));
    vbuf_free(sq.q.w->alloc, &sq.targets, 4U);
}

// query.c:585-605: the shape's own surface normal at a world point, the
// point and `toward` brought into the shape's frame (64 bits to its origin).
NphVec3 nph_physics_world_surface_normal(const NphPhysicsWorld *world, NphPhysicsShape shape,
                                         NphDVec3 point, u32 triangle, NphVec3 toward)
{
    const HShape *s = world != NULL ? world_shape(world, shape) : NULL;
    const int finite = __builtin_isfinite(point.x) && __builtin_isfinite(point.y) &&
                       __builtin_isfinite(point.z) &&
                       nph_isfinitef(toward.x) && nph_isfinitef(toward.y) && nph_isfinitef(toward.z);
    Query q;
    if (s == NULL || !finite || !query_begin(&q, world, NULL, point)) {
        return toward;
    }
    NphQuat rotation;
    NphDVec3 offset;
    shape_frame(&q, s, &rotation, &offset);
    const NphVec3 n = nph_physics_geometry_surface_normal(
        &s->geometry, narrow(origin_in_shape(rotation, offset)), triangle,
        q_rotate_inv(rotation, toward));
    return q_rotate(rotation, n);
}
