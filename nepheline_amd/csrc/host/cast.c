// SPDX-License-Identifier: Apache-2.0
// One geometry against a segment, a moving convex shape or a placed one
// (engine/physics/src/cast.c; collide.h:240-318), and the GJK glue that the
// continuous sweeps share: proxies, distances with poses and a simplex
// cache, the gap along an axis and the contact axis of two polytopes.
//
//  - Rays in the geometry's frame, one closed form a type: a ball and a
//    capsule's side by the closest approach then Pythagoras (no
//    cancellation for a far origin), a box by its slabs, a hull by its face
//    planes (Haines 1991); a mesh by its own ray query.  A ray that starts
//    inside a convex shape does not hit it.
//  - Shape casts by conservative advancement on GJK distances between the
//    cores (Mirtich 1996): each step moves the mover as far as the plane
//    between the closest points allows, in a frame at the mover's place,
//    until the rounded shapes are within NPH_PHYSICS_CAST_TOUCH; a last
//    step along the contact normal closes what is left of the gap.
//  - Overlaps by GJK on the rounded shapes (a mesh: on its triangles).
#include "world_impl.h"

#include <float.h>

#define ADVANCE_STEPS 32U
#define WELL_APART 1e-4F     // a core distance at which GJK's normal is well defined
#define NORMAL_BACKOFF 1e-3F // metres back along the motion for a polytope pair's normal
#define CAST_TOUCH 1e-4F     // physics_internal.h:168
#define CAST_SLIDE 1e-5F     // physics_internal.h:172
#define CAST_REACH NPH_PHYSICS_LINEAR_SLOP

// ---------------------------------------------------------------------------
// Proxies and distances
// ---------------------------------------------------------------------------

// geometry.c:446-508: a convex geometry as points and a radius (a box's
// corners in bit order x, y, z).
NphProxy proxy_of(const NphPhysicsGeometry *g, NphVec3 storage[8])
{
    NphProxy p;
    p.points = storage;
    p.radius = 0.0F;
    if (g->type == NPH_PHYSICS_SHAPE_SPHERE) {
        storage[0] = v3_zero();
        p.count = 1U;
        p.radius = g->sphere.radius;
    } else if (g->type == NPH_PHYSICS_SHAPE_CAPSULE) {
        const f32 h = g->capsule.half_height;
        storage[0] = v3(0.0F, -h, 0.0F);
        storage[1] = v3(0.0F, h, 0.0F);
        p.count = h > 0.0F ? 2U : 1U;
        p.radius = g->capsule.radius;
    } else if (g->type == NPH_PHYSICS_SHAPE_BOX) {
        const NphVec3 h = g->box.half_extents;
        for (u32 i = 0U; i < 8U; ++i) {
            storage[i] = v3((i & 1U) != 0U ? h.x : -h.x, (i & 2U) != 0U ? h.y : -h.y,
                            (i & 4U) != 0U ? h.z : -h.z);
        }
        p.count = 8U;
    } else { // a hull
        p.points = hull_vertices(g->hull);
        p.count = g->hull->vertex_count;
    }
    return p;
}

static u32 support(const NphProxy *p, NphVec3 d) { return points_support(p->points, p->count, d); }

// gjk.c:306-374 on the GJK stepper: supports searched in each proxy's own
// frame, vertices posed into the common one; the simplex starts from the
// cache when it names valid points, and the cache keeps the final one.
void proxy_gjk(const NphProxy *a, NphPhysicsPose pa, const NphProxy *b, NphPhysicsPose pb,
               NphSimplexCache *cache, NphGjk *g)
{
    gjk_reset(g);
    if (cache != NULL && cache->count >= 1U && cache->count <= 4U) {
        for (u32 i = 0U; i < cache->count; ++i) {
            const u32 ia = cache->index_a[i];
            const u32 ib = cache->index_b[i];
            if (ia >= a->count || ib >= b->count) {
                gjk_reset(g);
                break;
            }
            gjk_seed(g, ia, pose_point(pa, a->points[ia]), ib, pose_point(pb, b->points[ib]));
        }
    }
    if (g->count == 0U) {
        gjk_seed(g, 0U, pose_point(pa, a->points[0]), 0U, pose_point(pb, b->points[0]));
    }
    gjk_begin(g);
    while (!g->done) {
        const u32 ia = support(a, q_rotate_inv(pa.rotation, v3_neg(g->dir)));
        const u32 ib = support(b, q_rotate_inv(pb.rotation, g->dir));
        gjk_feed(g, ia, pose_point(pa, a->points[ia]), ib, pose_point(pb, b->points[ib]));
    }
    if (cache != NULL) {
        cache->count = g->count;
        for (u32 i = 0U; i < 4U; ++i) {
            cache->index_a[i] = i < g->count ? (u8)g->v[i].ia : 0U;
            cache->index_b[i] = i < g->count ? (u8)g->v[i].ib : 0U;
        }
    }
}

// gjk.c:376-416: the distance of the final simplex; with `rounded`, the
// radii round the cores (the points move onto the rounded surfaces, or
// both to the middle when those touch).
void proxy_distance(const NphProxy *a, NphPhysicsPose pa, const NphProxy *b, NphPhysicsPose pb,
                    int rounded, NphSimplexCache *cache, NphProxyDistance *out)
{
    NphGjk g;
    proxy_gjk(a, pa, b, pb, cache, &g);
    gjk_finish(&g, &out->d);
    out->simplex_count = g.count;
    out->iterations = g.iterations;
    if (rounded) {
        NphDistance *d = &out->d;
        const f32 ra = a->radius;
        const f32 rb = b->radius;
        if (d->distance > ra + rb) {
            d->distance -= ra + rb;
            d->point_a = v3_mul_add(d->point_a, d->normal, ra);
            d->point_b = v3_mul_add(d->point_b, d->normal, -rb);
        } else {
            const NphVec3 mid = v3_scale(
                v3_add(v3_mul_add(d->point_a, d->normal, ra), v3_mul_add(d->point_b, d->normal, -rb)),
                0.5F);
            d->point_a = mid;
            d->point_b = mid;
            d->distance = 0.0F;
        }
    }
}

// cast.c:262-272: how far apart m and t are along n (their rounded
// extremes: m's least projection less t's greatest).
f32 proxy_gap(const NphProxy *m, NphPhysicsPose pose_m, const NphProxy *t, NphPhysicsPose pose_t,
              NphVec3 n)
{
    const u32 im = support(m, q_rotate_inv(pose_m.rotation, v3_neg(n)));
    const u32 it = support(t, q_rotate_inv(pose_t.rotation, n));
    const f32 low = v3_dot(pose_point(pose_m, m->points[im]), n) - m->radius;
    const f32 high = v3_dot(pose_point(pose_t, t->points[it]), n) + t->radius;
    return low - high;
}

// The distinct points of one side of the simplex (its proxy's indices),
// posed, in the simplex's order.
static u32 simplex_side(const NphProxy *p, NphPhysicsPose pose, const u8 *index, u32 n,
                        NphVec3 out[4])
{
    u32 seen[4];
    u32 k = 0U;
    for (u32 i = 0U; i < n && i < 4U; ++i) {
        int again = 0;
        for (u32 j = 0U; j < k; ++j) {
            again = again || seen[j] == index[i];
        }
        if (!again && index[i] < p->count) {
            seen[k] = index[i];
            out[k] = pose_point(pose, p->points[index[i]]);
            k += 1U;
        }
    }
    return k;
}

// cast.c:388-446: the exact normal of the closest features the simplex
// ended on (a face of the target against a point of the mover, a point
// against a face of the mover, or two edges), turned as `hint`; none for a
// point or an edge against a point (GJK's estimate is inside their cone).
static int feature_axis(const NphProxy *m, NphPhysicsPose pose_m, const NphProxy *t,
                        NphPhysicsPose pose_t, const NphSimplexCache *c, NphVec3 hint, NphVec3 *out)
{
    NphVec3 tp[4];
    NphVec3 mp[4];
    const u32 nt = simplex_side(t, pose_t, c->index_a, c->count, tp);
    const u32 nm = simplex_side(m, pose_m, c->index_b, c->count, mp);
    NphVec3 n;
    if (nt >= 3U && nm == 1U) {
        n = v3_cross(v3_sub(tp[1], tp[0]), v3_sub(tp[2], tp[0]));
    } else if (nm >= 3U && nt == 1U) {
        n = v3_cross(v3_sub(mp[1], mp[0]), v3_sub(mp[2], mp[0]));
    } else if (nt == 2U && nm == 2U) {
        n = v3_cross(v3_sub(tp[1], tp[0]), v3_sub(mp[1], mp[0]));
    } else {
        return 0;
    }
    const f32 len = v3_len(n);
    if (!(len > 1e-12F)) {
        return 0;
    }
    n = v3_scale(n, 1.0F / len);
    *out = v3_dot(n, hint) < 0.0F ? v3_neg(n) : n;
    return 1;
}

// cast.c:448-459: the features' axis when it parts the two at least as far
// as `hint` does, else `hint`.
NphVec3 proxy_contact_axis(const NphProxy *m, NphPhysicsPose pose_m, const NphProxy *t,
                           NphPhysicsPose pose_t, const NphSimplexCache *features, NphVec3 hint)
{
    NphVec3 n;
    if (feature_axis(m, pose_m, t, pose_t, features, hint, &n) &&
        proxy_gap(m, pose_m, t, pose_t, n) >= proxy_gap(m, pose_m, t, pose_t, hint)) {
        return n;
    }
    return hint;
}

// ---------------------------------------------------------------------------
// Rays in a geometry's frame (cast.c:24-239)
// ---------------------------------------------------------------------------

// Entry of o + t d, t in [0, 1], into the ball of radius r at the origin.
static int ray_ball(NphVec3 o, NphVec3 d, f32 r, f32 *t_out)
{
    const f32 len = v3_len(d);
    if (!(len > 0.0F)) {
        return 0;
    }
    const NphVec3 u = v3_scale(d, 1.0F / len);
    const f32 closest = -v3_dot(o, u);
    const f32 h_sq = r * r - v3_len_sq(v3_mul_add(o, u, closest));
    if (h_sq < 0.0F) {
        return 0;
    }
    const f32 t = closest - nph_sqrtf(h_sq);
    if (t < 0.0F || t > len) {
        return 0; // behind the origin: the ball is behind it, or holds it
    }
    *t_out = t / len;
    return 1;
}

// A capsule on the Y axis: its side between the end circles (the closest
// approach in XZ, then Pythagoras), then its end balls, the first entry.
static int ray_capsule(NphVec3 o, NphVec3 d, f32 hh, f32 r, f32 *t_out, NphVec3 *n_out)
{
    const f32 yc = nph_clampf(o.y, -hh, hh);
    if (o.x * o.x + (o.y - yc) * (o.y - yc) + o.z * o.z < r * r) {
        return 0; // from inside
    }
    f32 best = FLT_MAX;
    NphVec3 n = v3(0.0F, 1.0F, 0.0F);
    const f32 len = nph_sqrtf(d.x * d.x + d.z * d.z);
    if (len > 0.0F) {
        const f32 ux = d.x / len;
        const f32 uz = d.z / len;
        const f32 closest = -(o.x * ux + o.z * uz);
        const f32 cx = o.x + closest * ux;
        const f32 cz = o.z + closest * uz;
        const f32 h_sq = r * r - (cx * cx + cz * cz);
        if (h_sq < 0.0F) {
            return 0; // past the cylinder, so past the capsule
        }
        const f32 t = (closest - nph_sqrtf(h_sq)) / len;
        if (t >= 0.0F && t <= 1.0F && nph_absf(o.y + t * d.y) <= hh) {
            best = t;
            n = v3_normalize_or(v3(o.x + t * d.x, 0.0F, o.z + t * d.z), v3(1.0F, 0.0F, 0.0F));
        }
    } else if (o.x * o.x + o.z * o.z > r * r) {
        return 0; // along the axis, outside its column
    }
    for (u32 end = 0U; end < 2U; ++end) {
        const NphVec3 c = v3(0.0F, end == 0U ? -hh : hh, 0.0F);
        f32 t = 0.0F;
        if (ray_ball(v3_sub(o, c), d, r, &t) && t < best) {
            best = t;
            n = v3_normalize_or(v3_sub(v3_mul_add(o, d, t), c), v3(0.0F, 1.0F, 0.0F));
        }
    }
    if (!(best <= 1.0F)) {
        return 0;
    }
    *t_out = best;
    *n_out = n;
    return 1;
}

// The slabs: the entry is the last slab entered, the exit the first left.
static int ray_box(NphVec3 o, NphVec3 d, NphVec3 half, f32 *t_out, NphVec3 *n_out)
{
    f32 enter = -FLT_MAX;
    f32 leave = 1.0F;
    u32 axis = 3U;
    f32 side = 0.0F;
    for (u32 i = 0U; i < 3U; ++i) {
        const f32 oi = v3_at(o, i);
        const f32 di = v3_at(d, i);
        const f32 hi = v3_at(half, i);
        if (di == 0.0F) {
            if (oi < -hi || oi > hi) {
                return 0;
            }
            continue;
        }
        const f32 near = di > 0.0F ? -hi : hi;
        const f32 t0 = (near - oi) / di;
        const f32 t1 = (-near - oi) / di;
        if (t0 > enter) {
            enter = t0;
            axis = i;
            side = di > 0.0F ? -1.0F : 1.0F;
        }
        leave = nph_minf(leave, t1);
        if (enter > leave) {
            return 0;
        }
    }
    if (axis == 3U || enter < 0.0F) {
        return 0;
    }
    *t_out = enter;
    *n_out = v3(axis == 0U ? side : 0.0F, axis == 1U ? side : 0.0F, axis == 2U ? side : 0.0F);
    return 1;
}

// The face planes: the entry is the last plane crossed from outside, the
// exit the first crossed from inside.
static int ray_hull(const NphPhysicsHull *h, NphVec3 o, NphVec3 d, f32 *t_out, NphVec3 *n_out)
{
    const NphPlane *planes = hull_planes(h);
    f32 enter = -FLT_MAX;
    f32 leave = 1.0F;
    u32 face = NPH_NULL32;
    for (u32 f = 0U; f < h->face_count; ++f) {
        const f32 dist = v3_dot(planes[f].normal, o) + planes[f].d;
        const f32 den = v3_dot(planes[f].normal, d);
        if (den == 0.0F) {
            if (dist > 0.0F) {
                return 0;
            }
            continue;
        }
        const f32 t = -dist / den;
        if (den < 0.0F) {
            if (t > enter) {
                enter = t;
                face = f;
            }
        } else {
            leave = nph_minf(leave, t);
        }
        if (enter > leave) {
            return 0;
        }
    }
    if (face == NPH_NULL32 || enter < 0.0F) {
        return 0;
    }
    *t_out = enter;
    *n_out = planes[face].normal;
    return 1;
}

int raycast_local(const NphPhysicsGeometry *g, NphVec3 origin, NphVec3 delta, NphPhysicsCastHit *out)
{
    if (!(v3_len_sq(delta) > 0.0F)) {
        return 0;
    }
    f32 t = 0.0F;
    NphVec3 n = v3(0.0F, 1.0F, 0.0F);
    u32 triangle = NPH_PHYSICS_NO_TRIANGLE;
    int hit = 0;
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        hit = ray_ball(origin, delta, g->sphere.radius, &t);
        if (hit) {
            n = v3_normalize_or(v3_mul_add(origin, delta, t), v3_neg(v3_normalize(delta)));
        }
        break;
    case NPH_PHYSICS_SHAPE_CAPSULE:
        hit = ray_capsule(origin, delta, g->capsule.half_height, g->capsule.radius, &t, &n);
        break;
    case NPH_PHYSICS_SHAPE_BOX:
        hit = ray_box(origin, delta, g->box.half_extents, &t, &n);
        break;
    case NPH_PHYSICS_SHAPE_HULL:
        hit = ray_hull(g->hull, origin, delta, &t, &n);
        break;
    case NPH_PHYSICS_SHAPE_MESH: {
        NphPhysicsMeshHit mh;
        hit = nph_physics_mesh_raycast(g->mesh, origin, delta, &mh);
        if (hit) {
            t = mh.fraction;
            n = mh.normal;
            triangle = mh.original;
        }
        break;
    }
    default:
        break;
    }
    if (!hit) {
        return 0;
    }
    out->point = v3_mul_add(origin, delta, t);
    out->normal = n;
    out->fraction = t;
    out->triangle = triangle;
    return 1;
}

bool nph_physics_raycast(const NphPhysicsGeometry *g, NphPhysicsPose pose, NphVec3 origin,
                         NphVec3 delta, NphPhysicsCastHit *out)
{
    if (g == NULL || out == NULL) {
        return false;
    }
    NphPhysicsCastHit local;
    if (!raycast_local(g, pose_inv_point(pose, origin), q_rotate_inv(pose.rotation, delta),
                       &local)) {
        return false;
    }
    out->point = v3_mul_add(origin, delta, local.fraction);
    out->normal = pose_vector(pose, local.normal);
    out->fraction = local.fraction;
    out->triangle = local.triangle;
    return true;
}

// ---------------------------------------------------------------------------
// Shape casts (cast.c:262-518)
// ---------------------------------------------------------------------------

// Where conservative advancement stopped: the fraction, the poses in a
// frame at the mover's place there, the last normal and target point GJK
// gave, the closest features, whether GJK's normal was well defined.
typedef struct Arrival {
    f32 fraction;
    NphPhysicsPose mover;
    NphPhysicsPose target;
    NphVec3 normal;
    NphVec3 point;
    NphVec3 base; // the target's place less the mover's start
    f32 approach; // the last step's closing speed along the normal
    NphSimplexCache features;
    int have_point;
    int known;
} Arrival;

// GJK with the target as A, the mover as B at fraction alpha of the motion,
// in the mover's frame there.
static void distance_at(const NphProxy *m, NphQuat rot_m, const NphProxy *t, NphQuat rot_t,
                        NphVec3 base, NphVec3 delta, f32 alpha, NphSimplexCache *cache,
                        NphProxyDistance *out)
{
    proxy_distance(t, pose_make(v3_mul_add(base, delta, -alpha), rot_t), m,
                   pose_make(v3_zero(), rot_m), 0, cache, out);
}

// cast.c:287-344: advance until the rounded shapes are within the touch
// distance; 0 when the motion never brings them there within `limit`.
static int arrive(const NphProxy *m, NphPhysicsPose pose_m, NphVec3 delta, const NphProxy *t,
                  NphPhysicsPose pose_t, f32 limit, Arrival *a)
{
    const f32 radius = m->radius + t->radius;
    NphSimplexCache cache;
    memset(&cache, 0, sizeof cache);
    f32 alpha = 0.0F;
    a->base = v3_sub(pose_t.position, pose_m.position);
    a->normal = v3(0.0F, 1.0F, 0.0F);
    a->point = pose_m.position;
    a->have_point = 0;
    a->approach = 0.0F;
    a->features = cache;
    for (u32 step = 0U; step < ADVANCE_STEPS; ++step) {
        NphProxyDistance pd;
        distance_at(m, pose_m.rotation, t, pose_t.rotation, a->base, delta, alpha, &cache, &pd);
        const NphDistance *d = &pd.d;
        if (pd.simplex_count < 4U) {
            a->features = cache; // the cores apart: the simplex names the closest features
        }
        const NphVec3 at = v3_mul_add(pose_m.position, delta, alpha);
        a->fraction = alpha;
        a->mover = pose_make(v3_zero(), pose_m.rotation);
        a->target = pose_make(v3_mul_add(a->base, delta, -alpha), pose_t.rotation);
        a->known = d->distance > WELL_APART;
        if (d->distance > 0.0F) {
            a->normal = d->normal;
            a->point = v3_add(at, v3_mul_add(d->point_a, d->normal, t->radius));
            a->have_point = 1;
        }
        const f32 separation = d->distance - radius;
        if (!(d->distance > 0.0F && separation > CAST_TOUCH)) {
            return 1;
        }
        const f32 approach = -v3_dot(delta, d->normal);
        if (!(approach > 0.0F)) {
            return 0;
        }
        alpha += separation / approach;
        if (!(alpha <= limit)) {
            return 0;
        }
        a->approach = approach;
    }
    a->known = 1; // a grazing approach still on its way stops short of contact
    return 1;
}

// cast.c:346-361: GJK's normal a millimetre of approach back, where the
// cores are apart.
static int normal_back(const NphProxy *m, NphPhysicsPose pose_m, NphVec3 delta, const NphProxy *t,
                       NphPhysicsPose pose_t, const Arrival *a, NphVec3 *normal)
{
    if (!(a->fraction > 0.0F) || !(a->approach > 0.0F)) {
        return 0;
    }
    const f32 back = nph_maxf(a->fraction - NORMAL_BACKOFF / a->approach, 0.0F);
    NphProxyDistance pd;
    distance_at(m, pose_m.rotation, t, pose_t.rotation, a->base, delta, back, NULL, &pd);
    if (!(pd.d.distance > WELL_APART)) {
        return 0;
    }
    *normal = pd.d.normal;
    return 1;
}

// cast.c:363-378: the last of the gap closed along the contact normal; -1
// when the motion slides along the contact or leaves it.
static f32 finish(const NphProxy *m, const NphProxy *t, const Arrival *a, NphVec3 delta,
                  NphVec3 normal)
{
    const f32 approach = -v3_dot(delta, normal);
    if (!(approach > CAST_SLIDE * v3_len(delta))) {
        return -1.0F;
    }
    const f32 g = proxy_gap(m, a->mover, t, a->target, normal);
    return g > 0.0F ? nph_minf(a->fraction + g / approach, 1.0F) : a->fraction;
}

static u32 deepest(const NphPhysicsManifold *m)
{
    u32 k = 0U;
    for (u32 i = 1U; i < m->point_count; ++i) {
        k = m->points[i].separation < m->points[k].separation ? i : k;
    }
    return k;
}

// cast.c:461-518: convex a moved by delta against convex b.
int cast_convex(const NphPhysicsGeometry *a, NphPhysicsPose pose_a, NphVec3 delta, f32 margin,
                const NphPhysicsGeometry *b, NphPhysicsPose pose_b, f32 limit,
                NphPhysicsCastHit *out)
{
    NphVec3 sa[8];
    NphVec3 sb[8];
    NphProxy pa = proxy_of(a, sa);
    pa.radius += margin;
    const NphProxy pb = proxy_of(b, sb);
    Arrival arr;
    if (!arrive(&pa, pose_a, delta, &pb, pose_b, limit, &arr)) {
        return 0;
    }
    NphVec3 normal = arr.normal;
    NphVec3 point = arr.point;
    if (!arr.known && arr.have_point) {
        // Polytopes whose cores touch: of GJK's last normal, the features'
        // exact one and GJK's a millimetre back, the one that parts them
        // farthest.
        normal = proxy_contact_axis(&pa, arr.mover, &pb, arr.target, &arr.features, normal);
        const f32 best = proxy_gap(&pa, arr.mover, &pb, arr.target, normal);
        NphVec3 n;
        if (normal_back(&pa, pose_a, delta, &pb, pose_b, &arr, &n) &&
            proxy_gap(&pa, arr.mover, &pb, arr.target, n) > best) {
            normal = n;
        }
    } else if (!arr.known) {
        // Touching at the start, cores and all: the narrow phase's normal.
        NphPhysicsManifold mf;
        nph_physics_collide(b, pose_b, a,
                            pose_make(v3_mul_add(pose_a.position, delta, arr.fraction),
                                      pose_a.rotation),
                            CAST_REACH + margin, NULL, &mf);
        if (mf.point_count == 0U) {
            return 0;
        }
        const NphPhysicsManifoldPoint p = mf.points[deepest(&mf)];
        normal = mf.normal;
        point = v3_mul_add(p.point, mf.normal, -0.5F * p.separation);
    }
    const f32 f = finish(&pa, &pb, &arr, delta, normal);
    if (f < 0.0F || f > limit) {
        return 0;
    }
    out->point = point;
    out->normal = normal;
    out->fraction = f;
    out->triangle = NPH_PHYSICS_NO_TRIANGLE;
    return 1;
}

static int is_convex(u32 type) { return type <= (u32)NPH_PHYSICS_SHAPE_HULL; }

bool nph_physics_shape_cast_margin(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                                   NphVec3 translation, f32 margin, const NphPhysicsGeometry *b,
                                   NphPhysicsPose pose_b, NphPhysicsCastHit *out)
{
    if (a == NULL || b == NULL || out == NULL || !is_convex((u32)a->type) ||
        !(v3_len_sq(translation) > 0.0F) || !(margin >= 0.0F) || !nph_isfinitef(margin) ||
        !is_convex((u32)b->type)) {
        return false;
    }
    return cast_convex(a, pose_a, translation, margin, b, pose_b, 1.0F, out) != 0;
}

bool nph_physics_shape_cast(const NphPhysicsGeometry *a, NphPhysicsPose pose_a, NphVec3 translation,
                            const NphPhysicsGeometry *b, NphPhysicsPose pose_b,
                            NphPhysicsCastHit *out)
{
    return nph_physics_shape_cast_margin(a, pose_a, translation, 0.0F, b, pose_b, out);
}

// ---------------------------------------------------------------------------
// Overlaps (cast.c:570-590; collide_mesh.c:1251-1294)
// ---------------------------------------------------------------------------
int overlap_convex(const NphPhysicsGeometry *a, NphPhysicsPose pose_a, const NphPhysicsGeometry *b,
                   NphPhysicsPose pose_b)
{
    NphVec3 sa[8];
    NphVec3 sb[8];
    const NphProxy pa = proxy_of(a, sa);
    const NphProxy pb = proxy_of(b, sb);
    NphProxyDistance pd;
    proxy_distance(&pa, pose_a, &pb, pose_b, 1, NULL, &pd);
    return pd.d.distance <= 0.0F;
}

// The mesh's triangles whose boxes meet `box` (the mesh's frame), depth
// first, while `visit` returns 1.
typedef int (*MeshTriangleFn)(void *user, const NphVec3 v[3]);

static void mesh_walk(const NphPhysicsMesh *mesh, NphAABB box, NphDVec3 origin, MeshTriangleFn visit,
                      void *user)
{
    const u32 *tri = mesh_triangles(mesh);
    const NphVec3 *v = mesh_vertices(mesh);
    NphBoxWalk w;
    box_walk_begin(&w, mesh, box);
    for (u32 t = box_walk_next(&w); t != NPH_NULL32; t = box_walk_next(&w)) {
        NphVec3 rv[3];
        for (u32 i = 0U; i < 3U; ++i) {
            rv[i] = relative_vertex(v[tri[3U * t + i]], origin);
        }
        if (!visit(user, rv)) {
            return;
        }
    }
}

typedef struct TriangleOverlap {
    NphProxy proxy;
    NphPhysicsPose pose;
    int hit;
} TriangleOverlap;

static int overlap_triangle(void *user, const NphVec3 v[3])
{
    TriangleOverlap *q = (TriangleOverlap *)user;
    const NphProxy tri = {v, 3U, 0.0F};
    NphProxyDistance pd;
    proxy_distance(&q->proxy, q->pose, &tri, pose_identity(), 1, NULL, &pd);
    q->hit = pd.d.distance <= 0.0F;
    return !q->hit;
}

// Convex a with `rotation` at `position` (the mesh's frame, 64 bits)
// touches one of the mesh's triangles, from either side.
int overlap_mesh_local(const NphPhysicsGeometry *a, NphQuat rotation, NphDVec3 position,
                       const NphPhysicsMesh *mesh)
{
    NphVec3 storage[8];
    TriangleOverlap q;
    q.proxy = proxy_of(a, storage);
    q.pose = pose_make(v3_zero(), rotation);
    q.hit = 0;
    const NphVec3 at = v3((f32)position.x, (f32)position.y, (f32)position.z);
    const NphVec3 aa = v3_abs(at);
    const f32 rounding = 2.4e-7F * nph_maxf(aa.x, nph_maxf(aa.y, aa.z));
    const NphAABB box = aabb_grow(nph_physics_geometry_aabb(a, pose_make(at, rotation)), rounding);
    mesh_walk(mesh, box, position, overlap_triangle, &q);
    return q.hit;
}

bool nph_physics_overlap(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                         const NphPhysicsGeometry *b, NphPhysicsPose pose_b)
{
    if (a == NULL || b == NULL || !is_convex((u32)a->type)) {
        return false;
    }
    if (b->type == NPH_PHYSICS_SHAPE_MESH) {
        // collide_mesh.c:1033-1041: a's place in the mesh's frame, 64 bits.
        const NphQuat rotation = q_mul(q_conj(pose_b.rotation), pose_a.rotation);
        const NphDVec3 d = dv3((f64)pose_a.position.x - (f64)pose_b.position.x,
                               (f64)pose_a.position.y - (f64)pose_b.position.y,
                               (f64)pose_a.position.z - (f64)pose_b.position.z);
        return overlap_mesh_local(a, rotation, drotate_inverse(pose_b.rotation, d), b->mesh) != 0;
    }
    return is_convex((u32)b->type) && overlap_convex(a, pose_a, b, pose_b) != 0;
}
