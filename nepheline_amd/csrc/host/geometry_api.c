// SPDX-License-Identifier: Apache-2.0
// The geometry and closest-point part of the public API (geometry.h:120-181,
// types.h:130-142, collide.h:181-235): host functions over the
// arithmetic the world step already uses (nph_geometry.h for mass
// properties and radii, nph_narrow.h for segment distance, the GJK stepper
// of nph_gjk.h through cast.c's proxies), with the reference's results.
#include "world_impl.h"

#include <float.h>

// ---------------------------------------------------------------------------
// Mass properties
// ---------------------------------------------------------------------------
NphPhysicsMassData nph_physics_mass_transform(NphPhysicsMassData m, NphPhysicsPose pose)
{
    return mass_transform(m, pose);
}

NphPhysicsMassData nph_physics_mass_add(NphPhysicsMassData a, NphPhysicsMassData b)
{
    return mass_add(a, b);
}

NphMat3 nph_physics_inertia_offset(f32 mass, NphVec3 r) { return inertia_offset(mass, r); }

// ---------------------------------------------------------------------------
// Radii and support points
// ---------------------------------------------------------------------------
static NphGeom api_geom(const NphPhysicsGeometry *g)
{
    const void *block = g->type == NPH_PHYSICS_SHAPE_HULL   ? (const void *)g->hull
                        : g->type == NPH_PHYSICS_SHAPE_MESH ? (const void *)g->mesh
                                                            : NULL;
    return geom_from(g, block);
}

f32 nph_physics_geometry_inner_radius(const NphPhysicsGeometry *g)
{
    const NphGeom x = api_geom(g);
    return geom_inner_radius(&x);
}

f32 nph_physics_geometry_outer_radius(const NphPhysicsGeometry *g)
{
    const NphGeom x = api_geom(g);
    return geom_outer_radius(&x);
}

// -1 for a negative component, else +1: a zero direction still names a
// corner.
static f32 side_of(f32 x) { return x < 0.0F ? -1.0F : 1.0F; }

// A ball's farthest point along d (straight up for a zero d).
static NphVec3 ball_reach(NphVec3 d, f32 radius)
{
    const f32 dd = v3_len_sq(d);
    return dd <= 0.0F ? v3(0.0F, radius, 0.0F) : v3_scale(d, radius / nph_sqrtf(dd));
}

NphVec3 nph_physics_geometry_support(const NphPhysicsGeometry *g, NphVec3 d)
{
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        return ball_reach(d, g->sphere.radius);
    case NPH_PHYSICS_SHAPE_CAPSULE:
        return v3_add(v3(0.0F, side_of(d.y) * g->capsule.half_height, 0.0F),
                      ball_reach(d, g->capsule.radius));
    case NPH_PHYSICS_SHAPE_BOX: {
        const NphVec3 h = g->box.half_extents;
        return v3(side_of(d.x) * h.x, side_of(d.y) * h.y, side_of(d.z) * h.z);
    }
    case NPH_PHYSICS_SHAPE_HULL: {
        const NphVec3 *v = (const NphVec3 *)((const u8 *)g->hull + g->hull->vertices_offset);
        return v[nph_physics_hull_support(g->hull, d)];
    }
    case NPH_PHYSICS_SHAPE_MESH: {
        // A scan of every vertex; the first of the farthest.
        const NphVec3 *v = mesh_vertices(g->mesh);
        u32 far = 0U;
        for (u32 i = 1U; i < g->mesh->vertex_count; ++i) {
            far = v3_dot(v[i], d) > v3_dot(v[far], d) ? i : far;
        }
        return v[far];
    }
    default:
        return v3_zero(); // heightfields are not built here
    }
}

// ---------------------------------------------------------------------------
// The surface's own normal at a point (geometry.c:304-395)
// ---------------------------------------------------------------------------
#define SURFACE_TOLERANCE 1e-4F

static NphVec3 unit_else(NphVec3 v, NphVec3 fallback)
{
    const f32 vv = v3_len_sq(v);
    return vv > 1e-24F ? v3_scale(v, 1.0F / nph_sqrtf(vv)) : fallback;
}

// Of the planes within `tolerance` of the nearest one to the point (faces
// meeting at an edge or corner), the one turned most towards `toward` (the
// first of equals). `dist(k)` is plane k's signed distance, `along(k)` its
// normal's projection on `toward`.
typedef struct FacePick {
    f32 (*dist)(const void *ctx, u32 k);
    f32 (*along)(const void *ctx, u32 k);
    const void *ctx;
    u32 count;
} FacePick;

static u32 face_at_point(const FacePick *fp, f32 tolerance)
{
    f32 nearest = -FLT_MAX;
    for (u32 k = 0U; k < fp->count; ++k) {
        nearest = nph_maxf(nearest, fp->dist(fp->ctx, k));
    }
    u32 best = 0U;
    f32 best_along = -FLT_MAX;
    for (u32 k = 0U; k < fp->count; ++k) {
        const f32 a = fp->along(fp->ctx, k);
        if (fp->dist(fp->ctx, k) >= nearest - tolerance && a > best_along) {
            best_along = a;
            best = k;
        }
    }
    return best;
}

typedef struct BoxAt {
    f32 p[3], e[3], t[3];
} BoxAt;

static f32 box_dist(const void *c, u32 k)
{
    const BoxAt *b = (const BoxAt *)c;
    return nph_absf(b->p[k]) - b->e[k];
}

static f32 box_along(const void *c, u32 k)
{
    const BoxAt *b = (const BoxAt *)c;
    return (b->p[k] < 0.0F ? -1.0F : 1.0F) * b->t[k];
}

typedef struct HullAt {
    const NphPlane *planes;
    NphVec3 p, t;
} HullAt;

static f32 hull_dist(const void *c, u32 k)
{
    const HullAt *h = (const HullAt *)c;
    return plane_distance(h->planes[k], h->p);
}

static f32 hull_along(const void *c, u32 k)
{
    const HullAt *h = (const HullAt *)c;
    return v3_dot(h->planes[k].normal, h->t);
}

NphVec3 nph_physics_geometry_surface_normal(const NphPhysicsGeometry *g, NphVec3 point,
                                            u32 triangle, NphVec3 toward)
{
    const NphVec3 ap = v3_abs(point);
    const f32 tolerance = SURFACE_TOLERANCE + 1e-6F * nph_maxf(ap.x, nph_maxf(ap.y, ap.z));
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        return unit_else(point, toward);
    case NPH_PHYSICS_SHAPE_CAPSULE: {
        const f32 h = g->capsule.half_height;
        const f32 y = point.y < -h ? -h : (point.y > h ? h : point.y);
        return unit_else(v3_sub(point, v3(0.0F, y, 0.0F)), toward);
    }
    case NPH_PHYSICS_SHAPE_BOX: {
        const BoxAt at = {{point.x, point.y, point.z},
                          {g->box.half_extents.x, g->box.half_extents.y, g->box.half_extents.z},
                          {toward.x, toward.y, toward.z}};
        const FacePick fp = {box_dist, box_along, &at, 3U};
        const u32 k = face_at_point(&fp, tolerance);
        f32 n[3] = {0.0F, 0.0F, 0.0F};
        n[k] = at.p[k] < 0.0F ? -1.0F : 1.0F;
        return v3(n[0], n[1], n[2]);
    }
    case NPH_PHYSICS_SHAPE_HULL: {
        const HullAt at = {(const NphPlane *)((const u8 *)g->hull + g->hull->planes_offset), point,
                           toward};
        const FacePick fp = {hull_dist, hull_along, &at, g->hull->face_count};
        return at.planes[face_at_point(&fp, tolerance)].normal;
    }
    case NPH_PHYSICS_SHAPE_MESH: {
        if (triangle >= g->mesh->triangle_count) {
            return toward;
        }
        NphVec3 v[3];
        mesh_corners(g->mesh, triangle, v);
        const NphVec3 n = unit_else(v3_cross(v3_sub(v[1], v[0]), v3_sub(v[2], v[0])), toward);
        const int two_sided = (g->mesh->flags & NPH_PHYSICS_MESH_TWO_SIDED) != 0U;
        return two_sided && v3_dot(n, toward) < 0.0F ? v3_neg(n) : n;
    }
    default:
        return toward;
    }
}

// ---------------------------------------------------------------------------
// Proxies and closest points
// ---------------------------------------------------------------------------
NphPhysicsProxy nph_physics_geometry_proxy(const NphPhysicsGeometry *g,
                                           NphVec3 storage[NPH_PHYSICS_PROXY_STORAGE])
{
    NphPhysicsProxy out;
    if (g->type > NPH_PHYSICS_SHAPE_HULL) { // not convex: a point at the origin
        storage[0] = v3_zero();
        out.points = storage;
        out.count = 1U;
        out.radius = 0.0F;
        return out;
    }
    const NphProxy p = proxy_of(g, storage);
    out.points = p.points;
    out.count = p.count;
    out.radius = p.radius;
    return out;
}

u32 nph_physics_proxy_support(const NphPhysicsProxy *p, NphVec3 direction)
{
    return points_support(p->points, p->count, direction);
}

NphPhysicsSegmentDistance nph_physics_segment_distance(NphVec3 p1, NphVec3 q1, NphVec3 p2,
                                                       NphVec3 q2)
{
    const NphSegDist s = segment_distance(p1, q1, p2, q2);
    NphPhysicsSegmentDistance out;
    out.closest1 = s.closest1;
    out.closest2 = s.closest2;
    out.fraction1 = s.fraction1;
    out.fraction2 = s.fraction2;
    out.distance_squared = s.distance_squared;
    return out;
}

void nph_physics_distance(const NphPhysicsProxy *a, NphPhysicsPose pose_a, const NphPhysicsProxy *b,
                          NphPhysicsPose pose_b, bool use_radii, NphPhysicsSimplexCache *cache,
                          NphPhysicsDistance *out)
{
    const NphProxy pa = {a->points, a->count, a->radius};
    const NphProxy pb = {b->points, b->count, b->radius};
    NphSimplexCache sc;
    memset(&sc, 0, sizeof sc);
    if (cache != NULL) {
        sc.count = cache->count;
        memcpy(sc.index_a, cache->index_a, 4U);
        memcpy(sc.index_b, cache->index_b, 4U);
    }
    NphProxyDistance pd;
    proxy_distance(&pa, pose_a, &pb, pose_b, use_radii ? 1 : 0, &sc, &pd);
    if (cache != NULL) {
        cache->count = sc.count;
        memcpy(cache->index_a, sc.index_a, 4U);
        memcpy(cache->index_b, sc.index_b, 4U);
    }
    out->point_a = pd.d.point_a;
    out->point_b = pd.d.point_b;
    out->normal = pd.d.normal;
    out->distance = pd.d.distance;
    out->iterations = pd.iterations;
    out->simplex_count = pd.simplex_count;
}
