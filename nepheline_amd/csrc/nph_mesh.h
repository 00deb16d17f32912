// SPDX-License-Identifier: Apache-2.0
// Convex shapes against triangle meshes (collide_mesh.c:1-1020, mesh.c:998-1050),
// __host__ __device__: the narrow phase runs it one contact a thread.
//
// Each triangle is collided on its own in a frame with the mesh's axes
// centred on the convex shape: spheres and capsules by closest points,
// boxes and hulls by the separating-axis pair function against the triangle
// as a flat two-faced hull, then the triangle's rules (one side, active
// edges).  The triangles' contacts merge by normal (within 5 degrees) into
// at most eight clusters, each reduced to four points, and the four clusters
// with the deepest points are the pair's manifolds.  Every step, constant
// and tie rule is the reference's: contact ids, hence warm starting, and the
// solver's results depend on them.
#ifndef NPH_MESH_H
#define NPH_MESH_H

#include "nph_bvh4.h"
#include "nph_polytope.h"

#define NPH_MESH_MERGE_COS 0.99619470F
#define NPH_MESH_MERGE_DISTANCE 1e-4F
#define NPH_MESH_CLUSTER_POINTS (8U * NPH_PHYSICS_MANIFOLD_MAX_POINTS)
#define NPH_MESH_MAX_CLUSTERS 8U
#define NPH_MESH_FACE_PREFERENCE (0.5F * NPH_PHYSICS_LINEAR_SLOP)
#define NPH_MESH_EDGE_TOLERANCE 1e-4F
#define NPH_MESH_RIDGE_REACH 0.0868F
#define NPH_MESH_AGAINST_FACE (-1e-4F)
#define NPH_MESH_FACE_COS 0.99999F
#define NPH_MESH_PARALLEL_SIN_SQ 1e-4F
#define NPH_MESH_COMPACT_NEAREST 8U
#define NPH_TOUCH_SEPARATION (0.1F * NPH_PHYSICS_LINEAR_SLOP)

enum { NPH_FEATURE_FACE = 0U, NPH_FEATURE_EDGE = 1U, NPH_FEATURE_VERTEX = 4U };
enum { NPH_RULE_DROP = 0, NPH_RULE_KEEP = 1, NPH_RULE_FACE = 2 };

// collide_mesh.c:95-102: the inverse rotation of a 64-bit vector, in 64 bits.
NPH_HD NphDVec3 drotate_inverse(NphQuat q, NphDVec3 v)
{
    const f64 ux = -(f64)q.x;
    const f64 uy = -(f64)q.y;
    const f64 uz = -(f64)q.z;
    const f64 tx = (uy * v.z - uz * v.y) * 2.0;
    const f64 ty = (uz * v.x - ux * v.z) * 2.0;
    const f64 tz = (ux * v.y - uy * v.x) * 2.0;
    const f64 w = (f64)q.w;
    const f64 ax = v.x + tx * w;
    const f64 ay = v.y + ty * w;
    const f64 az = v.z + tz * w;
    return dv3(ax + (uy * tz - uz * ty), ay + (uz * tx - ux * tz), az + (ux * ty - uy * tx));
}

// physics_internal.h:143-147
NPH_HD NphVec3 relative_vertex(NphVec3 p, NphDVec3 origin)
{
    return v3((f32)((f64)p.x - origin.x), (f32)((f64)p.y - origin.y), (f32)((f64)p.z - origin.z));
}

// ---------------------------------------------------------------------------
// The triangle (collide_mesh.c:104-314)
// ---------------------------------------------------------------------------
typedef struct NphTri {
    NphVec3 v[3];
    NphVec3 n;
    u32 active;
} NphTri;

// The unit normal in 64 bits; zero for a degenerate triangle.
NPH_HD NphVec3 triangle_normal(const NphVec3 v[3])
{
    const f64 ax = (f64)v[1].x - (f64)v[0].x;
    const f64 ay = (f64)v[1].y - (f64)v[0].y;
    const f64 az = (f64)v[1].z - (f64)v[0].z;
    const f64 bx = (f64)v[2].x - (f64)v[0].x;
    const f64 by = (f64)v[2].y - (f64)v[0].y;
    const f64 bz = (f64)v[2].z - (f64)v[0].z;
    const f64 cx = ay * bz - az * by;
    const f64 cy = az * bx - ax * bz;
    const f64 cz = ax * by - ay * bx;
    const f64 len = __builtin_sqrt((cx * cx + cy * cy) + cz * cz);
    if (!(len > 0.0)) {
        return v3_zero();
    }
    return v3((f32)(cx / len), (f32)(cy / len), (f32)(cz / len));
}

NPH_HD int tri_prepare(NphTri *t, const NphVec3 v[3], u32 active, int two_sided, NphVec3 center)
{
    const NphVec3 n = triangle_normal(v);
    if (v3_len_sq(n) == 0.0F) {
        return 0;
    }
    const int behind = v3_dot(n, v3_sub(center, v[0])) < 0.0F;
    if (behind && !two_sided) {
        return 0;
    }
    t->v[0] = v[0];
    if (!behind) {
        t->v[1] = v[1];
        t->v[2] = v[2];
        t->n = n;
        t->active = active;
    } else {
        t->v[1] = v[2];
        t->v[2] = v[1];
        t->n = v3_neg(n);
        t->active = ((active >> 2U) & 1U) | (active & 2U) | ((active & 1U) << 2U);
    }
    return 1;
}

NPH_HD NphVec3 edge_outward(const NphTri *t, u32 e)
{
    return v3_cross(v3_sub(t->v[(e + 1U) % 3U], t->v[e]), t->n);
}

NPH_HD int edge_active(const NphTri *t, u32 e) { return ((t->active >> e) & 1U) != 0U; }

// Ericson, Real-Time Collision Detection, 5.1.5.
NPH_HD NphVec3 closest_on_triangle(const NphTri *t, NphVec3 p, u32 *feature)
{
    const NphVec3 a = t->v[0];
    const NphVec3 b = t->v[1];
    const NphVec3 c = t->v[2];
    const NphVec3 ab = v3_sub(b, a);
    const NphVec3 ac = v3_sub(c, a);
    const NphVec3 ap = v3_sub(p, a);
    const f32 d1 = v3_dot(ab, ap);
    const f32 d2 = v3_dot(ac, ap);
    if (d1 <= 0.0F && d2 <= 0.0F) {
        *feature = NPH_FEATURE_VERTEX + 0U;
        return a;
    }
    const NphVec3 bp = v3_sub(p, b);
    const f32 d3 = v3_dot(ab, bp);
    const f32 d4 = v3_dot(ac, bp);
    if (d3 >= 0.0F && d4 <= d3) {
        *feature = NPH_FEATURE_VERTEX + 1U;
        return b;
    }
    const f32 vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0F && d1 >= 0.0F && d3 <= 0.0F) {
        *feature = NPH_FEATURE_EDGE + 0U;
        return v3_mul_add(a, ab, d1 / (d1 - d3));
    }
    const NphVec3 cp = v3_sub(p, c);
    const f32 d5 = v3_dot(ab, cp);
    const f32 d6 = v3_dot(ac, cp);
    if (d6 >= 0.0F && d5 <= d6) {
        *feature = NPH_FEATURE_VERTEX + 2U;
        return c;
    }
    const f32 vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0F && d2 >= 0.0F && d6 <= 0.0F) {
        *feature = NPH_FEATURE_EDGE + 2U;
        return v3_mul_add(a, ac, d2 / (d2 - d6));
    }
    const f32 va = d3 * d6 - d5 * d4;
    if (va <= 0.0F && d4 - d3 >= 0.0F && d5 - d6 >= 0.0F) {
        *feature = NPH_FEATURE_EDGE + 1U;
        return v3_mul_add(b, v3_sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6)));
    }
    *feature = NPH_FEATURE_FACE;
    return v3_mul_add(p, t->n, -v3_dot(t->n, ap));
}

NPH_HD u32 feature_at(const NphTri *t, NphVec3 p)
{
    u32 near = 0U;
    for (u32 e = 0U; e < 3U; ++e) {
        const NphVec3 m = edge_outward(t, e);
        if (v3_dot(m, v3_sub(p, t->v[e])) >= -NPH_MESH_EDGE_TOLERANCE * v3_len(m)) {
            near |= 1U << e;
        }
    }
    switch (near) {
    case 0U:
        return NPH_FEATURE_FACE;
    case 1U:
        return NPH_FEATURE_EDGE + 0U;
    case 2U:
        return NPH_FEATURE_EDGE + 1U;
    case 4U:
        return NPH_FEATURE_EDGE + 2U;
    case 3U:
        return NPH_FEATURE_VERTEX + 1U;
    case 6U:
        return NPH_FEATURE_VERTEX + 2U;
    case 5U:
        return NPH_FEATURE_VERTEX + 0U;
    default:
        break;
    }
    u32 best = 0U;
    for (u32 i = 1U; i < 3U; ++i) {
        if (v3_dist_sq(p, t->v[i]) < v3_dist_sq(p, t->v[best])) {
            best = i;
        }
    }
    return NPH_FEATURE_VERTEX + best;
}

NPH_HD int keeps_normal(const NphTri *t, u32 feature, NphVec3 n)
{
    if (feature >= NPH_FEATURE_VERTEX) {
        const u32 i = feature - NPH_FEATURE_VERTEX;
        const u32 out = i;
        const u32 in = (i + 2U) % 3U;
        return (edge_active(t, out) && v3_dot(n, edge_outward(t, out)) > 0.0F) ||
               (edge_active(t, in) && v3_dot(n, edge_outward(t, in)) > 0.0F);
    }
    if (feature >= NPH_FEATURE_EDGE) {
        const u32 e = feature - NPH_FEATURE_EDGE;
        return edge_active(t, e) && v3_dot(n, edge_outward(t, e)) >= 0.0F;
    }
    return 0;
}

NPH_HD int settle(const NphTri *t, u32 feature, NphVec3 n)
{
    if (v3_dot(n, t->n) < NPH_MESH_AGAINST_FACE) {
        return NPH_RULE_DROP;
    }
    return keeps_normal(t, feature, n) ? NPH_RULE_KEEP : NPH_RULE_FACE;
}

NPH_HD int over_triangle(const NphTri *t, NphVec3 p, f32 grow, f32 ridge)
{
    for (u32 e = 0U; e < 3U; ++e) {
        const NphVec3 m = edge_outward(t, e);
        const f32 out = edge_active(t, e) ? grow : grow + ridge;
        if (v3_dot(m, v3_sub(p, t->v[e])) > out * v3_len(m)) {
            return 0;
        }
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Spheres, capsules, boxes and hulls against one triangle (collide_mesh.c:316-629)
// ---------------------------------------------------------------------------
NPH_HD NphVec3 round_point(NphVec3 core, NphVec3 n, f32 radius, f32 separation)
{
    return v3_mul_add(core, n, -(radius + 0.5F * separation));
}

NPH_HD f32 tri_min_separation(const NphTriContact *m)
{
    f32 s = NPH_F32_MAX;
    for (u32 i = 0U; i < m->point_count; ++i) {
        s = nph_minf(s, m->points[i].separation);
    }
    return s;
}

NPH_HD void sphere_triangle(const NphTri *t, NphVec3 c, f32 radius, f32 speculative,
                            NphTriContact *out)
{
    u32 feature;
    const NphVec3 q = closest_on_triangle(t, c, &feature);
    const NphVec3 d = v3_sub(c, q);
    const f32 dist_sq = v3_len_sq(d);
    const f32 limit = radius + speculative;
    if (dist_sq > limit * limit) {
        return;
    }
    const f32 dist = nph_sqrtf(dist_sq);
    NphVec3 n = dist > 1e-9F ? v3_scale(d, 1.0F / dist) : t->n;
    f32 sep = dist - radius;
    const int rule = settle(t, feature, n);
    if (rule == NPH_RULE_DROP) {
        return;
    }
    if (rule == NPH_RULE_FACE) {
        const f32 h = v3_dot(t->n, v3_sub(c, t->v[0]));
        if (!over_triangle(t, v3_mul_add(c, t->n, -h), NPH_MESH_EDGE_TOLERANCE,
                           NPH_MESH_RIDGE_REACH * nph_maxf(h, 0.0F))) {
            return;
        }
        n = t->n;
        sep = h - radius;
    }
    out->normal = v3_neg(n);
    tri_push(out, round_point(c, n, radius, sep), sep, 0U);
}

NPH_HD int core_over(const NphTri *t, NphVec3 p0, NphVec3 d, f32 grow, f32 ridge, f32 *x_lo,
                     f32 *x_hi)
{
    f32 lo = 0.0F;
    f32 hi = 1.0F;
    for (u32 e = 0U; e < 3U; ++e) {
        const NphVec3 m = edge_outward(t, e);
        const f32 out = edge_active(t, e) ? grow : grow + ridge;
        const f32 s0 = v3_dot(m, v3_sub(p0, t->v[e])) - out * v3_len(m);
        const f32 sd = v3_dot(m, d);
        if (sd == 0.0F) {
            if (s0 > 0.0F) {
                return 0;
            }
        } else {
            const f32 x = -s0 / sd;
            if (sd > 0.0F) {
                hi = nph_minf(hi, x);
            } else {
                lo = nph_maxf(lo, x);
            }
            if (lo > hi) {
                return 0;
            }
        }
    }
    *x_lo = lo;
    *x_hi = hi;
    return 1;
}

NPH_HD void capsule_face(const NphTri *t, NphVec3 p0, NphVec3 d, f32 x_lo, f32 x_hi, f32 radius,
                         f32 speculative, NphTriContact *out)
{
    const f32 h0 = v3_dot(t->n, v3_sub(p0, t->v[0]));
    const f32 hd = v3_dot(t->n, d);
    const f32 xs[2] = {x_lo, x_hi};
    const f32 hs[2] = {h0 + x_lo * hd, h0 + x_hi * hd};
    const int both = (x_hi - x_lo) * v3_len(d) > NPH_MESH_MERGE_DISTANCE;
    u32 lower = hs[0] < hs[1] ? 0U : 1U;
    if (hs[0] == hs[1]) {
        lower = x_hi >= 1.0F ? 1U : 0U;
    }
    out->normal = v3_neg(t->n);
    for (u32 k = 0U; k < 2U; ++k) {
        const f32 sep = hs[k] - radius;
        if ((both || k == lower) && sep <= speculative) {
            const NphVec3 x = v3_mul_add(p0, d, xs[k]);
            tri_push(out, round_point(x, t->n, radius, sep), sep, k);
        }
    }
}

NPH_HD void capsule_triangle(const NphTri *t, NphVec3 p0, NphVec3 p1, f32 radius, f32 speculative,
                             NphTriContact *out)
{
    const NphVec3 d = v3_sub(p1, p0);
    NphSegDist best = segment_distance(p0, p1, t->v[0], t->v[1]);
    u32 edge = 0U;
    for (u32 e = 1U; e < 3U; ++e) {
        const NphSegDist sd = segment_distance(p0, p1, t->v[e], t->v[(e + 1U) % 3U]);
        if (sd.distance_squared < best.distance_squared) {
            best = sd;
            edge = e;
        }
    }
    const f32 edge_dist = nph_sqrtf(best.distance_squared);
    f32 x_lo = 0.0F;
    f32 x_hi = 0.0F;
    if (core_over(t, p0, d, 0.0F, 0.0F, &x_lo, &x_hi)) {
        const f32 h0 = v3_dot(t->n, v3_sub(p0, t->v[0]));
        const f32 hd = v3_dot(t->n, d);
        const f32 low = nph_minf(h0 + x_lo * hd, h0 + x_hi * hd);
        if (low <= edge_dist + NPH_MESH_FACE_PREFERENCE) {
            if (low - radius <= speculative) {
                capsule_face(t, p0, d, x_lo, x_hi, radius, speculative, out);
            }
            return;
        }
    }
    if (edge_dist > radius + speculative) {
        return;
    }
    u32 feature = NPH_FEATURE_EDGE + edge;
    if (best.fraction2 <= 0.0F) {
        feature = NPH_FEATURE_VERTEX + edge;
    } else if (best.fraction2 >= 1.0F) {
        feature = NPH_FEATURE_VERTEX + (edge + 1U) % 3U;
    }
    const NphVec3 n = edge_dist > 1e-9F
                          ? v3_scale(v3_sub(best.closest1, best.closest2), 1.0F / edge_dist)
                          : t->n;
    const int rule = settle(t, feature, n);
    if (rule == NPH_RULE_DROP) {
        return;
    }
    if (rule == NPH_RULE_FACE) {
        const f32 h0 = v3_dot(t->n, v3_sub(p0, t->v[0]));
        const f32 h1 = h0 + v3_dot(t->n, d);
        if (core_over(t, p0, d, NPH_MESH_EDGE_TOLERANCE,
                      NPH_MESH_RIDGE_REACH * nph_maxf(nph_maxf(h0, h1), 0.0F), &x_lo, &x_hi)) {
            capsule_face(t, p0, d, x_lo, x_hi, radius, speculative, out);
            if (tri_min_separation(out) < edge_dist - radius - NPH_PHYSICS_LINEAR_SLOP) {
                out->point_count = 0U;
            }
        }
        return;
    }
    out->normal = v3_neg(n);
    const NphVec3 a = t->v[edge];
    const NphVec3 u = v3_sub(t->v[(edge + 1U) % 3U], a);
    const f32 dd = v3_len_sq(d);
    const f32 uu = v3_len_sq(u);
    if (dd > 1e-12F && v3_len_sq(v3_cross(d, u)) <= NPH_MESH_PARALLEL_SIN_SQ * dd * uu) {
        const f32 len = nph_sqrtf(uu);
        const NphVec3 dir = v3_scale(u, 1.0F / len);
        const f32 a0 = v3_dot(v3_sub(p0, a), dir);
        const f32 a1 = v3_dot(v3_sub(p1, a), dir);
        const f32 lo = nph_maxf(0.0F, nph_minf(a0, a1));
        const f32 hi = nph_minf(len, nph_maxf(a0, a1));
        if (hi - lo > NPH_PHYSICS_LINEAR_SLOP) {
            const f32 ends[2] = {lo, hi};
            for (u32 k = 0U; k < 2U; ++k) {
                const NphVec3 q = v3_mul_add(a, dir, ends[k]);
                const NphVec3 x = closest_on_segment(p0, p1, q);
                const f32 sep = v3_dist(x, q) - radius;
                if (sep <= speculative) {
                    tri_push(out, round_point(x, n, radius, sep), sep, 2U + k);
                }
            }
            if (out->point_count > 0U) {
                return;
            }
        }
    }
    const f32 sep = edge_dist - radius;
    tri_push(out, round_point(best.closest1, n, radius, sep), sep, 4U);
}

// collide_hull.c:94-121: the triangle as a flat hull, faces front (0) and
// back (1), half-edges in twin pairs.
typedef struct NphTriView {
    NphVec3 vertices[3];
    NphPlane planes[2];
} NphTriView;

#if defined(__HIPCC__)
__device__ __constant__ static const u16 nph_tri_faces_d[2] = {0U, 1U};
__device__ __constant__ static const NphPhysicsHullEdge nph_tri_edges_d[6] = {
    {2U, 1U, 0U, 0U}, {5U, 0U, 1U, 1U}, {4U, 3U, 1U, 0U},
    {1U, 2U, 2U, 1U}, {0U, 5U, 2U, 0U}, {3U, 4U, 0U, 1U},
};
#endif
static const u16 nph_tri_faces_h[2] = {0U, 1U};
static const NphPhysicsHullEdge nph_tri_edges_h[6] = {
    {2U, 1U, 0U, 0U}, {5U, 0U, 1U, 1U}, {4U, 3U, 1U, 0U},
    {1U, 2U, 2U, 1U}, {0U, 5U, 2U, 0U}, {3U, 4U, 0U, 1U},
};

NPH_HD NphView triangle_view(const NphVec3 v[3], NphVec3 n, NphTriView *s)
{
    for (u32 i = 0U; i < 3U; ++i) {
        s->vertices[i] = v[i];
    }
    s->planes[0] = plane_make(n, -v3_dot(n, v[0]));
    s->planes[1] = plane_make(v3_neg(n), v3_dot(n, v[0]));
    NphView h;
    h.vertices = s->vertices;
    h.planes = s->planes;
#if defined(__HIP_DEVICE_COMPILE__)
    h.faces = nph_tri_faces_d;
    h.edges = nph_tri_edges_d;
#else
    h.faces = nph_tri_faces_h;
    h.edges = nph_tri_edges_h;
#endif
    h.vertex_count = 3U;
    h.face_count = 2U;
    h.edge_count = 6U;
    h.centroid = v3_scale(v3_add(v3_add(v[0], v[1]), v[2]), 1.0F / 3.0F);
    return h;
}

// The pair functions below run the polytope code (C++, nph_polytope.h).
#if defined(__cplusplus)

NPH_HD void hull_triangle(const NphView *a, NphPhysicsPose pa, const NphTri *t, f32 speculative,
                          NphTriContact *out)
{
    NphTriView storage;
    const NphView b = triangle_view(t->v, t->n, &storage);
    tri_clear(out);
    collide_hulls_into(a, pa, &b, pose_identity(), speculative, 0, 1, NPH_TRI_POINTS, out);
    if (out->point_count == 0U || v3_dot(out->normal, t->n) <= -NPH_MESH_FACE_COS) {
        return;
    }
    u32 deepest = 0U;
    for (u32 i = 1U; i < out->point_count; ++i) {
        if (out->points[i].separation < out->points[deepest].separation) {
            deepest = i;
        }
    }
    const f32 depth = out->points[deepest].separation;
    const NphVec3 on = v3_mul_add(out->points[deepest].point, out->normal, 0.5F * depth);
    const int rule = settle(t, feature_at(t, on), v3_neg(out->normal));
    if (rule == NPH_RULE_KEEP) {
        return;
    }
    if (rule == NPH_RULE_FACE) {
        collide_hull_face(a, pa, &b, pose_identity(), 0U, speculative, out);
        if (out->point_count == 0U ||
            tri_min_separation(out) >= depth - NPH_PHYSICS_LINEAR_SLOP) {
            return;
        }
    }
    out->point_count = 0U;
}

// The convex shape in the local frame: at the origin, with its rotation.
typedef struct NphProbe {
    const NphGeom *g;
    NphPhysicsPose pose;
    NphVec3 center;
    NphVec3 core[2];
    f32 radius;
    NphBoxStorage box;
    NphView view;
} NphProbe;

NPH_HD void probe_init(NphProbe *p, const NphGeom *g, NphQuat rotation)
{
    p->g = g;
    p->pose = pose_make(v3_zero(), rotation);
    p->center = v3_zero();
    p->radius = 0.0F;
    p->core[0] = v3_zero();
    p->core[1] = v3_zero();
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        p->radius = g->a;
        return;
    case NPH_PHYSICS_SHAPE_CAPSULE: {
        const NphVec3 axis = v3_scale(q_axis_y(rotation), g->a);
        p->core[0] = v3_neg(axis);
        p->core[1] = axis;
        p->radius = g->b;
        return;
    }
    case NPH_PHYSICS_SHAPE_BOX:
        p->view = box_view(v3(g->a, g->b, g->c), &p->box);
        return;
    default:
        p->view = hull_view(g->hull);
        p->center = q_rotate(rotation, g->hull->centroid);
        return;
    }
}

NPH_HD void probe_triangle(const NphProbe *p, const NphTri *t, f32 speculative,
                           NphTriContact *out)
{
    tri_clear(out);
    switch (p->g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        sphere_triangle(t, v3_zero(), p->radius, speculative, out);
        return;
    case NPH_PHYSICS_SHAPE_CAPSULE:
        if (p->g->a > 0.0F) {
            capsule_triangle(t, p->core[0], p->core[1], p->radius, speculative, out);
        } else {
            sphere_triangle(t, v3_zero(), p->radius, speculative, out);
        }
        return;
    default:
        hull_triangle(&p->view, p->pose, t, speculative, out);
        return;
    }
}

// ---------------------------------------------------------------------------
// Clusters of triangle manifolds (collide_mesh.c:664-840)
// ---------------------------------------------------------------------------
typedef struct NphCluster {
    NphVec3 first;
    NphVec3 normal;
    NphVec3 sum;
    f32 depth;
    u32 count;
    NphMPoint points[NPH_MESH_CLUSTER_POINTS];
} NphCluster;

typedef struct NphSurfaceQuery {
    f32 speculative;
    int two_sided;
    int single;
    u32 cluster_count;
    NphProbe probe;
    NphCluster clusters[NPH_MESH_MAX_CLUSTERS];
} NphSurfaceQuery;

NPH_HD void compact_direction(u32 k, f32 *x, f32 *y)
{
    // Every 45 degrees, starting along the first tangent.
    const f32 h = 0.70710678F;
    switch (k) {
    case 0U: *x = 1.0F; *y = 0.0F; break;
    case 1U: *x = h; *y = h; break;
    case 2U: *x = 0.0F; *y = 1.0F; break;
    case 3U: *x = -h; *y = h; break;
    case 4U: *x = -1.0F; *y = 0.0F; break;
    case 5U: *x = -h; *y = -h; break;
    case 6U: *x = 0.0F; *y = -1.0F; break;
    default: *x = h; *y = -h; break;
    }
}

NPH_HD void cluster_compact(NphCluster *c)
{
    const NphVec3 t1 = v3_perpendicular(c->normal);
    const NphVec3 t2 = v3_cross(c->normal, t1);
    u32 keep = 0U; // bit per point
    u32 nearest[NPH_MESH_COMPACT_NEAREST];
    u32 near_count = 0U;
    for (u32 i = 0U; i < c->count; ++i) {
        const f32 s = c->points[i].separation;
        if (near_count == NPH_MESH_COMPACT_NEAREST &&
            !(s < c->points[nearest[NPH_MESH_COMPACT_NEAREST - 1U]].separation)) {
            continue;
        }
        u32 j = near_count < NPH_MESH_COMPACT_NEAREST ? near_count++ : NPH_MESH_COMPACT_NEAREST - 1U;
        while (j > 0U && s < c->points[nearest[j - 1U]].separation) {
            nearest[j] = nearest[j - 1U];
            j -= 1U;
        }
        nearest[j] = i;
    }
    for (u32 k = 0U; k < near_count; ++k) {
        keep |= 1U << nearest[k];
    }
    for (u32 pass = 0U; pass < 2U; ++pass) {
        const int touching = pass == 0U;
        for (u32 k = 0U; k < 8U; k += touching ? 1U : 2U) {
            f32 cx, cy;
            compact_direction(k, &cx, &cy);
            const NphVec3 dir = v3_add(v3_scale(t1, cx), v3_scale(t2, cy));
            u32 far = NPH_NULL32;
            f32 far_d = -NPH_F32_MAX;
            for (u32 i = 0U; i < c->count; ++i) {
                const f32 d = v3_dot(c->points[i].point, dir);
                if ((c->points[i].separation <= NPH_TOUCH_SEPARATION) == touching && d > far_d) {
                    far_d = d;
                    far = i;
                }
            }
            if (far != NPH_NULL32) {
                keep |= 1U << far;
            }
        }
    }
    u32 n = 0U;
    for (u32 i = 0U; i < c->count; ++i) {
        if ((keep >> i) & 1U) {
            c->points[n++] = c->points[i];
        }
    }
    c->count = n;
}

NPH_HD void cluster_add(NphSurfaceQuery *q, const NphTriContact *m, u32 triangle)
{
    f32 depth = NPH_F32_MAX;
    for (u32 i = 0U; i < m->point_count; ++i) {
        depth = nph_minf(depth, m->points[i].separation);
    }
    NphCluster *c = 0;
    for (u32 k = 0U; k < q->cluster_count; ++k) {
        if (v3_dot(q->clusters[k].first, m->normal) >= NPH_MESH_MERGE_COS) {
            c = &q->clusters[k];
            break;
        }
    }
    if (c == 0) {
        if (q->cluster_count < NPH_MESH_MAX_CLUSTERS) {
            c = &q->clusters[q->cluster_count++];
        } else {
            u32 s = 0U;
            for (u32 k = 1U; k < NPH_MESH_MAX_CLUSTERS; ++k) {
                if (q->clusters[k].depth >= q->clusters[s].depth) {
                    s = k;
                }
            }
            if (!(depth < q->clusters[s].depth)) {
                return;
            }
            c = &q->clusters[s];
        }
        c->first = m->normal;
        c->sum = v3_zero();
        c->depth = depth;
        c->count = 0U;
    }
    c->sum = v3_mul_add(c->sum, m->normal, nph_maxf(q->speculative - depth, 0.0F) + 1e-6F);
    c->normal = v3_normalize_or(c->sum, m->normal);
    c->depth = nph_minf(c->depth, depth);
    if (q->single) {
        if (c->count == 0U || m->points[0].separation < c->points[0].separation) {
            c->points[0] = m->points[0];
            c->points[0].id = nph_mix_id(triangle, m->points[0].id);
            c->count = 1U;
        }
        return;
    }
    for (u32 i = 0U; i < m->point_count; ++i) {
        NphMPoint p = m->points[i];
        p.id = nph_mix_id(triangle, p.id);
        int merged = 0;
        for (u32 j = 0U; j < c->count && !merged; ++j) {
            if (v3_dist_sq(c->points[j].point, p.point) <=
                NPH_MESH_MERGE_DISTANCE * NPH_MESH_MERGE_DISTANCE) {
                c->points[j].separation = nph_minf(c->points[j].separation, p.separation);
                merged = 1;
            }
        }
        if (merged) {
            continue;
        }
        if (c->count == NPH_MESH_CLUSTER_POINTS) {
            cluster_compact(c);
        }
        c->points[c->count++] = p;
    }
}

NPH_HD void surface_triangle(NphSurfaceQuery *q, u32 t, const NphVec3 v[3], u32 active_edges)
{
    NphTri prepared;
    if (!tri_prepare(&prepared, v, active_edges, q->two_sided, q->probe.center)) {
        return;
    }
    NphTriContact m;
    probe_triangle(&q->probe, &prepared, q->speculative, &m);
    if (m.point_count > 0U) {
        cluster_add(q, &m, t);
    }
}

// ---------------------------------------------------------------------------
// The mesh's triangles under a box (mesh.c:998-1050), each moved next to the
// shape in 64 bits (collide_mesh.c:886-916)
// ---------------------------------------------------------------------------
NPH_HD void mesh_visit_box(const NphPhysicsMesh *mesh, NphAABB box, NphDVec3 origin,
                           NphSurfaceQuery *q)
{
    const u32 *tri = mesh_triangles(mesh);
    const NphVec3 *v = mesh_vertices(mesh);
    const u8 *edges = mesh_edges(mesh);
    NphBoxWalk w;
    box_walk_begin(&w, mesh, box);
    for (u32 t = box_walk_next(&w); t != NPH_NULL32; t = box_walk_next(&w)) {
        NphVec3 rv[3];
        for (u32 i = 0U; i < 3U; ++i) {
            rv[i] = relative_vertex(v[tri[3U * t + i]], origin);
        }
        surface_triangle(q, t, rv, edges[t]);
    }
}

// ---------------------------------------------------------------------------
// The pair's manifolds (collide_mesh.c:918-1020)
// ---------------------------------------------------------------------------
typedef struct NphMeshManifolds {
    u32 count;
    NphManifold manifolds[NPH_PHYSICS_MESH_MAX_MANIFOLDS];
} NphMeshManifolds;

NPH_HD u32 capsule_extremes(NphMPoint *points, u32 count, NphVec3 axis)
{
    if (count <= 2U) {
        return count;
    }
    u32 pick[5];
    u32 n = 0U;
    u32 deepest = 0U;
    for (u32 i = 1U; i < count; ++i) {
        if (points[i].separation < points[deepest].separation) {
            deepest = i;
        }
    }
    pick[n++] = deepest;
    for (u32 pass = 0U; pass < 2U; ++pass) {
        u32 lo = NPH_NULL32;
        u32 hi = NPH_NULL32;
        f32 d_lo = NPH_F32_MAX;
        f32 d_hi = -NPH_F32_MAX;
        for (u32 i = 0U; i < count; ++i) {
            if (pass == 0U && points[i].separation > NPH_TOUCH_SEPARATION) {
                continue;
            }
            const f32 d = v3_dot(points[i].point, axis);
            if (d < d_lo) {
                d_lo = d;
                lo = i;
            }
            if (d > d_hi) {
                d_hi = d;
                hi = i;
            }
        }
        const u32 ends[2] = {lo, hi};
        for (u32 e = 0U; e < 2U; ++e) {
            int have = ends[e] == NPH_NULL32;
            for (u32 k = 0U; k < n && !have; ++k) {
                have = pick[k] == ends[e];
            }
            if (!have) {
                pick[n++] = ends[e];
            }
        }
    }
    n = nph_minu(n, NPH_PHYSICS_MANIFOLD_MAX_POINTS);
    NphMPoint kept[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
    for (u32 k = 0U; k < n; ++k) {
        kept[k] = points[pick[k]];
    }
    for (u32 k = 0U; k < n; ++k) {
        points[k] = kept[k];
    }
    return n;
}

NPH_HD void surface_finish(NphSurfaceQuery *q, NphMeshManifolds *out)
{
    const int capsule = q->probe.g->type == NPH_PHYSICS_SHAPE_CAPSULE && !q->single;
    const NphVec3 axis = v3_sub(q->probe.core[1], q->probe.core[0]);
    u32 order[NPH_MESH_MAX_CLUSTERS];
    for (u32 k = 0U; k < q->cluster_count; ++k) {
        NphCluster *c = &q->clusters[k];
        c->count = capsule ? capsule_extremes(c->points, c->count, axis)
                           : manifold_reduce(c->points, c->count, c->normal);
        u32 j = k;
        while (j > 0U && q->clusters[order[j - 1U]].depth > c->depth) {
            order[j] = order[j - 1U];
            j -= 1U;
        }
        order[j] = k;
    }
    const u32 keep = nph_minu(q->cluster_count, NPH_PHYSICS_MESH_MAX_MANIFOLDS);
    for (u32 i = 0U; i < keep; ++i) {
        const NphCluster *c = &q->clusters[order[i]];
        NphManifold *m = &out->manifolds[i];
        m->normal = c->normal;
        m->point_count = c->count;
        for (u32 j = 0U; j < c->count; ++j) {
            m->points[j] = c->points[j];
        }
        if (q->single) {
            const f32 sep = m->points[0].separation;
            m->points[0].point = v3_scale(m->normal, q->probe.radius + 0.5F * sep);
        }
    }
    out->count = keep;
}

// collide_mesh.c:860-884, 1010-1020: `a` with `rotation` at `position` in the
// mesh's frame.
NPH_HD void collide_mesh_local(const NphGeom *a, NphQuat rotation, NphDVec3 position,
                               const NphPhysicsMesh *mesh, f32 speculative, NphSurfaceQuery *q,
                               NphMeshManifolds *out)
{
    q->speculative = nph_maxf(speculative, 0.0F);
    q->two_sided = (mesh->flags & NPH_PHYSICS_MESH_TWO_SIDED) != 0U;
    q->single = a->type == NPH_PHYSICS_SHAPE_SPHERE ||
                (a->type == NPH_PHYSICS_SHAPE_CAPSULE && !(a->a > 0.0F));
    q->cluster_count = 0U;
    probe_init(&q->probe, a, rotation);
    const NphVec3 at = v3((f32)position.x, (f32)position.y, (f32)position.z);
    const NphVec3 aa = v3_abs(at);
    const f32 rounding = 2.4e-7F * nph_maxf(aa.x, nph_maxf(aa.y, aa.z));
    const NphAABB box =
        aabb_grow(geom_aabb(a, pose_make(at, rotation)), q->speculative + rounding);
    mesh_visit_box(mesh, box, position, q);
    surface_finish(q, out);
}

#endif // __cplusplus

#endif
