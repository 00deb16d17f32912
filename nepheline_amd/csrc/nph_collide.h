// SPDX-License-Identifier: Apache-2.0
// The narrow phase for the kernels: the manifold of two convex shapes
// (spheres, capsules, boxes, hulls), in the reference's arithmetic.
//
//   dispatch / type order       collide.c:474-540
//   segment distance            collide.c:16-60 (Ericson 5.1.9)
//   manifold reduction          collide.c:62-273 (Gregorius 2015, catch angles)
//   balls, capsules, box-sphere collide.c:281-454
//   hull-sphere / hull-capsule  collide_hull.c:156-376
//   hull SAT + face clipping    collide_hull.c:378-1002 (Gregorius 2013)
//   box-box, 15 axes            collide_hull.c:1004-1266
//   GJK distance                gjk.c:26-416 (signed volumes, Montanari 2017)
//
// Everything is one thread's work over private arrays sized for hulls of at
// most NPH_HULL_MAX_VERTICES vertices. The step's kernels are built at 32
// (the arrays of the common hulls stay small); pairs with a larger hull run
// in kernels built from the same code at the reference's 256 with its clip
// capacity of 264 (collide_hull.c:25-26). At 32 a clipped face holds at
// most 32 + 32 points, so 68 never truncates, as 264 never does there; a
// mesh pair clips a triangle and a hull face, at most 64 + 3 points when the
// hull's faces have at most 64 vertices, so mesh pairs of any hull of such
// faces keep the 32-vertex build too.
#ifndef NPH_COLLIDE_H
#define NPH_COLLIDE_H

#include "nph_geometry.h"

#ifndef NPH_HULL_MAX_VERTICES
#define NPH_HULL_MAX_VERTICES 32U
#define NPH_CLIP_CAPACITY     68U
#endif
#define NPH_HULL_MAX_FACES    (2U * NPH_HULL_MAX_VERTICES)
// 1: the SAT's edge loop reads B's face normals (in A's frame) from a private
// array filled by the face loop; 0: it recomputes them (the same function of
// the same inputs, so the same bits) to keep the private frame small.
#ifndef NPH_CACHE_B_NORMALS
#define NPH_CACHE_B_NORMALS 1
#endif
// The largest hull face a world takes (see the mesh pairs above).
#define NPH_HULL_MAX_FACE_VERTICES 64U

#define NPH_FACE_BIAS         (0.1F * NPH_PHYSICS_LINEAR_SLOP)
#define NPH_EDGE_BIAS         (0.5F * NPH_PHYSICS_LINEAR_SLOP)
#define NPH_FEATURE_AGREEMENT 0.995F
#define NPH_REDUCE_LEVEL      (0.002F * NPH_PHYSICS_LINEAR_SLOP)
#define NPH_REDUCE_ANGLE_TIE  0.1F
#define NPH_REDUCE_ON_LINE    1e-3F
#define NPH_REDUCE_AT_POINT   1e-6F

enum { NPH_CACHE_NONE = 0, NPH_CACHE_FACE_A = 1, NPH_CACHE_FACE_B = 2, NPH_CACHE_EDGES = 3 };

typedef struct NphCache {
    u8 type;
    u8 reserved;
    u16 index_a;
    u16 index_b;
    u16 reserved2;
} NphCache; // collide.h:85-93

typedef struct NphMPoint {
    NphVec3 point;
    f32 separation;
    u32 id;
} NphMPoint;

typedef struct NphManifold {
    NphVec3 normal;
    u32 point_count;
    NphMPoint points[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
} NphManifold;

NPH_HD void manifold_clear(NphManifold *m)
{
    m->normal = v3(0.0F, 1.0F, 0.0F);
    m->point_count = 0U;
}

NPH_HD void manifold_push(NphManifold *m, NphVec3 point, f32 separation, u32 id)
{
    if (m->point_count < NPH_PHYSICS_MANIFOLD_MAX_POINTS) {
        NphMPoint *p = &m->points[m->point_count++];
        p->point = point;
        p->separation = separation;
        p->id = id;
    }
}

// ---------------------------------------------------------------------------
// Views: a hull block, or a box with the fixed topology of collide_hull.c:49-93
// ---------------------------------------------------------------------------
typedef struct NphView {
    const NphVec3 *vertices;
    const NphPlane *planes;
    const u16 *faces;
    const NphPhysicsHullEdge *edges;
    u32 vertex_count;
    u32 face_count;
    u32 edge_count;
    NphVec3 centroid;
} NphView;

typedef struct NphBoxStorage {
    NphVec3 vertices[8];
    NphPlane planes[6];
} NphBoxStorage;

// Corners by bit (x, y, z); faces -x +x -y +y -z +z; half-edges in twin pairs,
// each {next, twin, origin, face}, identical to the reference hull of a box.
#if defined(__HIPCC__)
__device__ __constant__ static const u16 nph_box_faces_d[6] = {0U, 8U, 16U, 5U, 7U, 19U};
__device__ __constant__ static const NphPhysicsHullEdge nph_box_edges_d[24] = {
#else
static const u16 nph_box_faces_h[6] = {0U, 8U, 16U, 5U, 7U, 19U};
static const NphPhysicsHullEdge nph_box_edges_h[24] = {
#endif
    {2U, 1U, 0U, 0U},   {16U, 0U, 4U, 2U},  {4U, 3U, 4U, 0U},   {19U, 2U, 6U, 5U},
    {6U, 5U, 6U, 0U},   {20U, 4U, 2U, 3U},  {0U, 7U, 2U, 0U},   {23U, 6U, 0U, 4U},
    {10U, 9U, 1U, 1U},  {17U, 8U, 3U, 4U},  {12U, 11U, 3U, 1U}, {22U, 10U, 7U, 3U},
    {14U, 13U, 7U, 1U}, {21U, 12U, 5U, 5U}, {8U, 15U, 5U, 1U},  {18U, 14U, 1U, 2U},
    {15U, 17U, 0U, 2U}, {7U, 16U, 1U, 4U},  {1U, 19U, 5U, 2U},  {13U, 18U, 4U, 5U},
    {11U, 21U, 6U, 3U}, {3U, 20U, 7U, 5U},  {5U, 23U, 3U, 3U},  {9U, 22U, 2U, 4U},
};

NPH_HD NphView box_view(NphVec3 h, NphBoxStorage *s)
{
    for (u32 i = 0U; i < 8U; ++i) {
        s->vertices[i] = v3((i & 1U) != 0U ? h.x : -h.x, (i & 2U) != 0U ? h.y : -h.y,
                            (i & 4U) != 0U ? h.z : -h.z);
    }
    s->planes[0] = plane_make(v3(-1.0F, 0.0F, 0.0F), -h.x);
    s->planes[1] = plane_make(v3(1.0F, 0.0F, 0.0F), -h.x);
    s->planes[2] = plane_make(v3(0.0F, -1.0F, 0.0F), -h.y);
    s->planes[3] = plane_make(v3(0.0F, 1.0F, 0.0F), -h.y);
    s->planes[4] = plane_make(v3(0.0F, 0.0F, -1.0F), -h.z);
    s->planes[5] = plane_make(v3(0.0F, 0.0F, 1.0F), -h.z);
    NphView v;
    v.vertices = s->vertices;
    v.planes = s->planes;
#if defined(__HIP_DEVICE_COMPILE__)
    v.faces = nph_box_faces_d;
    v.edges = nph_box_edges_d;
#elif defined(__HIPCC__)
    v.faces = 0;
    v.edges = 0;
#else
    v.faces = nph_box_faces_h;
    v.edges = nph_box_edges_h;
#endif
    v.vertex_count = 8U;
    v.face_count = 6U;
    v.edge_count = 24U;
    v.centroid = v3_zero();
    return v;
}

NPH_HD NphView hull_view(const NphPhysicsHull *h)
{
    NphView v;
    v.vertices = hull_vertices(h);
    v.planes = hull_planes(h);
    v.faces = hull_faces(h);
    v.edges = hull_edges(h);
    v.vertex_count = h->vertex_count;
    v.face_count = h->face_count;
    v.edge_count = h->edge_count;
    v.centroid = h->centroid;
    return v;
}

NPH_HD NphVec3 edge_head(const NphView *h, const NphVec3 *vertices, u32 e)
{
    return vertices[h->edges[h->edges[e].next].origin];
}

// ---------------------------------------------------------------------------
// Segment distance (collide.c:16-60) and closest point on a segment
// (math/src/geometry.c:250-262)
// ---------------------------------------------------------------------------
typedef struct NphSegDist {
    NphVec3 closest1;
    NphVec3 closest2;
    f32 fraction1;
    f32 fraction2;
    f32 distance_squared;
} NphSegDist;

NPH_HD NphSegDist segment_distance(NphVec3 p1, NphVec3 q1, NphVec3 p2, NphVec3 q2)
{
    const NphVec3 d1 = v3_sub(q1, p1);
    const NphVec3 d2 = v3_sub(q2, p2);
    const NphVec3 r = v3_sub(p1, p2);
    const f32 a = v3_dot(d1, d1);
    const f32 e = v3_dot(d2, d2);
    const f32 f = v3_dot(d2, r);
    const f32 eps = 1e-12F;
    f32 s = 0.0F;
    f32 t = 0.0F;
    if (a <= eps && e <= eps) {
        s = 0.0F;
        t = 0.0F;
    } else if (a <= eps) {
        t = nph_saturate(f / e);
    } else {
        const f32 c = v3_dot(d1, r);
        if (e <= eps) {
            s = nph_saturate(-c / a);
        } else {
            const f32 b = v3_dot(d1, d2);
            const f32 denom = a * e - b * b;
            s = denom > 0.0F ? nph_saturate((b * f - c * e) / denom) : 0.0F;
            t = (b * s + f) / e;
            if (t < 0.0F) {
                t = 0.0F;
                s = nph_saturate(-c / a);
            } else if (t > 1.0F) {
                t = 1.0F;
                s = nph_saturate((b - c) / a);
            }
        }
    }
    NphSegDist o;
    o.closest1 = v3_mul_add(p1, d1, s);
    o.closest2 = v3_mul_add(p2, d2, t);
    o.fraction1 = s;
    o.fraction2 = t;
    o.distance_squared = v3_dist_sq(o.closest1, o.closest2);
    return o;
}

NPH_HD NphVec3 closest_on_segment(NphVec3 a, NphVec3 b, NphVec3 p)
{
    const NphVec3 ab = v3_sub(b, a);
    const f32 len_sq = v3_len_sq(ab);
    f32 t = 0.0F;
    if (len_sq > 0.0F) {
        t = nph_saturate(v3_dot(v3_sub(p, a), ab) / len_sq);
    }
    return v3_mul_add(a, ab, t);
}

// ---------------------------------------------------------------------------
// GJK distance on point sets (gjk.c:26-416); proxies of at most 32 points
// ---------------------------------------------------------------------------
typedef struct NphGjkVertex {
    NphVec3 a;
    NphVec3 b;
    NphVec3 w;
    u32 ia;
    u32 ib;
} NphGjkVertex;

typedef struct NphGjkSimplex {
    NphGjkVertex v[4];
    f32 lambda[4];
    u32 count;
} NphGjkSimplex;

typedef struct NphDistance {
    NphVec3 point_a;
    NphVec3 point_b;
    NphVec3 normal;
    f32 distance;
} NphDistance;

// geometry.c:495-508: the first point of greatest projection.
NPH_HD u32 points_support(const NphVec3 *p, u32 count, NphVec3 d)
{
    u32 best = 0U;
    f32 best_dot = v3_dot(p[0], d);
    for (u32 i = 1U; i < count; ++i) {
        const f32 x = v3_dot(p[i], d);
        if (x > best_dot) {
            best_dot = x;
            best = i;
        }
    }
    return best;
}

// Identity poses (every caller in the narrow phase works in one frame):
// pose_point(identity, p) adds rotations by the unit quaternion, which
// return p exactly, and a zero translation.
NPH_HD NphVec3 gjk_point(NphVec3 p)
{
    return pose_point(pose_identity(), p);
}

NPH_HD NphGjkVertex gjk_vertex(const NphVec3 *a, u32 ia, const NphVec3 *b, u32 ib)
{
    NphGjkVertex v;
    v.a = gjk_point(a[ia]);
    v.b = gjk_point(b[ib]);
    v.w = v3_sub(v.a, v.b);
    v.ia = ia;
    v.ib = ib;
    return v;
}

NPH_HD int same_sign(f32 a, f32 b) { return (a > 0.0F && b > 0.0F) || (a < 0.0F && b < 0.0F); }

NPH_HD void gjk_s1d(NphVec3 s1, NphVec3 s2, f32 lambda[2])
{
    const NphVec3 t = v3_sub(s2, s1);
    const f32 tt = v3_dot(t, t);
    if (!(tt > 0.0F)) {
        lambda[0] = 1.0F;
        lambda[1] = 0.0F;
        return;
    }
    const NphVec3 p0 = v3_mul_add(s1, t, -v3_dot(s1, t) / tt);
    const u32 axis = v3_max_abs_index(t);
    const f32 mu = v3_at(s1, axis) - v3_at(s2, axis);
    const f32 c1 = v3_at(p0, axis) - v3_at(s2, axis);
    const f32 c2 = v3_at(s1, axis) - v3_at(p0, axis);
    if (same_sign(mu, c1) && same_sign(mu, c2)) {
        lambda[0] = c1 / mu;
        lambda[1] = c2 / mu;
    } else if (same_sign(mu, c1)) {
        lambda[0] = 1.0F;
        lambda[1] = 0.0F;
    } else {
        lambda[0] = 0.0F;
        lambda[1] = 1.0F;
    }
}

NPH_HD NphVec3 weighted2(NphVec3 s1, NphVec3 s2, const f32 l[2])
{
    return v3_add(v3_scale(s1, l[0]), v3_scale(s2, l[1]));
}

NPH_HD NphVec3 weighted3(NphVec3 s1, NphVec3 s2, NphVec3 s3, const f32 l[3])
{
    return v3_add(v3_add(v3_scale(s1, l[0]), v3_scale(s2, l[1])), v3_scale(s3, l[2]));
}

NPH_HD void gjk_s2d_edges(const NphVec3 s[3], const int test[3], f32 lambda[3])
{
    f32 best = NPH_F32_MAX;
    lambda[0] = 1.0F;
    lambda[1] = 0.0F;
    lambda[2] = 0.0F;
    for (u32 j = 0U; j < 3U; ++j) {
        if (!test[j]) {
            continue;
        }
        const u32 i0 = (j + 1U) % 3U;
        const u32 i1 = (j + 2U) % 3U;
        f32 l[2];
        gjk_s1d(s[i0], s[i1], l);
        const f32 d = v3_len_sq(weighted2(s[i0], s[i1], l));
        if (d < best) {
            best = d;
            lambda[j] = 0.0F;
            lambda[i0] = l[0];
            lambda[i1] = l[1];
        }
    }
}

NPH_HD f32 gjk_area2(NphVec3 p, NphVec3 q, NphVec3 r, u32 i, u32 j)
{
    return (v3_at(q, i) - v3_at(p, i)) * (v3_at(r, j) - v3_at(p, j)) -
           (v3_at(q, j) - v3_at(p, j)) * (v3_at(r, i) - v3_at(p, i));
}

NPH_HD void gjk_s2d(const NphVec3 s[3], f32 lambda[3])
{
    const NphVec3 n = v3_cross(v3_sub(s[1], s[0]), v3_sub(s[2], s[0]));
    const f32 nn = v3_dot(n, n);
    if (!(nn > 0.0F)) {
        const int all[3] = {1, 1, 1};
        gjk_s2d_edges(s, all, lambda);
        return;
    }
    const NphVec3 p0 = v3_scale(n, v3_dot(s[0], n) / nn);
    const u32 k = v3_max_abs_index(n);
    const u32 i = (k + 1U) % 3U;
    const u32 j = (k + 2U) % 3U;
    const f32 mu = v3_at(n, k);
    const f32 c[3] = {gjk_area2(p0, s[1], s[2], i, j), gjk_area2(s[0], p0, s[2], i, j),
                      gjk_area2(s[0], s[1], p0, i, j)};
    const int wrong[3] = {!same_sign(mu, c[0]), !same_sign(mu, c[1]), !same_sign(mu, c[2])};
    if (!wrong[0] && !wrong[1] && !wrong[2]) {
        lambda[0] = c[0] / mu;
        lambda[1] = c[1] / mu;
        lambda[2] = c[2] / mu;
        return;
    }
    gjk_s2d_edges(s, wrong, lambda);
}

NPH_HD f32 gjk_det3(NphVec3 a, NphVec3 b, NphVec3 c) { return v3_dot(a, v3_cross(b, c)); }

NPH_HD int gjk_s3d(const NphVec3 s[4], f32 lambda[4])
{
    const NphVec3 e1 = v3_sub(s[1], s[0]);
    const NphVec3 e2 = v3_sub(s[2], s[0]);
    const NphVec3 e3 = v3_sub(s[3], s[0]);
    const f32 det = gjk_det3(e1, e2, e3);
    const NphVec3 o0 = v3_neg(s[0]);
    const f32 c[4] = {gjk_det3(s[1], s[2], s[3]), gjk_det3(o0, e2, e3), gjk_det3(e1, o0, e3),
                      gjk_det3(e1, e2, o0)};
    int wrong[4];
    int inside = det != 0.0F;
    for (u32 i = 0U; i < 4U; ++i) {
        wrong[i] = !same_sign(det, c[i]);
        inside = inside && !wrong[i];
    }
    if (inside) {
        for (u32 i = 0U; i < 4U; ++i) {
            lambda[i] = c[i] / det;
        }
        return 1;
    }
    f32 best = NPH_F32_MAX;
    for (u32 i = 0U; i < 4U; ++i) {
        lambda[i] = i == 0U ? 1.0F : 0.0F;
    }
    for (u32 f = 0U; f < 4U; ++f) {
        if (!wrong[f] && det != 0.0F) {
            continue;
        }
        NphVec3 tri[3];
        u32 map[3];
        u32 n = 0U;
        for (u32 i = 0U; i < 4U; ++i) {
            if (i != f) {
                tri[n] = s[i];
                map[n] = i;
                n += 1U;
            }
        }
        f32 l[3];
        gjk_s2d(tri, l);
        const f32 d = v3_len_sq(weighted3(tri[0], tri[1], tri[2], l));
        if (d < best) {
            best = d;
            for (u32 i = 0U; i < 4U; ++i) {
                lambda[i] = 0.0F;
            }
            lambda[map[0]] = l[0];
            lambda[map[1]] = l[1];
            lambda[map[2]] = l[2];
        }
    }
    return 0;
}

NPH_HD NphVec3 gjk_simplex_solve(NphGjkSimplex *s, int *inside)
{
    *inside = 0;
    f32 lambda[4] = {1.0F, 0.0F, 0.0F, 0.0F};
    if (s->count == 2U) {
        gjk_s1d(s->v[0].w, s->v[1].w, lambda);
    } else if (s->count == 3U) {
        const NphVec3 tri[3] = {s->v[0].w, s->v[1].w, s->v[2].w};
        gjk_s2d(tri, lambda);
    } else if (s->count == 4U) {
        const NphVec3 tet[4] = {s->v[0].w, s->v[1].w, s->v[2].w, s->v[3].w};
        *inside = gjk_s3d(tet, lambda);
    }
    NphVec3 v = v3_zero();
    u32 kept = 0U;
    for (u32 i = 0U; i < s->count; ++i) {
        if (lambda[i] > 0.0F || (s->count == 1U)) {
            s->v[kept] = s->v[i];
            s->lambda[kept] = lambda[i];
            v = v3_mul_add(v, s->v[i].w, lambda[i]);
            kept += 1U;
        }
    }
    if (kept == 0U) {
        s->lambda[0] = 1.0F;
        v = s->v[0].w;
        kept = 1U;
    }
    s->count = kept;
    return v;
}

// nph_physics_distance with identity poses, no radii, no cache (every call
// of the narrow phase): GJK on the point sets, then the closest points.
NPH_HD void gjk_distance(const NphVec3 *a, u32 na, const NphVec3 *b, u32 nb, NphDistance *out)
{
    NphGjkSimplex s;
    s.v[0] = gjk_vertex(a, 0U, b, 0U);
    s.count = 1U;
    const f32 scale =
        nph_maxf(v3_len_sq(s.v[0].a), v3_len_sq(s.v[0].b)) + v3_len_sq(s.v[0].w) + 1e-12F;
    int inside = 0;
    NphVec3 v = gjk_simplex_solve(&s, &inside);
    f32 v_sq = v3_len_sq(v);
    u32 iterations = 0U;
    while (!inside && iterations < 40U) {
        if (v_sq <= 1e-12F * scale) {
            inside = 1;
            break;
        }
        iterations += 1U;
        const NphVec3 d = v3_neg(v);
        // proxy_support rotates the direction by the inverse identity first.
        const u32 ia = points_support(a, na, q_rotate_inv(q_identity(), d));
        const u32 ib = points_support(b, nb, q_rotate_inv(q_identity(), v));
        int duplicate = 0;
        for (u32 i = 0U; i < s.count; ++i) {
            duplicate = duplicate || (s.v[i].ia == ia && s.v[i].ib == ib);
        }
        if (duplicate) {
            break;
        }
        const NphGjkVertex w = gjk_vertex(a, ia, b, ib);
        if (v_sq - v3_dot(v, w.w) <= 1e-6F * v_sq) {
            break;
        }
        s.v[s.count++] = w;
        const NphVec3 v_new = gjk_simplex_solve(&s, &inside);
        const f32 v_new_sq = v3_len_sq(v_new);
        if (!inside && v_new_sq >= v_sq) {
            break;
        }
        v = v_new;
        v_sq = v_new_sq;
    }
    NphVec3 pa = v3_zero();
    NphVec3 pb = v3_zero();
    f32 total = 0.0F;
    for (u32 i = 0U; i < s.count; ++i) {
        pa = v3_mul_add(pa, s.v[i].a, s.lambda[i]);
        pb = v3_mul_add(pb, s.v[i].b, s.lambda[i]);
        total += s.lambda[i];
    }
    if (total > 0.0F && total != 1.0F) {
        pa = v3_scale(pa, 1.0F / total);
        pb = v3_scale(pb, 1.0F / total);
    }
    f32 distance = inside ? 0.0F : v3_dist(pa, pb);
    NphVec3 normal = v3(0.0F, 1.0F, 0.0F);
    if (distance > 0.0F) {
        normal = v3_scale(v3_sub(pb, pa), 1.0F / distance);
    }
    out->point_a = pa;
    out->point_b = pb;
    out->normal = normal;
    out->distance = distance;
}

// ---------------------------------------------------------------------------
// Manifold reduction (collide.c:62-273)
// ---------------------------------------------------------------------------
NPH_HD f32 signed_area(NphVec3 a, NphVec3 b, NphVec3 c, NphVec3 normal)
{
    return v3_dot(v3_cross(v3_sub(b, a), v3_sub(c, a)), normal);
}

NPH_HD f32 plane_distance_sq(NphVec3 a, NphVec3 b, NphVec3 normal)
{
    return v3_len_sq(v3_reject(v3_sub(a, b), normal));
}

NPH_HD void swap_points(NphMPoint *points, u32 i, u32 j)
{
    const NphMPoint t = points[i];
    points[i] = points[j];
    points[j] = t;
}

typedef struct NphReduceAxis {
    NphVec3 origin;
    NphVec3 across;
    NphVec3 normal;
    int about_point;
    f32 side;
    f32 min_arm;
} NphReduceAxis;

NPH_HD f32 reduce_arm(const NphReduceAxis *ax, NphVec3 p)
{
    const NphVec3 d = v3_sub(p, ax->origin);
    f32 arm;
    if (ax->about_point) {
        arm = v3_len_sq(v3_reject(d, ax->normal));
    } else {
        const f32 a = v3_dot(d, ax->across);
        arm = ax->side == 0.0F ? nph_absf(a) : ax->side * a;
    }
    return arm > ax->min_arm ? arm : 0.0F;
}

NPH_HD u32 catches_first(const NphMPoint *points, u32 from, u32 end, f32 level,
                         const NphReduceAxis *ax)
{
    u32 level_best = end;
    f32 level_arm = 0.0F;
    u32 least = end;
    f32 least_height = 0.0F;
    f32 least_arm = 0.0F;
    for (u32 i = from; i < end; ++i) {
        const f32 arm = reduce_arm(ax, points[i].point);
        if (!(arm > 0.0F)) {
            continue;
        }
        f32 height = points[i].separation - level;
        if (!(height > 0.0F)) {
            if (arm > level_arm) {
                level_arm = arm;
                level_best = i;
            }
            continue;
        }
        height = ax->about_point ? height * height : height;
        if (least == end || height * least_arm < least_height * arm) {
            least = i;
            least_height = height;
            least_arm = arm;
        }
    }
    if (level_best != end || least == end) {
        return level_best;
    }
    const f32 tie = ax->about_point ? (1.0F + NPH_REDUCE_ANGLE_TIE) * (1.0F + NPH_REDUCE_ANGLE_TIE)
                                    : 1.0F + NPH_REDUCE_ANGLE_TIE;
    u32 best = least;
    f32 best_arm = least_arm;
    for (u32 i = from; i < end; ++i) {
        const f32 arm = reduce_arm(ax, points[i].point);
        if (arm > best_arm) {
            f32 height = points[i].separation - level;
            height = ax->about_point ? height * height : height;
            if (height * least_arm <= tie * least_height * arm) {
                best_arm = arm;
                best = i;
            }
        }
    }
    return best;
}

NPH_HD u32 manifold_reduce(NphMPoint *points, u32 count, NphVec3 normal)
{
    if (count <= NPH_PHYSICS_MANIFOLD_MAX_POINTS) {
        return count;
    }
    f32 deepest = points[0].separation;
    NphVec3 centre = points[0].point;
    for (u32 i = 1U; i < count; ++i) {
        deepest = nph_minf(deepest, points[i].separation);
        centre = v3_add(centre, points[i].point);
    }
    centre = v3_scale(centre, 1.0F / (f32)count);
    {
        u32 best = 0U;
        f32 far = -1.0F;
        for (u32 i = 0U; i < count; ++i) {
            const f32 dd = plane_distance_sq(points[i].point, centre, normal);
            if (points[i].separation <= deepest + NPH_REDUCE_LEVEL && dd > far) {
                far = dd;
                best = i;
            }
        }
        swap_points(points, 0U, best);
    }
    const f32 level = nph_maxf(deepest, 0.0F) + NPH_REDUCE_LEVEL;
    const NphVec3 p0 = points[0].point;
    NphReduceAxis ax;
    ax.origin = p0;
    ax.across = v3_zero();
    ax.normal = normal;
    ax.about_point = 1;
    ax.side = 0.0F;
    ax.min_arm = NPH_REDUCE_AT_POINT * NPH_REDUCE_AT_POINT;
    const u32 second = catches_first(points, 1U, count, level, &ax);
    if (second == count) {
        return 1U;
    }
    swap_points(points, 1U, second);
    const NphVec3 p1 = points[1].point;
    const NphVec3 line = v3_reject(v3_sub(p1, p0), normal);
    const f32 line_length = v3_len(line);
    ax.about_point = 0;
    ax.across = v3_scale(v3_cross(normal, line), 1.0F / line_length);
    ax.min_arm = NPH_REDUCE_ON_LINE * line_length;
    const u32 third = catches_first(points, 2U, count, level, &ax);
    if (third == count) {
        return 2U;
    }
    swap_points(points, 2U, third);
    const NphVec3 p2 = points[2].point;
    ax.side = v3_dot(v3_sub(p2, p0), ax.across) < 0.0F ? 1.0F : -1.0F;
    u32 fourth = catches_first(points, 3U, count, level, &ax);
    if (fourth == count) {
        const f32 sign = signed_area(p0, p1, p2, normal) < 0.0F ? -1.0F : 1.0F;
        f32 best_out = -NPH_F32_MAX;
        fourth = 3U;
        for (u32 i = 3U; i < count; ++i) {
            const NphVec3 p = points[i].point;
            const f32 s0 = sign * signed_area(p0, p1, p, normal);
            const f32 s1 = sign * signed_area(p1, p2, p, normal);
            const f32 s2 = sign * signed_area(p2, p0, p, normal);
            const f32 out = -nph_minf(s0, nph_minf(s1, s2));
            if (out > best_out) {
                best_out = out;
                fourth = i;
            }
        }
    }
    swap_points(points, 3U, fourth);
    return 4U;
}

// ---------------------------------------------------------------------------
// Spheres, capsules, box-sphere (collide.c:277-454)
// ---------------------------------------------------------------------------
NPH_HD void collide_balls(NphVec3 ca, f32 ra, NphVec3 cb, f32 rb, f32 speculative, u32 id,
                          NphVec3 fallback_normal, NphManifold *out)
{
    const NphVec3 d = v3_sub(cb, ca);
    const f32 dist_sq = v3_len_sq(d);
    const f32 limit = ra + rb + speculative;
    if (dist_sq > limit * limit) {
        return;
    }
    const f32 dist = nph_sqrtf(dist_sq);
    const NphVec3 n = dist > 1e-9F ? v3_scale(d, 1.0F / dist) : fallback_normal;
    const NphVec3 sa = v3_mul_add(ca, n, ra);
    const NphVec3 sb = v3_mul_add(cb, n, -rb);
    out->normal = n;
    manifold_push(out, v3_scale(v3_add(sa, sb), 0.5F), dist - ra - rb, id);
}

NPH_HD void capsule_segment(f32 half_height, NphPhysicsPose pose, NphVec3 *p, NphVec3 *q)
{
    const NphVec3 axis = v3_scale(q_axis_y(pose.rotation), half_height);
    *p = v3_sub(pose.position, axis);
    *q = v3_add(pose.position, axis);
}

NPH_HD void collide_capsule_sphere(f32 hh, f32 ra, NphPhysicsPose pa, f32 rb, NphPhysicsPose pb,
                                   f32 speculative, NphManifold *out)
{
    NphVec3 p;
    NphVec3 q;
    capsule_segment(hh, pa, &p, &q);
    const NphVec3 c = closest_on_segment(p, q, pb.position);
    collide_balls(c, ra, pb.position, rb, speculative, 0U, q_axis_x(pa.rotation), out);
}

NPH_HD void collide_capsules(f32 hha, f32 ra, NphPhysicsPose pa, f32 hhb, f32 rb,
                             NphPhysicsPose pb, f32 speculative, NphManifold *out)
{
    NphVec3 a0;
    NphVec3 a1;
    NphVec3 b0;
    NphVec3 b1;
    capsule_segment(hha, pa, &a0, &a1);
    capsule_segment(hhb, pb, &b0, &b1);
    const NphSegDist sd = segment_distance(a0, a1, b0, b1);
    const f32 limit = ra + rb + speculative;
    if (sd.distance_squared > limit * limit) {
        return;
    }
    const NphVec3 da = v3_sub(a1, a0);
    const NphVec3 db = v3_sub(b1, b0);
    const f32 la_sq = v3_len_sq(da);
    const f32 lb_sq = v3_len_sq(db);
    const f32 dist = nph_sqrtf(sd.distance_squared);
    NphVec3 n;
    if (dist > 1e-6F) {
        n = v3_scale(v3_sub(sd.closest2, sd.closest1), 1.0F / dist);
    } else {
        NphVec3 c = v3_cross(da, db);
        if (v3_len_sq(c) <= 1e-12F * nph_maxf(la_sq * lb_sq, 1e-12F)) {
            c = la_sq > 0.0F ? v3_perpendicular(v3_normalize(da)) : v3(0.0F, 1.0F, 0.0F);
        }
        n = v3_normalize_or(c, v3(0.0F, 1.0F, 0.0F));
        if (v3_dot(n, v3_sub(pb.position, pa.position)) < 0.0F) {
            n = v3_neg(n);
        }
    }
    const f32 cross_sq = v3_len_sq(v3_cross(da, db));
    if (la_sq > 1e-12F && lb_sq > 1e-12F && cross_sq <= 1e-4F * la_sq * lb_sq) {
        const f32 la = nph_sqrtf(la_sq);
        const NphVec3 ua = v3_scale(da, 1.0F / la);
        const f32 t0 = v3_dot(v3_sub(b0, a0), ua);
        const f32 t1 = v3_dot(v3_sub(b1, a0), ua);
        const f32 lo = nph_maxf(0.0F, nph_minf(t0, t1));
        const f32 hi = nph_minf(la, nph_maxf(t0, t1));
        if (hi - lo > NPH_PHYSICS_LINEAR_SLOP) {
            const f32 ts[2] = {lo, hi};
            out->normal = n;
            for (u32 k = 0U; k < 2U; ++k) {
                const NphVec3 pa_core = v3_mul_add(a0, ua, ts[k]);
                const NphVec3 pb_core = closest_on_segment(b0, b1, pa_core);
                const f32 sep = v3_dot(v3_sub(pb_core, pa_core), n) - ra - rb;
                if (sep <= speculative) {
                    const NphVec3 sa = v3_mul_add(pa_core, n, ra);
                    const NphVec3 sb = v3_mul_add(pb_core, n, -rb);
                    manifold_push(out, v3_scale(v3_add(sa, sb), 0.5F), sep, k);
                }
            }
            if (out->point_count > 0U) {
                return;
            }
        }
    }
    const NphVec3 sa = v3_mul_add(sd.closest1, n, ra);
    const NphVec3 sb = v3_mul_add(sd.closest2, n, -rb);
    out->normal = n;
    manifold_push(out, v3_scale(v3_add(sa, sb), 0.5F), dist - ra - rb, 2U);
}

NPH_HD void collide_box_sphere(NphVec3 h, NphPhysicsPose pa, f32 rb, NphPhysicsPose pb,
                               f32 speculative, NphManifold *out)
{
    const NphVec3 c = pose_inv_point(pa, pb.position);
    const NphVec3 q = v3_clamp(c, v3_neg(h), h);
    const NphVec3 d = v3_sub(c, q);
    const f32 dist_sq = v3_len_sq(d);
    NphVec3 n_local;
    NphVec3 surface_a;
    f32 sep;
    u32 id;
    if (dist_sq > 0.0F) {
        const f32 limit = rb + speculative;
        if (dist_sq > limit * limit) {
            return;
        }
        const f32 dist = nph_sqrtf(dist_sq);
        n_local = v3_scale(d, 1.0F / dist);
        surface_a = q;
        sep = dist - rb;
        id = 0U;
    } else {
        const NphVec3 depth = v3_sub(h, v3_abs(c));
        u32 axis = 0U;
        if (depth.y < depth.x && depth.y <= depth.z) {
            axis = 1U;
        } else if (depth.z < depth.x && depth.z < depth.y) {
            axis = 2U;
        }
        const f32 s = v3_at(c, axis) < 0.0F ? -1.0F : 1.0F;
        n_local = v3_zero();
        surface_a = c;
        if (axis == 0U) {
            n_local.x = s;
            surface_a.x = s * h.x;
        } else if (axis == 1U) {
            n_local.y = s;
            surface_a.y = s * h.y;
        } else {
            n_local.z = s;
            surface_a.z = s * h.z;
        }
        sep = -v3_at(depth, axis) - rb;
        id = 1U + axis;
    }
    const NphVec3 surface_b = v3_mul_add(c, n_local, -rb);
    const NphVec3 mid = v3_scale(v3_add(surface_a, surface_b), 0.5F);
    out->normal = pose_vector(pa, n_local);
    manifold_push(out, pose_point(pa, mid), sep, id);
}

// ---------------------------------------------------------------------------
// Hull against sphere and capsule (collide_hull.c:156-376)
// ---------------------------------------------------------------------------
NPH_HD void collide_hull_sphere(const NphView *a, NphPhysicsPose pa, f32 radius,
                                NphPhysicsPose pb, f32 speculative, NphManifold *out)
{
    const NphVec3 c = pose_inv_point(pa, pb.position);
    f32 best = -NPH_F32_MAX;
    u32 face = 0U;
    for (u32 f = 0U; f < a->face_count; ++f) {
        const f32 s = plane_distance(a->planes[f], c);
        if (s > best) {
            best = s;
            face = f;
        }
    }
    if (best > radius + speculative) {
        return;
    }
    NphVec3 n;
    NphVec3 surface_a;
    f32 sep;
    u32 id;
    if (best <= 0.0F) {
        n = a->planes[face].normal;
        surface_a = v3_mul_add(c, n, -best);
        sep = best - radius;
        id = face;
    } else {
        NphDistance d;
        gjk_distance(a->vertices, a->vertex_count, &c, 1U, &d);
        if (d.distance > radius + speculative) {
            return;
        }
        n = d.distance > 1e-6F ? d.normal : a->planes[face].normal;
        surface_a = d.point_a;
        sep = d.distance - radius;
        id = 0x10000U | face;
    }
    const NphVec3 surface_b = v3_mul_add(c, n, -radius);
    out->normal = pose_vector(pa, n);
    manifold_push(out, pose_point(pa, v3_scale(v3_add(surface_a, surface_b), 0.5F)), sep, id);
}

NPH_HD int clip_segment_to_face(const NphView *h, u32 face, NphVec3 *p, NphVec3 *q)
{
    const NphVec3 n = h->planes[face].normal;
    f32 t0 = 0.0F;
    f32 t1 = 1.0F;
    const NphVec3 p0 = *p;
    const NphVec3 d = v3_sub(*q, *p);
    const u32 first = h->faces[face];
    u32 e = first;
    u32 guard = 0U;
    do {
        const NphVec3 a = h->vertices[h->edges[e].origin];
        const NphVec3 b = edge_head(h, h->vertices, e);
        const NphVec3 side = v3_cross(v3_sub(b, a), n);
        const f32 s0 = v3_dot(side, v3_sub(p0, a));
        const f32 sd = v3_dot(side, d);
        if (sd == 0.0F) {
            if (s0 > 0.0F) {
                return 0;
            }
        } else {
            const f32 t = -s0 / sd;
            if (sd > 0.0F) {
                t1 = nph_minf(t1, t);
            } else {
                t0 = nph_maxf(t0, t);
            }
            if (t0 > t1) {
                return 0;
            }
        }
        e = h->edges[e].next;
        guard += 1U;
    } while (e != first && guard < h->edge_count);
    *p = v3_mul_add(p0, d, t0);
    *q = v3_mul_add(p0, d, t1);
    return 1;
}

NPH_HD void collide_hull_capsule(const NphView *a, NphPhysicsPose pa, f32 half_height, f32 radius,
                                 NphPhysicsPose pb, f32 speculative, NphManifold *out)
{
    const NphPhysicsPose rel = pose_inv_mul(pa, pb);
    const NphVec3 axis = v3_scale(q_axis_y(rel.rotation), half_height);
    const NphVec3 s0 = v3_sub(rel.position, axis);
    const NphVec3 s1 = v3_add(rel.position, axis);
    const f32 limit = radius + speculative;
    f32 face_sep = -NPH_F32_MAX;
    u32 face = 0U;
    for (u32 f = 0U; f < a->face_count; ++f) {
        const NphPlane pl = a->planes[f];
        const f32 s = nph_minf(plane_distance(pl, s0), plane_distance(pl, s1));
        if (s > face_sep) {
            face_sep = s;
            face = f;
        }
    }
    if (face_sep > limit) {
        return;
    }
    const NphVec3 d = v3_sub(s1, s0);
    const f32 d_len_sq = v3_len_sq(d);
    f32 edge_sep = -NPH_F32_MAX;
    u32 edge = 0U;
    NphVec3 edge_axis = v3(0.0F, 1.0F, 0.0F);
    if (d_len_sq > 1e-12F) {
        for (u32 e = 0U; e < a->edge_count; e += 2U) {
            const NphVec3 n1 = a->planes[a->edges[e].face].normal;
            const NphVec3 n2 = a->planes[a->edges[e + 1U].face].normal;
            if (v3_dot(n1, d) * v3_dot(n2, d) >= 0.0F) {
                continue;
            }
            const NphVec3 p = a->vertices[a->edges[e].origin];
            const NphVec3 u = v3_sub(edge_head(a, a->vertices, e), p);
            NphVec3 x = v3_cross(u, d);
            const f32 len = v3_len(x);
            if (len < 0.005F * nph_sqrtf(v3_len_sq(u) * d_len_sq)) {
                continue;
            }
            x = v3_scale(x, 1.0F / len);
            if (v3_dot(x, v3_sub(p, a->centroid)) < 0.0F) {
                x = v3_neg(x);
            }
            const f32 s = v3_dot(x, v3_sub(s0, p));
            if (s > edge_sep) {
                edge_sep = s;
                edge = e;
                edge_axis = x;
            }
        }
    }
    if (edge_sep > limit) {
        return;
    }
    const int use_face = face_sep >= edge_sep - NPH_EDGE_BIAS;
    const f32 best = use_face ? face_sep : edge_sep;
    if (best > NPH_PHYSICS_LINEAR_SLOP) {
        const NphVec3 seg[2] = {s0, s1};
        NphDistance gd;
        gjk_distance(a->vertices, a->vertex_count, seg, d_len_sq > 0.0F ? 2U : 1U, &gd);
        if (gd.distance > limit) {
            return;
        }
        const NphVec3 sat_axis = use_face ? a->planes[face].normal : edge_axis;
        if (gd.distance > 0.0F &&
            (!use_face || v3_dot(gd.normal, sat_axis) < NPH_FEATURE_AGREEMENT)) {
            const NphVec3 surface_b = v3_mul_add(gd.point_b, gd.normal, -radius);
            const NphVec3 mid = v3_scale(v3_add(gd.point_a, surface_b), 0.5F);
            out->normal = pose_vector(pa, gd.normal);
            manifold_push(out, pose_point(pa, mid), gd.distance - radius, 0x20000U);
            return;
        }
    }
    if (use_face) {
        const NphPlane pl = a->planes[face];
        NphVec3 p = s0;
        NphVec3 q = s1;
        if (!clip_segment_to_face(a, face, &p, &q)) {
            p = plane_distance(pl, s0) <= plane_distance(pl, s1) ? s0 : s1;
            q = p;
        }
        out->normal = pose_vector(pa, pl.normal);
        const NphVec3 pts[2] = {p, q};
        const u32 count = v3_dist_sq(p, q) > 1e-12F ? 2U : 1U;
        for (u32 k = 0U; k < count; ++k) {
            const f32 dist = plane_distance(pl, pts[k]);
            const f32 sep = dist - radius;
            if (sep <= speculative) {
                const NphVec3 mid = v3_mul_add(pts[k], pl.normal, -0.5F * (dist + radius));
                manifold_push(out, pose_point(pa, mid), sep, (face << 8U) | k);
            }
        }
        return;
    }
    const NphVec3 ep = a->vertices[a->edges[edge].origin];
    const NphVec3 eq = edge_head(a, a->vertices, edge);
    const NphSegDist sd = segment_distance(ep, eq, s0, s1);
    const f32 sep = v3_dot(v3_sub(sd.closest2, sd.closest1), edge_axis) - radius;
    if (sep > speculative) {
        return;
    }
    const NphVec3 surface_b = v3_mul_add(sd.closest2, edge_axis, -radius);
    const NphVec3 mid = v3_scale(v3_add(sd.closest1, surface_b), 0.5F);
    out->normal = pose_vector(pa, edge_axis);
    manifold_push(out, pose_point(pa, mid), sep, 0x10000U | edge);
}

// ---------------------------------------------------------------------------
// Hull against hull (collide_hull.c:378-1002)
// ---------------------------------------------------------------------------
typedef struct NphSat {
    f32 face_a_sep;
    u32 face_a;
    f32 face_b_sep;
    u32 face_b;
    f32 edge_sep;
    u32 edge_a;
    u32 edge_b;
    NphVec3 edge_axis;
} NphSat;

typedef struct NphHullPair {
    const NphView *a;
    const NphView *b;
    NphPhysicsPose rel;
    NphVec3 bv[NPH_HULL_MAX_VERTICES];
} NphHullPair;

// The smallest projection in the reference's four-lane order
// (collide_hull.c:516-534): lane j keeps the minimum of points j, j+4, ...
// (the last point repeated to a multiple of four), then lanes (0,2),(1,3)
// and the two results fold, each min as (a < b ? a : b) in that argument
// order.  Equal to the scalar minimum except possibly in the sign of zero.
NPH_HD f32 min_projection4(const NphVec3 *v, u32 count, NphVec3 n)
{
    f32 m[4] = {NPH_F32_MAX, NPH_F32_MAX, NPH_F32_MAX, NPH_F32_MAX};
    const u32 count4 = (count + 3U) & ~3U;
    for (u32 i = 0U; i < count4; i += 4U) {
        for (u32 j = 0U; j < 4U; ++j) {
            const NphVec3 p = v[i + j < count ? i + j : count - 1U];
            const f32 d = n.x * p.x + n.y * p.y + n.z * p.z;
            m[j] = nph_minf(m[j], d);
        }
    }
    const f32 a0 = nph_minf(m[0], m[2]);
    const f32 a1 = nph_minf(m[1], m[3]);
    return nph_minf(a0, a1);
}

// Scalar loop of collide_hull.c:378-385 (the cache's path).
NPH_HD f32 min_projection(const NphVec3 *v, u32 count, NphVec3 n)
{
    f32 m = v3_dot(v[0], n);
    for (u32 i = 1U; i < count; ++i) {
        m = nph_minf(m, v3_dot(v[i], n));
    }
    return m;
}

NPH_HD NphPlane plane_b_in_a(const NphHullPair *hp, u32 f)
{
    const NphPlane pl = hp->b->planes[f];
    const NphVec3 n = q_rotate(hp->rel.rotation, pl.normal);
    return plane_make(n, pl.d - v3_dot(n, hp->rel.position));
}

NPH_HD NphVec3 arc_plane(NphVec3 c, NphVec3 d, NphVec3 origin, NphVec3 head)
{
    const NphVec3 u = v3_cross(d, c);
    return v3_len_sq(u) > 1e-12F ? u : v3_sub(origin, head);
}

NPH_HD f32 edge_pair_separation(NphVec3 pa, NphVec3 ua, NphVec3 pb, NphVec3 ub, NphVec3 centroid,
                                NphVec3 *out_axis)
{
    const NphVec3 x = v3_cross(ua, ub);
    const f32 len = v3_len(x);
    if (len < 0.005F * nph_sqrtf(v3_len_sq(ua) * v3_len_sq(ub))) {
        return -NPH_F32_MAX;
    }
    NphVec3 n = v3_scale(x, 1.0F / len);
    if (v3_dot(n, v3_sub(pa, centroid)) < 0.0F) {
        n = v3_neg(n);
    }
    *out_axis = n;
    return v3_dot(n, v3_sub(pb, pa));
}

NPH_HD f32 edge_pair_separation_above(NphVec3 pa, NphVec3 ua, NphVec3 pb, NphVec3 ub,
                                      NphVec3 centroid, f32 best, NphVec3 *out_axis)
{
    NphVec3 x = v3_cross(ua, ub);
    const f32 len_sq = v3_len_sq(x);
    if (len_sq < 2.5e-5F * v3_len_sq(ua) * v3_len_sq(ub)) {
        return -NPH_F32_MAX;
    }
    if (v3_dot(x, v3_sub(pa, centroid)) < 0.0F) {
        x = v3_neg(x);
    }
    const f32 sx = v3_dot(x, v3_sub(pb, pa));
    if (best > -NPH_F32_MAX) {
        const f32 bound = best * best * len_sq;
        const int wins =
            best < 0.0F ? (sx >= 0.0F || sx * sx < bound) : (sx > 0.0F && sx * sx > bound);
        if (!wins) {
            return -NPH_F32_MAX;
        }
    }
    const f32 inv = 1.0F / nph_sqrtf(len_sq);
    *out_axis = v3_scale(x, inv);
    return sx * inv;
}

// collide_hull.c:568-779 (the four-lane path: every hull here has at most 64
// vertices, so the reference takes it; its Gauss-map test is the scalar
// is_minkowski_face in the same operation order).
NPH_HD int sat_query(const NphHullPair *hp, f32 limit, NphCache *cache, NphSat *r)
{
    const NphView *a = hp->a;
    const NphView *b = hp->b;
    if (cache != 0 && cache->type != NPH_CACHE_NONE) {
        f32 s = -NPH_F32_MAX;
        if (cache->type == NPH_CACHE_FACE_A && cache->index_a < a->face_count) {
            const NphPlane pl = a->planes[cache->index_a];
            s = min_projection(hp->bv, b->vertex_count, pl.normal) + pl.d;
        } else if (cache->type == NPH_CACHE_FACE_B && cache->index_b < b->face_count) {
            const NphPlane pl = plane_b_in_a(hp, cache->index_b);
            s = min_projection(a->vertices, a->vertex_count, pl.normal) + pl.d;
        } else if (cache->type == NPH_CACHE_EDGES && cache->index_a < a->edge_count &&
                   cache->index_b < b->edge_count) {
            const u32 ea = cache->index_a;
            const u32 eb = cache->index_b;
            const NphVec3 pa = a->vertices[a->edges[ea].origin];
            const NphVec3 pb = hp->bv[b->edges[eb].origin];
            NphVec3 axis;
            s = edge_pair_separation(pa, v3_sub(edge_head(a, a->vertices, ea), pa), pb,
                                     v3_sub(edge_head(b, hp->bv, eb), pb), a->centroid, &axis);
        }
        if (s > limit) {
            return 0;
        }
    }
    r->face_a_sep = -NPH_F32_MAX;
    r->face_a = 0U;
    for (u32 f = 0U; f < a->face_count; ++f) {
        const NphPlane pl = a->planes[f];
        const f32 s = min_projection4(hp->bv, b->vertex_count, pl.normal) + pl.d;
        if (s > r->face_a_sep) {
            r->face_a_sep = s;
            r->face_a = f;
        }
    }
    if (r->face_a_sep > limit) {
        if (cache != 0) {
            cache->type = NPH_CACHE_FACE_A;
            cache->index_a = (u16)r->face_a;
        }
        return 0;
    }
#if NPH_CACHE_B_NORMALS
    NphVec3 bn[NPH_HULL_MAX_FACES];
#endif
    r->face_b_sep = -NPH_F32_MAX;
    r->face_b = 0U;
    for (u32 f = 0U; f < b->face_count; ++f) {
        const NphPlane pl = plane_b_in_a(hp, f);
#if NPH_CACHE_B_NORMALS
        bn[f] = pl.normal;
#endif
        const f32 s = min_projection4(a->vertices, a->vertex_count, pl.normal) + pl.d;
        if (s > r->face_b_sep) {
            r->face_b_sep = s;
            r->face_b = f;
        }
    }
    if (r->face_b_sep > limit) {
        if (cache != 0) {
            cache->type = NPH_CACHE_FACE_B;
            cache->index_b = (u16)r->face_b;
        }
        return 0;
    }
    r->edge_sep = -NPH_F32_MAX;
    r->edge_a = 0U;
    r->edge_b = 0U;
    r->edge_axis = v3(0.0F, 1.0F, 0.0F);
    for (u32 ea = 0U; ea < a->edge_count; ea += 2U) {
        const NphVec3 na1 = a->planes[a->edges[ea].face].normal;
        const NphVec3 na2 = a->planes[a->edges[ea + 1U].face].normal;
        const NphVec3 bxa = v3_cross(na2, na1);
        const NphVec3 pa = a->vertices[a->edges[ea].origin];
        const NphVec3 ua = v3_sub(edge_head(a, a->vertices, ea), pa);
        for (u32 eb = 0U; eb < b->edge_count; eb += 2U) {
#if NPH_CACHE_B_NORMALS
            const NphVec3 c = v3_neg(bn[b->edges[eb].face]);
            const NphVec3 d = v3_neg(bn[b->edges[eb + 1U].face]);
#else
            const NphVec3 c = v3_neg(plane_b_in_a(hp, b->edges[eb].face).normal);
            const NphVec3 d = v3_neg(plane_b_in_a(hp, b->edges[eb + 1U].face).normal);
#endif
            const NphVec3 pb = hp->bv[b->edges[eb].origin];
            const NphVec3 hb = edge_head(b, hp->bv, eb);
            const NphVec3 u = arc_plane(c, d, pb, hb);
            const f32 cba = v3_dot(c, bxa);
            const f32 dba = v3_dot(d, bxa);
            const f32 adc = v3_dot(na1, u);
            const f32 bdc = v3_dot(na2, u);
            if (!(cba * dba < 0.0F && adc * bdc < 0.0F && cba * bdc > 0.0F)) {
                continue;
            }
            const NphVec3 ub = v3_sub(hb, pb);
            NphVec3 axis;
            const f32 s = edge_pair_separation_above(pa, ua, pb, ub, a->centroid, r->edge_sep, &axis);
            if (s > r->edge_sep) {
                r->edge_sep = s;
                r->edge_a = ea;
                r->edge_b = eb;
                r->edge_axis = axis;
            }
        }
    }
    if (r->edge_sep > limit) {
        if (cache != 0) {
            cache->type = NPH_CACHE_EDGES;
            cache->index_a = (u16)r->edge_a;
            cache->index_b = (u16)r->edge_b;
        }
        return 0;
    }
    return 1;
}

// collide_hull.c:787-877: the reference face `ref` of R (A when !flip), the
// incident face clipped against its side planes; at most `max_points`.
// A convex shape's contact with one triangle, before its reduction: up to
// sixteen points (physics_internal.h:65-89).
#define NPH_TRI_POINTS 16U

typedef struct NphTriContact {
    NphVec3 normal; // from the shape towards the triangle
    u32 point_count;
    NphMPoint points[NPH_TRI_POINTS];
} NphTriContact;

NPH_HD void tri_clear(NphTriContact *c)
{
    c->normal = v3(0.0F, 1.0F, 0.0F);
    c->point_count = 0U;
}

NPH_HD void tri_push(NphTriContact *c, NphVec3 point, f32 separation, u32 id)
{
    if (c->point_count < NPH_TRI_POINTS) {
        NphMPoint *p = &c->points[c->point_count++];
        p->point = point;
        p->separation = separation;
        p->id = id;
    }
}

// collide_hull.c:786-877: more than `max_points` candidates are reduced to
// the manifold's four.
NPH_HD NphVec3 face_contact(const NphHullPair *hp, int flip, u32 ref, f32 speculative,
                            u32 max_points, NphMPoint *points, u32 *count)
{
    const NphView *rh = flip ? hp->b : hp->a;
    const NphView *ih = flip ? hp->a : hp->b;
    const NphVec3 *rv = flip ? hp->bv : hp->a->vertices;
    const NphVec3 *iv = flip ? hp->a->vertices : hp->bv;
    const NphPlane ref_plane = flip ? plane_b_in_a(hp, ref) : hp->a->planes[ref];
    const NphVec3 n = ref_plane.normal;
    u32 inc = 0U;
    f32 best = NPH_F32_MAX;
    for (u32 f = 0U; f < ih->face_count; ++f) {
        const NphVec3 fn = flip ? ih->planes[f].normal : plane_b_in_a(hp, f).normal;
        const f32 d = v3_dot(fn, n);
        if (d < best) {
            best = d;
            inc = f;
        }
    }
    NphMPoint buf[2][NPH_CLIP_CAPACITY];
    u32 len = 0U;
    {
        const u32 first = ih->faces[inc];
        u32 e = first;
        do {
            buf[0][len].point = iv[ih->edges[e].origin];
            buf[0][len].separation = 0.0F;
            buf[0][len].id = ih->edges[e].origin;
            len += 1U;
            e = ih->edges[e].next;
        } while (e != first && len < NPH_CLIP_CAPACITY && len < ih->edge_count);
    }
    u32 cur = 0U;
    const u32 first = rh->faces[ref];
    u32 e = first;
    u32 side = 0U;
    do {
        const NphVec3 a = rv[rh->edges[e].origin];
        const NphVec3 b = edge_head(rh, rv, e);
        const NphVec3 sn = v3_cross(v3_sub(b, a), n);
        const f32 so = v3_dot(sn, a);
        const NphMPoint *in = buf[cur];
        NphMPoint *outp = buf[cur ^ 1U];
        u32 out_len = 0U;
        for (u32 i = 0U; i < len && out_len + 1U < NPH_CLIP_CAPACITY; ++i) {
            const NphMPoint p = in[i];
            const NphMPoint q = in[(i + 1U) % len];
            const f32 dp = v3_dot(sn, p.point) - so;
            const f32 dq = v3_dot(sn, q.point) - so;
            if (dp <= 0.0F) {
                outp[out_len++] = p;
            }
            if ((dp <= 0.0F) != (dq <= 0.0F)) {
                const f32 t = dp / (dp - dq);
                outp[out_len].point = v3_lerp(p.point, q.point, t);
                outp[out_len].separation = 0.0F;
                outp[out_len].id = nph_mix_id(nph_mix_id(p.id, q.id), 0x100U + side);
                out_len += 1U;
            }
        }
        len = out_len;
        cur ^= 1U;
        e = rh->edges[e].next;
        side += 1U;
    } while (e != first && len > 0U && side < rh->edge_count);
    const u32 tag = nph_mix_id((flip ? 0x80000000U : 0U) | (ref << 16U), inc);
    const NphMPoint *clipped = buf[cur];
    NphMPoint *cand = buf[cur ^ 1U];
    u32 n_out = 0U;
    for (u32 i = 0U; i < len; ++i) {
        const f32 sep = plane_distance(ref_plane, clipped[i].point);
        if (sep <= speculative) {
            cand[n_out].point = v3_mul_add(clipped[i].point, n, -0.5F * sep);
            cand[n_out].separation = sep;
            cand[n_out].id = nph_mix_id(tag, clipped[i].id);
            n_out += 1U;
        }
    }
    if (n_out > max_points) {
        n_out = manifold_reduce(cand, n_out, n);
    }
    for (u32 i = 0U; i < n_out; ++i) {
        points[i] = cand[i];
    }
    *count = n_out;
    return flip ? v3_neg(n) : n;
}

// collide_hull.c:921-1002: the pair function, keeping up to `max_points`
// points; `prefer_b` turns the tie between faces to B's (a triangle's face).
// `out` is left as it was when nothing touches.
NPH_HD void collide_hulls_into(const NphView *a, NphPhysicsPose pa, const NphView *b,
                               NphPhysicsPose pb, f32 speculative, NphCache *cache, int prefer_b,
                               u32 max_points, NphTriContact *out)
{
    NphHullPair hp;
    hp.a = a;
    hp.b = b;
    hp.rel = pose_inv_mul(pa, pb);
    for (u32 i = 0U; i < b->vertex_count; ++i) {
        hp.bv[i] = pose_point(hp.rel, b->vertices[i]);
    }
    NphSat r;
    if (!sat_query(&hp, speculative, cache, &r)) {
        return;
    }
    const f32 face_sep = nph_maxf(r.face_a_sep, r.face_b_sep);
    const int use_edge = r.edge_sep > face_sep + NPH_EDGE_BIAS;
    const int flip = !use_edge && (prefer_b ? r.face_b_sep >= r.face_a_sep - NPH_FACE_BIAS
                                            : r.face_b_sep > r.face_a_sep + NPH_FACE_BIAS);
    if (cache != 0) {
        cache->type = (u8)(use_edge ? NPH_CACHE_EDGES : (flip ? NPH_CACHE_FACE_B : NPH_CACHE_FACE_A));
        cache->index_a = (u16)(use_edge ? r.edge_a : r.face_a);
        cache->index_b = (u16)(use_edge ? r.edge_b : r.face_b);
    }
    const f32 best = use_edge ? r.edge_sep : (flip ? r.face_b_sep : r.face_a_sep);
    if (best > NPH_PHYSICS_LINEAR_SLOP) {
        NphDistance gd;
        gjk_distance(a->vertices, a->vertex_count, hp.bv, b->vertex_count, &gd);
        if (gd.distance > speculative) {
            return;
        }
        const NphVec3 sat_axis = use_edge ? r.edge_axis
                                          : (flip ? v3_neg(plane_b_in_a(&hp, r.face_b).normal)
                                                  : a->planes[r.face_a].normal);
        if (gd.distance > 0.0F &&
            (use_edge || v3_dot(gd.normal, sat_axis) < NPH_FEATURE_AGREEMENT)) {
            const NphVec3 mid = v3_scale(v3_add(gd.point_a, gd.point_b), 0.5F);
            out->normal = pose_vector(pa, gd.normal);
            tri_push(out, pose_point(pa, mid), gd.distance, 0x40000000U);
            return;
        }
    }
    if (use_edge) {
        const NphVec3 pa0 = a->vertices[a->edges[r.edge_a].origin];
        const NphVec3 pa1 = edge_head(a, a->vertices, r.edge_a);
        const NphVec3 pb0 = hp.bv[b->edges[r.edge_b].origin];
        const NphVec3 pb1 = edge_head(b, hp.bv, r.edge_b);
        const NphSegDist sd = segment_distance(pa0, pa1, pb0, pb1);
        const f32 sep = v3_dot(v3_sub(sd.closest2, sd.closest1), r.edge_axis);
        if (sep > speculative) {
            return;
        }
        const NphVec3 mid = v3_scale(v3_add(sd.closest1, sd.closest2), 0.5F);
        out->normal = pose_vector(pa, r.edge_axis);
        tri_push(out, pose_point(pa, mid), sep, nph_mix_id(0x20000000U | r.edge_a, r.edge_b));
        return;
    }
    NphMPoint pts[NPH_TRI_POINTS];
    u32 count = 0U;
    const NphVec3 n =
        face_contact(&hp, flip, flip ? r.face_b : r.face_a, speculative, max_points, pts, &count);
    out->normal = pose_vector(pa, n);
    for (u32 i = 0U; i < count; ++i) {
        tri_push(out, pose_point(pa, pts[i].point), pts[i].separation, pts[i].id);
    }
}

// collide_hull.c:1004-1017: a convex pair, four points at most.
NPH_HD void collide_hulls(const NphView *a, NphPhysicsPose pa, const NphView *b,
                          NphPhysicsPose pb, f32 speculative, NphCache *cache, NphManifold *out)
{
    NphTriContact c;
    tri_clear(&c);
    collide_hulls_into(a, pa, b, pb, speculative, cache, 0, NPH_PHYSICS_MANIFOLD_MAX_POINTS, &c);
    if (c.point_count > 0U) {
        out->normal = c.normal;
        for (u32 i = 0U; i < c.point_count; ++i) {
            manifold_push(out, c.points[i].point, c.points[i].separation, c.points[i].id);
        }
    }
}

// collide_hull.c:879-904: B's face `face` against A, as the face contact of
// the pair with B's face the reference.
NPH_HD void collide_hull_face(const NphView *a, NphPhysicsPose pa, const NphView *b,
                              NphPhysicsPose pb, u32 face, f32 speculative, NphTriContact *out)
{
    NphHullPair hp;
    hp.a = a;
    hp.b = b;
    hp.rel = pose_inv_mul(pa, pb);
    for (u32 i = 0U; i < b->vertex_count; ++i) {
        hp.bv[i] = pose_point(hp.rel, b->vertices[i]);
    }
    NphMPoint pts[NPH_TRI_POINTS];
    u32 count = 0U;
    const NphVec3 n = face_contact(&hp, 1, face, speculative, NPH_TRI_POINTS, pts, &count);
    tri_clear(out);
    out->normal = pose_vector(pa, n);
    for (u32 i = 0U; i < count; ++i) {
        tri_push(out, pose_point(pa, pts[i].point), pts[i].separation, pts[i].id);
    }
}

// ---------------------------------------------------------------------------
// Box against box (collide_hull.c:1004-1266)
// ---------------------------------------------------------------------------
typedef struct NphBoxAxes {
    f32 r[3][3];
    f32 ar[3][3];
    f32 t[3];
    f32 ha[3];
    f32 hb[3];
} NphBoxAxes;

NPH_HD f32 box_face_a(const NphBoxAxes *x, u32 i, f32 *sign)
{
    const f32 rb = x->hb[0] * x->ar[i][0] + x->hb[1] * x->ar[i][1] + x->hb[2] * x->ar[i][2];
    *sign = x->t[i] >= 0.0F ? 1.0F : -1.0F;
    return nph_absf(x->t[i]) - (x->ha[i] + rb);
}

NPH_HD f32 box_face_b(const NphBoxAxes *x, u32 j, f32 *sign)
{
    const f32 tb = x->t[0] * x->r[0][j] + x->t[1] * x->r[1][j] + x->t[2] * x->r[2][j];
    const f32 ra = x->ha[0] * x->ar[0][j] + x->ha[1] * x->ar[1][j] + x->ha[2] * x->ar[2][j];
    *sign = tb >= 0.0F ? 1.0F : -1.0F;
    return nph_absf(tb) - (ra + x->hb[j]);
}

NPH_HD f32 box_edge(const NphBoxAxes *x, u32 i, u32 j, NphVec3 *axis)
{
    const u32 i1 = (i + 1U) % 3U;
    const u32 i2 = (i + 2U) % 3U;
    const u32 j1 = (j + 1U) % 3U;
    const u32 j2 = (j + 2U) % 3U;
    const f32 len_sq = 1.0F - x->r[i][j] * x->r[i][j];
    if (len_sq < 1e-6F) {
        return -NPH_F32_MAX;
    }
    const f32 ra = x->ha[i1] * x->ar[i2][j] + x->ha[i2] * x->ar[i1][j];
    const f32 rb = x->hb[j1] * x->ar[i][j2] + x->hb[j2] * x->ar[i][j1];
    const f32 tl = x->t[i2] * x->r[i1][j] - x->t[i1] * x->r[i2][j];
    const f32 inv_len = 1.0F / nph_sqrtf(len_sq);
    f32 l[3];
    l[i] = 0.0F;
    l[i1] = -x->r[i2][j];
    l[i2] = x->r[i1][j];
    const f32 s = tl >= 0.0F ? inv_len : -inv_len;
    *axis = v3(l[0] * s, l[1] * s, l[2] * s);
    return (nph_absf(tl) - (ra + rb)) * inv_len;
}

NPH_HD void collide_boxes(NphVec3 half_a, NphPhysicsPose pa, NphVec3 half_b, NphPhysicsPose pb,
                          f32 speculative, NphCache *cache, NphManifold *out)
{
    const NphPhysicsPose rel = pose_inv_mul(pa, pb);
    const NphMat3 rm = m3_from_quat(rel.rotation);
    NphBoxAxes x;
    for (u32 i = 0U; i < 3U; ++i) {
        for (u32 j = 0U; j < 3U; ++j) {
            x.r[i][j] = v3_at(rm.c[j], i);
            x.ar[i][j] = nph_absf(x.r[i][j]) + 1e-6F;
        }
    }
    x.t[0] = rel.position.x;
    x.t[1] = rel.position.y;
    x.t[2] = rel.position.z;
    x.ha[0] = half_a.x;
    x.ha[1] = half_a.y;
    x.ha[2] = half_a.z;
    x.hb[0] = half_b.x;
    x.hb[1] = half_b.y;
    x.hb[2] = half_b.z;
    if (cache != 0 && cache->type != NPH_CACHE_NONE) {
        f32 sign;
        NphVec3 axis;
        f32 s = -NPH_F32_MAX;
        if (cache->type == NPH_CACHE_FACE_A && cache->index_a < 6U) {
            s = box_face_a(&x, cache->index_a / 2U, &sign);
        } else if (cache->type == NPH_CACHE_FACE_B && cache->index_b < 6U) {
            s = box_face_b(&x, cache->index_b / 2U, &sign);
        } else if (cache->type == NPH_CACHE_EDGES && cache->index_a < 3U && cache->index_b < 3U) {
            s = box_edge(&x, cache->index_a, cache->index_b, &axis);
        }
        if (s > speculative) {
            return;
        }
    }
    f32 face_a_sep = -NPH_F32_MAX;
    u32 face_a = 0U;
    for (u32 i = 0U; i < 3U; ++i) {
        f32 sign;
        const f32 s = box_face_a(&x, i, &sign);
        if (s > face_a_sep) {
            face_a_sep = s;
            face_a = 2U * i + (sign > 0.0F ? 1U : 0U);
        }
    }
    f32 face_b_sep = -NPH_F32_MAX;
    u32 face_b = 0U;
    for (u32 j = 0U; j < 3U; ++j) {
        f32 sign;
        const f32 s = box_face_b(&x, j, &sign);
        if (s > face_b_sep) {
            face_b_sep = s;
            face_b = 2U * j + (sign > 0.0F ? 0U : 1U);
        }
    }
    f32 edge_sep = -NPH_F32_MAX;
    u32 edge_i = 0U;
    u32 edge_j = 0U;
    NphVec3 edge_axis = v3(0.0F, 1.0F, 0.0F);
    for (u32 i = 0U; i < 3U; ++i) {
        for (u32 j = 0U; j < 3U; ++j) {
            NphVec3 axis;
            const f32 s = box_edge(&x, i, j, &axis);
            if (s > edge_sep) {
                edge_sep = s;
                edge_i = i;
                edge_j = j;
                edge_axis = axis;
            }
        }
    }
    const f32 face_sep = nph_maxf(face_a_sep, face_b_sep);
    const int use_edge = edge_sep > face_sep + NPH_EDGE_BIAS;
    const int flip = !use_edge && face_b_sep > face_a_sep + NPH_FACE_BIAS;
    const f32 best = use_edge ? edge_sep : (flip ? face_b_sep : face_a_sep);
    if (cache != 0) {
        if (nph_maxf(face_sep, edge_sep) > speculative) {
            const int by_edge = edge_sep > face_sep;
            const int by_b = !by_edge && face_b_sep > face_a_sep;
            cache->type = (u8)(by_edge ? NPH_CACHE_EDGES : (by_b ? NPH_CACHE_FACE_B : NPH_CACHE_FACE_A));
            cache->index_a = (u16)(by_edge ? edge_i : face_a);
            cache->index_b = (u16)(by_edge ? edge_j : face_b);
        } else {
            cache->type = NPH_CACHE_NONE;
        }
    }
    if (nph_maxf(face_sep, edge_sep) > speculative) {
        return;
    }
    NphBoxStorage sa;
    NphBoxStorage sb;
    const NphView va = box_view(half_a, &sa);
    const NphView vb = box_view(half_b, &sb);
    NphHullPair hp;
    hp.a = &va;
    hp.b = &vb;
    hp.rel = rel;
    for (u32 i = 0U; i < 8U; ++i) {
        hp.bv[i] = pose_point(rel, vb.vertices[i]);
    }
    if (best > NPH_PHYSICS_LINEAR_SLOP) {
        NphDistance gd;
        gjk_distance(va.vertices, 8U, hp.bv, 8U, &gd);
        if (gd.distance > speculative) {
            return;
        }
        const NphVec3 sat_axis = use_edge ? edge_axis
                                          : (flip ? v3_neg(plane_b_in_a(&hp, face_b).normal)
                                                  : va.planes[face_a].normal);
        if (gd.distance > 0.0F &&
            (use_edge || v3_dot(gd.normal, sat_axis) < NPH_FEATURE_AGREEMENT)) {
            const NphVec3 mid = v3_scale(v3_add(gd.point_a, gd.point_b), 0.5F);
            out->normal = pose_vector(pa, gd.normal);
            manifold_push(out, pose_point(pa, mid), gd.distance, 0x40000000U);
            return;
        }
    }
    if (use_edge) {
        NphVec3 ca = v3_zero();
        for (u32 k = 0U; k < 3U; ++k) {
            if (k != edge_i) {
                const f32 s = v3_at(edge_axis, k) >= 0.0F ? x.ha[k] : -x.ha[k];
                if (k == 0U) {
                    ca.x = s;
                } else if (k == 1U) {
                    ca.y = s;
                } else {
                    ca.z = s;
                }
            }
        }
        NphVec3 ea = v3_zero();
        if (edge_i == 0U) {
            ea.x = x.ha[0];
        } else if (edge_i == 1U) {
            ea.y = x.ha[1];
        } else {
            ea.z = x.ha[2];
        }
        NphVec3 cb = rel.position;
        for (u32 m = 0U; m < 3U; ++m) {
            if (m != edge_j) {
                const f32 d = v3_dot(rm.c[m], edge_axis);
                cb = v3_mul_add(cb, rm.c[m], d >= 0.0F ? -x.hb[m] : x.hb[m]);
            }
        }
        const NphVec3 eb = v3_scale(rm.c[edge_j], x.hb[edge_j]);
        const NphSegDist sd =
            segment_distance(v3_sub(ca, ea), v3_add(ca, ea), v3_sub(cb, eb), v3_add(cb, eb));
        const f32 sep = v3_dot(v3_sub(sd.closest2, sd.closest1), edge_axis);
        if (sep > speculative) {
            return;
        }
        const NphVec3 mid = v3_scale(v3_add(sd.closest1, sd.closest2), 0.5F);
        out->normal = pose_vector(pa, edge_axis);
        manifold_push(out, pose_point(pa, mid), sep, nph_mix_id(0x20000000U | edge_i, edge_j));
        return;
    }
    NphMPoint pts[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
    u32 count = 0U;
    const NphVec3 n = face_contact(&hp, flip, flip ? face_b : face_a, speculative,
                                   NPH_PHYSICS_MANIFOLD_MAX_POINTS, pts, &count);
    out->normal = pose_vector(pa, n);
    for (u32 i = 0U; i < count; ++i) {
        manifold_push(out, pose_point(pa, pts[i].point), pts[i].separation, pts[i].id);
    }
}

// ---------------------------------------------------------------------------
// Dispatch (collide.c:474-540): the pair in type order (a >= b), the normal
// reversed when the call's order was the other.
// ---------------------------------------------------------------------------
NPH_HD void collide_ordered(const NphGeom *a, NphPhysicsPose pa, const NphGeom *b,
                            NphPhysicsPose pb, f32 speculative, NphCache *cache, NphManifold *out)
{
    switch (a->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        collide_balls(pa.position, a->a, pb.position, b->a, speculative, 0U, v3(0.0F, 1.0F, 0.0F),
                      out);
        return;
    case NPH_PHYSICS_SHAPE_CAPSULE:
        if (b->type == NPH_PHYSICS_SHAPE_SPHERE) {
            collide_capsule_sphere(a->a, a->b, pa, b->a, pb, speculative, out);
        } else {
            collide_capsules(a->a, a->b, pa, b->a, b->b, pb, speculative, out);
        }
        return;
    default: {
        if (a->type == NPH_PHYSICS_SHAPE_BOX && b->type == NPH_PHYSICS_SHAPE_SPHERE) {
            collide_box_sphere(v3(a->a, a->b, a->c), pa, b->a, pb, speculative, out);
            return;
        }
        if (a->type == NPH_PHYSICS_SHAPE_BOX && b->type == NPH_PHYSICS_SHAPE_BOX) {
            collide_boxes(v3(a->a, a->b, a->c), pa, v3(b->a, b->b, b->c), pb, speculative, cache,
                          out);
            return;
        }
        NphBoxStorage sa;
        const NphView va = a->type == NPH_PHYSICS_SHAPE_BOX ? box_view(v3(a->a, a->b, a->c), &sa)
                                                            : hull_view(a->hull);
        if (b->type == NPH_PHYSICS_SHAPE_SPHERE) {
            collide_hull_sphere(&va, pa, b->a, pb, speculative, out);
        } else if (b->type == NPH_PHYSICS_SHAPE_CAPSULE) {
            collide_hull_capsule(&va, pa, b->a, b->b, pb, speculative, out);
        } else {
            NphBoxStorage sb;
            const NphView vb = b->type == NPH_PHYSICS_SHAPE_BOX
                                   ? box_view(v3(b->a, b->b, b->c), &sb)
                                   : hull_view(b->hull);
            collide_hulls(&va, pa, &vb, pb, speculative, cache, out);
        }
        return;
    }
    }
}

NPH_HD void nph_collide(const NphGeom *a, NphPhysicsPose pose_a, const NphGeom *b,
                        NphPhysicsPose pose_b, f32 speculative_distance, NphCache *cache,
                        NphManifold *out)
{
    manifold_clear(out);
    const f32 speculative = nph_maxf(speculative_distance, 0.0F);
    if (a->type >= b->type) {
        collide_ordered(a, pose_a, b, pose_b, speculative, cache, out);
        return;
    }
    collide_ordered(b, pose_b, a, pose_a, speculative, cache, out);
    out->normal = v3_neg(out->normal);
}

#endif // NPH_COLLIDE_H
