// SPDX-License-Identifier: Apache-2.0
// Convex shapes against triangle meshes: one mesh pair on a team of
// NPH_SURFACE_TEAM lanes of a wavefront. The reference takes a pair on one
// thread (collide_mesh.c:104-1020); the results here are its bits, the
// work is spread as follows.
//
//  1. Gather. The mesh block's four-wide tree is walked depth first with
//     the slots in order, the order in which the reference visits
//     triangles (it fixes the clusters below): a popped node's four slots
//     are tested by four lanes at once and pushed by one ballot, the last
//     slot deepest; a leaf's (up to four) triangles are tested the same way
//     and appended to a queue in LDS in slot order.
//  2. Contact. A full queue (one triangle a lane) is collided. Spheres and
//     capsules take a triangle a lane (closed forms: closest points and
//     segment distances). Boxes and hulls take the queued triangles one at
//     a time on the whole team: the separating-axis test and face clipping
//     of nph_polytope.h, the triangle as a flat polytope of two faces.
//     Either way a triangle's contact then passes its rules: one-sided
//     meshes drop contacts from behind, and a contact at an edge or corner
//     keeps its own normal only across an active edge (else it is taken
//     again along the face normal, or dropped).
//  3. Merge, in queue order. A triangle's contact joins the first cluster
//     whose first normal is within 5 degrees of its own (the lanes test
//     the clusters at once; the lowest set bit of a ballot wins) or starts
//     one; with eight clusters, the shallowest gives way to a deeper
//     contact. Each point joins the cluster unless it lies within 0.1 mm
//     of a point there (the lanes test the cluster's points at once; the
//     first match keeps the deeper separation). A cluster that reaches 32
//     points is thinned to what the last reduction can want: its eight
//     deepest points, and its farthest points along eight directions in
//     the contact plane (touching points) and along four (speculative
//     ones); ranks and farthest points are team reductions, and the
//     survivors are packed in order by ballot prefix counts.
//  4. Finish. Each cluster is reduced to four points (a cluster a lane;
//     capsules keep their deepest point and the ends along their axis),
//     the clusters ranked by depth, the four deepest kept.
//
// Every comparison keeps the reference's tie rule (strict "less" keeps the
// first of equals, "at least" the last), so the chosen points, hence
// contact ids and warm starting, are the reference's.
#ifndef NPH_SURFACE_H
#define NPH_SURFACE_H

#include "nph_bvh4.h"
#include "nph_polytope.h"

// ---------------------------------------------------------------------------
// Mesh-frame arithmetic shared with the host's world queries and casts
// ---------------------------------------------------------------------------

// v rotated by the inverse of q, in 64 bits: v + w t + u x t with u the
// conjugate's vector part and t = 2 u x v (collide_mesh.c:95-102).
NPH_HD NphDVec3 drotate_inverse(NphQuat q, NphDVec3 v)
{
    const f64 u[3] = {-(f64)q.x, -(f64)q.y, -(f64)q.z};
    const f64 w = (f64)q.w;
    const f64 t0 = (u[1] * v.z - u[2] * v.y) * 2.0;
    const f64 t1 = (u[2] * v.x - u[0] * v.z) * 2.0;
    const f64 t2 = (u[0] * v.y - u[1] * v.x) * 2.0;
    const f64 x = (v.x + t0 * w) + (u[1] * t2 - u[2] * t1);
    const f64 y = (v.y + t1 * w) + (u[2] * t0 - u[0] * t2);
    const f64 z = (v.z + t2 * w) + (u[0] * t1 - u[1] * t0);
    return dv3(x, y, z);
}

// A mesh vertex relative to a 64-bit origin, rounded once to f32
// (physics_internal.h:143-147).
NPH_HD NphVec3 relative_vertex(NphVec3 p, NphDVec3 origin)
{
    return v3((f32)((f64)p.x - origin.x), (f32)((f64)p.y - origin.y), (f32)((f64)p.z - origin.z));
}

#if defined(__cplusplus)

// ---------------------------------------------------------------------------
// Constants (collide_mesh.c:20-92)
// ---------------------------------------------------------------------------
#define NPH_SURFACE_TEAM 16U
#define NPH_SURFACE_MERGE_COS 0.99619470F // cos 5 degrees: one cluster
#define NPH_SURFACE_WELD 1e-4F            // points this close are one point
#define NPH_SURFACE_CLUSTERS 8U
#define NPH_SURFACE_CLUSTER_POINTS (8U * NPH_PHYSICS_MANIFOLD_MAX_POINTS)
#define NPH_SURFACE_KEEP_DEEPEST 8U // a thinned cluster's deepest points
#define NPH_SURFACE_FACE_PREFERENCE (0.5F * NPH_PHYSICS_LINEAR_SLOP)
#define NPH_SURFACE_EDGE_TOLERANCE 1e-4F
#define NPH_SURFACE_RIDGE 0.0868F // tan 5 degrees: reach past an inactive edge per unit of height
#define NPH_SURFACE_AGAINST (-1e-4F)
#define NPH_SURFACE_FACE_COS 0.99999F
#define NPH_SURFACE_PARALLEL 1e-4F // sin^2 below which a capsule lies along an edge
#define NPH_TOUCH_SEPARATION (0.1F * NPH_PHYSICS_LINEAR_SLOP)

// Where on a triangle a closest point lies: its face, edge k (from corner
// k to corner k + 1) or corner k.
enum { NPH_AT_FACE = 0U, NPH_AT_EDGE = 1U, NPH_AT_CORNER = 4U };
// What a contact at an edge or corner may do (collide_mesh.c:240-262).
enum { NPH_VERDICT_DROP = 0, NPH_VERDICT_KEEP = 1, NPH_VERDICT_FACE = 2 };

// ---------------------------------------------------------------------------
// A triangle seen from the shape
// ---------------------------------------------------------------------------
struct NphFacet {
    NphVec3 v[3];  // corners, ordered so the normal faces the shape
    NphVec3 n;     // unit normal
    NphVec3 m[3];  // edge k's outward normal in the plane, (v[k+1] - v[k]) x n
    f32 mlen[3];   // |m[k]|
    u32 active;    // bit k: edge k may give a contact its own normal
};

// The facet of corners c (already next to the shape) and active-edge bits,
// for a shape whose centre is `center`; false for a degenerate triangle
// and, on a one-sided mesh, for one the shape is behind.
NPH_HD bool facet_make(NphFacet *f, const NphVec3 c[3], u32 active, int two_sided, NphVec3 center)
{
    // The normal in 64 bits (collide_mesh.c:104-125).
    f64 e[2][3];
    for (u32 k = 0U; k < 2U; ++k) {
        const NphVec3 p = c[k + 1U];
        e[k][0] = (f64)p.x - (f64)c[0].x;
        e[k][1] = (f64)p.y - (f64)c[0].y;
        e[k][2] = (f64)p.z - (f64)c[0].z;
    }
    const f64 nx = e[0][1] * e[1][2] - e[0][2] * e[1][1];
    const f64 ny = e[0][2] * e[1][0] - e[0][0] * e[1][2];
    const f64 nz = e[0][0] * e[1][1] - e[0][1] * e[1][0];
    const f64 len = __builtin_sqrt((nx * nx + ny * ny) + nz * nz);
    const NphVec3 n = len > 0.0 ? v3((f32)(nx / len), (f32)(ny / len), (f32)(nz / len)) : v3_zero();
    if (v3_len_sq(n) == 0.0F) {
        return false;
    }
    // Seen from behind, the corners run the other way round: edge 0 is
    // the old edge 2 reversed, edge 2 the old edge 0.
    const bool flip = v3_dot(n, v3_sub(center, c[0])) < 0.0F;
    if (flip && !two_sided) {
        return false;
    }
    f->v[0] = c[0];
    f->v[1] = flip ? c[2] : c[1];
    f->v[2] = flip ? c[1] : c[2];
    f->n = flip ? v3_neg(n) : n;
    f->active = flip ? (((active >> 2U) & 1U) | (active & 2U) | ((active & 1U) << 2U)) : active;
    for (u32 k = 0U; k < 3U; ++k) {
        f->m[k] = v3_cross(v3_sub(f->v[k == 2U ? 0U : k + 1U], f->v[k]), f->n);
        f->mlen[k] = v3_len(f->m[k]);
    }
    return true;
}

NPH_HD bool facet_edge_active(const NphFacet *f, u32 k) { return ((f->active >> k) & 1U) != 0U; }

// How far p lies out across edge k, times |m[k]|.
NPH_HD f32 facet_edge_height(const NphFacet *f, u32 k, NphVec3 p)
{
    return v3_dot(f->m[k], v3_sub(p, f->v[k]));
}

// The point of the triangle closest to p, and where it lies: Voronoi
// regions tested corner, edge, corner, edge, corner, edge, then the face
// (Ericson, Real-Time Collision Detection, 5.1.5).
NPH_HD NphVec3 facet_closest(const NphFacet *f, NphVec3 p, u32 *where)
{
    const NphVec3 a = f->v[0], b = f->v[1], c = f->v[2];
    const NphVec3 ab = v3_sub(b, a), ac = v3_sub(c, a);
    const NphVec3 ap = v3_sub(p, a);
    const f32 d1 = v3_dot(ab, ap), d2 = v3_dot(ac, ap);
    if (d1 <= 0.0F && d2 <= 0.0F) {
        *where = NPH_AT_CORNER + 0U;
        return a;
    }
    const NphVec3 bp = v3_sub(p, b);
    const f32 d3 = v3_dot(ab, bp), d4 = v3_dot(ac, bp);
    if (d3 >= 0.0F && d4 <= d3) {
        *where = NPH_AT_CORNER + 1U;
        return b;
    }
    const f32 vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0F && d1 >= 0.0F && d3 <= 0.0F) {
        *where = NPH_AT_EDGE + 0U;
        return v3_mul_add(a, ab, d1 / (d1 - d3));
    }
    const NphVec3 cp = v3_sub(p, c);
    const f32 d5 = v3_dot(ab, cp), d6 = v3_dot(ac, cp);
    if (d6 >= 0.0F && d5 <= d6) {
        *where = NPH_AT_CORNER + 2U;
        return c;
    }
    const f32 vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0F && d2 >= 0.0F && d6 <= 0.0F) {
        *where = NPH_AT_EDGE + 2U;
        return v3_mul_add(a, ac, d2 / (d2 - d6));
    }
    const f32 va = d3 * d6 - d5 * d4;
    const f32 u = d4 - d3, w = d5 - d6;
    if (va <= 0.0F && u >= 0.0F && w >= 0.0F) {
        *where = NPH_AT_EDGE + 1U;
        return v3_mul_add(b, v3_sub(c, b), u / (u + w));
    }
    *where = NPH_AT_FACE;
    return v3_mul_add(p, f->n, -v3_dot(f->n, ap));
}

// Where a point already on (or near) the triangle lies, from which edge
// lines it is within the tolerance of: none, the face; one, that edge; two,
// their shared corner; all three (a sliver), the nearest corner.
NPH_HD u32 facet_locate(const NphFacet *f, NphVec3 p)
{
    u32 near = 0U;
    for (u32 k = 0U; k < 3U; ++k) {
        if (facet_edge_height(f, k, p) >= -NPH_SURFACE_EDGE_TOLERANCE * f->mlen[k]) {
            near |= 1U << k;
        }
    }
    // Edge pairs: {0,1} meet at corner 1, {1,2} at corner 2, {0,2} at 0.
    const u8 by_mask[7] = {NPH_AT_FACE,       NPH_AT_EDGE + 0U,   NPH_AT_EDGE + 1U,
                           NPH_AT_CORNER + 1U, NPH_AT_EDGE + 2U,  NPH_AT_CORNER + 0U,
                           NPH_AT_CORNER + 2U};
    if (near < 7U) {
        return by_mask[near];
    }
    u32 best = 0U;
    f32 best_d = v3_dist_sq(p, f->v[0]);
    for (u32 k = 1U; k < 3U; ++k) {
        const f32 d = v3_dist_sq(p, f->v[k]);
        if (d < best_d) {
            best_d = d;
            best = k;
        }
    }
    return NPH_AT_CORNER + best;
}

// What a contact at `where` with normal n (from the triangle towards the
// shape) may do: pointing into the face, it is dropped; across an active
// edge at that place (an edge strictly, a corner if either of its edges
// faces n), it keeps n; otherwise the face normal must decide.
NPH_HD int facet_verdict(const NphFacet *f, u32 where, NphVec3 n)
{
    if (v3_dot(n, f->n) < NPH_SURFACE_AGAINST) {
        return NPH_VERDICT_DROP;
    }
    bool own = false;
    if (where >= NPH_AT_CORNER) {
        const u32 k = where - NPH_AT_CORNER;
        const u32 before = k == 0U ? 2U : k - 1U; // the edge ending at corner k
        own = (facet_edge_active(f, k) && v3_dot(n, f->m[k]) > 0.0F) ||
              (facet_edge_active(f, before) && v3_dot(n, f->m[before]) > 0.0F);
    } else if (where >= NPH_AT_EDGE) {
        const u32 k = where - NPH_AT_EDGE;
        own = facet_edge_active(f, k) && v3_dot(n, f->m[k]) >= 0.0F;
    }
    return own ? NPH_VERDICT_KEEP : NPH_VERDICT_FACE;
}

// Whether p lies over the triangle grown by `grow` across every edge and
// by `ridge` more across inactive ones.
NPH_HD bool facet_covers(const NphFacet *f, NphVec3 p, f32 grow, f32 ridge)
{
    for (u32 k = 0U; k < 3U; ++k) {
        const f32 reach = facet_edge_active(f, k) ? grow : grow + ridge;
        if (facet_edge_height(f, k, p) > reach * f->mlen[k]) {
            return false;
        }
    }
    return true;
}

// The part [lo, hi] of the segment p0 + x d (x in [0, 1]) over the grown
// triangle (as facet_covers); false when no part is.
NPH_HD bool facet_clip_segment(const NphFacet *f, NphVec3 p0, NphVec3 d, f32 grow, f32 ridge,
                               f32 *lo, f32 *hi)
{
    f32 x0 = 0.0F, x1 = 1.0F;
    for (u32 k = 0U; k < 3U; ++k) {
        const f32 reach = facet_edge_active(f, k) ? grow : grow + ridge;
        const f32 h0 = facet_edge_height(f, k, p0) - reach * f->mlen[k];
        const f32 slope = v3_dot(f->m[k], d);
        if (slope == 0.0F) {
            if (h0 > 0.0F) {
                return false;
            }
            continue;
        }
        const f32 x = -h0 / slope;
        if (slope > 0.0F) {
            x1 = nph_minf(x1, x);
        } else {
            x0 = nph_maxf(x0, x);
        }
        if (x0 > x1) {
            return false;
        }
    }
    *lo = x0;
    *hi = x1;
    return true;
}

// ---------------------------------------------------------------------------
// A round shape's contact with one facet (at most two points)
// ---------------------------------------------------------------------------
struct NphFacetHit {
    NphVec3 normal; // from the shape towards the mesh
    u32 count;
    NphMPoint p[2];
};

NPH_HD void hit_add(NphFacetHit *h, NphVec3 core, NphVec3 n, f32 radius, f32 separation, u32 id)
{
    // The point half way into the overlap, from the core along -n.
    NphMPoint *q = &h->p[h->count++];
    q->point = v3_mul_add(core, n, -(radius + 0.5F * separation));
    q->separation = separation;
    q->id = id;
}

NPH_HD f32 hit_depth(const NphFacetHit *h)
{
    f32 s = NPH_F32_MAX;
    for (u32 i = 0U; i < h->count; ++i) {
        s = nph_minf(s, h->p[i].separation);
    }
    return s;
}

// A ball of `radius` at the origin (collide_mesh.c:316-356).
NPH_HD void ball_facet(const NphFacet *f, f32 radius, f32 speculative, NphFacetHit *out)
{
    const NphVec3 c = v3_zero();
    u32 where;
    const NphVec3 q = facet_closest(f, c, &where);
    const NphVec3 d = v3_sub(c, q);
    const f32 d2 = v3_len_sq(d);
    const f32 reach = radius + speculative;
    if (d2 > reach * reach) {
        return;
    }
    const f32 dist = nph_sqrtf(d2);
    NphVec3 n = dist > 1e-9F ? v3_scale(d, 1.0F / dist) : f->n;
    f32 sep = dist - radius;
    const int verdict = facet_verdict(f, where, n);
    if (verdict == NPH_VERDICT_DROP) {
        return;
    }
    if (verdict == NPH_VERDICT_FACE) {
        // Along the face normal, if the foot of the centre is over the
        // triangle (past an inactive edge by the ridge allowance).
        const f32 h = v3_dot(f->n, v3_sub(c, f->v[0]));
        if (!facet_covers(f, v3_mul_add(c, f->n, -h), NPH_SURFACE_EDGE_TOLERANCE,
                          NPH_SURFACE_RIDGE * nph_maxf(h, 0.0F))) {
            return;
        }
        n = f->n;
        sep = h - radius;
    }
    out->normal = v3_neg(n);
    hit_add(out, c, n, radius, sep, 0U);
}

// A capsule's core p0 + x d over the face: the ends of [lo, hi] (ids 0, 1),
// or only the lower one when the part is a point.
NPH_HD void capsule_over_face(const NphFacet *f, NphVec3 p0, NphVec3 d, f32 lo, f32 hi, f32 radius,
                              f32 speculative, NphFacetHit *out)
{
    const f32 h0 = v3_dot(f->n, v3_sub(p0, f->v[0]));
    const f32 slope = v3_dot(f->n, d);
    const f32 x[2] = {lo, hi};
    const f32 h[2] = {h0 + lo * slope, h0 + hi * slope};
    const bool both = (hi - lo) * v3_len(d) > NPH_SURFACE_WELD;
    u32 lower = h[0] < h[1] ? 0U : 1U;
    if (h[0] == h[1]) {
        lower = hi >= 1.0F ? 1U : 0U;
    }
    out->normal = v3_neg(f->n);
    for (u32 k = 0U; k < 2U; ++k) {
        const f32 sep = h[k] - radius;
        if ((both || k == lower) && sep <= speculative) {
            hit_add(out, v3_mul_add(p0, d, x[k]), f->n, radius, sep, k);
        }
    }
}

// A capsule's core along an edge it runs parallel to: the ends of the
// overlap of core and edge (ids 2, 3); false when the overlap is too short.
NPH_HD bool capsule_along_edge(const NphFacet *f, u32 k, NphVec3 p0, NphVec3 p1, NphVec3 n,
                               f32 radius, f32 speculative, NphFacetHit *out)
{
    const NphVec3 d = v3_sub(p1, p0);
    const NphVec3 a = f->v[k];
    const NphVec3 u = v3_sub(f->v[k == 2U ? 0U : k + 1U], a);
    const f32 dd = v3_len_sq(d);
    const f32 uu = v3_len_sq(u);
    if (!(dd > 1e-12F && v3_len_sq(v3_cross(d, u)) <= NPH_SURFACE_PARALLEL * dd * uu)) {
        return false;
    }
    const f32 len = nph_sqrtf(uu);
    const NphVec3 dir = v3_scale(u, 1.0F / len);
    const f32 s0 = v3_dot(v3_sub(p0, a), dir);
    const f32 s1 = v3_dot(v3_sub(p1, a), dir);
    const f32 ends[2] = {nph_maxf(0.0F, nph_minf(s0, s1)), nph_minf(len, nph_maxf(s0, s1))};
    if (!(ends[1] - ends[0] > NPH_PHYSICS_LINEAR_SLOP)) {
        return false;
    }
    for (u32 e = 0U; e < 2U; ++e) {
        const NphVec3 q = v3_mul_add(a, dir, ends[e]);
        const NphVec3 x = closest_on_segment(p0, p1, q);
        const f32 sep = v3_dist(x, q) - radius;
        if (sep <= speculative) {
            hit_add(out, x, n, radius, sep, 2U + e);
        }
    }
    return out->count > 0U;
}

// A capsule of core p0-p1 (collide_mesh.c:358-536): over the face when its
// core is no farther from the face than from the nearest edge (by the face
// preference), else at the nearest edge or corner by the rules.
NPH_HD void capsule_facet(const NphFacet *f, NphVec3 p0, NphVec3 p1, f32 radius, f32 speculative,
                          NphFacetHit *out)
{
    const NphVec3 d = v3_sub(p1, p0);
    u32 k = 0U;
    NphSegDist near = segment_distance(p0, p1, f->v[0], f->v[1]);
    for (u32 e = 1U; e < 3U; ++e) {
        const NphSegDist s = segment_distance(p0, p1, f->v[e], f->v[e == 2U ? 0U : e + 1U]);
        if (s.distance_squared < near.distance_squared) {
            near = s;
            k = e;
        }
    }
    const f32 edge_dist = nph_sqrtf(near.distance_squared);
    f32 lo = 0.0F, hi = 0.0F;
    if (facet_clip_segment(f, p0, d, 0.0F, 0.0F, &lo, &hi)) {
        const f32 h0 = v3_dot(f->n, v3_sub(p0, f->v[0]));
        const f32 slope = v3_dot(f->n, d);
        const f32 low = nph_minf(h0 + lo * slope, h0 + hi * slope);
        if (low <= edge_dist + NPH_SURFACE_FACE_PREFERENCE) {
            if (low - radius <= speculative) {
                capsule_over_face(f, p0, d, lo, hi, radius, speculative, out);
            }
            return;
        }
    }
    if (edge_dist > radius + speculative) {
        return;
    }
    u32 where = NPH_AT_EDGE + k;
    if (near.fraction2 <= 0.0F) {
        where = NPH_AT_CORNER + k;
    } else if (near.fraction2 >= 1.0F) {
        where = NPH_AT_CORNER + (k == 2U ? 0U : k + 1U);
    }
    const NphVec3 n = edge_dist > 1e-9F
                          ? v3_scale(v3_sub(near.closest1, near.closest2), 1.0F / edge_dist)
                          : f->n;
    const int verdict = facet_verdict(f, where, n);
    if (verdict == NPH_VERDICT_DROP) {
        return;
    }
    if (verdict == NPH_VERDICT_FACE) {
        const f32 h0 = v3_dot(f->n, v3_sub(p0, f->v[0]));
        const f32 h1 = h0 + v3_dot(f->n, d);
        if (facet_clip_segment(f, p0, d, NPH_SURFACE_EDGE_TOLERANCE,
                               NPH_SURFACE_RIDGE * nph_maxf(nph_maxf(h0, h1), 0.0F), &lo, &hi)) {
            capsule_over_face(f, p0, d, lo, hi, radius, speculative, out);
            // Not along the face if that would be much deeper than the edge.
            if (hit_depth(out) < edge_dist - radius - NPH_PHYSICS_LINEAR_SLOP) {
                out->count = 0U;
            }
        }
        return;
    }
    out->normal = v3_neg(n);
    if (capsule_along_edge(f, k, p0, p1, n, radius, speculative, out)) {
        return;
    }
    hit_add(out, near.closest1, n, radius, edge_dist - radius, 4U);
}

// ---------------------------------------------------------------------------
// The team's working memory
// ---------------------------------------------------------------------------
struct NphClusterHead {
    NphVec3 first;  // the first member's normal: membership
    NphVec3 normal; // the weighted sum's direction, from the shape to the mesh
    NphVec3 sum;    // members' normals weighted by how near they come
    f32 depth;      // the least separation
    u32 count;
};

struct NphMeshManifolds {
    u32 count;
    NphManifold manifolds[NPH_PHYSICS_MESH_MAX_MANIFOLDS];
};

// A facet as a flat polytope: two faces back to back, three edges in twin
// pairs. The numbering is the reference's (collide_hull.c:94-121), which
// feature ids, hence contact ids, come from.
struct NphFacetView {
    NphVec3 v[3];
    NphPlane planes[2];
};

#if defined(__HIPCC__)
__device__ __constant__ static const u16 nph_facet_faces_d[2] = {0U, 1U};
__device__ __constant__ static const NphPhysicsHullEdge nph_facet_edges_d[6] = {
    {2U, 1U, 0U, 0U}, {5U, 0U, 1U, 1U}, {4U, 3U, 1U, 0U},
    {1U, 2U, 2U, 1U}, {0U, 5U, 2U, 0U}, {3U, 4U, 0U, 1U},
};
#endif
static const u16 nph_facet_faces_h[2] = {0U, 1U};
static const NphPhysicsHullEdge nph_facet_edges_h[6] = {
    {2U, 1U, 0U, 0U}, {5U, 0U, 1U, 1U}, {4U, 3U, 1U, 0U},
    {1U, 2U, 2U, 1U}, {0U, 5U, 2U, 0U}, {3U, 4U, 0U, 1U},
};

NPH_TD NphView facet_view(const NphFacetView *s)
{
    NphView h;
    h.vertices = s->v;
    h.planes = s->planes;
#if defined(__HIP_DEVICE_COMPILE__)
    h.faces = nph_facet_faces_d;
    h.edges = nph_facet_edges_d;
#else
    h.faces = nph_facet_faces_h;
    h.edges = nph_facet_edges_h;
#endif
    h.vertex_count = 3U;
    h.face_count = 2U;
    h.edge_count = 6U;
    h.centroid = v3_scale(v3_add(v3_add(s->v[0], s->v[1]), s->v[2]), 1.0F / 3.0F);
    return h;
}

template <u32 T, u32 CAP> struct NphSurfaceLds {
    NphClusterHead head[NPH_SURFACE_CLUSTERS];
    NphMPoint pts[NPH_SURFACE_CLUSTERS][NPH_SURFACE_CLUSTER_POINTS];
    u32 stack[NPH_BVH4_STACK];
    u32 queue[T + 4U];
    NphFacetHit hits[T];
    NphFacetView facet;
    NphMeshManifolds out;
    NphPolyScratch<4U, CAP> poly; // the facet's corners, clip polygons
};

// The convex shape in the mesh's frame: at the origin, turned by `rotation`.
struct NphSurfaceShape {
    u32 kind; // 0 a ball, 1 a capsule with a core, 2 a polytope
    f32 radius;
    NphVec3 core0, core1;
    NphVec3 center;
    NphPhysicsPose pose;
    NphView view;
};

// ---------------------------------------------------------------------------
// Clusters
// ---------------------------------------------------------------------------
// The eight directions a full cluster keeps its farthest points along,
// 45 degrees apart from the first tangent, so each corner of a rectangle is
// farthest along at least two (collide_mesh.c:697-700).
NPH_TD void thin_direction(u32 k, f32 *x, f32 *y)
{
    const f32 h = 0.70710678F;
    const f32 xs[8] = {1.0F, h, 0.0F, -h, -1.0F, -h, 0.0F, h};
    *x = xs[k];
    *y = xs[(k + 6U) & 7U];
}

// Thins cluster c (full) in place, keeping its order.
template <u32 T, u32 CAP>
NPH_TD void cluster_thin(const NphTeam<T> &tm, NphSurfaceLds<T, CAP> *S, u32 c)
{
    constexpr u32 R = NPH_SURFACE_CLUSTER_POINTS / T; // points a lane
    NphMPoint *pts = S->pts[c];
    const u32 count = S->head[c].count;
    const NphVec3 normal = S->head[c].normal;
    bool keep[R];
    // The deepest: a point is kept when fewer than eight come before it in
    // (separation, place) order.
    for (u32 r = 0U; r < R; ++r) {
        const u32 i = r * T + tm.lane;
        keep[r] = false;
        if (i < count) {
            const f32 s = pts[i].separation;
            u32 before = 0U;
            for (u32 j = 0U; j < count; ++j) {
                const f32 t = pts[j].separation;
                before += (t < s || (t == s && j < i)) ? 1U : 0U;
            }
            keep[r] = before < NPH_SURFACE_KEEP_DEEPEST;
        }
    }
    // The farthest along each direction: touching points along all eight,
    // speculative ones along every other.
    const NphVec3 t1 = v3_perpendicular(normal);
    const NphVec3 t2 = v3_cross(normal, t1);
    for (u32 touching = 1U;; touching = 0U) {
        for (u32 k = 0U; k < 8U; k += touching ? 1U : 2U) {
            f32 x, y;
            thin_direction(k, &x, &y);
            const NphVec3 dir = v3_add(v3_scale(t1, x), v3_scale(t2, y));
            const NphPick far = team_first_max(tm, count, [&](u32 i) {
                const bool touches = pts[i].separation <= NPH_TOUCH_SEPARATION;
                return touches == (touching != 0U) ? v3_dot(pts[i].point, dir) : -NPH_F32_MAX;
            });
            if (far.index != NPH_NULL32 && far.index % T == tm.lane) {
                keep[far.index / T] = true;
            }
        }
        if (touching == 0U) {
            break;
        }
    }
    // Packed in order.
    NphMPoint mine[R];
    u32 at[R];
    u32 kept = 0U;
    for (u32 r = 0U; r < R; ++r) {
        const u64 m = tm.ballot(keep[r]);
        at[r] = kept + tm.below(m);
        kept += (u32)__builtin_popcountll(m);
        if (keep[r]) {
            mine[r] = pts[r * T + tm.lane];
        }
    }
    tm.sync();
    for (u32 r = 0U; r < R; ++r) {
        if (keep[r]) {
            pts[at[r]] = mine[r];
        }
    }
    if (tm.lane == 0U) {
        S->head[c].count = kept;
    }
    tm.sync();
}

// One facet's contact (normal, `count` points at `pts`, the same on every
// lane) merged into the clusters; `nc` is the cluster count (uniform).
template <u32 T, u32 CAP>
NPH_TD void cluster_merge(const NphTeam<T> &tm, NphSurfaceLds<T, CAP> *S, u32 *nc, NphVec3 normal,
                          u32 count, const NphMPoint *pts, u32 facet, f32 speculative, bool single)
{
    f32 depth = NPH_F32_MAX;
    for (u32 i = 0U; i < count; ++i) {
        depth = nph_minf(depth, pts[i].separation);
    }
    const u64 joins = tm.ballot(tm.lane < *nc && v3_dot(S->head[tm.lane].first, normal) >=
                                                     NPH_SURFACE_MERGE_COS);
    u32 c;
    if (joins != 0ULL) {
        c = nph_lowest_bit(joins);
    } else {
        if (*nc < NPH_SURFACE_CLUSTERS) {
            c = *nc;
            *nc += 1U;
        } else {
            // The shallowest gives way (the last of equals), and only to a
            // deeper contact.
            const bool mine = tm.lane < NPH_SURFACE_CLUSTERS;
            const f32 dk = mine ? S->head[tm.lane].depth : -NPH_F32_MAX;
            f32 top = dk;
            for (u32 k = 0U; k < NPH_SURFACE_CLUSTERS; ++k) {
                const f32 o = S->head[k].depth;
                top = o > top ? o : top;
            }
            const u64 tied = tm.ballot(mine && dk == top);
            c = 63U - (u32)__builtin_clzll(tied);
            if (!(depth < S->head[c].depth)) {
                return;
            }
        }
        tm.sync();
        if (tm.lane == 0U) {
            S->head[c].first = normal;
            S->head[c].sum = v3_zero();
            S->head[c].depth = depth;
            S->head[c].count = 0U;
        }
        tm.sync();
    }
    const NphClusterHead h = S->head[c];
    const NphVec3 sum = v3_mul_add(h.sum, normal, nph_maxf(speculative - depth, 0.0F) + 1e-6F);
    tm.sync();
    if (tm.lane == 0U) {
        S->head[c].sum = sum;
        S->head[c].normal = v3_normalize_or(sum, normal);
        S->head[c].depth = nph_minf(h.depth, depth);
    }
    tm.sync();
    NphMPoint *cp = S->pts[c];
    if (single) {
        // A ball touches at one point: the deepest (the first of equals).
        if (h.count == 0U || pts[0].separation < cp[0].separation) {
            tm.sync();
            if (tm.lane == 0U) {
                cp[0] = pts[0];
                cp[0].id = nph_mix_id(facet, pts[0].id);
                S->head[c].count = 1U;
            }
        }
        tm.sync();
        return;
    }
    for (u32 i = 0U; i < count; ++i) {
        NphMPoint p = pts[i];
        p.id = nph_mix_id(facet, p.id);
        u32 held = S->head[c].count;
        u64 welded = 0ULL; // bit j: held point j is within the weld distance
        for (u32 base = 0U; base < held; base += T) {
            const u32 j = base + tm.lane;
            const bool w = j < held && v3_dist_sq(cp[j].point, p.point) <=
                                           NPH_SURFACE_WELD * NPH_SURFACE_WELD;
            welded |= tm.ballot(w) << base;
        }
        tm.sync();
        if (welded != 0ULL) {
            const u32 j = nph_lowest_bit(welded);
            if (tm.lane == 0U) {
                cp[j].separation = nph_minf(cp[j].separation, p.separation);
            }
            tm.sync();
            continue;
        }
        if (held == NPH_SURFACE_CLUSTER_POINTS) {
            cluster_thin(tm, S, c);
            held = S->head[c].count;
        }
        if (tm.lane == 0U) {
            cp[held] = p;
            S->head[c].count = held + 1U;
        }
        tm.sync();
    }
}

// A capsule's cluster: its deepest point (the first of equals), then the
// first-found ends of the points along the core's axis, of the touching
// points and then of all, to four points at most; fewer than three stay.
NPH_TD u32 capsule_keep(NphMPoint *pts, u32 count, NphVec3 axis)
{
    if (count <= 2U) {
        return count;
    }
    u32 order[5];
    u32 n = 0U;
    u64 taken = 0ULL;
    u32 deepest = 0U;
    for (u32 i = 1U; i < count; ++i) {
        deepest = pts[i].separation < pts[deepest].separation ? i : deepest;
    }
    order[n++] = deepest;
    taken |= 1ULL << deepest;
    for (u32 all = 0U; all < 2U; ++all) {
        u32 end[2] = {NPH_NULL32, NPH_NULL32}; // least, greatest
        f32 ext[2] = {NPH_F32_MAX, -NPH_F32_MAX};
        for (u32 i = 0U; i < count; ++i) {
            if (all == 0U && pts[i].separation > NPH_TOUCH_SEPARATION) {
                continue;
            }
            const f32 d = v3_dot(pts[i].point, axis);
            if (d < ext[0]) {
                ext[0] = d;
                end[0] = i;
            }
            if (d > ext[1]) {
                ext[1] = d;
                end[1] = i;
            }
        }
        for (u32 e = 0U; e < 2U; ++e) {
            if (end[e] != NPH_NULL32 && ((taken >> end[e]) & 1ULL) == 0ULL) {
                taken |= 1ULL << end[e];
                order[n++] = end[e];
            }
        }
    }
    n = n < NPH_PHYSICS_MANIFOLD_MAX_POINTS ? n : NPH_PHYSICS_MANIFOLD_MAX_POINTS;
    NphMPoint kept[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
    for (u32 k = 0U; k < n; ++k) {
        kept[k] = pts[order[k]];
    }
    for (u32 k = 0U; k < n; ++k) {
        pts[k] = kept[k];
    }
    return n;
}

// ---------------------------------------------------------------------------
// The pair
// ---------------------------------------------------------------------------

// A polytope shape against one facet on the whole team: the separating
// axes and face clipping of nph_polytope.h (the facet a flat polytope whose
// face wins ties), then the facet's rules at the deepest point: a contact
// from behind is dropped; one at an edge or corner that may not keep its
// normal is clipped against the face itself instead, unless that reaches
// much deeper than the pair's own contact (collide_mesh.c:590-629).
template <u32 T, u32 CAP>
NPH_TD void facet_polytope(const NphTeam<T> &tm, NphSurfaceLds<T, CAP> *S,
                           const NphSurfaceShape &sh, const NphFacet &f, f32 speculative,
                           NphTriContact *hit)
{
    tm.sync();
    if (tm.lane == 0U) {
        for (u32 k = 0U; k < 3U; ++k) {
            S->facet.v[k] = f.v[k];
        }
        S->facet.planes[0] = plane_make(f.n, -v3_dot(f.n, f.v[0]));
        S->facet.planes[1] = plane_make(v3_neg(f.n), v3_dot(f.n, f.v[0]));
    }
    tm.sync();
    const NphView fv = facet_view(&S->facet);
    tri_clear(hit);
    team_polytopes(tm, &S->poly, &sh.view, sh.pose, &fv, pose_identity(), speculative, 0, 1,
                   NPH_TRI_POINTS, hit);
    if (hit->point_count == 0U || v3_dot(hit->normal, f.n) <= -NPH_SURFACE_FACE_COS) {
        return;
    }
    u32 deepest = 0U;
    for (u32 i = 1U; i < hit->point_count; ++i) {
        deepest = hit->points[i].separation < hit->points[deepest].separation ? i : deepest;
    }
    const f32 depth = hit->points[deepest].separation;
    const NphVec3 on = v3_mul_add(hit->points[deepest].point, hit->normal, 0.5F * depth);
    const int verdict = facet_verdict(&f, facet_locate(&f, on), v3_neg(hit->normal));
    if (verdict == NPH_VERDICT_KEEP) {
        return;
    }
    if (verdict == NPH_VERDICT_FACE) {
        team_hull_face(tm, &S->poly, &sh.view, sh.pose, &fv, pose_identity(), 0U, speculative,
                       hit);
        f32 s = NPH_F32_MAX;
        for (u32 i = 0U; i < hit->point_count; ++i) {
            s = nph_minf(s, hit->points[i].separation);
        }
        if (hit->point_count == 0U || s >= depth - NPH_PHYSICS_LINEAR_SLOP) {
            return;
        }
    }
    hit->point_count = 0U;
}

// The queued facets q[0..n) collided and merged in order.
template <u32 T, u32 CAP>
NPH_TD void surface_chunk(const NphTeam<T> &tm, NphSurfaceLds<T, CAP> *S, const NphSurfaceShape &sh,
                          const NphPhysicsMesh *mesh, NphDVec3 origin, u32 n, f32 speculative,
                          int two_sided, u32 *nc)
{
    const u32 *tri = mesh_triangles(mesh);
    const NphVec3 *verts = mesh_vertices(mesh);
    const u8 *edges = mesh_edges(mesh);
    const bool single = sh.kind == 0U;
    if (sh.kind != 2U) {
        // Round shapes: a facet a lane, then the merge in queue order.
        if (tm.lane < n) {
            const u32 t = S->queue[tm.lane];
            NphVec3 c[3];
            for (u32 k = 0U; k < 3U; ++k) {
                c[k] = relative_vertex(verts[tri[3U * t + k]], origin);
            }
            NphFacetHit h;
            h.normal = v3(0.0F, 1.0F, 0.0F);
            h.count = 0U;
            NphFacet f;
            if (facet_make(&f, c, edges[t], two_sided, sh.center)) {
                if (sh.kind == 0U) {
                    ball_facet(&f, sh.radius, speculative, &h);
                } else {
                    capsule_facet(&f, sh.core0, sh.core1, sh.radius, speculative, &h);
                }
            }
            S->hits[tm.lane] = h;
        }
        tm.sync();
        for (u32 j = 0U; j < n; ++j) {
            const u32 count = S->hits[j].count;
            if (count > 0U) {
                NphMPoint p[2];
                for (u32 i = 0U; i < count; ++i) {
                    p[i] = S->hits[j].p[i];
                }
                cluster_merge(tm, S, nc, S->hits[j].normal, count, p, S->queue[j], speculative,
                              single);
            }
        }
        tm.sync();
        return;
    }
    // Polytopes: the team on one facet at a time.
    for (u32 j = 0U; j < n; ++j) {
        const u32 t = S->queue[j];
        NphVec3 c[3];
        for (u32 k = 0U; k < 3U; ++k) {
            c[k] = relative_vertex(verts[tri[3U * t + k]], origin);
        }
        NphFacet f;
        if (!facet_make(&f, c, edges[t], two_sided, sh.center)) {
            continue;
        }
        NphTriContact hit;
        facet_polytope(tm, S, sh, f, speculative, &hit);
        if (hit.point_count > 0U) {
            cluster_merge(tm, S, nc, hit.normal, hit.point_count, hit.points, t, speculative,
                          false);
        }
    }
    tm.sync();
}

// The convex shape of geometry `g` at the origin, turned by `rotation`; a
// box's corners go to the team's LDS.
template <u32 T, u32 CAP>
NPH_TD NphSurfaceShape surface_shape(const NphTeam<T> &tm, NphSurfaceLds<T, CAP> *S,
                                     const NphGeom *g, NphQuat rotation)
{
    NphSurfaceShape sh;
    sh.radius = 0.0F;
    sh.core0 = sh.core1 = v3_zero();
    sh.center = v3_zero();
    sh.pose = pose_make(v3_zero(), rotation);
    if (g->type == NPH_PHYSICS_SHAPE_SPHERE) {
        sh.kind = 0U;
        sh.radius = g->a;
    } else if (g->type == NPH_PHYSICS_SHAPE_CAPSULE) {
        const NphVec3 half = v3_scale(q_axis_y(rotation), g->a);
        sh.kind = g->a > 0.0F ? 1U : 0U;
        sh.radius = g->b;
        sh.core0 = v3_neg(half);
        sh.core1 = half;
    } else {
        sh.kind = 2U;
        sh.view = team_view(tm, &S->poly, g, 0U);
        if (g->type == NPH_PHYSICS_SHAPE_HULL) {
            sh.center = q_rotate(rotation, g->hull->centroid);
        }
    }
    return sh;
}

// One facet's contact with the shape on the whole team (every lane ends
// with it): a ball's or a capsule's closed form, or a polytope's.
template <u32 T, u32 CAP>
NPH_TD void facet_contact(const NphTeam<T> &tm, NphSurfaceLds<T, CAP> *S,
                          const NphSurfaceShape &sh, const NphFacet &f, f32 speculative,
                          NphTriContact *out)
{
    if (sh.kind == 2U) {
        facet_polytope(tm, S, sh, f, speculative, out);
        return;
    }
    NphFacetHit h;
    h.normal = v3(0.0F, 1.0F, 0.0F);
    h.count = 0U;
    if (sh.kind == 0U) {
        ball_facet(&f, sh.radius, speculative, &h);
    } else {
        capsule_facet(&f, sh.core0, sh.core1, sh.radius, speculative, &h);
    }
    tri_clear(out);
    out->normal = h.normal;
    for (u32 i = 0U; i < h.count; ++i) {
        tri_push(out, h.p[i].point, h.p[i].separation, h.p[i].id);
    }
}

// Shape `g`, turned by `rotation` and placed at `position` in the mesh's
// frame, against the mesh: the pair's manifolds (at most four, deepest
// first) in S->out, in the mesh's frame relative to `position`.
template <u32 T, u32 CAP>
NPH_TD void surface_collide(const NphTeam<T> &tm, NphSurfaceLds<T, CAP> *S, const NphGeom *g,
                            NphQuat rotation, NphDVec3 position, const NphPhysicsMesh *mesh,
                            f32 speculative_distance)
{
    const f32 speculative = nph_maxf(speculative_distance, 0.0F);
    const int two_sided = (mesh->flags & NPH_PHYSICS_MESH_TWO_SIDED) != 0U;
    const NphSurfaceShape sh = surface_shape(tm, S, g, rotation);
    // The query box: the shape's box at its rounded place, grown by the
    // speculative distance and the rounding of that place.
    const NphVec3 at = v3((f32)position.x, (f32)position.y, (f32)position.z);
    const NphVec3 aa = v3_abs(at);
    const f32 rounding = 2.4e-7F * nph_maxf(aa.x, nph_maxf(aa.y, aa.z));
    const NphAABB box = aabb_grow(geom_aabb(g, pose_make(at, rotation)), speculative + rounding);
    // Gather, collide and merge (steps 1-3).
    const NphPhysicsMeshNode *nodes = mesh_nodes(mesh);
    u32 nc = 0U;
    u32 top = 1U;
    u32 queued = 0U;
    if (tm.lane == 0U) {
        S->stack[0] = 0U;
    }
    tm.sync();
    for (;;) {
        while (queued >= T) {
            surface_chunk(tm, S, sh, mesh, position, T, speculative, two_sided, &nc);
            const u32 rest = queued - T;
            const u32 moved = tm.lane < rest ? S->queue[T + tm.lane] : 0U;
            tm.sync();
            if (tm.lane < rest) {
                S->queue[tm.lane] = moved;
            }
            tm.sync();
            queued = rest;
        }
        if (top == 0U) {
            break;
        }
        const u32 code = S->stack[--top];
        tm.sync();
        if (mesh_child_is_leaf(code)) {
            const u32 t = mesh_leaf_first(code) + tm.lane;
            bool in = false;
            if (tm.lane < mesh_leaf_count(code)) {
                NphVec3 p[3];
                mesh_corners(mesh, t, p);
                in = aabb_overlaps(corners_box(p), box);
            }
            const u64 m = tm.ballot(in);
            if (in) {
                S->queue[queued + tm.below(m)] = t;
            }
            queued += (u32)__builtin_popcountll(m);
        } else {
            const NphPhysicsMeshNode *nd = &nodes[code];
            const u32 child = tm.lane < 4U ? nd->child[tm.lane] : NPH_PHYSICS_MESH_CHILD_EMPTY;
            const bool in = child != NPH_PHYSICS_MESH_CHILD_EMPTY &&
                            aabb_overlaps(bvh4_slot_box(nd, tm.lane), box);
            const u64 m = tm.ballot(in);
            const u32 pushes = (u32)__builtin_popcountll(m);
            if (top + pushes > NPH_BVH4_STACK) {
                top = 0U; // cannot happen for a validated block: the walk ends
            } else {
                // Slot k lands below the slots after it: the first slot on top.
                if (in) {
                    const u64 after = m & ~((2ULL << tm.lane) - 1ULL);
                    S->stack[top + (u32)__builtin_popcountll(after)] = child;
                }
                top += pushes;
            }
        }
        tm.sync();
    }
    if (queued > 0U) {
        surface_chunk(tm, S, sh, mesh, position, queued, speculative, two_sided, &nc);
    }
    // Finish (step 4): each cluster reduced on its own lane, then ranked.
    const bool capsule = sh.kind == 1U;
    const bool single = sh.kind == 0U;
    if (tm.lane < nc) {
        NphClusterHead *h = &S->head[tm.lane];
        h->count = capsule ? capsule_keep(S->pts[tm.lane], h->count, v3_sub(sh.core1, sh.core0))
                           : manifold_reduce(S->pts[tm.lane], h->count, h->normal);
    }
    tm.sync();
    if (tm.lane < nc) {
        const f32 d = S->head[tm.lane].depth;
        u32 rank = 0U;
        for (u32 k = 0U; k < nc; ++k) {
            const f32 o = S->head[k].depth;
            rank += (o < d || (o == d && k < tm.lane)) ? 1U : 0U;
        }
        if (rank < NPH_PHYSICS_MESH_MAX_MANIFOLDS) {
            const NphClusterHead *h = &S->head[tm.lane];
            NphManifold *m = &S->out.manifolds[rank];
            m->normal = h->normal;
            m->point_count = h->count;
            for (u32 j = 0U; j < h->count; ++j) {
                m->points[j] = S->pts[tm.lane][j];
            }
            if (single) {
                // Placed again along the merged normal.
                m->points[0].point = v3_scale(m->normal, sh.radius + 0.5F * m->points[0].separation);
            }
        }
    }
    if (tm.lane == 0U) {
        S->out.count = nc < NPH_PHYSICS_MESH_MAX_MANIFOLDS ? nc : NPH_PHYSICS_MESH_MAX_MANIFOLDS;
    }
    tm.sync();
}

#endif // __cplusplus
#endif // NPH_SURFACE_H
