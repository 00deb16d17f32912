// SPDX-License-Identifier: Apache-2.0
// Polytope pairs on a team of lanes: boxes and hulls against spheres,
// capsules, each other and mesh triangles, with the separating-axis test
// of Gregorius (GDC 2013) and incident-face clipping, in the reference's
// arithmetic (collide_hull.c:156-1266) so that manifolds, feature ids and
// caches are its bits.
//
// How the work spreads over a team of T lanes (nph_team.h):
//  * B's vertices go to A's frame once, a vertex a lane, into the team's
//    LDS; face queries take a face a lane and keep the first best face by a
//    team reduction;
//  * edge pairs (A's edge k, B's edge j) are numbered k-major and taken T
//    at a time: each lane runs the Gauss-map test of its pair and prepares
//    its separation; the pairs that pass and could still win are visited in
//    number order by the whole team (a ballot's set bits), which applies the
//    reference's running-best rule to them, bit for bit;
//  * the clip polygon lives in LDS: each clipping plane is applied by every
//    lane to its own polygon edge at once, the output places coming from an
//    exclusive scan of what each lane emits (0, 1 or 2 points), so the
//    output order is the serial clipper's; candidates below the speculative
//    distance are compacted the same way;
//  * GJK's support searches take a vertex a lane and reduce.
// Everything else (the decisions between axes, the closed forms of the edge
// contacts, the reduction of a clipped face) is uniform: every lane computes
// it, and lane 0 writes what leaves the team.
//
// With T == 1 the same code runs one pair a thread (the mesh pairs) or on
// the host.
#ifndef NPH_POLYTOPE_H
#define NPH_POLYTOPE_H

#include "nph_narrow.h"
#include "nph_team.h"

#if defined(__cplusplus)

// A team's working memory for hulls of at most V vertices and clip
// polygons of at most CAP points (LDS for a team of several lanes).
template <u32 V, u32 CAP> struct NphPolyScratch {
    NphVec3 bv[V]; // B's vertices in A's frame
    union {
        NphVec3 bn[2U * V];     // B's face normals in A's frame (the axis test)
        NphMPoint poly[2][CAP]; // clip polygons (the face contact)
    };
    NphBoxStorage box[2]; // box views
};

// Two polytopes in A's frame.
struct NphPolyPair {
    const NphView *a;
    const NphView *b;
    NphPhysicsPose rel; // B in A's frame
};

// What the separating-axis test found.
struct NphSat {
    f32 face_a_sep;
    u32 face_a;
    f32 face_b_sep;
    u32 face_b;
    f32 edge_sep;
    u32 edge_a;
    u32 edge_b;
    NphVec3 edge_axis; // A's frame, from A towards B
};

// B's face f as a plane in A's frame.
NPH_TD NphPlane face_b_plane(const NphPolyPair &pp, u32 f)
{
    const NphPlane pl = pp.b->planes[f];
    const NphVec3 n = q_rotate(pp.rel.rotation, pl.normal);
    return plane_make(n, pl.d - v3_dot(n, pp.rel.position));
}

// B's vertices into A's frame (scratch.bv), a vertex a lane.
template <u32 T, u32 V, u32 CAP>
NPH_TD void place_b(const NphTeam<T> &tm, const NphPolyPair &pp, NphPolyScratch<V, CAP> *sc)
{
    for (u32 i = tm.lane; i < pp.b->vertex_count; i += T) {
        sc->bv[i] = pose_point(pp.rel, pp.b->vertices[i]);
    }
    tm.sync();
}

// The smallest projection of a point set on n.
NPH_TD f32 lowest_projection(const NphVec3 *v, u32 count, NphVec3 n)
{
    f32 m = v3_dot(v[0], n);
    for (u32 i = 1U; i < count; ++i) {
        m = nph_minf(m, v3_dot(v[i], n));
    }
    return m;
}

// The first index of the largest value over [0, count) of value(i) (values
// not above -FLT_MAX never win): index NPH_NULL32 when none is.
template <u32 T, class F> NPH_TD NphPick team_first_max(const NphTeam<T> &tm, u32 count, F value)
{
    NphPick run = {-NPH_F32_MAX, NPH_NULL32};
    for (u32 base = 0U; base < count; base += T) {
        const u32 i = base + tm.lane;
        NphPick mine = {-NPH_F32_MAX, NPH_NULL32};
        if (i < count) {
            const f32 s = value(i);
            if (s > -NPH_F32_MAX) {
                mine.value = s;
                mine.index = i;
            }
        }
        run = pick_later_max(run, tm.first_max(mine));
    }
    return run;
}

// The first index of the smallest value below FLT_MAX.
template <u32 T, class F> NPH_TD NphPick team_first_min(const NphTeam<T> &tm, u32 count, F value)
{
    NphPick run = {NPH_F32_MAX, NPH_NULL32};
    for (u32 base = 0U; base < count; base += T) {
        const u32 i = base + tm.lane;
        NphPick mine = {NPH_F32_MAX, NPH_NULL32};
        if (i < count) {
            const f32 s = value(i);
            if (s < NPH_F32_MAX) {
                mine.value = s;
                mine.index = i;
            }
        }
        run = pick_later_min(run, tm.first_min(mine));
    }
    return run;
}

// The support point of a set along d: the first of greatest projection.
template <u32 T> NPH_TD u32 team_support(const NphTeam<T> &tm, const NphVec3 *p, u32 count, NphVec3 d)
{
    NphPick run = {0.0F, NPH_NULL32};
    for (u32 base = 0U; base < count; base += T) {
        const u32 i = base + tm.lane;
        NphPick mine = {0.0F, NPH_NULL32};
        if (i < count) {
            mine.value = v3_dot(p[i], d);
            mine.index = i;
        }
        run = pick_later_max(run, tm.first_max(mine));
    }
    return run.index;
}

// GJK between two point sets, supports searched by the team.
template <u32 T>
NPH_TD void team_distance(const NphTeam<T> &tm, const NphVec3 *a, u32 na, const NphVec3 *b, u32 nb,
                          NphDistance *out)
{
    NphGjk g;
    gjk_reset(&g);
    gjk_seed(&g, 0U, gjk_through_identity(a[0]), 0U, gjk_through_identity(b[0]));
    gjk_begin(&g);
    while (!g.done) {
        const u32 ia = team_support(tm, a, na, v3_neg(g.dir));
        const u32 ib = team_support(tm, b, nb, g.dir);
        gjk_feed(&g, ia, gjk_through_identity(a[ia]), ib, gjk_through_identity(b[ib]));
    }
    gjk_finish(&g, out);
}

// ---------------------------------------------------------------------------
// Hull against sphere (collide_hull.c:156-201)
// ---------------------------------------------------------------------------
template <u32 T>
NPH_TD void team_hull_sphere(const NphTeam<T> &tm, const NphView *a, NphPhysicsPose pa, f32 radius,
                             NphPhysicsPose pb, f32 speculative, NphManifold *out)
{
    const NphVec3 c = pose_inv_point(pa, pb.position);
    const NphPick top =
        team_first_max(tm, a->face_count, [&](u32 f) { return plane_distance(a->planes[f], c); });
    const f32 best = top.index == NPH_NULL32 ? -NPH_F32_MAX : top.value;
    const u32 face = top.index == NPH_NULL32 ? 0U : top.index;
    if (best > radius + speculative) {
        return;
    }
    NphVec3 n, on_a;
    f32 sep;
    u32 id;
    if (best <= 0.0F) {
        // Inside: out through the shallowest face.
        n = a->planes[face].normal;
        on_a = v3_mul_add(c, n, -best);
        sep = best - radius;
        id = face;
    } else {
        // Outside: the nearest point may be on an edge or a corner.
        NphDistance d;
        team_distance(tm, a->vertices, a->vertex_count, &c, 1U, &d);
        if (d.distance > radius + speculative) {
            return;
        }
        n = d.distance > 1e-6F ? d.normal : a->planes[face].normal;
        on_a = d.point_a;
        sep = d.distance - radius;
        id = 0x10000U | face;
    }
    const NphVec3 on_b = v3_mul_add(c, n, -radius);
    out->normal = pose_vector(pa, n);
    manifold_push(out, pose_point(pa, v3_scale(v3_add(on_a, on_b), 0.5F)), sep, id);
}

// ---------------------------------------------------------------------------
// Hull against capsule (collide_hull.c:203-376)
// ---------------------------------------------------------------------------

// The segment [p, q] clipped to the prism over face `face` (its side
// planes); false when nothing is left. One lane's loop: the clamps keep
// the serial order (signed zeros included).
NPH_TD int clip_segment_to_face(const NphView *h, u32 face, NphVec3 *p, NphVec3 *q)
{
    const NphVec3 n = h->planes[face].normal;
    const NphVec3 p0 = *p;
    const NphVec3 d = v3_sub(*q, *p);
    f32 lo = 0.0F;
    f32 hi = 1.0F;
    const u32 first = h->faces[face];
    u32 e = first;
    u32 sides = 0U;
    do {
        const NphVec3 a = h->vertices[h->edges[e].origin];
        const NphVec3 out = v3_cross(v3_sub(h->vertices[edge_end(h, e)], a), n);
        const f32 at0 = v3_dot(out, v3_sub(p0, a));
        const f32 rate = v3_dot(out, d);
        // Inside where at0 + t rate <= 0.
        if (rate == 0.0F) {
            if (at0 > 0.0F) {
                return 0;
            }
        } else {
            const f32 t = -at0 / rate;
            if (rate > 0.0F) {
                hi = nph_minf(hi, t);
            } else {
                lo = nph_maxf(lo, t);
            }
            if (lo > hi) {
                return 0;
            }
        }
        e = h->edges[e].next;
        sides += 1U;
    } while (e != first && sides < h->edge_count);
    *p = v3_mul_add(p0, d, lo);
    *q = v3_mul_add(p0, d, hi);
    return 1;
}

template <u32 T>
NPH_TD void team_hull_capsule(const NphTeam<T> &tm, const NphView *a, NphPhysicsPose pa,
                              f32 half_height, f32 radius, NphPhysicsPose pb, f32 speculative,
                              NphManifold *out)
{
    const NphPhysicsPose rel = pose_inv_mul(pa, pb);
    const NphVec3 up = v3_scale(q_axis_y(rel.rotation), half_height);
    NphVec3 seg[2] = {v3_sub(rel.position, up), v3_add(rel.position, up)};
    const f32 reach = radius + speculative;
    // Faces: the segment's lower end against each.
    const NphPick fp = team_first_max(tm, a->face_count, [&](u32 f) {
        const NphPlane pl = a->planes[f];
        return nph_minf(plane_distance(pl, seg[0]), plane_distance(pl, seg[1]));
    });
    const f32 face_sep = fp.index == NPH_NULL32 ? -NPH_F32_MAX : fp.value;
    const u32 face = fp.index == NPH_NULL32 ? 0U : fp.index;
    if (face_sep > reach) {
        return;
    }
    // Edges whose faces lie on opposite sides of the plane perpendicular to
    // the segment (the segment's Gauss map).
    const NphVec3 d = v3_sub(seg[1], seg[0]);
    const f32 d_sq = v3_len_sq(d);
    f32 edge_sep = -NPH_F32_MAX;
    u32 edge = 0U;
    NphVec3 edge_axis = v3(0.0F, 1.0F, 0.0F);
    if (d_sq > 1e-12F) {
        NphPick run = {-NPH_F32_MAX, NPH_NULL32};
        const u32 pairs = a->edge_count / 2U;
        for (u32 base = 0U; base < pairs; base += T) {
            const u32 k = base + tm.lane;
            NphPick mine = {-NPH_F32_MAX, NPH_NULL32};
            NphVec3 axis = v3_zero();
            if (k < pairs) {
                const u32 e = 2U * k;
                const NphVec3 n1 = a->planes[a->edges[e].face].normal;
                const NphVec3 n2 = a->planes[a->edges[e + 1U].face].normal;
                if (!(v3_dot(n1, d) * v3_dot(n2, d) >= 0.0F)) {
                    const NphVec3 p = a->vertices[a->edges[e].origin];
                    const NphVec3 u = v3_sub(a->vertices[edge_end(a, e)], p);
                    const NphVec3 x = v3_cross(u, d);
                    const f32 len = v3_len(x);
                    if (!(len < 0.005F * nph_sqrtf(v3_len_sq(u) * d_sq))) {
                        axis = v3_scale(x, 1.0F / len);
                        if (v3_dot(axis, v3_sub(p, a->centroid)) < 0.0F) {
                            axis = v3_neg(axis);
                        }
                        const f32 s = v3_dot(axis, v3_sub(seg[0], p));
                        if (s > -NPH_F32_MAX) {
                            mine.value = s;
                            mine.index = k;
                        }
                    }
                }
            }
            const NphPick chunk = tm.first_max(mine);
            const NphPick next = pick_later_max(run, chunk);
            if (next.index != run.index) {
                edge_axis = tm.get(axis, next.index - base);
            }
            run = next;
        }
        if (run.index != NPH_NULL32) {
            edge_sep = run.value;
            edge = 2U * run.index;
        }
    }
    if (edge_sep > reach) {
        return;
    }
    const int use_face = face_sep >= edge_sep - NPH_EDGE_BIAS;
    const f32 best = use_face ? face_sep : edge_sep;
    if (best > NPH_PHYSICS_LINEAR_SLOP) {
        // Apart: GJK's exact gap, and its features when they are not the
        // axis test's.
        NphDistance gd;
        team_distance(tm, a->vertices, a->vertex_count, seg, d_sq > 0.0F ? 2U : 1U, &gd);
        if (gd.distance > reach) {
            return;
        }
        const NphVec3 sat_axis = use_face ? a->planes[face].normal : edge_axis;
        if (gd.distance > 0.0F &&
            (!use_face || v3_dot(gd.normal, sat_axis) < NPH_FEATURE_AGREEMENT)) {
            const NphVec3 on_b = v3_mul_add(gd.point_b, gd.normal, -radius);
            out->normal = pose_vector(pa, gd.normal);
            manifold_push(out, pose_point(pa, v3_scale(v3_add(gd.point_a, on_b), 0.5F)),
                          gd.distance - radius, 0x20000U);
            return;
        }
    }
    if (use_face) {
        const NphPlane pl = a->planes[face];
        NphVec3 p = seg[0];
        NphVec3 q = seg[1];
        if (!clip_segment_to_face(a, face, &p, &q)) {
            // Beside the face: its lower end decides.
            p = plane_distance(pl, seg[0]) <= plane_distance(pl, seg[1]) ? seg[0] : seg[1];
            q = p;
        }
        out->normal = pose_vector(pa, pl.normal);
        const u32 count = v3_dist_sq(p, q) > 1e-12F ? 2U : 1U;
        for (u32 k = 0U; k < count; ++k) {
            const NphVec3 x = k == 0U ? p : q;
            const f32 dist = plane_distance(pl, x);
            const f32 sep = dist - radius;
            if (sep <= speculative) {
                manifold_push(out, pose_point(pa, v3_mul_add(x, pl.normal, -0.5F * (dist + radius))),
                              sep, (face << 8U) | k);
            }
        }
        return;
    }
    // An edge against the core: one point.
    const NphSegDist sd = segment_distance(a->vertices[a->edges[edge].origin],
                                           a->vertices[edge_end(a, edge)], seg[0], seg[1]);
    const f32 sep = v3_dot(v3_sub(sd.closest2, sd.closest1), edge_axis) - radius;
    if (sep > speculative) {
        return;
    }
    const NphVec3 on_b = v3_mul_add(sd.closest2, edge_axis, -radius);
    out->normal = pose_vector(pa, edge_axis);
    manifold_push(out, pose_point(pa, v3_scale(v3_add(sd.closest1, on_b), 0.5F)), sep,
                  0x10000U | edge);
}

// ---------------------------------------------------------------------------
// The separating-axis test of two polytopes (collide_hull.c:378-783)
// ---------------------------------------------------------------------------

// The plane of the Gauss-map arc of B's edge between face normals c and d
// (negated): d x c, or for a flat edge (opposite faces, a triangle's) the
// limit of that cross product, the edge reversed.
NPH_TD NphVec3 arc_normal(NphVec3 c, NphVec3 d, NphVec3 origin, NphVec3 head)
{
    const NphVec3 u = v3_cross(d, c);
    return v3_len_sq(u) > 1e-12F ? u : v3_sub(origin, head);
}

// The separation along an edge pair's cross product, oriented away from
// A's centroid, with its axis; -FLT_MAX for nearly parallel edges.
NPH_TD f32 edge_axis_separation(NphVec3 pa, NphVec3 ua, NphVec3 pb, NphVec3 ub, NphVec3 centroid,
                                NphVec3 *axis)
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
    *axis = n;
    return v3_dot(n, v3_sub(pb, pa));
}

// One edge pair of the search, prepared by its lane: the unnormalised axis
// x (oriented), |x|^2, the separation along it times |x| (sx), and the
// separation and unit axis. `viable` is false for nearly parallel edges.
struct NphEdgeCand {
    f32 len_sq;
    f32 sx;
    f32 sep;
    NphVec3 axis;
    int viable;
};

NPH_TD NphEdgeCand edge_candidate(NphVec3 pa, NphVec3 ua, NphVec3 pb, NphVec3 ub, NphVec3 centroid)
{
    NphEdgeCand c;
    NphVec3 x = v3_cross(ua, ub);
    c.len_sq = v3_len_sq(x);
    c.viable = !(c.len_sq < 2.5e-5F * v3_len_sq(ua) * v3_len_sq(ub));
    if (v3_dot(x, v3_sub(pa, centroid)) < 0.0F) {
        x = v3_neg(x);
    }
    c.sx = v3_dot(x, v3_sub(pb, pa));
    const f32 inv = 1.0F / nph_sqrtf(c.len_sq);
    c.axis = v3_scale(x, inv);
    c.sep = c.sx * inv;
    return c;
}

// Whether a candidate passes the search's cheap test against the running
// best (collide_hull.c:530-564): the separation s = sx / |x| compared with
// `best` through squares, without the division.
NPH_TD int edge_beats(f32 best, f32 sx, f32 len_sq)
{
    if (!(best > -NPH_F32_MAX)) {
        return 1;
    }
    const f32 bound = best * best * len_sq;
    return best < 0.0F ? (sx >= 0.0F || sx * sx < bound) : (sx > 0.0F && sx * sx > bound);
}

// The test proper: false when an axis separates the pair by more than
// `limit` (the cache then names it). B's vertices are in sc->bv.
template <u32 T, u32 V, u32 CAP>
NPH_TD int team_sat(const NphTeam<T> &tm, const NphPolyPair &pp, NphPolyScratch<V, CAP> *sc,
                    f32 limit, NphCache *cache, NphSat *r)
{
    const NphView *a = pp.a;
    const NphView *b = pp.b;
    const NphVec3 *bv = sc->bv;
    // The axis that decided the pair last time.
    if (cache != 0 && cache->type != NPH_CACHE_NONE) {
        f32 s = -NPH_F32_MAX;
        if (cache->type == NPH_CACHE_FACE_A && cache->index_a < a->face_count) {
            const NphPlane pl = a->planes[cache->index_a];
            s = lowest_projection(bv, b->vertex_count, pl.normal) + pl.d;
        } else if (cache->type == NPH_CACHE_FACE_B && cache->index_b < b->face_count) {
            const NphPlane pl = face_b_plane(pp, cache->index_b);
            s = lowest_projection(a->vertices, a->vertex_count, pl.normal) + pl.d;
        } else if (cache->type == NPH_CACHE_EDGES && cache->index_a < a->edge_count &&
                   cache->index_b < b->edge_count) {
            const u32 ea = cache->index_a;
            const u32 eb = cache->index_b;
            const NphVec3 pa = a->vertices[a->edges[ea].origin];
            const NphVec3 pb = bv[b->edges[eb].origin];
            NphVec3 axis;
            s = edge_axis_separation(pa, v3_sub(a->vertices[edge_end(a, ea)], pa), pb,
                                     v3_sub(bv[edge_end(b, eb)], pb), a->centroid, &axis);
        }
        if (s > limit) {
            return 0;
        }
    }
    // A's faces against B's vertices.
    const NphPick fa = team_first_max(tm, a->face_count, [&](u32 f) {
        const NphPlane pl = a->planes[f];
        return lowest_projection(bv, b->vertex_count, pl.normal) + pl.d;
    });
    r->face_a_sep = fa.index == NPH_NULL32 ? -NPH_F32_MAX : fa.value;
    r->face_a = fa.index == NPH_NULL32 ? 0U : fa.index;
    if (r->face_a_sep > limit) {
        if (cache != 0) {
            cache->type = NPH_CACHE_FACE_A;
            cache->index_a = (u16)r->face_a;
        }
        return 0;
    }
    // B's faces against A's vertices; their normals stay for the edges.
    const NphPick fb = team_first_max(tm, b->face_count, [&](u32 f) {
        const NphPlane pl = face_b_plane(pp, f);
        sc->bn[f] = pl.normal;
        return lowest_projection(a->vertices, a->vertex_count, pl.normal) + pl.d;
    });
    tm.sync();
    r->face_b_sep = fb.index == NPH_NULL32 ? -NPH_F32_MAX : fb.value;
    r->face_b = fb.index == NPH_NULL32 ? 0U : fb.index;
    if (r->face_b_sep > limit) {
        if (cache != 0) {
            cache->type = NPH_CACHE_FACE_B;
            cache->index_b = (u16)r->face_b;
        }
        return 0;
    }
    // Edge pairs, A's edge k (half-edge 2k) major; B's edges across the
    // lanes, T at a time.
    const u32 na = a->edge_count / 2U;
    const u32 nb = b->edge_count / 2U;
    f32 best = -NPH_F32_MAX;
    u32 best_a = 0U, best_b = 0U;
    NphVec3 best_axis = v3(0.0F, 1.0F, 0.0F);
    for (u32 k = 0U; k < na; ++k) {
        const u32 ea = 2U * k;
        const NphVec3 n1 = a->planes[a->edges[ea].face].normal;
        const NphVec3 n2 = a->planes[a->edges[ea + 1U].face].normal;
        const NphVec3 w = v3_cross(n2, n1);
        const NphVec3 pa = a->vertices[a->edges[ea].origin];
        const NphVec3 ua = v3_sub(a->vertices[edge_end(a, ea)], pa);
        for (u32 base = 0U; base < nb; base += T) {
            const u32 j = base + tm.lane;
            int keep = 0;
            NphEdgeCand cand;
            cand.len_sq = cand.sx = cand.sep = 0.0F;
            cand.axis = v3_zero();
            if (j < nb) {
                const u32 eb = 2U * j;
                const NphVec3 c = v3_neg(sc->bn[b->edges[eb].face]);
                const NphVec3 dn = v3_neg(sc->bn[b->edges[eb + 1U].face]);
                const NphVec3 pb = bv[b->edges[eb].origin];
                const NphVec3 hb = bv[edge_end(b, eb)];
                // Gregorius: the arcs n1 n2 (A) and c d (B, negated) cross.
                const NphVec3 u = arc_normal(c, dn, pb, hb);
                const f32 cw = v3_dot(c, w);
                const f32 dw = v3_dot(dn, w);
                const f32 nu1 = v3_dot(n1, u);
                const f32 nu2 = v3_dot(n2, u);
                if (cw * dw < 0.0F && nu1 * nu2 < 0.0F && cw * nu2 > 0.0F) {
                    cand = edge_candidate(pa, ua, pb, v3_sub(hb, pb), a->centroid);
                    // Not above the best so far: never taken later either
                    // (the best only grows).
                    keep = cand.viable && cand.sep > best;
                }
            }
            u64 pass = tm.ballot(keep != 0);
            while (pass != 0ULL) {
                const u32 src = nph_lowest_bit(pass);
                pass &= pass - 1ULL;
                const f32 len_sq = tm.get(cand.len_sq, src);
                const f32 sx = tm.get(cand.sx, src);
                const f32 sep = tm.get(cand.sep, src);
                const NphVec3 axis = tm.get(cand.axis, src);
                if (edge_beats(best, sx, len_sq) && sep > best) {
                    best = sep;
                    best_a = ea;
                    best_b = 2U * (base + src);
                    best_axis = axis;
                }
            }
        }
    }
    r->edge_sep = best;
    r->edge_a = best_a;
    r->edge_b = best_b;
    r->edge_axis = best_axis;
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

// ---------------------------------------------------------------------------
// The face contact: the incident face clipped by the reference face's side
// planes (collide_hull.c:785-877)
// ---------------------------------------------------------------------------

// The polygon poly[from] (len points) clipped by the plane sn.x = so (inside
// where sn.x - so <= 0) into poly[1 - from]; returns the new length. Side
// `side` tags the points it makes.
template <u32 T, u32 V, u32 CAP>
NPH_TD u32 team_clip(const NphTeam<T> &tm, NphPolyScratch<V, CAP> *sc, u32 from, u32 len, NphVec3 sn,
                     f32 so, u32 side)
{
    const NphMPoint *in = sc->poly[from];
    NphMPoint *outp = sc->poly[from ^ 1U];
    u32 written = 0U;
    int full = 0;
    for (u32 base = 0U; base < len && !full; base += T) {
        const u32 i = base + tm.lane;
        u32 emit = 0U;
        NphMPoint p, x;
        int keep_p = 0, cross = 0;
        if (i < len) {
            p = in[i];
            const NphMPoint q = in[i + 1U == len ? 0U : i + 1U];
            const f32 dp = v3_dot(sn, p.point) - so;
            const f32 dq = v3_dot(sn, q.point) - so;
            keep_p = dp <= 0.0F;
            cross = (dp <= 0.0F) != (dq <= 0.0F);
            if (cross) {
                x.point = v3_lerp(p.point, q.point, dp / (dp - dq));
                x.separation = 0.0F;
                x.id = nph_mix_id(nph_mix_id(p.id, q.id), 0x100U + side);
            }
            emit = (u32)keep_p + (u32)cross;
        }
        // Output places: an exclusive scan of the lanes' counts (at most two
        // each, so the ballots of the two bits add up to it).
        const u64 b0 = tm.ballot((emit & 1U) != 0U);
        const u64 b1 = tm.ballot((emit & 2U) != 0U);
        const u32 at = written + tm.below(b0) + 2U * tm.below(b1);
        // The serial clipper stops taking points once its output holds
        // CAP - 1: point i is taken when at + 1 < CAP.
        const int taken = i < len && at + 1U < CAP;
        if (taken) {
            u32 k = at;
            if (keep_p) {
                outp[k++] = p;
            }
            if (cross) {
                outp[k] = x;
            }
        }
        const u64 took = tm.ballot(taken != 0);
        const u64 lanes = (T == 64U) ? ~0ULL : ((1ULL << T) - 1ULL);
        const u32 in_chunk = len - base < T ? len - base : T;
        const u64 present = in_chunk >= 64U ? ~0ULL : ((1ULL << in_chunk) - 1ULL);
        // Points taken are a prefix; the chunk's written count is its
        // taken lanes' emits.
        written += tm.add(taken ? emit : 0U);
        full = (took & present & lanes) != (present & lanes);
    }
    tm.sync();
    return written;
}

// The face contact of reference face `ref` of R (A, or B when `flip`).
// Writes at most max_points points (A's frame) to pts and returns the normal
// from A to B.
template <u32 T, u32 V, u32 CAP>
NPH_TD NphVec3 team_face_contact(const NphTeam<T> &tm, const NphPolyPair &pp,
                                 NphPolyScratch<V, CAP> *sc, int flip, u32 ref, f32 speculative,
                                 u32 max_points, NphMPoint *pts, u32 *count)
{
    const NphView *rh = flip ? pp.b : pp.a;
    const NphView *ih = flip ? pp.a : pp.b;
    const NphVec3 *rv = flip ? sc->bv : pp.a->vertices;
    const NphVec3 *iv = flip ? pp.a->vertices : sc->bv;
    const NphPlane ref_plane = flip ? face_b_plane(pp, ref) : pp.a->planes[ref];
    const NphVec3 n = ref_plane.normal;
    // The incident face: the most anti-parallel one.
    const NphPick ip = team_first_min(tm, ih->face_count, [&](u32 f) {
        const NphVec3 fn = flip ? ih->planes[f].normal : face_b_plane(pp, f).normal;
        return v3_dot(fn, n);
    });
    const u32 inc = ip.index == NPH_NULL32 ? 0U : ip.index;
    tm.sync(); // the axis test's normals share the polygons' memory
    u32 len = 0U;
    if (tm.lane == 0U) {
        const u32 first = ih->faces[inc];
        u32 e = first;
        do {
            NphMPoint *p = &sc->poly[0][len];
            p->point = iv[ih->edges[e].origin];
            p->separation = 0.0F;
            p->id = ih->edges[e].origin;
            len += 1U;
            e = ih->edges[e].next;
        } while (e != first && len < CAP && len < ih->edge_count);
    }
    len = tm.get(len, 0U);
    tm.sync();
    u32 cur = 0U;
    const u32 first = rh->faces[ref];
    u32 e = first;
    u32 side = 0U;
    do {
        const NphVec3 a = rv[rh->edges[e].origin];
        const NphVec3 sn = v3_cross(v3_sub(rv[edge_end(rh, e)], a), n); // outward
        len = team_clip(tm, sc, cur, len, sn, v3_dot(sn, a), side);
        cur ^= 1U;
        e = rh->edges[e].next;
        side += 1U;
    } while (e != first && len > 0U && side < rh->edge_count);
    // Points within the speculative distance, compacted in order.
    const u32 tag = nph_mix_id((flip ? 0x80000000U : 0U) | (ref << 16U), inc);
    const NphMPoint *clipped = sc->poly[cur];
    NphMPoint *cand = sc->poly[cur ^ 1U];
    u32 kept = 0U;
    for (u32 base = 0U; base < len; base += T) {
        const u32 i = base + tm.lane;
        NphMPoint c;
        int keep = 0;
        if (i < len) {
            const f32 sep = plane_distance(ref_plane, clipped[i].point);
            if (sep <= speculative) {
                c.point = v3_mul_add(clipped[i].point, n, -0.5F * sep);
                c.separation = sep;
                c.id = nph_mix_id(tag, clipped[i].id);
                keep = 1;
            }
        }
        const u64 m = tm.ballot(keep != 0);
        if (keep) {
            cand[kept + tm.below(m)] = c;
        }
        kept += (u32)__builtin_popcountll(m);
    }
    tm.sync();
    if (kept > max_points) {
        if (tm.lane == 0U) {
            kept = manifold_reduce(cand, kept, n);
        }
        kept = tm.get(kept, 0U);
        tm.sync();
    }
    for (u32 i = 0U; i < kept; ++i) {
        pts[i] = cand[i];
    }
    *count = kept;
    return flip ? v3_neg(n) : n;
}

// ---------------------------------------------------------------------------
// Two polytopes (collide_hull.c:921-1017)
// ---------------------------------------------------------------------------

// The pair's points (A's frame turned to pa's) appended to `out`, up to
// `max_points`; `prefer_b` turns the tie between faces to B's (a triangle's
// face). `out` is left as it was when nothing touches.
template <u32 T, u32 V, u32 CAP>
NPH_TD void team_polytopes(const NphTeam<T> &tm, NphPolyScratch<V, CAP> *sc, const NphView *a,
                           NphPhysicsPose pa, const NphView *b, NphPhysicsPose pb, f32 speculative,
                           NphCache *cache, int prefer_b, u32 max_points, NphTriContact *out)
{
    NphPolyPair pp;
    pp.a = a;
    pp.b = b;
    pp.rel = pose_inv_mul(pa, pb);
    place_b(tm, pp, sc);
    NphSat r;
    if (!team_sat(tm, pp, sc, speculative, cache, &r)) {
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
        // Apart: GJK measures the gap; corner contacts take its one point.
        NphDistance gd;
        team_distance(tm, a->vertices, a->vertex_count, sc->bv, b->vertex_count, &gd);
        if (gd.distance > speculative) {
            return;
        }
        const NphVec3 sat_axis = use_edge ? r.edge_axis
                                          : (flip ? v3_neg(face_b_plane(pp, r.face_b).normal)
                                                  : a->planes[r.face_a].normal);
        if (gd.distance > 0.0F &&
            (use_edge || v3_dot(gd.normal, sat_axis) < NPH_FEATURE_AGREEMENT)) {
            out->normal = pose_vector(pa, gd.normal);
            tri_push(out, pose_point(pa, v3_scale(v3_add(gd.point_a, gd.point_b), 0.5F)),
                     gd.distance, NPH_ID_GJK_POINT);
            return;
        }
    }
    if (use_edge) {
        const NphSegDist sd = segment_distance(
            a->vertices[a->edges[r.edge_a].origin], a->vertices[edge_end(a, r.edge_a)],
            sc->bv[b->edges[r.edge_b].origin], sc->bv[edge_end(b, r.edge_b)]);
        const f32 sep = v3_dot(v3_sub(sd.closest2, sd.closest1), r.edge_axis);
        if (sep > speculative) {
            return;
        }
        out->normal = pose_vector(pa, r.edge_axis);
        tri_push(out, pose_point(pa, v3_scale(v3_add(sd.closest1, sd.closest2), 0.5F)), sep,
                 nph_mix_id(NPH_ID_EDGE_PAIR | r.edge_a, r.edge_b));
        return;
    }
    NphMPoint pts[NPH_TRI_POINTS];
    u32 count = 0U;
    const NphVec3 n = team_face_contact(tm, pp, sc, flip, flip ? r.face_b : r.face_a, speculative,
                                        max_points, pts, &count);
    out->normal = pose_vector(pa, n);
    for (u32 i = 0U; i < count; ++i) {
        tri_push(out, pose_point(pa, pts[i].point), pts[i].separation, pts[i].id);
    }
}

// B's face `face` as the reference face against A (collide_hull.c:879-904).
template <u32 T, u32 V, u32 CAP>
NPH_TD void team_hull_face(const NphTeam<T> &tm, NphPolyScratch<V, CAP> *sc, const NphView *a,
                           NphPhysicsPose pa, const NphView *b, NphPhysicsPose pb, u32 face,
                           f32 speculative, NphTriContact *out)
{
    NphPolyPair pp;
    pp.a = a;
    pp.b = b;
    pp.rel = pose_inv_mul(pa, pb);
    place_b(tm, pp, sc);
    NphMPoint pts[NPH_TRI_POINTS];
    u32 count = 0U;
    const NphVec3 n =
        team_face_contact(tm, pp, sc, 1, face, speculative, NPH_TRI_POINTS, pts, &count);
    tri_clear(out);
    out->normal = pose_vector(pa, n);
    for (u32 i = 0U; i < count; ++i) {
        tri_push(out, pose_point(pa, pts[i].point), pts[i].separation, pts[i].id);
    }
}

// ---------------------------------------------------------------------------
// Box against box: the fifteen axes of the oriented-box test (Gottschalk,
// Lin and Manocha 1996), then the contact as any polytope pair's
// (collide_hull.c:1019-1266)
// ---------------------------------------------------------------------------
struct NphBoxFrame {
    f32 r[3][3];  // A_i . B_j
    f32 ar[3][3]; // |r| + 1e-6
    f32 t[3];     // B's centre in A's frame
    f32 ha[3];
    f32 hb[3];
};

// Separation along A's axis i (*side: the side B is on).
NPH_TD f32 box_axis_a(const NphBoxFrame *x, u32 i, f32 *side)
{
    const f32 rb = x->hb[0] * x->ar[i][0] + x->hb[1] * x->ar[i][1] + x->hb[2] * x->ar[i][2];
    *side = x->t[i] >= 0.0F ? 1.0F : -1.0F;
    return nph_absf(x->t[i]) - (x->ha[i] + rb);
}

NPH_TD f32 box_axis_b(const NphBoxFrame *x, u32 j, f32 *side)
{
    const f32 tb = x->t[0] * x->r[0][j] + x->t[1] * x->r[1][j] + x->t[2] * x->r[2][j];
    const f32 ra = x->ha[0] * x->ar[0][j] + x->ha[1] * x->ar[1][j] + x->ha[2] * x->ar[2][j];
    *side = tb >= 0.0F ? 1.0F : -1.0F;
    return nph_absf(tb) - (ra + x->hb[j]);
}

// Separation along A_i x B_j (unit, from A towards B); -FLT_MAX when the
// axes are parallel.
NPH_TD f32 box_axis_cross(const NphBoxFrame *x, u32 i, u32 j, NphVec3 *axis)
{
    const u32 i1 = i == 2U ? 0U : i + 1U;
    const u32 i2 = i == 0U ? 2U : i - 1U;
    const u32 j1 = j == 2U ? 0U : j + 1U;
    const u32 j2 = j == 0U ? 2U : j - 1U;
    const f32 len_sq = 1.0F - x->r[i][j] * x->r[i][j];
    if (len_sq < 1e-6F) {
        return -NPH_F32_MAX;
    }
    const f32 ra = x->ha[i1] * x->ar[i2][j] + x->ha[i2] * x->ar[i1][j];
    const f32 rb = x->hb[j1] * x->ar[i][j2] + x->hb[j2] * x->ar[i][j1];
    const f32 tl = x->t[i2] * x->r[i1][j] - x->t[i1] * x->r[i2][j];
    const f32 inv = 1.0F / nph_sqrtf(len_sq);
    f32 l[3];
    l[i] = 0.0F;
    l[i1] = -x->r[i2][j];
    l[i2] = x->r[i1][j];
    const f32 s = tl >= 0.0F ? inv : -inv;
    *axis = v3(l[0] * s, l[1] * s, l[2] * s);
    return (nph_absf(tl) - (ra + rb)) * inv;
}

template <u32 T, u32 V, u32 CAP>
NPH_TD void team_boxes(const NphTeam<T> &tm, NphPolyScratch<V, CAP> *sc, NphVec3 half_a,
                       NphPhysicsPose pa, NphVec3 half_b, NphPhysicsPose pb, f32 speculative,
                       NphCache *cache, NphManifold *out)
{
    const NphPhysicsPose rel = pose_inv_mul(pa, pb);
    const NphMat3 rm = m3_from_quat(rel.rotation);
    NphBoxFrame x;
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
    // The cache holds face 2i + side, or the edge axes i and j.
    if (cache != 0 && cache->type != NPH_CACHE_NONE) {
        f32 side;
        NphVec3 axis;
        f32 s = -NPH_F32_MAX;
        if (cache->type == NPH_CACHE_FACE_A && cache->index_a < 6U) {
            s = box_axis_a(&x, cache->index_a / 2U, &side);
        } else if (cache->type == NPH_CACHE_FACE_B && cache->index_b < 6U) {
            s = box_axis_b(&x, cache->index_b / 2U, &side);
        } else if (cache->type == NPH_CACHE_EDGES && cache->index_a < 3U && cache->index_b < 3U) {
            s = box_axis_cross(&x, cache->index_a, cache->index_b, &axis);
        }
        if (s > speculative) {
            return;
        }
    }
    f32 face_a_sep = -NPH_F32_MAX, face_b_sep = -NPH_F32_MAX, edge_sep = -NPH_F32_MAX;
    u32 face_a = 0U, face_b = 0U, edge_i = 0U, edge_j = 0U;
    NphVec3 edge_axis = v3(0.0F, 1.0F, 0.0F);
    for (u32 i = 0U; i < 3U; ++i) {
        f32 side;
        const f32 s = box_axis_a(&x, i, &side);
        if (s > face_a_sep) {
            face_a_sep = s;
            face_a = 2U * i + (side > 0.0F ? 1U : 0U);
        }
    }
    for (u32 j = 0U; j < 3U; ++j) {
        f32 side;
        const f32 s = box_axis_b(&x, j, &side);
        if (s > face_b_sep) {
            face_b_sep = s;
            face_b = 2U * j + (side > 0.0F ? 0U : 1U); // A is on B's -B_j side
        }
    }
    for (u32 i = 0U; i < 3U; ++i) {
        for (u32 j = 0U; j < 3U; ++j) {
            NphVec3 axis;
            const f32 s = box_axis_cross(&x, i, j, &axis);
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
    const int apart = nph_maxf(face_sep, edge_sep) > speculative;
    if (cache != 0) {
        // A separated pair keeps the axis that separates it most.
        if (apart) {
            const int by_edge = edge_sep > face_sep;
            const int by_b = !by_edge && face_b_sep > face_a_sep;
            cache->type = (u8)(by_edge ? NPH_CACHE_EDGES : (by_b ? NPH_CACHE_FACE_B : NPH_CACHE_FACE_A));
            cache->index_a = (u16)(by_edge ? edge_i : face_a);
            cache->index_b = (u16)(by_edge ? edge_j : face_b);
        } else {
            cache->type = NPH_CACHE_NONE;
        }
    }
    if (apart) {
        return;
    }
    if (tm.lane == 0U) {
        box_fill(half_a, &sc->box[0]);
        box_fill(half_b, &sc->box[1]);
    }
    tm.sync();
    const NphView va = box_view_at(&sc->box[0]);
    const NphView vb = box_view_at(&sc->box[1]);
    NphPolyPair pp;
    pp.a = &va;
    pp.b = &vb;
    pp.rel = rel;
    place_b(tm, pp, sc);
    if (best > NPH_PHYSICS_LINEAR_SLOP) {
        NphDistance gd;
        team_distance(tm, va.vertices, 8U, sc->bv, 8U, &gd);
        if (gd.distance > speculative) {
            return;
        }
        const NphVec3 sat_axis = use_edge ? edge_axis
                                          : (flip ? v3_neg(face_b_plane(pp, face_b).normal)
                                                  : va.planes[face_a].normal);
        if (gd.distance > 0.0F &&
            (use_edge || v3_dot(gd.normal, sat_axis) < NPH_FEATURE_AGREEMENT)) {
            out->normal = pose_vector(pa, gd.normal);
            manifold_push(out, pose_point(pa, v3_scale(v3_add(gd.point_a, gd.point_b), 0.5F)),
                          gd.distance, NPH_ID_GJK_POINT);
            return;
        }
    }
    if (use_edge) {
        // The supporting edges: A's along A_i at its corner farthest along
        // the axis, B's along B_j at its corner farthest against it.
        NphVec3 ca = v3_zero();
        NphVec3 ea = v3_zero();
        for (u32 k = 0U; k < 3U; ++k) {
            f32 *cak = k == 0U ? &ca.x : (k == 1U ? &ca.y : &ca.z);
            f32 *eak = k == 0U ? &ea.x : (k == 1U ? &ea.y : &ea.z);
            if (k != edge_i) {
                *cak = v3_at(edge_axis, k) >= 0.0F ? x.ha[k] : -x.ha[k];
            } else {
                *eak = x.ha[k];
            }
        }
        NphVec3 cb = rel.position;
        for (u32 m = 0U; m < 3U; ++m) {
            if (m != edge_j) {
                const f32 along = v3_dot(rm.c[m], edge_axis);
                cb = v3_mul_add(cb, rm.c[m], along >= 0.0F ? -x.hb[m] : x.hb[m]);
            }
        }
        const NphVec3 eb = v3_scale(rm.c[edge_j], x.hb[edge_j]);
        const NphSegDist sd =
            segment_distance(v3_sub(ca, ea), v3_add(ca, ea), v3_sub(cb, eb), v3_add(cb, eb));
        const f32 sep = v3_dot(v3_sub(sd.closest2, sd.closest1), edge_axis);
        if (sep > speculative) {
            return;
        }
        out->normal = pose_vector(pa, edge_axis);
        manifold_push(out, pose_point(pa, v3_scale(v3_add(sd.closest1, sd.closest2), 0.5F)), sep,
                      nph_mix_id(NPH_ID_EDGE_PAIR | edge_i, edge_j));
        return;
    }
    NphMPoint pts[NPH_TRI_POINTS];
    u32 count = 0U;
    const NphVec3 n = team_face_contact(tm, pp, sc, flip, flip ? face_b : face_a, speculative,
                                        NPH_PHYSICS_MANIFOLD_MAX_POINTS, pts, &count);
    out->normal = pose_vector(pa, n);
    for (u32 i = 0U; i < count; ++i) {
        manifold_push(out, pose_point(pa, pts[i].point), pts[i].separation, pts[i].id);
    }
}

// ---------------------------------------------------------------------------
// Dispatch (collide.c:474-540): every convex pair, closed forms included
// ---------------------------------------------------------------------------

// A's view: its hull block, or its box in the team's scratch.
template <u32 T, u32 V, u32 CAP>
NPH_TD NphView team_view(const NphTeam<T> &tm, NphPolyScratch<V, CAP> *sc, const NphGeom *g, u32 slot)
{
    if (g->type == NPH_PHYSICS_SHAPE_BOX) {
        if (tm.lane == 0U) {
            box_fill(v3(g->a, g->b, g->c), &sc->box[slot]);
        }
        tm.sync();
        return box_view_at(&sc->box[slot]);
    }
    return hull_view(g->hull);
}

// The pair with a's type at least b's.
template <u32 T, u32 V, u32 CAP>
NPH_TD void team_collide_ordered(const NphTeam<T> &tm, NphPolyScratch<V, CAP> *sc, const NphGeom *a,
                                 NphPhysicsPose pa, const NphGeom *b, NphPhysicsPose pb,
                                 f32 speculative, NphCache *cache, NphManifold *out)
{
    if (closed_form_ordered(a, pa, b, pb, speculative, out)) {
        return;
    }
    switch (a->type) {
    case NPH_PHYSICS_SHAPE_BOX:
    case NPH_PHYSICS_SHAPE_HULL: {
        if (a->type == NPH_PHYSICS_SHAPE_BOX && b->type == NPH_PHYSICS_SHAPE_BOX) {
            team_boxes(tm, sc, v3(a->a, a->b, a->c), pa, v3(b->a, b->b, b->c), pb, speculative,
                       cache, out);
            return;
        }
        const NphView va = team_view(tm, sc, a, 0U);
        if (b->type == NPH_PHYSICS_SHAPE_SPHERE) {
            team_hull_sphere(tm, &va, pa, b->a, pb, speculative, out);
        } else if (b->type == NPH_PHYSICS_SHAPE_CAPSULE) {
            team_hull_capsule(tm, &va, pa, b->a, b->b, pb, speculative, out);
        } else {
            const NphView vb = team_view(tm, sc, b, 1U);
            NphTriContact c;
            tri_clear(&c);
            team_polytopes(tm, sc, &va, pa, &vb, pb, speculative, cache, 0,
                           NPH_PHYSICS_MANIFOLD_MAX_POINTS, &c);
            if (c.point_count > 0U) {
                out->normal = c.normal;
                for (u32 i = 0U; i < c.point_count; ++i) {
                    manifold_push(out, c.points[i].point, c.points[i].separation, c.points[i].id);
                }
            }
        }
        return;
    }
    default:
        return;
    }
}

// nph_physics_collide: the manifold of two convex shapes, A's normal towards
// B; the other order of types gives the same points, the normal reversed.
template <u32 T, u32 V, u32 CAP>
NPH_TD void team_collide(const NphTeam<T> &tm, NphPolyScratch<V, CAP> *sc, const NphGeom *a,
                         NphPhysicsPose pa, const NphGeom *b, NphPhysicsPose pb,
                         f32 speculative_distance, NphCache *cache, NphManifold *out)
{
    manifold_clear(out);
    const f32 speculative = nph_maxf(speculative_distance, 0.0F);
    if (a->type >= b->type) {
        team_collide_ordered(tm, sc, a, pa, b, pb, speculative, cache, out);
        return;
    }
    team_collide_ordered(tm, sc, b, pb, a, pa, speculative, cache, out);
    out->normal = v3_neg(out->normal);
}

// ---------------------------------------------------------------------------
// The lane pass's share of the polytope pairs
// ---------------------------------------------------------------------------

// Whether a convex pair runs on a team: a hull against a box or a hull (the
// pairs whose edge-pair search is long). Boxes against boxes and capsules,
// and hulls against spheres and capsules, are a lane's work.
NPH_TD int pair_wants_team(u32 ta, u32 tb)
{
    const int hull_a = ta == NPH_PHYSICS_SHAPE_HULL, hull_b = tb == NPH_PHYSICS_SHAPE_HULL;
    const int poly_a = hull_a || ta == NPH_PHYSICS_SHAPE_BOX;
    const int poly_b = hull_b || tb == NPH_PHYSICS_SHAPE_BOX;
    return (hull_a || hull_b) && poly_a && poly_b;
}

// For a pair that runs on a team: whether the axis in the cache still
// separates it by more than the speculative distance, which ends the pair
// with no points and the cache as it is (collide_hull.c:571-594). The same
// arithmetic as team_sat's first test, one lane.
NPH_TD int cached_axis_separates(const NphGeom *ga, NphPhysicsPose pose_a, const NphGeom *gb,
                                 NphPhysicsPose pose_b, f32 speculative_distance,
                                 const NphCache *cache)
{
    if (cache->type == NPH_CACHE_NONE) {
        return 0;
    }
    // The team orders the pair by type (hulls first).
    const int swap = ga->type < gb->type;
    const NphGeom *a = swap ? gb : ga;
    const NphGeom *b = swap ? ga : gb;
    const NphPhysicsPose pa = swap ? pose_b : pose_a;
    const NphPhysicsPose pb = swap ? pose_a : pose_b;
    const f32 limit = nph_maxf(speculative_distance, 0.0F);
    NphBoxStorage box;
    const NphView va = a->type == NPH_PHYSICS_SHAPE_BOX ? box_view(v3(a->a, a->b, a->c), &box)
                                                        : hull_view(a->hull);
    NphBoxStorage boxb;
    const NphView vb = b->type == NPH_PHYSICS_SHAPE_BOX ? box_view(v3(b->a, b->b, b->c), &boxb)
                                                        : hull_view(b->hull);
    NphPolyPair pp;
    pp.a = &va;
    pp.b = &vb;
    pp.rel = pose_inv_mul(pa, pb);
    f32 s = -NPH_F32_MAX;
    if (cache->type == NPH_CACHE_FACE_A && cache->index_a < va.face_count) {
        const NphPlane pl = va.planes[cache->index_a];
        f32 m = v3_dot(pose_point(pp.rel, vb.vertices[0]), pl.normal);
        for (u32 i = 1U; i < vb.vertex_count; ++i) {
            m = nph_minf(m, v3_dot(pose_point(pp.rel, vb.vertices[i]), pl.normal));
        }
        s = m + pl.d;
    } else if (cache->type == NPH_CACHE_FACE_B && cache->index_b < vb.face_count) {
        const NphPlane pl = face_b_plane(pp, cache->index_b);
        s = lowest_projection(va.vertices, va.vertex_count, pl.normal) + pl.d;
    } else if (cache->type == NPH_CACHE_EDGES && cache->index_a < va.edge_count &&
               cache->index_b < vb.edge_count) {
        const u32 ea = cache->index_a;
        const u32 eb = cache->index_b;
        const NphVec3 p_a = va.vertices[va.edges[ea].origin];
        const NphVec3 p_b = pose_point(pp.rel, vb.vertices[vb.edges[eb].origin]);
        const NphVec3 h_b = pose_point(pp.rel, vb.vertices[edge_end(&vb, eb)]);
        NphVec3 axis;
        s = edge_axis_separation(p_a, v3_sub(va.vertices[edge_end(&va, ea)], p_a), p_b,
                                 v3_sub(h_b, p_b), va.centroid, &axis);
    }
    return s > limit;
}

// nph_physics_collide for the pairs the lane pass keeps: closed forms,
// boxes against boxes and capsules, hulls against spheres and capsules.
// The scratch holds two box views and one box's vertices and clip polygons.
typedef NphPolyScratch<8U, 12U> NphLaneScratch;

NPH_TD void lane_collide(const NphGeom *a, NphPhysicsPose pa, const NphGeom *b, NphPhysicsPose pb,
                         f32 speculative_distance, NphCache *cache, NphLaneScratch *sc,
                         NphManifold *out)
{
    const NphTeam<1U> one;
    manifold_clear(out);
    const f32 speculative = nph_maxf(speculative_distance, 0.0F);
    const int swap = a->type < b->type; // as team_collide: the higher type first
    const NphGeom *x = swap ? b : a;
    const NphGeom *y = swap ? a : b;
    const NphPhysicsPose px = swap ? pb : pa;
    const NphPhysicsPose py = swap ? pa : pb;
    if (!closed_form_ordered(x, px, y, py, speculative, out)) {
        if (x->type == NPH_PHYSICS_SHAPE_BOX && y->type == NPH_PHYSICS_SHAPE_BOX) {
            team_boxes(one, sc, v3(x->a, x->b, x->c), px, v3(y->a, y->b, y->c), py, speculative,
                       cache, out);
        } else {
            const NphView vx = team_view(one, sc, x, 0U);
            if (y->type == NPH_PHYSICS_SHAPE_SPHERE) {
                team_hull_sphere(one, &vx, px, y->a, py, speculative, out);
            } else {
                team_hull_capsule(one, &vx, px, y->a, y->b, py, speculative, out);
            }
        }
    }
    if (swap) {
        out->normal = v3_neg(out->normal);
    }
}

// One lane, its own scratch (private memory).
NPH_TD void nph_collide(const NphGeom *a, NphPhysicsPose pa, const NphGeom *b, NphPhysicsPose pb,
                        f32 speculative, NphCache *cache, NphManifold *out)
{
    NphTeam<1U> one;
    NphPolyScratch<NPH_HULL_MAX_VERTICES, NPH_CLIP_CAPACITY> sc;
    team_collide(one, &sc, a, pa, b, pb, speculative, cache, out);
}

// The mesh path's one-lane entry points (a shape against one triangle view).
NPH_TD void collide_hulls_into(const NphView *a, NphPhysicsPose pa, const NphView *b,
                               NphPhysicsPose pb, f32 speculative, NphCache *cache, int prefer_b,
                               u32 max_points, NphTriContact *out)
{
    NphTeam<1U> one;
    NphPolyScratch<NPH_HULL_MAX_VERTICES, NPH_CLIP_CAPACITY> sc;
    team_polytopes(one, &sc, a, pa, b, pb, speculative, cache, prefer_b, max_points, out);
}

NPH_TD void collide_hull_face(const NphView *a, NphPhysicsPose pa, const NphView *b,
                              NphPhysicsPose pb, u32 face, f32 speculative, NphTriContact *out)
{
    NphTeam<1U> one;
    NphPolyScratch<NPH_HULL_MAX_VERTICES, NPH_CLIP_CAPACITY> sc;
    team_hull_face(one, &sc, a, pa, b, pb, face, speculative, out);
}

#endif // __cplusplus
#endif // NPH_POLYTOPE_H
