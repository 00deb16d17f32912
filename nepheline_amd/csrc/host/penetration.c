// SPDX-License-Identifier: Apache-2.0
// How deep two rounded convex proxies overlap (collide.h:223-236;
// gjk.c:419-701): GJK on the cores first (cast.c's proxy_gjk); when the
// cores are apart only the rounding can overlap and their closest points
// give the answer; when the cores overlap, an expanding polytope (van den
// Bergen, "Proximity Queries and Penetration Depth Computation on 3D Game
// Objects", GDC 2001) grows over the Minkowski difference A - B from GJK's
// final simplex until its facet nearest the origin lies on the boundary.
//
// The polytope keeps its facets in one array in the order they were made.
// A step takes the nearest facet (the first of equals), asks both proxies
// for the support along its normal, and stops when the new point gains
// less than max(1e-5, 1e-4 |distance|) on it. Otherwise every facet that
// sees the point leaves (the rest keep their order), the rim those facets
// leave open is collected as directed edges (an edge met again reversed
// lies between two leaving facets and is dropped by moving the rim's last
// edge into its place), and one facet per rim edge closes the polytope
// over the point. The reference's capacities stay: 72 points, 144 facets,
// a 96-edge rim, 64 steps; a rim that cannot be closed ends the search on
// the facet the step began from.
#include "world_impl.h"

#include <float.h>

#define HULL_POINTS 72U
#define HULL_FACETS 144U
#define RIM_EDGES 96U
#define EXPAND_STEPS 64U
#define GROW_ATTEMPTS 16U
#define FLAT 1e-6F // how far a new point must stand off a flat simplex

typedef struct Pair {
    const NphProxy *a;
    NphPhysicsPose pa;
    const NphProxy *b;
    NphPhysicsPose pb;
} Pair;

// The point of A - B farthest along d: A's support along d, B's against it.
static NphGjkVertex pair_support(const Pair *p, NphVec3 d)
{
    const u32 ia = points_support(p->a->points, p->a->count, q_rotate_inv(p->pa.rotation, d));
    const u32 ib = points_support(p->b->points, p->b->count, q_rotate_inv(p->pb.rotation, v3_neg(d)));
    return gjk_vertex(ia, pose_point(p->pa, p->a->points[ia]), ib, pose_point(p->pb, p->b->points[ib]));
}

// ---------------------------------------------------------------------------
// A simplex through the origin grown to a tetrahedron (gjk.c:494-550)
// ---------------------------------------------------------------------------

// Duff, Burgess, Christensen, Hery, Kensler, Liani and Villemin, "Building
// an Orthonormal Basis, Revisited" (JCGT 2017): two unit vectors across n.
static void across(NphVec3 n, NphVec3 *t, NphVec3 *u)
{
    const f32 s = copysignf(1.0F, n.z);
    const f32 k = -1.0F / (s + n.z);
    const f32 m = n.x * n.y * k;
    *t = v3(1.0F + s * n.x * n.x * k, s * m, -s * n.x);
    *u = v3(m, s + n.y * n.y * k, -n.y);
}

// Where to look for a point off the simplex's span: the six axis directions
// about a point, four across a segment, both sides of a triangle.
static u32 off_span_directions(const NphGjk *g, NphVec3 d[6])
{
    const NphVec3 o = g->v[0].w;
    if (g->count == 1U) {
        d[0] = v3(1.0F, 0.0F, 0.0F);
        d[1] = v3(-1.0F, 0.0F, 0.0F);
        d[2] = v3(0.0F, 1.0F, 0.0F);
        d[3] = v3(0.0F, -1.0F, 0.0F);
        d[4] = v3(0.0F, 0.0F, 1.0F);
        d[5] = v3(0.0F, 0.0F, -1.0F);
        return 6U;
    }
    if (g->count == 2U) {
        NphVec3 t, u;
        across(v3_normalize_or(v3_sub(g->v[1].w, o), v3(1.0F, 0.0F, 0.0F)), &t, &u);
        d[0] = t;
        d[1] = v3_neg(t);
        d[2] = u;
        d[3] = v3_neg(u);
        return 4U;
    }
    d[0] = v3_cross(v3_sub(g->v[1].w, o), v3_sub(g->v[2].w, o));
    d[1] = v3_neg(d[0]);
    return 2U;
}

// Whether w stands off the simplex's span: apart from its point, off its
// segment's line, off its triangle's plane.
static int off_span(const NphGjk *g, NphVec3 w)
{
    const NphVec3 o = g->v[0].w;
    const NphVec3 rel = v3_sub(w, o);
    if (g->count == 1U) {
        return v3_len_sq(rel) > FLAT * FLAT;
    }
    if (g->count == 2U) {
        const NphVec3 e = v3_sub(g->v[1].w, o);
        return v3_len_sq(v3_cross(e, rel)) > FLAT * FLAT * v3_len_sq(e);
    }
    const NphVec3 n =
        v3_normalize_or(v3_cross(v3_sub(g->v[1].w, o), v3_sub(g->v[2].w, o)), v3_zero());
    return nph_absf(v3_dot(n, rel)) > FLAT;
}

// GJK can end on a point, segment or triangle when the cores only touch;
// a point off its span is added until four stand with volume. Fails on a
// flat Minkowski difference.
static int grow_tetrahedron(const Pair *p, NphGjk *g)
{
    for (u32 attempt = 0U; attempt < GROW_ATTEMPTS && g->count < 4U; ++attempt) {
        NphVec3 d[6];
        const u32 n = off_span_directions(g, d);
        u32 k = 0U;
        while (k < n) {
            const NphGjkVertex w = pair_support(p, d[k]);
            if (off_span(g, w.w)) {
                g->v[g->count] = w;
                g->count += 1U;
                break;
            }
            k += 1U;
        }
        if (k == n) {
            return 0;
        }
    }
    return g->count == 4U;
}

// ---------------------------------------------------------------------------
// The expanding polytope (gjk.c:419-657)
// ---------------------------------------------------------------------------

typedef struct Facet {
    NphVec3 normal; // unit, outwards
    f32 distance;   // of its plane from the origin
    u8 corner[3];   // counter-clockwise seen from outside
} Facet;

typedef struct Polytope {
    NphGjkVertex point[HULL_POINTS];
    Facet facet[HULL_FACETS];
    u8 rim[RIM_EDGES][2];
    u32 points;
    u32 facets;
    u32 rim_edges;
    NphVec3 centre; // inside the first tetrahedron, to orient its facets
} Polytope;

// Appends facet (a, b, c), turned to face away from the centre; fails when
// the facets are full or the triangle has no area.
static int facet_add(Polytope *t, u32 a, u32 b, u32 c)
{
    if (t->facets >= HULL_FACETS) {
        return 0;
    }
    const NphVec3 pa = t->point[a].w;
    NphVec3 n = v3_cross(v3_sub(t->point[b].w, pa), v3_sub(t->point[c].w, pa));
    const f32 len = v3_len(n);
    if (!(len > 1e-12F)) {
        return 0;
    }
    n = v3_scale(n, 1.0F / len);
    Facet *f = &t->facet[t->facets];
    t->facets += 1U;
    const int flip = v3_dot(n, v3_sub(pa, t->centre)) < 0.0F;
    f->corner[0] = (u8)a;
    f->corner[1] = (u8)(flip ? c : b);
    f->corner[2] = (u8)(flip ? b : c);
    f->normal = flip ? v3_neg(n) : n;
    f->distance = v3_dot(f->normal, pa);
    return 1;
}

// Adds the directed edge (a, b) to the rim, or drops its reverse.
static void rim_toggle(Polytope *t, u32 a, u32 b)
{
    for (u32 i = 0U; i < t->rim_edges; ++i) {
        if (t->rim[i][0] == b && t->rim[i][1] == a) {
            t->rim_edges -= 1U;
            t->rim[i][0] = t->rim[t->rim_edges][0];
            t->rim[i][1] = t->rim[t->rim_edges][1];
            return;
        }
    }
    if (t->rim_edges < RIM_EDGES) {
        t->rim[t->rim_edges][0] = (u8)a;
        t->rim[t->rim_edges][1] = (u8)b;
        t->rim_edges += 1U;
    }
}

// The first facet of least distance; HULL_FACETS when none is below
// FLT_MAX (every distance infinite or NaN).
static u32 nearest_facet(const Polytope *t)
{
    u32 best = HULL_FACETS;
    f32 least = FLT_MAX;
    for (u32 f = 0U; f < t->facets; ++f) {
        if (t->facet[f].distance < least) {
            least = t->facet[f].distance;
            best = f;
        }
    }
    return best;
}

// Point w joins: the facets that see it leave, the rest keep their order,
// and the rim they leave is closed over w. False when it cannot be.
static int add_point(Polytope *t, NphGjkVertex w)
{
    const u32 wi = t->points;
    t->point[wi] = w;
    t->points += 1U;
    t->rim_edges = 0U;
    u32 kept = 0U;
    for (u32 f = 0U; f < t->facets; ++f) {
        const Facet x = t->facet[f];
        if (v3_dot(x.normal, v3_sub(w.w, t->point[x.corner[0]].w)) > 0.0F) {
            rim_toggle(t, x.corner[0], x.corner[1]);
            rim_toggle(t, x.corner[1], x.corner[2]);
            rim_toggle(t, x.corner[2], x.corner[0]);
        } else {
            t->facet[kept] = x;
            kept += 1U;
        }
    }
    t->facets = kept;
    if (t->rim_edges == 0U) {
        return 0;
    }
    for (u32 k = 0U; k < t->rim_edges; ++k) {
        if (!facet_add(t, t->rim[k][0], t->rim[k][1], wi)) {
            return 0;
        }
    }
    return 1;
}

typedef struct Depth {
    NphVec3 normal;
    f32 depth;
    NphVec3 point_a;
    NphVec3 point_b;
} Depth;

// The origin's foot on facet f, as weights of its corners, carried to both
// proxies' points.
static void facet_witness(const Polytope *t, const Facet *f, Depth *out)
{
    const NphGjkVertex *c0 = &t->point[f->corner[0]];
    const NphGjkVertex *c1 = &t->point[f->corner[1]];
    const NphGjkVertex *c2 = &t->point[f->corner[2]];
    const NphVec3 foot = v3_scale(f->normal, f->distance);
    const NphVec3 e1 = v3_sub(c1->w, c0->w);
    const NphVec3 e2 = v3_sub(c2->w, c0->w);
    const NphVec3 n = v3_cross(e1, e2);
    const f32 nn = v3_dot(n, n);
    f32 w0 = 1.0F, w1 = 0.0F, w2 = 0.0F;
    if (nn > 0.0F) {
        const NphVec3 r = v3_sub(foot, c0->w);
        w1 = v3_dot(v3_cross(r, e2), n) / nn;
        w2 = v3_dot(v3_cross(e1, r), n) / nn;
        w0 = 1.0F - w1 - w2;
    }
    out->point_a = v3_add(v3_add(v3_scale(c0->a, w0), v3_scale(c1->a, w1)), v3_scale(c2->a, w2));
    out->point_b = v3_add(v3_add(v3_scale(c0->b, w0), v3_scale(c1->b, w1)), v3_scale(c2->b, w2));
    out->normal = f->normal;
    out->depth = nph_maxf(f->distance, 0.0F);
}

static int expand(const Pair *p, NphGjk *g, Depth *out)
{
    if (!grow_tetrahedron(p, g)) {
        return 0;
    }
    Polytope t;
    t.points = 4U;
    t.facets = 0U;
    t.rim_edges = 0U;
    t.centre = v3_zero();
    for (u32 i = 0U; i < 4U; ++i) {
        t.point[i] = g->v[i];
        t.centre = v3_add(t.centre, g->v[i].w);
    }
    t.centre = v3_scale(t.centre, 0.25F);
    if (!facet_add(&t, 0U, 1U, 2U) || !facet_add(&t, 0U, 3U, 1U) || !facet_add(&t, 0U, 2U, 3U) ||
        !facet_add(&t, 1U, 3U, 2U)) {
        return 0;
    }
    // After a step that closes its rim the chosen slot is kept: when the
    // steps run out, the answer is the facet now in that slot.
    u32 best = 0U;
    for (u32 step = 0U; step < EXPAND_STEPS; ++step) {
        best = nearest_facet(&t);
        if (best == HULL_FACETS) {
            return 0;
        }
        const Facet near = t.facet[best];
        const NphGjkVertex w = pair_support(p, near.normal);
        const f32 gain = v3_dot(w.w, near.normal) - near.distance;
        if (gain <= nph_maxf(1e-5F, 1e-4F * nph_absf(near.distance)) || t.points >= HULL_POINTS) {
            break;
        }
        if (!add_point(&t, w)) {
            t.facet[0] = near;
            t.facets = 1U;
            best = 0U;
            break;
        }
    }
    facet_witness(&t, &t.facet[best], out);
    return 1;
}

bool nph_physics_penetration(const NphPhysicsProxy *a, NphPhysicsPose pose_a,
                             const NphPhysicsProxy *b, NphPhysicsPose pose_b,
                             NphPhysicsPenetration *out)
{
    const NphProxy core_a = {a->points, a->count, a->radius};
    const NphProxy core_b = {b->points, b->count, b->radius};
    const Pair pair = {&core_a, pose_a, &core_b, pose_b};
    NphGjk g;
    // Weights the search never writes read as zero, should a flat
    // difference leave grown points in the simplex.
    memset(g.lambda, 0, sizeof g.lambda);
    proxy_gjk(&core_a, pose_a, &core_b, pose_b, NULL, &g);
    NphDistance near;
    gjk_finish(&g, &near);
    const f32 reach = a->radius + b->radius;
    Depth d;
    if (!g.inside) {
        if (near.distance >= reach || !(near.distance > 0.0F)) {
            return false;
        }
        // Only the rounding overlaps.
        out->normal = near.normal;
        out->depth = reach - near.distance;
        out->point_a = v3_mul_add(near.point_a, near.normal, a->radius);
        out->point_b = v3_mul_add(near.point_b, near.normal, -b->radius);
        return true;
    }
    if (!expand(&pair, &g, &d)) {
        // A difference without volume (flat shapes in one plane): a touch
        // along +y at the simplex's points.
        gjk_finish(&g, &near);
        d.point_a = near.point_a;
        d.point_b = near.point_b;
        d.normal = v3(0.0F, 1.0F, 0.0F);
        d.depth = 0.0F;
    }
    // The cores' depth, then each radius in turn.
    out->normal = d.normal;
    out->depth = d.depth + a->radius + b->radius;
    out->point_a = v3_mul_add(d.point_a, d.normal, a->radius);
    out->point_b = v3_mul_add(d.point_b, d.normal, -b->radius);
    return true;
}
