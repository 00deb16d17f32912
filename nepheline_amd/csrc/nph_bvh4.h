// SPDX-License-Identifier: Apache-2.0
// The mesh block's four-wide box tree (mesh.h:99-150) and the two walks over
// it, the same code on the host (mesh queries, world queries) and on the
// device (the narrow phase against mesh triangles):
//   - the box walk yields the triangles whose boxes overlap a query box in
//     tree order: depth first, slots in order, a leaf's triangles in order.
//     That order fixes the order of visits and of contact clusters
//     (mesh.c:998-1050).
//   - the segment walk yields the triangles of the leaves a segment (or a
//     box swept along it) reaches, nearest child box first, skipping boxes
//     it enters beyond the best fraction the caller has found so far
//     (mesh.c:1052-1274).
// Both keep their pending child codes in a stack of NPH_BVH4_STACK entries.
// A tree the builder writes needs at most 3 d + 4 of them at depth d <= 46,
// and the validator refuses any tree that could need more; a walk over a
// block that needs more stops and raises its `overflow` flag rather than
// run past its stack.
//
// The ray against a triangle is the watertight test of Woop, Benthin and
// Wald ("Watertight Ray/Triangle Intersection", JCGT 2(1), 2013): shear the
// triangle so the ray runs along its dominant axis, take the three edge
// functions in the plane across it, decide a zero one in f64 (so two
// triangles sharing an edge never both miss), then the fraction from the
// scaled heights. The arithmetic keeps the reference's operation order
// (physics_internal.h:313-372), whose fractions the queries report.
#ifndef NPH_BVH4_H
#define NPH_BVH4_H

#include "nph_math.h"

#define NPH_BVH4_STACK 160U

// ---------------------------------------------------------------------------
// The block
// ---------------------------------------------------------------------------
NPH_HD const NphPhysicsMeshNode *mesh_nodes(const NphPhysicsMesh *m)
{
    return (const NphPhysicsMeshNode *)((const u8 *)m + m->nodes_offset);
}
NPH_HD const NphVec3 *mesh_vertices(const NphPhysicsMesh *m)
{
    return (const NphVec3 *)((const u8 *)m + m->vertices_offset);
}
NPH_HD const u32 *mesh_triangles(const NphPhysicsMesh *m)
{
    return (const u32 *)((const u8 *)m + m->triangles_offset);
}
NPH_HD const u8 *mesh_edges(const NphPhysicsMesh *m)
{
    return (const u8 *)m + m->edges_offset;
}
NPH_HD const u32 *mesh_originals(const NphPhysicsMesh *m)
{
    return (const u32 *)((const u8 *)m + m->originals_offset);
}
// A child code: a node index, a leaf (the top bit, then the first triangle
// and the count less one in the low two bits), or the empty slot.
NPH_HD int mesh_child_is_leaf(u32 child)
{
    return child != NPH_PHYSICS_MESH_CHILD_EMPTY && (child & NPH_PHYSICS_MESH_CHILD_LEAF) != 0U;
}
NPH_HD u32 mesh_leaf_code(u32 first, u32 count)
{
    return NPH_PHYSICS_MESH_CHILD_LEAF | (first << 2U) | (count - 1U);
}
NPH_HD u32 mesh_leaf_first(u32 child)
{
    return (child & ~NPH_PHYSICS_MESH_CHILD_LEAF) >> 2U;
}
NPH_HD u32 mesh_leaf_count(u32 child)
{
    return (child & 3U) + 1U;
}
NPH_HD NphAABB bvh4_slot_box(const NphPhysicsMeshNode *n, u32 k)
{
    return aabb_make(v3(n->min_x[k], n->min_y[k], n->min_z[k]),
                     v3(n->max_x[k], n->max_y[k], n->max_z[k]));
}
NPH_HD void mesh_corners(const NphPhysicsMesh *m, u32 t, NphVec3 out[3])
{
    const NphVec3 *v = mesh_vertices(m);
    const u32 *tri = mesh_triangles(m) + 3U * t;
    out[0] = v[tri[0]];
    out[1] = v[tri[1]];
    out[2] = v[tri[2]];
}
NPH_HD NphAABB corners_box(const NphVec3 p[3])
{
    return aabb_make(v3_min(v3_min(p[0], p[1]), p[2]), v3_max(v3_max(p[0], p[1]), p[2]));
}

// ---------------------------------------------------------------------------
// The box walk
// ---------------------------------------------------------------------------
typedef struct NphBoxWalk {
    const NphPhysicsMesh *mesh;
    NphAABB box;
    u32 next; // the open leaf's next triangle
    u32 end;  // one past its last; next == end when no leaf is open
    u32 top;
    u32 overflow;
    u32 stack[NPH_BVH4_STACK];
} NphBoxWalk;

NPH_HD void box_walk_begin(NphBoxWalk *w, const NphPhysicsMesh *m, NphAABB box)
{
    w->mesh = m;
    w->box = box;
    w->next = 0U;
    w->end = 0U;
    w->top = 1U;
    w->overflow = 0U;
    w->stack[0] = 0U;
}

// The next triangle whose box overlaps the query box; NPH_NULL32 when the
// walk is over.
NPH_HD u32 box_walk_next(NphBoxWalk *w)
{
    for (;;) {
        while (w->next < w->end) {
            const u32 t = w->next++;
            NphVec3 p[3];
            mesh_corners(w->mesh, t, p);
            if (aabb_overlaps(corners_box(p), w->box)) {
                return t;
            }
        }
        if (w->top == 0U) {
            return NPH_NULL32;
        }
        const u32 code = w->stack[--w->top];
        if (mesh_child_is_leaf(code)) {
            w->next = mesh_leaf_first(code);
            w->end = w->next + mesh_leaf_count(code);
            continue;
        }
        const NphPhysicsMeshNode *n = &mesh_nodes(w->mesh)[code];
        // The last slot deepest, so the slots come off in order.
        for (u32 k = 4U; k-- > 0U;) {
            if (n->child[k] == NPH_PHYSICS_MESH_CHILD_EMPTY ||
                !aabb_overlaps(bvh4_slot_box(n, k), w->box)) {
                continue;
            }
            if (w->top == NPH_BVH4_STACK) {
                w->overflow = 1U;
                w->top = 0U;
                return NPH_NULL32;
            }
            w->stack[w->top++] = n->child[k];
        }
    }
}

// ---------------------------------------------------------------------------
// The segment walk
// ---------------------------------------------------------------------------
// 1 / d, but +-1e30 for |d| < 1e-30: a segment parallel to an axis then
// gets huge slab fractions of the right sign instead of 0 * inf.
NPH_HD f32 bvh4_recip(f32 d)
{
    if (nph_absf(d) < 1e-30F) {
        return d < 0.0F ? -1e30F : 1e30F;
    }
    return 1.0F / d;
}

// Where the segment o + t d, t in [0, limit], enters slot k's box grown by
// `grow` on every side (slab test, Kay and Kajiya 1986); NPH_F32_MAX for a
// miss. `inv` holds bvh4_recip of d's components.
NPH_HD f32 slot_entry(const NphPhysicsMeshNode *n, u32 k, NphVec3 o, NphVec3 inv, NphVec3 grow,
                      f32 limit)
{
    const f32 x0 = ((n->min_x[k] - grow.x) - o.x) * inv.x;
    const f32 x1 = ((n->max_x[k] + grow.x) - o.x) * inv.x;
    const f32 y0 = ((n->min_y[k] - grow.y) - o.y) * inv.y;
    const f32 y1 = ((n->max_y[k] + grow.y) - o.y) * inv.y;
    const f32 z0 = ((n->min_z[k] - grow.z) - o.z) * inv.z;
    const f32 z1 = ((n->max_z[k] + grow.z) - o.z) * inv.z;
    const f32 in = nph_maxf(nph_maxf(nph_minf(x0, x1), nph_minf(y0, y1)),
                            nph_maxf(nph_minf(z0, z1), 0.0F));
    const f32 out = nph_minf(nph_minf(nph_maxf(x0, x1), nph_maxf(y0, y1)),
                             nph_minf(nph_maxf(z0, z1), limit));
    return in <= out ? in : NPH_F32_MAX;
}

typedef struct NphSegmentWalk {
    const NphPhysicsMesh *mesh;
    NphVec3 origin;
    NphVec3 inv;
    NphVec3 grow;
    f32 best; // the caller lowers it as it finds hits
    u32 next;
    u32 end;
    u32 top;
    u32 overflow;
    u32 code[NPH_BVH4_STACK];
    f32 entry[NPH_BVH4_STACK];
} NphSegmentWalk;

// A walk along origin + t delta (t up to 1) of a box with half extents
// `grow` (zero for a ray); `best` is the fraction beyond which nothing is
// wanted yet.
NPH_HD void segment_walk_begin(NphSegmentWalk *w, const NphPhysicsMesh *m, NphVec3 origin,
                               NphVec3 delta, NphVec3 grow, f32 best)
{
    w->mesh = m;
    w->origin = origin;
    w->inv = v3(bvh4_recip(delta.x), bvh4_recip(delta.y), bvh4_recip(delta.z));
    w->grow = grow;
    w->best = best;
    w->next = 0U;
    w->end = 0U;
    w->top = 1U;
    w->overflow = 0U;
    w->code[0] = 0U;
    w->entry[0] = 0.0F;
}

// The next triangle of a leaf the segment reaches before `best`;
// NPH_NULL32 when the walk is over.
NPH_HD u32 segment_walk_next(NphSegmentWalk *w)
{
    for (;;) {
        if (w->next < w->end) {
            return w->next++;
        }
        if (w->top == 0U) {
            return NPH_NULL32;
        }
        w->top -= 1U;
        if (w->entry[w->top] > w->best) {
            continue; // entered beyond the best hit found since it was pushed
        }
        const u32 code = w->code[w->top];
        if (mesh_child_is_leaf(code)) {
            w->next = mesh_leaf_first(code);
            w->end = w->next + mesh_leaf_count(code);
            continue;
        }
        const NphPhysicsMeshNode *n = &mesh_nodes(w->mesh)[code];
        const f32 limit = nph_minf(w->best, 1.0F);
        f32 e[4];
        for (u32 k = 0U; k < 4U; ++k) {
            e[k] = n->child[k] == NPH_PHYSICS_MESH_CHILD_EMPTY
                       ? NPH_F32_MAX
                       : slot_entry(n, k, w->origin, w->inv, w->grow, limit);
        }
        // Pushed farthest first (equal entries: the lower slot first), so
        // the nearest comes off next.
        u32 taken = 0U;
        for (u32 round = 0U; round < 4U; ++round) {
            u32 pick = 4U;
            for (u32 k = 0U; k < 4U; ++k) {
                if ((taken & (1U << k)) == 0U && (pick == 4U || e[k] > e[pick])) {
                    pick = k;
                }
            }
            taken |= 1U << pick;
            if (!(e[pick] < NPH_F32_MAX) || !(e[pick] <= w->best)) {
                continue;
            }
            if (w->top == NPH_BVH4_STACK) {
                w->overflow = 1U;
                w->top = 0U;
                w->end = w->next;
                return NPH_NULL32;
            }
            w->code[w->top] = n->child[pick];
            w->entry[w->top] = e[pick];
            w->top += 1U;
        }
    }
}

// ---------------------------------------------------------------------------
// The watertight ray
// ---------------------------------------------------------------------------
typedef struct NphShearRay {
    NphVec3 origin;
    u32 axis[3];  // the two axes across the ray, then its dominant axis
    f32 shear[2]; // the across components of the direction per unit along it
    f32 scale;    // 1 / the direction's dominant component
} NphShearRay;

NPH_HD NphShearRay shear_ray(NphVec3 origin, NphVec3 delta)
{
    NphShearRay r;
    r.origin = origin;
    const u32 z = v3_max_abs_index(delta);
    u32 x = z == 2U ? 0U : z + 1U;
    u32 y = x == 2U ? 0U : x + 1U;
    const f32 dz = v3_at(delta, z);
    if (dz < 0.0F) {
        // Running down the axis: swap the other two to keep the winding.
        const u32 t = x;
        x = y;
        y = t;
    }
    r.axis[0] = x;
    r.axis[1] = y;
    r.axis[2] = z;
    r.shear[0] = v3_at(delta, x) / dz;
    r.shear[1] = v3_at(delta, y) / dz;
    r.scale = 1.0F / dz;
    return r;
}

// Point p in the ray's sheared frame: its offset across the ray and its
// height along the dominant axis.
typedef struct NphSheared {
    f32 x, y, h;
} NphSheared;

NPH_HD NphSheared shear_point(const NphShearRay *r, NphVec3 p)
{
    const NphVec3 q = v3_sub(p, r->origin);
    NphSheared s;
    s.h = v3_at(q, r->axis[2]);
    s.x = v3_at(q, r->axis[0]) - r->shear[0] * s.h;
    s.y = v3_at(q, r->axis[1]) - r->shear[1] * s.h;
    return s;
}

// p x q in the plane across the ray: twice the signed area of the ray's
// trace and the edge p -> q.
NPH_HD f32 shear_edge(NphSheared p, NphSheared q)
{
    return p.x * q.y - p.y * q.x;
}
NPH_HD f32 shear_edge_wide(NphSheared p, NphSheared q)
{
    return (f32)((f64)p.x * (f64)q.y - (f64)p.y * (f64)q.x);
}

// The fraction where the ray crosses triangle (pa, pb, pc), negative for a
// miss; *back is set when it crosses from behind (clockwise as the ray sees
// it).
NPH_HD f32 shear_ray_cross(const NphShearRay *r, NphVec3 pa, NphVec3 pb, NphVec3 pc, int *back)
{
    const NphSheared a = shear_point(r, pa);
    const NphSheared b = shear_point(r, pb);
    const NphSheared c = shear_point(r, pc);
    f32 u = shear_edge(c, b);
    f32 v = shear_edge(a, c);
    f32 w = shear_edge(b, a);
    if (u == 0.0F || v == 0.0F || w == 0.0F) {
        u = shear_edge_wide(c, b);
        v = shear_edge_wide(a, c);
        w = shear_edge_wide(b, a);
    }
    const int some_negative = u < 0.0F || v < 0.0F || w < 0.0F;
    const int some_positive = u > 0.0F || v > 0.0F || w > 0.0F;
    if (some_negative && some_positive) {
        return -1.0F; // the trace is outside an edge
    }
    const f32 det = u + v + w;
    if (det == 0.0F) {
        return -1.0F; // seen edge on
    }
    const f32 t = (u * r->scale * a.h + v * r->scale * b.h + w * r->scale * c.h) / det;
    if (!(t >= 0.0F && t <= 1.0F)) {
        return -1.0F;
    }
    *back = det < 0.0F;
    return t;
}

// The closest triangle the segment origin -> origin + delta crosses: for a
// one-sided mesh only from the front; equal fractions keep the first found.
typedef struct NphBvh4Hit {
    f32 fraction;
    u32 triangle; // NPH_NULL32: none
    int back;
} NphBvh4Hit;

NPH_HD NphBvh4Hit bvh4_raycast(const NphPhysicsMesh *m, NphVec3 origin, NphVec3 delta,
                               u32 *overflow)
{
    NphBvh4Hit hit;
    hit.fraction = NPH_F32_MAX;
    hit.triangle = NPH_NULL32;
    hit.back = 0;
    *overflow = 0U;
    if (!(v3_len_sq(delta) > 0.0F)) {
        return hit;
    }
    const int two_sided = (m->flags & NPH_PHYSICS_MESH_TWO_SIDED) != 0U;
    const NphShearRay ray = shear_ray(origin, delta);
    NphSegmentWalk w;
    segment_walk_begin(&w, m, origin, delta, v3_zero(), NPH_F32_MAX);
    for (u32 t = segment_walk_next(&w); t != NPH_NULL32; t = segment_walk_next(&w)) {
        NphVec3 p[3];
        mesh_corners(m, t, p);
        int back = 0;
        const f32 f = shear_ray_cross(&ray, p[0], p[1], p[2], &back);
        if (f >= 0.0F && f < w.best && (two_sided || !back)) {
            w.best = f;
            hit.fraction = f;
            hit.triangle = t;
            hit.back = back;
        }
    }
    *overflow = w.overflow;
    return hit;
}

#endif
