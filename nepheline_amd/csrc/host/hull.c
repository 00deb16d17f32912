// SPDX-License-Identifier: Apache-2.0
// Hull blocks on the host (engine/physics/include/nepheline/physics/hull.h):
// the block packer, the box hull, the Quickhull builder, and the block
// queries and validator.
//
// A hull's topology is part of every contact it makes: face and half-edge
// numbers go into the manifold points' ids, which key the warm start. So the
// builder must write the reference's bytes for the same cloud, not merely
// some valid hull of it. It takes the reference's decisions in the
// reference's order (hull.c:350-1197): the initial tetrahedron from the
// extreme points, the farthest outside point first, the depth-first horizon,
// John Lloyd's two merge passes (QuickHull3D: non-convex with respect to the
// larger face, then with respect to either), orphaned points re-assigned to
// the new faces, every distance in f64 on the f32 input; and it packs faces,
// half-edges and vertices in the same order with the same f64 arithmetic
// for planes and mass (hull.c:128-309). The block is host data: worlds copy
// it to the device when a shape first uses it.

#include "world_impl.h"

#include <float.h>
#include <math.h>
#include <stdio.h>

#define HNONE 0xFFFFFFFFu

static void *hull_take(NphAllocator *a, u64 size)
{
    NphCallsite cs = {__FILE__, (u32)__LINE__};
    return a->vtable->alloc(a, size > 0U ? size : 16U, 16U, cs);
}

static void hull_give(NphAllocator *a, void *p, u64 size)
{
    NphCallsite cs = {__FILE__, (u32)__LINE__};
    if (p != NULL) {
        a->vtable->free(a, p, size > 0U ? size : 16U, 16U, cs);
    }
}

static u32 up16(u32 n)
{
    return (n + 15U) & ~15U;
}

// ---------------------------------------------------------------------------
// f64 vectors (vec.h:899-964: component-wise, dot left to right, scaling by
// a reciprocal)
// ---------------------------------------------------------------------------
static NphDVec3 d3(f64 x, f64 y, f64 z)
{
    NphDVec3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
static NphDVec3 d3_of(NphVec3 v)
{
    return d3((f64)v.x, (f64)v.y, (f64)v.z);
}
static NphVec3 d3_narrow(NphDVec3 v)
{
    NphVec3 r;
    r.x = (f32)v.x;
    r.y = (f32)v.y;
    r.z = (f32)v.z;
    return r;
}
static NphDVec3 d3_add(NphDVec3 a, NphDVec3 b)
{
    return d3(a.x + b.x, a.y + b.y, a.z + b.z);
}
static NphDVec3 d3_sub(NphDVec3 a, NphDVec3 b)
{
    return d3(a.x - b.x, a.y - b.y, a.z - b.z);
}
static NphDVec3 d3_scale(NphDVec3 v, f64 s)
{
    return d3(v.x * s, v.y * s, v.z * s);
}
static f64 d3_dot(NphDVec3 a, NphDVec3 b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
static NphDVec3 d3_cross(NphDVec3 a, NphDVec3 b)
{
    return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static f64 d3_comp(NphDVec3 v, u32 i)
{
    return i == 0U ? v.x : (i == 1U ? v.y : v.z);
}
// The unit vector along v; zero for a zero or non-finite v; a v whose
// squared length overflows is scaled down by its largest component first.
static NphDVec3 d3_unit_or_zero(NphDVec3 v)
{
    const f64 sq = d3_dot(v, v);
    if (sq > 0.0 && sq <= DBL_MAX) {
        return d3_scale(v, 1.0 / sqrt(sq));
    }
    if (!(sq > DBL_MAX) || !isfinite(v.x) || !isfinite(v.y) || !isfinite(v.z)) {
        return d3(0.0, 0.0, 0.0);
    }
    const f64 ax = fabs(v.x), ay = fabs(v.y), az = fabs(v.z);
    const f64 m = ax > (ay > az ? ay : az) ? ax : (ay > az ? ay : az);
    const NphDVec3 u = d3_scale(v, 1.0 / m);
    return d3_scale(u, 1.0 / sqrt(d3_dot(u, u)));
}

// f32 helpers in the reference's order (vec.h:300-368, geometry.h:289).
static f32 f3_dot(NphVec3 a, NphVec3 b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
static NphVec3 f3_sub(NphVec3 a, NphVec3 b)
{
    NphVec3 r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    r.z = a.z - b.z;
    return r;
}
static NphVec3 f3_cross(NphVec3 a, NphVec3 b)
{
    NphVec3 r;
    r.x = a.y * b.z - a.z * b.y;
    r.y = a.z * b.x - a.x * b.z;
    r.z = a.x * b.y - a.y * b.x;
    return r;
}
static f32 f3_len(NphVec3 v)
{
    return sqrtf(f3_dot(v, v));
}
static f32 f_min(f32 a, f32 b)
{
    return a < b ? a : b;
}
static f32 f_max(f32 a, f32 b)
{
    return a > b ? a : b;
}
static f32 plane_height(NphPlane p, NphVec3 x)
{
    return f3_dot(p.normal, x) + p.d;
}
static bool f3_finite(NphVec3 v)
{
    return isfinite(v.x) && isfinite(v.y) && isfinite(v.z);
}

// ---------------------------------------------------------------------------
// Mass properties of a closed polyhedron: Eberly, "Polyhedral Mass
// Properties (Revisited)", Geometric Tools 2002 -- the published
// subexpressions per coordinate and the ten integrals (1, x, y, z, x^2, y^2,
// z^2, xy, yz, zx) summed over a fan of triangles per face. The operation
// order is the publication's, which the reference also keeps (hull.c:69-126).
// ---------------------------------------------------------------------------
typedef struct Eberly {
    f64 f1, f2, f3, g0, g1, g2;
} Eberly;

static Eberly eberly(f64 w0, f64 w1, f64 w2)
{
    Eberly t;
    const f64 s01 = w0 + w1;
    const f64 w0sq = w0 * w0;
    t.f1 = s01 + w2;
    const f64 q = w0sq + w1 * s01;
    t.f2 = q + w2 * t.f1;
    t.f3 = w0 * w0sq + w1 * q + w2 * t.f2;
    t.g0 = t.f2 + w0 * (t.f1 + w0);
    t.g1 = t.f2 + w1 * (t.f1 + w1);
    t.g2 = t.f2 + w2 * (t.f1 + w2);
    return t;
}

static void mass_triangle(f64 acc[10], NphDVec3 p0, NphDVec3 p1, NphDVec3 p2)
{
    const NphDVec3 u = d3_sub(p1, p0);
    const NphDVec3 v = d3_sub(p2, p0);
    // The unnormalised triangle normal (b1 c2 - b2 c1, a2 c1 - a1 c2, a1 b2 - a2 b1).
    const f64 nx = u.y * v.z - v.y * u.z;
    const f64 ny = v.x * u.z - u.x * v.z;
    const f64 nz = u.x * v.y - v.x * u.y;
    const Eberly ex = eberly(p0.x, p1.x, p2.x);
    const Eberly ey = eberly(p0.y, p1.y, p2.y);
    const Eberly ez = eberly(p0.z, p1.z, p2.z);
    acc[0] += nx * ex.f1;
    acc[1] += nx * ex.f2;
    acc[2] += ny * ey.f2;
    acc[3] += nz * ez.f2;
    acc[4] += nx * ex.f3;
    acc[5] += ny * ey.f3;
    acc[6] += nz * ez.f3;
    acc[7] += nx * (p0.y * ex.g0 + p1.y * ex.g1 + p2.y * ex.g2);
    acc[8] += ny * (p0.z * ey.g0 + p1.z * ey.g1 + p2.z * ey.g2);
    acc[9] += nz * (p0.x * ez.g0 + p1.x * ez.g1 + p2.x * ez.g2);
}

// ---------------------------------------------------------------------------
// Packing (hull.c:128-309): vertices and counter-clockwise face loops into a
// block. Face f's loop is loop[offsets[f]] .. loop[offsets[f + 1] - 1].
// ---------------------------------------------------------------------------
static int key_order(const void *pa, const void *pb)
{
    const u64 a = *(const u64 *)pa;
    const u64 b = *(const u64 *)pb;
    return a < b ? -1 : (a > b ? 1 : 0);
}

static NphResult hull_pack(const NphVec3 *vertices, u32 nv, const u32 *loop, const u32 *offsets,
                           u32 nf, f32 tolerance, NphAllocator *a, NphPhysicsHull **out)
{
    const u32 ne = offsets[nf];
    if (nv < 4U || nf < 4U || ne > 0xFFFFU || nv > NPH_PHYSICS_HULL_MAX_VERTICES ||
        (ne & 1U) != 0U) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    // Twins: sort the half-edges by (lower vertex, higher vertex, direction),
    // so each undirected edge is two neighbours, the one leaving the lower
    // vertex first; anything else is a loop set that is not a closed
    // two-manifold.
    u64 *keys = (u64 *)hull_take(a, (u64)ne * 8U);
    u32 *twin = (u32 *)hull_take(a, (u64)ne * 4U);
    u16 *slot = (u16 *)hull_take(a, (u64)ne * 2U);
    NphResult r = NPH_RESULT_OK;
    if (keys == NULL || twin == NULL || slot == NULL) {
        r = NPH_RESULT_OUT_OF_MEMORY;
    } else {
        for (u32 f = 0U; f < nf; ++f) {
            const u32 first = offsets[f];
            const u32 n = offsets[f + 1U] - first;
            for (u32 i = 0U; i < n; ++i) {
                const u32 tail = loop[first + i];
                const u32 head = loop[first + (i + 1U) % n];
                const u64 lo = tail < head ? tail : head;
                const u64 hi = tail < head ? head : tail;
                keys[first + i] = (lo << 44) | (hi << 24) | ((u64)(tail != lo) << 16) |
                                  (u64)(first + i);
            }
        }
        qsort(keys, ne, sizeof(u64), key_order);
        for (u32 k = 0U; k < ne; k += 2U) {
            const u64 x = keys[k];
            const u64 y = keys[k + 1U];
            if ((x >> 24) != (y >> 24) || ((x >> 16) & 1U) != 0U || ((y >> 16) & 1U) != 1U) {
                r = NPH_RESULT_INVALID_ARGUMENT;
                break;
            }
            twin[x & 0xFFFFU] = (u32)(y & 0xFFFFU);
            twin[y & 0xFFFFU] = (u32)(x & 0xFFFFU);
        }
    }
    hull_give(a, keys, (u64)ne * 8U);
    if (r != NPH_RESULT_OK) {
        hull_give(a, twin, (u64)ne * 4U);
        hull_give(a, slot, (u64)ne * 2U);
        return r;
    }
    // Output numbering: each edge's first half-edge in face and loop order
    // gets the next even index, its twin the odd one after it.
    for (u32 h = 0U; h < ne; ++h) {
        slot[h] = 0xFFFFU;
    }
    u32 pair = 0U;
    for (u32 h = 0U; h < ne; ++h) {
        if (slot[h] == 0xFFFFU) {
            slot[h] = (u16)pair;
            slot[twin[h]] = (u16)(pair + 1U);
            pair += 2U;
        }
    }

    const u32 v_off = (u32)sizeof(NphPhysicsHull);
    const u32 p_off = v_off + up16(nv * (u32)sizeof(NphVec3));
    const u32 f_off = p_off + nf * (u32)sizeof(NphPlane);
    const u32 e_off = f_off + up16(nf * 2U);
    const u32 size = e_off + up16(ne * (u32)sizeof(NphPhysicsHullEdge));
    NphPhysicsHull *hull = (NphPhysicsHull *)hull_take(a, size);
    if (hull == NULL) {
        hull_give(a, twin, (u64)ne * 4U);
        hull_give(a, slot, (u64)ne * 2U);
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    memset(hull, 0, size);
    hull->byte_size = size;
    hull->vertex_count = nv;
    hull->edge_count = ne;
    hull->face_count = nf;
    hull->vertices_offset = v_off;
    hull->planes_offset = p_off;
    hull->faces_offset = f_off;
    hull->edges_offset = e_off;
    hull->tolerance = tolerance;
    NphVec3 *vo = (NphVec3 *)((u8 *)hull + v_off);
    NphPlane *po = (NphPlane *)((u8 *)hull + p_off);
    u16 *fo = (u16 *)((u8 *)hull + f_off);
    NphPhysicsHullEdge *eo = (NphPhysicsHullEdge *)((u8 *)hull + e_off);

    memcpy(vo, vertices, (size_t)nv * sizeof(NphVec3));
    NphAABB box;
    box.min = vertices[0];
    box.max = vertices[0];
    NphDVec3 mid = d3(0.0, 0.0, 0.0);
    for (u32 v = 0U; v < nv; ++v) {
        box.min.x = f_min(box.min.x, vertices[v].x);
        box.min.y = f_min(box.min.y, vertices[v].y);
        box.min.z = f_min(box.min.z, vertices[v].z);
        box.max.x = f_max(box.max.x, vertices[v].x);
        box.max.y = f_max(box.max.y, vertices[v].y);
        box.max.z = f_max(box.max.z, vertices[v].z);
        mid = d3_add(mid, d3_of(vertices[v]));
    }
    // Planes and integrals are taken relative to the vertices' mean, for
    // precision far from the origin.
    mid = d3_scale(mid, 1.0 / (f64)nv);
    hull->bounds = box;

    f64 acc[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (u32 f = 0U; f < nf; ++f) {
        const u32 first = offsets[f];
        const u32 n = offsets[f + 1U] - first;
        // Newell's normal of the loop and the loop's mean.
        NphDVec3 nrm = d3(0.0, 0.0, 0.0);
        NphDVec3 mean = d3(0.0, 0.0, 0.0);
        for (u32 i = 0U; i < n; ++i) {
            const NphDVec3 c = d3_sub(d3_of(vertices[loop[first + i]]), mid);
            const NphDVec3 d = d3_sub(d3_of(vertices[loop[first + (i + 1U) % n]]), mid);
            nrm.x += (c.y - d.y) * (c.z + d.z);
            nrm.y += (c.z - d.z) * (c.x + d.x);
            nrm.z += (c.x - d.x) * (c.y + d.y);
            mean = d3_add(mean, c);
        }
        mean = d3_scale(mean, 1.0 / (f64)n);
        nrm = d3_unit_or_zero(nrm);
        po[f].normal = d3_narrow(nrm);
        po[f].d = (f32)(-d3_dot(nrm, d3_add(mean, mid)));
        fo[f] = slot[first];
        for (u32 i = 0U; i < n; ++i) {
            const u32 h = first + i;
            NphPhysicsHullEdge *e = &eo[slot[h]];
            e->next = slot[first + (i + 1U) % n];
            e->twin = slot[twin[h]];
            e->origin = (u16)loop[h];
            e->face = (u16)f;
        }
        const NphDVec3 p0 = d3_sub(d3_of(vertices[loop[first]]), mid);
        for (u32 i = 1U; i + 1U < n; ++i) {
            mass_triangle(acc, p0, d3_sub(d3_of(vertices[loop[first + i]]), mid),
                          d3_sub(d3_of(vertices[loop[first + i + 1U]]), mid));
        }
    }
    hull_give(a, twin, (u64)ne * 4U);
    hull_give(a, slot, (u64)ne * 2U);

    // Eberly's final scale factors.
    static const f64 scale[10] = {1.0 / 6.0,  1.0 / 24.0, 1.0 / 24.0,  1.0 / 24.0,  1.0 / 60.0,
                                  1.0 / 60.0, 1.0 / 60.0, 1.0 / 120.0, 1.0 / 120.0, 1.0 / 120.0};
    for (u32 i = 0U; i < 10U; ++i) {
        acc[i] *= scale[i];
    }
    const f64 vol = acc[0];
    if (!(vol > 0.0)) {
        hull_give(a, hull, size);
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const f64 cx = acc[1] / vol;
    const f64 cy = acc[2] / vol;
    const f64 cz = acc[3] / vol;
    // Moments about the centroid (parallel axis), products negated.
    const f64 ixx = acc[5] + acc[6] - vol * (cy * cy + cz * cz);
    const f64 iyy = acc[4] + acc[6] - vol * (cz * cz + cx * cx);
    const f64 izz = acc[4] + acc[5] - vol * (cx * cx + cy * cy);
    const f64 ixy = -(acc[7] - vol * cx * cy);
    const f64 iyz = -(acc[8] - vol * cy * cz);
    const f64 ixz = -(acc[9] - vol * cz * cx);
    hull->volume = (f32)vol;
    hull->centroid = d3_narrow(d3_add(d3(cx, cy, cz), mid));
    hull->inertia.c[0] = d3_narrow(d3(ixx, ixy, ixz));
    hull->inertia.c[1] = d3_narrow(d3(ixy, iyy, iyz));
    hull->inertia.c[2] = d3_narrow(d3(ixz, iyz, izz));
    f32 inner = FLT_MAX;
    for (u32 f = 0U; f < nf; ++f) {
        inner = f_min(inner, -plane_height(po[f], hull->centroid));
    }
    f32 outer = 0.0F;
    for (u32 v = 0U; v < nv; ++v) {
        outer = f_max(outer, f3_len(f3_sub(vertices[v], hull->centroid)));
    }
    hull->inner_radius = f_max(inner, 0.0F);
    hull->outer_radius = outer;
    *out = hull;
    return NPH_RESULT_OK;
}

void nph_physics_hull_free(NphPhysicsHull *hull, NphAllocator *alloc)
{
    if (hull != NULL && alloc != NULL) {
        hull_give(alloc, hull, hull->byte_size);
    }
}

NphResult nph_physics_hull_make_box(NphVec3 h, NphAllocator *alloc, NphPhysicsHull **out)
{
    const f32 lim = NPH_PHYSICS_MAX_COORDINATE;
    if (alloc == NULL || out == NULL || !(h.x > 0.0F && h.x <= lim) ||
        !(h.y > 0.0F && h.y <= lim) || !(h.z > 0.0F && h.z <= lim)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    // Corner k: bit 0 +x, bit 1 +y, bit 2 +z. Faces -x, +x, -y, +y, -z, +z,
    // each loop counter-clockwise seen from outside (the reference's box
    // numbering, hull.c:333-341, which the box's contact ids depend on).
    NphVec3 corner[8];
    for (u32 k = 0U; k < 8U; ++k) {
        corner[k].x = (k & 1U) != 0U ? h.x : -h.x;
        corner[k].y = (k & 2U) != 0U ? h.y : -h.y;
        corner[k].z = (k & 4U) != 0U ? h.z : -h.z;
    }
    static const u32 loops[24] = {0, 4, 6, 2, 1, 3, 7, 5, 0, 1, 5, 4,
                                  2, 6, 7, 3, 0, 2, 3, 1, 4, 5, 7, 6};
    static const u32 offsets[7] = {0, 4, 8, 12, 16, 20, 24};
    const f32 tol = 3.0F * FLT_EPSILON * (h.x + h.y + h.z);
    return hull_pack(corner, 8U, loops, offsets, 6U, tol, alloc, out);
}

// ---------------------------------------------------------------------------
// Quickhull (hull.c:350-1139)
// ---------------------------------------------------------------------------

// A half-edge of the working mesh. It stores the vertex it ENDS at; its start
// is the previous half-edge's end, so dropping a vertex while merging is a
// relinking only.
typedef struct QEdge {
    u32 head;
    u32 next;
    u32 prev;
    u32 twin;
    u32 face;
} QEdge;

enum { QF_LIVE = 0, QF_NONCONVEX = 1, QF_GONE = 2 };

typedef struct QFace {
    NphDVec3 normal;
    NphDVec3 centroid;
    f64 offset; // height(p) = normal . p - offset
    f64 area;
    f64 far_dist;
    u32 far_point; // the outside point farthest above the face (HNONE: none)
    u32 edge;      // a half-edge of the face
    u32 nverts;
    u32 outside; // the face's outside set, linked through Qh.link
    u32 state;
    u32 pad;
} QFace;

typedef struct QFrame {
    u32 face;
    u32 stop; // the frame ends when its cursor comes back here
    u32 at;
} QFrame;

typedef struct Qh {
    NphAllocator *a;
    NphDVec3 *pts;
    u32 n;
    u32 *link; // per point: the next point of the same outside set
    NphVecBuf edges, faces, horizon, fresh, orphans, frames;
    f64 tol;
    int oom;
} Qh;

#define QE(q) VB((q)->edges, QEdge)
#define QF(q) VB((q)->faces, QFace)

static f64 height(const QFace *f, NphDVec3 p)
{
    return d3_dot(f->normal, p) - f->offset;
}

static u32 tail_of(const Qh *q, u32 e)
{
    return QE(q)[QE(q)[e].prev].head;
}

// The face's plane from its loop: the fan of cross products from the loop's
// first vertex, the mean of its vertices, its area (hull.c:414-465). A face
// of almost no area has an unreliable normal; its component along the
// face's longest edge is removed.
static void face_refit(Qh *q, u32 fi)
{
    const QEdge *E = QE(q);
    QFace *f = &QF(q)[fi];
    const u32 first = f->edge;
    const NphDVec3 origin = q->pts[E[first].head];
    u32 e = E[first].next;
    NphDVec3 last = d3_sub(q->pts[E[e].head], origin);
    NphDVec3 sum = d3_add(origin, q->pts[E[e].head]);
    NphDVec3 nrm = d3(0.0, 0.0, 0.0);
    u32 count = 2U;
    for (e = E[e].next; e != first; e = E[e].next) {
        const NphDVec3 p = q->pts[E[e].head];
        const NphDVec3 d = d3_sub(p, origin);
        nrm = d3_add(nrm, d3_cross(last, d));
        sum = d3_add(sum, p);
        last = d;
        count += 1U;
    }
    f->area = sqrt(d3_dot(nrm, nrm));
    f->nverts = count;
    f->centroid = d3_scale(sum, 1.0 / (f64)count);
    if (f->area > 0.0) {
        nrm = d3_scale(nrm, 1.0 / f->area);
    }
    if (f->area < q->tol * q->tol * 16.0 || !(f->area > 0.0)) {
        u32 longest = first;
        f64 best = 0.0;
        e = first;
        do {
            const NphDVec3 s = d3_sub(q->pts[E[e].head], q->pts[tail_of(q, e)]);
            const f64 len = d3_dot(s, s);
            if (len > best) {
                best = len;
                longest = e;
            }
            e = E[e].next;
        } while (e != first);
        if (best > 0.0) {
            const NphDVec3 s = d3_sub(q->pts[E[longest].head], q->pts[tail_of(q, longest)]);
            const NphDVec3 u = d3_scale(s, 1.0 / sqrt(best));
            nrm = d3_unit_or_zero(d3_sub(nrm, d3_scale(u, d3_dot(nrm, u))));
        }
    }
    f->normal = nrm;
    f->offset = d3_dot(nrm, f->centroid);
}

// A triangle whose loop ends its half-edges at v0, v1, v2 in turn; its
// first half-edge ends at v0.
static u32 add_tri(Qh *q, u32 v0, u32 v1, u32 v2)
{
    const u32 e = q->edges.len;
    const u32 fi = q->faces.len;
    if (vbuf_reserve(q->a, &q->edges, (u32)sizeof(QEdge), e + 3U) != NPH_RESULT_OK ||
        vbuf_reserve(q->a, &q->faces, (u32)sizeof(QFace), fi + 1U) != NPH_RESULT_OK) {
        q->oom = 1;
        return HNONE;
    }
    const u32 heads[3] = {v0, v1, v2};
    QEdge *E = QE(q);
    for (u32 i = 0U; i < 3U; ++i) {
        E[e + i].head = heads[i];
        E[e + i].next = e + (i + 1U) % 3U;
        E[e + i].prev = e + (i + 2U) % 3U;
        E[e + i].twin = HNONE;
        E[e + i].face = fi;
    }
    q->edges.len = e + 3U;
    QFace *f = &QF(q)[fi];
    memset(f, 0, sizeof *f);
    f->edge = e;
    f->outside = HNONE;
    f->far_point = HNONE;
    f->state = QF_LIVE;
    q->faces.len = fi + 1U;
    face_refit(q, fi);
    return fi;
}

// Outside sets: singly linked, newest first; the far point is the first of
// the largest heights in list order.
static void outside_push(Qh *q, u32 fi, u32 p, f64 d)
{
    QFace *f = &QF(q)[fi];
    q->link[p] = f->outside;
    f->outside = p;
    if (f->far_point == HNONE || d > f->far_dist) {
        f->far_point = p;
        f->far_dist = d;
    }
}

static void outside_rescan(Qh *q, u32 fi)
{
    QFace *f = &QF(q)[fi];
    f->far_point = HNONE;
    f->far_dist = 0.0;
    for (u32 p = f->outside; p != HNONE; p = q->link[p]) {
        const f64 d = height(f, q->pts[p]);
        if (f->far_point == HNONE || d > f->far_dist) {
            f->far_point = p;
            f->far_dist = d;
        }
    }
}

static void outside_remove(Qh *q, u32 fi, u32 p)
{
    u32 *at = &QF(q)[fi].outside;
    while (*at != HNONE && *at != p) {
        at = &q->link[*at];
    }
    if (*at == p) {
        *at = q->link[p];
    }
    outside_rescan(q, fi);
}

// Empties a face's outside set, in list order: into face `into` for the
// points above it by more than the tolerance, onto the orphans otherwise.
static void outside_release(Qh *q, u32 fi, u32 into)
{
    u32 p = QF(q)[fi].outside;
    QF(q)[fi].outside = HNONE;
    QF(q)[fi].far_point = HNONE;
    while (p != HNONE) {
        const u32 after = q->link[p];
        bool kept = false;
        if (into != HNONE) {
            const f64 d = height(&QF(q)[into], q->pts[p]);
            if (d > q->tol) {
                outside_push(q, into, p, d);
                kept = true;
            }
        }
        if (!kept && vbuf_push(q->a, &q->orphans, 4U, &p) != NPH_RESULT_OK) {
            q->oom = 1;
        }
        p = after;
    }
}

static u32 across(const Qh *q, u32 e)
{
    return QE(q)[QE(q)[e].twin].face;
}

// The height of the centroid of the face across e above e's own face.
static f64 across_height(const Qh *q, u32 e)
{
    return height(&QF(q)[QE(q)[e].face], QF(q)[across(q, e)].centroid);
}

// Rejoins face fi's loop between `before` and `at` after a merge
// (hull.c:581-621). When both border the same face, the vertex between them
// is redundant: the two half-edges become one, and a triangle across them
// vanishes (returned; HNONE otherwise).
static u32 rejoin(Qh *q, u32 fi, u32 before, u32 at)
{
    QEdge *E = QE(q);
    QFace *F = QF(q);
    if (across(q, before) != across(q, at)) {
        E[before].next = at;
        E[at].prev = before;
        return HNONE;
    }
    const u32 other = across(q, at);
    u32 gone = HNONE;
    u32 mate;
    if (before == F[fi].edge) {
        F[fi].edge = at;
    }
    if (F[other].nverts == 3U) {
        mate = E[E[E[at].twin].prev].twin;
        F[other].state = QF_GONE;
        gone = other;
    } else {
        mate = E[E[at].twin].next;
        if (F[other].edge == E[mate].prev) {
            F[other].edge = mate;
        }
        E[mate].prev = E[E[mate].prev].prev;
        E[E[mate].prev].next = mate;
    }
    E[at].prev = E[before].prev;
    E[E[at].prev].next = at;
    E[at].twin = mate;
    E[mate].twin = at;
    if (F[other].state != QF_GONE) {
        face_refit(q, other);
        outside_rescan(q, other);
    }
    return gone;
}

// Face fi takes in the face across its half-edge e, with every further
// half-edge the two share (hull.c:623-664). The faces that vanish go to
// `gone`; returns their count (1 to 3).
static u32 absorb(Qh *q, u32 fi, u32 e, u32 gone[3])
{
    QEdge *E = QE(q);
    QFace *F = QF(q);
    const u32 other = across(q, e);
    u32 n = 0U;
    gone[n++] = other;
    F[other].state = QF_GONE;
    const u32 mate = E[e].twin;
    u32 own_prev = E[e].prev;
    u32 own_next = E[e].next;
    u32 mate_prev = E[mate].prev;
    u32 mate_next = E[mate].next;
    while (across(q, own_prev) == other) {
        own_prev = E[own_prev].prev;
        mate_next = E[mate_next].next;
    }
    while (across(q, own_next) == other) {
        mate_prev = E[mate_prev].prev;
        own_next = E[own_next].next;
    }
    for (u32 h = mate_next; h != E[mate_prev].next; h = E[h].next) {
        E[h].face = fi;
    }
    if (e == F[fi].edge) {
        F[fi].edge = own_next;
    }
    u32 g = rejoin(q, fi, mate_prev, own_next);
    if (g != HNONE) {
        gone[n++] = g;
    }
    g = rejoin(q, fi, own_prev, mate_next);
    if (g != HNONE) {
        gone[n++] = g;
    }
    face_refit(q, fi);
    outside_rescan(q, fi);
    return n;
}

enum { MERGE_WRT_LARGER = 0, MERGE_WRT_EITHER = 1 };

// One merge of a non-convex neighbour into fi, if any (hull.c:671-711);
// true when it merged, so the caller looks again. A face left with a
// neighbour non-convex only with respect to the neighbour is marked for the
// second pass.
static bool merge_once(Qh *q, u32 fi, int rule)
{
    const f64 lim = -q->tol;
    u32 e = QF(q)[fi].edge;
    bool convex = true;
    do {
        const u32 other = across(q, e);
        const u32 t = QE(q)[e].twin;
        bool merge = false;
        if (rule == MERGE_WRT_EITHER) {
            merge = across_height(q, e) > lim || across_height(q, t) > lim;
        } else if (QF(q)[fi].area > QF(q)[other].area) {
            if (across_height(q, e) > lim) {
                merge = true;
            } else if (across_height(q, t) > lim) {
                convex = false;
            }
        } else {
            if (across_height(q, t) > lim) {
                merge = true;
            } else if (across_height(q, e) > lim) {
                convex = false;
            }
        }
        if (merge) {
            u32 gone[3];
            const u32 n = absorb(q, fi, e, gone);
            for (u32 i = 0U; i < n; ++i) {
                outside_release(q, gone[i], fi);
            }
            return true;
        }
        e = QE(q)[e].next;
    } while (e != QF(q)[fi].edge);
    if (!convex) {
        QF(q)[fi].state = QF_NONCONVEX;
    }
    return false;
}

// Removes every face that sees `eye`, depth first from `start`, and lists
// the horizon: the half-edges of removed faces whose twins stay, in order
// around the eye (hull.c:717-752). An explicit stack of frames stands in for
// the recursion; a frame walks its face's loop from the half-edge after the
// one it was entered by (the start face: from its first half-edge).
static void carve(Qh *q, NphDVec3 eye, u32 start)
{
    q->horizon.len = 0U;
    q->frames.len = 0U;
    outside_release(q, start, HNONE);
    QF(q)[start].state = QF_GONE;
    const QFrame root = {start, QF(q)[start].edge, QF(q)[start].edge};
    if (vbuf_push(q->a, &q->frames, (u32)sizeof(QFrame), &root) != NPH_RESULT_OK) {
        q->oom = 1;
        return;
    }
    while (q->frames.len > 0U && !q->oom) {
        QFrame *top = &VB(q->frames, QFrame)[q->frames.len - 1U];
        const u32 e = top->at;
        top->at = QE(q)[e].next;
        if (top->at == top->stop) {
            q->frames.len -= 1U;
        }
        const u32 t = QE(q)[e].twin;
        const u32 nb = QE(q)[t].face;
        if (QF(q)[nb].state == QF_GONE) {
            continue;
        }
        if (height(&QF(q)[nb], eye) > q->tol) {
            outside_release(q, nb, HNONE);
            QF(q)[nb].state = QF_GONE;
            const QFrame child = {nb, t, QE(q)[t].next};
            if (vbuf_push(q->a, &q->frames, (u32)sizeof(QFrame), &child) != NPH_RESULT_OK) {
                q->oom = 1;
            }
        } else if (vbuf_push(q->a, &q->horizon, 4U, &e) != NPH_RESULT_OK) {
            q->oom = 1;
        }
    }
}

// A triangle from each horizon half-edge to the eye, twinned to the face
// beyond the horizon and to its neighbours in the cone (hull.c:756-804).
static void cone(Qh *q, u32 eye)
{
    q->fresh.len = 0U;
    u32 side_prev = HNONE;
    u32 side_first = HNONE;
    for (u32 i = 0U; i < q->horizon.len && !q->oom; ++i) {
        const u32 h = VB(q->horizon, u32)[i];
        const u32 tail = tail_of(q, h);
        const u32 head = QE(q)[h].head;
        const u32 fi = add_tri(q, eye, tail, head);
        if (fi == HNONE) {
            return;
        }
        QEdge *E = QE(q);
        const u32 to_eye = QF(q)[fi].edge;   // head -> eye
        const u32 from_eye = E[to_eye].next; // eye -> tail
        const u32 base = E[from_eye].next;   // tail -> head, on the horizon
        const u32 beyond = E[h].twin;
        E[base].twin = beyond;
        E[beyond].twin = base;
        if (side_prev != HNONE) {
            E[from_eye].twin = side_prev;
            E[side_prev].twin = from_eye;
        } else {
            side_first = to_eye;
        }
        if (vbuf_push(q->a, &q->fresh, 4U, &fi) != NPH_RESULT_OK) {
            q->oom = 1;
            return;
        }
        side_prev = to_eye;
    }
    if (side_first != HNONE) {
        QEdge *E = QE(q);
        const u32 from_eye = E[side_first].next;
        E[from_eye].twin = side_prev;
        E[side_prev].twin = from_eye;
    }
}

// Each orphan goes to the live new face it is highest above, beyond the
// tolerance; the first such face on ties (hull.c:806-829).
static void adopt_orphans(Qh *q)
{
    for (u32 i = 0U; i < q->orphans.len; ++i) {
        const u32 p = VB(q->orphans, u32)[i];
        f64 best = q->tol;
        u32 best_face = HNONE;
        for (u32 k = 0U; k < q->fresh.len; ++k) {
            const u32 fi = VB(q->fresh, u32)[k];
            if (QF(q)[fi].state != QF_LIVE) {
                continue;
            }
            const f64 d = height(&QF(q)[fi], q->pts[p]);
            if (d > best) {
                best = d;
                best_face = fi;
            }
        }
        if (best_face != HNONE) {
            outside_push(q, best_face, p, best);
        }
    }
    q->orphans.len = 0U;
}

static void insert_point(Qh *q, u32 eye, u32 eye_face)
{
    outside_remove(q, eye_face, eye);
    q->orphans.len = 0U;
    carve(q, q->pts[eye], eye_face);
    if (q->oom) {
        return;
    }
    cone(q, eye);
    if (q->oom) {
        return;
    }
    for (u32 k = 0U; k < q->fresh.len; ++k) {
        const u32 fi = VB(q->fresh, u32)[k];
        if (QF(q)[fi].state == QF_LIVE) {
            while (merge_once(q, fi, MERGE_WRT_LARGER)) {
            }
        }
    }
    for (u32 k = 0U; k < q->fresh.len; ++k) {
        const u32 fi = VB(q->fresh, u32)[k];
        if (QF(q)[fi].state == QF_NONCONVEX) {
            QF(q)[fi].state = QF_LIVE;
            while (merge_once(q, fi, MERGE_WRT_EITHER)) {
            }
        }
    }
    adopt_orphans(q);
}

// The first tetrahedron (hull.c:856-962): the two extreme points along the
// axis of largest spread, the point farthest from their line, the point
// farthest from the plane of the three; then every other point to the face
// it is highest above. False for a cloud without volume.
static bool seed_tetrahedron(Qh *q)
{
    const NphDVec3 *P = q->pts;
    u32 lo[3] = {0U, 0U, 0U};
    u32 hi[3] = {0U, 0U, 0U};
    for (u32 i = 1U; i < q->n; ++i) {
        for (u32 ax = 0U; ax < 3U; ++ax) {
            const f64 c = d3_comp(P[i], ax);
            if (c < d3_comp(P[lo[ax]], ax)) {
                lo[ax] = i;
            }
            if (c > d3_comp(P[hi[ax]], ax)) {
                hi[ax] = i;
            }
        }
    }
    f64 span = 0.0;
    u32 axis = 0U;
    for (u32 ax = 0U; ax < 3U; ++ax) {
        const f64 s = d3_comp(P[hi[ax]], ax) - d3_comp(P[lo[ax]], ax);
        if (s > span) {
            span = s;
            axis = ax;
        }
    }
    if (span <= q->tol) {
        return false;
    }
    u32 v[4] = {hi[axis], lo[axis], HNONE, HNONE};
    const NphDVec3 dir = d3_unit_or_zero(d3_sub(P[v[1]], P[v[0]]));
    f64 best = 0.0;
    NphDVec3 nrm = d3(0.0, 0.0, 0.0);
    for (u32 i = 0U; i < q->n; ++i) {
        const NphDVec3 x = d3_cross(dir, d3_sub(P[i], P[v[0]]));
        const f64 len = d3_dot(x, x);
        if (len > best && i != v[0] && i != v[1]) {
            best = len;
            v[2] = i;
            nrm = x;
        }
    }
    if (v[2] == HNONE || sqrt(best) <= 100.0 * q->tol) {
        return false;
    }
    nrm = d3_unit_or_zero(nrm);
    const f64 d0 = d3_dot(P[v[2]], nrm);
    best = 0.0;
    for (u32 i = 0U; i < q->n; ++i) {
        const f64 dist = fabs(d3_dot(P[i], nrm) - d0);
        if (dist > best && i != v[0] && i != v[1] && i != v[2]) {
            best = dist;
            v[3] = i;
        }
    }
    if (v[3] == HNONE || best <= 100.0 * q->tol) {
        return false;
    }
    // Loops counter-clockwise from outside: the base faces away from v3.
    u32 tri[4];
    if (d3_dot(P[v[3]], nrm) - d0 < 0.0) {
        tri[0] = add_tri(q, v[0], v[1], v[2]);
        tri[1] = add_tri(q, v[3], v[1], v[0]);
        tri[2] = add_tri(q, v[3], v[2], v[1]);
        tri[3] = add_tri(q, v[3], v[0], v[2]);
    } else {
        tri[0] = add_tri(q, v[0], v[2], v[1]);
        tri[1] = add_tri(q, v[3], v[0], v[1]);
        tri[2] = add_tri(q, v[3], v[1], v[2]);
        tri[3] = add_tri(q, v[3], v[2], v[0]);
    }
    if (q->oom) {
        return false;
    }
    // Twins of the twelve half-edges: the one running the other way.
    QEdge *E = QE(q);
    const u32 base = QF(q)[tri[0]].edge;
    for (u32 i = base; i < base + 12U; ++i) {
        for (u32 j = base; j < base + 12U; ++j) {
            if (E[j].head == tail_of(q, i) && tail_of(q, j) == E[i].head) {
                E[i].twin = j;
            }
        }
    }
    for (u32 i = 0U; i < q->n; ++i) {
        if (i == v[0] || i == v[1] || i == v[2] || i == v[3]) {
            continue;
        }
        f64 top = q->tol;
        u32 face = HNONE;
        for (u32 k = 0U; k < 4U; ++k) {
            const f64 d = height(&QF(q)[tri[k]], P[i]);
            if (d > top) {
                top = d;
                face = tri[k];
            }
        }
        if (face != HNONE) {
            outside_push(q, face, i, top);
        }
    }
    return true;
}

static void qh_release(Qh *q)
{
    hull_give(q->a, q->pts, (u64)q->n * sizeof(NphDVec3));
    hull_give(q->a, q->link, (u64)q->n * 4U);
    vbuf_free(q->a, &q->edges, (u32)sizeof(QEdge));
    vbuf_free(q->a, &q->faces, (u32)sizeof(QFace));
    vbuf_free(q->a, &q->horizon, 4U);
    vbuf_free(q->a, &q->fresh, 4U);
    vbuf_free(q->a, &q->orphans, 4U);
    vbuf_free(q->a, &q->frames, (u32)sizeof(QFrame));
}

// The live faces, in creation order, as loops over the vertices numbered in
// order of first use; packed (hull.c:977-1066).
static NphResult emit(Qh *q, const NphVec3 *input, f32 tol, NphAllocator *a,
                      NphPhysicsHull **out)
{
    u32 nf = 0U;
    u32 nl = 0U;
    for (u32 f = 0U; f < q->faces.len; ++f) {
        if (QF(q)[f].state != QF_GONE) {
            nf += 1U;
            nl += QF(q)[f].nverts;
        }
    }
    const u64 b_remap = (u64)q->n * 4U;
    const u64 b_loop = (u64)nl * 4U;
    const u64 b_offs = (u64)(nf + 1U) * 4U;
    const u64 b_verts = (u64)NPH_PHYSICS_HULL_MAX_VERTICES * sizeof(NphVec3);
    u32 *remap = (u32 *)hull_take(a, b_remap);
    u32 *loop = (u32 *)hull_take(a, b_loop);
    u32 *offs = (u32 *)hull_take(a, b_offs);
    NphVec3 *verts = (NphVec3 *)hull_take(a, b_verts);
    NphResult r = NPH_RESULT_OK;
    if (remap == NULL || loop == NULL || offs == NULL || verts == NULL) {
        r = NPH_RESULT_OUT_OF_MEMORY;
    } else {
        for (u32 i = 0U; i < q->n; ++i) {
            remap[i] = HNONE;
        }
        u32 nv = 0U;
        u32 face = 0U;
        u32 cursor = 0U;
        const QEdge *E = QE(q);
        for (u32 f = 0U; f < q->faces.len && r == NPH_RESULT_OK; ++f) {
            const QFace *qf = &QF(q)[f];
            if (qf->state == QF_GONE) {
                continue;
            }
            offs[face++] = cursor;
            u32 e = qf->edge;
            do {
                const u32 p = E[e].head;
                if (remap[p] == HNONE) {
                    if (nv == NPH_PHYSICS_HULL_MAX_VERTICES) {
                        r = NPH_RESULT_INVALID_ARGUMENT;
                        break;
                    }
                    remap[p] = nv;
                    verts[nv++] = input[p];
                }
                loop[cursor++] = remap[p];
                e = E[e].next;
            } while (e != qf->edge);
        }
        offs[face] = cursor;
        if (r == NPH_RESULT_OK) {
            r = hull_pack(verts, nv, loop, offs, nf, tol, a, out);
        }
    }
    hull_give(a, remap, b_remap);
    hull_give(a, loop, b_loop);
    hull_give(a, offs, b_offs);
    hull_give(a, verts, b_verts);
    return r;
}

static bool hull_check(const void *block, u32 block_size, f32 tolerance, bool loud);

// One build at `tol` (hull.c:1076-1139).
static NphResult quickhull(const NphPhysicsHullDesc *desc, u32 budget, f64 tol, NphAllocator *a,
                           NphPhysicsHull **out)
{
    Qh q;
    memset(&q, 0, sizeof q);
    q.a = a;
    q.n = desc->point_count;
    q.tol = tol;
    q.pts = (NphDVec3 *)hull_take(a, (u64)q.n * sizeof(NphDVec3));
    q.link = (u32 *)hull_take(a, (u64)q.n * 4U);
    if (q.pts == NULL || q.link == NULL) {
        qh_release(&q);
        return NPH_RESULT_OUT_OF_MEMORY;
    }
    for (u32 i = 0U; i < q.n; ++i) {
        q.pts[i] = d3_of(desc->points[i]);
        q.link[i] = HNONE;
    }
    if (!seed_tetrahedron(&q)) {
        const int oom = q.oom;
        qh_release(&q);
        return oom ? NPH_RESULT_OUT_OF_MEMORY : NPH_RESULT_INVALID_ARGUMENT;
    }
    u32 used = 4U;
    while (!q.oom && used < budget) {
        // The highest outside point over all faces; the lowest face index
        // on ties.
        u32 eye_face = HNONE;
        f64 top = q.tol;
        for (u32 f = 0U; f < q.faces.len; ++f) {
            const QFace *face = &QF(&q)[f];
            if (face->state != QF_GONE && face->far_point != HNONE && face->far_dist > top) {
                top = face->far_dist;
                eye_face = f;
            }
        }
        if (eye_face == HNONE) {
            break;
        }
        insert_point(&q, QF(&q)[eye_face].far_point, eye_face);
        used += 1U;
    }
    NphResult r = NPH_RESULT_OUT_OF_MEMORY;
    if (!q.oom) {
        r = emit(&q, desc->points, (f32)tol, a, out);
    }
    qh_release(&q);
    return r;
}

NphResult nph_physics_hull_build(const NphPhysicsHullDesc *desc, NphAllocator *alloc,
                                 NphPhysicsHull **out)
{
    if (desc == NULL || alloc == NULL || out == NULL || desc->points == NULL ||
        desc->point_count < 4U) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    const u32 budget =
        desc->max_vertices == 0U ? NPH_PHYSICS_HULL_DEFAULT_MAX_VERTICES : desc->max_vertices;
    if (budget < 4U || budget > NPH_PHYSICS_HULL_MAX_VERTICES || !(desc->tolerance >= 0.0F) ||
        !isfinite(desc->tolerance)) {
        return NPH_RESULT_INVALID_ARGUMENT;
    }
    f64 mx = 0.0;
    f64 my = 0.0;
    f64 mz = 0.0;
    for (u32 i = 0U; i < desc->point_count; ++i) {
        const NphVec3 p = desc->points[i];
        if (!f3_finite(p) || fabsf(p.x) > NPH_PHYSICS_MAX_COORDINATE ||
            fabsf(p.y) > NPH_PHYSICS_MAX_COORDINATE || fabsf(p.z) > NPH_PHYSICS_MAX_COORDINATE) {
            return NPH_RESULT_INVALID_ARGUMENT;
        }
        const f64 ax = fabs((f64)p.x), ay = fabs((f64)p.y), az = fabs((f64)p.z);
        mx = mx > ax ? mx : ax;
        my = my > ay ? my : ay;
        mz = mz > az ? mz : az;
    }
    // The round-off bound of a plane distance over the cloud's extent
    // (Barber et al. 1996, section 4).
    f64 tol = desc->tolerance > 0.0F ? (f64)desc->tolerance
                                     : 3.0 * (f64)FLT_EPSILON * (mx + my + mz);
    // A cloud with features near the tolerance's scale can leave a merged
    // face reflex beyond it; such a hull is built again at 4x, then 16x the
    // tolerance, and refused after that.
    for (u32 attempt = 0U; attempt < 3U; ++attempt) {
        const NphResult r = quickhull(desc, budget, tol, alloc, out);
        if (r != NPH_RESULT_OK) {
            return r;
        }
        if (hull_check(*out, (*out)->byte_size, 0.0F, false)) {
            return NPH_RESULT_OK;
        }
        nph_physics_hull_free(*out, alloc);
        *out = NULL;
        tol *= 4.0;
    }
    fprintf(stderr, "nepheline_amd: hull build: the %u points give no hull valid within %g\n",
            desc->point_count, tol / 4.0);
    return NPH_RESULT_INVALID_ARGUMENT;
}

// ---------------------------------------------------------------------------
// Queries (hull.c:1203-1230)
// ---------------------------------------------------------------------------
u32 nph_physics_hull_support(const NphPhysicsHull *hull, NphVec3 direction)
{
    if (hull == NULL || hull->vertex_count == 0U) {
        return 0U;
    }
    const NphVec3 *v = (const NphVec3 *)((const u8 *)hull + hull->vertices_offset);
    u32 best = 0U;
    f32 top = f3_dot(v[0], direction);
    for (u32 i = 1U; i < hull->vertex_count; ++i) {
        const f32 d = f3_dot(v[i], direction);
        if (d > top) {
            top = d;
            best = i;
        }
    }
    return best;
}

u32 nph_physics_hull_face_vertex_count(const NphPhysicsHull *hull, u32 face)
{
    if (hull == NULL || face >= hull->face_count) {
        return 0U;
    }
    const NphPhysicsHullEdge *E =
        (const NphPhysicsHullEdge *)((const u8 *)hull + hull->edges_offset);
    const u32 first = ((const u16 *)((const u8 *)hull + hull->faces_offset))[face];
    u32 n = 0U;
    u32 e = first;
    do {
        n += 1U;
        e = E[e].next;
    } while (e != first && n <= hull->edge_count);
    return n;
}

// ---------------------------------------------------------------------------
// Validation (hull.c:1254-1377): the same checks in the same order, so a
// build is accepted or retried exactly when the reference's is.
// ---------------------------------------------------------------------------
#define REJECT(...)                                                                                \
    do {                                                                                           \
        if (loud) {                                                                                \
            fprintf(stderr, "nepheline_amd: hull validate: " __VA_ARGS__);                        \
            fputc('\n', stderr);                                                                   \
        }                                                                                          \
        return false;                                                                              \
    } while (0)

static bool in_range(f32 v)
{
    return fabsf(v) <= NPH_PHYSICS_MAX_COORDINATE; // false for NaN
}

static bool v_in_range(NphVec3 v)
{
    return in_range(v.x) && in_range(v.y) && in_range(v.z);
}

static bool hull_check(const void *block, u32 block_size, f32 tolerance, bool loud)
{
    const u32 head = (u32)sizeof(NphPhysicsHull);
    if (block == NULL || block_size < head || ((uintptr_t)block & 15U) != 0U) {
        REJECT("no 16-byte aligned header in %u bytes", block_size);
    }
    const NphPhysicsHull *hull = (const NphPhysicsHull *)block;
    if (hull->byte_size < head || hull->byte_size > block_size) {
        REJECT("byte_size %u outside the %u-byte block", hull->byte_size, block_size);
    }
    const u64 size = hull->byte_size;
    const u64 v_end = (u64)hull->vertices_offset + (u64)hull->vertex_count * sizeof(NphVec3);
    const u64 p_end = (u64)hull->planes_offset + (u64)hull->face_count * sizeof(NphPlane);
    const u64 f_end = (u64)hull->faces_offset + (u64)hull->face_count * 2U;
    const u64 e_end = (u64)hull->edges_offset + (u64)hull->edge_count * sizeof(NphPhysicsHullEdge);
    if (hull->vertices_offset < head || hull->planes_offset < head || hull->faces_offset < head ||
        hull->edges_offset < head || v_end > size || p_end > size || f_end > size ||
        e_end > size || (hull->vertices_offset & 3U) != 0U || (hull->planes_offset & 3U) != 0U ||
        (hull->faces_offset & 1U) != 0U || (hull->edges_offset & 1U) != 0U) {
        REJECT("arrays outside the %u-byte block", hull->byte_size);
    }
    const u32 nv = hull->vertex_count;
    const u32 ne = hull->edge_count;
    const u32 nf = hull->face_count;
    // At most 2V - 4 faces and 3V - 6 edges for V vertices (Euler).
    if (nv < 4U || nv > NPH_PHYSICS_HULL_MAX_VERTICES || nf < 4U || nf > 2U * nv - 4U ||
        ne < 12U || (ne & 1U) != 0U || ne / 2U > 3U * nv - 6U) {
        REJECT("counts nv=%u ne=%u nf=%u", nv, ne, nf);
    }
    const int64_t chi = (int64_t)nv - (int64_t)(ne / 2U) + (int64_t)nf;
    if (chi != 2) {
        REJECT("Euler characteristic nv - ne + nf = %lld", (long long)chi);
    }
    const NphAABB box = hull->bounds;
    bool ok = v_in_range(box.min) && v_in_range(box.max) && box.min.x <= box.max.x &&
              box.min.y <= box.max.y && box.min.z <= box.max.z && v_in_range(hull->centroid) &&
              hull->volume > 0.0F && isfinite(hull->volume) && hull->tolerance > 0.0F &&
              hull->tolerance <= NPH_PHYSICS_MAX_COORDINATE && hull->inner_radius >= 0.0F &&
              hull->inner_radius <= hull->outer_radius &&
              hull->outer_radius <= 4.0F * NPH_PHYSICS_MAX_COORDINATE;
    for (u32 i = 0U; i < 3U; ++i) {
        ok = ok && f3_finite(hull->inertia.c[i]);
    }
    if (!ok) {
        REJECT("header numbers not finite or out of range");
    }
    const f32 tol = tolerance > 0.0F ? tolerance : f_max(hull->tolerance, 1e-6F) * 8.0F;
    const NphVec3 *V = (const NphVec3 *)((const u8 *)hull + hull->vertices_offset);
    for (u32 v = 0U; v < nv; ++v) {
        const NphVec3 p = V[v];
        if (!f3_finite(p) || !(p.x >= box.min.x && p.x <= box.max.x && p.y >= box.min.y &&
                               p.y <= box.max.y && p.z >= box.min.z && p.z <= box.max.z)) {
            REJECT("vertex %u not finite or outside the bounds", v);
        }
    }
    const NphPlane *PL = (const NphPlane *)((const u8 *)hull + hull->planes_offset);
    const u16 *FA = (const u16 *)((const u8 *)hull + hull->faces_offset);
    const NphPhysicsHullEdge *E =
        (const NphPhysicsHullEdge *)((const u8 *)hull + hull->edges_offset);
    for (u32 i = 0U; i < ne; ++i) {
        const NphPhysicsHullEdge *e = &E[i];
        if (e->twin != (i ^ 1U) || e->next >= ne || e->origin >= nv || e->face >= nf) {
            REJECT("half-edge %u has bad links", i);
        }
        if (E[e->next].face != e->face) {
            REJECT("half-edge %u: next leaves the face", i);
        }
        if (E[e->twin].origin != E[e->next].origin) {
            REJECT("half-edge %u: twin does not start at its head", i);
        }
        if (E[e->twin].face == e->face) {
            REJECT("half-edge %u: both sides on face %u", i, (u32)e->face);
        }
    }
    u32 seen = 0U;
    for (u32 f = 0U; f < nf; ++f) {
        if (FA[f] >= ne || E[FA[f]].face != f) {
            REJECT("face %u: bad first edge", f);
        }
        const NphPlane pl = PL[f];
        if (!f3_finite(pl.normal) || fabsf(f3_len(pl.normal) - 1.0F) > 1e-4F || !isfinite(pl.d)) {
            REJECT("face %u: plane normal not unit", f);
        }
        u32 n = 0U;
        u32 e = FA[f];
        do {
            const u32 nx = E[e].next;
            const NphVec3 a = V[E[e].origin];
            const NphVec3 b = V[E[nx].origin];
            const NphVec3 c = V[E[E[nx].next].origin];
            if (fabsf(plane_height(pl, a)) > tol) {
                REJECT("face %u: vertex %u is %g off its plane", f, (u32)E[e].origin,
                       (f64)plane_height(pl, a));
            }
            // A convex corner: c left of a -> b in the face's plane.
            const NphVec3 ab = f3_sub(b, a);
            const NphVec3 bc = f3_sub(c, b);
            if (f3_dot(f3_cross(ab, bc), pl.normal) < -tol * f3_len(ab)) {
                REJECT("face %u: reflex corner at vertex %u", f, (u32)E[nx].origin);
            }
            n += 1U;
            e = nx;
        } while (e != FA[f] && n <= ne);
        if (e != FA[f]) {
            REJECT("face %u: loop does not close", f);
        }
        seen += n;
    }
    if (seen != ne) {
        REJECT("face loops cover %u of %u half-edges", seen, ne);
    }
    for (u32 f = 0U; f < nf; ++f) {
        for (u32 v = 0U; v < nv; ++v) {
            if (plane_height(PL[f], V[v]) > tol) {
                REJECT("vertex %u is %g in front of face %u", v, (f64)plane_height(PL[f], V[v]),
                       f);
            }
        }
    }
    return true;
}

bool nph_physics_hull_validate(const void *block, u32 size, f32 tolerance)
{
    return hull_check(block, size, tolerance, true);
}
