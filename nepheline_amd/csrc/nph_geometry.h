// SPDX-License-Identifier: Apache-2.0
// Shape geometry for the host and the kernels: mass closed forms, boxes,
// inner/outer radii, support points (engine/physics/src/geometry.c).
#ifndef NPH_GEOMETRY_H
#define NPH_GEOMETRY_H

#include "nph_math.h"

// The geometry as the device keeps it: type, up to three sizes, and the hull
// or mesh block (device address on the device, host address on the host).
typedef struct NphGeom {
    u32 type;
    f32 a, b, c; // sphere: r; capsule: half_height, radius; box: half extents
    union {
        const NphPhysicsHull *hull;
        const NphPhysicsMesh *mesh;
    };
} NphGeom;

// `block_address`: the hull or mesh block as the caller will read it.
NPH_HD NphGeom geom_from(const NphPhysicsGeometry *g, const void *block_address)
{
    NphGeom o;
    o.type = (u32)g->type;
    o.a = 0.0F;
    o.b = 0.0F;
    o.c = 0.0F;
    o.hull = 0;
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        o.a = g->sphere.radius;
        break;
    case NPH_PHYSICS_SHAPE_CAPSULE:
        o.a = g->capsule.half_height;
        o.b = g->capsule.radius;
        break;
    case NPH_PHYSICS_SHAPE_BOX:
        o.a = g->box.half_extents.x;
        o.b = g->box.half_extents.y;
        o.c = g->box.half_extents.z;
        break;
    case NPH_PHYSICS_SHAPE_MESH:
        o.mesh = (const NphPhysicsMesh *)block_address;
        break;
    default:
        o.hull = (const NphPhysicsHull *)block_address;
        break;
    }
    return o;
}

// geometry.c:16-25: mass (|r|^2 E - r r^T)
NPH_HD NphMat3 inertia_offset(f32 mass, NphVec3 r)
{
    const f32 rr = v3_dot(r, r);
    const NphVec3 c0 = v3(rr - r.x * r.x, -r.y * r.x, -r.z * r.x);
    const NphVec3 c1 = v3(-r.x * r.y, rr - r.y * r.y, -r.z * r.y);
    const NphVec3 c2 = v3(-r.x * r.z, -r.y * r.z, rr - r.z * r.z);
    return m3_scale(m3_cols(c0, c1, c2), mass);
}

// geometry.c:27-35
NPH_HD NphPhysicsMassData mass_transform(NphPhysicsMassData m, NphPhysicsPose pose)
{
    const NphMat3 r = m3_from_quat(pose.rotation);
    NphPhysicsMassData o;
    o.mass = m.mass;
    o.center = pose_point(pose, m.center);
    o.inertia = m3_mul(m3_mul(r, m.inertia), m3_transpose(r));
    return o;
}

// geometry.c:37-56
NPH_HD NphPhysicsMassData mass_add(NphPhysicsMassData a, NphPhysicsMassData b)
{
    if (a.mass <= 0.0F) {
        return b;
    }
    if (b.mass <= 0.0F) {
        return a;
    }
    NphPhysicsMassData o;
    o.mass = a.mass + b.mass;
    o.center = v3_scale(v3_add(v3_scale(a.center, a.mass), v3_scale(b.center, b.mass)),
                        1.0F / o.mass);
    const NphMat3 ia = m3_add(a.inertia, inertia_offset(a.mass, v3_sub(a.center, o.center)));
    const NphMat3 ib = m3_add(b.inertia, inertia_offset(b.mass, v3_sub(b.center, o.center)));
    o.inertia = m3_add(ia, ib);
    return o;
}

NPH_HD NphPhysicsMassData mass_diag(f32 mass, f32 ixx, f32 iyy, f32 izz)
{
    NphPhysicsMassData m;
    m.mass = mass;
    m.center = v3_zero();
    m.inertia = m3_diag(v3(ixx, iyy, izz));
    return m;
}

// geometry.c:66-127: spheres, capsules and boxes in f64, rounded once.
NPH_HD NphPhysicsMassData geom_mass(const NphGeom *g, f32 density)
{
    const f64 pi = 3.14159265358979323846;
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE: {
        const f64 r = (f64)g->a;
        const f64 mass = (f64)density * (4.0 / 3.0) * pi * r * r * r;
        const f64 i = 0.4 * mass * r * r;
        return mass_diag((f32)mass, (f32)i, (f32)i, (f32)i);
    }
    case NPH_PHYSICS_SHAPE_CAPSULE: {
        const f64 h = (f64)g->a;
        const f64 r = (f64)g->b;
        const f64 rho = (f64)density;
        const f64 m_cyl = rho * pi * r * r * (2.0 * h);
        const f64 m_hemi = rho * (2.0 / 3.0) * pi * r * r * r;
        const f64 axial = m_cyl * r * r * 0.5 + 2.0 * (0.4 * m_hemi * r * r);
        const f64 cyl_perp = m_cyl * ((2.0 * h) * (2.0 * h) / 12.0 + r * r * 0.25);
        const f64 hemi_perp = 0.4 * m_hemi * r * r + m_hemi * (h * h + 0.75 * h * r);
        const f64 perp = cyl_perp + 2.0 * hemi_perp;
        return mass_diag((f32)(m_cyl + 2.0 * m_hemi), (f32)perp, (f32)axial, (f32)perp);
    }
    case NPH_PHYSICS_SHAPE_BOX: {
        const f64 x = (f64)g->a;
        const f64 y = (f64)g->b;
        const f64 z = (f64)g->c;
        const f64 mass = (f64)density * 8.0 * x * y * z;
        const f64 k = mass / 3.0;
        return mass_diag((f32)mass, (f32)(k * (y * y + z * z)), (f32)(k * (x * x + z * z)),
                         (f32)(k * (x * x + y * y)));
    }
    case NPH_PHYSICS_SHAPE_HULL: {
        NphPhysicsMassData m;
        m.mass = g->hull->volume * density;
        m.center = g->hull->centroid;
        m.inertia = m3_scale(g->hull->inertia, density);
        return m;
    }
    default:
        return mass_diag(0.0F, 0.0F, 0.0F, 0.0F);
    }
}

// geometry.c:180-190 (Arvo 1990)
NPH_HD NphAABB transform_box(NphAABB b, NphPhysicsPose pose)
{
    const NphMat3 m = m3_from_quat(pose.rotation);
    const NphVec3 h = v3_scale(v3_sub(b.max, b.min), 0.5F);
    const NphVec3 c = pose_point(pose, aabb_center(b));
    const NphVec3 e = v3_add(v3_add(v3_scale(v3_abs(m.c[0]), h.x), v3_scale(v3_abs(m.c[1]), h.y)),
                             v3_scale(v3_abs(m.c[2]), h.z));
    return aabb_make(v3_sub(c, e), v3_add(c, e));
}

// geometry.c:192-229
NPH_HD NphAABB geom_aabb(const NphGeom *g, NphPhysicsPose pose)
{
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE: {
        const NphVec3 r = v3_splat(g->a);
        return aabb_make(v3_sub(pose.position, r), v3_add(pose.position, r));
    }
    case NPH_PHYSICS_SHAPE_CAPSULE: {
        const NphVec3 axis = v3_scale(q_axis_y(pose.rotation), g->a);
        const NphVec3 a = v3_sub(pose.position, axis);
        const NphVec3 b = v3_add(pose.position, axis);
        const NphVec3 r = v3_splat(g->b);
        return aabb_make(v3_sub(v3_min(a, b), r), v3_add(v3_max(a, b), r));
    }
    case NPH_PHYSICS_SHAPE_BOX: {
        const NphVec3 h = v3(g->a, g->b, g->c);
        return transform_box(aabb_make(v3_neg(h), h), pose);
    }
    case NPH_PHYSICS_SHAPE_MESH:
        return transform_box(g->mesh->bounds, pose);
    default: {
        const NphPhysicsHull *hull = g->hull;
        const NphVec3 *v = hull_vertices(hull);
        const NphVec3 p0 = pose_point(pose, v[0]);
        NphAABB box = aabb_make(p0, p0);
        for (u32 i = 1U; i < hull->vertex_count; ++i) {
            const NphVec3 p = pose_point(pose, v[i]);
            box.min = v3_min(box.min, p);
            box.max = v3_max(box.max, p);
        }
        return box;
    }
    }
}

// geometry.c:397-441 (convex types)
NPH_HD f32 geom_inner_radius(const NphGeom *g)
{
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        return g->a;
    case NPH_PHYSICS_SHAPE_CAPSULE:
        return g->b;
    case NPH_PHYSICS_SHAPE_BOX:
        return nph_minf(g->a, nph_minf(g->b, g->c));
    case NPH_PHYSICS_SHAPE_HULL:
        return g->hull->inner_radius;
    default:
        return 0.0F;
    }
}

NPH_HD f32 geom_outer_radius(const NphGeom *g)
{
    switch (g->type) {
    case NPH_PHYSICS_SHAPE_SPHERE:
        return g->a;
    case NPH_PHYSICS_SHAPE_CAPSULE:
        return g->a + g->b;
    case NPH_PHYSICS_SHAPE_BOX:
        return v3_len(v3(g->a, g->b, g->c));
    case NPH_PHYSICS_SHAPE_HULL:
        return g->hull->outer_radius;
    case NPH_PHYSICS_SHAPE_MESH: {
        // The farthest corner of the vertices' box.
        const NphAABB b = g->mesh->bounds;
        return v3_len(v3_max(v3_abs(b.min), v3_abs(b.max)));
    }
    default:
        return 0.0F;
    }
}

#endif // NPH_GEOMETRY_H
