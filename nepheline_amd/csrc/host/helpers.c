// SPDX-License-Identifier: Apache-2.0
// The small helpers the reference's headers define inline (types.h:80-130,
// geometry.h:49-120, hull.h:80-110, mesh.h:60-150), exported here so that a
// caller of this library's header, or a binding that cannot see inline
// code, finds them with the same results: pose arithmetic on the
// reference's quaternion formulas (nph_math.h), geometry constructors,
// and the hull and mesh blocks' array accessors.
#include "world_impl.h"

NphPhysicsPose nph_physics_pose(NphVec3 position, NphQuat rotation)
{
    return pose_make(position, rotation);
}

NphPhysicsPose nph_physics_pose_identity(void) { return pose_identity(); }

NphVec3 nph_physics_pose_transform_point(NphPhysicsPose t, NphVec3 p) { return pose_point(t, p); }

NphVec3 nph_physics_pose_transform_vector(NphPhysicsPose t, NphVec3 v)
{
    return q_rotate(t.rotation, v);
}

NphVec3 nph_physics_pose_inverse_transform_point(NphPhysicsPose t, NphVec3 p)
{
    return q_rotate_inv(t.rotation, v3_sub(p, t.position));
}

NphVec3 nph_physics_pose_inverse_transform_vector(NphPhysicsPose t, NphVec3 v)
{
    return q_rotate_inv(t.rotation, v);
}

// a then b: b's frame carried by a.
NphPhysicsPose nph_physics_pose_mul(NphPhysicsPose a, NphPhysicsPose b)
{
    return pose_make(pose_point(a, b.position), q_mul(a.rotation, b.rotation));
}

// b in a's frame.
NphPhysicsPose nph_physics_pose_inv_mul(NphPhysicsPose a, NphPhysicsPose b)
{
    return pose_make(nph_physics_pose_inverse_transform_point(a, b.position),
                     q_mul(q_conj(a.rotation), b.rotation));
}

NphPhysicsPose nph_physics_pose_inverse(NphPhysicsPose t)
{
    const NphQuat back = q_conj(t.rotation);
    return pose_make(v3_neg(q_rotate(back, t.position)), back);
}

bool nph_physics_shape_type_is_convex(NphPhysicsShapeType type)
{
    return (u32)type <= (u32)NPH_PHYSICS_SHAPE_HULL;
}

static NphPhysicsGeometry geometry_of(NphPhysicsShapeType type)
{
    NphPhysicsGeometry g;
    memset(&g, 0, sizeof g);
    g.type = type;
    return g;
}

NphPhysicsGeometry nph_physics_geometry_sphere(f32 radius)
{
    NphPhysicsGeometry g = geometry_of(NPH_PHYSICS_SHAPE_SPHERE);
    g.sphere.radius = radius;
    return g;
}

NphPhysicsGeometry nph_physics_geometry_capsule(f32 half_height, f32 radius)
{
    NphPhysicsGeometry g = geometry_of(NPH_PHYSICS_SHAPE_CAPSULE);
    g.capsule.half_height = half_height;
    g.capsule.radius = radius;
    return g;
}

NphPhysicsGeometry nph_physics_geometry_box(NphVec3 half_extents)
{
    NphPhysicsGeometry g = geometry_of(NPH_PHYSICS_SHAPE_BOX);
    g.box.half_extents = half_extents;
    return g;
}

NphPhysicsGeometry nph_physics_geometry_hull(const NphPhysicsHull *hull)
{
    NphPhysicsGeometry g = geometry_of(NPH_PHYSICS_SHAPE_HULL);
    g.hull = hull;
    return g;
}

NphPhysicsGeometry nph_physics_geometry_mesh(const NphPhysicsMesh *mesh)
{
    NphPhysicsGeometry g = geometry_of(NPH_PHYSICS_SHAPE_MESH);
    g.mesh = mesh;
    return g;
}

NphPhysicsGeometry nph_physics_geometry_heightfield(const NphPhysicsHeightfield *h)
{
    NphPhysicsGeometry g = geometry_of(NPH_PHYSICS_SHAPE_HEIGHTFIELD);
    g.heightfield = h;
    return g;
}

const void *nph_physics_hull_bytes(const NphPhysicsHull *h, u32 offset)
{
    return (const u8 *)h + offset;
}

const NphVec3 *nph_physics_hull_vertices(const NphPhysicsHull *h) { return hull_vertices(h); }

const NphPlane *nph_physics_hull_planes(const NphPhysicsHull *h)
{
    return (const NphPlane *)nph_physics_hull_bytes(h, h->planes_offset);
}

const u16 *nph_physics_hull_faces(const NphPhysicsHull *h)
{
    return (const u16 *)nph_physics_hull_bytes(h, h->faces_offset);
}

const NphPhysicsHullEdge *nph_physics_hull_edges(const NphPhysicsHull *h) { return hull_edges(h); }

// A leaf child: the flag, the first triangle, the count less one.
u32 nph_physics_mesh_leaf(u32 first, u32 count)
{
    return NPH_PHYSICS_MESH_CHILD_LEAF | (first << 2U) | (count - 1U);
}

bool nph_physics_mesh_child_is_leaf(u32 child)
{
    return child != NPH_PHYSICS_MESH_CHILD_EMPTY && (child & NPH_PHYSICS_MESH_CHILD_LEAF) != 0U;
}

u32 nph_physics_mesh_leaf_first(u32 child) { return (child & ~NPH_PHYSICS_MESH_CHILD_LEAF) >> 2U; }

u32 nph_physics_mesh_leaf_count(u32 child) { return (child & 3U) + 1U; }

const void *nph_physics_mesh_bytes(const NphPhysicsMesh *m, u32 offset)
{
    return (const u8 *)m + offset;
}

const NphPhysicsMeshNode *nph_physics_mesh_nodes(const NphPhysicsMesh *m)
{
    return (const NphPhysicsMeshNode *)nph_physics_mesh_bytes(m, m->nodes_offset);
}

const NphVec3 *nph_physics_mesh_vertices(const NphPhysicsMesh *m) { return mesh_vertices(m); }

const u32 *nph_physics_mesh_triangles(const NphPhysicsMesh *m)
{
    return (const u32 *)nph_physics_mesh_bytes(m, m->triangles_offset);
}

const u8 *nph_physics_mesh_edges(const NphPhysicsMesh *m) { return mesh_edges(m); }

const u32 *nph_physics_mesh_originals(const NphPhysicsMesh *m)
{
    return (const u32 *)nph_physics_mesh_bytes(m, m->originals_offset);
}
