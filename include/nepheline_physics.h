// SPDX-License-Identifier: Apache-2.0
// nepheline-mi355x: the C ABI of the rigid-body world step, drop-in for the
// reference engine's physics API (engine/physics/include/nepheline/physics/).
//
// Every declaration below replaces the reference function or type cited
// next to it, with the same name, argument meaning, struct layout and error
// behaviour, so a caller written against the reference links against
// libnepheline_amd.so unchanged.  The arithmetic of a step runs on an
// MI355X (gfx950) as HIP kernels; the serial bookkeeping around it (handles,
// contact records, the constraint graph, the broadphase tree) runs on the
// host in C17, as in the reference.
//
// Additions of this library, with no counterpart in the reference's public
// headers, are marked "EXTENSION".
#ifndef NEPHELINE_PHYSICS_H
#define NEPHELINE_PHYSICS_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// ---------------------------------------------------------------------------
// Base types (engine/core/include/nepheline/core/types.h; math/vec.h:21-67,
// math/quat.h:20-24, math/mat.h:24-33, math/geometry.h)
// ---------------------------------------------------------------------------
#ifndef NPH_BASE_TYPES_DEFINED
#define NPH_BASE_TYPES_DEFINED
typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;
typedef float f32;
typedef double f64;
typedef int32_t b32;
#endif

typedef struct NphVec3 {
    f32 x, y, z;
} NphVec3; // 12 B, unpadded (vec.h:63)
typedef struct NphDVec3 {
    f64 x, y, z;
} NphDVec3; // 24 B (vec.h:67)
typedef struct NphQuat {
    f32 x, y, z, w;
} NphQuat; // 16 B (quat.h:24)
typedef struct NphMat3 {
    NphVec3 c[3]; // columns (mat.h:24-33)
} NphMat3;
typedef struct NphAABB {
    NphVec3 min, max;
} NphAABB; // 24 B
typedef struct NphPlane {
    NphVec3 normal;
    f32 d; // n.p + d = 0
} NphPlane;

// engine/core/include/nepheline/core/result.h:15-37 (same values).
typedef enum NphResult {
    NPH_RESULT_OK = 0,
    NPH_RESULT_OUT_OF_MEMORY,
    NPH_RESULT_INVALID_ARGUMENT,
    NPH_RESULT_NOT_FOUND,
    NPH_RESULT_ALREADY_EXISTS,
    NPH_RESULT_FULL,
    NPH_RESULT_EMPTY,
    NPH_RESULT_OVERFLOW,
    NPH_RESULT_UNSUPPORTED,
    NPH_RESULT_CORRUPT,
    NPH_RESULT_VERSION_MISMATCH,
    NPH_RESULT_TYPE_MISMATCH,
    NPH_RESULT_PARSE_ERROR,
    NPH_RESULT_END_OF_STREAM,
    NPH_RESULT_IO_ERROR,
    NPH_RESULT_ACCESS_DENIED,
    NPH_RESULT_BUSY,
    NPH_RESULT_DEVICE_LOST,
    NPH_RESULT_INVALID_STATE,
    NPH_RESULT_UNKNOWN,
    NPH_RESULT_COUNT
} NphResult;

// engine/core/include/nepheline/core/alloc.h:55-67: the allocator vtable the
// world allocates its host memory through.
typedef struct NphCallsite {
    const char *file;
    u32 line;
} NphCallsite;
typedef struct NphAllocator NphAllocator;
typedef struct NphAllocatorVtable {
    void *(*alloc)(NphAllocator *self, u64 size, u32 align, NphCallsite site);
    void *(*realloc)(NphAllocator *self, void *ptr, u64 old_size, u64 new_size, u32 align,
                     NphCallsite site);
    void (*free)(NphAllocator *self, void *ptr, u64 size, u32 align, NphCallsite site);
} NphAllocatorVtable;
struct NphAllocator {
    const NphAllocatorVtable *vtable;
    const char *name;
};

// engine/jobs/include/nepheline/jobs/jobs.h: opaque; the step accepts it
// for signature compatibility and ignores it (the device does the parallel
// work), as the reference's results do not depend on it (world.h:45-46).
typedef struct NphJobSystem NphJobSystem;

// EXTENSION: a malloc-backed allocator for callers without an engine core
// (the reference's nph_allocator_system, core/alloc.h:143, lives in a module
// this library does not replace).
NphAllocator *nph_amd_allocator_system(void);

// ---------------------------------------------------------------------------
// Tolerances and poses (physics/types.h:38-70)
// ---------------------------------------------------------------------------
#define NPH_PHYSICS_LINEAR_SLOP          0.005F
#define NPH_PHYSICS_MAX_COORDINATE       1.0e5F
#define NPH_PHYSICS_SPECULATIVE_DISTANCE (4.0F * NPH_PHYSICS_LINEAR_SLOP)
#define NPH_PHYSICS_AABB_MARGIN          0.1F
#define NPH_PHYSICS_MAX_OBJECTS          (1U << 20U)
#define NPH_PHYSICS_MANIFOLD_MAX_POINTS  4U

typedef struct NphPhysicsPose {
    NphVec3 position;
    NphQuat rotation;
} NphPhysicsPose; // 28 B (types.h:70)

typedef struct NphPhysicsMassData {
    f32 mass;
    NphVec3 center;
    NphMat3 inertia;
} NphPhysicsMassData; // types.h

// ---------------------------------------------------------------------------
// Hull blocks (physics/hull.h:49-78): caller-owned, immutable, pointer-free;
// the same bytes the reference's nph_physics_hull_build writes.
// ---------------------------------------------------------------------------
typedef struct NphPhysicsHullEdge {
    u16 next;
    u16 twin;
    u16 origin;
    u16 face;
} NphPhysicsHullEdge; // 8 B (hull.h:56)

typedef struct NphPhysicsHull {
    u32 byte_size;
    u32 vertex_count;
    u32 edge_count;
    u32 face_count;
    u32 vertices_offset;
    u32 planes_offset;
    u32 faces_offset;
    u32 edges_offset;
    NphAABB bounds;
    NphVec3 centroid;
    f32 volume;
    NphMat3 inertia;
    f32 inner_radius;
    f32 outer_radius;
    f32 tolerance;
    u32 reserved[2];
} NphPhysicsHull; // 128 B (hull.h:78)

// The builder and the block queries (hull.h:41-170): host code, the same
// bytes as the reference's for the same input.
#define NPH_PHYSICS_HULL_MAX_VERTICES         256U
#define NPH_PHYSICS_HULL_DEFAULT_MAX_VERTICES 64U

typedef struct NphPhysicsHullDesc {
    const NphVec3 *points;
    u32 point_count;  // at least 4
    u32 max_vertices; // 0: NPH_PHYSICS_HULL_DEFAULT_MAX_VERTICES; 4..256
    f32 tolerance;    // 0: 3 * FLT_EPSILON * (max|x| + max|y| + max|z|)
} NphPhysicsHullDesc; // hull.h:111-132

// Quickhull over desc->points into one block from `alloc` (which also holds
// the build's temporaries); built again at 4x, then 16x the tolerance when
// the first hull fails validation. Replaces hull.h:135-136 (hull.c:1141-1197).
NphResult nph_physics_hull_build(const NphPhysicsHullDesc *desc, NphAllocator *alloc,
                                 NphPhysicsHull **out);
// The box hull, made directly (hull.h:142-143, hull.c:321-346).
NphResult nph_physics_hull_make_box(NphVec3 half_extents, NphAllocator *alloc,
                                    NphPhysicsHull **out);
// Releases a block of the two above; NULL is ignored (hull.h:147, hull.c:311-319).
void nph_physics_hull_free(NphPhysicsHull *hull, NphAllocator *alloc);
// The first vertex farthest along `direction` (hull.h:150, hull.c:1203-1216).
u32 nph_physics_hull_support(const NphPhysicsHull *hull, NphVec3 direction);
// The vertex count of face `face` (hull.h:153, hull.c:1218-1230).
u32 nph_physics_hull_face_vertex_count(const NphPhysicsHull *hull, u32 face);
// Checks a block of `size` bytes: layout, counts, Euler, header numbers,
// twins, closed convex loops, every vertex behind every plane, within
// `tolerance` (0: the hull's own). Replaces hull.h:166 (hull.c:1254-1377).
bool nph_physics_hull_validate(const void *block, u32 size, f32 tolerance);

// ---------------------------------------------------------------------------
// Mesh blocks (physics/mesh.h:44-150): caller-owned, immutable, pointer-free;
// the same bytes the reference's nph_physics_mesh_build writes (a four-wide
// tree over triangles in leaf order, active edges, input indices).
// ---------------------------------------------------------------------------
#define NPH_PHYSICS_MESH_MAX_TRIANGLES  (1U << 28U)
#define NPH_PHYSICS_MESH_LEAF_TRIANGLES 4U
#define NPH_PHYSICS_MESH_EDGE_0         (1U << 0U)
#define NPH_PHYSICS_MESH_EDGE_1         (1U << 1U)
#define NPH_PHYSICS_MESH_EDGE_2         (1U << 2U)
#define NPH_PHYSICS_MESH_EDGES          7U
#define NPH_PHYSICS_MESH_TWO_SIDED      (1U << 0U)
#define NPH_PHYSICS_MESH_CHILD_LEAF     0x80000000U
#define NPH_PHYSICS_MESH_CHILD_EMPTY    0xFFFFFFFFU
#define NPH_PHYSICS_MESH_MAX_MANIFOLDS  4U // collide.h:152

typedef struct NphPhysicsMeshNode {
    f32 min_x[4];
    f32 min_y[4];
    f32 min_z[4];
    f32 max_x[4];
    f32 max_y[4];
    f32 max_z[4];
    u32 child[4];
} NphPhysicsMeshNode; // 112 B (mesh.h:95)

typedef struct NphPhysicsMesh {
    u32 byte_size;
    u32 vertex_count;
    u32 triangle_count;
    u32 node_count;
    u32 nodes_offset;
    u32 vertices_offset;
    u32 triangles_offset;
    u32 edges_offset;
    u32 originals_offset;
    u32 flags;
    NphAABB bounds;
} NphPhysicsMesh; // 64 B (mesh.h:113)

// The builder (mesh.h:154-201): host code, the same bytes as the
// reference's for the same input.
typedef struct NphPhysicsMeshDesc {
    const NphVec3 *vertices;
    u32 vertex_count;
    const u32 *indices; // three per triangle
    u32 triangle_count;
    f32 weld_tolerance;    // 0: 1e-6 of the largest coordinate; negative: exact copies only
    f32 active_edge_angle; // radians; 0: NPH_PHYSICS_MESH_DEFAULT_ACTIVE_ANGLE
    b32 two_sided;
} NphPhysicsMeshDesc; // mesh.h:158-176
#define NPH_PHYSICS_MESH_DEFAULT_ACTIVE_ANGLE (5.0F * 3.14159265F / 180.0F)

// Welds, drops degenerate triangles, marks active edges, builds the binned
// SAH tree and collapses it four wide, into one block from `alloc`.
// Replaces mesh.h:189-190 (mesh.c:675-798).
NphResult nph_physics_mesh_build(const NphPhysicsMeshDesc *desc, NphAllocator *alloc,
                                 NphPhysicsMesh **out);
// Releases a block of nph_physics_mesh_build; NULL is ignored (mesh.h:193).
void nph_physics_mesh_free(NphPhysicsMesh *mesh, NphAllocator *alloc);
// Checks a block of `size` bytes: layout, counts, indices, flags, the tree
// (order, boxes, leaves covering the triangles in order, depth). Replaces
// mesh.h:205 (mesh.c:872-996).
bool nph_physics_mesh_validate(const void *block, u32 size);
// Triangle t's vertices in leaf order (mesh.h:151, mesh.c:34-41).
void nph_physics_mesh_triangle(const NphPhysicsMesh *m, u32 t, NphVec3 out[3]);

// The mesh's box and ray queries (mesh.h:207-227), host code.
typedef bool (*NphPhysicsMeshVisit)(void *user, u32 triangle);
// Visits the triangles whose boxes overlap `box`, in tree order, until a
// visit returns false. Replaces mesh.h:213-214 (mesh.c:1013-1049).
void nph_physics_mesh_overlap(const NphPhysicsMesh *m, NphAABB box, NphPhysicsMeshVisit visit,
                              void *user);
typedef struct NphPhysicsMeshHit {
    f32 fraction;
    NphVec3 normal; // towards the side the ray came from
    u32 triangle;   // in leaf order
    u32 original;   // the input's index
    b32 back;       // hit from behind (two-sided meshes only)
} NphPhysicsMeshHit; // mesh.h:216-222
// The closest triangle the segment origin -> origin + delta crosses
// (watertight test; one-sided meshes from the front only). Replaces
// mesh.h:224-225 (mesh.c:1104-1192).
bool nph_physics_mesh_raycast(const NphPhysicsMesh *m, NphVec3 origin, NphVec3 delta,
                              NphPhysicsMeshHit *hit);

typedef struct NphPhysicsHeightfield NphPhysicsHeightfield; // heightfield.h (not supported)

// ---------------------------------------------------------------------------
// Geometry (physics/geometry.h:37-78)
// ---------------------------------------------------------------------------
typedef enum NphPhysicsShapeType {
    NPH_PHYSICS_SHAPE_SPHERE = 0,
    NPH_PHYSICS_SHAPE_CAPSULE = 1,
    NPH_PHYSICS_SHAPE_BOX = 2,
    NPH_PHYSICS_SHAPE_HULL = 3,
    NPH_PHYSICS_SHAPE_MESH = 4,
    NPH_PHYSICS_SHAPE_HEIGHTFIELD = 5,
    NPH_PHYSICS_SHAPE_TYPE_COUNT = 6,
} NphPhysicsShapeType;

typedef struct NphPhysicsSphere {
    f32 radius;
} NphPhysicsSphere;
typedef struct NphPhysicsCapsule {
    f32 half_height;
    f32 radius;
} NphPhysicsCapsule;
typedef struct NphPhysicsBox {
    NphVec3 half_extents;
} NphPhysicsBox;

typedef struct NphPhysicsGeometry {
    NphPhysicsShapeType type;
    union {
        NphPhysicsSphere sphere;
        NphPhysicsCapsule capsule;
        NphPhysicsBox box;
        const NphPhysicsHull *hull;
        const NphPhysicsMesh *mesh;
        const NphPhysicsHeightfield *heightfield;
    };
} NphPhysicsGeometry;

// The helpers the reference's headers define inline, exported with the same
// results (types.h:80-130, geometry.h:49-120, hull.h:80-110,
// mesh.h:60-150): pose arithmetic, geometry constructors, the blocks'
// array accessors and mesh child codes.
NphPhysicsPose nph_physics_pose(NphVec3 position, NphQuat rotation);
NphPhysicsPose nph_physics_pose_identity(void);
NphVec3 nph_physics_pose_transform_point(NphPhysicsPose t, NphVec3 p);
NphVec3 nph_physics_pose_transform_vector(NphPhysicsPose t, NphVec3 v);
NphVec3 nph_physics_pose_inverse_transform_point(NphPhysicsPose t, NphVec3 p);
NphVec3 nph_physics_pose_inverse_transform_vector(NphPhysicsPose t, NphVec3 v);
NphPhysicsPose nph_physics_pose_mul(NphPhysicsPose a, NphPhysicsPose b);
NphPhysicsPose nph_physics_pose_inv_mul(NphPhysicsPose a, NphPhysicsPose b);
NphPhysicsPose nph_physics_pose_inverse(NphPhysicsPose t);
bool nph_physics_shape_type_is_convex(NphPhysicsShapeType type);
NphPhysicsGeometry nph_physics_geometry_sphere(f32 radius);
NphPhysicsGeometry nph_physics_geometry_capsule(f32 half_height, f32 radius);
NphPhysicsGeometry nph_physics_geometry_box(NphVec3 half_extents);
NphPhysicsGeometry nph_physics_geometry_hull(const NphPhysicsHull *hull);
NphPhysicsGeometry nph_physics_geometry_mesh(const NphPhysicsMesh *mesh);
NphPhysicsGeometry nph_physics_geometry_heightfield(const NphPhysicsHeightfield *h);
const void *nph_physics_hull_bytes(const NphPhysicsHull *h, u32 offset);
const NphVec3 *nph_physics_hull_vertices(const NphPhysicsHull *h);
const NphPlane *nph_physics_hull_planes(const NphPhysicsHull *h);
const u16 *nph_physics_hull_faces(const NphPhysicsHull *h);
const NphPhysicsHullEdge *nph_physics_hull_edges(const NphPhysicsHull *h);
u32 nph_physics_mesh_leaf(u32 first, u32 count);
bool nph_physics_mesh_child_is_leaf(u32 child);
u32 nph_physics_mesh_leaf_first(u32 child);
u32 nph_physics_mesh_leaf_count(u32 child);
const void *nph_physics_mesh_bytes(const NphPhysicsMesh *m, u32 offset);
const NphPhysicsMeshNode *nph_physics_mesh_nodes(const NphPhysicsMesh *m);
const NphVec3 *nph_physics_mesh_vertices(const NphPhysicsMesh *m);
const u32 *nph_physics_mesh_triangles(const NphPhysicsMesh *m);
const u8 *nph_physics_mesh_edges(const NphPhysicsMesh *m);
const u32 *nph_physics_mesh_originals(const NphPhysicsMesh *m);

// geometry.c:57-140 (mass closed forms), :178-229 (boxes).
NphPhysicsMassData nph_physics_geometry_mass(const NphPhysicsGeometry *g, f32 density);
NphAABB nph_physics_geometry_aabb(const NphPhysicsGeometry *g, NphPhysicsPose pose);
bool nph_physics_geometry_is_valid(const NphPhysicsGeometry *g);

// types.h:132-142: mass properties carried to a parent frame, added, and
// the parallel-axis tensor of a point mass.
NphPhysicsMassData nph_physics_mass_transform(NphPhysicsMassData m, NphPhysicsPose pose);
NphPhysicsMassData nph_physics_mass_add(NphPhysicsMassData a, NphPhysicsMassData b);
NphMat3 nph_physics_inertia_offset(f32 mass, NphVec3 r);

// geometry.h:125-151: the farthest point along a direction (rounding
// included), the inner and outer radii about the centre of mass, and the
// outward normal of the geometry's own surface at a point.
NphVec3 nph_physics_geometry_support(const NphPhysicsGeometry *g, NphVec3 direction);
f32 nph_physics_geometry_inner_radius(const NphPhysicsGeometry *g);
f32 nph_physics_geometry_outer_radius(const NphPhysicsGeometry *g);
NphVec3 nph_physics_geometry_surface_normal(const NphPhysicsGeometry *g, NphVec3 point,
                                            u32 triangle, NphVec3 toward);

// geometry.h:163-181: a convex shape as the hull of a few points rounded by
// a radius (spheres one point, capsules two, boxes eight corners written to
// `storage`, hulls their vertices).
typedef struct NphPhysicsProxy {
    const NphVec3 *points;
    u32 count;
    f32 radius;
} NphPhysicsProxy;
#define NPH_PHYSICS_PROXY_STORAGE 8U
NphPhysicsProxy nph_physics_geometry_proxy(const NphPhysicsGeometry *g,
                                           NphVec3 storage[NPH_PHYSICS_PROXY_STORAGE]);
u32 nph_physics_proxy_support(const NphPhysicsProxy *p, NphVec3 direction);

// ---------------------------------------------------------------------------
// Pair queries (physics/collide.h:62-121): the narrow phase of one pair,
// run on the device by the same code as the world step's.
// ---------------------------------------------------------------------------
typedef struct NphPhysicsManifoldPoint {
    NphVec3 point;  // midway between the two surfaces
    f32 separation; // along the normal; negative when the shapes overlap
    u32 id;         // the features that made the point
} NphPhysicsManifoldPoint;

typedef struct NphPhysicsManifold {
    NphVec3 normal;  // unit, from shape A towards shape B
    u32 point_count; // 0: farther apart than the speculative distance
    NphPhysicsManifoldPoint points[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
} NphPhysicsManifold;

// What the previous call on the pair found (zero: no history).
typedef struct NphPhysicsCollideCache {
    u8 type;
    u8 reserved;
    u16 index_a;
    u16 index_b;
    u16 reserved2;
} NphPhysicsCollideCache;

// collide.h:96: every pair of convex types.
bool nph_physics_collide_supported(NphPhysicsShapeType a, NphPhysicsShapeType b);
// collide.h:103-106: the manifold of a at pose_a and b at pose_b, points
// at most speculative_distance apart, at most four; cache may be NULL.
// Without a HIP device: no points, and a message on stderr.
void nph_physics_collide(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                         const NphPhysicsGeometry *b, NphPhysicsPose pose_b,
                         f32 speculative_distance, NphPhysicsCollideCache *cache,
                         NphPhysicsManifold *out);
// collide.h:121: `count` candidates reduced to at most four, in place.
u32 nph_physics_manifold_reduce(NphPhysicsManifoldPoint *points, u32 count, NphVec3 normal);

// collide.h:129-168: a convex shape against one triangle and against a
// mesh (up to four manifolds, the deepest first), on the device by the
// world step's mesh narrow phase.
typedef struct NphPhysicsTriangle {
    NphVec3 vertices[3];
    u32 active_edges; // NPH_PHYSICS_MESH_EDGE_* bits
    b32 two_sided;
} NphPhysicsTriangle;
typedef struct NphPhysicsMeshManifolds {
    u32 count;
    NphPhysicsManifold manifolds[NPH_PHYSICS_MESH_MAX_MANIFOLDS];
} NphPhysicsMeshManifolds;
void nph_physics_collide_triangle(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                                  const NphPhysicsTriangle *t, f32 speculative_distance,
                                  NphPhysicsManifold *out);
void nph_physics_collide_mesh(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                              const NphPhysicsMesh *mesh, NphPhysicsPose pose_mesh,
                              f32 speculative_distance, NphPhysicsMeshManifolds *out);

// collide.h:181-235: closest points of two segments and of two proxies
// (GJK, with an optional simplex cache).
typedef struct NphPhysicsSegmentDistance {
    NphVec3 closest1;
    NphVec3 closest2;
    f32 fraction1;
    f32 fraction2;
    f32 distance_squared;
} NphPhysicsSegmentDistance;
NphPhysicsSegmentDistance nph_physics_segment_distance(NphVec3 p1, NphVec3 q1, NphVec3 p2,
                                                       NphVec3 q2);
typedef struct NphPhysicsSimplexCache {
    u32 count;
    u8 index_a[4];
    u8 index_b[4];
} NphPhysicsSimplexCache;
typedef struct NphPhysicsDistance {
    NphVec3 point_a;
    NphVec3 point_b;
    NphVec3 normal; // from A towards B; +Y when the shapes touch or overlap
    f32 distance;
    u32 iterations;
    u32 simplex_count;
} NphPhysicsDistance;
void nph_physics_distance(const NphPhysicsProxy *a, NphPhysicsPose pose_a, const NphPhysicsProxy *b,
                          NphPhysicsPose pose_b, bool use_radii, NphPhysicsSimplexCache *cache,
                          NphPhysicsDistance *out);

// collide.h:223-236: how deep two rounded proxies overlap (GJK, then an
// expanding polytope when the cores overlap); false when they do not.
typedef struct NphPhysicsPenetration {
    NphVec3 normal;  // unit, from A towards B: B moved by depth * normal separates them
    f32 depth;       // > 0
    NphVec3 point_a; // the deepest point of A in B, on A's rounded surface
    NphVec3 point_b; // the deepest point of B in A
} NphPhysicsPenetration;
bool nph_physics_penetration(const NphPhysicsProxy *a, NphPhysicsPose pose_a,
                             const NphPhysicsProxy *b, NphPhysicsPose pose_b,
                             NphPhysicsPenetration *out);

// ---------------------------------------------------------------------------
// Casts and overlaps against one geometry (collide.h:240-318, cast.c), host
// code.  Meshes: rays and overlaps; a shape cast against a mesh is not
// built here (it returns false).
// ---------------------------------------------------------------------------
#define NPH_PHYSICS_NO_TRIANGLE 0xFFFFFFFFU

typedef struct NphPhysicsCastHit {
    NphVec3 point;  // on the geometry's surface
    NphVec3 normal; // unit: the surface's there, out of it (against the motion)
    f32 fraction;   // of the motion, in [0, 1]
    u32 triangle;   // a mesh's index in its input, or NPH_PHYSICS_NO_TRIANGLE
} NphPhysicsCastHit;

bool nph_physics_raycast(const NphPhysicsGeometry *g, NphPhysicsPose pose, NphVec3 origin,
                         NphVec3 delta, NphPhysicsCastHit *out);                 // cast.c:241-257
bool nph_physics_shape_cast(const NphPhysicsGeometry *a, NphPhysicsPose pose_a, NphVec3 translation,
                            const NphPhysicsGeometry *b, NphPhysicsPose pose_b,
                            NphPhysicsCastHit *out);                             // cast.c:525-530
bool nph_physics_shape_cast_margin(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                                   NphVec3 translation, f32 margin, const NphPhysicsGeometry *b,
                                   NphPhysicsPose pose_b, NphPhysicsCastHit *out); // :532-564
bool nph_physics_overlap(const NphPhysicsGeometry *a, NphPhysicsPose pose_a,
                         const NphPhysicsGeometry *b, NphPhysicsPose pose_b);    // :570-590

// No counterpart in the reference: many pairs in one launch (pairs of
// convex shapes only); caches, when given, are read and written per pair.
typedef struct NphAmdCollideQuery {
    NphPhysicsGeometry a;
    NphPhysicsPose pose_a;
    NphPhysicsGeometry b;
    NphPhysicsPose pose_b;
    f32 speculative_distance;
    u32 pad;
} NphAmdCollideQuery;
NphResult nph_amd_collide_many(const NphAmdCollideQuery *queries, u32 count,
                               NphPhysicsManifold *out, NphPhysicsCollideCache *caches);

// No counterpart in the reference: nph_amd_collide_many for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation). It stages a call's
// pairs, manifolds, caches and hull blocks in one fixed device array of
// NPH_AMD_STATIC_SCRATCH_BYTES per GPU (held until the process ends),
// runs on the null stream and calls no allocator. A call whose pairs and
// distinct hull blocks do not fit, or that names more than
// NPH_AMD_STATIC_MAX_HULLS distinct hull blocks, returns
// NPH_RESULT_INVALID_ARGUMENT before touching the device. Every call on a
// GPU shares that GPU's array: calls that overlap on one GPU (from two
// host threads) race. Results equal nph_amd_collide_many's bit for bit.
#define NPH_AMD_STATIC_SCRATCH_BYTES (16U << 20)
#define NPH_AMD_STATIC_MAX_HULLS 256U
#define NPH_AMD_MAX_DEVICES 64U
NphResult nph_amd_collide_many_static(const NphAmdCollideQuery *queries, u32 count,
                                      NphPhysicsManifold *out, NphPhysicsCollideCache *caches);

// ---------------------------------------------------------------------------
// World (physics/world.h:57-326)
// ---------------------------------------------------------------------------
typedef struct NphPhysicsWorld NphPhysicsWorld;
typedef struct NphPhysicsBody {
    u32 value;
} NphPhysicsBody;
typedef struct NphPhysicsShape {
    u32 value;
} NphPhysicsShape;

typedef enum NphPhysicsSolver {
    NPH_PHYSICS_SOLVER_SOFT_STEP = 0,
    NPH_PHYSICS_SOLVER_TGS = 1,  // rejected: NPH_RESULT_UNSUPPORTED
    NPH_PHYSICS_SOLVER_XPBD = 2, // rejected: NPH_RESULT_UNSUPPORTED
    NPH_PHYSICS_SOLVER_COUNT = 3,
} NphPhysicsSolver;

typedef struct NphPhysicsWorldDesc {
    NphAllocator *alloc;
    NphVec3 gravity;
    NphPhysicsSolver solver;
    u32 substeps;
    f32 contact_hertz;
    f32 contact_damping_ratio;
    f32 contact_push_max_velocity;
    f32 restitution_threshold;
    f32 max_linear_speed;
    f32 joint_hertz;
    f32 joint_damping_ratio;
    f32 time_to_sleep;
    b32 enable_sleep;
    b32 enable_continuous;
    u32 body_capacity;
    u32 shape_capacity;
    u32 contact_capacity;
} NphPhysicsWorldDesc; // world.h:104-143

typedef struct NphPhysicsContactBeginEvent {
    NphPhysicsShape shape_a;
    NphPhysicsShape shape_b;
} NphPhysicsContactBeginEvent;
typedef struct NphPhysicsContactEndEvent {
    NphPhysicsShape shape_a;
    NphPhysicsShape shape_b;
} NphPhysicsContactEndEvent;
typedef struct NphPhysicsContactHitEvent {
    NphPhysicsShape shape_a;
    NphPhysicsShape shape_b;
    NphDVec3 point;
    NphVec3 normal;
    f32 approach_speed;
} NphPhysicsContactHitEvent;
typedef struct NphPhysicsBodyMoveEvent {
    NphPhysicsBody body;
    u32 fell_asleep;
    u64 user_data;
    NphDVec3 position;
    NphQuat rotation;
    NphVec3 linear_velocity;
    NphVec3 angular_velocity;
} NphPhysicsBodyMoveEvent;
typedef struct NphPhysicsSensorEvent {
    NphPhysicsShape sensor;
    NphPhysicsShape visitor;
    u64 sensor_user_data;
    u64 visitor_user_data;
} NphPhysicsSensorEvent;
typedef struct NphPhysicsEvents {
    const NphPhysicsContactBeginEvent *begin;
    u32 begin_count;
    const NphPhysicsContactEndEvent *end;
    u32 end_count;
    const NphPhysicsContactHitEvent *hit;
    u32 hit_count;
    const NphPhysicsBodyMoveEvent *move;
    u32 move_count;
    const NphPhysicsSensorEvent *sensor_begin;
    u32 sensor_begin_count;
    const NphPhysicsSensorEvent *sensor_end;
    u32 sensor_end_count;
} NphPhysicsEvents; // world.h:175-240

typedef struct NphPhysicsStats {
    u32 bodies;
    u32 awake_bodies;
    u32 shapes;
    u32 contacts;
    u32 touching_contacts;
    u32 manifold_points;
    u32 joints;
    u32 islands;
    u32 awake_islands;
    u32 colors_used;
    u32 overflow_constraints;
    u32 moved_proxies;
    u32 new_pairs;
    u32 swept_bodies;
    u32 impact_bodies;
    u32 sensors;
    u32 sensor_visitors;
    u32 workers;
    u64 steps;
    f64 step_ms;
    f64 pairs_ms;
    f64 collide_ms;
    f64 solve_ms;
    f64 finalize_ms;
    f64 sleep_ms;
    f64 sensors_ms;
} NphPhysicsStats; // world.h:250-280

NphPhysicsWorldDesc nph_physics_world_desc_default(void);                   // world.c:85-105
NphResult nph_physics_world_create(const NphPhysicsWorldDesc *desc,
                                   NphPhysicsWorld **out);                  // world.c:112-215
void nph_physics_world_destroy(NphPhysicsWorld *world);                     // world.c:217
NphResult nph_physics_world_step(NphPhysicsWorld *world, f32 dt,
                                 NphJobSystem *jobs);                       // world.c:307-383
void nph_physics_world_set_gravity(NphPhysicsWorld *world, NphVec3 gravity); // world.h:170
NphVec3 nph_physics_world_gravity(const NphPhysicsWorld *world);             // world.h:171
NphPhysicsEvents nph_physics_world_events(const NphPhysicsWorld *world);     // world.h:240
void nph_physics_world_set_hit_threshold(NphPhysicsWorld *world, f32 speed); // world.h:246
NphPhysicsStats nph_physics_world_stats(const NphPhysicsWorld *world);       // world.h:280
u64 nph_physics_world_hash(const NphPhysicsWorld *world);                    // world.c:436-462
bool nph_physics_world_validate(const NphPhysicsWorld *world);               // world.c:482-695

// ---------------------------------------------------------------------------
// Bodies (physics/body.h:45-164)
// ---------------------------------------------------------------------------
typedef enum NphPhysicsBodyType {
    NPH_PHYSICS_BODY_STATIC = 0,
    NPH_PHYSICS_BODY_KINEMATIC = 1,
    NPH_PHYSICS_BODY_DYNAMIC = 2,
} NphPhysicsBodyType;

enum {
    NPH_PHYSICS_LOCK_LINEAR_X = 1U << 0U,
    NPH_PHYSICS_LOCK_LINEAR_Y = 1U << 1U,
    NPH_PHYSICS_LOCK_LINEAR_Z = 1U << 2U,
    NPH_PHYSICS_LOCK_ANGULAR_X = 1U << 3U,
    NPH_PHYSICS_LOCK_ANGULAR_Y = 1U << 4U,
    NPH_PHYSICS_LOCK_ANGULAR_Z = 1U << 5U,
};

typedef struct NphPhysicsBodyDesc {
    NphPhysicsBodyType type;
    NphDVec3 position;
    NphQuat rotation;
    NphVec3 linear_velocity;
    NphVec3 angular_velocity;
    f32 linear_damping;
    f32 angular_damping;
    f32 gravity_scale;
    f32 sleep_threshold;
    u32 locks;
    b32 enable_sleep;
    b32 awake;
    b32 gyroscopic;
    b32 bullet;
    u64 user_data;
} NphPhysicsBodyDesc; // body.h:45-72

NphPhysicsBodyDesc nph_physics_body_desc_default(void);                          // body.c:12-24
NphResult nph_physics_body_create(NphPhysicsWorld *world, const NphPhysicsBodyDesc *desc,
                                  NphPhysicsBody *out);                          // body.c:56-130
bool nph_physics_body_is_valid(const NphPhysicsWorld *world, NphPhysicsBody body);
NphPhysicsBodyType nph_physics_body_type(const NphPhysicsWorld *world, NphPhysicsBody body);
NphDVec3 nph_physics_body_position(const NphPhysicsWorld *world, NphPhysicsBody body);
NphQuat nph_physics_body_rotation(const NphPhysicsWorld *world, NphPhysicsBody body);
NphDVec3 nph_physics_body_center_of_mass(const NphPhysicsWorld *world, NphPhysicsBody body);
NphVec3 nph_physics_body_linear_velocity(const NphPhysicsWorld *world, NphPhysicsBody body);
NphVec3 nph_physics_body_angular_velocity(const NphPhysicsWorld *world, NphPhysicsBody body);
void nph_physics_body_set_linear_velocity(NphPhysicsWorld *world, NphPhysicsBody body,
                                          NphVec3 velocity);
void nph_physics_body_set_angular_velocity(NphPhysicsWorld *world, NphPhysicsBody body,
                                           NphVec3 velocity);
NphPhysicsMassData nph_physics_body_mass_data(const NphPhysicsWorld *world, NphPhysicsBody body);
u64 nph_physics_body_user_data(const NphPhysicsWorld *world, NphPhysicsBody body);
bool nph_physics_body_is_awake(const NphPhysicsWorld *world, NphPhysicsBody body);
// world.h:70-77 (inline there; exported here).
bool nph_physics_body_is_null(NphPhysicsBody b);
bool nph_physics_shape_is_null(NphPhysicsShape s);
// body.c:132-180: its joints, contacts (neighbours that touched wake) and
// shapes go first.
void nph_physics_body_destroy(NphPhysicsWorld *world, NphPhysicsBody body);
// body.c:236-262: a teleport; the shapes' proxies move now.
void nph_physics_body_set_transform(NphPhysicsWorld *world, NphPhysicsBody body, NphDVec3 position,
                                    NphQuat rotation);
// body.c:316-386: forces and torques until the end of the next step;
// impulses now.  Dynamic bodies only.
void nph_physics_body_apply_force(NphPhysicsWorld *world, NphPhysicsBody body, NphVec3 force,
                                  NphDVec3 point);
void nph_physics_body_apply_force_at_center(NphPhysicsWorld *world, NphPhysicsBody body,
                                            NphVec3 force);
void nph_physics_body_apply_torque(NphPhysicsWorld *world, NphPhysicsBody body, NphVec3 torque);
void nph_physics_body_apply_linear_impulse(NphPhysicsWorld *world, NphPhysicsBody body,
                                           NphVec3 impulse, NphDVec3 point);
void nph_physics_body_apply_angular_impulse(NphPhysicsWorld *world, NphPhysicsBody body,
                                            NphVec3 impulse);
// body.c:463-481: set mass properties (mass <= 0: back to the shapes').
void nph_physics_body_set_mass_data(NphPhysicsWorld *world, NphPhysicsBody body,
                                    const NphPhysicsMassData *mass);
// body.c:483-604
bool nph_physics_body_is_bullet(const NphPhysicsWorld *world, NphPhysicsBody body);
void nph_physics_body_set_bullet(NphPhysicsWorld *world, NphPhysicsBody body, bool bullet);
void nph_physics_body_set_damping(NphPhysicsWorld *world, NphPhysicsBody body, f32 linear,
                                  f32 angular);
void nph_physics_body_set_gravity_scale(NphPhysicsWorld *world, NphPhysicsBody body, f32 scale);
void nph_physics_body_set_locks(NphPhysicsWorld *world, NphPhysicsBody body, u32 locks);
void nph_physics_body_set_sleep(NphPhysicsWorld *world, NphPhysicsBody body, bool enable,
                                f32 threshold);
// body.c:567-593: awake wakes the body's island; asleep puts its whole
// island to sleep now (a kinematic body alone).
void nph_physics_body_set_awake(NphPhysicsWorld *world, NphPhysicsBody body, bool awake);
void nph_physics_body_set_user_data(NphPhysicsWorld *world, NphPhysicsBody body, u64 user_data);
u32 nph_physics_body_shapes(const NphPhysicsWorld *world, NphPhysicsBody body, NphPhysicsShape *out,
                            u32 capacity);

// ---------------------------------------------------------------------------
// Shapes (physics/shape.h:36-143)
// ---------------------------------------------------------------------------
typedef struct NphPhysicsFilter {
    u64 category;
    u64 mask;
    i32 group;
} NphPhysicsFilter;

typedef enum NphPhysicsCombine {
    NPH_PHYSICS_COMBINE_DEFAULT = 0,
    NPH_PHYSICS_COMBINE_AVERAGE = 1,
    NPH_PHYSICS_COMBINE_MIN = 2,
    NPH_PHYSICS_COMBINE_MULTIPLY = 3,
    NPH_PHYSICS_COMBINE_MAX = 4,
    NPH_PHYSICS_COMBINE_COUNT = 5,
} NphPhysicsCombine;

typedef struct NphPhysicsShapeDesc {
    NphPhysicsGeometry geometry;
    NphPhysicsPose pose;
    f32 density;
    f32 friction;
    f32 restitution;
    NphPhysicsCombine friction_combine;
    NphPhysicsCombine restitution_combine;
    NphPhysicsFilter filter;
    b32 contact_events;
    b32 hit_events;
    b32 sensor;
    b32 sensor_events;
    b32 ghost;
    u64 user_data;
} NphPhysicsShapeDesc; // shape.h:59-81

NphPhysicsFilter nph_physics_filter_default(void);                                   // shape.c:12
bool nph_physics_filter_should_collide(NphPhysicsFilter a, NphPhysicsFilter b);      // shape.c:21-27
NphPhysicsShapeDesc nph_physics_shape_desc_default(NphPhysicsGeometry geometry);     // shape.h:83-87
NphResult nph_physics_shape_create(NphPhysicsWorld *world, NphPhysicsBody body,
                                   const NphPhysicsShapeDesc *desc,
                                   NphPhysicsShape *out);                            // shape.c:179-263
bool nph_physics_shape_is_valid(const NphPhysicsWorld *world, NphPhysicsShape shape);
NphPhysicsBody nph_physics_shape_body(const NphPhysicsWorld *world, NphPhysicsShape shape);
NphAABB nph_physics_shape_aabb(const NphPhysicsWorld *world, NphPhysicsShape shape);
// shape.c:29-42: the pair's friction and restitution for these values and modes.
f32 nph_physics_combine_friction(f32 a, NphPhysicsCombine mode_a, f32 b, NphPhysicsCombine mode_b);
f32 nph_physics_combine_restitution(f32 a, NphPhysicsCombine mode_a, f32 b,
                                    NphPhysicsCombine mode_b);
// shape.c:285-311: its contacts (touching neighbours wake) and proxy first.
void nph_physics_shape_destroy(NphPhysicsWorld *world, NphPhysicsShape shape);
// shape.c:320-440: the shape's settings; new surfaces apply from the next
// step's narrow phase, a new density to the body's mass now, a new filter
// ends the contacts it no longer allows now.
NphPhysicsGeometry nph_physics_shape_geometry(const NphPhysicsWorld *world, NphPhysicsShape shape);
NphPhysicsPose nph_physics_shape_pose(const NphPhysicsWorld *world, NphPhysicsShape shape);
f32 nph_physics_shape_friction(const NphPhysicsWorld *world, NphPhysicsShape shape);
f32 nph_physics_shape_restitution(const NphPhysicsWorld *world, NphPhysicsShape shape);
NphPhysicsCombine nph_physics_shape_friction_combine(const NphPhysicsWorld *world,
                                                     NphPhysicsShape shape);
NphPhysicsCombine nph_physics_shape_restitution_combine(const NphPhysicsWorld *world,
                                                        NphPhysicsShape shape);
void nph_physics_shape_set_friction(NphPhysicsWorld *world, NphPhysicsShape shape, f32 friction);
void nph_physics_shape_set_restitution(NphPhysicsWorld *world, NphPhysicsShape shape,
                                       f32 restitution);
void nph_physics_shape_set_friction_combine(NphPhysicsWorld *world, NphPhysicsShape shape,
                                            NphPhysicsCombine mode);
void nph_physics_shape_set_restitution_combine(NphPhysicsWorld *world, NphPhysicsShape shape,
                                               NphPhysicsCombine mode);
f32 nph_physics_shape_density(const NphPhysicsWorld *world, NphPhysicsShape shape);
void nph_physics_shape_set_density(NphPhysicsWorld *world, NphPhysicsShape shape, f32 density);
NphPhysicsFilter nph_physics_shape_filter(const NphPhysicsWorld *world, NphPhysicsShape shape);
void nph_physics_shape_set_filter(NphPhysicsWorld *world, NphPhysicsShape shape,
                                  NphPhysicsFilter filter);
u64 nph_physics_shape_user_data(const NphPhysicsWorld *world, NphPhysicsShape shape);
void nph_physics_shape_set_user_data(NphPhysicsWorld *world, NphPhysicsShape shape, u64 user_data);

// ---------------------------------------------------------------------------
// World queries (physics/query.h:51-172, query.c): the broadphase's trees
// walked (static, kinematic, then dynamic shapes) and each shape found
// tested exactly in a frame centred on the query's origin; hits in the
// trees' order.  Host code over the host's trees; collide_shape's manifolds
// come from the device's narrow phase, all of a query's pairs in one
// launch.  Against mesh shapes: rays, box and shape overlaps; casts and
// collide_shape pass them over.
// ---------------------------------------------------------------------------
enum {
    NPH_PHYSICS_QUERY_STATIC = 1U << NPH_PHYSICS_BODY_STATIC,
    NPH_PHYSICS_QUERY_KINEMATIC = 1U << NPH_PHYSICS_BODY_KINEMATIC,
    NPH_PHYSICS_QUERY_DYNAMIC = 1U << NPH_PHYSICS_BODY_DYNAMIC,
    NPH_PHYSICS_QUERY_ALL_BODIES = 7U,
};

typedef struct NphPhysicsQueryFilter {
    u64 category;
    u64 mask;
    u32 body_types;        // NPH_PHYSICS_QUERY_*
    NphPhysicsBody ignore; // 0: none
    b32 sensors;
} NphPhysicsQueryFilter; // query.h:61-67

typedef struct NphPhysicsQueryHit {
    NphPhysicsShape shape;
    NphPhysicsBody body;
    NphDVec3 point;
    NphVec3 normal;
    f32 fraction;
    u32 triangle;
} NphPhysicsQueryHit; // query.h:73-80

typedef f32 (*NphPhysicsQueryCastFn)(void *user, const NphPhysicsQueryHit *hit);
typedef bool (*NphPhysicsQueryShapeFn)(void *user, NphPhysicsShape shape, NphPhysicsBody body);

typedef struct NphPhysicsQueryContact {
    NphPhysicsShape shape;
    NphPhysicsBody body;
    NphPhysicsManifold manifold;
} NphPhysicsQueryContact; // query.h:91-100
typedef bool (*NphPhysicsQueryContactFn)(void *user, const NphPhysicsQueryContact *contact);

NphPhysicsQueryFilter nph_physics_query_filter_default(void);                    // query.c:13-22
bool nph_physics_world_raycast_closest(const NphPhysicsWorld *world, NphDVec3 origin,
                                       NphVec3 translation, const NphPhysicsQueryFilter *filter,
                                       NphPhysicsQueryHit *out);                // query.c:271-283
void nph_physics_world_raycast(const NphPhysicsWorld *world, NphDVec3 origin, NphVec3 translation,
                               const NphPhysicsQueryFilter *filter, NphPhysicsQueryCastFn fn,
                               void *user);                                     // query.c:249-269
bool nph_physics_world_cast_shape_closest(const NphPhysicsWorld *world,
                                          const NphPhysicsGeometry *geometry, NphDVec3 position,
                                          NphQuat rotation, NphVec3 translation,
                                          const NphPhysicsQueryFilter *filter,
                                          NphPhysicsQueryHit *out);
void nph_physics_world_cast_shape(const NphPhysicsWorld *world, const NphPhysicsGeometry *geometry,
                                  NphDVec3 position, NphQuat rotation, NphVec3 translation,
                                  const NphPhysicsQueryFilter *filter, NphPhysicsQueryCastFn fn,
                                  void *user);
bool nph_physics_world_cast_shape_margin_closest(const NphPhysicsWorld *world,
                                                 const NphPhysicsGeometry *geometry,
                                                 NphDVec3 position, NphQuat rotation,
                                                 NphVec3 translation, f32 margin,
                                                 const NphPhysicsQueryFilter *filter,
                                                 NphPhysicsQueryHit *out);
void nph_physics_world_cast_shape_margin(const NphPhysicsWorld *world,
                                         const NphPhysicsGeometry *geometry, NphDVec3 position,
                                         NphQuat rotation, NphVec3 translation, f32 margin,
                                         const NphPhysicsQueryFilter *filter,
                                         NphPhysicsQueryCastFn fn, void *user); // query.c:294-329
void nph_physics_world_overlap_shape(const NphPhysicsWorld *world,
                                     const NphPhysicsGeometry *geometry, NphDVec3 position,
                                     NphQuat rotation, const NphPhysicsQueryFilter *filter,
                                     NphPhysicsQueryShapeFn fn, void *user);   // query.c:492-508
void nph_physics_world_overlap_aabb(const NphPhysicsWorld *world, NphAABB box,
                                    const NphPhysicsQueryFilter *filter, NphPhysicsQueryShapeFn fn,
                                    void *user);                                // query.c:533-546
void nph_physics_world_collide_shape(const NphPhysicsWorld *world,
                                     const NphPhysicsGeometry *geometry, NphDVec3 position,
                                     NphQuat rotation, f32 speculative_distance,
                                     const NphPhysicsQueryFilter *filter,
                                     NphPhysicsQueryContactFn fn, void *user);  // query.c:548-566
// query.h:180-181: the outward normal of a shape's own surface at a hit's
// world point (nph_physics_geometry_surface_normal in the shape's frame).
NphVec3 nph_physics_world_surface_normal(const NphPhysicsWorld *world, NphPhysicsShape shape,
                                         NphDVec3 point, u32 triangle, NphVec3 toward);

// ---------------------------------------------------------------------------
// Debug drawing (world.h:291-325): lines through a callback, in 32-bit
// coordinates relative to `origin`; colours 0xAABBGGRR.
// ---------------------------------------------------------------------------
enum {
    NPH_PHYSICS_DRAW_SHAPES = 1U << 0U,
    NPH_PHYSICS_DRAW_AABBS = 1U << 1U,    // the broadphase's fattened boxes
    NPH_PHYSICS_DRAW_CONTACTS = 1U << 2U, // touching points and normals
    NPH_PHYSICS_DRAW_MASS = 1U << 3U,     // centres of mass
    NPH_PHYSICS_DRAW_SLEEP = 1U << 4U,    // sleeping bodies' shapes in their own colour
    NPH_PHYSICS_DRAW_JOINTS = 1U << 5U,   // anchors to the centres, and the frames' x axes
};
#define NPH_PHYSICS_DRAW_SENSOR_RGBA 0xFF40D0E0U
#define NPH_PHYSICS_DRAW_GHOST_RGBA 0xFFE0A0C0U
typedef struct NphPhysicsDebugDraw {
    void (*line)(void *user, NphVec3 a, NphVec3 b, u32 rgba);
    void *user;
    NphDVec3 origin;
    u32 flags; // NPH_PHYSICS_DRAW_*
} NphPhysicsDebugDraw;
// The world through `draw`, between steps (reads the device's poses and
// contact records).
void nph_physics_world_draw(const NphPhysicsWorld *world, const NphPhysicsDebugDraw *draw);
// One shape's outline, as the world draws its shapes: `geometry` at `pose`
// relative to `base` (draw's flags are not read).
void nph_physics_geometry_draw(const NphPhysicsDebugDraw *draw, const NphPhysicsGeometry *geometry,
                               NphDVec3 base, NphPhysicsPose pose, u32 rgba);

// ---------------------------------------------------------------------------
// Joints (joint.h:51-272).  Ball, hinge and fixed joints are solved; the
// slider, distance and 6-DOF types are declared for binding compatibility and
// rejected by nph_physics_joint_create with NPH_RESULT_UNSUPPORTED.
// ---------------------------------------------------------------------------
typedef struct NphPhysicsJoint {
    u32 value;
} NphPhysicsJoint; // joint.h:51-53

typedef enum NphPhysicsJointType {
    NPH_PHYSICS_JOINT_BALL = 0,
    NPH_PHYSICS_JOINT_HINGE = 1,
    NPH_PHYSICS_JOINT_FIXED = 2,
    NPH_PHYSICS_JOINT_SLIDER = 3,
    NPH_PHYSICS_JOINT_DISTANCE = 4,
    NPH_PHYSICS_JOINT_D6 = 5,
    NPH_PHYSICS_JOINT_TYPE_COUNT = 6,
} NphPhysicsJointType; // joint.h:62-70

typedef struct NphPhysicsBallJoint {
    b32 enable_swing_limit;
    f32 swing_y;
    f32 swing_z;
    b32 enable_twist_limit;
    f32 twist_lower;
    f32 twist_upper;
    f32 friction_torque;
} NphPhysicsBallJoint; // joint.h:73-86

typedef struct NphPhysicsHingeJoint {
    b32 enable_limit;
    f32 lower;
    f32 upper;
    b32 enable_motor;
    f32 motor_speed;
    f32 max_motor_torque;
    b32 enable_spring;
    f32 spring_hertz;
    f32 spring_damping_ratio;
    f32 spring_target;
} NphPhysicsHingeJoint; // joint.h:90-104

typedef struct NphPhysicsSliderJoint {
    b32 enable_limit;
    f32 lower;
    f32 upper;
    b32 enable_motor;
    f32 motor_speed;
    f32 max_motor_force;
    b32 enable_spring;
    f32 spring_hertz;
    f32 spring_damping_ratio;
    f32 spring_target;
} NphPhysicsSliderJoint; // joint.h:108-122

typedef struct NphPhysicsDistanceJoint {
    f32 length;
    b32 enable_limit;
    f32 min_length;
    f32 max_length;
    b32 enable_spring;
    f32 spring_hertz;
    f32 spring_damping_ratio;
    b32 enable_motor;
    f32 motor_speed;
    f32 max_motor_force;
} NphPhysicsDistanceJoint; // joint.h:128-141

typedef enum NphPhysicsAxisMotion {
    NPH_PHYSICS_AXIS_LOCKED = 0,
    NPH_PHYSICS_AXIS_LIMITED = 1,
    NPH_PHYSICS_AXIS_FREE = 2,
} NphPhysicsAxisMotion;

typedef struct NphPhysicsD6Axis {
    NphPhysicsAxisMotion motion;
    f32 lower;
    f32 upper;
    b32 enable_motor;
    f32 motor_speed;
    f32 max_motor_force;
    b32 enable_spring;
    f32 spring_hertz;
    f32 spring_damping_ratio;
    f32 spring_target;
} NphPhysicsD6Axis; // joint.h:152-163

enum {
    NPH_PHYSICS_D6_LINEAR_X = 0,
    NPH_PHYSICS_D6_LINEAR_Y = 1,
    NPH_PHYSICS_D6_LINEAR_Z = 2,
    NPH_PHYSICS_D6_TWIST = 3,
    NPH_PHYSICS_D6_SWING_Y = 4,
    NPH_PHYSICS_D6_SWING_Z = 5,
    NPH_PHYSICS_D6_AXES = 6,
};

typedef struct NphPhysicsD6Joint {
    NphPhysicsD6Axis axes[NPH_PHYSICS_D6_AXES];
} NphPhysicsD6Joint;

typedef struct NphPhysicsJointDesc {
    NphPhysicsJointType type;
    NphPhysicsBody body_a;
    NphPhysicsBody body_b;
    NphPhysicsPose frame_a; // in body A's frame (its origin)
    NphPhysicsPose frame_b;
    b32 collide_connected;
    u64 user_data;
    union {
        NphPhysicsBallJoint ball;
        NphPhysicsHingeJoint hinge;
        NphPhysicsSliderJoint slider;
        NphPhysicsDistanceJoint distance;
        NphPhysicsD6Joint d6;
    };
} NphPhysicsJointDesc; // joint.h:180-195

NphPhysicsJointDesc nph_physics_joint_desc_default(NphPhysicsJointType type, NphPhysicsBody a,
                                                   NphPhysicsBody b);              // joint.c:37-81
void nph_physics_joint_frames_at(const NphPhysicsWorld *world, NphPhysicsBody a, NphPhysicsBody b,
                                 NphDVec3 anchor, NphVec3 axis, NphPhysicsPose *frame_a,
                                 NphPhysicsPose *frame_b);                         // joint.c:83-107
NphResult nph_physics_joint_create(NphPhysicsWorld *world, const NphPhysicsJointDesc *desc,
                                   NphPhysicsJoint *out);                          // joint.c:357-443
void nph_physics_joint_destroy(NphPhysicsWorld *world, NphPhysicsJoint joint);     // joint.c:466-473
bool nph_physics_joint_is_null(NphPhysicsJoint j); // joint.h:55-58 (inline there)
bool nph_physics_joint_is_valid(const NphPhysicsWorld *world, NphPhysicsJoint joint);
NphPhysicsJointDesc nph_physics_joint_desc(const NphPhysicsWorld *world,
                                           NphPhysicsJoint joint);                 // joint.c:475-509
void nph_physics_joint_set_ball(NphPhysicsWorld *world, NphPhysicsJoint joint,
                                const NphPhysicsBallJoint *ball);                  // joint.c:517-531
void nph_physics_joint_set_hinge(NphPhysicsWorld *world, NphPhysicsJoint joint,
                                 const NphPhysicsHingeJoint *hinge);               // joint.c:533-547
void nph_physics_joint_set_slider(NphPhysicsWorld *world, NphPhysicsJoint joint,
                                  const NphPhysicsSliderJoint *slider);
void nph_physics_joint_set_distance(NphPhysicsWorld *world, NphPhysicsJoint joint,
                                    const NphPhysicsDistanceJoint *distance);
void nph_physics_joint_set_d6(NphPhysicsWorld *world, NphPhysicsJoint joint,
                              const NphPhysicsD6Joint *d6);
f32 nph_physics_joint_hinge_angle(const NphPhysicsWorld *world, NphPhysicsJoint joint); // :627-631
f32 nph_physics_joint_hinge_speed(const NphPhysicsWorld *world, NphPhysicsJoint joint); // :633-644
f32 nph_physics_joint_slider_position(const NphPhysicsWorld *world, NphPhysicsJoint joint);
f32 nph_physics_joint_slider_speed(const NphPhysicsWorld *world, NphPhysicsJoint joint);
f32 nph_physics_joint_length(const NphPhysicsWorld *world, NphPhysicsJoint joint);     // :621-625
void nph_physics_joint_reaction(const NphPhysicsWorld *world, NphPhysicsJoint joint,
                                NphVec3 *force, NphVec3 *torque);                  // :646-665
u64 nph_physics_joint_user_data(const NphPhysicsWorld *world, NphPhysicsJoint joint);
void nph_physics_joint_set_user_data(NphPhysicsWorld *world, NphPhysicsJoint joint,
                                     u64 user_data);
u32 nph_physics_body_joints(const NphPhysicsWorld *world, NphPhysicsBody body,
                            NphPhysicsJoint *out, u32 capacity);                   // joint.c:675-693

// ---------------------------------------------------------------------------
// EXTENSION: contact readback.  The reference exposes contacts only through
// events, stats and debug drawing; its tests read the contact records
// white-box (test_world.c:123-124).  This copies every live contact record
// (the pool's order; child records of mesh pairs included) as the record
// below, whose point layout is PhysContactPoint's (world_internal.h:196-205).
// Returns the number of live records; writes at most `capacity`.
// ---------------------------------------------------------------------------
typedef struct NphPhysicsContactPointInfo {
    NphVec3 anchor_a;
    NphVec3 anchor_b;
    f32 separation;
    u32 id;
    f32 normal_impulse;
    f32 max_normal_impulse;
    NphVec3 tangent_impulse;
    f32 relative_velocity;
} NphPhysicsContactPointInfo; // 56 B

typedef struct NphPhysicsContactInfo {
    u32 record;  // index in the contact pool
    u32 shape_a; // slots (handle & 0xFFFFF)
    u32 shape_b;
    u32 body_a;
    u32 body_b;
    u32 flags;      // bit 0: touching
    u32 color;      // constraint-graph colour; 24 = overflow; 0xFFFFFFFF: not in the graph
    u32 color_slot; // position in the colour's list
    u32 parent;
    u32 child;
    f32 friction;
    f32 restitution;
    NphVec3 normal;
    u32 point_count;
    NphPhysicsContactPointInfo points[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
} NphPhysicsContactInfo;

u32 nph_physics_world_contacts(const NphPhysicsWorld *world, NphPhysicsContactInfo *out,
                               u32 capacity);

// EXTENSION: the name of the compute backend the world steps on ("hip:gfx950"
// for this library).
const char *nph_amd_backend_name(const NphPhysicsWorld *world);

// EXTENSION: the contacts the last step's solver took: every colour's and the
// overflow's (the lanes of solver.c:810-843); times substeps, the step's
// contact solves.
u32 nph_amd_solved_contacts(const NphPhysicsWorld *world);
// EXTENSION: the mechanical energy of the world's dynamic bodies now, in
// f64 joules: out[0] linear kinetic, out[1] angular kinetic (body-axis
// inertia), out[2] potential in the world's gravity (zero at the origin).
void nph_amd_world_energy(const NphPhysicsWorld *world, f64 out[3]);

// EXTENSION: device time of the last step's stages, in milliseconds, from HIP
// events on the world's stream: [0] narrow phase, [1] solver and finalize,
// [2] the constraint passes alone (warm start, solve, relax, restitution).
void nph_amd_device_times(const NphPhysicsWorld *world, f64 out[3]);

// EXTENSION: per-launch timing of the constraint-pass kernel (one launch per
// colour per pass), for measurement. Enabling (or disabling) clears the
// totals; while enabled each launch is bracketed by two HIP events on the
// world's stream. out[0] = device milliseconds summed over the timed
// launches, out[1] = how many.
void nph_amd_time_kernels(NphPhysicsWorld *world, bool enable);
void nph_amd_kernel_times(const NphPhysicsWorld *world, f64 out[2]);

// ---------------------------------------------------------------------------
// EXTENSION: the learned warm start's network (no counterpart in the
// reference).  A 16 -> 64 -> 64 -> 16 MLP with GELU: 14 contact features in
// (zero-padded to 16), the normal and two friction impulses out (columns
// 0..2 of 4), BF16 weights and activations on the matrix cores with FP32
// accumulation.  All pointers are device addresses; `stream` is a
// hipStream_t (NULL: the null stream).  Parameters are one FP32 array of
// NPH_AMD_MLP_PARAMS: W1 [16][64], b1 [64], W2 [64][64], b2 [64],
// W3 [64][16], b3 [16], row-major, input index first.
// ---------------------------------------------------------------------------
#define NPH_AMD_MLP_FEATURES 14U
#define NPH_AMD_MLP_INPUTS 16U
#define NPH_AMD_MLP_OUTPUTS 3U
#define NPH_AMD_MLP_PARAMS 6288U

// y[r][0..3] = the network at x[r][0..15] (BF16), for r < rows. x must be
// 8-byte and y 16-byte aligned (INVALID_ARGUMENT otherwise).
NphResult nph_amd_mlp_forward(const void *x_bf16, u32 rows, const f32 *params, f32 *y,
                              void *stream);
// The network's GELU at x[0..n): out[4 i + 0..3] = GELU and its derivative
// by the approximation the kernels use, then both through the device's
// erff and expf (a check of the approximation; no counterpart in the
// reference).
NphResult nph_amd_mlp_gelu(const f32 *x, u32 n, f32 *out, void *stream);
// Adds the gradient of scale * sum (y - target)^2 over rows and outputs
// 0..2 (targets as y's layout, 4 floats a row) into grad[NPH_AMD_MLP_PARAMS],
// and scale * sum (y - target)^2 into *loss.
NphResult nph_amd_mlp_grad(const void *x_bf16, const f32 *target, u32 rows, const f32 *params,
                           f32 scale, f32 *grad, f32 *loss, void *stream);
// One Adam step (beta 0.9, 0.999, eps 1e-8) at `step` (from 1).
NphResult nph_amd_mlp_adam(f32 *params, const f32 *grad, f32 *m, f32 *v, u32 step, f32 lr,
                           void *stream);

// Device memory for the network's callers (no counterpart in the
// reference): allocation (NULL on failure), release, and a synchronous copy
// between any two host or device addresses.
void *nph_amd_device_alloc(u64 bytes);
void nph_amd_device_free(void *p);
NphResult nph_amd_device_copy(void *dst, const void *src, u64 bytes);

// ---------------------------------------------------------------------------
// EXTENSION: the learned warm start of a world (no counterpart in the
// reference, whose warm start is the impulse carried by point id or place,
// contact.c:296-374).  With a network set, the step computes every contact
// point's 14 features after the constraint rows are prepared (approach speed
// along the normal and both tangents, separation, normal, the two
// inverse masses, friction, normal mass, the three tangent mass terms, in
// that order), runs the network, and starts the chosen points from its
// impulses: impulse = output * output_scale * normal mass, the normal one
// clamped at zero and the friction ones to the friction circle.  The plain
// solver then runs every substep as before.  With the gate on, a contact
// whose post-solve residual exceeds the largest residual among the contacts
// that started plainly that step is started plainly (carried impulses) for
// the rest of its life.  (With NPH_AMD_WARM_ALL_POINTS, contact records whose
// index is a multiple of 16 always start plainly: that control sample gives
// the plain residual.)  The residual of a point is the relax pass's
// condition at the step's end: with s the separation and
// e = vn + (s > 0 ? s / h : 0), |e| if its normal impulse is positive, else
// max(0, -e); a contact's is its points' largest.
// ---------------------------------------------------------------------------
#define NPH_AMD_WARM_NEW_POINTS 0U // points that carried no impulse
#define NPH_AMD_WARM_ALL_POINTS 1U // every point

typedef struct NphAmdWarmStart {
    const f32 *params;        // host, NPH_AMD_MLP_PARAMS (copied)
    f32 feature_shift[16];    // network input = (feature - shift) * scale
    f32 feature_scale[16];
    f32 output_scale[4];      // [0..2]: normal and two friction outputs
    u32 mode;                 // NPH_AMD_WARM_*
    bool gate;                // the residual gate
} NphAmdWarmStart;

// Sets (a copy of) the network, or with NULL returns to the plain warm start.
NphResult nph_amd_world_set_warm_start(NphPhysicsWorld *world, const NphAmdWarmStart *ws);

// Residual statistics of the last step (computed while a network is set or
// measuring is enabled), over contacts (lanes) that started plainly and
// from the network.
typedef struct NphAmdWarmStats {
    f32 residual_max_plain;
    f32 residual_max_learned;
    f32 residual_sum_plain;
    f32 residual_sum_learned;
    f32 penetration_max;
    u32 contacts_plain;
    u32 contacts_learned;
    u32 gated; // contacts the gate handed to the plain warm start this step
} NphAmdWarmStats;
void nph_amd_world_measure_residual(NphPhysicsWorld *world, bool enable);
NphResult nph_amd_world_warm_stats(const NphPhysicsWorld *world, NphAmdWarmStats *out);

// The contact log, the network's training data: while enabled (capacity > 0
// rows; 0 disables and frees it) every step appends, for each contact point,
// its 16 features as above (slot 14 the normal impulse it carried, slot 15
// one when it carried no impulse) and its 4 targets (the normal and two friction
// impulses the step converged to, after restitution, then the same flag).
// Rows past the capacity are dropped.  The arrays are device memory owned by
// the world, valid until the log is disabled or the world destroyed.
NphResult nph_amd_world_log_contacts(NphPhysicsWorld *world, u32 capacity);
NphResult nph_amd_world_contact_log(const NphPhysicsWorld *world, u32 *rows,
                                    const f32 **features, const f32 **targets);

// Training on the device with the network's kernels: rows appended from
// device arrays (a contact log), a normalisation fixed from them (inputs to
// zero mean and unit deviation; targets = impulses / normal mass, scaled to
// unit deviation), then epochs of minibatch Adam in a fresh pseudo-random
// order each.
typedef struct NphAmdTrainer NphAmdTrainer;
NphResult nph_amd_trainer_create(u32 capacity, u64 seed, NphAmdTrainer **out);
void nph_amd_trainer_destroy(NphAmdTrainer *t);
// Only points whose flag (slot 15) is at least `min_flag` are kept.
NphResult nph_amd_trainer_add(NphAmdTrainer *t, const f32 *features, const f32 *targets,
                              u32 rows, f32 min_flag);
u32 nph_amd_trainer_rows(const NphAmdTrainer *t);
NphResult nph_amd_trainer_prepare(NphAmdTrainer *t);
// One epoch; *mean_loss = its batches' mean loss (normalised units).
NphResult nph_amd_trainer_epoch(NphAmdTrainer *t, u32 batch, f32 lr, f32 *mean_loss);
// The loss over every row.
NphResult nph_amd_trainer_loss(NphAmdTrainer *t, f32 *loss);
// The trained network: its parameters into params (host,
// NPH_AMD_MLP_PARAMS), the rest into *out (out->params = params).
NphResult nph_amd_trainer_warm_start(const NphAmdTrainer *t, f32 *params, NphAmdWarmStart *out);

// ---------------------------------------------------------------------------
// Batched worlds (no counterpart in the reference; for callers that step
// many small independent worlds, such as parameter sweeps and training
// environments): a batch creates `world_count` worlds that share one
// NphPhysicsWorldDesc and one device stream; each is filled and read through
// the world API above, and nph_amd_batch_step steps all of them, each
// stage of the step one launch over every world's items.  Each world's
// results are bit for bit those it gets stepped alone.  The batch owns its
// worlds: destroy the batch, not the worlds.  Continuous collision and sleep
// are not batched (creation returns NPH_RESULT_UNSUPPORTED), nor is the contact log
// (the step returns NPH_RESULT_UNSUPPORTED while a world logs).  The learned
// warm start is: one network for the batch (nph_amd_batch_set_warm_start,
// or the same one set on each world), its rows for every world in one
// forward launch; the step returns NPH_RESULT_UNSUPPORTED when two worlds
// have different networks.
// ---------------------------------------------------------------------------
typedef struct NphAmdBatch NphAmdBatch;

// Launch choices of the batched step.
typedef struct NphAmdBatchTuning {
    u32 pass_block; // threads a block of the colour passes: 64, 128 or 256
    u32 body_block; // threads a block of the per-body and per-lane stages
    u32 use_graph;  // the substep loop launched as a captured HIP graph
    u32 groups;     // world groups the step is pipelined over (1 to 4), each
                    // on its own stream, balanced by awake bodies every step
} NphAmdBatchTuning;

typedef struct NphAmdBatchStats {
    u64 solved_contacts; // the last step's constraint-graph contacts, all worlds
    u64 awake_bodies;    // the last step's awake bodies, all worlds
    f64 step_ms;         // wall time of the last step
    f64 narrow_ms;       // both narrow-phase calls: staging, launches, waiting for codes
    f64 pairs_ms;        // new pairs: trees and change lists (host), the search (device)
    f64 changes_ms;      // host: narrow-phase changes and solver input
    f64 solve_ms;        // the solve: lists up, device stages, grown proxies back
    f64 post_ms;         // host: grown proxies into the trees, events
    f64 device_solve_ms; // device time of the solve (events)
    f64 device_passes_ms;
    f64 device_collide_ms;
    f64 wait_narrow_ms;  // host blocked on the device for the change codes
    f64 wait_solve_ms;   // and for the solve's grown proxies
    f64 device_pairs_ms; // device time of the pair searches (events)
    f64 pairs_trees_ms;  // of pairs_ms: the trees' rebuilds and the change lists (wall)
    f64 pairs_search_ms; // the searches: upload, launches, pairs back (wall)
    f64 pairs_create_ms; // the new contacts (wall)
    f64 rebuild_cpu_ms;  // the trees' rebuilds, summed over the worlds' threads
    u64 colored_joints;  // the last step's joints in colours, all worlds
    u64 colored_joint_ends; // their dynamic bodies
    f64 pairs_gather_ms;    // of pairs_search_ms: packing the trees' written nodes (host)
    f64 pairs_upload_ms;    // the upload's copy (device events)
    u64 pairs_upload_bytes; // its bytes
} NphAmdBatchStats;

NphResult nph_amd_batch_create(const NphPhysicsWorldDesc *desc, u32 world_count,
                               NphAmdBatch **out);
void nph_amd_batch_destroy(NphAmdBatch *batch);
u32 nph_amd_batch_world_count(const NphAmdBatch *batch);
NphPhysicsWorld *nph_amd_batch_world(NphAmdBatch *batch, u32 index);
NphResult nph_amd_batch_step(NphAmdBatch *batch, f32 dt);
NphAmdBatchStats nph_amd_batch_stats(const NphAmdBatch *batch);
void nph_amd_batch_set_tuning(NphAmdBatch *batch, const NphAmdBatchTuning *tuning);
NphAmdBatchTuning nph_amd_batch_tuning(const NphAmdBatch *batch);
// The autotuner: steps the batch under every candidate launch choice in
// turn, `rounds` times round-robin (the worlds advance rounds x candidates
// steps), keeps the one whose steps took the least wall time in total, and
// returns it in *chosen.
NphResult nph_amd_batch_autotune(NphAmdBatch *batch, f32 dt, u32 rounds,
                                 NphAmdBatchTuning *chosen);
// The learned warm start on every world of the batch (NULL: off), and
// residual measurement on every world.
NphResult nph_amd_batch_set_warm_start(NphAmdBatch *batch, const NphAmdWarmStart *ws);
void nph_amd_batch_measure_residual(NphAmdBatch *batch, bool enable);
// The residual statistics summed over every measured world and step since
// the last reset (NphAmdWarmStats' quantities, as sums and maxima).
typedef struct NphAmdBatchWarmStats {
    f64 residual_sum_plain;   // contacts' post-solve residuals (m/s), plain start
    f64 residual_sum_learned; // and learned start
    f64 contacts_plain;       // contact-steps measured
    f64 contacts_learned;
    f64 residual_max_plain;
    f64 residual_max_learned;
    f64 penetration_max;
    f64 gated;
    u32 steps;
    u32 pad;
} NphAmdBatchWarmStats;
NphResult nph_amd_batch_warm_stats(NphAmdBatch *batch, NphAmdBatchWarmStats *out, bool reset);
// HIP events around every colour-pass launch (not in graph mode): out[0] =
// device milliseconds summed since enabling, out[1] = launches.
void nph_amd_batch_time_kernels(NphAmdBatch *batch, bool enable);
void nph_amd_batch_kernel_times(const NphAmdBatch *batch, f64 out[2]);
// With kernel timing on: the convex narrow-phase kernel's dispatch-stamped
// durations since enabling: out[0] = device milliseconds, out[1] = launches,
// out[2] = pairs those launches took.
void nph_amd_batch_collide_times(const NphAmdBatch *batch, f64 out[3]);

#ifdef __cplusplus
}
#endif

#endif // NEPHELINE_PHYSICS_H
