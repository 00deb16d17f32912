# SPDX-License-Identifier: Apache-2.0
"""Python mirror of the reference's physics C API (engine/physics/include/
nepheline/physics/{world,body,shape,geometry,types}.h) over ctypes.

One binding class drives any library that exports that API: this package's
own libnepheline_amd.so (the HIP world step) or, in tests, the reference's
own code built under oracle/_ref.  Struct layouts follow the reference's
headers field for field (world.h:104-143, body.h:45-72, shape.h:59-81,
geometry.h:68-78, types.h:65-70); the readback record is the extension
declared in include/nepheline_physics.h.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PRODUCT_LIB = os.path.join(HERE, "libnepheline_amd.so")

# NphResult (engine/core/include/nepheline/core/result.h:15-37)
OK, OUT_OF_MEMORY, INVALID_ARGUMENT, NOT_FOUND = 0, 1, 2, 3
UNSUPPORTED = 8

SHAPE_SPHERE, SHAPE_CAPSULE, SHAPE_BOX, SHAPE_HULL, SHAPE_MESH = 0, 1, 2, 3, 4
BODY_STATIC, BODY_KINEMATIC, BODY_DYNAMIC = 0, 1, 2
NULL32 = 0xFFFFFFFF


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class DVec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Quat(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Mat3(C.Structure):
    _fields_ = [("c", Vec3 * 3)]


class AABB(C.Structure):
    _fields_ = [("min", Vec3), ("max", Vec3)]


class Pose(C.Structure):
    _fields_ = [("position", Vec3), ("rotation", Quat)]


class MassData(C.Structure):
    _fields_ = [("mass", C.c_float), ("center", Vec3), ("inertia", Mat3)]


class _GeomUnion(C.Union):
    _fields_ = [("radius", C.c_float), ("capsule", C.c_float * 2), ("half_extents", Vec3),
                ("ptr", C.c_void_p)]


class Geometry(C.Structure):
    _fields_ = [("type", C.c_int), ("u", _GeomUnion)]


class ManifoldPoint(C.Structure):
    """collide.h:66-73"""
    _fields_ = [("point", Vec3), ("separation", C.c_float), ("id", C.c_uint32)]


class Manifold(C.Structure):
    """collide.h:75-79"""
    _fields_ = [("normal", Vec3), ("point_count", C.c_uint32), ("points", ManifoldPoint * 4)]


class CollideCache(C.Structure):
    """collide.h:85-91"""
    _fields_ = [("type", C.c_uint8), ("reserved", C.c_uint8), ("index_a", C.c_uint16),
                ("index_b", C.c_uint16), ("reserved2", C.c_uint16)]


class CollideQuery(C.Structure):
    """nph_amd_collide_many's pair (this library's extension)."""
    _fields_ = [("a", Geometry), ("pose_a", Pose), ("b", Geometry), ("pose_b", Pose),
                ("speculative_distance", C.c_float), ("pad", C.c_uint32)]


def bind_collide(a):
    """The pair queries of either library (collide.h:96-121)."""
    if hasattr(a, "collide"):
        return a
    L = a.lib
    a.collide_supported = _bind(L, "nph_physics_collide_supported", C.c_bool, [C.c_int, C.c_int])
    a.collide = _bind(L, "nph_physics_collide", None,
                      [C.POINTER(Geometry), Pose, C.POINTER(Geometry), Pose, C.c_float,
                       C.POINTER(CollideCache), C.POINTER(Manifold)])
    a.manifold_reduce = _bind(L, "nph_physics_manifold_reduce", C.c_uint32,
                              [C.POINTER(ManifoldPoint), C.c_uint32, Vec3])
    if hasattr(L, "nph_amd_collide_many"):
        a.collide_many = _bind(L, "nph_amd_collide_many", C.c_int,
                               [C.POINTER(CollideQuery), C.c_uint32, C.POINTER(Manifold),
                                C.POINTER(CollideCache)])
    if hasattr(L, "nph_amd_collide_many_static"):
        a.collide_many_static = _bind(L, "nph_amd_collide_many_static", C.c_int,
                                      [C.POINTER(CollideQuery), C.c_uint32, C.POINTER(Manifold),
                                       C.POINTER(CollideCache)])
    return a


class Proxy(C.Structure):
    """geometry.h:163-167"""
    _fields_ = [("points", C.c_void_p), ("count", C.c_uint32), ("radius", C.c_float)]


class SegmentDistance(C.Structure):
    _fields_ = [("closest1", Vec3), ("closest2", Vec3), ("fraction1", C.c_float),
                ("fraction2", C.c_float), ("distance_squared", C.c_float)]


class SimplexCache(C.Structure):
    _fields_ = [("count", C.c_uint32), ("index_a", C.c_uint8 * 4), ("index_b", C.c_uint8 * 4)]


class Distance(C.Structure):
    _fields_ = [("point_a", Vec3), ("point_b", Vec3), ("normal", Vec3), ("distance", C.c_float),
                ("iterations", C.c_uint32), ("simplex_count", C.c_uint32)]


class Penetration(C.Structure):
    _fields_ = [("normal", Vec3), ("depth", C.c_float), ("point_a", Vec3), ("point_b", Vec3)]


class Triangle(C.Structure):
    _fields_ = [("vertices", Vec3 * 3), ("active_edges", C.c_uint32), ("two_sided", C.c_int32)]


class MeshManifolds(C.Structure):
    _fields_ = [("count", C.c_uint32), ("manifolds", Manifold * 4)]


def bind_geometry(a):
    """The geometry, mass and closest-point functions (geometry.h:120-181,
    types.h:130-142, collide.h:129-235)."""
    L = a.lib
    G = C.POINTER(Geometry)
    a.geometry_support = _bind(L, "nph_physics_geometry_support", Vec3, [G, Vec3])
    a.geometry_inner_radius = _bind(L, "nph_physics_geometry_inner_radius", C.c_float, [G])
    a.geometry_outer_radius = _bind(L, "nph_physics_geometry_outer_radius", C.c_float, [G])
    a.geometry_surface_normal = _bind(L, "nph_physics_geometry_surface_normal", Vec3,
                                      [G, Vec3, C.c_uint32, Vec3])
    a.geometry_proxy = _bind(L, "nph_physics_geometry_proxy", Proxy, [G, C.POINTER(Vec3)])
    a.proxy_support = _bind(L, "nph_physics_proxy_support", C.c_uint32,
                            [C.POINTER(Proxy), Vec3])
    a.mass_transform = _bind(L, "nph_physics_mass_transform", MassData, [MassData, Pose])
    a.mass_add = _bind(L, "nph_physics_mass_add", MassData, [MassData, MassData])
    a.inertia_offset = _bind(L, "nph_physics_inertia_offset", Mat3, [C.c_float, Vec3])
    a.segment_distance = _bind(L, "nph_physics_segment_distance", SegmentDistance,
                               [Vec3, Vec3, Vec3, Vec3])
    a.distance = _bind(L, "nph_physics_distance", None,
                       [C.POINTER(Proxy), Pose, C.POINTER(Proxy), Pose, C.c_bool,
                        C.POINTER(SimplexCache), C.POINTER(Distance)])
    a.penetration = _bind(L, "nph_physics_penetration", C.c_bool,
                          [C.POINTER(Proxy), Pose, C.POINTER(Proxy), Pose,
                           C.POINTER(Penetration)])
    return a


DRAW_SHAPES, DRAW_AABBS, DRAW_CONTACTS, DRAW_MASS, DRAW_SLEEP, DRAW_JOINTS = (
    1, 2, 4, 8, 16, 32)
DrawLineFn = C.CFUNCTYPE(None, C.c_void_p, Vec3, Vec3, C.c_uint32)


class DebugDraw(C.Structure):
    """world.h:310-315"""
    _fields_ = [("line", DrawLineFn), ("user", C.c_void_p), ("origin", DVec3),
                ("flags", C.c_uint32)]


class LineRecorder:
    """A debug draw whose lines are kept as (a, b, rgba) byte strings."""

    def __init__(self, origin=(0.0, 0.0, 0.0), flags=0):
        self.lines = []

        def line(_user, a, b, rgba):
            self.lines.append(bytes(a) + bytes(b) + int(rgba).to_bytes(4, "little"))

        self._fn = DrawLineFn(line)
        self.draw = DebugDraw(self._fn, None, DVec3(*origin), flags)


def bind_draw(a):
    """Debug drawing (world.h:291-325)."""
    L = a.lib
    a.world_draw = _bind(L, "nph_physics_world_draw", None, [C.c_void_p, C.POINTER(DebugDraw)])
    a.geometry_draw = _bind(L, "nph_physics_geometry_draw", None,
                            [C.POINTER(DebugDraw), C.POINTER(Geometry), DVec3, Pose, C.c_uint32])
    return a


def bind_surface_queries(a):
    """A convex shape against one triangle and against a mesh
    (collide.h:129-168)."""
    L = a.lib
    G = C.POINTER(Geometry)
    a.collide_triangle = _bind(L, "nph_physics_collide_triangle", None,
                               [G, Pose, C.POINTER(Triangle), C.c_float, C.POINTER(Manifold)])
    a.collide_mesh = _bind(L, "nph_physics_collide_mesh", None,
                           [G, Pose, C.c_void_p, Pose, C.c_float, C.POINTER(MeshManifolds)])
    return a


class HullDesc(C.Structure):
    """hull.h:111-132."""
    _fields_ = [("points", C.c_void_p), ("point_count", C.c_uint32),
                ("max_vertices", C.c_uint32), ("tolerance", C.c_float)]


def bind_hull(a):
    """The hull builder and block queries of either library (hull.h:135-166)."""
    if hasattr(a, "hull_build"):
        return a
    L = a.lib
    a.hull_build = _bind(L, "nph_physics_hull_build", C.c_int,
                         [C.POINTER(HullDesc), C.c_void_p, C.POINTER(C.c_void_p)])
    a.hull_make_box = _bind(L, "nph_physics_hull_make_box", C.c_int,
                            [Vec3, C.c_void_p, C.POINTER(C.c_void_p)])
    a.hull_free = _bind(L, "nph_physics_hull_free", None, [C.c_void_p, C.c_void_p])
    a.hull_support = _bind(L, "nph_physics_hull_support", C.c_uint32, [C.c_void_p, Vec3])
    a.hull_face_vertex_count = _bind(L, "nph_physics_hull_face_vertex_count", C.c_uint32,
                                     [C.c_void_p, C.c_uint32])
    a.hull_validate = _bind(L, "nph_physics_hull_validate", C.c_bool,
                            [C.c_void_p, C.c_uint32, C.c_float])
    return a


class MeshDesc(C.Structure):
    """mesh.h:158-176."""
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint32),
                ("indices", C.c_void_p), ("triangle_count", C.c_uint32),
                ("weld_tolerance", C.c_float), ("active_edge_angle", C.c_float),
                ("two_sided", C.c_int32)]


def bind_mesh(a):
    """The mesh builder and validator of either library (mesh.h:189-205)."""
    if hasattr(a, "mesh_build"):
        return a
    L = a.lib
    a.mesh_build = _bind(L, "nph_physics_mesh_build", C.c_int,
                         [C.POINTER(MeshDesc), C.c_void_p, C.POINTER(C.c_void_p)])
    a.mesh_free = _bind(L, "nph_physics_mesh_free", None, [C.c_void_p, C.c_void_p])
    a.mesh_validate = _bind(L, "nph_physics_mesh_validate", C.c_bool, [C.c_void_p, C.c_uint32])
    a.mesh_overlap = _bind(L, "nph_physics_mesh_overlap", None,
                           [C.c_void_p, AABB, MESH_VISIT, C.c_void_p])
    a.mesh_raycast = _bind(L, "nph_physics_mesh_raycast", C.c_bool,
                           [C.c_void_p, Vec3, Vec3, C.POINTER(MeshHit)])
    return a


# mesh.h:207-222: the overlap's visit and the ray's hit.
MESH_VISIT = C.CFUNCTYPE(C.c_bool, C.c_void_p, C.c_uint32)


class MeshHit(C.Structure):
    _fields_ = [("fraction", C.c_float), ("normal", Vec3), ("triangle", C.c_uint32),
                ("original", C.c_uint32), ("back", C.c_int32)]


class CastHit(C.Structure):
    """collide.h:246-253"""
    _fields_ = [("point", Vec3), ("normal", Vec3), ("fraction", C.c_float),
                ("triangle", C.c_uint32)]


class QueryFilter(C.Structure):
    """query.h:61-67"""
    _fields_ = [("category", C.c_uint64), ("mask", C.c_uint64), ("body_types", C.c_uint32),
                ("ignore", C.c_uint32), ("sensors", C.c_int32)]


class QueryHit(C.Structure):
    """query.h:73-80"""
    _fields_ = [("shape", C.c_uint32), ("body", C.c_uint32), ("point", DVec3), ("normal", Vec3),
                ("fraction", C.c_float), ("triangle", C.c_uint32)]


class QueryContact(C.Structure):
    """query.h:91-100"""
    _fields_ = [("shape", C.c_uint32), ("body", C.c_uint32), ("manifold", Manifold)]


QUERY_CAST_FN = C.CFUNCTYPE(C.c_float, C.c_void_p, C.POINTER(QueryHit))
QUERY_SHAPE_FN = C.CFUNCTYPE(C.c_bool, C.c_void_p, C.c_uint32, C.c_uint32)
QUERY_CONTACT_FN = C.CFUNCTYPE(C.c_bool, C.c_void_p, C.POINTER(QueryContact))


def bind_query(a):
    """The casts, overlaps and world queries of either library (collide.h:240-318,
    query.h:51-172)."""
    if hasattr(a, "world_raycast"):
        return a
    L = a.lib
    PG = C.POINTER(Geometry)
    PF = C.POINTER(QueryFilter)
    a.raycast = _bind(L, "nph_physics_raycast", C.c_bool, [PG, Pose, Vec3, Vec3, C.POINTER(CastHit)])
    a.shape_cast_margin = _bind(L, "nph_physics_shape_cast_margin", C.c_bool,
                                [PG, Pose, Vec3, C.c_float, PG, Pose, C.POINTER(CastHit)])
    a.overlap = _bind(L, "nph_physics_overlap", C.c_bool, [PG, Pose, PG, Pose])
    a.query_filter_default = _bind(L, "nph_physics_query_filter_default", QueryFilter, [])
    a.world_raycast = _bind(L, "nph_physics_world_raycast", None,
                            [C.c_void_p, DVec3, Vec3, PF, QUERY_CAST_FN, C.c_void_p])
    a.world_raycast_closest = _bind(L, "nph_physics_world_raycast_closest", C.c_bool,
                                    [C.c_void_p, DVec3, Vec3, PF, C.POINTER(QueryHit)])
    a.world_cast_shape_margin = _bind(L, "nph_physics_world_cast_shape_margin", None,
                                      [C.c_void_p, PG, DVec3, Quat, Vec3, C.c_float, PF,
                                       QUERY_CAST_FN, C.c_void_p])
    a.world_cast_shape_closest = _bind(L, "nph_physics_world_cast_shape_closest", C.c_bool,
                                       [C.c_void_p, PG, DVec3, Quat, Vec3, PF,
                                        C.POINTER(QueryHit)])
    a.world_overlap_shape = _bind(L, "nph_physics_world_overlap_shape", None,
                                  [C.c_void_p, PG, DVec3, Quat, PF, QUERY_SHAPE_FN, C.c_void_p])
    a.world_overlap_aabb = _bind(L, "nph_physics_world_overlap_aabb", None,
                                 [C.c_void_p, AABB, PF, QUERY_SHAPE_FN, C.c_void_p])
    a.world_collide_shape = _bind(L, "nph_physics_world_collide_shape", None,
                                  [C.c_void_p, PG, DVec3, Quat, C.c_float, PF, QUERY_CONTACT_FN,
                                   C.c_void_p])
    return a


def _hit_tuple(h):
    return (h.shape, h.body, h.point.x, h.point.y, h.point.z, h.normal.x, h.normal.y, h.normal.z,
            h.fraction, h.triangle)


def world_raycast_all(a, world_ptr, origin, translation, flt=None, answer=1.0, stop_after=None):
    """Every hit nph_physics_world_raycast reports, in order, as tuples; the
    callback answers `answer` (1: all hits; use the hit's fraction for the
    closest's search), or -1 after `stop_after` hits."""
    bind_query(a)
    out = []

    def fn(_u, hp):
        h = hp.contents
        out.append(_hit_tuple(h))
        if stop_after is not None and len(out) >= stop_after:
            return -1.0
        return h.fraction if answer == "fraction" else answer

    cb = QUERY_CAST_FN(fn)
    a.world_raycast(world_ptr, DVec3(*origin), Vec3(*translation),
                    C.byref(flt) if flt is not None else None, cb, None)
    return out


def world_cast_all(a, world_ptr, geometry, position, rotation, translation, margin=0.0, flt=None,
                   answer=1.0):
    bind_query(a)
    out = []

    def fn(_u, hp):
        h = hp.contents
        out.append(_hit_tuple(h))
        return h.fraction if answer == "fraction" else answer

    cb = QUERY_CAST_FN(fn)
    a.world_cast_shape_margin(world_ptr, C.byref(geometry), DVec3(*position), Quat(*rotation),
                              Vec3(*translation), margin,
                              C.byref(flt) if flt is not None else None, cb, None)
    return out


def world_overlaps(a, world_ptr, geometry=None, position=None, rotation=None, box=None, flt=None):
    """The (shape, body) pairs an overlap query reports, in order: a shape's
    with `geometry`, else the box's."""
    bind_query(a)
    out = []

    def fn(_u, shape, body):
        out.append((shape, body))
        return True

    cb = QUERY_SHAPE_FN(fn)
    f = C.byref(flt) if flt is not None else None
    if geometry is not None:
        a.world_overlap_shape(world_ptr, C.byref(geometry), DVec3(*position), Quat(*rotation), f,
                              cb, None)
    else:
        a.world_overlap_aabb(world_ptr, AABB(Vec3(*box[0]), Vec3(*box[1])), f, cb, None)
    return out


def world_contacts_with(a, world_ptr, geometry, position, rotation, speculative, flt=None):
    """The manifolds nph_physics_world_collide_shape reports, in order."""
    bind_query(a)
    out = []

    def fn(_u, cp):
        c = cp.contents
        m = c.manifold
        pts = tuple((m.points[k].point.x, m.points[k].point.y, m.points[k].point.z,
                     m.points[k].separation, m.points[k].id) for k in range(m.point_count))
        out.append((c.shape, c.body, m.normal.x, m.normal.y, m.normal.z, pts))
        return True

    cb = QUERY_CONTACT_FN(fn)
    a.world_collide_shape(world_ptr, C.byref(geometry), DVec3(*position), Quat(*rotation),
                          speculative, C.byref(flt) if flt is not None else None, cb, None)
    return out


def mesh_overlap(a, mesh_ptr, lo, hi, stop_after=None):
    """The triangles nph_physics_mesh_overlap visits for the box (lo, hi),
    in order; with `stop_after`, the visit returns false at that count."""
    bind_mesh(a)
    seen = []

    def visit(_user, t):
        seen.append(int(t))
        return stop_after is None or len(seen) < stop_after

    cb = MESH_VISIT(visit)
    a.mesh_overlap(mesh_ptr, AABB(Vec3(*lo), Vec3(*hi)), cb, None)
    return seen


def build_mesh(a, vertices, indices, weld_tolerance: float = 0.0, active_edge_angle: float = 0.0,
               two_sided: bool = False):
    """(result code, the block's bytes or None) of the library's
    nph_physics_mesh_build over float32 (n, 3) vertices and uint32 (m, 3)
    indices; the block is freed."""
    bind_mesh(a)
    v = np.ascontiguousarray(vertices, dtype=np.float32)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    d = MeshDesc(v.ctypes.data, int(v.shape[0]), idx.ctypes.data, int(idx.size // 3),
                 float(weld_tolerance), float(active_edge_angle), int(bool(two_sided)))
    out = C.c_void_p()
    rc = a.mesh_build(C.byref(d), a.allocator(), C.byref(out))
    if rc != 0:
        return rc, None
    size = C.cast(out, C.POINTER(C.c_uint32))[0]
    block = C.string_at(out, size)
    a.mesh_free(out, a.allocator())
    return rc, block


def build_hull(a, points, max_vertices: int = 0, tolerance: float = 0.0):
    """(result code, the block's bytes or None) of the library's
    nph_physics_hull_build over an (n, 3) float32 array; the block is freed."""
    bind_hull(a)
    pts = np.ascontiguousarray(points, dtype=np.float32)
    d = HullDesc(pts.ctypes.data, int(pts.shape[0]), int(max_vertices), float(tolerance))
    out = C.c_void_p()
    rc = a.hull_build(C.byref(d), a.allocator(), C.byref(out))
    if rc != 0:
        return rc, None
    size = C.cast(out, C.POINTER(C.c_uint32))[0]
    block = C.string_at(out, size)
    a.hull_free(out, a.allocator())
    return rc, block


class WorldDesc(C.Structure):
    _fields_ = [("alloc", C.c_void_p), ("gravity", Vec3), ("solver", C.c_int),
                ("substeps", C.c_uint32), ("contact_hertz", C.c_float),
                ("contact_damping_ratio", C.c_float), ("contact_push_max_velocity", C.c_float),
                ("restitution_threshold", C.c_float), ("max_linear_speed", C.c_float),
                ("joint_hertz", C.c_float), ("joint_damping_ratio", C.c_float),
                ("time_to_sleep", C.c_float), ("enable_sleep", C.c_int32),
                ("enable_continuous", C.c_int32), ("body_capacity", C.c_uint32),
                ("shape_capacity", C.c_uint32), ("contact_capacity", C.c_uint32)]


class BodyDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("position", DVec3), ("rotation", Quat),
                ("linear_velocity", Vec3), ("angular_velocity", Vec3),
                ("linear_damping", C.c_float), ("angular_damping", C.c_float),
                ("gravity_scale", C.c_float), ("sleep_threshold", C.c_float),
                ("locks", C.c_uint32), ("enable_sleep", C.c_int32), ("awake", C.c_int32),
                ("gyroscopic", C.c_int32), ("bullet", C.c_int32), ("user_data", C.c_uint64)]


class Filter(C.Structure):
    _fields_ = [("category", C.c_uint64), ("mask", C.c_uint64), ("group", C.c_int32)]


class ShapeDesc(C.Structure):
    _fields_ = [("geometry", Geometry), ("pose", Pose), ("density", C.c_float),
                ("friction", C.c_float), ("restitution", C.c_float),
                ("friction_combine", C.c_int), ("restitution_combine", C.c_int),
                ("filter", Filter), ("contact_events", C.c_int32), ("hit_events", C.c_int32),
                ("sensor", C.c_int32), ("sensor_events", C.c_int32), ("ghost", C.c_int32),
                ("user_data", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "bodies", "awake_bodies", "shapes", "contacts", "touching_contacts", "manifold_points",
        "joints", "islands", "awake_islands", "colors_used", "overflow_constraints",
        "moved_proxies", "new_pairs", "swept_bodies", "impact_bodies", "sensors",
        "sensor_visitors", "workers")] + [("steps", C.c_uint64)] + [(n, C.c_double) for n in (
        "step_ms", "pairs_ms", "collide_ms", "solve_ms", "finalize_ms", "sleep_ms",
        "sensors_ms")]


class ContactPointInfo(C.Structure):
    _fields_ = [("anchor_a", Vec3), ("anchor_b", Vec3), ("separation", C.c_float),
                ("id", C.c_uint32), ("normal_impulse", C.c_float),
                ("max_normal_impulse", C.c_float), ("tangent_impulse", Vec3),
                ("relative_velocity", C.c_float)]


class ContactInfo(C.Structure):
    _fields_ = [("record", C.c_uint32), ("shape_a", C.c_uint32), ("shape_b", C.c_uint32),
                ("body_a", C.c_uint32), ("body_b", C.c_uint32), ("flags", C.c_uint32),
                ("color", C.c_uint32), ("color_slot", C.c_uint32), ("parent", C.c_uint32),
                ("child", C.c_uint32), ("friction", C.c_float), ("restitution", C.c_float),
                ("normal", Vec3), ("point_count", C.c_uint32),
                ("points", ContactPointInfo * 4)]


class PairEvent(C.Structure):
    """Contact begin and end events (world.h:175-186)."""
    _fields_ = [("shape_a", C.c_uint32), ("shape_b", C.c_uint32)]


class HitEvent(C.Structure):
    _fields_ = [("shape_a", C.c_uint32), ("shape_b", C.c_uint32), ("point", DVec3),
                ("normal", Vec3), ("approach_speed", C.c_float)]


class MoveEvent(C.Structure):
    _fields_ = [("body", C.c_uint32), ("fell_asleep", C.c_uint32), ("user_data", C.c_uint64),
                ("position", DVec3), ("rotation", Quat), ("linear_velocity", Vec3),
                ("angular_velocity", Vec3)]


class SensorEvent(C.Structure):
    _fields_ = [("sensor", C.c_uint32), ("visitor", C.c_uint32),
                ("sensor_user_data", C.c_uint64), ("visitor_user_data", C.c_uint64)]


class Events(C.Structure):
    _fields_ = [("begin", C.POINTER(PairEvent)), ("begin_count", C.c_uint32),
                ("end", C.POINTER(PairEvent)), ("end_count", C.c_uint32),
                ("hit", C.POINTER(HitEvent)), ("hit_count", C.c_uint32),
                ("move", C.POINTER(MoveEvent)), ("move_count", C.c_uint32),
                ("sensor_begin", C.POINTER(SensorEvent)), ("sensor_begin_count", C.c_uint32),
                ("sensor_end", C.POINTER(SensorEvent)), ("sensor_end_count", C.c_uint32)]


class BallJoint(C.Structure):
    _fields_ = [("enable_swing_limit", C.c_int32), ("swing_y", C.c_float), ("swing_z", C.c_float),
                ("enable_twist_limit", C.c_int32), ("twist_lower", C.c_float),
                ("twist_upper", C.c_float), ("friction_torque", C.c_float)]


class HingeJoint(C.Structure):
    _fields_ = [("enable_limit", C.c_int32), ("lower", C.c_float), ("upper", C.c_float),
                ("enable_motor", C.c_int32), ("motor_speed", C.c_float),
                ("max_motor_torque", C.c_float), ("enable_spring", C.c_int32),
                ("spring_hertz", C.c_float), ("spring_damping_ratio", C.c_float),
                ("spring_target", C.c_float)]


class SliderJoint(C.Structure):
    _fields_ = [("enable_limit", C.c_int32), ("lower", C.c_float), ("upper", C.c_float),
                ("enable_motor", C.c_int32), ("motor_speed", C.c_float),
                ("max_motor_force", C.c_float), ("enable_spring", C.c_int32),
                ("spring_hertz", C.c_float), ("spring_damping_ratio", C.c_float),
                ("spring_target", C.c_float)]


class DistanceJoint(C.Structure):
    _fields_ = [("length", C.c_float), ("enable_limit", C.c_int32), ("min_length", C.c_float),
                ("max_length", C.c_float), ("enable_spring", C.c_int32),
                ("spring_hertz", C.c_float), ("spring_damping_ratio", C.c_float),
                ("enable_motor", C.c_int32), ("motor_speed", C.c_float),
                ("max_motor_force", C.c_float)]


class D6Axis(C.Structure):
    _fields_ = [("motion", C.c_int), ("lower", C.c_float), ("upper", C.c_float),
                ("enable_motor", C.c_int32), ("motor_speed", C.c_float),
                ("max_motor_force", C.c_float), ("enable_spring", C.c_int32),
                ("spring_hertz", C.c_float), ("spring_damping_ratio", C.c_float),
                ("spring_target", C.c_float)]


class D6Joint(C.Structure):
    _fields_ = [("axes", D6Axis * 6)]


class _JointUnion(C.Union):
    _fields_ = [("ball", BallJoint), ("hinge", HingeJoint), ("slider", SliderJoint),
                ("distance", DistanceJoint), ("d6", D6Joint)]


class JointDesc(C.Structure):
    _anonymous_ = ("u",)
    _fields_ = [("type", C.c_int), ("body_a", C.c_uint32), ("body_b", C.c_uint32),
                ("frame_a", Pose), ("frame_b", Pose), ("collide_connected", C.c_int32),
                ("user_data", C.c_uint64), ("u", _JointUnion)]


JOINT_BALL, JOINT_HINGE, JOINT_FIXED, JOINT_SLIDER, JOINT_DISTANCE, JOINT_D6 = range(6)

assert C.sizeof(Pose) == 28 and C.sizeof(Geometry) == 24
assert C.sizeof(ContactPointInfo) == 56 and C.sizeof(ContactInfo) == 288

_U32 = C.c_uint32


def _bind(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


class Api:
    """The bound functions of one library exporting the reference's API."""

    def __init__(self, path: str, allocator_symbol: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"physics library not built: {path}")
        # RTLD_LOCAL: the product and the reference export the same names.
        self.lib = C.CDLL(path, mode=os.RTLD_NOW | os.RTLD_LOCAL)
        L = self.lib
        self.path = path
        self.allocator = _bind(L, allocator_symbol, C.c_void_p, [])
        self.world_desc_default = _bind(L, "nph_physics_world_desc_default", WorldDesc, [])
        self.world_create = _bind(L, "nph_physics_world_create", C.c_int,
                                  [C.POINTER(WorldDesc), C.POINTER(C.c_void_p)])
        self.world_destroy = _bind(L, "nph_physics_world_destroy", None, [C.c_void_p])
        self.world_step = _bind(L, "nph_physics_world_step", C.c_int,
                                [C.c_void_p, C.c_float, C.c_void_p])
        self.world_hash = _bind(L, "nph_physics_world_hash", C.c_uint64, [C.c_void_p])
        self.world_stats = _bind(L, "nph_physics_world_stats", Stats, [C.c_void_p])
        self.world_events = _bind(L, "nph_physics_world_events", Events, [C.c_void_p])
        self.world_set_hit_threshold = _bind(L, "nph_physics_world_set_hit_threshold", None,
                                             [C.c_void_p, C.c_float])
        self.world_validate = _bind(L, "nph_physics_world_validate", C.c_bool, [C.c_void_p])
        self.world_contacts = _bind(L, "nph_physics_world_contacts", _U32,
                                    [C.c_void_p, C.POINTER(ContactInfo), _U32])
        self.body_desc_default = _bind(L, "nph_physics_body_desc_default", BodyDesc, [])
        self.body_create = _bind(L, "nph_physics_body_create", C.c_int,
                                 [C.c_void_p, C.POINTER(BodyDesc), C.POINTER(_U32)])
        self.body_position = _bind(L, "nph_physics_body_position", DVec3, [C.c_void_p, _U32])
        self.body_rotation = _bind(L, "nph_physics_body_rotation", Quat, [C.c_void_p, _U32])
        self.body_center = _bind(L, "nph_physics_body_center_of_mass", DVec3,
                                 [C.c_void_p, _U32])
        self.body_linear_velocity = _bind(L, "nph_physics_body_linear_velocity", Vec3,
                                          [C.c_void_p, _U32])
        self.body_angular_velocity = _bind(L, "nph_physics_body_angular_velocity", Vec3,
                                           [C.c_void_p, _U32])
        self.body_mass_data = _bind(L, "nph_physics_body_mass_data", MassData,
                                    [C.c_void_p, _U32])
        self.body_set_linear_velocity = _bind(L, "nph_physics_body_set_linear_velocity", None,
                                              [C.c_void_p, _U32, Vec3])
        self.body_set_angular_velocity = _bind(L, "nph_physics_body_set_angular_velocity", None,
                                               [C.c_void_p, _U32, Vec3])
        self.body_destroy = _bind(L, "nph_physics_body_destroy", None, [C.c_void_p, _U32])
        self.body_set_transform = _bind(L, "nph_physics_body_set_transform", None,
                                        [C.c_void_p, _U32, DVec3, Quat])
        self.body_apply_force = _bind(L, "nph_physics_body_apply_force", None,
                                      [C.c_void_p, _U32, Vec3, DVec3])
        self.body_apply_force_at_center = _bind(L, "nph_physics_body_apply_force_at_center", None,
                                                [C.c_void_p, _U32, Vec3])
        self.body_apply_torque = _bind(L, "nph_physics_body_apply_torque", None,
                                       [C.c_void_p, _U32, Vec3])
        self.body_apply_linear_impulse = _bind(L, "nph_physics_body_apply_linear_impulse", None,
                                               [C.c_void_p, _U32, Vec3, DVec3])
        self.body_apply_angular_impulse = _bind(L, "nph_physics_body_apply_angular_impulse", None,
                                                [C.c_void_p, _U32, Vec3])
        self.body_set_mass_data = _bind(L, "nph_physics_body_set_mass_data", None,
                                        [C.c_void_p, _U32, C.POINTER(MassData)])
        self.body_set_bullet = _bind(L, "nph_physics_body_set_bullet", None,
                                     [C.c_void_p, _U32, C.c_bool])
        self.body_is_bullet = _bind(L, "nph_physics_body_is_bullet", C.c_bool, [C.c_void_p, _U32])
        self.body_set_damping = _bind(L, "nph_physics_body_set_damping", None,
                                      [C.c_void_p, _U32, C.c_float, C.c_float])
        self.body_set_gravity_scale = _bind(L, "nph_physics_body_set_gravity_scale", None,
                                            [C.c_void_p, _U32, C.c_float])
        self.body_set_locks = _bind(L, "nph_physics_body_set_locks", None,
                                    [C.c_void_p, _U32, _U32])
        self.body_set_sleep = _bind(L, "nph_physics_body_set_sleep", None,
                                    [C.c_void_p, _U32, C.c_bool, C.c_float])
        self.body_set_awake = _bind(L, "nph_physics_body_set_awake", None,
                                    [C.c_void_p, _U32, C.c_bool])
        self.body_is_awake = _bind(L, "nph_physics_body_is_awake", C.c_bool, [C.c_void_p, _U32])
        self.body_set_user_data = _bind(L, "nph_physics_body_set_user_data", None,
                                        [C.c_void_p, _U32, C.c_uint64])
        self.body_user_data = _bind(L, "nph_physics_body_user_data", C.c_uint64,
                                    [C.c_void_p, _U32])
        self.body_shapes = _bind(L, "nph_physics_body_shapes", _U32,
                                 [C.c_void_p, _U32, C.POINTER(_U32), _U32])
        self.shape_destroy = _bind(L, "nph_physics_shape_destroy", None, [C.c_void_p, _U32])
        self.shape_friction = _bind(L, "nph_physics_shape_friction", C.c_float, [C.c_void_p, _U32])
        self.shape_set_friction = _bind(L, "nph_physics_shape_set_friction", None,
                                        [C.c_void_p, _U32, C.c_float])
        self.shape_set_restitution = _bind(L, "nph_physics_shape_set_restitution", None,
                                           [C.c_void_p, _U32, C.c_float])
        self.shape_set_friction_combine = _bind(L, "nph_physics_shape_set_friction_combine", None,
                                                [C.c_void_p, _U32, C.c_int])
        self.shape_friction_combine = _bind(L, "nph_physics_shape_friction_combine", C.c_int,
                                            [C.c_void_p, _U32])
        self.shape_set_density = _bind(L, "nph_physics_shape_set_density", None,
                                       [C.c_void_p, _U32, C.c_float])
        self.shape_density = _bind(L, "nph_physics_shape_density", C.c_float, [C.c_void_p, _U32])
        self.shape_set_filter = _bind(L, "nph_physics_shape_set_filter", None,
                                      [C.c_void_p, _U32, Filter])
        self.shape_filter = _bind(L, "nph_physics_shape_filter", Filter, [C.c_void_p, _U32])
        self.combine_friction = _bind(L, "nph_physics_combine_friction", C.c_float,
                                      [C.c_float, C.c_int, C.c_float, C.c_int])
        self.combine_restitution = _bind(L, "nph_physics_combine_restitution", C.c_float,
                                         [C.c_float, C.c_int, C.c_float, C.c_int])
        self.shape_desc_default = _bind(L, "nph_physics_shape_desc_default", ShapeDesc,
                                        [Geometry])
        self.shape_create = _bind(L, "nph_physics_shape_create", C.c_int,
                                  [C.c_void_p, _U32, C.POINTER(ShapeDesc), C.POINTER(_U32)])
        self.shape_aabb = _bind(L, "nph_physics_shape_aabb", AABB, [C.c_void_p, _U32])
        self.geometry_mass = _bind(L, "nph_physics_geometry_mass", MassData,
                                   [C.POINTER(Geometry), C.c_float])
        self.joint_desc_default = _bind(L, "nph_physics_joint_desc_default", JointDesc,
                                        [C.c_int, _U32, _U32])
        self.joint_frames_at = _bind(L, "nph_physics_joint_frames_at", None,
                                     [C.c_void_p, _U32, _U32, DVec3, Vec3, C.POINTER(Pose),
                                      C.POINTER(Pose)])
        self.joint_create = _bind(L, "nph_physics_joint_create", C.c_int,
                                  [C.c_void_p, C.POINTER(JointDesc), C.POINTER(_U32)])
        self.joint_destroy = _bind(L, "nph_physics_joint_destroy", None, [C.c_void_p, _U32])
        self.joint_hinge_angle = _bind(L, "nph_physics_joint_hinge_angle", C.c_float,
                                       [C.c_void_p, _U32])
        self.joint_length = _bind(L, "nph_physics_joint_length", C.c_float, [C.c_void_p, _U32])
        self.joint_reaction = _bind(L, "nph_physics_joint_reaction", None,
                                    [C.c_void_p, _U32, C.POINTER(Vec3), C.POINTER(Vec3)])


_APIS: dict = {}


def api(path: str = PRODUCT_LIB, allocator_symbol: str = "nph_amd_allocator_system") -> Api:
    key = (path, allocator_symbol)
    if key not in _APIS:
        _APIS[key] = Api(path, allocator_symbol)
    return _APIS[key]


def product_api() -> Api:
    """The HIP world step.  Raises if the extension is not built: there is no
    CPU fallback."""
    return api(PRODUCT_LIB, "nph_amd_allocator_system")


def geometry_sphere(radius):
    g = Geometry()
    g.type = SHAPE_SPHERE
    g.u.radius = radius
    return g


def geometry_capsule(half_height, radius):
    g = Geometry()
    g.type = SHAPE_CAPSULE
    g.u.capsule[0] = half_height
    g.u.capsule[1] = radius
    return g


def geometry_box(hx, hy, hz):
    g = Geometry()
    g.type = SHAPE_BOX
    g.u.half_extents = Vec3(hx, hy, hz)
    return g


def geometry_hull(block_ptr):
    g = Geometry()
    g.type = SHAPE_HULL
    g.u.ptr = block_ptr
    return g


def geometry_mesh(block_ptr):
    g = Geometry()
    g.type = SHAPE_MESH
    g.u.ptr = block_ptr
    return g


class MeshNode(C.Structure):
    """mesh.h:88-96"""
    _fields_ = [("min_x", C.c_float * 4), ("min_y", C.c_float * 4), ("min_z", C.c_float * 4),
                ("max_x", C.c_float * 4), ("max_y", C.c_float * 4), ("max_z", C.c_float * 4),
                ("child", C.c_uint32 * 4)]


class Mesh(C.Structure):
    """The mesh block's header (mesh.h:100-113)."""
    _fields_ = [("byte_size", C.c_uint32), ("vertex_count", C.c_uint32),
                ("triangle_count", C.c_uint32), ("node_count", C.c_uint32),
                ("nodes_offset", C.c_uint32), ("vertices_offset", C.c_uint32),
                ("triangles_offset", C.c_uint32), ("edges_offset", C.c_uint32),
                ("originals_offset", C.c_uint32), ("flags", C.c_uint32), ("bounds", AABB)]


def world_desc(a: Api, overrides: dict | None = None) -> WorldDesc:
    d = a.world_desc_default()
    d.alloc = a.allocator()
    for k, v in (overrides or {}).items():
        if k == "gravity":
            d.gravity = Vec3(*v)
        else:
            setattr(d, k, v)
    return d


class World:
    """A world of either library, with the calls the tests and the bench use.
    With `ptr`, it wraps a world a batch made and owns (close() leaves it)."""

    def __init__(self, a: Api, desc_overrides: dict | None = None, ptr=None):
        self.api = a
        self.desc = world_desc(a, desc_overrides)
        self.owned = ptr is None
        if ptr is None:
            w = C.c_void_p()
            r = a.world_create(C.byref(self.desc), C.byref(w))
            if r != OK:
                raise RuntimeError(f"nph_physics_world_create failed: {r}")
            ptr = w
        self.ptr = ptr
        self.bodies: list[int] = []
        self.shapes: list[int] = []
        self._keep = []  # hull blocks and other memory the world points at

    def close(self):
        if self.ptr:
            if self.owned:
                self.api.world_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_body(self, body_type, position, rotation=(0.0, 0.0, 0.0, 1.0), velocity=None,
                 angular_velocity=None, **fields):
        a = self.api
        bd = a.body_desc_default()
        bd.type = body_type
        bd.position = DVec3(*position)
        bd.rotation = Quat(*rotation)
        if velocity is not None:
            bd.linear_velocity = Vec3(*velocity)
        if angular_velocity is not None:
            bd.angular_velocity = Vec3(*angular_velocity)
        for k, v in fields.items():
            setattr(bd, k, v)
        h = _U32()
        r = a.body_create(self.ptr, C.byref(bd), C.byref(h))
        if r != OK:
            raise RuntimeError(f"nph_physics_body_create failed: {r}")
        self.bodies.append(h.value)
        return h.value

    def add_shape(self, body, geometry, friction=0.6, restitution=0.0, density=None,
                  pose=None, **fields):
        a = self.api
        sd = a.shape_desc_default(geometry)
        sd.friction = friction
        sd.restitution = restitution
        if density is not None:
            sd.density = density
        if pose is not None:
            sd.pose = Pose(Vec3(*pose[0]), Quat(*pose[1]))
        for k, v in fields.items():
            setattr(sd, k, v)
        h = _U32()
        r = a.shape_create(self.ptr, body, C.byref(sd), C.byref(h))
        if r != OK:
            raise RuntimeError(f"nph_physics_shape_create failed: {r}")
        self.shapes.append(h.value)
        return h.value

    def add_joint(self, joint_type, body_a, body_b, anchor, axis, **settings):
        """A joint whose frames meet at the world point `anchor` with X along
        `axis` (nph_physics_joint_frames_at), then `settings` applied: a
        dict per union member, e.g. ball={"enable_swing_limit": 1, ...}."""
        a = self.api
        d = a.joint_desc_default(joint_type, body_a, body_b)
        a.joint_frames_at(self.ptr, body_a, body_b, DVec3(*anchor), Vec3(*axis),
                          C.byref(d.frame_a), C.byref(d.frame_b))
        for member, values in settings.items():
            if isinstance(values, dict):
                target = getattr(d, member)
                for k, v in values.items():
                    setattr(target, k, v)
            else:
                setattr(d, member, values)
        h = _U32()
        r = a.joint_create(self.ptr, C.byref(d), C.byref(h))
        if r != OK:
            raise RuntimeError(f"nph_physics_joint_create failed: {r}")
        return h.value

    def step(self, dt, jobs=None):
        r = self.api.world_step(self.ptr, dt, jobs)
        if r != OK:
            raise RuntimeError(f"nph_physics_world_step failed: {r}")

    def hash(self) -> int:
        return int(self.api.world_hash(self.ptr))

    def stats(self) -> Stats:
        return self.api.world_stats(self.ptr)

    def events(self):
        """The last step's events as tuples: (begin, end, hit, move)."""
        e = self.api.world_events(self.ptr)
        begin = [(e.begin[i].shape_a, e.begin[i].shape_b) for i in range(e.begin_count)]
        end = [(e.end[i].shape_a, e.end[i].shape_b) for i in range(e.end_count)]
        hit = [(h.shape_a, h.shape_b, h.point.x, h.point.y, h.point.z, h.normal.x, h.normal.y,
                h.normal.z, h.approach_speed) for h in (e.hit[i] for i in range(e.hit_count))]
        move = [(m.body, m.fell_asleep, m.position.x, m.position.y, m.position.z, m.rotation.w,
                 m.linear_velocity.y, m.angular_velocity.x)
                for m in (e.move[i] for i in range(e.move_count))]
        return begin, end, hit, move

    def contacts(self):
        n = self.api.world_contacts(self.ptr, None, 0)
        buf = (ContactInfo * max(n, 1))()
        m = self.api.world_contacts(self.ptr, buf, n)
        assert m == n
        return buf[:n]

    def body_state(self) -> np.ndarray:
        """Per body (creation order): origin xyz (f64), rotation xyzw, v, w."""
        out = np.zeros((len(self.bodies), 13), dtype=np.float64)
        a = self.api
        for i, h in enumerate(self.bodies):
            p = a.body_position(self.ptr, h)
            q = a.body_rotation(self.ptr, h)
            v = a.body_linear_velocity(self.ptr, h)
            w = a.body_angular_velocity(self.ptr, h)
            out[i] = (p.x, p.y, p.z, q.x, q.y, q.z, q.w, v.x, v.y, v.z, w.x, w.y, w.z)
        return out


def contacts_as_arrays(contacts):
    """Canonical views of a contact readback: the pair list (min, max shape
    slot) and the contact-id list (slot_a, slot_b, chain index, point id) of
    touching manifolds, both sorted, plus a dict of per-contact values keyed
    by (shape_a, shape_b)."""
    pairs = sorted((min(c.shape_a, c.shape_b), max(c.shape_a, c.shape_b))
                   for c in contacts if c.parent == NULL32)
    ids = []
    by_pair = {}
    for c in contacts:
        if c.point_count > 0:
            for k in range(c.point_count):
                ids.append((c.shape_a, c.shape_b, 0 if c.parent == NULL32 else 1,
                            c.points[k].id))
        by_pair[(c.shape_a, c.shape_b)] = c
    ids.sort()
    return pairs, ids, by_pair


def canonical_digests(world) -> dict:
    """The canonical pair and contact-id lists of contacts_as_arrays, built
    with numpy from one contact readback (for worlds of 10^5 bodies), and the
    SHA-256 of each as little-endian u32 rows: pairs (min, max shape slot) of
    parent records, sorted; ids (shape_a, shape_b, chained, point id) of
    every touching manifold's points, sorted."""
    import hashlib
    n = world.api.world_contacts(world.ptr, None, 0)
    buf = (ContactInfo * max(n, 1))()
    world.api.world_contacts(world.ptr, buf, n)
    arr = np.frombuffer(buf, dtype=np.dtype(ContactInfo), count=n)
    parent = arr["parent"] == NULL32
    a, b = arr["shape_a"].astype(np.uint32), arr["shape_b"].astype(np.uint32)
    pairs = np.stack([np.minimum(a, b), np.maximum(a, b)], 1)[parent]
    pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))] if len(pairs) else pairs
    rows = []
    for k in range(4):
        sel = arr["point_count"] > k
        ids = arr["points"]["id"][sel, k].astype(np.uint32)
        rows.append(np.stack([a[sel], b[sel], (~parent[sel]).astype(np.uint32), ids], 1))
    ids = np.concatenate(rows) if rows else np.zeros((0, 4), np.uint32)
    if len(ids):
        ids = ids[np.lexsort((ids[:, 3], ids[:, 2], ids[:, 1], ids[:, 0]))]
    return {"pairs": int(len(pairs)), "ids": int(len(ids)),
            "pairs_sha256": hashlib.sha256(np.ascontiguousarray(pairs, "<u4").tobytes()).hexdigest(),
            "ids_sha256": hashlib.sha256(np.ascontiguousarray(ids, "<u4").tobytes()).hexdigest()}


# ---------------------------------------------------------------------------
# Batched worlds (nph_amd_batch_*: this library's extension)
# ---------------------------------------------------------------------------
class BatchTuning(C.Structure):
    _fields_ = [("pass_block", _U32), ("body_block", _U32), ("use_graph", _U32),
                ("groups", _U32)]


class BatchStats(C.Structure):
    _fields_ = [("solved_contacts", C.c_uint64), ("awake_bodies", C.c_uint64),
                ("step_ms", C.c_double), ("narrow_ms", C.c_double), ("pairs_ms", C.c_double),
                ("changes_ms", C.c_double),
                ("solve_ms", C.c_double), ("post_ms", C.c_double),
                ("device_solve_ms", C.c_double), ("device_passes_ms", C.c_double),
                ("device_collide_ms", C.c_double), ("wait_narrow_ms", C.c_double),
                ("wait_solve_ms", C.c_double), ("device_pairs_ms", C.c_double),
                ("pairs_trees_ms", C.c_double), ("pairs_search_ms", C.c_double),
                ("pairs_create_ms", C.c_double), ("rebuild_cpu_ms", C.c_double),
                ("colored_joints", C.c_uint64), ("colored_joint_ends", C.c_uint64),
                ("pairs_gather_ms", C.c_double), ("pairs_upload_ms", C.c_double),
                ("pairs_upload_bytes", C.c_uint64)]


def _bind_batch(a: Api):
    if hasattr(a, "batch_create"):
        return
    L = a.lib
    a.batch_create = _bind(L, "nph_amd_batch_create", C.c_int,
                           [C.POINTER(WorldDesc), _U32, C.POINTER(C.c_void_p)])
    a.batch_destroy = _bind(L, "nph_amd_batch_destroy", None, [C.c_void_p])
    a.batch_world = _bind(L, "nph_amd_batch_world", C.c_void_p, [C.c_void_p, _U32])
    a.batch_step = _bind(L, "nph_amd_batch_step", C.c_int, [C.c_void_p, C.c_float])
    a.batch_stats = _bind(L, "nph_amd_batch_stats", BatchStats, [C.c_void_p])
    a.batch_set_tuning = _bind(L, "nph_amd_batch_set_tuning", None,
                               [C.c_void_p, C.POINTER(BatchTuning)])
    a.batch_tuning = _bind(L, "nph_amd_batch_tuning", BatchTuning, [C.c_void_p])
    a.batch_autotune = _bind(L, "nph_amd_batch_autotune", C.c_int,
                             [C.c_void_p, C.c_float, _U32, C.POINTER(BatchTuning)])
    a.batch_time_kernels = _bind(L, "nph_amd_batch_time_kernels", None, [C.c_void_p, C.c_bool])
    a.batch_kernel_times = _bind(L, "nph_amd_batch_kernel_times", None,
                                 [C.c_void_p, C.POINTER(C.c_double)])


class Batch:
    """`count` worlds of one description stepped together; `worlds[k]` is
    filled and read like any World."""

    def __init__(self, a: Api, count: int, desc_overrides: dict | None = None):
        _bind_batch(a)
        self.api = a
        self.desc = world_desc(a, desc_overrides)
        b = C.c_void_p()
        r = a.batch_create(C.byref(self.desc), count, C.byref(b))
        if r != OK:
            raise RuntimeError(f"nph_amd_batch_create failed: {r}")
        self.ptr = b
        self.worlds = [World(a, desc_overrides, ptr=C.c_void_p(a.batch_world(b, k)))
                       for k in range(count)]

    def step(self, dt):
        r = self.api.batch_step(self.ptr, dt)
        if r != OK:
            raise RuntimeError(f"nph_amd_batch_step failed: {r}")

    def stats(self) -> BatchStats:
        return self.api.batch_stats(self.ptr)

    def tuning(self) -> BatchTuning:
        return self.api.batch_tuning(self.ptr)

    def set_tuning(self, pass_block=256, body_block=256, use_graph=0, groups=2):
        t = BatchTuning(pass_block, body_block, use_graph, groups)
        self.api.batch_set_tuning(self.ptr, C.byref(t))

    def autotune(self, dt, rounds=2) -> BatchTuning:
        t = BatchTuning()
        r = self.api.batch_autotune(self.ptr, dt, rounds, C.byref(t))
        if r != OK:
            raise RuntimeError(f"nph_amd_batch_autotune failed: {r}")
        return t

    def close(self):
        if self.ptr:
            for w in self.worlds:
                w.ptr = None
            self.api.batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
