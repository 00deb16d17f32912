// SPDX-License-Identifier: Apache-2.0
// Test-infrastructure shim compiled into oracle/_ref/libnph_ref.so next to
// the reference's own physics sources (oracle/Makefile).  It gives the
// reference world the readback that include/nepheline_physics.h declares as
// an extension (nph_physics_world_contacts), read white-box from the
// reference's records the way its own tests do (test_world.c:123-124), and
// a few helpers a ctypes caller needs (the system allocator, a job system,
// the solver's forced thread count, world_internal.h:653).
//
// Only tests, smoke() and bench.py's CPU-baseline leg load this library.

#include <nepheline/core/alloc.h>
#include <nepheline/jobs/jobs.h>

#include "world_internal.h"

#include <string.h>

// Same layout as NphPhysicsContactInfo in include/nepheline_physics.h.
typedef struct ShimPoint {
    NphVec3 anchor_a;
    NphVec3 anchor_b;
    f32 separation;
    u32 id;
    f32 normal_impulse;
    f32 max_normal_impulse;
    NphVec3 tangent_impulse;
    f32 relative_velocity;
} ShimPoint;

typedef struct ShimContact {
    u32 record;
    u32 shape_a;
    u32 shape_b;
    u32 body_a;
    u32 body_b;
    u32 flags;
    u32 color;
    u32 color_slot;
    u32 parent;
    u32 child;
    f32 friction;
    f32 restitution;
    NphVec3 normal;
    u32 point_count;
    ShimPoint points[4];
} ShimContact;

_Static_assert(sizeof(ShimPoint) == 56, "PhysContactPoint layout");
_Static_assert(sizeof(ShimContact) == 64 + 4 * 56, "contact info layout");

u32 nph_physics_world_contacts(const NphPhysicsWorld *w, ShimContact *out, u32 capacity)
{
    u32 n = 0U;
    for (u32 i = 0U; i < w->contacts.len; ++i) {
        const PhysContact *c = &w->contacts.data[i];
        if (c->shape_a == PHYS_NULL) {
            continue;
        }
        if (n < capacity) {
            ShimContact *o = &out[n];
            memset(o, 0, sizeof *o);
            o->record = i;
            o->shape_a = c->shape_a;
            o->shape_b = c->shape_b;
            o->body_a = c->body_a;
            o->body_b = c->body_b;
            o->flags = c->flags;
            o->color = c->color;
            o->color_slot = c->color_slot;
            o->parent = c->parent;
            o->child = c->child;
            o->friction = c->friction;
            o->restitution = c->restitution;
            o->normal = c->normal;
            o->point_count = c->point_count;
            for (u32 k = 0U; k < c->point_count && k < 4U; ++k) {
                const PhysContactPoint *p = &c->points[k];
                o->points[k].anchor_a = p->anchor_a;
                o->points[k].anchor_b = p->anchor_b;
                o->points[k].separation = p->separation;
                o->points[k].id = p->id;
                o->points[k].normal_impulse = p->normal_impulse;
                o->points[k].max_normal_impulse = p->max_normal_impulse;
                o->points[k].tangent_impulse = p->tangent_impulse;
                o->points[k].relative_velocity = p->relative_velocity;
            }
        }
        n += 1U;
    }
    return n;
}

NphAllocator *ref_shim_allocator(void)
{
    return nph_allocator_system();
}

NphJobSystem *ref_shim_jobs_create(u32 workers)
{
    NphJobSystemDesc d;
    memset(&d, 0, sizeof d);
    d.alloc = nph_allocator_system();
    d.worker_count = workers;
    NphJobSystem *js = NULL;
    if (nph_job_system_create(&d, &js) != NPH_RESULT_OK) {
        return NULL;
    }
    return js;
}

void ref_shim_jobs_destroy(NphJobSystem *js)
{
    nph_job_system_destroy(js);
}

void ref_shim_force_participants(NphPhysicsWorld *w, u32 n)
{
    w->forced_participants = n;
}

// The awake contact list, in its order (the order the narrow phase's changes
// are applied in, contact.c:677-687).
u32 ref_shim_awake_contacts(const NphPhysicsWorld *w, u32 *out, u32 capacity)
{
    for (u32 i = 0U; i < w->awake_contacts.len && i < capacity; ++i) {
        out[i] = w->awake_contacts.data[i];
    }
    return w->awake_contacts.len;
}

// A colour's contact list in order (c == 24: the overflow).
u32 ref_shim_color_contacts(const NphPhysicsWorld *w, u32 c, u32 *out, u32 capacity)
{
    const NphArrayU32 *list = c == PHYS_GRAPH_OVERFLOW ? &w->overflow : &w->colors[c].contacts;
    for (u32 i = 0U; i < list->len && i < capacity; ++i) {
        out[i] = list->data[i];
    }
    return list->len;
}

// Contacts the solver took in the last step: every colour's and the
// overflow's (the lanes build_groups fills, solver.c:810-843).
u32 ref_shim_graph_contacts(const NphPhysicsWorld *w)
{
    u32 n = w->overflow.len;
    for (u32 c = 0U; c < PHYS_GRAPH_COLORS; ++c) {
        n += w->colors[c].contacts.len;
    }
    return n;
}

// The reference's eight-lane contact kernels (solver_impl.h, the table of
// solver_sse4.c / solver_avx2.c) on caller-owned lanes, for pinning the
// oracle's restatement of them bit for bit.
void ref_shim_prepare(const PhysSolverBody *bodies, PhysContactLanes *g, PhysPointLanes *points,
                      const PhysPrepareLanes *in)
{
    phys_solver_kernels_best()->prepare(bodies, g, points, in);
}

void ref_shim_warm_start(PhysSolverBody *bodies, PhysContactLanes *g, PhysPointLanes *points)
{
    phys_solver_kernels_best()->warm_start(bodies, g, points);
}

void ref_shim_solve(PhysSolverBody *bodies, PhysContactLanes *g, PhysPointLanes *points, f32 inv_h,
                    f32 push_max, f32 slop, u32 use_bias)
{
    phys_solver_kernels_best()->solve(bodies, g, points, inv_h, push_max, slop, use_bias != 0U);
}

void ref_shim_restitution(PhysSolverBody *bodies, PhysContactLanes *g, PhysPointLanes *points,
                          f32 threshold)
{
    phys_solver_kernels_best()->restitution(bodies, g, points, threshold);
}

void ref_shim_make_soft(f32 hertz, f32 zeta, f32 h, f32 out[3])
{
    const PhysSoftness s = phys_make_soft(hertz, zeta, h);
    out[0] = s.bias_rate;
    out[1] = s.mass_scale;
    out[2] = s.impulse_scale;
}

u32 ref_shim_sizes(u32 which)
{
    switch (which) {
    case 0:
        return (u32)sizeof(PhysSolverBody);
    case 1:
        return (u32)sizeof(PhysContactLanes);
    case 2:
        return (u32)sizeof(PhysPointLanes);
    default:
        return (u32)sizeof(PhysPrepareLanes);
    }
}

// The reference's header-inline helpers (types.h:80-130, geometry.h:49-120,
// hull.h:80-110, mesh.h:60-150) as callable functions, for comparing this
// library's exported versions with them.
void ref_shim_pose_ops(NphPhysicsPose a, NphPhysicsPose b, NphVec3 p, NphPhysicsPose out_pose[3],
                       NphVec3 out_vec[4])
{
    out_pose[0] = nph_physics_pose_mul(a, b);
    out_pose[1] = nph_physics_pose_inv_mul(a, b);
    out_pose[2] = nph_physics_pose_inverse(a);
    out_vec[0] = nph_physics_pose_transform_point(a, p);
    out_vec[1] = nph_physics_pose_transform_vector(a, p);
    out_vec[2] = nph_physics_pose_inverse_transform_point(a, p);
    out_vec[3] = nph_physics_pose_inverse_transform_vector(a, p);
}

u32 ref_shim_mesh_leaf_ops(u32 first, u32 count, u32 child, u32 out[3])
{
    out[0] = nph_physics_mesh_child_is_leaf(child) ? 1U : 0U;
    out[1] = nph_physics_mesh_leaf_first(child);
    out[2] = nph_physics_mesh_leaf_count(child);
    return nph_physics_mesh_leaf(first, count);
}
