// SPDX-License-Identifier: Apache-2.0
// The per-item work of the world step, as the kernels run it: one body, one
// contact, or one lane of an eight-contact group at a time.  The data live
// in HBM as structure-of-arrays (NphStepData); the functions are
// __host__ __device__ so that every kernel is a thin loop over them.
//
//   body state, prepare, integration, finalize   solver.c:131-369
//   contact update (narrow phase + warm carry)   contact.c:296-577
//   the soft-step lane kernels                   solver_impl.h:107-668
//   group gather / store                         solver.c:394-499
//
// A group is eight consecutive lanes (the reference's PHYS_LANES); its
// decisions that span lanes (the points of the longest manifold, impact
// sweeps, restitution's bounced and loaded points) are computed once per
// group into NphGroup before the lanes run, so a lane never reads another
// lane's state while it is being written.
#ifndef NPH_STEP_H
#define NPH_STEP_H

#include "nph_surface.h"

#define NPH_LANES 8U
#define NPH_IMPACT_SPEED 1.0F
#define NPH_IMPACT_SWEEPS 7U
#define NPH_WARM_DISTANCE (2.0F * NPH_PHYSICS_LINEAR_SLOP)

// ---------------------------------------------------------------------------
// Records in HBM
// ---------------------------------------------------------------------------

enum {
    NPH_BODY_GYROSCOPIC = 1U << 2U, // world_internal.h:58-68
    NPH_BODY_BULLET = 1U << 5U,
    NPH_BODY_ENABLE_SLEEP = 1U << 1U,
    NPH_BODY_FAST = 1U << 6U, // moved far enough this step to be swept (continuous collision)
};

// A body slot's numeric state (PhysBody's numbers, world_internal.h:71-107).
typedef struct NphBodyRec {
    NphDVec3 origin;
    NphDVec3 center;
    NphQuat rotation;
    NphVec3 local_center;
    NphVec3 v;
    NphVec3 w;
    NphVec3 force;
    NphVec3 torque;
    f32 mass;
    f32 inv_mass;
    NphMat3 inertia;
    NphMat3 inv_inertia;
    f32 linear_damping;
    f32 angular_damping;
    f32 gravity_scale;
    f32 sleep_threshold;
    f32 sleep_time;
    f32 max_extent;
    f32 min_extent;
    u32 type;
    u32 flags; // NPH_BODY_* bits
    u32 locks;
    u32 shape_head; // first shape slot (NPH_NULL32: none)
    u32 enlarged;   // set by finalize when a shape left its fat box
} NphBodyRec;

// NPH_SHAPE_BIG_HULL: a hull of more vertices than the step's kernels are
// built for; its convex pairs run in the large-hull kernels.
enum { NPH_SHAPE_ENLARGED = 1U, NPH_SHAPE_BIG_HULL = 2U };

typedef struct NphShapeRec {
    NphGeom geom;
    NphPhysicsPose pose;
    NphAABB aabb;
    NphAABB fat_aabb;
    f32 friction;
    f32 restitution;
    u32 body;
    u32 next; // the body's next shape
    u32 combine; // friction_combine | restitution_combine << 8
    u32 flags;   // NPH_SHAPE_ENLARGED | NPH_SHAPE_BIG_HULL
} NphShapeRec;

typedef struct NphPointRec {
    NphVec3 anchor_a;
    NphVec3 anchor_b;
    f32 separation;
    u32 id;
    f32 normal_impulse;
    f32 max_normal_impulse;
    NphVec3 tangent_impulse;
    f32 relative_velocity;
} NphPointRec; // PhysContactPoint, 56 B

enum {
    NPH_CHANGE_STARTED = 1U,
    NPH_CHANGE_STOPPED = 2U,
    NPH_CHANGE_DISJOINT = 4U,
    NPH_CHANGE_CHILDREN = 8U, // a mesh pair's child record started or stopped
};
// A mesh pair's child k (1..3) in its parent's change code: started at bit
// NPH_CHANGE_CHILD_SHIFT + 2 (k - 1), stopped at the bit after.
#define NPH_CHANGE_CHILD_SHIFT 16U

typedef struct NphContactRec {
    u32 shape_a;
    u32 shape_b;
    u32 body_a;
    u32 body_b;
    u32 change; // NPH_CHANGE_*
    f32 friction;
    f32 restitution;
    NphCache cache;
    NphVec3 normal;
    u32 point_count;
    NphPointRec points[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
    u32 child; // a mesh pair's next record, or NPH_NULL32
} NphContactRec;

// The solver's body: four 16-byte rows, v | w | dp | dq (world_internal.h:387-400).
typedef struct NphSolverBody {
    f32 v[4];
    f32 w[4];
    f32 dp[4];
    f32 dq[4];
} NphSolverBody;

enum { NPH_PARAM_DYNAMIC = 1U, NPH_PARAM_GYROSCOPIC = 2U, NPH_PARAM_BULLET = 4U };

typedef struct NphSolverParams {
    NphVec3 linear_accel;
    f32 linear_damping;
    NphVec3 angular_accel;
    f32 angular_damping;
    NphMat3 inv_inertia;
    NphMat3 inertia_local;
    NphQuat q0;
    f32 max_angular_speed_sq;
    f32 max_linear_speed_sq;
    u32 body;
    u32 flags;
    u32 locks;
    u32 pad;
} NphSolverParams; // world_internal.h:403-417

// Per group: what spans its eight lanes.
typedef struct NphGroup {
    u32 lane_count;
    u32 point_count; // the longest manifold's
    u32 first_point; // its rows in the point arrays
    u32 impact;      // some lane has an impact (solve's extra sweeps)
    u32 loaded;      // restitution: bit k*8+l, lane l's point k bounces
} NphGroup;

// Lane fields (f32 unless noted), each an array of groups * 8.
enum {
    NPH_L_IMA = 0,
    NPH_L_IMB,
    NPH_L_NX,
    NPH_L_NY,
    NPH_L_NZ,
    NPH_L_T1X,
    NPH_L_T1Y,
    NPH_L_T1Z,
    NPH_L_T2X,
    NPH_L_T2Y,
    NPH_L_T2Z,
    NPH_L_FRICTION,
    NPH_L_RESTITUTION,
    NPH_L_BIAS,
    NPH_L_MASS_SCALE,
    NPH_L_IMPULSE_SCALE,
    NPH_L_FIELDS
};

// Point-row fields, each an array of rows * 8.
enum {
    NPH_P_RAX = 0,
    NPH_P_RAY,
    NPH_P_RAZ,
    NPH_P_RBX,
    NPH_P_RBY,
    NPH_P_RBZ,
    NPH_P_SEP,
    NPH_P_NMASS,
    NPH_P_NIMP,
    NPH_P_NMAX,
    NPH_P_RELV,
    NPH_P_TM0,
    NPH_P_TM1,
    NPH_P_TM2,
    NPH_P_TI1,
    NPH_P_TI2,
    NPH_P_AAN, // 3: I^-1 (r x axis), A, normal
    NPH_P_ABN = NPH_P_AAN + 3,
    NPH_P_AAT1 = NPH_P_ABN + 3,
    NPH_P_ABT1 = NPH_P_AAT1 + 3,
    NPH_P_AAT2 = NPH_P_ABT1 + 3,
    NPH_P_ABT2 = NPH_P_AAT2 + 3,
    NPH_P_FIELDS = NPH_P_ABT2 + 3
};

enum { NPH_LANE_WRITE_A = 1U, NPH_LANE_WRITE_B = 2U, NPH_LANE_IMPACT = 4U };

typedef struct NphJointConf NphJointConf;   // nph_joint.h
typedef struct NphJointState NphJointState;

// Everything a step's kernels touch, by device address.
typedef struct NphStepData {
    NphBodyRec *bodies;       // by body slot
    NphShapeRec *shapes;      // by shape slot
    NphContactRec *contacts;  // by contact record
    u32 *solver_of_body;      // body slot -> solver index (dummy for the rest)
    const u32 *awake_bodies;  // solver index -> body slot
    NphSolverBody *sbodies;   // awake bodies, then the dummy
    NphSolverParams *params;  // per awake body
    NphGroup *groups;
    u32 *lane_contact;        // groups * 8
    u32 *lane_body_a;
    u32 *lane_body_b;
    u32 *lane_flags;
    f32 *lanes;               // NPH_L_FIELDS * lane_stride
    f32 *points;              // NPH_P_FIELDS * point_stride
    u32 lane_stride;          // capacity in lanes
    u32 point_stride;         // capacity in point lanes (rows * 8)
    u32 body_count;           // awake bodies (the dummy's index)
    const NphJointConf *joint_conf; // by joint slot
    NphJointState *joint_state;     // by joint slot
    const u32 *step_joints;         // the graph's joints, colour by colour
    u32 joint_count;
} NphStepData;

// The world's numbers for one step (solver.c:961-1017).
typedef struct NphStepParams {
    NphVec3 gravity;
    f32 dt;
    f32 h;
    f32 inv_h;
    f32 push_max;
    f32 restitution_threshold;
    f32 max_linear_speed;
    f32 contact_bias, contact_mass, contact_impulse; // phys_make_soft(contact)
    f32 static_bias, static_mass, static_impulse;    // phys_make_soft(2x contact)
    f32 joint_bias, joint_mass, joint_impulse;       // phys_make_soft(joint)
    u32 substeps;
    u32 continuous; // enable_continuous: finalize marks fast bodies
} NphStepParams;

// solver.c:48-65
NPH_HD void make_soft(f32 hertz, f32 zeta, f32 h, f32 *bias, f32 *mass, f32 *impulse)
{
    if (hertz <= 0.0F) {
        *bias = 0.0F;
        *mass = 1.0F;
        *impulse = 0.0F;
        return;
    }
    const f32 omega = 2.0F * 3.14159265F * hertz;
    const f32 a1 = 2.0F * zeta + h * omega;
    const f32 a2 = h * omega * a1;
    const f32 a3 = 1.0F / (1.0F + a2);
    *bias = omega / a1;
    *mass = a2 * a3;
    *impulse = a3;
}

// ---------------------------------------------------------------------------
// Bodies (solver.c:82-369)
// ---------------------------------------------------------------------------
NPH_HD NphVec3 vec_without(NphVec3 v, u32 axis)
{
    if (axis == 0U) {
        v.x = 0.0F;
    } else if (axis == 1U) {
        v.y = 0.0F;
    } else {
        v.z = 0.0F;
    }
    return v;
}

NPH_HD NphMat3 lock_inertia(NphMat3 m, u32 locks)
{
    for (u32 axis = 0U; axis < 3U; ++axis) {
        if ((locks & ((u32)NPH_PHYSICS_LOCK_ANGULAR_X << axis)) == 0U) {
            continue;
        }
        m.c[axis] = v3_zero();
        for (u32 k = 0U; k < 3U; ++k) {
            m.c[k] = vec_without(m.c[k], axis);
        }
    }
    return m;
}

NPH_HD int inertia_is_isotropic(NphMat3 m)
{
    const f32 tol = 1e-6F * (nph_absf(m.c[0].x) + nph_absf(m.c[1].y) + nph_absf(m.c[2].z));
    return nph_absf(m.c[1].y - m.c[0].x) <= tol && nph_absf(m.c[2].z - m.c[0].x) <= tol &&
           nph_absf(m.c[0].y) <= tol && nph_absf(m.c[0].z) <= tol && nph_absf(m.c[1].z) <= tol;
}

NPH_HD void sb_set_v(NphSolverBody *s, NphVec3 v)
{
    s->v[0] = v.x;
    s->v[1] = v.y;
    s->v[2] = v.z;
}
NPH_HD void sb_set_w(NphSolverBody *s, NphVec3 w)
{
    s->w[0] = w.x;
    s->w[1] = w.y;
    s->w[2] = w.z;
}
NPH_HD NphVec3 sb_v(const NphSolverBody *s) { return v3(s->v[0], s->v[1], s->v[2]); }
NPH_HD NphVec3 sb_w(const NphSolverBody *s) { return v3(s->w[0], s->w[1], s->w[2]); }
NPH_HD NphVec3 sb_dp(const NphSolverBody *s) { return v3(s->dp[0], s->dp[1], s->dp[2]); }
NPH_HD NphQuat sb_dq(const NphSolverBody *s) { return q4(s->dq[0], s->dq[1], s->dq[2], s->dq[3]); }

NPH_HD void solver_body_dummy(NphSolverBody *s)
{
    for (u32 i = 0U; i < 4U; ++i) {
        s->v[i] = 0.0F;
        s->w[i] = 0.0F;
        s->dp[i] = 0.0F;
        s->dq[i] = 0.0F;
    }
    s->dq[3] = 1.0F;
}

// solver.c:131-168
NPH_HD void prepare_body(const NphStepParams *sp, const NphBodyRec *b, u32 slot,
                         NphSolverBody *sb, NphSolverParams *p)
{
    solver_body_dummy(sb);
    sb_set_v(sb, b->v);
    sb_set_w(sb, b->w);
    memset(p, 0, sizeof *p);
    p->body = slot;
    p->q0 = b->rotation;
    p->locks = b->locks;
    const f32 max_rotation = 0.25F * 3.14159265F * sp->inv_h;
    p->max_angular_speed_sq = max_rotation * max_rotation;
    p->max_linear_speed_sq = sp->max_linear_speed * sp->max_linear_speed;
    if (b->type == NPH_PHYSICS_BODY_DYNAMIC) {
        p->flags = NPH_PARAM_DYNAMIC;
        if ((b->flags & NPH_BODY_GYROSCOPIC) != 0U && !inertia_is_isotropic(b->inertia)) {
            p->flags |= NPH_PARAM_GYROSCOPIC;
        }
        if ((b->flags & NPH_BODY_BULLET) != 0U) {
            p->flags |= NPH_PARAM_BULLET;
        }
        p->linear_accel =
            v3_add(v3_scale(sp->gravity, b->gravity_scale), v3_scale(b->force, b->inv_mass));
        const NphMat3 r = m3_from_quat(b->rotation);
        p->inv_inertia = lock_inertia(m3_mul(m3_mul(r, b->inv_inertia), m3_transpose(r)), b->locks);
        p->angular_accel = m3_mul_v3(p->inv_inertia, b->torque);
        p->inertia_local = b->inertia;
        p->linear_damping = 1.0F / (1.0F + sp->h * b->linear_damping);
        p->angular_damping = 1.0F / (1.0F + sp->h * b->angular_damping);
    }
}

// solver.c:176-202 (Catto, GDC 2015): one Newton step in body space.
NPH_HD NphVec3 gyroscopic(NphVec3 w_world, NphQuat q, NphMat3 inertia, f32 h)
{
    const NphVec3 w = q_rotate_inv(q, w_world);
    const NphVec3 iw = m3_mul_v3(inertia, w);
    const NphVec3 f = v3_scale(v3_cross(w, iw), h);
    const NphVec3 e[3] = {v3(1.0F, 0.0F, 0.0F), v3(0.0F, 1.0F, 0.0F), v3(0.0F, 0.0F, 1.0F)};
    NphVec3 j[3];
    for (u32 k = 0U; k < 3U; ++k) {
        j[k] = v3_mul_add(inertia.c[k], v3_sub(v3_cross(w, inertia.c[k]), v3_cross(iw, e[k])), h);
    }
    const NphVec3 c12 = v3_cross(j[1], j[2]);
    const f32 det = v3_dot(j[0], c12);
    if (!(det > 0.0F)) {
        return w_world;
    }
    const f32 inv_det = 1.0F / det;
    const NphVec3 x = v3(v3_dot(f, c12) * inv_det, v3_dot(j[0], v3_cross(f, j[2])) * inv_det,
                         v3_dot(j[0], v3_cross(j[1], f)) * inv_det);
    return q_rotate(q, v3_sub(w, x));
}

NPH_HD NphVec3 lock_linear(NphVec3 v, u32 locks)
{
    if ((locks & NPH_PHYSICS_LOCK_LINEAR_X) != 0U) {
        v.x = 0.0F;
    }
    if ((locks & NPH_PHYSICS_LOCK_LINEAR_Y) != 0U) {
        v.y = 0.0F;
    }
    if ((locks & NPH_PHYSICS_LOCK_LINEAR_Z) != 0U) {
        v.z = 0.0F;
    }
    return v;
}

NPH_HD NphVec3 lock_angular(NphVec3 w, u32 locks)
{
    if ((locks & NPH_PHYSICS_LOCK_ANGULAR_X) != 0U) {
        w.x = 0.0F;
    }
    if ((locks & NPH_PHYSICS_LOCK_ANGULAR_Y) != 0U) {
        w.y = 0.0F;
    }
    if ((locks & NPH_PHYSICS_LOCK_ANGULAR_Z) != 0U) {
        w.z = 0.0F;
    }
    return w;
}

// solver.c:232-259
NPH_HD void integrate_velocity(const NphStepParams *sp, const NphSolverParams *p,
                               NphSolverBody *sb)
{
    if ((p->flags & NPH_PARAM_DYNAMIC) == 0U) {
        return;
    }
    const f32 h = sp->h;
    NphVec3 v = v3_scale(v3_mul_add(sb_v(sb), p->linear_accel, h), p->linear_damping);
    NphVec3 av = v3_scale(v3_mul_add(sb_w(sb), p->angular_accel, h), p->angular_damping);
    if ((p->flags & NPH_PARAM_GYROSCOPIC) != 0U) {
        av = gyroscopic(av, q_mul(sb_dq(sb), p->q0), p->inertia_local, h);
    }
    const f32 v_sq = v3_len_sq(v);
    if (v_sq > p->max_linear_speed_sq) {
        v = v3_scale(v, nph_sqrtf(p->max_linear_speed_sq / v_sq));
    }
    const f32 w_sq = v3_len_sq(av);
    if (w_sq > p->max_angular_speed_sq) {
        av = v3_scale(av, nph_sqrtf(p->max_angular_speed_sq / w_sq));
    }
    sb_set_v(sb, lock_linear(v, p->locks));
    sb_set_w(sb, lock_angular(av, p->locks));
}

// solver.c:261-270
NPH_HD void integrate_position(const NphStepParams *sp, const NphSolverParams *p,
                               NphSolverBody *sb)
{
    const NphVec3 v = lock_linear(sb_v(sb), p->locks);
    const NphVec3 w = lock_angular(sb_w(sb), p->locks);
    sb_set_v(sb, v);
    sb_set_w(sb, w);
    const NphVec3 dp = v3_mul_add(sb_dp(sb), v, sp->h);
    const NphQuat dq = q_integrate(sb_dq(sb), w, sp->h);
    sb->dp[0] = dp.x;
    sb->dp[1] = dp.y;
    sb->dp[2] = dp.z;
    sb->dq[0] = dq.x;
    sb->dq[1] = dq.y;
    sb->dq[2] = dq.z;
    sb->dq[3] = dq.w;
}

// shape.c:62-67: a shape's pose relative to `origin`.
NPH_HD NphPhysicsPose shape_pose_relative(const NphBodyRec *b, const NphShapeRec *s,
                                          NphDVec3 origin)
{
    const NphVec3 body = dv3_relative(b->origin, origin);
    return pose_make(v3_add(body, q_rotate(b->rotation, s->pose.position)),
                     q_mul(b->rotation, s->pose.rotation));
}

// shape.c:99-103
NPH_HD NphAABB shape_world_aabb(const NphBodyRec *b, const NphShapeRec *s)
{
    const NphPhysicsPose pose = shape_pose_relative(b, s, b->origin);
    return nph_world_box(b->origin, geom_aabb(&s->geom, pose));
}

NPH_HD u32 body_refresh_boxes(NphBodyRec *b, NphShapeRec *shapes);

// solver.c:304-369 without sleep's island logic (the world rejects sleep):
// the pose, velocities, forces cleared, the sleep timer, and the shapes'
// boxes.  Returns nonzero when a shape grew.
NPH_HD u32 finalize_body(const NphStepParams *sp, const NphSolverParams *p,
                         const NphSolverBody *sb, NphBodyRec *b, NphShapeRec *shapes)
{
    b->enlarged = 0U;
    b->v = sb_v(sb);
    b->w = sb_w(sb);
    b->rotation = q_normalize(q_mul(sb_dq(sb), p->q0));
    b->center = dv3_add_v3(b->center, sb_dp(sb));
    b->origin = dv3_add_v3(b->center, v3_neg(q_rotate(b->rotation, b->local_center)));
    b->force = v3_zero();
    b->torque = v3_zero();
    const f32 speed = nph_maxf(v3_len(sb_v(sb)), b->max_extent * v3_len(sb_w(sb)));
    if ((b->flags & NPH_BODY_ENABLE_SLEEP) == 0U || speed > b->sleep_threshold) {
        b->sleep_time = 0.0F;
    } else {
        b->sleep_time += sp->dt;
    }
    // Moved farther than half its thinnest shape's inner radius, or than
    // speculative contacts reach (solver.c:326-347): swept on the host
    // against the static shapes before its boxes are refreshed, unless a
    // bullet (refreshed now, swept after every body is final).
    if (sp->continuous != 0U && (p->flags & NPH_PARAM_DYNAMIC) != 0U) {
        const f32 limit =
            nph_minf(0.5F * b->min_extent, NPH_PHYSICS_AABB_MARGIN - NPH_PHYSICS_LINEAR_SLOP);
        const f32 half = 0.25F * limit * limit;
        const f32 arm = 3.14159265F * b->max_extent;
        const NphQuat dq = sb_dq(sb);
        const f32 turned = v3_len_sq(q_xyz(dq)) * arm * arm;
        const NphVec3 dp = sb_dp(sb);
        if ((v3_len_sq(dp) > half || turned > half) &&
            v3_len(dp) + 2.0F * nph_atan2f(v3_len(q_xyz(dq)), nph_absf(dq.w)) * b->max_extent >
                limit) {
            b->flags |= NPH_BODY_FAST;
            if ((p->flags & NPH_PARAM_BULLET) == 0U) {
                return 0U;
            }
        }
    }
    return body_refresh_boxes(b, shapes);
}

// solver.c:351-369: the shapes' boxes; a shape that left its fat box gets a
// new one and is flagged for the tree.
NPH_HD u32 body_refresh_boxes(NphBodyRec *b, NphShapeRec *shapes)
{
    u32 grew = 0U;
    for (u32 s = b->shape_head; s != NPH_NULL32; s = shapes[s].next) {
        NphShapeRec *shape = &shapes[s];
        shape->aabb = shape_world_aabb(b, shape);
        const NphAABB fat = shape->fat_aabb;
        const NphAABB box = shape->aabb;
        const int inside = fat.min.x <= box.min.x && fat.min.y <= box.min.y &&
                           fat.min.z <= box.min.z && box.max.x <= fat.max.x &&
                           box.max.y <= fat.max.y && box.max.z <= fat.max.z;
        if (!inside) {
            shape->fat_aabb = aabb_grow(box, NPH_PHYSICS_AABB_MARGIN);
            shape->flags |= NPH_SHAPE_ENLARGED;
            grew = 1U;
        }
    }
    b->enlarged = b->enlarged != 0U || grew != 0U ? 1U : 0U;
    return grew;
}

// ---------------------------------------------------------------------------
// Contacts: the narrow phase and the warm-start carry (contact.c:296-577)
// ---------------------------------------------------------------------------

// contact.c:485-514
NPH_HD f32 pair_speculative(const NphBodyRec *a, const NphBodyRec *b, f32 step_dt)
{
    const NphVec3 dv = v3_sub(b->v, a->v);
    const f32 turn_a = a->max_extent * a->max_extent;
    const f32 turn_b = b->max_extent * b->max_extent;
    const f32 room = (NPH_PHYSICS_SPECULATIVE_DISTANCE - NPH_PHYSICS_LINEAR_SLOP) /
                     nph_maxf(step_dt, 1e-9F);
    if (3.0F * (v3_len_sq(dv) + v3_len_sq(a->w) * turn_a + v3_len_sq(b->w) * turn_b) <=
        room * room) {
        return NPH_PHYSICS_SPECULATIVE_DISTANCE;
    }
    const f32 closing =
        step_dt * (v3_len(dv) + v3_len(a->w) * a->max_extent + v3_len(b->w) * b->max_extent);
    const f32 reach = closing + NPH_PHYSICS_LINEAR_SLOP;
    if (!(reach > NPH_PHYSICS_SPECULATIVE_DISTANCE)) {
        return NPH_PHYSICS_SPECULATIVE_DISTANCE;
    }
    return nph_minf(reach, NPH_PHYSICS_AABB_MARGIN);
}

// world_internal.h:119-155: the pair's friction and restitution.
NPH_HD f32 combine_values(f32 a, u32 mode_a, f32 b, u32 mode_b, int restitution)
{
    switch (mode_a > mode_b ? mode_a : mode_b) {
    case NPH_PHYSICS_COMBINE_AVERAGE:
        return 0.5F * (a + b);
    case NPH_PHYSICS_COMBINE_MIN:
        return nph_minf(a, b);
    case NPH_PHYSICS_COMBINE_MULTIPLY:
        return a * b;
    case NPH_PHYSICS_COMBINE_MAX:
        return nph_maxf(a, b);
    default:
        return restitution ? nph_maxf(a, b) : nph_sqrtf(a * b);
    }
}

NPH_HD void pair_surface(const NphShapeRec *a, const NphShapeRec *b, f32 *friction,
                         f32 *restitution)
{
    if ((a->combine | b->combine) == 0U) {
        *friction = nph_sqrtf(a->friction * b->friction);
        *restitution = nph_maxf(a->restitution, b->restitution);
        return;
    }
    *friction = combine_values(a->friction, a->combine & 0xFFU, b->friction, b->combine & 0xFFU, 0);
    *restitution =
        combine_values(a->restitution, a->combine >> 8U, b->restitution, b->combine >> 8U, 1);
}

#if defined(__cplusplus) // the narrow phase of a pair (nph_polytope.h is C++)
// A mesh pair (contact.c:378-483): the other shape in the mesh shape's
// frame, and what carrying its manifolds back to world anchors needs.
struct NphMeshFrame {
    int mesh_a;           // the mesh is the pair's shape A
    const NphGeom *other; // the convex shape
    const NphPhysicsMesh *mesh;
    NphQuat qm;           // the mesh shape's world rotation
    NphQuat rotation;     // the other shape's, in the mesh's frame
    NphDVec3 position;    // the other shape's place in the mesh's frame
    NphVec3 shape_from_x; // the other shape from its body's centre
    NphVec3 x_from_m;     // that centre from the mesh body's centre
    f32 speculative;
};

NPH_HD void mesh_frame(const NphContactRec *c, const NphShapeRec *shapes, const NphBodyRec *bodies,
                       f32 step_dt, NphMeshFrame *f)
{
    const NphShapeRec *sa = &shapes[c->shape_a];
    const NphShapeRec *sb = &shapes[c->shape_b];
    f->mesh_a = sa->geom.type == NPH_PHYSICS_SHAPE_MESH;
    const NphShapeRec *sm = f->mesh_a ? sa : sb;
    const NphShapeRec *sx = f->mesh_a ? sb : sa;
    const NphBodyRec *bm = &bodies[sm->body];
    const NphBodyRec *bx = &bodies[sx->body];
    f->other = &sx->geom;
    f->mesh = sm->geom.mesh;
    f->qm = q_mul(bm->rotation, sm->pose.rotation);
    const NphQuat qx = q_mul(bx->rotation, sx->pose.rotation);
    const NphDVec3 pm = dv3_add_v3(bm->origin, q_rotate(bm->rotation, sm->pose.position));
    const NphDVec3 px = dv3_add_v3(bx->origin, q_rotate(bx->rotation, sx->pose.position));
    f->rotation = q_mul(q_conj(f->qm), qx);
    f->position = drotate_inverse(f->qm, dv3(px.x - pm.x, px.y - pm.y, px.z - pm.z));
    f->shape_from_x = dv3_relative(px, bx->center);
    f->x_from_m = dv3_relative(bx->center, bm->center);
    f->speculative = pair_speculative(bm, bx, step_dt);
}

// The pair's manifolds (mesh frame, deepest first) into its chain of
// records: the deepest in the pair's record and the rest in its children;
// impulses carried over the whole chain, by id then by place (with the
// manifolds' normals within the merge angle). Change codes: the record's
// own, and each child's that started or stopped.
NPH_HD void mesh_store(NphContactRec *contacts, u32 contact, const NphShapeRec *shapes,
                       const NphMeshFrame *f, const NphMeshManifolds *mm)
{
    NphContactRec *c = &contacts[contact];
    // The chain's points of the last step.
    NphPointRec old[NPH_PHYSICS_MESH_MAX_MANIFOLDS * NPH_PHYSICS_MANIFOLD_MAX_POINTS];
    NphVec3 old_normals[NPH_PHYSICS_MESH_MAX_MANIFOLDS * NPH_PHYSICS_MANIFOLD_MAX_POINTS];
    u32 old_count = 0U;
    for (u32 k = contact; k != NPH_NULL32; k = contacts[k].child) {
        const NphContactRec *r = &contacts[k];
        for (u32 j = 0U; j < r->point_count; ++j) {
            old_normals[old_count] = r->normal;
            old[old_count++] = r->points[j];
        }
    }
    u32 claimed = 0U;
    u32 unmatched = 0U; // bit 4i + j: point j of manifold i
    f32 friction = 0.0F;
    f32 restitution = 0.0F;
    pair_surface(&shapes[c->shape_a], &shapes[c->shape_b], &friction, &restitution);
    u32 children = 0U;
    u32 rec = contact;
    for (u32 i = 0U; rec != NPH_NULL32; ++i) {
        NphContactRec *r = &contacts[rec];
        const int was = r->point_count > 0U;
        r->point_count = 0U;
        if (i < mm->count) {
            const NphManifold *m = &mm->manifolds[i];
            const NphVec3 n = q_rotate(f->qm, m->normal);
            r->normal = f->mesh_a ? v3_neg(n) : n;
            for (u32 j = 0U; j < m->point_count; ++j) {
                NphPointRec *p = &r->points[j];
                memset(p, 0, sizeof *p);
                const NphVec3 ax = v3_add(f->shape_from_x, q_rotate(f->qm, m->points[j].point));
                const NphVec3 am = v3_add(ax, f->x_from_m);
                p->anchor_a = f->mesh_a ? am : ax;
                p->anchor_b = f->mesh_a ? ax : am;
                p->separation = m->points[j].separation;
                p->id = m->points[j].id;
                int found = 0;
                for (u32 o = 0U; o < old_count; ++o) { // warm_by_id
                    if (old[o].id == p->id) {
                        p->normal_impulse = old[o].normal_impulse;
                        p->tangent_impulse = old[o].tangent_impulse;
                        claimed |= 1U << o;
                        found = 1;
                        break;
                    }
                }
                if (!found) {
                    unmatched |= 1U << (i * NPH_PHYSICS_MANIFOLD_MAX_POINTS + j);
                }
            }
            r->point_count = m->point_count;
        }
        r->friction = friction;
        r->restitution = restitution;
        const int touching = r->point_count > 0U;
        r->change = 0U;
        if (touching && !was) {
            r->change = NPH_CHANGE_STARTED;
        } else if (!touching && was) {
            r->change = NPH_CHANGE_STOPPED;
        }
        if (i > 0U && r->change != 0U) {
            children |= r->change << (NPH_CHANGE_CHILD_SHIFT + 2U * (i - 1U));
        }
        rec = r->child;
    }
    if (unmatched != 0U && claimed != (1U << old_count) - 1U) {
        rec = contact;
        for (u32 i = 0U; rec != NPH_NULL32 && i < mm->count; ++i) {
            NphContactRec *r = &contacts[rec];
            for (u32 j = 0U; j < r->point_count; ++j) {
                if ((unmatched & (1U << (i * NPH_PHYSICS_MANIFOLD_MAX_POINTS + j))) == 0U) {
                    continue;
                }
                NphPointRec *p = &r->points[j];
                u32 best = NPH_NULL32;
                f32 best_d = NPH_WARM_DISTANCE * NPH_WARM_DISTANCE;
                for (u32 o = 0U; o < old_count; ++o) { // warm_by_place
                    if ((claimed & (1U << o)) != 0U ||
                        v3_dot(old_normals[o], r->normal) < NPH_SURFACE_MERGE_COS) {
                        continue;
                    }
                    const f32 d = nph_maxf(v3_dist_sq(old[o].anchor_a, p->anchor_a),
                                           v3_dist_sq(old[o].anchor_b, p->anchor_b));
                    if (d < best_d) {
                        best_d = d;
                        best = o;
                    }
                }
                if (best != NPH_NULL32) {
                    p->normal_impulse = old[best].normal_impulse;
                    p->tangent_impulse = old[best].tangent_impulse;
                    claimed |= 1U << best;
                }
            }
            rec = r->child;
        }
    }
    if (children != 0U) {
        c->change |= NPH_CHANGE_CHILDREN | children;
    }
}

// Whether a contact's pair has a hull beyond the step kernels' capacity
// (the large-hull kernels take it).
NPH_HD int contact_has_big_hull(const NphContactRec *c, const NphShapeRec *shapes)
{
    return ((shapes[c->shape_a].flags | shapes[c->shape_b].flags) & NPH_SHAPE_BIG_HULL) != 0U;
}

// A convex pair's narrow-phase input (contact.c:516-530): both shapes'
// geometry and poses in A's centre-of-mass frame, the speculative distance,
// B's centre in that frame and the pair's cache.
typedef struct NphPairJob {
    NphGeom ga;
    NphGeom gb;
    NphPhysicsPose pa;
    NphPhysicsPose pb;
    NphVec3 ab;
    f32 speculative;
    NphCache cache;
} NphPairJob;

// False when the pair's fat boxes are apart (its change is then
// NPH_CHANGE_DISJOINT and the record is left as it is).
NPH_HD int contact_job(const NphContactRec *c, const NphShapeRec *shapes, const NphBodyRec *bodies,
                       f32 step_dt, NphPairJob *job)
{
    const NphShapeRec *sa = &shapes[c->shape_a];
    const NphShapeRec *sb = &shapes[c->shape_b];
    if (!aabb_overlaps(sa->fat_aabb, sb->fat_aabb)) {
        return 0;
    }
    const NphBodyRec *ba = &bodies[c->body_a];
    const NphBodyRec *bb = &bodies[c->body_b];
    const NphDVec3 origin = ba->center;
    job->ga = sa->geom;
    job->gb = sb->geom;
    job->pa = shape_pose_relative(ba, sa, origin);
    job->pb = shape_pose_relative(bb, sb, origin);
    job->ab = dv3_relative(bb->center, origin);
    job->speculative = pair_speculative(ba, bb, step_dt);
    job->cache = c->cache;
    return 1;
}

// contact.c:531-577: the new manifold into the record, impulses carried by
// id then by place, and the change of touching.
NPH_HD void contact_store(NphContactRec *c, const NphShapeRec *shapes, const NphPairJob *job,
                          const NphManifold *mp)
{
    const NphManifold m = *mp;
    const NphShapeRec *sa = &shapes[c->shape_a];
    const NphShapeRec *sb = &shapes[c->shape_b];
    const NphVec3 ab = job->ab;
    const int was_touching = c->point_count > 0U;
    c->change = 0U;
    c->cache = job->cache;
    NphPointRec points[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
    const u32 old_count = c->point_count;
    u32 claimed = 0U;
    u32 unmatched = 0U;
    for (u32 i = 0U; i < m.point_count; ++i) {
        NphPointRec *p = &points[i];
        memset(p, 0, sizeof *p);
        p->anchor_a = m.points[i].point;
        p->anchor_b = v3_sub(m.points[i].point, ab);
        p->separation = m.points[i].separation;
        p->id = m.points[i].id;
        int found = 0;
        for (u32 o = 0U; o < old_count; ++o) { // warm_by_id, contact.c:340-349
            if (c->points[o].id == p->id) {
                p->normal_impulse = c->points[o].normal_impulse;
                p->tangent_impulse = c->points[o].tangent_impulse;
                claimed |= 1U << o;
                found = 1;
                break;
            }
        }
        if (!found) {
            unmatched |= 1U << i;
        }
    }
    if (unmatched != 0U && claimed != (1U << old_count) - 1U) {
        for (u32 i = 0U; i < m.point_count; ++i) { // warm_by_place, contact.c:355-374
            if ((unmatched & (1U << i)) == 0U) {
                continue;
            }
            NphPointRec *p = &points[i];
            u32 best = NPH_NULL32;
            f32 best_d = NPH_WARM_DISTANCE * NPH_WARM_DISTANCE;
            for (u32 o = 0U; o < old_count; ++o) {
                if ((claimed & (1U << o)) != 0U) {
                    continue;
                }
                const f32 d = nph_maxf(v3_dist_sq(c->points[o].anchor_a, p->anchor_a),
                                       v3_dist_sq(c->points[o].anchor_b, p->anchor_b));
                if (d < best_d) {
                    best_d = d;
                    best = o;
                }
            }
            if (best != NPH_NULL32) {
                p->normal_impulse = c->points[best].normal_impulse;
                p->tangent_impulse = c->points[best].tangent_impulse;
                claimed |= 1U << best;
            }
        }
    }
    c->normal = m.normal;
    c->point_count = m.point_count;
    for (u32 i = 0U; i < m.point_count; ++i) {
        c->points[i] = points[i];
    }
    pair_surface(sa, sb, &c->friction, &c->restitution);
    const int touching = m.point_count > 0U;
    if (touching && !was_touching) {
        c->change = NPH_CHANGE_STARTED;
    } else if (!touching && was_touching) {
        c->change = NPH_CHANGE_STOPPED;
    }
}
#endif // __cplusplus

// ---------------------------------------------------------------------------
// The soft-step lanes (solver_impl.h:107-668), one lane of a group
// ---------------------------------------------------------------------------
typedef struct NphLaneBody {
    NphVec3 v;
    NphVec3 w;
    NphVec3 dp;
    NphQuat q;
} NphLaneBody;

NPH_HD NphLaneBody lane_gather(const NphSolverBody *s, int positions)
{
    NphLaneBody b;
    b.v = sb_v(s);
    b.w = sb_w(s);
    if (positions) {
        b.dp = sb_dp(s);
        b.q = sb_dq(s);
    } else {
        b.dp = v3_zero();
        b.q = q_identity();
    }
    return b;
}

// solver_impl.h:74-105: velocity rows written back, padding zeroed.
NPH_HD void lane_scatter(NphSolverBody *s, const NphLaneBody *b)
{
    s->v[0] = b->v.x;
    s->v[1] = b->v.y;
    s->v[2] = b->v.z;
    s->v[3] = 0.0F;
    s->w[0] = b->w.x;
    s->w[1] = b->w.y;
    s->w[2] = b->w.z;
    s->w[3] = 0.0F;
}

// solver_impl.h:122-142: vB + wB x rB - (vA + wA x rA)
NPH_HD NphVec3 lane_relative_velocity(const NphLaneBody *a, const NphLaneBody *b, NphVec3 ra,
                                      NphVec3 rb)
{
    const NphVec3 ca = v3_cross(a->w, ra);
    const NphVec3 cb = v3_cross(b->w, rb);
    return v3(b->v.x + cb.x - (a->v.x + ca.x), b->v.y + cb.y - (a->v.y + ca.y),
              b->v.z + cb.z - (a->v.z + ca.z));
}

// solver_impl.h:147-168
NPH_HD void lane_apply(NphLaneBody *a, NphLaneBody *b, f32 ma, f32 mb, f32 lambda, NphVec3 axis,
                       NphVec3 ang_a, NphVec3 ang_b)
{
    const f32 la = ma * lambda;
    const f32 lb = mb * lambda;
    a->v.x = a->v.x - la * axis.x;
    a->v.y = a->v.y - la * axis.y;
    a->v.z = a->v.z - la * axis.z;
    b->v.x = b->v.x + lb * axis.x;
    b->v.y = b->v.y + lb * axis.y;
    b->v.z = b->v.z + lb * axis.z;
    a->w.x = a->w.x - lambda * ang_a.x;
    a->w.y = a->w.y - lambda * ang_a.y;
    a->w.z = a->w.z - lambda * ang_a.z;
    b->w.x = b->w.x + lambda * ang_b.x;
    b->w.y = b->w.y + lambda * ang_b.y;
    b->w.z = b->w.z + lambda * ang_b.z;
}

// Field accessors: lane field f of lane L; point field f of row-lane R.
#define LF(d, f, L) ((d)->lanes[(u64)(f) * (d)->lane_stride + (L)])
#define PF(d, f, R) ((d)->points[(u64)(f) * (d)->point_stride + (R)])
#define PV3(d, f, R) v3(PF(d, f, R), PF(d, (f) + 1U, R), PF(d, (f) + 2U, R))

NPH_HD void pv3_store(const NphStepData *d, u32 f, u64 R, NphVec3 v)
{
    PF(d, f, R) = v.x;
    PF(d, f + 1U, R) = v.y;
    PF(d, f + 2U, R) = v.z;
}

// solver.c:394-473 + solver_impl.h:276-385 for lane l of group g: gathers the
// contact into the lane and derives its point rows.
NPH_HD void lane_prepare(const NphStepData *d, const NphStepParams *sp, u32 g, u32 l)
{
    const NphGroup *grp = &d->groups[g];
    const u32 L = g * NPH_LANES + l;
    const u32 dummy = d->body_count;
    const u32 P = grp->point_count;
    u32 ia = dummy;
    u32 ib = dummy;
    u32 flags = 0U;
    f32 ma = 0.0F;
    f32 mb = 0.0F;
    NphVec3 n = v3_zero();
    NphVec3 t1 = v3_zero();
    NphVec3 t2 = v3_zero();
    f32 friction = 0.0F;
    f32 restitution = 0.0F;
    f32 bias = 0.0F;
    f32 mass_scale = 0.0F;
    f32 impulse_scale = 0.0F;
    NphMat3 ii_a = m3_zero();
    NphMat3 ii_b = m3_zero();
    const NphContactRec *c = 0;
    if (l < grp->lane_count) {
        c = &d->contacts[d->lane_contact[L]];
        const NphBodyRec *a = &d->bodies[c->body_a];
        const NphBodyRec *b = &d->bodies[c->body_b];
        const int dyn_a = a->type == NPH_PHYSICS_BODY_DYNAMIC;
        const int dyn_b = b->type == NPH_PHYSICS_BODY_DYNAMIC;
        ia = d->solver_of_body[c->body_a];
        ib = d->solver_of_body[c->body_b];
        flags = (dyn_a ? NPH_LANE_WRITE_A : 0U) | (dyn_b ? NPH_LANE_WRITE_B : 0U);
        ma = dyn_a ? a->inv_mass : 0.0F;
        mb = dyn_b ? b->inv_mass : 0.0F;
        ii_a = dyn_a ? d->params[ia].inv_inertia : m3_zero();
        ii_b = dyn_b ? d->params[ib].inv_inertia : m3_zero();
        n = c->normal;
        t1 = v3_perpendicular(n);
        t2 = v3_cross(n, t1);
        friction = c->friction;
        restitution = c->restitution;
        const int stat = !dyn_a || !dyn_b;
        bias = stat ? sp->static_bias : sp->contact_bias;
        mass_scale = stat ? sp->static_mass : sp->contact_mass;
        impulse_scale = stat ? sp->static_impulse : sp->contact_impulse;
    }
    d->lane_body_a[L] = ia;
    d->lane_body_b[L] = ib;
    LF(d, NPH_L_IMA, L) = ma;
    LF(d, NPH_L_IMB, L) = mb;
    LF(d, NPH_L_NX, L) = n.x;
    LF(d, NPH_L_NY, L) = n.y;
    LF(d, NPH_L_NZ, L) = n.z;
    LF(d, NPH_L_T1X, L) = t1.x;
    LF(d, NPH_L_T1Y, L) = t1.y;
    LF(d, NPH_L_T1Z, L) = t1.z;
    LF(d, NPH_L_T2X, L) = t2.x;
    LF(d, NPH_L_T2Y, L) = t2.y;
    LF(d, NPH_L_T2Z, L) = t2.z;
    LF(d, NPH_L_FRICTION, L) = friction;
    LF(d, NPH_L_RESTITUTION, L) = restitution;
    LF(d, NPH_L_BIAS, L) = bias;
    LF(d, NPH_L_MASS_SCALE, L) = mass_scale;
    LF(d, NPH_L_IMPULSE_SCALE, L) = impulse_scale;
    // The prepare kernel proper.
    const NphLaneBody A = lane_gather(&d->sbodies[ia], 0);
    const NphLaneBody B = lane_gather(&d->sbodies[ib], 0);
    const f32 mass = ma + mb;
    const NphVec3 axes[3] = {n, t1, t2};
    u32 impact = 0U;
    for (u32 k = 0U; k < P; ++k) {
        const u64 R = (u64)(grp->first_point + k) * NPH_LANES + l;
        const int active = c != 0 && k < c->point_count;
        NphVec3 ra = v3_zero();
        NphVec3 rb = v3_zero();
        f32 sep = 0.0F;
        f32 nimp = 0.0F;
        NphVec3 timp = v3_zero();
        if (active) {
            ra = c->points[k].anchor_a;
            rb = c->points[k].anchor_b;
            sep = c->points[k].separation;
            nimp = c->points[k].normal_impulse;
            timp = c->points[k].tangent_impulse;
        }
        pv3_store(d, NPH_P_RAX, R, ra);
        pv3_store(d, NPH_P_RBX, R, rb);
        const f32 gap = v3_dot(v3(rb.x - ra.x, rb.y - ra.y, rb.z - ra.z), n);
        PF(d, NPH_P_SEP, R) = sep - gap;
        NphVec3 rxa[3];
        NphVec3 rxb[3];
        NphVec3 aa[3];
        NphVec3 ab[3];
        const u32 fa[3] = {NPH_P_AAN, NPH_P_AAT1, NPH_P_AAT2};
        const u32 fb[3] = {NPH_P_ABN, NPH_P_ABT1, NPH_P_ABT2};
        for (u32 x = 0U; x < 3U; ++x) {
            rxa[x] = v3_cross(ra, axes[x]);
            rxb[x] = v3_cross(rb, axes[x]);
            aa[x] = m3_mul_v3(ii_a, rxa[x]);
            ab[x] = m3_mul_v3(ii_b, rxb[x]);
            pv3_store(d, fa[x], R, aa[x]);
            pv3_store(d, fb[x], R, ab[x]);
        }
        const f32 kn = mass + v3_dot(rxa[0], aa[0]) + v3_dot(rxb[0], ab[0]);
        PF(d, NPH_P_NMASS, R) = (active && kn > 0.0F) ? 1.0F / kn : 0.0F;
        const f32 k11 = mass + v3_dot(rxa[1], aa[1]) + v3_dot(rxb[1], ab[1]);
        const f32 k22 = mass + v3_dot(rxa[2], aa[2]) + v3_dot(rxb[2], ab[2]);
        const f32 k12 = v3_dot(rxa[1], aa[2]) + v3_dot(rxb[1], ab[2]);
        const f32 det = k11 * k22 - k12 * k12;
        const f32 inv = (active && det > 0.0F) ? 1.0F / det : 0.0F;
        PF(d, NPH_P_TM0, R) = k22 * inv;
        PF(d, NPH_P_TM1, R) = -k12 * inv;
        PF(d, NPH_P_TM2, R) = k11 * inv;
        const NphVec3 dv = lane_relative_velocity(&A, &B, ra, rb);
        const f32 relv = active ? v3_dot(dv, n) : 0.0F;
        PF(d, NPH_P_RELV, R) = relv;
        PF(d, NPH_P_NIMP, R) = nimp;
        PF(d, NPH_P_NMAX, R) = 0.0F;
        PF(d, NPH_P_TI1, R) = v3_dot(timp, t1);
        PF(d, NPH_P_TI2, R) = v3_dot(timp, t2);
        if (relv < -NPH_IMPACT_SPEED) {
            impact = 1U;
        }
    }
    d->lane_flags[L] = flags | (impact ? NPH_LANE_IMPACT : 0U);
}

// After every lane of a group is prepared: does any lane have an impact.
NPH_HD void group_finish_prepare(const NphStepData *d, u32 g)
{
    u32 any = 0U;
    for (u32 l = 0U; l < NPH_LANES; ++l) {
        any |= d->lane_flags[g * NPH_LANES + l] & NPH_LANE_IMPACT;
    }
    d->groups[g].impact = any != 0U ? 1U : 0U;
}

NPH_HD NphVec3 lane_axis(const NphStepData *d, u32 f, u32 L)
{
    return v3(LF(d, f, L), LF(d, f + 1U, L), LF(d, f + 2U, L));
}

// A lane's point rows held in registers for a pass: every load issued before
// any arithmetic, every store after it (the rows are private to the lane,
// so nothing in between can change them).
typedef struct NphPointRegs {
    NphVec3 ra;
    NphVec3 rb;
    f32 sep;
    f32 nmass;
    f32 nimp;
    f32 nmax;
    f32 relv;
    f32 tm0;
    f32 tm1;
    f32 tm2;
    f32 ti1;
    f32 ti2;
    NphVec3 aan;
    NphVec3 abn;
    NphVec3 aat1;
    NphVec3 abt1;
    NphVec3 aat2;
    NphVec3 abt2;
} NphPointRegs;

// Which fields a pass reads.
enum {
    NPH_LOAD_WARM = 1U,  // impulses and angular terms
    NPH_LOAD_SOLVE = 2U, // everything but the approach speed
    NPH_LOAD_REST = 4U,  // anchors, mass, normal impulses, approach speed, normal terms
};

NPH_HD void point_load(const NphStepData *d, u64 R, u32 what, NphPointRegs *p)
{
    if ((what & (NPH_LOAD_SOLVE | NPH_LOAD_REST)) != 0U) {
        p->ra = PV3(d, NPH_P_RAX, R);
        p->rb = PV3(d, NPH_P_RBX, R);
        p->nmass = PF(d, NPH_P_NMASS, R);
        p->nmax = PF(d, NPH_P_NMAX, R);
    }
    if ((what & NPH_LOAD_SOLVE) != 0U) {
        p->sep = PF(d, NPH_P_SEP, R);
        p->tm0 = PF(d, NPH_P_TM0, R);
        p->tm1 = PF(d, NPH_P_TM1, R);
        p->tm2 = PF(d, NPH_P_TM2, R);
    }
    if ((what & NPH_LOAD_REST) != 0U) {
        p->relv = PF(d, NPH_P_RELV, R);
    }
    p->nimp = PF(d, NPH_P_NIMP, R);
    p->aan = PV3(d, NPH_P_AAN, R);
    p->abn = PV3(d, NPH_P_ABN, R);
    if ((what & (NPH_LOAD_WARM | NPH_LOAD_SOLVE)) != 0U) {
        p->ti1 = PF(d, NPH_P_TI1, R);
        p->ti2 = PF(d, NPH_P_TI2, R);
        p->aat1 = PV3(d, NPH_P_AAT1, R);
        p->abt1 = PV3(d, NPH_P_ABT1, R);
        p->aat2 = PV3(d, NPH_P_AAT2, R);
        p->abt2 = PV3(d, NPH_P_ABT2, R);
    }
}

// solver_impl.h:181-258 (the soft step's warm start: nothing scales it).
NPH_HD void lane_warm_start(const NphStepData *d, u32 g, u32 l)
{
    const NphGroup *grp = &d->groups[g];
    const u32 L = g * NPH_LANES + l;
    const u32 P = grp->point_count;
    const u32 ia = d->lane_body_a[L];
    const u32 ib = d->lane_body_b[L];
    NphPointRegs pt[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            point_load(d, (u64)(grp->first_point + k) * NPH_LANES + l, NPH_LOAD_WARM, &pt[k]);
        }
    }
    NphLaneBody a = lane_gather(&d->sbodies[ia], 0);
    NphLaneBody b = lane_gather(&d->sbodies[ib], 0);
    const f32 ma = LF(d, NPH_L_IMA, L);
    const f32 mb = LF(d, NPH_L_IMB, L);
    const NphVec3 n = lane_axis(d, NPH_L_NX, L);
    const NphVec3 t1 = lane_axis(d, NPH_L_T1X, L);
    const NphVec3 t2 = lane_axis(d, NPH_L_T2X, L);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            lane_apply(&a, &b, ma, mb, pt[k].nimp, n, pt[k].aan, pt[k].abn);
            lane_apply(&a, &b, ma, mb, pt[k].ti1, t1, pt[k].aat1, pt[k].abt1);
            lane_apply(&a, &b, ma, mb, pt[k].ti2, t2, pt[k].aat2, pt[k].abt2);
        }
    }
    const u32 flags = d->lane_flags[L];
    if ((flags & NPH_LANE_WRITE_A) != 0U) {
        lane_scatter(&d->sbodies[ia], &a);
    }
    if ((flags & NPH_LANE_WRITE_B) != 0U) {
        lane_scatter(&d->sbodies[ib], &b);
    }
}

// solver_impl.h:425-464: a point's current separation.
NPH_HD f32 lane_separation(const NphLaneBody *a, const NphLaneBody *b, NphVec3 ra, NphVec3 rb,
                           NphVec3 n, f32 adjusted)
{
    const NphVec3 pa = q_rotate(a->q, ra);
    const NphVec3 pb = q_rotate(b->q, rb);
    const NphVec3 dd = v3((b->dp.x - a->dp.x) + (pb.x - pa.x), (b->dp.y - a->dp.y) + (pb.y - pa.y),
                          (b->dp.z - a->dp.z) + (pb.z - pa.z));
    return v3_dot(dd, n) + adjusted;
}

// solver_impl.h:471-482
NPH_HD void normal_row(f32 s, f32 inv_h, f32 slop, f32 soft_bias, f32 min_bias, f32 soft_mass,
                       f32 soft_impulse, f32 *bias, f32 *mass_scale, f32 *impulse_scale)
{
    const int apart = s > 0.0F;
    const f32 overlap = nph_minf(s + slop, 0.0F);
    *bias = apart ? s * inv_h : nph_maxf(soft_bias * overlap, min_bias);
    *mass_scale = apart ? 1.0F : soft_mass;
    *impulse_scale = apart ? 0.0F : soft_impulse;
}

// solver_impl.h:400-421
NPH_HD void friction_circle(f32 vt1, f32 vt2, f32 limit, f32 *new1, f32 *new2)
{
    const f32 len = nph_sqrtf(*new1 * *new1 + *new2 * *new2);
    const int over = len > limit;
    const f32 slip = nph_sqrtf(vt1 * vt1 + vt2 * vt2);
    const int sliding = over && slip > 0.0F;
    const f32 r = limit / (over ? len : 1.0F);
    const f32 stop_share = r / (over ? len : 1.0F);
    const f32 slip_share = (1.0F - r) / (sliding ? slip : 1.0F);
    const f32 d1 = stop_share * *new1 - slip_share * vt1;
    const f32 d2 = stop_share * *new2 - slip_share * vt2;
    const f32 dlen = nph_sqrtf(d1 * d1 + d2 * d2);
    const f32 to_edge = limit / (sliding ? dlen : 1.0F);
    const f32 o1 = sliding ? d1 * to_edge : (over ? *new1 * r : *new1);
    const f32 o2 = sliding ? d2 * to_edge : (over ? *new2 * r : *new2);
    *new1 = o1;
    *new2 = o2;
}

// solver_impl.h:484-602: the solve (use_bias) and relax passes of one lane.
NPH_HD void lane_solve(const NphStepData *d, u32 g, u32 l, f32 inv_h, f32 push_max, f32 slop,
                       int use_bias)
{
    const NphGroup *grp = &d->groups[g];
    const u32 L = g * NPH_LANES + l;
    const u32 P = grp->point_count;
    const u32 ia = d->lane_body_a[L];
    const u32 ib = d->lane_body_b[L];
    NphPointRegs pt[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            point_load(d, (u64)(grp->first_point + k) * NPH_LANES + l, NPH_LOAD_SOLVE, &pt[k]);
        }
    }
    NphLaneBody a = lane_gather(&d->sbodies[ia], 1);
    NphLaneBody b = lane_gather(&d->sbodies[ib], 1);
    const f32 ma = LF(d, NPH_L_IMA, L);
    const f32 mb = LF(d, NPH_L_IMB, L);
    const NphVec3 n = lane_axis(d, NPH_L_NX, L);
    const NphVec3 t1 = lane_axis(d, NPH_L_T1X, L);
    const NphVec3 t2 = lane_axis(d, NPH_L_T2X, L);
    const f32 friction = LF(d, NPH_L_FRICTION, L);
    const u32 lane_flags = d->lane_flags[L];
    const f32 min_bias = -push_max;
    const f32 soft_bias = use_bias ? LF(d, NPH_L_BIAS, L) : 0.0F;
    const f32 soft_mass = use_bias ? LF(d, NPH_L_MASS_SCALE, L) : 1.0F;
    const f32 soft_impulse = use_bias ? LF(d, NPH_L_IMPULSE_SCALE, L) : 0.0F;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            NphPointRegs *p = &pt[k];
            f32 bias;
            f32 mscale;
            f32 iscale;
            normal_row(lane_separation(&a, &b, p->ra, p->rb, n, p->sep), inv_h, slop, soft_bias,
                       min_bias, soft_mass, soft_impulse, &bias, &mscale, &iscale);
            const f32 vn = v3_dot(lane_relative_velocity(&a, &b, p->ra, p->rb), n);
            const f32 row = p->nmass * mscale;
            const f32 accumulated = p->nimp;
            const f32 impulse = -(row * (vn + bias)) - iscale * accumulated;
            const f32 total = nph_maxf(accumulated + impulse, 0.0F);
            const f32 delta = total - accumulated;
            p->nimp = total;
            p->nmax = nph_maxf(p->nmax, total);
            lane_apply(&a, &b, ma, mb, delta, n, p->aan, p->abn);
        }
    }
    if (P > 1U && grp->impact != 0U) {
        const int impact = (lane_flags & NPH_LANE_IMPACT) != 0U;
        f32 row_bias[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
        f32 row_mass[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
        f32 row_soft[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
        for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
            if (k < P) {
                f32 mscale;
                normal_row(lane_separation(&a, &b, pt[k].ra, pt[k].rb, n, pt[k].sep), inv_h, slop,
                           soft_bias, min_bias, soft_mass, soft_impulse, &row_bias[k], &mscale,
                           &row_soft[k]);
                row_mass[k] = pt[k].nmass * mscale;
            }
        }
        for (u32 sweep = 0U; sweep < NPH_IMPACT_SWEEPS; ++sweep) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
            for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
                if (k < P) {
                    NphPointRegs *p = &pt[k];
                    const f32 vn = v3_dot(lane_relative_velocity(&a, &b, p->ra, p->rb), n);
                    const f32 accumulated = p->nimp;
                    const f32 impulse =
                        -(row_mass[k] * (vn + row_bias[k])) - row_soft[k] * accumulated;
                    const f32 total = impact ? nph_maxf(accumulated + impulse, 0.0F) : accumulated;
                    const f32 delta = total - accumulated;
                    p->nimp = total;
                    p->nmax = nph_maxf(p->nmax, total);
                    lane_apply(&a, &b, ma, mb, delta, n, p->aan, p->abn);
                }
            }
        }
    }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            NphPointRegs *p = &pt[k];
            const NphVec3 dv = lane_relative_velocity(&a, &b, p->ra, p->rb);
            const f32 vt1 = v3_dot(dv, t1);
            const f32 vt2 = v3_dot(dv, t2);
            const f32 old1 = p->ti1;
            const f32 old2 = p->ti2;
            f32 new1 = old1 - (p->tm0 * vt1 + p->tm1 * vt2);
            f32 new2 = old2 - (p->tm1 * vt1 + p->tm2 * vt2);
            friction_circle(vt1, vt2, friction * p->nimp, &new1, &new2);
            p->ti1 = new1;
            p->ti2 = new2;
            lane_apply(&a, &b, ma, mb, new1 - old1, t1, p->aat1, p->abt1);
            lane_apply(&a, &b, ma, mb, new2 - old2, t2, p->aat2, p->abt2);
        }
    }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            const u64 R = (u64)(grp->first_point + k) * NPH_LANES + l;
            PF(d, NPH_P_NIMP, R) = pt[k].nimp;
            PF(d, NPH_P_NMAX, R) = pt[k].nmax;
            PF(d, NPH_P_TI1, R) = pt[k].ti1;
            PF(d, NPH_P_TI2, R) = pt[k].ti2;
        }
    }
    const u32 flags = lane_flags;
    if ((flags & NPH_LANE_WRITE_A) != 0U) {
        lane_scatter(&d->sbodies[ia], &a);
    }
    if ((flags & NPH_LANE_WRITE_B) != 0U) {
        lane_scatter(&d->sbodies[ib], &b);
    }
}

// solver_impl.h:624-629, before the restitution passes: which of a group's
// points bounce (a point of some lane in the group must, or the group is
// skipped; per point row, some lane must).
NPH_HD void group_restitution_masks(const NphStepData *d, f32 threshold, u32 g)
{
    const NphGroup *grp = &d->groups[g];
    u32 loaded = 0U;
    for (u32 k = 0U; k < grp->point_count; ++k) {
        for (u32 l = 0U; l < NPH_LANES; ++l) {
            const u64 R = (u64)(grp->first_point + k) * NPH_LANES + l;
            if (PF(d, NPH_P_RELV, R) < -threshold && PF(d, NPH_P_NMAX, R) > 0.0F) {
                loaded |= 1U << (k * NPH_LANES + l);
            }
        }
    }
    d->groups[g].loaded = loaded;
}

// solver_impl.h:611-668
NPH_HD void lane_restitution(const NphStepData *d, u32 g, u32 l)
{
    const NphGroup *grp = &d->groups[g];
    const u32 loaded_all = grp->loaded;
    if (loaded_all == 0U) {
        return;
    }
    const u32 L = g * NPH_LANES + l;
    const u32 P = grp->point_count;
    const u32 ia = d->lane_body_a[L];
    const u32 ib = d->lane_body_b[L];
    NphPointRegs pt[NPH_PHYSICS_MANIFOLD_MAX_POINTS];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            point_load(d, (u64)(grp->first_point + k) * NPH_LANES + l, NPH_LOAD_REST, &pt[k]);
        }
    }
    NphLaneBody a = lane_gather(&d->sbodies[ia], 0);
    NphLaneBody b = lane_gather(&d->sbodies[ib], 0);
    const f32 ma = LF(d, NPH_L_IMA, L);
    const f32 mb = LF(d, NPH_L_IMB, L);
    const NphVec3 n = lane_axis(d, NPH_L_NX, L);
    const f32 e = LF(d, NPH_L_RESTITUTION, L);
    const u32 sweeps = P > 1U ? 1U + NPH_IMPACT_SWEEPS : 1U;
    for (u32 sweep = 0U; sweep < sweeps; ++sweep) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
        for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
            const u32 row_mask = (loaded_all >> (k * NPH_LANES)) & 0xFFU;
            if (k >= P || row_mask == 0U) {
                continue;
            }
            NphPointRegs *p = &pt[k];
            const int loaded = (row_mask >> l) & 1U;
            const f32 vn = v3_dot(lane_relative_velocity(&a, &b, p->ra, p->rb), n);
            const f32 accumulated = p->nimp;
            const f32 impulse = -(p->nmass * (vn + e * p->relv));
            const f32 total = nph_maxf(accumulated + impulse, 0.0F);
            const f32 delta = loaded ? total - accumulated : 0.0F;
            p->nimp = accumulated + delta;
            p->nmax = nph_maxf(p->nmax, accumulated + delta);
            lane_apply(&a, &b, ma, mb, delta, n, p->aan, p->abn);
        }
    }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (u32 k = 0U; k < NPH_PHYSICS_MANIFOLD_MAX_POINTS; ++k) {
        if (k < P) {
            const u64 R = (u64)(grp->first_point + k) * NPH_LANES + l;
            PF(d, NPH_P_NIMP, R) = pt[k].nimp;
            PF(d, NPH_P_NMAX, R) = pt[k].nmax;
        }
    }
    const u32 flags = d->lane_flags[L];
    if ((flags & NPH_LANE_WRITE_A) != 0U) {
        lane_scatter(&d->sbodies[ia], &a);
    }
    if ((flags & NPH_LANE_WRITE_B) != 0U) {
        lane_scatter(&d->sbodies[ib], &b);
    }
}

// solver.c:475-499
NPH_HD void lane_store(const NphStepData *d, u32 g, u32 l)
{
    const NphGroup *grp = &d->groups[g];
    if (l >= grp->lane_count) {
        return;
    }
    const u32 L = g * NPH_LANES + l;
    NphContactRec *c = &d->contacts[d->lane_contact[L]];
    const NphVec3 t1 = lane_axis(d, NPH_L_T1X, L);
    const NphVec3 t2 = lane_axis(d, NPH_L_T2X, L);
    for (u32 k = 0U; k < c->point_count; ++k) {
        const u64 R = (u64)(grp->first_point + k) * NPH_LANES + l;
        NphPointRec *cp = &c->points[k];
        cp->normal_impulse = PF(d, NPH_P_NIMP, R);
        cp->max_normal_impulse = PF(d, NPH_P_NMAX, R);
        cp->relative_velocity = PF(d, NPH_P_RELV, R);
        cp->tangent_impulse = v3_add(v3_scale(t1, PF(d, NPH_P_TI1, R)), v3_scale(t2, PF(d, NPH_P_TI2, R)));
    }
}

#endif // NPH_STEP_H
