// SPDX-License-Identifier: Apache-2.0
// Joints on the device: the ball, hinge and fixed joints' preparation, warm
// start and soft/rigid solve (joint_solver.c:32-394, 726-907), one thread per
// joint, in the colour passes beside the contacts.
//
// A joint's record is two parts kept by joint slot in HBM: the host-written
// configuration (type, bodies, frames, limits, motor, spring) and the
// device-written state (accumulated impulses, kept between steps for the warm
// start, and what the step's preparation fixes: solver indices, masses,
// anchors and frames at the step's start).  The rows run in the reference's
// order (motor or spring, limits, locked axes, the anchors last) with its
// arithmetic, operation for operation; joints are scalar in the reference
// too, so nothing here is lane-shaped.
#ifndef NPH_JOINT_H
#define NPH_JOINT_H

#include "nph_step.h"

enum {
    NPH_JOINT_COLLIDE_CONNECTED = 1U, // joint.c:376 (PHYS_JOINT_COLLIDE_CONNECTED)
    NPH_JOINT_WRITE_A = 1U << 8U,      // joint_solver.c:101-103
    NPH_JOINT_WRITE_B = 1U << 9U,
};

typedef union NphJointSettings {
    NphPhysicsBallJoint ball;
    NphPhysicsHingeJoint hinge;
} NphJointSettings;

struct NphJointConf {
    u32 type;  // NphPhysicsJointType: ball, hinge or fixed
    u32 flags; // NPH_JOINT_COLLIDE_CONNECTED
    u32 body_a; // body slots
    u32 body_b;
    NphPhysicsPose frame_a; // in body A's frame, rotation normalised
    NphPhysicsPose frame_b;
    NphJointSettings s;
};

struct NphJointState {
    // Accumulated impulses (per substep), kept between steps.
    NphVec3 linear_impulse;
    NphVec3 angular_impulse;
    NphVec3 friction_impulse;
    f32 lower_impulse;
    f32 upper_impulse;
    f32 swing_impulse;
    f32 motor_impulse;
    f32 spring_impulse;
    // Fixed by the step's preparation.
    u32 solver_a;
    u32 solver_b;
    u32 write; // NPH_JOINT_WRITE_A | NPH_JOINT_WRITE_B
    f32 inv_mass_a;
    f32 inv_mass_b;
    NphMat3 inv_inertia_a;
    NphMat3 inv_inertia_b;
    NphVec3 anchor_a;     // from A's centre of mass, world axes
    NphVec3 anchor_b;
    NphVec3 delta_center; // B's centre less A's
    NphQuat frame_q_a;    // the frames in the world
    NphQuat frame_q_b;
    f32 spring_bias;
    f32 spring_mass;
    f32 spring_impulse_scale;
};

// ---------------------------------------------------------------------------
// Small linear algebra (joint_solver.c:32-96)
// ---------------------------------------------------------------------------

// K x = b for a symmetric K by columns, by Cramer's rule; zero unless the
// determinant is positive.
NPH_HD NphVec3 jnt_solve3(const NphVec3 k[3], NphVec3 b)
{
    const NphVec3 c12 = v3_cross(k[1], k[2]);
    const f32 det = v3_dot(k[0], c12);
    if (!(det > 0.0F)) {
        return v3_zero();
    }
    const f32 inv = 1.0F / det;
    return v3(v3_dot(b, c12) * inv, v3_dot(k[0], v3_cross(b, k[2])) * inv,
              v3_dot(k[0], v3_cross(k[1], b)) * inv);
}

// (mA + mB) I - [rA]x IA^-1 [rA]x - [rB]x IB^-1 [rB]x, column by column.
NPH_HD void jnt_point_mass(const NphJointState *j, NphVec3 ra, NphVec3 rb, NphVec3 k[3])
{
    const f32 m = j->inv_mass_a + j->inv_mass_b;
    for (u32 c = 0U; c < 3U; ++c) {
        const NphVec3 e = v3(c == 0U ? 1.0F : 0.0F, c == 1U ? 1.0F : 0.0F, c == 2U ? 1.0F : 0.0F);
        const NphVec3 ta = m3_mul_v3(j->inv_inertia_a, v3_cross(ra, e));
        const NphVec3 tb = m3_mul_v3(j->inv_inertia_b, v3_cross(rb, e));
        k[c] = v3_sub(v3_sub(v3_scale(e, m), v3_cross(ra, ta)), v3_cross(rb, tb));
    }
}

NPH_HD void jnt_angular_mass(const NphJointState *j, NphVec3 k[3])
{
    for (u32 c = 0U; c < 3U; ++c) {
        k[c] = v3_add(j->inv_inertia_a.c[c], j->inv_inertia_b.c[c]);
    }
}

NPH_HD f32 jnt_axial_mass(const NphJointState *j, NphVec3 axis)
{
    const f32 k = v3_dot(axis, m3_mul_v3(j->inv_inertia_a, axis)) +
                  v3_dot(axis, m3_mul_v3(j->inv_inertia_b, axis));
    return k > 0.0F ? 1.0F / k : 0.0F;
}

NPH_HD NphQuat jnt_qneg(NphQuat q) { return q4(-q.x, -q.y, -q.z, -q.w); }

// B's twist from A's frame about A's X axis, in [-pi, pi].
NPH_HD f32 jnt_twist_angle(NphQuat qa, NphQuat qb)
{
    NphQuat r = q_mul(q_conj(qa), qb);
    if (r.w < 0.0F) {
        r = jnt_qneg(r);
    }
    return 2.0F * nph_atan2f(r.x, r.w);
}

NPH_HD f32 jnt_unwind(f32 a)
{
    const f32 pi = 3.14159265F;
    if (a > pi) {
        return a - 2.0F * pi;
    }
    if (a < -pi) {
        return a + 2.0F * pi;
    }
    return a;
}

// ---------------------------------------------------------------------------
// Bodies (joint_solver.c:114-165)
// ---------------------------------------------------------------------------
typedef struct NphJointNow {
    NphSolverBody *a;
    NphSolverBody *b;
    NphVec3 ra; // anchors from the centres of mass, now
    NphVec3 rb;
    NphQuat qa; // frames, now
    NphQuat qb;
} NphJointNow;

NPH_HD NphJointNow jnt_now(NphSolverBody *bodies, const NphJointState *j)
{
    NphJointNow s;
    s.a = &bodies[j->solver_a];
    s.b = &bodies[j->solver_b];
    s.ra = q_rotate(sb_dq(s.a), j->anchor_a);
    s.rb = q_rotate(sb_dq(s.b), j->anchor_b);
    s.qa = q_mul(sb_dq(s.a), j->frame_q_a);
    s.qb = q_mul(sb_dq(s.b), j->frame_q_b);
    return s;
}

// A linear impulse p at the anchors: B gains it, A loses it.
NPH_HD void jnt_apply_linear(const NphJointState *j, NphJointNow *s, NphVec3 p)
{
    if ((j->write & NPH_JOINT_WRITE_A) != 0U) {
        sb_set_v(s->a, v3_mul_add(sb_v(s->a), p, -j->inv_mass_a));
        sb_set_w(s->a, v3_sub(sb_w(s->a), m3_mul_v3(j->inv_inertia_a, v3_cross(s->ra, p))));
    }
    if ((j->write & NPH_JOINT_WRITE_B) != 0U) {
        sb_set_v(s->b, v3_mul_add(sb_v(s->b), p, j->inv_mass_b));
        sb_set_w(s->b, v3_add(sb_w(s->b), m3_mul_v3(j->inv_inertia_b, v3_cross(s->rb, p))));
    }
}

NPH_HD void jnt_apply_angular(const NphJointState *j, NphJointNow *s, NphVec3 l)
{
    if ((j->write & NPH_JOINT_WRITE_A) != 0U) {
        sb_set_w(s->a, v3_sub(sb_w(s->a), m3_mul_v3(j->inv_inertia_a, l)));
    }
    if ((j->write & NPH_JOINT_WRITE_B) != 0U) {
        sb_set_w(s->b, v3_add(sb_w(s->b), m3_mul_v3(j->inv_inertia_b, l)));
    }
}

NPH_HD NphVec3 jnt_relative_angular(const NphJointNow *s) { return v3_sub(sb_w(s->b), sb_w(s->a)); }

// ---------------------------------------------------------------------------
// Rows (joint_solver.c:170-394)
// ---------------------------------------------------------------------------
typedef struct NphJointSoft {
    f32 bias;
    f32 mass;
    f32 impulse;
} NphJointSoft;

typedef struct NphJointStep {
    f32 h;
    f32 inv_h;
    NphJointSoft soft; // the world's joint stiffness at the substep
} NphJointStep;

// Speculative while the gap c is open, soft while solving, rigid relaxing.
NPH_HD NphJointSoft jnt_limit_soft(f32 c, const NphJointStep *step, int use_bias)
{
    NphJointSoft r = {0.0F, 1.0F, 0.0F};
    if (c > 0.0F) {
        r.bias = c * step->inv_h;
    } else if (use_bias) {
        r.bias = step->soft.bias * c;
        r.mass = step->soft.mass;
        r.impulse = step->soft.impulse;
    }
    return r;
}

// One-sided angular row: the gap c closes at dot(wA - wB, axis).
NPH_HD void jnt_angular_limit(const NphJointState *j, NphJointNow *s, NphVec3 axis, f32 c,
                              f32 *accumulated, const NphJointStep *step, int use_bias)
{
    const NphJointSoft r = jnt_limit_soft(c, step, use_bias);
    const f32 cdot = -v3_dot(jnt_relative_angular(s), axis);
    const f32 impulse =
        -jnt_axial_mass(j, axis) * r.mass * (cdot + r.bias) - r.impulse * *accumulated;
    const f32 total = nph_maxf(*accumulated + impulse, 0.0F);
    const f32 delta = total - *accumulated;
    *accumulated = total;
    jnt_apply_angular(j, s, v3_scale(axis, -delta));
}

// The anchors meet: three rows at once.
NPH_HD void jnt_point(NphJointState *j, NphJointNow *s, const NphJointStep *step, int use_bias)
{
    const NphVec3 va = v3_add(sb_v(s->a), v3_cross(sb_w(s->a), s->ra));
    const NphVec3 vb = v3_add(sb_v(s->b), v3_cross(sb_w(s->b), s->rb));
    const NphVec3 cdot = v3_sub(vb, va);
    NphVec3 bias = v3_zero();
    f32 mass_scale = 1.0F;
    f32 impulse_scale = 0.0F;
    if (use_bias) {
        const NphVec3 c = v3_add(v3_add(v3_sub(sb_dp(s->b), sb_dp(s->a)), v3_sub(s->rb, s->ra)),
                                 j->delta_center);
        bias = v3_scale(c, step->soft.bias);
        mass_scale = step->soft.mass;
        impulse_scale = step->soft.impulse;
    }
    NphVec3 k[3];
    jnt_point_mass(j, s->ra, s->rb, k);
    const NphVec3 x = jnt_solve3(k, v3_add(cdot, bias));
    const NphVec3 impulse =
        v3_sub(v3_scale(x, -mass_scale), v3_scale(j->linear_impulse, impulse_scale));
    j->linear_impulse = v3_add(j->linear_impulse, impulse);
    jnt_apply_linear(j, s, impulse);
}

// The frames align (fixed): the rotation vector from A's frame to B's.
NPH_HD void jnt_rotation_lock(NphJointState *j, NphJointNow *s, const NphJointStep *step,
                              int use_bias)
{
    const NphVec3 cdot = jnt_relative_angular(s);
    NphVec3 bias = v3_zero();
    f32 mass_scale = 1.0F;
    f32 impulse_scale = 0.0F;
    if (use_bias) {
        NphQuat e = q_mul(s->qb, q_conj(s->qa));
        if (e.w < 0.0F) {
            e = jnt_qneg(e);
        }
        bias = v3_scale(v3(e.x, e.y, e.z), 2.0F * step->soft.bias);
        mass_scale = step->soft.mass;
        impulse_scale = step->soft.impulse;
    }
    NphVec3 k[3];
    jnt_angular_mass(j, k);
    const NphVec3 x = jnt_solve3(k, v3_add(cdot, bias));
    const NphVec3 impulse =
        v3_sub(v3_scale(x, -mass_scale), v3_scale(j->angular_impulse, impulse_scale));
    j->angular_impulse = v3_add(j->angular_impulse, impulse);
    jnt_apply_angular(j, s, impulse);
}

// The X axes align (hinge): two rows about A's Y and Z axes, their impulses
// kept as scalars in angular_impulse.x and .y.
NPH_HD void jnt_axis_lock(NphJointState *j, NphJointNow *s, const NphJointStep *step,
                          int use_bias)
{
    const NphVec3 p1 = q_axis_y(s->qa);
    const NphVec3 p2 = q_axis_z(s->qa);
    const NphVec3 dw = jnt_relative_angular(s);
    f32 c1 = 0.0F;
    f32 c2 = 0.0F;
    f32 mass_scale = 1.0F;
    f32 impulse_scale = 0.0F;
    if (use_bias) {
        const NphVec3 e = v3_cross(q_axis_x(s->qa), q_axis_x(s->qb));
        c1 = step->soft.bias * v3_dot(e, p1);
        c2 = step->soft.bias * v3_dot(e, p2);
        mass_scale = step->soft.mass;
        impulse_scale = step->soft.impulse;
    }
    const NphVec3 ip1 = v3_add(m3_mul_v3(j->inv_inertia_a, p1), m3_mul_v3(j->inv_inertia_b, p1));
    const NphVec3 ip2 = v3_add(m3_mul_v3(j->inv_inertia_a, p2), m3_mul_v3(j->inv_inertia_b, p2));
    const f32 k11 = v3_dot(p1, ip1);
    const f32 k12 = v3_dot(p1, ip2);
    const f32 k22 = v3_dot(p2, ip2);
    const f32 det = k11 * k22 - k12 * k12;
    if (!(det > 0.0F)) {
        return;
    }
    const f32 inv = 1.0F / det;
    const f32 b1 = v3_dot(dw, p1) + c1;
    const f32 b2 = v3_dot(dw, p2) + c2;
    const f32 x1 = (k22 * b1 - k12 * b2) * inv;
    const f32 x2 = (k11 * b2 - k12 * b1) * inv;
    const f32 i1 = -mass_scale * x1 - impulse_scale * j->angular_impulse.x;
    const f32 i2 = -mass_scale * x2 - impulse_scale * j->angular_impulse.y;
    j->angular_impulse.x += i1;
    j->angular_impulse.y += i2;
    jnt_apply_angular(j, s, v3_add(v3_scale(p1, i1), v3_scale(p2, i2)));
}

NPH_HD void jnt_hinge(const NphPhysicsHingeJoint *h, NphJointState *j, NphJointNow *s,
                      const NphJointStep *step, int use_bias)
{
    const NphVec3 axis = q_axis_x(s->qa);
    const f32 angle = jnt_twist_angle(s->qa, s->qb);
    const f32 mass = jnt_axial_mass(j, axis);
    // The spring once a substep, in the solve.
    if (h->enable_spring != 0 && use_bias) {
        const f32 c = jnt_unwind(angle - h->spring_target);
        const f32 cdot = v3_dot(jnt_relative_angular(s), axis);
        const f32 impulse = -j->spring_mass * mass * (cdot + j->spring_bias * c) -
                            j->spring_impulse_scale * j->spring_impulse;
        j->spring_impulse += impulse;
        jnt_apply_angular(j, s, v3_scale(axis, impulse));
    }
    if (h->enable_motor != 0) {
        const f32 cdot = v3_dot(jnt_relative_angular(s), axis) - h->motor_speed;
        const f32 max_impulse = step->h * h->max_motor_torque;
        const f32 old = j->motor_impulse;
        const f32 want = old - mass * cdot;
        j->motor_impulse = want < -max_impulse ? -max_impulse : (want > max_impulse ? max_impulse : want);
        jnt_apply_angular(j, s, v3_scale(axis, j->motor_impulse - old));
    }
    if (h->enable_limit != 0) {
        jnt_angular_limit(j, s, v3_neg(axis), angle - h->lower, &j->lower_impulse, step, use_bias);
        jnt_angular_limit(j, s, axis, h->upper - angle, &j->upper_impulse, step, use_bias);
    }
    jnt_axis_lock(j, s, step, use_bias);
}

// The ball's cone: the X axes' angle within the ellipse of swing_y, swing_z.
NPH_HD void jnt_swing(const NphPhysicsBallJoint *b, NphJointState *j, NphJointNow *s,
                      const NphJointStep *step, int use_bias)
{
    const NphVec3 xa = q_axis_x(s->qa);
    const NphVec3 xb = q_axis_x(s->qb);
    const NphVec3 n = v3_cross(xa, xb);
    const f32 sin_phi = v3_len(n);
    if (!(sin_phi > 1e-6F)) {
        return;
    }
    const NphVec3 axis = v3_scale(n, 1.0F / sin_phi);
    const NphVec3 local = q_rotate_inv(s->qa, axis);
    const f32 ey = local.y / b->swing_y;
    const f32 ez = local.z / b->swing_z;
    const f32 limit = 1.0F / nph_sqrtf(ey * ey + ez * ez);
    const f32 phi = nph_atan2f(sin_phi, v3_dot(xa, xb));
    jnt_angular_limit(j, s, axis, limit - phi, &j->swing_impulse, step, use_bias);
}

NPH_HD void jnt_ball(const NphPhysicsBallJoint *b, NphJointState *j, NphJointNow *s,
                     const NphJointStep *step, int use_bias)
{
    if (b->friction_torque > 0.0F) {
        NphVec3 k[3];
        jnt_angular_mass(j, k);
        const NphVec3 x = jnt_solve3(k, jnt_relative_angular(s));
        const NphVec3 old = j->friction_impulse;
        NphVec3 total = v3_sub(old, x);
        const f32 max_impulse = step->h * b->friction_torque;
        const f32 len_sq = v3_len_sq(total);
        if (len_sq > max_impulse * max_impulse) {
            total = v3_scale(total, max_impulse / nph_sqrtf(len_sq));
        }
        j->friction_impulse = total;
        jnt_apply_angular(j, s, v3_sub(total, old));
    }
    if (b->enable_swing_limit != 0) {
        jnt_swing(b, j, s, step, use_bias);
    }
    if (b->enable_twist_limit != 0) {
        const NphVec3 sum = v3_add(q_axis_x(s->qa), q_axis_x(s->qb));
        const f32 len = v3_len(sum);
        if (len > 1e-3F) {
            const NphVec3 axis = v3_scale(sum, 1.0F / len);
            const f32 twist = jnt_twist_angle(s->qa, s->qb);
            jnt_angular_limit(j, s, v3_neg(axis), twist - b->twist_lower, &j->lower_impulse, step,
                              use_bias);
            jnt_angular_limit(j, s, axis, b->twist_upper - twist, &j->upper_impulse, step,
                              use_bias);
        }
    }
}

// ---------------------------------------------------------------------------
// Step entry points (joint_solver.c:726-907)
// ---------------------------------------------------------------------------

// Before the passes: masses, anchors and frames at the step's start; what is
// switched off starts from nothing.
NPH_HD void joint_prepare(const NphJointConf *cf, NphJointState *j, const NphBodyRec *a,
                          const NphBodyRec *b, u32 solver_a, u32 solver_b,
                          const NphSolverParams *params, f32 h)
{
    const int dyn_a = a->type == (u32)NPH_PHYSICS_BODY_DYNAMIC;
    const int dyn_b = b->type == (u32)NPH_PHYSICS_BODY_DYNAMIC;
    j->solver_a = solver_a;
    j->solver_b = solver_b;
    j->write = (dyn_a ? NPH_JOINT_WRITE_A : 0U) | (dyn_b ? NPH_JOINT_WRITE_B : 0U);
    j->inv_mass_a = dyn_a ? a->inv_mass : 0.0F;
    j->inv_mass_b = dyn_b ? b->inv_mass : 0.0F;
    j->inv_inertia_a = dyn_a ? params[solver_a].inv_inertia : m3_zero();
    j->inv_inertia_b = dyn_b ? params[solver_b].inv_inertia : m3_zero();
    j->anchor_a = q_rotate(a->rotation, v3_sub(cf->frame_a.position, a->local_center));
    j->anchor_b = q_rotate(b->rotation, v3_sub(cf->frame_b.position, b->local_center));
    j->delta_center = dv3_relative(b->center, a->center);
    j->frame_q_a = q_mul(a->rotation, cf->frame_a.rotation);
    j->frame_q_b = q_mul(b->rotation, cf->frame_b.rotation);
    j->spring_bias = 0.0F;
    j->spring_mass = 1.0F;
    j->spring_impulse_scale = 0.0F;
    if (cf->type == (u32)NPH_PHYSICS_JOINT_HINGE) {
        const NphPhysicsHingeJoint *hj = &cf->s.hinge;
        if (hj->enable_limit == 0) {
            j->lower_impulse = 0.0F;
            j->upper_impulse = 0.0F;
        }
        if (hj->enable_motor == 0) {
            j->motor_impulse = 0.0F;
        }
        if (hj->enable_spring == 0) {
            j->spring_impulse = 0.0F;
        }
        make_soft(hj->enable_spring != 0 ? hj->spring_hertz : 0.0F, hj->spring_damping_ratio, h,
                  &j->spring_bias, &j->spring_mass, &j->spring_impulse_scale);
    } else if (cf->type == (u32)NPH_PHYSICS_JOINT_BALL) {
        const NphPhysicsBallJoint *bj = &cf->s.ball;
        if (bj->enable_swing_limit == 0) {
            j->swing_impulse = 0.0F;
        }
        if (bj->enable_twist_limit == 0) {
            j->lower_impulse = 0.0F;
            j->upper_impulse = 0.0F;
        }
        if (!(bj->friction_torque > 0.0F)) {
            j->friction_impulse = v3_zero();
        }
    }
}

// The accumulated angular impulse as a world vector (joint_solver.c:826-853).
NPH_HD NphVec3 joint_angular_impulse(const NphJointConf *cf, const NphJointState *j, NphQuat qa,
                                     NphQuat qb)
{
    NphVec3 angular = j->angular_impulse;
    if (cf->type == (u32)NPH_PHYSICS_JOINT_HINGE) {
        angular = v3_add(v3_scale(q_axis_y(qa), j->angular_impulse.x),
                         v3_scale(q_axis_z(qa), j->angular_impulse.y));
        const f32 axial = j->lower_impulse - j->upper_impulse + j->motor_impulse + j->spring_impulse;
        angular = v3_mul_add(angular, q_axis_x(qa), axial);
    } else if (cf->type == (u32)NPH_PHYSICS_JOINT_BALL) {
        angular = j->friction_impulse;
        const NphVec3 xa = q_axis_x(qa);
        const NphVec3 xb = q_axis_x(qb);
        const NphVec3 n = v3_cross(xa, xb);
        const f32 sin_phi = v3_len(n);
        if (j->swing_impulse != 0.0F && sin_phi > 1e-6F) {
            angular = v3_mul_add(angular, n, -j->swing_impulse / sin_phi);
        }
        const NphVec3 sum = v3_add(xa, xb);
        const f32 len = v3_len(sum);
        if (len > 1e-3F) {
            angular = v3_mul_add(angular, sum, (j->lower_impulse - j->upper_impulse) / len);
        }
    }
    return angular;
}

NPH_HD void joint_warm_start(NphSolverBody *bodies, const NphJointConf *cf, NphJointState *j)
{
    NphJointNow s = jnt_now(bodies, j);
    jnt_apply_angular(j, &s, joint_angular_impulse(cf, j, s.qa, s.qb));
    jnt_apply_linear(j, &s, j->linear_impulse);
}

NPH_HD void joint_solve(NphSolverBody *bodies, const NphJointStep *step, const NphJointConf *cf,
                        NphJointState *j, int use_bias)
{
    NphJointNow s = jnt_now(bodies, j);
    if (cf->type == (u32)NPH_PHYSICS_JOINT_BALL) {
        jnt_ball(&cf->s.ball, j, &s, step, use_bias);
    } else if (cf->type == (u32)NPH_PHYSICS_JOINT_HINGE) {
        jnt_hinge(&cf->s.hinge, j, &s, step, use_bias);
    } else {
        jnt_rotation_lock(j, &s, step, use_bias);
    }
    jnt_point(j, &s, step, use_bias);
}

#endif // NPH_JOINT_H
