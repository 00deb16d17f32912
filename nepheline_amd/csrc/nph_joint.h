// SPDX-License-Identifier: Apache-2.0
// Joints on the device: ball, hinge and fixed joints as ordered lists of
// constraint rows, solved one joint a thread in the colour passes beside the
// contacts, with the reference's arithmetic (joint_solver.c:32-394,726-907).
//
// A joint's record is two parts kept by joint slot in HBM: the host-written
// configuration (type, bodies, frames, limits, motor, spring) and the
// device-written state (accumulated impulses, kept between steps for the
// warm start, and what the step's preparation fixes: solver indices,
// masses, anchors and frames at the step's start).
//
// A solve first measures the joint's geometry from the bodies' substep
// rotations (the hinge's axis and angle, the ball's swing and twist axes
// and gaps; positions do not change within a pass), then runs the rows of
// the joint's type in order. Each row applies its impulse to the bodies'
// velocities at once, so a later row sees the earlier rows' effect:
//   hinge  spring, motor, lower limit, upper limit, the two locked axes,
//          the anchors
//   ball   friction, swing cone, lower twist, upper twist, the anchors
//   fixed  the orientation, the anchors
// Rows are soft while solving (the world's joint softness), rigid while
// relaxing, and limits are speculative while their gap is open.
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

// A row's soft-constraint coefficients: velocity bias per unit of error,
// mass scale and accumulated-impulse scale.
typedef struct NphJointSoft {
    f32 bias;
    f32 mass;
    f32 impulse;
} NphJointSoft;

typedef struct NphJointStep {
    f32 h;
    f32 inv_h;
    NphJointSoft soft; // the world's joint softness at the substep
} NphJointStep;

// ---------------------------------------------------------------------------
// Effective masses
// ---------------------------------------------------------------------------

// x with K x = rhs for a symmetric positive K given by columns (Cramer's
// rule); zero when the determinant is not positive.
NPH_HD NphVec3 joint_solve_sym3(const NphVec3 k[3], NphVec3 rhs)
{
    const NphVec3 c12 = v3_cross(k[1], k[2]);
    const f32 det = v3_dot(k[0], c12);
    if (!(det > 0.0F)) {
        return v3_zero();
    }
    const f32 inv = 1.0F / det;
    return v3(v3_dot(rhs, c12) * inv, v3_dot(k[0], v3_cross(rhs, k[2])) * inv,
              v3_dot(k[0], v3_cross(k[1], rhs)) * inv);
}

// The anchors' inverse mass matrix, (mA + mB) 1 - [rA] IA^-1 [rA] - [rB] IB^-1
// [rB] (cross-product matrices), by columns.
NPH_HD void joint_anchor_matrix(const NphJointState *j, NphVec3 ra, NphVec3 rb, NphVec3 k[3])
{
    const f32 m = j->inv_mass_a + j->inv_mass_b;
    for (u32 col = 0U; col < 3U; ++col) {
        NphVec3 e = v3_zero();
        if (col == 0U) {
            e.x = 1.0F;
        } else if (col == 1U) {
            e.y = 1.0F;
        } else {
            e.z = 1.0F;
        }
        const NphVec3 turn_a = m3_mul_v3(j->inv_inertia_a, v3_cross(ra, e));
        const NphVec3 turn_b = m3_mul_v3(j->inv_inertia_b, v3_cross(rb, e));
        k[col] = v3_sub(v3_sub(v3_scale(e, m), v3_cross(ra, turn_a)), v3_cross(rb, turn_b));
    }
}

// The rotations' inverse mass matrix IA^-1 + IB^-1, by columns.
NPH_HD void joint_rotation_matrix(const NphJointState *j, NphVec3 k[3])
{
    for (u32 col = 0U; col < 3U; ++col) {
        k[col] = v3_add(j->inv_inertia_a.c[col], j->inv_inertia_b.c[col]);
    }
}

// The effective mass of a rotation about a unit axis (0 when immovable).
NPH_HD f32 joint_axial_mass(const NphJointState *j, NphVec3 axis)
{
    const f32 inv = v3_dot(axis, m3_mul_v3(j->inv_inertia_a, axis)) +
                    v3_dot(axis, m3_mul_v3(j->inv_inertia_b, axis));
    return inv > 0.0F ? 1.0F / inv : 0.0F;
}

NPH_HD NphQuat joint_flip(NphQuat q) { return q4(-q.x, -q.y, -q.z, -q.w); }

// The angle of B's frame about A's frame's X axis, in [-pi, pi].
NPH_HD f32 joint_twist_angle(NphQuat qa, NphQuat qb)
{
    NphQuat rel = q_mul(q_conj(qa), qb);
    if (rel.w < 0.0F) {
        rel = joint_flip(rel);
    }
    return 2.0F * nph_atan2f(rel.x, rel.w);
}

// An angle difference brought back into [-pi, pi] (one turn at most).
NPH_HD f32 joint_wrap(f32 a)
{
    const f32 pi = 3.14159265F;
    return a > pi ? a - 2.0F * pi : (a < -pi ? a + 2.0F * pi : a);
}

// ---------------------------------------------------------------------------
// The bodies as a pass sees them
// ---------------------------------------------------------------------------
typedef struct NphJointPose {
    NphSolverBody *a;
    NphSolverBody *b;
    NphVec3 ra; // anchors from the centres of mass, rotated by the substep
    NphVec3 rb;
    NphQuat qa; // the frames, rotated by the substep
    NphQuat qb;
} NphJointPose;

NPH_HD NphJointPose joint_pose(NphSolverBody *bodies, const NphJointState *j)
{
    NphJointPose p;
    p.a = &bodies[j->solver_a];
    p.b = &bodies[j->solver_b];
    p.ra = q_rotate(sb_dq(p.a), j->anchor_a);
    p.rb = q_rotate(sb_dq(p.b), j->anchor_b);
    p.qa = q_mul(sb_dq(p.a), j->frame_q_a);
    p.qb = q_mul(sb_dq(p.b), j->frame_q_b);
    return p;
}

// An impulse p at the anchors: B gains it, A loses it (only moving bodies
// are written).
NPH_HD void joint_push(const NphJointState *j, NphJointPose *p, NphVec3 impulse)
{
    if ((j->write & NPH_JOINT_WRITE_A) != 0U) {
        sb_set_v(p->a, v3_mul_add(sb_v(p->a), impulse, -j->inv_mass_a));
        sb_set_w(p->a, v3_sub(sb_w(p->a), m3_mul_v3(j->inv_inertia_a, v3_cross(p->ra, impulse))));
    }
    if ((j->write & NPH_JOINT_WRITE_B) != 0U) {
        sb_set_v(p->b, v3_mul_add(sb_v(p->b), impulse, j->inv_mass_b));
        sb_set_w(p->b, v3_add(sb_w(p->b), m3_mul_v3(j->inv_inertia_b, v3_cross(p->rb, impulse))));
    }
}

// An angular impulse: B gains it, A loses it.
NPH_HD void joint_twist(const NphJointState *j, NphJointPose *p, NphVec3 impulse)
{
    if ((j->write & NPH_JOINT_WRITE_A) != 0U) {
        sb_set_w(p->a, v3_sub(sb_w(p->a), m3_mul_v3(j->inv_inertia_a, impulse)));
    }
    if ((j->write & NPH_JOINT_WRITE_B) != 0U) {
        sb_set_w(p->b, v3_add(sb_w(p->b), m3_mul_v3(j->inv_inertia_b, impulse)));
    }
}

NPH_HD NphVec3 joint_spin(const NphJointPose *p) { return v3_sub(sb_w(p->b), sb_w(p->a)); }

// ---------------------------------------------------------------------------
// The joint's geometry at a pass (positions only)
// ---------------------------------------------------------------------------
typedef struct NphJointShape {
    // Hinge: the axis (A's X), the angle about it and its effective mass.
    NphVec3 hinge_axis;
    f32 hinge_angle;
    f32 hinge_mass;
    // Ball: the swing row's axis and its gap to the cone (swing_on when the
    // X axes are apart), the twist rows' axis and angle (twist_on when the
    // X axes are not opposite).
    int swing_on;
    NphVec3 swing_axis;
    f32 swing_gap;
    int twist_on;
    NphVec3 twist_axis;
    f32 twist_angle;
} NphJointShape;

NPH_HD NphJointShape joint_shape(const NphJointConf *cf, const NphJointState *j,
                                 const NphJointPose *p)
{
    NphJointShape g;
    g.swing_on = 0;
    g.twist_on = 0;
    g.hinge_angle = g.hinge_mass = g.swing_gap = g.twist_angle = 0.0F;
    g.hinge_axis = g.swing_axis = g.twist_axis = v3_zero();
    if (cf->type == (u32)NPH_PHYSICS_JOINT_HINGE) {
        g.hinge_axis = q_axis_x(p->qa);
        g.hinge_angle = joint_twist_angle(p->qa, p->qb);
        g.hinge_mass = joint_axial_mass(j, g.hinge_axis);
    } else if (cf->type == (u32)NPH_PHYSICS_JOINT_BALL) {
        const NphPhysicsBallJoint *ball = &cf->s.ball;
        const NphVec3 xa = q_axis_x(p->qa);
        const NphVec3 xb = q_axis_x(p->qb);
        if (ball->enable_swing_limit != 0) {
            // The cone is the ellipse of half-angles swing_y, swing_z about
            // A's X; the swing turns about xa x xb.
            const NphVec3 across = v3_cross(xa, xb);
            const f32 s = v3_len(across);
            if (s > 1e-6F) {
                g.swing_on = 1;
                g.swing_axis = v3_scale(across, 1.0F / s);
                const NphVec3 in_a = q_rotate_inv(p->qa, g.swing_axis);
                const f32 uy = in_a.y / ball->swing_y;
                const f32 uz = in_a.z / ball->swing_z;
                const f32 cone = 1.0F / nph_sqrtf(uy * uy + uz * uz);
                g.swing_gap = cone - nph_atan2f(s, v3_dot(xa, xb));
            }
        }
        if (ball->enable_twist_limit != 0) {
            const NphVec3 mid = v3_add(xa, xb);
            const f32 len = v3_len(mid);
            if (len > 1e-3F) {
                g.twist_on = 1;
                g.twist_axis = v3_scale(mid, 1.0F / len);
                g.twist_angle = joint_twist_angle(p->qa, p->qb);
            }
        }
    }
    return g;
}

// ---------------------------------------------------------------------------
// Rows
// ---------------------------------------------------------------------------

// A one-sided angular row: the gap closes at -(wB - wA).axis; its
// accumulated impulse stays non-negative. Speculative while the gap is
// open (it may close by the end of the substep), soft or rigid when shut.
NPH_HD void joint_limit_row(const NphJointState *j, NphJointPose *p, NphVec3 axis, f32 gap,
                            f32 *sum, const NphJointStep *step, int soft)
{
    NphJointSoft k = {0.0F, 1.0F, 0.0F};
    if (gap > 0.0F) {
        k.bias = gap * step->inv_h;
    } else if (soft) {
        k.bias = step->soft.bias * gap;
        k.mass = step->soft.mass;
        k.impulse = step->soft.impulse;
    }
    const f32 closing = -v3_dot(joint_spin(p), axis);
    const f32 delta = -joint_axial_mass(j, axis) * k.mass * (closing + k.bias) - k.impulse * *sum;
    const f32 now = nph_maxf(*sum + delta, 0.0F);
    const f32 applied = now - *sum;
    *sum = now;
    joint_twist(j, p, v3_scale(axis, -applied));
}

// The hinge's spring towards its target angle (once a substep: solve only).
NPH_HD void joint_spring_row(const NphPhysicsHingeJoint *hinge, NphJointState *j, NphJointPose *p,
                             const NphJointShape *g)
{
    const f32 error = joint_wrap(g->hinge_angle - hinge->spring_target);
    const f32 rate = v3_dot(joint_spin(p), g->hinge_axis);
    const f32 delta = -j->spring_mass * g->hinge_mass * (rate + j->spring_bias * error) -
                      j->spring_impulse_scale * j->spring_impulse;
    j->spring_impulse += delta;
    joint_twist(j, p, v3_scale(g->hinge_axis, delta));
}

// The hinge's motor: the relative speed about the axis driven towards the
// motor speed, within the torque limit.
NPH_HD void joint_motor_row(const NphPhysicsHingeJoint *hinge, NphJointState *j, NphJointPose *p,
                            const NphJointShape *g, const NphJointStep *step)
{
    const f32 error = v3_dot(joint_spin(p), g->hinge_axis) - hinge->motor_speed;
    const f32 cap = step->h * hinge->max_motor_torque;
    const f32 before = j->motor_impulse;
    const f32 asked = before - g->hinge_mass * error;
    j->motor_impulse = nph_clampf(asked, -cap, cap);
    joint_twist(j, p, v3_scale(g->hinge_axis, j->motor_impulse - before));
}

// Three-row blocks solve K x = error + bias at once; their impulse is
// -mass_scale x - impulse_scale (accumulated).
NPH_HD NphVec3 joint_block_impulse(const NphVec3 k[3], NphVec3 error, NphVec3 bias, NphVec3 sum,
                                   f32 mass_scale, f32 impulse_scale)
{
    const NphVec3 x = joint_solve_sym3(k, v3_add(error, bias));
    return v3_sub(v3_scale(x, -mass_scale), v3_scale(sum, impulse_scale));
}

// The anchors meet (every joint, last).
NPH_HD void joint_anchor_rows(NphJointState *j, NphJointPose *p, const NphJointStep *step, int soft)
{
    const NphVec3 at_a = v3_add(sb_v(p->a), v3_cross(sb_w(p->a), p->ra));
    const NphVec3 at_b = v3_add(sb_v(p->b), v3_cross(sb_w(p->b), p->rb));
    NphVec3 bias = v3_zero();
    f32 mass_scale = 1.0F, impulse_scale = 0.0F;
    if (soft) {
        const NphVec3 apart =
            v3_add(v3_add(v3_sub(sb_dp(p->b), sb_dp(p->a)), v3_sub(p->rb, p->ra)), j->delta_center);
        bias = v3_scale(apart, step->soft.bias);
        mass_scale = step->soft.mass;
        impulse_scale = step->soft.impulse;
    }
    NphVec3 k[3];
    joint_anchor_matrix(j, p->ra, p->rb, k);
    const NphVec3 delta = joint_block_impulse(k, v3_sub(at_b, at_a), bias, j->linear_impulse,
                                              mass_scale, impulse_scale);
    j->linear_impulse = v3_add(j->linear_impulse, delta);
    joint_push(j, p, delta);
}

// The frames align (fixed joints): the error is the rotation vector from
// A's frame to B's (twice the vector part of qb qa*, the short way).
NPH_HD void joint_orientation_rows(NphJointState *j, NphJointPose *p, const NphJointStep *step,
                                   int soft)
{
    NphVec3 bias = v3_zero();
    f32 mass_scale = 1.0F, impulse_scale = 0.0F;
    if (soft) {
        NphQuat off = q_mul(p->qb, q_conj(p->qa));
        if (off.w < 0.0F) {
            off = joint_flip(off);
        }
        bias = v3_scale(q_xyz(off), 2.0F * step->soft.bias);
        mass_scale = step->soft.mass;
        impulse_scale = step->soft.impulse;
    }
    NphVec3 k[3];
    joint_rotation_matrix(j, k);
    const NphVec3 delta = joint_block_impulse(k, joint_spin(p), bias, j->angular_impulse,
                                              mass_scale, impulse_scale);
    j->angular_impulse = v3_add(j->angular_impulse, delta);
    joint_twist(j, p, delta);
}

// The hinge's two locked axes (A's Y and Z): a 2 x 2 block whose impulses
// accumulate in angular_impulse.x and .y; the error is how far B's X has
// turned from A's, xa x xb, on each.
NPH_HD void joint_hinge_axes_rows(NphJointState *j, NphJointPose *p, const NphJointStep *step,
                                  int soft)
{
    const NphVec3 u = q_axis_y(p->qa);
    const NphVec3 v = q_axis_z(p->qa);
    f32 bu = 0.0F, bv = 0.0F, mass_scale = 1.0F, impulse_scale = 0.0F;
    if (soft) {
        const NphVec3 tilt = v3_cross(q_axis_x(p->qa), q_axis_x(p->qb));
        bu = step->soft.bias * v3_dot(tilt, u);
        bv = step->soft.bias * v3_dot(tilt, v);
        mass_scale = step->soft.mass;
        impulse_scale = step->soft.impulse;
    }
    const NphVec3 iu = v3_add(m3_mul_v3(j->inv_inertia_a, u), m3_mul_v3(j->inv_inertia_b, u));
    const NphVec3 iv = v3_add(m3_mul_v3(j->inv_inertia_a, v), m3_mul_v3(j->inv_inertia_b, v));
    const f32 kuu = v3_dot(u, iu);
    const f32 kuv = v3_dot(u, iv);
    const f32 kvv = v3_dot(v, iv);
    const f32 det = kuu * kvv - kuv * kuv;
    if (!(det > 0.0F)) {
        return;
    }
    const f32 inv = 1.0F / det;
    const NphVec3 spin = joint_spin(p);
    const f32 eu = v3_dot(spin, u) + bu;
    const f32 ev = v3_dot(spin, v) + bv;
    const f32 du = -mass_scale * ((kvv * eu - kuv * ev) * inv) - impulse_scale * j->angular_impulse.x;
    const f32 dv = -mass_scale * ((kuu * ev - kuv * eu) * inv) - impulse_scale * j->angular_impulse.y;
    j->angular_impulse.x += du;
    j->angular_impulse.y += dv;
    joint_twist(j, p, v3_add(v3_scale(u, du), v3_scale(v, dv)));
}

// The ball's friction: the relative spin stopped, within the friction
// torque (the accumulated impulse clamped to a ball).
NPH_HD void joint_friction_rows(NphJointState *j, NphJointPose *p, const NphJointStep *step,
                                f32 torque)
{
    NphVec3 k[3];
    joint_rotation_matrix(j, k);
    const NphVec3 stop = joint_solve_sym3(k, joint_spin(p));
    const NphVec3 before = j->friction_impulse;
    NphVec3 now = v3_sub(before, stop);
    const f32 cap = step->h * torque;
    const f32 size_sq = v3_len_sq(now);
    if (size_sq > cap * cap) {
        now = v3_scale(now, cap / nph_sqrtf(size_sq));
    }
    j->friction_impulse = now;
    joint_twist(j, p, v3_sub(now, before));
}

// The rows of each type, in the order they are solved.
enum {
    NPH_ROW_END = 0,
    NPH_ROW_SPRING,
    NPH_ROW_MOTOR,
    NPH_ROW_LOWER,
    NPH_ROW_UPPER,
    NPH_ROW_HINGE_AXES,
    NPH_ROW_FRICTION,
    NPH_ROW_SWING,
    NPH_ROW_TWIST_LOWER,
    NPH_ROW_TWIST_UPPER,
    NPH_ROW_ORIENTATION,
    NPH_ROW_ANCHORS,
};

NPH_HD u32 joint_row(u32 type, u32 k)
{
    if (type == (u32)NPH_PHYSICS_JOINT_HINGE) { // spring .. hinge axes, then the anchors
        return k < 5U ? NPH_ROW_SPRING + k : (k == 5U ? NPH_ROW_ANCHORS : NPH_ROW_END);
    }
    if (type == (u32)NPH_PHYSICS_JOINT_BALL) {
        return k < 4U ? NPH_ROW_FRICTION + k : (k == 4U ? NPH_ROW_ANCHORS : NPH_ROW_END);
    }
    return k == 0U ? NPH_ROW_ORIENTATION : (k == 1U ? NPH_ROW_ANCHORS : NPH_ROW_END);
}

NPH_HD void joint_run_row(u32 row, const NphJointConf *cf, NphJointState *j, NphJointPose *p,
                          const NphJointShape *g, const NphJointStep *step, int soft)
{
    const NphPhysicsHingeJoint *hinge = &cf->s.hinge;
    const NphPhysicsBallJoint *ball = &cf->s.ball;
    switch (row) {
    case NPH_ROW_SPRING:
        if (hinge->enable_spring != 0 && soft) {
            joint_spring_row(hinge, j, p, g);
        }
        return;
    case NPH_ROW_MOTOR:
        if (hinge->enable_motor != 0) {
            joint_motor_row(hinge, j, p, g, step);
        }
        return;
    case NPH_ROW_LOWER:
        if (hinge->enable_limit != 0) {
            joint_limit_row(j, p, v3_neg(g->hinge_axis), g->hinge_angle - hinge->lower,
                            &j->lower_impulse, step, soft);
        }
        return;
    case NPH_ROW_UPPER:
        if (hinge->enable_limit != 0) {
            joint_limit_row(j, p, g->hinge_axis, hinge->upper - g->hinge_angle, &j->upper_impulse,
                            step, soft);
        }
        return;
    case NPH_ROW_HINGE_AXES:
        joint_hinge_axes_rows(j, p, step, soft);
        return;
    case NPH_ROW_FRICTION:
        if (ball->friction_torque > 0.0F) {
            joint_friction_rows(j, p, step, ball->friction_torque);
        }
        return;
    case NPH_ROW_SWING:
        if (g->swing_on) {
            joint_limit_row(j, p, g->swing_axis, g->swing_gap, &j->swing_impulse, step, soft);
        }
        return;
    case NPH_ROW_TWIST_LOWER:
        if (g->twist_on) {
            joint_limit_row(j, p, v3_neg(g->twist_axis), g->twist_angle - ball->twist_lower,
                            &j->lower_impulse, step, soft);
        }
        return;
    case NPH_ROW_TWIST_UPPER:
        if (g->twist_on) {
            joint_limit_row(j, p, g->twist_axis, ball->twist_upper - g->twist_angle,
                            &j->upper_impulse, step, soft);
        }
        return;
    case NPH_ROW_ORIENTATION:
        joint_orientation_rows(j, p, step, soft);
        return;
    default:
        joint_anchor_rows(j, p, step, soft);
        return;
    }
}

// ---------------------------------------------------------------------------
// Step entry points (joint_solver.c:726-907)
// ---------------------------------------------------------------------------

// Before the passes: masses, anchors and frames at the step's start; what is
// switched off starts from nothing when switched on again.
NPH_HD void joint_prepare(const NphJointConf *cf, NphJointState *j, const NphBodyRec *a,
                          const NphBodyRec *b, u32 solver_a, u32 solver_b,
                          const NphSolverParams *params, f32 h)
{
    const int moves_a = a->type == (u32)NPH_PHYSICS_BODY_DYNAMIC;
    const int moves_b = b->type == (u32)NPH_PHYSICS_BODY_DYNAMIC;
    j->solver_a = solver_a;
    j->solver_b = solver_b;
    j->write = (moves_a ? NPH_JOINT_WRITE_A : 0U) | (moves_b ? NPH_JOINT_WRITE_B : 0U);
    j->inv_mass_a = moves_a ? a->inv_mass : 0.0F;
    j->inv_mass_b = moves_b ? b->inv_mass : 0.0F;
    j->inv_inertia_a = moves_a ? params[solver_a].inv_inertia : m3_zero();
    j->inv_inertia_b = moves_b ? params[solver_b].inv_inertia : m3_zero();
    j->anchor_a = q_rotate(a->rotation, v3_sub(cf->frame_a.position, a->local_center));
    j->anchor_b = q_rotate(b->rotation, v3_sub(cf->frame_b.position, b->local_center));
    j->delta_center = dv3_relative(b->center, a->center);
    j->frame_q_a = q_mul(a->rotation, cf->frame_a.rotation);
    j->frame_q_b = q_mul(b->rotation, cf->frame_b.rotation);
    j->spring_bias = 0.0F;
    j->spring_mass = 1.0F;
    j->spring_impulse_scale = 0.0F;
    if (cf->type == (u32)NPH_PHYSICS_JOINT_HINGE) {
        const NphPhysicsHingeJoint *hinge = &cf->s.hinge;
        if (hinge->enable_limit == 0) {
            j->lower_impulse = j->upper_impulse = 0.0F;
        }
        if (hinge->enable_motor == 0) {
            j->motor_impulse = 0.0F;
        }
        if (hinge->enable_spring == 0) {
            j->spring_impulse = 0.0F;
        }
        make_soft(hinge->enable_spring != 0 ? hinge->spring_hertz : 0.0F,
                  hinge->spring_damping_ratio, h, &j->spring_bias, &j->spring_mass,
                  &j->spring_impulse_scale);
    } else if (cf->type == (u32)NPH_PHYSICS_JOINT_BALL) {
        const NphPhysicsBallJoint *ball = &cf->s.ball;
        if (ball->enable_swing_limit == 0) {
            j->swing_impulse = 0.0F;
        }
        if (ball->enable_twist_limit == 0) {
            j->lower_impulse = j->upper_impulse = 0.0F;
        }
        if (!(ball->friction_torque > 0.0F)) {
            j->friction_impulse = v3_zero();
        }
    }
}

// The accumulated angular impulse as one world vector, for the warm start
// and the reaction torque (joint_solver.c:826-853): the hinge's locked axes
// and its axial rows; the ball's friction, swing and twist rows.
NPH_HD NphVec3 joint_angular_impulse(const NphJointConf *cf, const NphJointState *j, NphQuat qa,
                                     NphQuat qb)
{
    if (cf->type == (u32)NPH_PHYSICS_JOINT_HINGE) {
        const NphVec3 locked = v3_add(v3_scale(q_axis_y(qa), j->angular_impulse.x),
                                      v3_scale(q_axis_z(qa), j->angular_impulse.y));
        const f32 axial = j->lower_impulse - j->upper_impulse + j->motor_impulse + j->spring_impulse;
        return v3_mul_add(locked, q_axis_x(qa), axial);
    }
    if (cf->type == (u32)NPH_PHYSICS_JOINT_BALL) {
        NphVec3 total = j->friction_impulse;
        const NphVec3 xa = q_axis_x(qa);
        const NphVec3 xb = q_axis_x(qb);
        const NphVec3 across = v3_cross(xa, xb);
        const f32 s = v3_len(across);
        if (j->swing_impulse != 0.0F && s > 1e-6F) {
            total = v3_mul_add(total, across, -j->swing_impulse / s);
        }
        const NphVec3 mid = v3_add(xa, xb);
        const f32 len = v3_len(mid);
        if (len > 1e-3F) {
            total = v3_mul_add(total, mid, (j->lower_impulse - j->upper_impulse) / len);
        }
        return total;
    }
    return j->angular_impulse;
}

NPH_HD void joint_warm_start(NphSolverBody *bodies, const NphJointConf *cf, NphJointState *j)
{
    NphJointPose p = joint_pose(bodies, j);
    joint_twist(j, &p, joint_angular_impulse(cf, j, p.qa, p.qb));
    joint_push(j, &p, j->linear_impulse);
}

// One pass of the joint's rows: soft (solve) or rigid (relax).
NPH_HD void joint_solve(NphSolverBody *bodies, const NphJointStep *step, const NphJointConf *cf,
                        NphJointState *j, int soft)
{
    NphJointPose p = joint_pose(bodies, j);
    const NphJointShape g = joint_shape(cf, j, &p);
    for (u32 k = 0U;; ++k) {
        const u32 row = joint_row(cf->type, k);
        if (row == NPH_ROW_END) {
            break;
        }
        joint_run_row(row, cf, j, &p, &g, step, soft);
    }
}

#endif // NPH_JOINT_H
