// SPDX-License-Identifier: Apache-2.0
// The learned warm start's per-point quantities, shared by the step's
// kernels: the network's 14 input features of a contact point (read after
// the lanes are prepared, before the first substep), and the post-solve
// residual that decides whether the network keeps a contact.
//
// The reference has no learned component and no residual metric; its warm
// start is the impulse carried from the previous step by point id or place
// (contact.c:296-374), applied by solver_impl.h:181-258, and the impulses the
// solver converges to are what store_group keeps (solver.c:475-499).  The
// features are the constants that warm start and solve read for the point
// (world_internal.h:430-479).
#ifndef NPH_LEARNED_H
#define NPH_LEARNED_H

#include "nph_step.h"

// The network's inputs: approach speed along the normal and the two
// tangents, separation, normal, the two inverse masses, friction,
// normal mass and the tangent mass terms.
enum {
    NPH_F_RELV = 0,
    NPH_F_VT1,
    NPH_F_VT2,
    NPH_F_SEP,
    NPH_F_NX,
    NPH_F_NY,
    NPH_F_NZ,
    NPH_F_IMA,
    NPH_F_IMB,
    NPH_F_FRICTION,
    NPH_F_NMASS,
    NPH_F_TM0,
    NPH_F_TM1,
    NPH_F_TM2,
    NPH_F_COUNT
};

// Network rows: four per lane (one per point slot), lane-major.
#define NPH_LEARNED_ROWS_PER_LANE 4U
#define NPH_LEARNED_IN 16U  // inputs per row (14 used)
#define NPH_LEARNED_OUT 4U  // outputs per row (3 used)

// A lane that some point's warm start came from the network.
#define NPH_LANE_LEARNED 8U

// In the every-point mode, contact records whose index is a multiple of this
// keep the plain warm start (the gate's control sample).
#define NPH_LEARNED_CONTROL 16U

// The features of point k of lane L (group g, lane l), after lane_prepare.
NPH_HD void learned_features(const NphStepData *d, u32 g, u32 l, u32 k, f32 f[NPH_F_COUNT])
{
    const NphGroup *grp = &d->groups[g];
    const u32 L = g * NPH_LANES + l;
    const u64 R = (u64)(grp->first_point + k) * NPH_LANES + l;
    const NphLaneBody A = lane_gather(&d->sbodies[d->lane_body_a[L]], 0);
    const NphLaneBody B = lane_gather(&d->sbodies[d->lane_body_b[L]], 0);
    const NphVec3 ra = PV3(d, NPH_P_RAX, R);
    const NphVec3 rb = PV3(d, NPH_P_RBX, R);
    const NphVec3 dv = lane_relative_velocity(&A, &B, ra, rb);
    f[NPH_F_RELV] = PF(d, NPH_P_RELV, R);
    f[NPH_F_VT1] = v3_dot(dv, lane_axis(d, NPH_L_T1X, L));
    f[NPH_F_VT2] = v3_dot(dv, lane_axis(d, NPH_L_T2X, L));
    // The point's separation at the step's start (the row keeps it less
    // the anchors' offset along the normal).
    const NphVec3 n = lane_axis(d, NPH_L_NX, L);
    f[NPH_F_SEP] = v3_dot(v3(rb.x - ra.x, rb.y - ra.y, rb.z - ra.z), n) + PF(d, NPH_P_SEP, R);
    f[NPH_F_NX] = LF(d, NPH_L_NX, L);
    f[NPH_F_NY] = LF(d, NPH_L_NY, L);
    f[NPH_F_NZ] = LF(d, NPH_L_NZ, L);
    f[NPH_F_IMA] = LF(d, NPH_L_IMA, L);
    f[NPH_F_IMB] = LF(d, NPH_L_IMB, L);
    f[NPH_F_FRICTION] = LF(d, NPH_L_FRICTION, L);
    f[NPH_F_NMASS] = PF(d, NPH_P_NMASS, R);
    f[NPH_F_TM0] = PF(d, NPH_P_TM0, R);
    f[NPH_F_TM1] = PF(d, NPH_P_TM1, R);
    f[NPH_F_TM2] = PF(d, NPH_P_TM2, R);
}

// The network's outputs are impulses divided by the point's normal mass
// (velocities), scaled: impulse = y * out_scale * normal mass.  The normal
// impulse is clamped at zero and the friction impulse to the friction
// circle of that normal impulse, as the solver's rows clamp them.
NPH_HD void learned_impulses(const f32 y[3], const f32 out_scale[3], f32 nmass, f32 friction,
                             f32 *nimp, f32 *ti1, f32 *ti2)
{
    const f32 n = nph_maxf(y[0] * out_scale[0] * nmass, 0.0F);
    f32 a = y[1] * out_scale[1] * nmass;
    f32 b = y[2] * out_scale[2] * nmass;
    const f32 limit = friction * n;
    const f32 len = nph_sqrtf(a * a + b * b);
    if (len > limit) {
        const f32 r = len > 0.0F ? limit / len : 0.0F;
        a *= r;
        b *= r;
    }
    *nimp = n;
    *ti1 = a;
    *ti2 = b;
}

// The relax pass's contact conditions at the end of the step, per point:
// with s the current separation and b = s / h for a speculative point
// (s > 0) else 0, the relax rows drive vn + b to zero while the normal
// impulse is positive and keep vn + b >= 0 when it is zero.  The residual is
// |vn + b| for a loaded point and max(0, -(vn + b)) for an unloaded one, in
// m/s; a lane's residual is its points' maximum, and its penetration the
// largest max(0, -s).
NPH_HD void learned_residual(const NphStepData *d, f32 inv_h, u32 g, u32 l, const NphContactRec *c,
                             f32 *residual, f32 *penetration)
{
    const NphGroup *grp = &d->groups[g];
    const u32 L = g * NPH_LANES + l;
    const NphLaneBody A = lane_gather(&d->sbodies[d->lane_body_a[L]], 1);
    const NphLaneBody B = lane_gather(&d->sbodies[d->lane_body_b[L]], 1);
    const NphVec3 n = lane_axis(d, NPH_L_NX, L);
    f32 res = 0.0F;
    f32 pen = 0.0F;
    for (u32 k = 0U; k < c->point_count; ++k) {
        const u64 R = (u64)(grp->first_point + k) * NPH_LANES + l;
        const NphVec3 ra = PV3(d, NPH_P_RAX, R);
        const NphVec3 rb = PV3(d, NPH_P_RBX, R);
        const f32 s = lane_separation(&A, &B, ra, rb, n, PF(d, NPH_P_SEP, R));
        const f32 vn = v3_dot(lane_relative_velocity(&A, &B, ra, rb), n);
        const f32 e = vn + (s > 0.0F ? s * inv_h : 0.0F);
        const f32 r = PF(d, NPH_P_NIMP, R) > 0.0F ? nph_absf(e) : nph_maxf(-e, 0.0F);
        res = nph_maxf(res, r);
        pen = nph_maxf(pen, -s);
    }
    *residual = res;
    *penetration = pen;
}

#endif
