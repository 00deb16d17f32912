// SPDX-License-Identifier: Apache-2.0
// TEST INFRASTRUCTURE ONLY: the oracle of the solver's arithmetic (see
// oracle.h).  A restatement of engine/physics/src/solver_impl.h, solver.c:48-65
// and engine/core/src/hash.c:73-120 written lane by lane: every eight-wide
// operation of the reference is a loop over the eight lanes doing the same
// IEEE operation in the same order, masks are 0/1 integers, and selects are
// conditionals.  Built with the reference's flags (-ffp-contract=off, SSE2
// scalar arithmetic) so each operation rounds as the reference's.
#include "oracle.h"

#include <math.h>
#include <string.h>

typedef float W[OR_LANES];

#define FOR_LANES for (int l = 0; l < OR_LANES; ++l)

static float fmin_ref(float a, float b) { return a < b ? a : b; } // minps
static float fmax_ref(float a, float b) { return a > b ? a : b; } // maxps

// A lane's body rows (solver_impl.h:27-32).
typedef struct Bodies8 {
    W vx, vy, vz, wx, wy, wz, dpx, dpy, dpz, qx, qy, qz, qw;
} Bodies8;

// solver_impl.h:43-69: rows of the lanes' bodies.
static void gather(const OrSolverBody *bodies, const uint32_t *index, int positions, Bodies8 *o)
{
    FOR_LANES
    {
        const OrSolverBody *b = &bodies[index[l]];
        o->vx[l] = b->v[0];
        o->vy[l] = b->v[1];
        o->vz[l] = b->v[2];
        o->wx[l] = b->w[0];
        o->wy[l] = b->w[1];
        o->wz[l] = b->w[2];
        if (positions) {
            o->dpx[l] = b->dp[0];
            o->dpy[l] = b->dp[1];
            o->dpz[l] = b->dp[2];
            o->qx[l] = b->dq[0];
            o->qy[l] = b->dq[1];
            o->qz[l] = b->dq[2];
            o->qw[l] = b->dq[3];
        }
    }
}

// solver_impl.h:74-105: masked lanes' v and w back, in lane order, pads zero.
static void scatter(OrSolverBody *bodies, const uint32_t *index, uint32_t mask, const Bodies8 *b)
{
    FOR_LANES
    {
        if ((mask & (1U << l)) == 0U) {
            continue;
        }
        OrSolverBody *o = &bodies[index[l]];
        o->v[0] = b->vx[l];
        o->v[1] = b->vy[l];
        o->v[2] = b->vz[l];
        o->v[3] = 0.0F;
        o->w[0] = b->wx[l];
        o->w[1] = b->wy[l];
        o->w[2] = b->wz[l];
        o->w[3] = 0.0F;
    }
}

static float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx + ay * by) + az * bz;
}

// (ay bz - az by, az bx - ax bz, ax by - ay bx)
static void cross3(float ax, float ay, float az, float bx, float by, float bz, float *ox, float *oy,
                   float *oz)
{
    *ox = ay * bz - az * by;
    *oy = az * bx - ax * bz;
    *oz = ax * by - ay * bx;
}

// solver_impl.h:122-142 for lane l: (vB + wB x rB) - (vA + wA x rA)
static void rel_velocity(const Bodies8 *a, const Bodies8 *b, const OrPoints *p, int l, float *dx,
                         float *dy, float *dz)
{
    float ax, ay, az, bx, by, bz;
    cross3(a->wx[l], a->wy[l], a->wz[l], p->anchor_a[0][l], p->anchor_a[1][l], p->anchor_a[2][l],
           &ax, &ay, &az);
    cross3(b->wx[l], b->wy[l], b->wz[l], p->anchor_b[0][l], p->anchor_b[1][l], p->anchor_b[2][l],
           &bx, &by, &bz);
    *dx = (b->vx[l] + bx) - (a->vx[l] + ax);
    *dy = (b->vy[l] + by) - (a->vy[l] + ay);
    *dz = (b->vz[l] + bz) - (a->vz[l] + az);
}

// solver_impl.h:147-168 for lane l
static void apply(Bodies8 *a, Bodies8 *b, float ma, float mb, float lambda, float x, float y,
                  float z, const float (*ang_a)[OR_LANES], const float (*ang_b)[OR_LANES], int l)
{
    const float la = ma * lambda;
    const float lb = mb * lambda;
    a->vx[l] = a->vx[l] - la * x;
    a->vy[l] = a->vy[l] - la * y;
    a->vz[l] = a->vz[l] - la * z;
    b->vx[l] = b->vx[l] + lb * x;
    b->vy[l] = b->vy[l] + lb * y;
    b->vz[l] = b->vz[l] + lb * z;
    a->wx[l] = a->wx[l] - lambda * ang_a[0][l];
    a->wy[l] = a->wy[l] - lambda * ang_a[1][l];
    a->wz[l] = a->wz[l] - lambda * ang_a[2][l];
    b->wx[l] = b->wx[l] + lambda * ang_b[0][l];
    b->wy[l] = b->wy[l] + lambda * ang_b[1][l];
    b->wz[l] = b->wz[l] + lambda * ang_b[2][l];
}

static int mask_set(const float *m, int l)
{
    uint32_t u;
    memcpy(&u, &m[l], 4);
    return (u >> 31U) != 0U; // blendv/movemask read the sign bit
}

static void mask_store(float *m, int l, int on)
{
    const uint32_t u = on ? 0xFFFFFFFFU : 0U;
    memcpy(&m[l], &u, 4);
}

void oracle_warm_start(OrSolverBody *bodies, OrGroup *g, OrPoints *points)
{
    Bodies8 a, b;
    gather(bodies, g->body_a, 0, &a);
    gather(bodies, g->body_b, 0, &b);
    int any_hit = 0;
    FOR_LANES { any_hit |= mask_set(g->impacted, l); }
    W scale;
    FOR_LANES { scale[l] = 1.0F; }
    if (any_hit) {
        Bodies8 a1 = a;
        Bodies8 b1 = b;
        for (uint32_t k = 0; k < g->point_count; ++k) {
            const OrPoints *p = &points[k];
            FOR_LANES
            {
                apply(&a1, &b1, g->inv_mass_a[l], g->inv_mass_b[l], p->normal_impulse[l],
                      g->normal[0][l], g->normal[1][l], g->normal[2][l], p->ang_a_n, p->ang_b_n, l);
                apply(&a1, &b1, g->inv_mass_a[l], g->inv_mass_b[l], p->tangent_impulse[0][l],
                      g->tangent1[0][l], g->tangent1[1][l], g->tangent1[2][l], p->ang_a_t1,
                      p->ang_b_t1, l);
                apply(&a1, &b1, g->inv_mass_a[l], g->inv_mass_b[l], p->tangent_impulse[1][l],
                      g->tangent2[0][l], g->tangent2[1][l], g->tangent2[2][l], p->ang_a_t2,
                      p->ang_b_t2, l);
            }
        }
        W share;
        FOR_LANES { share[l] = 1.0F; }
        for (uint32_t k = 0; k < g->point_count; ++k) {
            const OrPoints *p = &points[k];
            FOR_LANES
            {
                float vx, vy, vz;
                rel_velocity(&a, &b, p, l, &vx, &vy, &vz);
                const float before =
                    dot3(vx, vy, vz, g->normal[0][l], g->normal[1][l], g->normal[2][l]);
                rel_velocity(&a1, &b1, p, l, &vx, &vy, &vz);
                const float after =
                    dot3(vx, vy, vz, g->normal[0][l], g->normal[1][l], g->normal[2][l]);
                const float allowed = fmax_ref(before, 0.0F);
                const int pushed = after > allowed;
                const float part = (allowed - before) / (pushed ? after - before : 1.0F);
                share[l] = fmin_ref(share[l], pushed ? part : 1.0F);
            }
        }
        FOR_LANES { scale[l] = mask_set(g->impacted, l) ? share[l] : 1.0F; }
        FOR_LANES { g->impacted[l] = 0.0F; }
    }
    for (uint32_t k = 0; k < g->point_count; ++k) {
        OrPoints *p = &points[k];
        FOR_LANES
        {
            float normal = p->normal_impulse[l];
            float t1 = p->tangent_impulse[0][l];
            float t2 = p->tangent_impulse[1][l];
            if (any_hit) {
                normal = normal * scale[l];
                t1 = t1 * scale[l];
                t2 = t2 * scale[l];
                p->normal_impulse[l] = normal;
                p->tangent_impulse[0][l] = t1;
                p->tangent_impulse[1][l] = t2;
            }
            apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], normal, g->normal[0][l],
                  g->normal[1][l], g->normal[2][l], p->ang_a_n, p->ang_b_n, l);
            apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], t1, g->tangent1[0][l],
                  g->tangent1[1][l], g->tangent1[2][l], p->ang_a_t1, p->ang_b_t1, l);
            apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], t2, g->tangent2[0][l],
                  g->tangent2[1][l], g->tangent2[2][l], p->ang_a_t2, p->ang_b_t2, l);
        }
    }
    scatter(bodies, g->body_a, g->write_mask_a, &a);
    scatter(bodies, g->body_b, g->write_mask_b, &b);
}

// m v for column-major m (column * 3 + row), lane l
static void mat_mul(const float m[9][OR_LANES], const float v[3], float out[3], int l)
{
    for (int r = 0; r < 3; ++r) {
        out[r] = (m[r][l] * v[0] + m[3 + r][l] * v[1]) + m[6 + r][l] * v[2];
    }
}

void oracle_prepare(const OrSolverBody *bodies, OrGroup *g, OrPoints *points, const OrPrepareIn *in)
{
    Bodies8 a, b;
    gather(bodies, g->body_a, 0, &a);
    gather(bodies, g->body_b, 0, &b);
    for (uint32_t k = 0; k < g->point_count; ++k) {
        OrPoints *p = &points[k];
        FOR_LANES
        {
            const int active = in->active[k][l] > 0.0F;
            const float mass = g->inv_mass_a[l] + g->inv_mass_b[l];
            const float ax[3][3] = {{g->normal[0][l], g->normal[1][l], g->normal[2][l]},
                                    {g->tangent1[0][l], g->tangent1[1][l], g->tangent1[2][l]},
                                    {g->tangent2[0][l], g->tangent2[1][l], g->tangent2[2][l]}};
            const float ra[3] = {p->anchor_a[0][l], p->anchor_a[1][l], p->anchor_a[2][l]};
            const float rb[3] = {p->anchor_b[0][l], p->anchor_b[1][l], p->anchor_b[2][l]};
            const float gap = dot3(rb[0] - ra[0], rb[1] - ra[1], rb[2] - ra[2], ax[0][0], ax[0][1],
                                   ax[0][2]);
            p->adjusted_separation[l] = p->adjusted_separation[l] - gap;
            float (*const ang_a[3])[OR_LANES] = {p->ang_a_n, p->ang_a_t1, p->ang_a_t2};
            float (*const ang_b[3])[OR_LANES] = {p->ang_b_n, p->ang_b_t1, p->ang_b_t2};
            float rxa[3][3], rxb[3][3], aa[3][3], ab[3][3];
            for (int x = 0; x < 3; ++x) {
                cross3(ra[0], ra[1], ra[2], ax[x][0], ax[x][1], ax[x][2], &rxa[x][0], &rxa[x][1],
                       &rxa[x][2]);
                cross3(rb[0], rb[1], rb[2], ax[x][0], ax[x][1], ax[x][2], &rxb[x][0], &rxb[x][1],
                       &rxb[x][2]);
                mat_mul(in->inv_inertia_a, rxa[x], aa[x], l);
                mat_mul(in->inv_inertia_b, rxb[x], ab[x], l);
                for (int c = 0; c < 3; ++c) {
                    ang_a[x][c][l] = aa[x][c];
                    ang_b[x][c][l] = ab[x][c];
                }
            }
            const float kn = (mass + dot3(rxa[0][0], rxa[0][1], rxa[0][2], aa[0][0], aa[0][1],
                                          aa[0][2])) +
                             dot3(rxb[0][0], rxb[0][1], rxb[0][2], ab[0][0], ab[0][1], ab[0][2]);
            p->normal_mass[l] = (active && kn > 0.0F) ? 1.0F / kn : 0.0F;
            const float k11 = (mass + dot3(rxa[1][0], rxa[1][1], rxa[1][2], aa[1][0], aa[1][1],
                                           aa[1][2])) +
                              dot3(rxb[1][0], rxb[1][1], rxb[1][2], ab[1][0], ab[1][1], ab[1][2]);
            const float k22 = (mass + dot3(rxa[2][0], rxa[2][1], rxa[2][2], aa[2][0], aa[2][1],
                                           aa[2][2])) +
                              dot3(rxb[2][0], rxb[2][1], rxb[2][2], ab[2][0], ab[2][1], ab[2][2]);
            const float k12 = dot3(rxa[1][0], rxa[1][1], rxa[1][2], aa[2][0], aa[2][1], aa[2][2]) +
                              dot3(rxb[1][0], rxb[1][1], rxb[1][2], ab[2][0], ab[2][1], ab[2][2]);
            const float det = k11 * k22 - k12 * k12;
            const float inv = (active && det > 0.0F) ? 1.0F / det : 0.0F;
            p->tangent_mass[0][l] = k22 * inv;
            p->tangent_mass[1][l] = (-k12) * inv;
            p->tangent_mass[2][l] = k11 * inv;
            float vx, vy, vz;
            rel_velocity(&a, &b, p, l, &vx, &vy, &vz);
            p->relative_velocity[l] =
                active ? dot3(vx, vy, vz, ax[0][0], ax[0][1], ax[0][2]) : 0.0F;
            const float tx = in->tangent_impulse[k][0][l];
            const float ty = in->tangent_impulse[k][1][l];
            const float tz = in->tangent_impulse[k][2][l];
            p->tangent_impulse[0][l] = dot3(tx, ty, tz, ax[1][0], ax[1][1], ax[1][2]);
            p->tangent_impulse[1][l] = dot3(tx, ty, tz, ax[2][0], ax[2][1], ax[2][2]);
        }
    }
    FOR_LANES
    {
        int impact = 0;
        for (uint32_t k = 0; k < g->point_count; ++k) {
            impact |= points[k].relative_velocity[l] < -1.0F;
        }
        mask_store(g->impact_lanes, l, impact);
    }
}

// solver_impl.h:400-421
static void friction_circle(float vt1, float vt2, float limit, float *new1, float *new2)
{
    const float len = sqrtf(*new1 * *new1 + *new2 * *new2);
    const int over = len > limit;
    const float slip = sqrtf(vt1 * vt1 + vt2 * vt2);
    const int sliding = over && slip > 0.0F;
    const float r = limit / (over ? len : 1.0F);
    const float stop_share = r / (over ? len : 1.0F);
    const float slip_share = (1.0F - r) / (sliding ? slip : 1.0F);
    const float d1 = stop_share * *new1 - slip_share * vt1;
    const float d2 = stop_share * *new2 - slip_share * vt2;
    const float dlen = sqrtf(d1 * d1 + d2 * d2);
    const float to_edge = limit / (sliding ? dlen : 1.0F);
    const float n1 = sliding ? d1 * to_edge : (over ? *new1 * r : *new1);
    const float n2 = sliding ? d2 * to_edge : (over ? *new2 * r : *new2);
    *new1 = n1;
    *new2 = n2;
}

// solver_impl.h:425-464 for lane l: anchors rotated by the bodies' dq.
static void rotate_lane(float qx, float qy, float qz, float qw, float vx, float vy, float vz,
                        float *ox, float *oy, float *oz)
{
    float tx, ty, tz;
    cross3(qx, qy, qz, vx, vy, vz, &tx, &ty, &tz);
    tx = tx * 2.0F;
    ty = ty * 2.0F;
    tz = tz * 2.0F;
    float cx, cy, cz;
    cross3(qx, qy, qz, tx, ty, tz, &cx, &cy, &cz);
    *ox = (vx + tx * qw) + cx;
    *oy = (vy + ty * qw) + cy;
    *oz = (vz + tz * qw) + cz;
}

static float separation(const Bodies8 *a, const Bodies8 *b, const OrPoints *p, const OrGroup *g,
                        int l)
{
    float pax, pay, paz, pbx, pby, pbz;
    rotate_lane(a->qx[l], a->qy[l], a->qz[l], a->qw[l], p->anchor_a[0][l], p->anchor_a[1][l],
                p->anchor_a[2][l], &pax, &pay, &paz);
    rotate_lane(b->qx[l], b->qy[l], b->qz[l], b->qw[l], p->anchor_b[0][l], p->anchor_b[1][l],
                p->anchor_b[2][l], &pbx, &pby, &pbz);
    const float dx = (b->dpx[l] - a->dpx[l]) + (pbx - pax);
    const float dy = (b->dpy[l] - a->dpy[l]) + (pby - pay);
    const float dz = (b->dpz[l] - a->dpz[l]) + (pbz - paz);
    return dot3(dx, dy, dz, g->normal[0][l], g->normal[1][l], g->normal[2][l]) +
           p->adjusted_separation[l];
}

// solver_impl.h:471-482
static void normal_row(float s, float inv_h, float slop, float soft_bias, float push_max,
                       float soft_mass, float soft_impulse, float *bias, float *mass_scale,
                       float *impulse_scale)
{
    const int apart = s > 0.0F;
    const float overlap = fmin_ref(s + slop, 0.0F);
    *bias = apart ? s * inv_h : fmax_ref(soft_bias * overlap, push_max);
    *mass_scale = apart ? 1.0F : soft_mass;
    *impulse_scale = apart ? 0.0F : soft_impulse;
}

void oracle_solve(OrSolverBody *bodies, OrGroup *g, OrPoints *points, float inv_h, float push_max,
                  float slop, int use_bias)
{
    Bodies8 a, b;
    gather(bodies, g->body_a, 1, &a);
    gather(bodies, g->body_b, 1, &b);
    const float min_bias = -push_max;
    for (uint32_t k = 0; k < g->point_count; ++k) {
        OrPoints *p = &points[k];
        FOR_LANES
        {
            const float soft_bias = use_bias ? g->bias_rate[l] : 0.0F;
            const float soft_mass = use_bias ? g->mass_scale[l] : 1.0F;
            const float soft_impulse = use_bias ? g->impulse_scale[l] : 0.0F;
            float bias, ms, is;
            normal_row(separation(&a, &b, p, g, l), inv_h, slop, soft_bias, min_bias, soft_mass,
                       soft_impulse, &bias, &ms, &is);
            float vx, vy, vz;
            rel_velocity(&a, &b, p, l, &vx, &vy, &vz);
            const float vn = dot3(vx, vy, vz, g->normal[0][l], g->normal[1][l], g->normal[2][l]);
            const float row = p->normal_mass[l] * ms;
            const float acc = p->normal_impulse[l];
            const float impulse = -(row * (vn + bias)) - is * acc;
            const float total = fmax_ref(acc + impulse, 0.0F);
            const float delta = total - acc;
            p->normal_impulse[l] = total;
            p->max_normal_impulse[l] = fmax_ref(p->max_normal_impulse[l], total);
            apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], delta, g->normal[0][l],
                  g->normal[1][l], g->normal[2][l], p->ang_a_n, p->ang_b_n, l);
        }
    }
    int any_impact = 0;
    FOR_LANES { any_impact |= mask_set(g->impact_lanes, l); }
    if (g->scale_hits != 0U && any_impact) {
        FOR_LANES
        {
            mask_store(g->impacted, l, mask_set(g->impacted, l) || mask_set(g->impact_lanes, l));
        }
    }
    if (g->point_count > 1U && any_impact) {
        float row_bias[OR_MAX_POINTS][OR_LANES], row_mass[OR_MAX_POINTS][OR_LANES],
            row_soft[OR_MAX_POINTS][OR_LANES];
        for (uint32_t k = 0; k < g->point_count; ++k) {
            FOR_LANES
            {
                const float soft_bias = use_bias ? g->bias_rate[l] : 0.0F;
                const float soft_mass = use_bias ? g->mass_scale[l] : 1.0F;
                const float soft_impulse = use_bias ? g->impulse_scale[l] : 0.0F;
                float ms;
                normal_row(separation(&a, &b, &points[k], g, l), inv_h, slop, soft_bias, min_bias,
                           soft_mass, soft_impulse, &row_bias[k][l], &ms, &row_soft[k][l]);
                row_mass[k][l] = points[k].normal_mass[l] * ms;
            }
        }
        for (int sweep = 0; sweep < 7; ++sweep) {
            for (uint32_t k = 0; k < g->point_count; ++k) {
                OrPoints *p = &points[k];
                FOR_LANES
                {
                    float vx, vy, vz;
                    rel_velocity(&a, &b, p, l, &vx, &vy, &vz);
                    const float vn =
                        dot3(vx, vy, vz, g->normal[0][l], g->normal[1][l], g->normal[2][l]);
                    const float acc = p->normal_impulse[l];
                    const float impulse =
                        -(row_mass[k][l] * (vn + row_bias[k][l])) - row_soft[k][l] * acc;
                    const float total =
                        mask_set(g->impact_lanes, l) ? fmax_ref(acc + impulse, 0.0F) : acc;
                    const float delta = total - acc;
                    p->normal_impulse[l] = total;
                    p->max_normal_impulse[l] = fmax_ref(p->max_normal_impulse[l], total);
                    apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], delta, g->normal[0][l],
                          g->normal[1][l], g->normal[2][l], p->ang_a_n, p->ang_b_n, l);
                }
            }
        }
    }
    for (uint32_t k = 0; k < g->point_count; ++k) {
        OrPoints *p = &points[k];
        FOR_LANES
        {
            float vx, vy, vz;
            rel_velocity(&a, &b, p, l, &vx, &vy, &vz);
            const float vt1 =
                dot3(vx, vy, vz, g->tangent1[0][l], g->tangent1[1][l], g->tangent1[2][l]);
            const float vt2 =
                dot3(vx, vy, vz, g->tangent2[0][l], g->tangent2[1][l], g->tangent2[2][l]);
            const float m00 = p->tangent_mass[0][l];
            const float m01 = p->tangent_mass[1][l];
            const float m11 = p->tangent_mass[2][l];
            const float old1 = p->tangent_impulse[0][l];
            const float old2 = p->tangent_impulse[1][l];
            float new1 = old1 - (m00 * vt1 + m01 * vt2);
            float new2 = old2 - (m01 * vt1 + m11 * vt2);
            friction_circle(vt1, vt2, g->friction[l] * p->normal_impulse[l], &new1, &new2);
            p->tangent_impulse[0][l] = new1;
            p->tangent_impulse[1][l] = new2;
            apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], new1 - old1, g->tangent1[0][l],
                  g->tangent1[1][l], g->tangent1[2][l], p->ang_a_t1, p->ang_b_t1, l);
            apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], new2 - old2, g->tangent2[0][l],
                  g->tangent2[1][l], g->tangent2[2][l], p->ang_a_t2, p->ang_b_t2, l);
        }
    }
    scatter(bodies, g->body_a, g->write_mask_a, &a);
    scatter(bodies, g->body_b, g->write_mask_b, &b);
}

void oracle_restitution(OrSolverBody *bodies, OrGroup *g, OrPoints *points, float threshold)
{
    int loaded[OR_MAX_POINTS][OR_LANES];
    int bounced[OR_LANES] = {0};
    int any = 0;
    for (uint32_t k = 0; k < g->point_count; ++k) {
        FOR_LANES
        {
            loaded[k][l] = points[k].relative_velocity[l] < -threshold &&
                           points[k].max_normal_impulse[l] > 0.0F;
            bounced[l] |= loaded[k][l];
            any |= loaded[k][l];
        }
    }
    if (!any) {
        return;
    }
    Bodies8 a, b;
    gather(bodies, g->body_a, 0, &a);
    gather(bodies, g->body_b, 0, &b);
    const uint32_t sweeps = g->point_count > 1U ? 8U : 1U;
    for (uint32_t sweep = 0; sweep < sweeps; ++sweep) {
        for (uint32_t k = 0; k < g->point_count; ++k) {
            int row_any = 0;
            FOR_LANES { row_any |= loaded[k][l]; }
            if (!row_any) {
                continue;
            }
            OrPoints *p = &points[k];
            FOR_LANES
            {
                float vx, vy, vz;
                rel_velocity(&a, &b, p, l, &vx, &vy, &vz);
                const float vn = dot3(vx, vy, vz, g->normal[0][l], g->normal[1][l], g->normal[2][l]);
                const float rel = p->relative_velocity[l];
                const float acc = p->normal_impulse[l];
                const float impulse = -(p->normal_mass[l] * (vn + g->restitution[l] * rel));
                const float total = fmax_ref(acc + impulse, 0.0F);
                const float delta = loaded[k][l] ? total - acc : 0.0F;
                p->normal_impulse[l] = acc + delta;
                p->max_normal_impulse[l] = fmax_ref(p->max_normal_impulse[l], acc + delta);
                apply(&a, &b, g->inv_mass_a[l], g->inv_mass_b[l], delta, g->normal[0][l],
                      g->normal[1][l], g->normal[2][l], p->ang_a_n, p->ang_b_n, l);
            }
        }
    }
    if (g->scale_hits != 0U) {
        FOR_LANES { mask_store(g->impacted, l, mask_set(g->impacted, l) || bounced[l]); }
    }
    scatter(bodies, g->body_a, g->write_mask_a, &a);
    scatter(bodies, g->body_b, g->write_mask_b, &b);
}

void oracle_make_soft(float hertz, float zeta, float h, float out[3])
{
    if (hertz <= 0.0F) {
        out[0] = 0.0F;
        out[1] = 1.0F;
        out[2] = 0.0F;
        return;
    }
    const float omega = 2.0F * 3.14159265F * hertz;
    const float a1 = 2.0F * zeta + h * omega;
    const float a2 = h * omega * a1;
    const float a3 = 1.0F / (1.0F + a2);
    out[0] = omega / a1;
    out[1] = a2 * a3;
    out[2] = a3;
}

// ---------------------------------------------------------------------------
// nph_hash64: 64x64 -> 128 multiplies folded (wyhash), 48-byte blocks in
// three lanes, short inputs by overlapping reads.
// ---------------------------------------------------------------------------
static uint64_t mix(uint64_t a, uint64_t b)
{
    const unsigned __int128 r = (unsigned __int128)a * b;
    return (uint64_t)r ^ (uint64_t)(r >> 64);
}

static uint64_t rd64(const unsigned char *p)
{
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}

static uint64_t rd32(const unsigned char *p)
{
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

uint64_t oracle_hash64(const void *data, uint64_t len, uint64_t seed)
{
    static const uint64_t k[4] = {0x9E3779B97F4A7C15ULL, 0x6A09E667F3BCC909ULL,
                                  0xBB67AE8584CAA73BULL, 0x3C6EF372FE94F82BULL};
    const unsigned char *p = (const unsigned char *)data;
    seed ^= mix(seed ^ k[0], k[1]);
    uint64_t a = 0;
    uint64_t b = 0;
    if (len <= 16) {
        if (len >= 4) {
            const uint64_t step = (len >> 3) << 2;
            a = (rd32(p) << 32) | rd32(p + step);
            b = (rd32(p + len - 4) << 32) | rd32(p + len - 4 - step);
        } else if (len > 0) {
            a = ((uint64_t)p[0] << 16) | ((uint64_t)p[len >> 1] << 8) | p[len - 1];
        }
    } else {
        uint64_t i = len;
        if (i > 48) {
            uint64_t s1 = seed;
            uint64_t s2 = seed;
            do {
                seed = mix(rd64(p) ^ k[1], rd64(p + 8) ^ seed);
                s1 = mix(rd64(p + 16) ^ k[2], rd64(p + 24) ^ s1);
                s2 = mix(rd64(p + 32) ^ k[3], rd64(p + 40) ^ s2);
                p += 48;
                i -= 48;
            } while (i > 48);
            seed ^= s1 ^ s2;
        }
        while (i > 16) {
            seed = mix(rd64(p) ^ k[1], rd64(p + 8) ^ seed);
            i -= 16;
            p += 16;
        }
        a = rd64(p + i - 16);
        b = rd64(p + i - 8);
    }
    a ^= k[1];
    b ^= seed;
    const unsigned __int128 r = (unsigned __int128)a * b;
    return mix((uint64_t)r ^ k[0] ^ len, (uint64_t)(r >> 64) ^ k[1]);
}
